"""The fused single-launch partition batch (gf_transform_batch): classify, device plan, device output
acquisition and stable scatter in one kernel, compared with the reference's three-step protocol
(classify -> host plan -> scatter, flow/Execution.hpp:2363-2725) and with the oracle."""
import hashlib

import numpy as np
import pytest

import graphflow_amd as g
import oracle as O
from gpu_util import SEED, assert_same_trace, feedback_flow, run_flow, run_oracle, torch_cuda

pytestmark = pytest.mark.gpu


def _filled_pool(C, fills, seed=7, S=512):
    pool = g.Pool(g.PAYLOAD_PARTICLE, S, C)
    handles = [pool.acquire() for _ in fills]
    rng = np.random.default_rng(seed)
    ids_all = []
    for h, fill in zip(handles, fills):
        base = pool.base(h)
        ids = rng.integers(0, 2**32, fill, dtype=np.uint64).astype(np.uint32)
        pool.write(0, base, rng.random(fill, dtype=np.float32))
        pool.write(1, base, ids)
        pool.write(2, base, np.arange(fill, dtype=np.uint32))
        ids_all.append(ids)
    return pool, handles, np.concatenate(ids_all)


def _read_outputs(pool, outs, C):
    cols = []
    for col, dt in ((0, np.float32), (1, np.uint32), (2, np.uint32)):
        parts = []
        for o in outs:
            a = np.zeros(o[2], dt)
            pool.read(col, o[0] * C, a)
            parts.append(a)
        cols.append(np.concatenate(parts) if parts else np.zeros(0, dt))
    return cols


LONG = [1000, 37, 613, 999, 1000, 512, 1, 800, 1000, 250]


@pytest.mark.parametrize("B,T,fills", [(7, 1024, [1000, 37, 613]), (3, 256, [1000, 37, 613]),
                                       (300, 1024, [1000, 37, 613]),
                                       # 28 tiles: the grid form (count, plan, scatter launches)
                                       (7, 256, LONG), (3, 256, LONG), (300, 256, LONG)])
def test_one_ragged_batch_equals_three_step_protocol_and_oracle(B, T, fills):
    C = 1000
    torch_cuda()
    pa, ha, ids_all = _filled_pool(C, fills)
    pb, hb, _ = _filled_pool(C, fills)
    node = g.IdMod(B, 0).desc()
    # reference protocol: launch (classify) -> advance (host plan + scatter) -> advance -> finish
    sa = g.Stage(pa, node, node_key=1, seed=SEED, batch_chunks=16, tile_items=T)
    sa.launch([(h.index, h.generation, f, 0) for h, f in zip(ha, fills)])
    assert sa.advance()[0] and not sa.advance()[0]
    outs_a = sa.finish()
    # one launch -> one advance -> finish
    sb = g.Stage(pb, node, node_key=1, seed=SEED, batch_chunks=16, tile_items=T)
    sb.transform_batch([(h.index, h.generation, f, 0) for h, f in zip(hb, fills)])
    more, allocated = sb.advance()
    assert not more and allocated == len(outs_a)
    outs_b = sb.finish()
    assert outs_b == outs_a  # same chunk indices, generations, fills and bins
    ca, cb = _read_outputs(pa, outs_a, C), _read_outputs(pb, outs_b, C)
    for x, y in zip(ca, cb):
        assert (x.view(np.uint32) == y.view(np.uint32)).all()
    # oracle: counts, fills and the stable slot of every item
    rc, bins, tot, rank = O.partition_items(dict(role=1, op="CLS_ID_MOD", bins=B, u=[0]), len(ids_all), ids=ids_all)
    rc, base, out_fills, R = O.plan_dense_bins(tot.tolist(), C)
    assert [o[2] for o in outs_b] == out_fills and len(outs_b) == R
    flat = np.concatenate([[0], np.cumsum([o[2] for o in outs_b])])
    pos = np.array([flat[base[b] + j // C] + j % C for b, j in zip(bins.tolist(), rank.tolist())])
    assert (cb[1][pos] == ids_all).all()
    # the unused reservation went back: both pools hold the same live chunks
    assert pa.counts()["free"] == pb.counts()["free"] == 512 - len(fills) - R
    for h in hb:
        pb.release(h)
    for o in outs_b:
        pb.release(g.ChunkHandle(o[0], o[1]))
    assert pb.counts()["free"] == 512


@pytest.mark.parametrize("depth,Q,T", [(1, 4, None), (2, 4, None), (1, 8, 256), (2, 8, 256)])
def test_runner_with_fused_batches_trace(depth, Q, T):
    """The host GraphRunner driving every partition batch as one fused batch (one launch, or the grid
    form's three for batches of many tiles): every batch record equals the oracle's runner with
    fused stages."""
    torch = torch_cuda()
    N, C, P = 20000, 1024, 2048 if Q == 4 else 8192
    flow, src, rec = feedback_flow(N)
    stats, trace = run_flow(flow, src, N, C=C, P=P, Q=Q, schedule=g.SCHEDULE_HOST, fused_batches=True,
                            inflight_batches=depth, tile_items=T)
    ref = run_oracle(flow, N, C=C, P=P, Q=Q, schedule=g.SCHEDULE_HOST, record=True, fused_stages=True,
                     inflight=depth)
    assert ref["status"] == O.OK, ref["message"]
    assert_same_trace(trace, ref["trace"])
    w, _ = O.feedback_scalar(SEED, N)
    torch.cuda.synchronize()
    assert (rec.cpu().numpy().view(np.uint32) == w).all()


def test_fused_batches_65536_bins_sha():
    """The hash path inside the fused kernel: tools/BENCH_TABLE.md:216 (N=262144, B=65536) through the
    host runner with one launch per partition batch."""
    torch = torch_cuda()
    N, B = 262144, 65536
    counts = torch.zeros(B, dtype=torch.int64, device="cuda")
    flow = g.Flow(g.PAYLOAD_PARTICLE)
    src = flow.source("generate", g.UniformParticles())
    part = flow.partition("partition", g.IdMod(B, 0, update="GEN_TO_BIN"))
    sink = flow.sink("histogram", g.WeightHistogram(counts.data_ptr(), 0, B, field_=2))
    flow.connect(src, part)
    for b in range(B):
        flow.connect(part, sink, b)
    flow.configure(g.Tuning(seed=SEED, batch_chunks=8, chunk_capacity=4096, live_particles=16384,
                            schedule=g.SCHEDULE_HOST, fused_batches=True))
    assert flow.run(src, N)["sunk"] == N
    rows = g.hist_rows(counts.data_ptr(), B, 0)
    assert hashlib.sha256(g.rows_text(rows)).hexdigest() == \
        "968f4a008df5ad231cfb5578843033a9d4f2a28233f8568f0c7c6ed51fd638c0"


def test_fused_batch_invalid_bin_returns_the_reservation():
    torch = torch_cuda()
    lut = torch.tensor([0, 1, 5, 1], dtype=torch.int32, device="cuda")  # id 2 -> bin 5 of 2
    pool = g.Pool(g.PAYLOAD_PARTICLE, 16, 4)
    h = pool.acquire()
    pool.write(1, h.index * 4, np.arange(4, dtype=np.uint32))
    pool.write(2, h.index * 4, np.zeros(4, np.uint32))
    stage = g.Stage(pool, g.Lut(2, lut.data_ptr(), 4, 1, 0).desc(), 1, SEED, 1)
    stage.transform_batch([(h.index, h.generation, 4, 0)])
    with pytest.raises(g.BatchError):
        stage.advance()
    assert pool.counts()["free"] == 15
    assert [pool.generation(i) for i in range(16)] == [0] * 16


def test_fused_grid_batch_invalid_bin_returns_the_reservation():
    """The same error in the grid form (12 tiles of 256 items: a count launch over the grid reports
    it, the plan publishes nothing, the scatter launch does nothing); a new stage on the same pool
    then runs a valid batch of the same shape."""
    torch = torch_cuda()
    lut = torch.tensor([0, 1, 5, 1], dtype=torch.int32, device="cuda")  # id 2 -> bin 5 of 2
    ok = torch.tensor([0, 1, 1, 0], dtype=torch.int32, device="cuda")
    C, n = 256, 12
    pool = g.Pool(g.PAYLOAD_PARTICLE, 64, C)
    hs = [pool.acquire() for _ in range(n)]
    for h in hs:
        pool.write(0, h.index * C, np.ones(C, np.float32))
        pool.write(1, h.index * C, (np.arange(C) % 4).astype(np.uint32))
        pool.write(2, h.index * C, np.zeros(C, np.uint32))
    chunks = [(h.index, h.generation, C, 0) for h in hs]
    stage = g.Stage(pool, g.Lut(2, lut.data_ptr(), 4, 1, 0).desc(), 1, SEED, 16, tile_items=256)
    stage.transform_batch(chunks)
    with pytest.raises(g.BatchError):
        stage.advance()
    assert pool.counts()["free"] == 64 - n
    assert [pool.generation(i) for i in range(64)] == [0] * 64
    good = g.Stage(pool, g.Lut(2, ok.data_ptr(), 4, 1, 0).desc(), 1, SEED, 16, tile_items=256)
    good.transform_batch(chunks)
    more, allocated = good.advance()
    outs = good.finish()
    assert not more and allocated == len(outs) == 12 and [o[3] for o in outs] == [0] * 6 + [1] * 6
    assert all(o[2] == C for o in outs)
