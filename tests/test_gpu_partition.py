"""GPU parity of the partition hot path beyond the feedback network: the kernel-axis histograms and
their published SHA-256 values, the scripted routing fixture, the murmur3/Zipf hash pipeline,
the stage-level C ABI (launch / advance / finish, device tile flattening) and the error contract."""
import hashlib

import numpy as np
import pytest

import graphflow_amd as g
import oracle as O
from gpu_util import SEED, assert_same_trace, decode_trace, run_oracle, torch_cuda

pytestmark = pytest.mark.gpu


def _kernel_axis_flow(N, bins, dist):
    torch = torch_cuda()
    counts = torch.zeros(bins, dtype=torch.int64, device="cuda")
    flow = g.Flow(g.PAYLOAD_PARTICLE)
    src = flow.source("generate", g.UniformParticles())
    part = flow.partition("partition", g.IdMod(bins, dist, update="GEN_TO_BIN"))
    sink = flow.sink("histogram", g.WeightHistogram(counts.data_ptr(), 0, bins, field_=2))
    flow.connect(src, part)
    for b in range(bins):
        flow.connect(part, sink, b)
    return flow, src, counts


@pytest.mark.parametrize("N,bins,dist,sha", [
    (1000000, 64, 0, "f3a977d720ab00c985bb39e0fffab715f57696fe6a762412e14e3f3810ce3505"),
    (1000000, 64, 1, "e59ef1c49b3c32c8714b7d10b1aabc17be313c9124ebee3c53e1337afe7678b4"),
    (1000000, 64, 2, "ab5932dacd593dc0ff4734c4903dd396da7d07f00c375aeefc0f83e6224200d2"),
    (1000000, 1024, 0, "8b946fbb08966136cf9f17d9547ceb262187e11781345e184da1c0798d2de9a6"),
    (262144, 4096, 0, "06f1bd73ff252b1a8f2f51aa879ad8195ac6816d0627d8edb5071fe4dc441e55"),
])
def test_kernel_axis_histogram_sha(N, bins, dist, sha):
    """tools/BENCH_TABLE.md:215-222: the reference's published histogram SHA-256 values, reproduced
    through the device engine (N, bins and distributions of the reference's kernel-axis runs,
    C=4096, P=16384, Q=8)."""
    flow, src, counts = _kernel_axis_flow(N, bins, dist)
    flow.configure(g.Tuning(seed=SEED, batch_chunks=8, chunk_capacity=4096, live_particles=16384,
                            schedule=g.SCHEDULE_DEVICE))
    st = flow.run(src, N)
    assert st["sunk"] == N
    rows = g.hist_rows(counts.data_ptr(), bins, 0)
    assert hashlib.sha256(g.rows_text(rows)).hexdigest() == sha


def test_kernel_axis_trace_parity_plain():
    N, bins = 20000, 64
    flow, src, counts = _kernel_axis_flow(N, bins, 2)
    flow.configure(g.Tuning(seed=SEED, batch_chunks=2, chunk_capacity=1024, live_particles=4096,
                            schedule=g.SCHEDULE_HOST, trace=True))
    flow.run(src, N)
    ref = run_oracle(flow, N, C=1024, P=4096, Q=2, schedule=g.SCHEDULE_HOST, hist_size=bins)
    assert ref["status"] == O.OK
    assert_same_trace(flow.trace(), ref["trace"])


@pytest.mark.parametrize("schedule", [g.SCHEDULE_HOST, g.SCHEDULE_DEVICE])
@pytest.mark.parametrize("capacity", [1, 3])
def test_scripted_routing_fixture(schedule, capacity):
    """tests/integration/test_scripted_routing.cpp:172-183: weights {0,1000,1,1001} over the U(0,1]
    start weight, rows {(0,1),(1,1),(1000,1),(1001,1)}."""
    torch = torch_cuda()
    filt = [0, 0, 0, 0, 1, 1, 1, 1, 1, 0, 1, 0]
    a = [0, 0, 0, 1, 0, 0, 0, 0, 0, 1, 0, 0]
    b = [0, 0, 0, 0, 0, 0, 1, 0, 0, 0, 1, 0]
    lut = torch.tensor(filt + a + b, dtype=torch.int32, device="cuda")
    rec = torch.full((4,), -1, dtype=torch.int32, device="cuda")
    flow = g.Flow(g.PAYLOAD_PARTICLE)
    src = flow.source("generate", g.UniformParticles())
    f = flow.partition("filter", g.Lut(2, lut.data_ptr(), 4, 3, 0))
    pa = flow.partition("A", g.Lut(2, lut.data_ptr(), 4, 3, 12, update="ADD_ON_BIN", add_bin=1, add=1000.0))
    pb = flow.partition("B", g.Lut(2, lut.data_ptr(), 4, 3, 24, update="ADD_ON_BIN", add_bin=1, add=1.0))
    sink = flow.sink("sink", g.RecordWeight(rec.data_ptr(), 4))
    flow.connect(src, f)
    flow.connect(f.out(0), pa)
    flow.connect(f.out(1), pb)
    for p in (pa, pb):
        flow.connect(p.out(0), sink)
        flow.connect(p.out(1), f)
    flow.configure(g.Tuning(seed=SEED, batch_chunks=2, chunk_capacity=capacity, live_particles=4,
                            schedule=schedule, trace=True))
    flow.run(src, 4)
    torch.cuda.synchronize()
    final = rec.cpu().numpy().view(np.float32)
    w0 = np.array([O.lib().or_uniform_f32(O.stream_words(SEED, i, 0, 0)[0]) for i in range(4)], np.float32)
    assert (final - w0).round().tolist() == [0.0, 1000.0, 1.0, 1001.0]
    ref = run_oracle(flow, 4, C=capacity, P=4, Q=2, schedule=schedule, record=True,
                     lut=filt + a + b)
    assert_same_trace(flow.trace(), ref["trace"])


@pytest.mark.parametrize("Q", [1, 4, 8, 16])
def test_hash_pipeline_zipf(Q):
    """Config 3 shape: Zipf(1.1) keys over 2^20 values, murmur3 fmix32 into 4096 buckets, C=8192.
    Exact per-batch counts and slots (trace), the per-chunk order-sensitive digest log (within-chunk
    order) and every bucket's multiset hash and count against the oracle."""
    torch = torch_cuda()
    N, B, C = 200000, 4096, 8192
    cap = 1 << 17
    log = torch.zeros(4 * cap, dtype=torch.int32, device="cuda")
    mset = torch.zeros(2 * B, dtype=torch.int64, device="cuda")
    flow = g.Flow(g.PAYLOAD_KEYED)
    src = flow.source("keys", g.ZipfKeys(1 << 20, 1.1))
    part = flow.partition("bucket", g.KeyBucket(B))
    sink = flow.sink("digest", g.ChunkDigest(log.data_ptr(), cap, mset.data_ptr(), B))
    flow.connect(src, part)
    for b in range(B):
        flow.connect(part, sink, b)
    flow.configure(g.Tuning(seed=SEED, batch_chunks=Q, chunk_capacity=C, schedule=g.SCHEDULE_DEVICE, trace=True))
    st = flow.run(src, N)
    ref = run_oracle(flow, N, C=C, P=0, Q=Q, schedule=g.SCHEDULE_DEVICE, digest_cap=cap, multiset_bins=B)
    assert ref["status"] == O.OK, ref["message"]
    assert_same_trace(flow.trace(), ref["trace"])
    torch.cuda.synchronize()
    n = ref["digest_len"]
    got = log.cpu().numpy().view(np.uint32).reshape(-1, 4)[:n]
    assert (got == ref["digest"]).all()
    ms = mset.cpu().numpy().view(np.uint64).reshape(-1, 2)
    assert (ms[:, 0] == ref["multiset"]).all()
    assert (ms[:, 1] == ref["multiset_items"]).all()
    assert int(ms[:, 1].sum()) == N == st["sunk"]


def test_stage_abi_partition_batch():
    """The operator boundary alone: a pool, one partition stage, a ragged batch of three chunks.
    launch -> advance (plan + scatter) -> advance (finalise) -> finish; per-bin counts, output
    descriptors, every output slot and the device tile flattening against the oracle."""
    C, B, T = 1000, 7, 1024
    pool = g.Pool(g.PAYLOAD_PARTICLE, 64, C)
    fills = [1000, 37, 613]
    handles = [pool.acquire() for _ in fills]
    rng = np.random.default_rng(7)
    ids_all, gen_all = [], []
    for h, fill in zip(handles, fills):
        w = rng.random(fill, dtype=np.float32)
        ids = rng.integers(0, 2**32, fill, dtype=np.uint64).astype(np.uint32)
        gens = np.arange(fill, dtype=np.uint32)
        base = pool.base(h)
        pool.write(0, base, w)
        pool.write(1, base, ids)
        pool.write(2, base, gens)
        ids_all.append(ids)
        gen_all.append(gens)
    ids_all = np.concatenate(ids_all)
    node = g.IdMod(B, 0).desc()
    stage = g.Stage(pool, node, node_key=1, seed=SEED, batch_chunks=4, tile_items=T)
    chunks = [(h.index, h.generation, f, 0) for h, f in zip(handles, fills)]
    stage.launch(chunks)
    tiles = stage.tiles()
    rc, ref_tiles = O.flatten_tiles([h.index for h in handles], fills, 256 * 4)
    assert [list(t) for t in tiles] == ref_tiles
    more, allocated = stage.advance()
    assert more and allocated > 0
    more, _ = stage.advance()
    assert not more
    outs = stage.finish()
    counts = stage.counts(B)
    rc, bins, tot, rank = O.partition_items(dict(role=1, op="CLS_ID_MOD", bins=B, u=[0]), len(ids_all),
                                             ids=ids_all)
    assert counts == tot.tolist()
    rc, base, out_fills, R = O.plan_dense_bins(tot.tolist(), C)
    assert [o[2] for o in outs] == out_fills and len(outs) == R == allocated
    assert [o[3] for o in outs] == [b for b in range(B) for _ in range(-(-int(tot[b]) // C))]
    # every item sits at table[binBase[bin] + j / C], j % C with j its stable rank in the bin
    got_ids = np.zeros(len(ids_all), np.uint32)
    out_ids = {}
    for r, o in enumerate(outs):
        a = np.zeros(o[2], np.uint32)
        pool.read(1, o[0] * C, a)
        out_ids[r] = a
    for i in range(len(ids_all)):
        b, j = int(bins[i]), int(rank[i])
        got_ids[i] = out_ids[base[b] + j // C][j % C]
    assert (got_ids == ids_all).all()
    for (h, _) in zip(handles, fills):
        pool.release(h)
    for o in outs:
        pool.release(g.ChunkHandle(o[0], o[1]))
    assert pool.counts()["free"] == 64


@pytest.mark.parametrize("schedule,engine", [(g.SCHEDULE_HOST, 0), (g.SCHEDULE_DEVICE, 1), (g.SCHEDULE_DEVICE, 2)])
def test_invalid_classifier_bin_is_a_batch_error(schedule, engine):
    """test_dense_bins.cpp:220-258: an out-of-range classifier result fails the batch (BatchError);
    no partial output is published and the flow can run again."""
    torch = torch_cuda()
    lut = torch.tensor([0, 1, 5, 1], dtype=torch.int32, device="cuda")  # id 2 -> bin 5 of 2
    rec = torch.full((64,), -1, dtype=torch.int32, device="cuda")
    flow = g.Flow(g.PAYLOAD_PARTICLE)
    src = flow.source("generate", g.UniformParticles())
    f = flow.partition("filter", g.Lut(2, lut.data_ptr(), 4, 1, 0))
    sink = flow.sink("sink", g.RecordWeight(rec.data_ptr(), 64))
    flow.connect(src, f)
    flow.connect(f.out(0), sink)
    flow.connect(f.out(1), sink)
    flow.configure(g.Tuning(seed=SEED, batch_chunks=2, chunk_capacity=16, live_particles=64, schedule=schedule,
                            device_engine=engine))
    with pytest.raises(g.BatchError):
        flow.run(src, 64)
    # ids 0 and 1 only: valid, the same flow object runs again
    assert flow.run(src, 2)["sunk"] == 2


@pytest.mark.parametrize("schedule,engine,fused", [(g.SCHEDULE_HOST, 0, False), (g.SCHEDULE_HOST, 0, True),
                                                   (g.SCHEDULE_DEVICE, 1, False), (g.SCHEDULE_DEVICE, 2, False),
                                                   (g.SCHEDULE_DEVICE, 3, False)])
def test_sink_record_out_of_range_is_a_batch_error(schedule, engine, fused):
    """A sink record outside its buffer fails the run (BatchError), also where the host runner
    completes sink batches without waiting for them; the pool is whole again and the same flow
    runs again (test_failure_paths.cpp:228-366: a failed batch fails the run and restores the pool)."""
    torch = torch_cuda()
    N = 6000
    rec = torch.full((N // 2,), -1, dtype=torch.int32, device="cuda")
    flow = g.Flow(g.PAYLOAD_PARTICLE)
    src = flow.source("generate", g.UniformParticles())
    part = flow.partition("split", g.IdMod(2, 0))
    sink = flow.sink("record", g.RecordWeight(rec.data_ptr(), N // 2))
    flow.connect(src, part)
    flow.connect(part, sink, 0)
    flow.connect(part, sink, 1)
    flow.configure(g.Tuning(seed=SEED, batch_chunks=2, chunk_capacity=512, live_particles=2048, schedule=schedule,
                            device_engine=engine, fused_batches=fused))
    with pytest.raises(g.BatchError):
        flow.run(src, N)
    st = flow.run(src, N // 2)
    torch.cuda.synchronize()
    assert st["sunk"] == N // 2 and (rec.cpu().numpy() != -1).all()


@pytest.mark.parametrize("schedule", [g.SCHEDULE_HOST, g.SCHEDULE_DEVICE])
def test_updater_overflow_is_a_batch_error(schedule):
    """test_failure_paths.cpp:1061-1085: a checked updater overflow fails the batch."""
    torch = torch_cuda()
    lut = torch.tensor([1], dtype=torch.int32, device="cuda")
    rec = torch.full((16,), -1, dtype=torch.int32, device="cuda")
    flow = g.Flow(g.PAYLOAD_PARTICLE)
    src = flow.source("generate", g.UniformParticles())
    f = flow.partition("boost", g.Lut(2, lut.data_ptr(), 1, 1, 0, update="ADD_ON_BIN", add_bin=1, add=3.0e38))
    f2 = flow.partition("boost again", g.Lut(2, lut.data_ptr(), 1, 1, 0, update="ADD_ON_BIN", add_bin=1, add=3.0e38))
    sink = flow.sink("sink", g.RecordWeight(rec.data_ptr(), 16))
    flow.connect(src, f)
    flow.connect(f.out(0), sink)
    flow.connect(f.out(1), f2)
    flow.connect(f2.out(0), sink)
    flow.connect(f2.out(1), sink)
    flow.configure(g.Tuning(seed=SEED, batch_chunks=1, chunk_capacity=16, live_particles=16, schedule=schedule))
    with pytest.raises(g.BatchError):
        flow.run(src, 16)


@pytest.mark.parametrize("schedule", [g.SCHEDULE_HOST, g.SCHEDULE_DEVICE])
def test_pool_too_small_is_a_stall(schedule):
    """A pool below S = P + K*R cannot make progress: SchedulerStallError, never a hang."""
    torch = torch_cuda()
    counts = torch.zeros(8, dtype=torch.int64, device="cuda")
    flow = g.Flow(g.PAYLOAD_PARTICLE)
    src = flow.source("generate", g.UniformParticles())
    p = flow.partition("split", g.IdMod(8, 0))
    sink = flow.sink("sink", g.WeightHistogram(counts.data_ptr(), 0, 8, field_=2))
    flow.connect(src, p)
    for b in range(8):
        flow.connect(p, sink, b)
    flow.configure(g.Tuning(seed=SEED, batch_chunks=1, chunk_capacity=8, live_particles=64, num_chunks=3,
                            schedule=schedule))
    with pytest.raises(g.SchedulerStallError):
        flow.run(src, 64)


def test_registration_errors():
    flow = g.Flow(g.PAYLOAD_PARTICLE)
    src = flow.source("generate", g.UniformParticles())
    with pytest.raises(ValueError):
        flow.source("second", g.UniformParticles())
    f = flow.partition("filter", g.WeightRoute())
    with pytest.raises(ValueError):
        flow.connect(f, src, 0)  # a route may not target the source
    with pytest.raises(ValueError):
        flow.connect(f, f)  # multi-output node needs out<Bin>()
    flow.connect(src, f)
    with pytest.raises(ValueError):
        flow.run(src, 10)  # missing routes of the filter


def test_descriptor_buffers_are_validated():
    """A table classifier whose table is shorter than what it indexes, and a photon step whose tally
    buffer holds fewer than the run's tally bins, are rejected before any kernel runs."""
    torch = torch_cuda()
    lut = torch.zeros(4, dtype=torch.int32, device="cuda")
    flow = g.Flow(g.PAYLOAD_PARTICLE)
    src = flow.source("generate", g.UniformParticles())
    f = flow.partition("filter", g.Lut(2, lut.data_ptr(), 4, 3, 0, table_len=4))
    rec = torch.zeros(16, dtype=torch.int32, device="cuda")
    sink = flow.sink("sink", g.RecordWeight(rec.data_ptr(), 16))
    flow.connect(src, f)
    flow.connect(f.out(0), sink)
    flow.connect(f.out(1), sink)
    flow.configure(g.Tuning(seed=SEED, chunk_capacity=16, schedule=g.SCHEDULE_DEVICE))
    with pytest.raises(ValueError, match="table"):
        flow.run(src, 16)
    pool = g.Pool(g.PAYLOAD_PARTICLE, 4, 16)
    with pytest.raises(ValueError):
        g.Stage(pool, g.Lut(2, lut.data_ptr(), 4, 3, 0, table_len=4).desc(), 1, SEED, 1)
    tally = torch.zeros(16, dtype=torch.float64, device="cuda")
    pf = g.Flow(g.PAYLOAD_PHOTON)
    psrc = pf.source("generate", g.PhotonSource())
    step = pf.transform("step", g.PhotonStep(tally.data_ptr(), 16))
    alive = pf.partition("alive", g.PhotonAlive())
    done = pf.sink("done", g.PhotonDone())
    pf.connect(psrc, step)
    pf.connect(step, alive)
    pf.connect(alive.out(0), step)
    pf.connect(alive.out(1), done)
    pf.configure(g.Tuning(seed=SEED, chunk_capacity=64, schedule=g.SCHEDULE_DEVICE))  # 1024 tally bins
    with pytest.raises(ValueError, match="tally"):
        pf.run(psrc, 64)
