"""Lanes on the GPU (Tuning::inflightBatches, flow/Tuning.hpp:34-41; sched/GraphRunner.hpp:766-771,
1318-1330): the host runner with D lanes per node, and the wide and one-workgroup step engines
running up to D batches per node and step. Every batch record (inputs, counts, output chunks,
generations, fills) must equal the oracle's schedule with the same depth."""
import numpy as np
import pytest

import graphflow_amd as g
import oracle as O
from gpu_util import SEED, assert_same_trace, feedback_flow, run_flow, run_oracle, torch_cuda

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("depth", [1, 2, 4])
@pytest.mark.parametrize("schedule,engine,fuse", [(g.SCHEDULE_HOST, 0, False), (g.SCHEDULE_DEVICE, 1, False),
                                                  (g.SCHEDULE_DEVICE, 1, True), (g.SCHEDULE_DEVICE, 2, True)])
def test_feedback_lanes_trace(depth, schedule, engine, fuse):
    torch = torch_cuda()
    N, C, P, Q = 20000, 512, 4096, 4
    flow, src, rec = feedback_flow(N)
    stats, trace = run_flow(flow, src, N, C=C, P=P, Q=Q, schedule=schedule, fuse=fuse, engine=engine,
                            inflight_batches=depth)
    ref = run_oracle(flow, N, C=C, P=P, Q=Q, schedule=schedule, fuse=fuse, record=True, inflight=depth)
    assert ref["status"] == O.OK, ref["message"]
    assert_same_trace(trace, ref["trace"])
    assert stats["steps"] == ref["steps"]
    w, _ = O.feedback_scalar(SEED, N)
    torch.cuda.synchronize()
    assert (rec.cpu().numpy().view(np.uint32) == w).all()


@pytest.mark.parametrize("depth,chunks", [(2, 1), (4, 1), (2, 4)])
def test_hash_pipeline_lanes_trace(depth, chunks):
    """Config 3 shape with several batches of the 4096-bucket partition per step (and source lanes
    of several chunks)."""
    torch = torch_cuda()
    N, B, C, Q = 300000, 4096, 8192, 4
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
    P = 1 << 17
    flow.configure(g.Tuning(seed=SEED, batch_chunks=Q, chunk_capacity=C, live_particles=P, schedule=g.SCHEDULE_DEVICE,
                            trace=True, inflight_batches=depth, device_engine=1, source_chunks=chunks))
    st = flow.run(src, N)
    ref = run_oracle(flow, N, C=C, P=P, Q=Q, schedule=g.SCHEDULE_DEVICE, digest_cap=cap, multiset_bins=B,
                     inflight=depth, source_chunks=chunks)
    assert ref["status"] == O.OK, ref["message"]
    assert_same_trace(flow.trace(), ref["trace"])
    torch.cuda.synchronize()
    ms = mset.cpu().numpy().view(np.uint64).reshape(-1, 2)
    assert (ms[:, 0] == ref["multiset"]).all() and int(ms[:, 1].sum()) == N == st["sunk"]


@pytest.mark.parametrize("chunks,depth,N,P", [(2, 1, 30011, 4096), (8, 2, 30011, 4096), (4, 2, 5000, 300),
                                              (8, 2, 4097, 8192)])
@pytest.mark.parametrize("engine", [1, 2])
def test_source_chunks_trace(chunks, depth, N, P, engine):
    """Source lanes admitting several chunks per step (wide and one-workgroup engines): every batch
    record and every final weight equals the oracle's step schedule with the same setting, also
    with a live bound below C and a partial tail chunk."""
    torch = torch_cuda()
    C, Q = 512, 4
    flow, src, rec = feedback_flow(N)
    stats, trace = run_flow(flow, src, N, C=C, P=P, Q=Q, schedule=g.SCHEDULE_DEVICE, fuse=True, engine=engine,
                            inflight_batches=depth, source_chunks=chunks)
    ref = run_oracle(flow, N, C=C, P=P, Q=Q, schedule=g.SCHEDULE_DEVICE, fuse=True, record=True, inflight=depth,
                     source_chunks=chunks)
    assert ref["status"] == O.OK, ref["message"]
    assert_same_trace(trace, ref["trace"])
    assert stats["steps"] == ref["steps"]
    w, _ = O.feedback_scalar(SEED, N)
    torch.cuda.synchronize()
    assert (rec.cpu().numpy().view(np.uint32) == w).all()


def test_team_engine_rejects_lanes():
    flow, src, _ = feedback_flow(1000)
    with pytest.raises(ValueError):
        flow.configure(g.Tuning(seed=SEED, schedule=g.SCHEDULE_DEVICE, device_engine=3, inflight_batches=2))
    with pytest.raises(ValueError):
        flow.configure(g.Tuning(seed=SEED, schedule=g.SCHEDULE_DEVICE, device_engine=3, source_chunks=2))


@pytest.mark.parametrize("Q", [1, 16])
def test_config3_full_size_bucket_multisets(Q):
    """Config 3 at full size: 1e8 Zipf(1.1) keys over 2^20 values into 4096 murmur3 buckets, C=8192,
    32 lanes per node; every bucket's item count and item-hash sum equal the oracle's schedule-free
    model on the same seed."""
    torch = torch_cuda()
    N, B, C, D, P = 100_000_000, 4096, 8192, 32, 1 << 21
    mset = torch.zeros(2 * B, dtype=torch.int64, device="cuda")
    flow = g.Flow(g.PAYLOAD_KEYED)
    src = flow.source("keys", g.ZipfKeys(1 << 20, 1.1))
    part = flow.partition("bucket", g.KeyBucket(B))
    sink = flow.sink("digest", g.ChunkDigest(0, 0, mset.data_ptr(), B))
    flow.connect(src, part)
    for b in range(B):
        flow.connect(part, sink, b)
    S = min(524000, P // C + 2 * D * (Q + B) + 64)
    flow.configure(g.Tuning(seed=SEED, batch_chunks=Q, chunk_capacity=C, live_particles=P, num_chunks=S,
                            schedule=g.SCHEDULE_DEVICE, device_engine=1, max_workgroups=1024, tile_items=4096,
                            inflight_batches=D))
    st = flow.run(src, N)
    assert st["sunk"] == N
    ms, ct = O.keyed_scalar(SEED, N, threads=16)
    got = mset.cpu().numpy().view(np.uint64).reshape(-1, 2)
    assert (got[:, 1] == ct).all() and (got[:, 0] == ms).all()
