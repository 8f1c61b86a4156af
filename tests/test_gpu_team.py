"""GPU parity of the team engine (one persistent launch, a team of workgroups on one XCD, a
replicated scheduler and producer-side segment counts): every batch of the trace equals the
oracle's step-mode trace, across team sizes, relaunches, hand-over to the wide engine, the keyed
hash pipeline and the error contract."""
import numpy as np
import pytest

import graphflow_amd as g
import oracle as O
from gpu_util import SEED, assert_same_trace, feedback_flow, log, run_flow, run_oracle, torch_cuda

pytestmark = pytest.mark.gpu


def _feedback_parity(N, C, P, Q, fuse, **kw):
    torch = torch_cuda()
    log(f"[team] feedback N={N} C={C} P={P} Q={Q} fuse={fuse} {kw}")
    flow, src, rec = feedback_flow(N)
    stats, trace = run_flow(flow, src, N, C=C, P=P, Q=Q, schedule=g.SCHEDULE_DEVICE, fuse=fuse, engine=3, **kw)
    ref = run_oracle(flow, N, C=C, P=P, Q=Q, schedule=g.SCHEDULE_DEVICE, fuse=fuse, record=True)
    assert ref["status"] == O.OK, ref["message"]
    assert_same_trace(trace, ref["trace"])
    torch.cuda.synchronize()
    w, _ = O.feedback_scalar(SEED, N)
    assert (rec.cpu().numpy().view(np.uint32) == w).all()
    assert stats["sunk"] == N and stats["generated"] == N
    assert stats["steps"] == ref["steps"]
    return stats


@pytest.mark.parametrize("team", [1, 2, 7, 16, 32])
def test_team_sizes(team):
    _feedback_parity(40000, 8192, 16384, 8, True, max_workgroups=team)


@pytest.mark.parametrize("steps_per_launch", [1, 7, 64])
def test_relaunch_keeps_state(steps_per_launch):
    """Short launches: the state (rings, free-stack window, counters, pending count-row clears)
    round-trips through global memory between launches."""
    _feedback_parity(12000, 1024, 2048, 4, True, steps_per_launch=steps_per_launch)


@pytest.mark.parametrize("C", [1, 3, 127, 128, 129, 255, 256, 257, 1000])
def test_segment_edges(C):
    """Capacities below, at and around the 128-slot segment and twice it (a chunk of one
    segment, of one segment and a one-item tail, of two), ragged tails, non-power-of-two C."""
    N = 3000 if C >= 255 else 600
    _feedback_parity(N, C, max(2 * C, 8), 3, False)


def test_handoff_to_wide_engine():
    """C=1 with Q=600: batches of hundreds of chunks exceed the team's per-step limits, so the run
    continues on the wide engine from the flushed state, with the same trace."""
    st = _feedback_parity(3000, 1, 600, 600, False)
    assert st["steps"] > 0


@pytest.mark.parametrize("bins", [4, 16, 31])
def test_keyed_hash_pipeline(bins):
    """Keyed payload through a narrow murmur3 bucket partition: per-batch counts and slots, the
    per-chunk order digests and every bucket's multiset."""
    torch = torch_cuda()
    N, C, Q = 50000, 1024, 4
    cap = 1 << 15
    logb = torch.zeros(4 * cap, dtype=torch.int32, device="cuda")
    mset = torch.zeros(2 * bins, dtype=torch.int64, device="cuda")
    flow = g.Flow(g.PAYLOAD_KEYED)
    src = flow.source("keys", g.ZipfKeys(1 << 20, 1.1))
    part = flow.partition("bucket", g.KeyBucket(bins))
    sink = flow.sink("digest", g.ChunkDigest(logb.data_ptr(), cap, mset.data_ptr(), bins))
    flow.connect(src, part)
    for b in range(bins):
        flow.connect(part, sink, b)
    log(f"[team] keyed bins={bins}")
    flow.configure(g.Tuning(seed=SEED, batch_chunks=Q, chunk_capacity=C, schedule=g.SCHEDULE_DEVICE, trace=True,
                            device_engine=3))
    st = flow.run(src, N)
    ref = run_oracle(flow, N, C=C, P=0, Q=Q, schedule=g.SCHEDULE_DEVICE, digest_cap=cap, multiset_bins=bins)
    assert ref["status"] == O.OK, ref["message"]
    assert_same_trace(flow.trace(), ref["trace"])
    torch.cuda.synchronize()
    n = ref["digest_len"]
    got = logb.cpu().numpy().view(np.uint32).reshape(-1, 4)[:n]
    assert (got == ref["digest"]).all()
    ms = mset.cpu().numpy().view(np.uint64).reshape(-1, 2)
    assert (ms[:, 0] == ref["multiset"]).all() and (ms[:, 1] == ref["multiset_items"]).all()
    assert st["sunk"] == N


def test_headline_shape_histogram():
    """C=8192, P=16384, Q=8, fused, N=2e6: sorted histogram rows equal the schedule-free model."""
    torch = torch_cuda()
    N = 2000000
    flow, src, counts = feedback_flow(N, sink="hist")
    log("[team] headline shape N=2e6")
    run_flow(flow, src, N, C=8192, P=16384, Q=8, schedule=g.SCHEDULE_DEVICE, fuse=True, engine=3, trace=False)
    rows = g.hist_rows(counts.data_ptr(), g.FEEDBACK_HIST_N, g.FEEDBACK_HIST_LO)
    w, _ = O.feedback_scalar(SEED, N, threads=8)
    v, c = np.unique(w, return_counts=True)
    assert len(rows) == len(v) and (rows[:, 0] == v).all() and (rows[:, 1] == c).all()
    assert int(rows[:, 1].sum()) == N


def test_invalid_bin_is_a_batch_error_and_nothing_is_published():
    """A classifier bin outside the partition's range is counted by the producer and fails the
    consuming batch at plan time."""
    torch = torch_cuda()
    lut = torch.tensor([0, 5], dtype=torch.int32, device="cuda")  # id 1 -> bin 5 of a 2-bin partition
    rec = torch.full((64,), -1, dtype=torch.int32, device="cuda")
    flow = g.Flow(g.PAYLOAD_PARTICLE)
    src = flow.source("generate", g.UniformParticles())
    part = flow.partition("p", g.Lut(2, lut.data_ptr(), 2, 1, 0))
    sink = flow.sink("sink", g.RecordWeight(rec.data_ptr(), 64))
    flow.connect(src, part)
    flow.connect(part, sink, 0)
    flow.connect(part, sink, 1)
    flow.configure(g.Tuning(seed=SEED, batch_chunks=2, chunk_capacity=16, live_particles=64,
                            schedule=g.SCHEDULE_DEVICE, device_engine=3))
    with pytest.raises(g.BatchError):
        flow.run(src, 64)
    torch.cuda.synchronize()
    assert (rec.cpu().numpy() == -1).all()


@pytest.mark.parametrize("fuse", [False, True])
def test_two_partitions_and_later_transforms_in_one_step(fuse):
    """Two partitions and (unfused) two transforms after them in node order are active in the same
    steps, so a scatter wave's output range includes the outputs of the partitions before its node,
    which it derives from their count rows. Trace and final weights equal the oracle's."""
    torch = torch_cuda()
    N = 20000
    rec = torch.full((N,), -1, dtype=torch.int32, device="cuda")
    flow = g.Flow(g.PAYLOAD_PARTICLE)
    src = flow.source("generate", g.UniformParticles())
    p1 = flow.partition("route", g.WeightRoute(1e-3, "parity", 2))
    p2 = flow.partition("split", g.IdMod(2))
    a = flow.transform("A", g.ScaleWeight(0.5))
    b = flow.transform("B", g.ScaleWeight(0.75))
    sink = flow.sink("sink", g.RecordWeight(rec.data_ptr(), N))
    flow.connect(src, p1)
    flow.connect(p1.out(0), sink)
    flow.connect(p1.out(1), a)
    flow.connect(p1.out(2), p2)
    flow.connect(p2.out(0), a)
    flow.connect(p2.out(1), b)
    flow.connect(a, p1)
    flow.connect(b, p1)
    log(f"[team] two partitions fuse={fuse}")
    stats, trace = run_flow(flow, src, N, C=256, P=2048, Q=4, schedule=g.SCHEDULE_DEVICE, fuse=fuse, engine=3)
    ref = run_oracle(flow, N, C=256, P=2048, Q=4, schedule=g.SCHEDULE_DEVICE, fuse=fuse, record=True)
    assert ref["status"] == O.OK, ref["message"]
    assert stats["engine"] == 3
    assert_same_trace(trace, ref["trace"])
    torch.cuda.synchronize()
    assert (rec.cpu().numpy().view(np.uint32) == ref["record"]).all()
    assert stats["sunk"] == N


@pytest.mark.parametrize("fuse", [False, True])
def test_batches_of_more_than_64_inputs(fuse):
    """C=1, Q=200: partition batches of up to 200 one-item chunks (more than one wave of inputs and
    more than 128 segments per batch), still within the team's per-step limits; the trace and the
    final weights equal the oracle's."""
    st = _feedback_parity(3000, 1, 600, 200, fuse)
    assert st["engine"] == 3
