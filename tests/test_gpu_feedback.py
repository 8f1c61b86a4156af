"""GPU parity of the fp32 feedback network (N=1e6 plain and N=1e7 batched workloads): every planned batch
(inputs, per-bin counts, output chunk indices, generations, fills, bins) and every final weight
must equal the oracle's, for the plain host-planned form and both device engines."""
import numpy as np
import pytest

import graphflow_amd as g
import oracle as O
from gpu_util import (SEED, assert_same_trace, decode_trace, feedback_flow, run_flow, run_oracle, torch_cuda)

pytestmark = pytest.mark.gpu

CASES = [
    # (N, C, P, Q, schedule, fuse, engine)
    (5000, 1024, 2048, 1, g.SCHEDULE_HOST, False, 0),
    (5000, 1024, 2048, 4, g.SCHEDULE_HOST, False, 0),
    (20000, 1024, 2048, 8, g.SCHEDULE_DEVICE, False, 1),
    (20000, 1024, 2048, 8, g.SCHEDULE_DEVICE, True, 1),
    (20000, 1024, 2048, 8, g.SCHEDULE_DEVICE, False, 2),
    (20000, 1024, 2048, 8, g.SCHEDULE_DEVICE, True, 2),
    (100000, 8192, 16384, 8, g.SCHEDULE_DEVICE, True, 2),
    (3000, 1000, 1500, 3, g.SCHEDULE_DEVICE, False, 1),  # non-power-of-two C, tail tiles
    (20000, 1024, 2048, 8, g.SCHEDULE_DEVICE, False, 3),
    (20000, 1024, 2048, 8, g.SCHEDULE_DEVICE, True, 3),
    (100000, 8192, 16384, 8, g.SCHEDULE_DEVICE, True, 3),
    (3000, 1000, 1500, 3, g.SCHEDULE_DEVICE, False, 3),  # non-power-of-two C, tail segments
]


@pytest.mark.parametrize("N,C,P,Q,schedule,fuse,engine", CASES)
def test_feedback_trace_and_weights(N, C, P, Q, schedule, fuse, engine):
    torch = torch_cuda()
    flow, src, rec = feedback_flow(N)
    stats, trace = run_flow(flow, src, N, C=C, P=P, Q=Q, schedule=schedule, fuse=fuse, engine=engine)
    ref = run_oracle(flow, N, C=C, P=P, Q=Q, schedule=schedule, fuse=fuse, record=True)
    assert ref["status"] == O.OK, ref["message"]
    assert_same_trace(trace, ref["trace"])
    w, _ = O.feedback_scalar(SEED, N)
    torch.cuda.synchronize()
    got = rec.cpu().numpy().view(np.uint32)
    assert (got == w).all()
    assert stats["sunk"] == N and stats["generated"] == N
    assert stats["pool_chunks"] == ref["pool_chunks"]
    # conservation of every partition batch: counts sum to the batch, fills sum to the counts
    for node, step, ins, counts, outs in decode_trace(trace):
        if len(counts) and len(ins):
            assert int(counts.sum()) == int(ins[:, 2].sum())
            assert int(outs[:, 2].sum()) == int(counts.sum())


def test_config1_histogram_rows_and_sha():
    """Workload N=1e6, C=1024, P=2048, Q=1, seed 42 on the plain form; sorted histogram
    rows bit-exact with the oracle, rows summing to N."""
    N = 1000000
    flow, src, counts = feedback_flow(N, sink="hist")
    run_flow(flow, src, N, C=1024, P=2048, Q=1, schedule=g.SCHEDULE_HOST, trace=False)
    rows = g.hist_rows(counts.data_ptr(), g.FEEDBACK_HIST_N, g.FEEDBACK_HIST_LO)
    w, _ = O.feedback_scalar(SEED, N, threads=8)
    v, c = np.unique(w, return_counts=True)
    assert len(rows) == len(v)
    assert (rows[:, 0] == v).all() and (rows[:, 1] == c).all()
    assert int(rows[:, 1].sum()) == N
    ref_rows = np.stack([v.astype(np.uint64), c.astype(np.uint64)], axis=1)
    assert g.rows_sha256(rows) == g.rows_sha256(ref_rows)


@pytest.mark.parametrize("engine", [1, 2, 3])
def test_config2_shape_histogram(engine):
    """Config 2 shape (C=8192, P=16384, Q=8) at N=2e6 on the batched form."""
    N = 2000000
    flow, src, counts = feedback_flow(N, sink="hist")
    run_flow(flow, src, N, C=8192, P=16384, Q=8, schedule=g.SCHEDULE_DEVICE, fuse=True, engine=engine, trace=False)
    rows = g.hist_rows(counts.data_ptr(), g.FEEDBACK_HIST_N, g.FEEDBACK_HIST_LO)
    w, _ = O.feedback_scalar(SEED, N, threads=8)
    v, c = np.unique(w, return_counts=True)
    assert len(rows) == len(v) and (rows[:, 0] == v).all() and (rows[:, 1] == c).all()


@pytest.mark.parametrize("distribution,destinations", [("parity", 2), ("skew_90_10", 2), ("all_one", 2),
                                                       ("uniform", 16)])
@pytest.mark.parametrize("fuse", [False, True])
@pytest.mark.parametrize("engine", [1, 3])
def test_routing_distributions(distribution, destinations, fuse, engine):
    """The four routing distributions of config 2 at C=8192, P=16384, Q=8 (reduced N)."""
    torch = torch_cuda()
    N = 60000
    flow, src, rec = feedback_flow(N, distribution=distribution, destinations=destinations)
    stats, trace = run_flow(flow, src, N, C=8192, P=16384, Q=8, schedule=g.SCHEDULE_DEVICE, fuse=fuse, engine=engine)
    ref = run_oracle(flow, N, C=8192, P=16384, Q=8, schedule=g.SCHEDULE_DEVICE, fuse=fuse, record=True)
    assert ref["status"] == O.OK, ref["message"]
    assert_same_trace(trace, ref["trace"])
    mult = [0.5 if d % 2 == 0 else 0.75 for d in range(destinations)]
    dist = {"parity": 0, "skew_90_10": 1, "all_one": 2, "uniform": 3}[distribution]
    w, _ = O.feedback_scalar(SEED, N, 1e-3, dist, destinations, mult)
    torch.cuda.synchronize()
    assert (rec.cpu().numpy().view(np.uint32) == w).all()


def test_device_engines_agree_on_large_run():
    """wide, micro and team engines produce the same trace on a config-2 shaped run (unfused)."""
    N = 50000
    traces = []
    for engine in (1, 2, 3):
        flow, src, _ = feedback_flow(N)
        _, trace = run_flow(flow, src, N, C=8192, P=16384, Q=8, schedule=g.SCHEDULE_DEVICE, engine=engine)
        traces.append(trace)
    assert_same_trace(traces[0], traces[1])
    assert_same_trace(traces[0], traces[2])


@pytest.mark.parametrize("distribution", ["parity", "skew_90_10", "all_one", "uniform"])
def test_config2_full_size_histogram_equals_golden(distribution):
    """Config 2 at full size: N=1e8, C=8192, P=16384, Q=8, seed 42, batched team engine with A/B
    fused; the sorted histogram must equal the oracle's (SHA-256 of the rows, row count, sum = N)
    recorded in tests/golden/feedback_N1e8.json."""
    import json
    import os

    fx = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "feedback_N1e8.json")))
    case = fx["cases"][distribution]
    N = fx["items"]
    flow, src, counts = feedback_flow(N, sink="hist", distribution=distribution, destinations=case["destinations"])
    stats, _ = run_flow(flow, src, N, C=8192, P=16384, Q=8, schedule=g.SCHEDULE_DEVICE, fuse=True, engine=3,
                        trace=False)
    rows = g.hist_rows(counts.data_ptr(), g.FEEDBACK_HIST_N, g.FEEDBACK_HIST_LO)
    assert stats["sunk"] == N
    assert len(rows) == case["rows"] and int(rows[:, 1].sum()) == N
    assert g.rows_sha256(rows) == case["sha256"]
