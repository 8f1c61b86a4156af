"""Helpers shared by the GPU parity tests: build a graph on the device path and the same graph on
the oracle, run both, return comparable outputs."""
from __future__ import annotations

import numpy as np

import graphflow_amd as g
import oracle as O

SEED = 42


def log(*a):
    print(*a, flush=True)


def torch_cuda():
    import torch

    return torch


def oracle_graph(flow: g.Flow):
    return flow.neutral_graph()


def feedback_flow(N, *, distribution="parity", destinations=2, sink="record", multipliers=None):
    torch = torch_cuda()
    flow = g.Flow(g.PAYLOAD_PARTICLE)
    if sink == "record":
        buf = torch.full((N,), -1, dtype=torch.int32, device="cuda")
        op = g.RecordWeight(buf.data_ptr(), N)
    else:
        buf = torch.zeros(g.FEEDBACK_HIST_N, dtype=torch.int64, device="cuda")
        op = g.WeightHistogram(buf.data_ptr(), g.FEEDBACK_HIST_LO, g.FEEDBACK_HIST_N)
    src, filt, dests, snk = g.feedback_network(flow, op, 1e-3, distribution, destinations, multipliers)
    return flow, src, buf


def run_flow(flow, src, N, *, C, P, Q, schedule, fuse=False, engine=0, trace=True, tile_items=None, **kw):
    flow.configure(g.Tuning(seed=SEED, batch_chunks=Q, chunk_capacity=C, live_particles=P, schedule=schedule,
                            trace=trace, fuse_transforms=fuse, device_engine=engine, tile_items=tile_items, **kw))
    stats = flow.run(src, N)
    return stats, (flow.trace() if trace else None)


def run_oracle(flow, N, *, C, P, Q, schedule, fuse=False, **kw):
    mode = O.MODE_RUNNER if schedule == g.SCHEDULE_HOST else O.MODE_STEP
    return O.run(oracle_graph(flow), seed=SEED, items=N, capacity=C, live=P, batch=Q, mode=mode,
                 fuse=fuse, trace_cap=1 << 25, **kw)


def first_diff(a, b):
    n = min(len(a), len(b))
    d = np.nonzero(a[:n] != b[:n])[0]
    i = int(d[0]) if len(d) else n
    return i, a[max(0, i - 6):i + 6].tolist(), b[max(0, i - 6):i + 6].tolist()


def assert_same_trace(gpu, ref):
    assert len(gpu) == len(ref) and (gpu == ref).all(), first_diff(gpu, ref)


def decode_trace(tr):
    """Yield (node, step, inputs[(index,gen,fill)], counts, outputs[(index,gen,fill,bin)]) records."""
    i = 0
    while i < len(tr):
        node, step, n_in = int(tr[i]) & 0xFFFF, int(tr[i + 1]), int(tr[i + 2])
        ins = tr[i + 3:i + 3 + 3 * n_in].reshape(-1, 3)
        nb = int(tr[i + 3 + 3 * n_in])
        counts = tr[i + 4 + 3 * n_in:i + 4 + 3 * n_in + nb]
        R = int(tr[i + 4 + 3 * n_in + nb])
        o0 = i + 5 + 3 * n_in + nb
        outs = tr[o0:o0 + 4 * R].reshape(-1, 4)
        yield node, step, ins, counts, outs
        i = o0 + 4 * R
