"""The autotuner picks a device-schedule configuration by probe runs; whatever it picks, results
stay exact: per-particle weights equal the oracle's scalar model, bucket multisets equal the
oracle's."""
import numpy as np
import pytest

import graphflow_amd as g
import oracle as O
from gpu_util import SEED, torch_cuda

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("Q", [8, None])
def test_feedback_autotuned_is_exact(Q):
    torch = torch_cuda()
    N = 200_000
    rec = torch.full((N,), -1, dtype=torch.int32, device="cuda")
    flow = g.Flow(g.PAYLOAD_PARTICLE)
    src, *_ = g.feedback_network(flow, g.RecordWeight(rec.data_ptr(), N))
    flow.configure(g.Tuning(seed=SEED, batch_chunks=Q, chunk_capacity=8192, live_particles=16384,
                            schedule=g.SCHEDULE_DEVICE, fuse_transforms=True))
    pick = flow.autotune(src, probe_items=100_000)
    assert pick["device_engine"] in (1, 2, 3)
    if Q is None:
        assert pick["batch_chunks"] in (4, 8, 16)
    else:
        assert pick["batch_chunks"] == Q
    rec.fill_(-1)
    st = flow.run(src, N)
    assert st["sunk"] == N and int(st["engine"]) == pick["device_engine"]
    w, _ = O.feedback_scalar(SEED, N, threads=16)
    assert (rec.cpu().numpy().view(np.uint32) == w).all()


def test_hash_pipeline_autotuned_is_exact():
    torch = torch_cuda()
    N, bins = 300_000, 4096
    ms = torch.zeros(2 * bins, dtype=torch.int64, device="cuda")
    flow = g.Flow(g.PAYLOAD_KEYED)
    src = flow.source("keys", g.ZipfKeys(1 << 20, 1.1))
    part = flow.partition("bucket", g.KeyBucket(bins))
    snk = flow.sink("digest", g.ChunkDigest(0, 0, ms.data_ptr(), bins))
    flow.connect(src, part)
    for b in range(bins):
        flow.connect(part, snk, b)
    flow.configure(g.Tuning(seed=SEED, batch_chunks=8, chunk_capacity=8192, schedule=g.SCHEDULE_DEVICE))
    pick = flow.autotune(src, probe_items=100_000)
    assert pick["device_engine"] in (1, 2)  # the team engine does not take 4096-bin partitions
    ms.zero_()
    flow.run(src, N)
    graph = flow.neutral_graph()
    ref = O.run(graph, seed=SEED, items=N, capacity=8192, live=0, batch=8, mode=O.MODE_STEP, multiset_bins=bins)
    got = ms.cpu().numpy().view(np.uint64).reshape(-1, 2)
    assert (got[:, 0] == ref["multiset"]).all() and (got[:, 1] == ref["multiset_items"]).all()
