"""The hash bin path on the GPU (BinPath::hash, core/BinPath.hpp; kernels flow/Execution.hpp:1207-1502,
table graph/BinTable.hpp:69-137): per-tile open-addressing tables in LDS, partitions up to 65536
bins, the table-full batch error, and results identical to the dense path and the oracle."""
import hashlib

import numpy as np
import pytest

import graphflow_amd as g
import oracle as O
from gpu_util import SEED, assert_same_trace, run_oracle, torch_cuda

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


@pytest.mark.parametrize("schedule,engine", [(g.SCHEDULE_DEVICE, 1), (g.SCHEDULE_HOST, 0)])
def test_sparse_65536_bins_published_sha(schedule, engine):
    """tools/BENCH_TABLE.md:216: N=262144 into 65536 bins (balanced id % B) has the rows SHA-256
    968f4a00...; C=4096, P=16384, Q=8 as the reference's kernel-axis run."""
    N, B = 262144, 65536
    flow, src, counts = _kernel_axis_flow(N, B, 0)
    flow.configure(g.Tuning(seed=SEED, batch_chunks=8, chunk_capacity=4096, live_particles=16384,
                            schedule=schedule, device_engine=engine))
    st = flow.run(src, N)
    assert st["sunk"] == N
    rows = g.hist_rows(counts.data_ptr(), B, 0)
    assert hashlib.sha256(g.rows_text(rows)).hexdigest() == \
        "968f4a008df5ad231cfb5578843033a9d4f2a28233f8568f0c7c6ed51fd638c0"


@pytest.mark.parametrize("schedule,engine,tile", [(g.SCHEDULE_HOST, 0, 1024), (g.SCHEDULE_DEVICE, 1, 1024),
                                                  (g.SCHEDULE_DEVICE, 2, 256)])
def test_hash_path_trace_equals_oracle(schedule, engine, tile):
    """The 4096-bucket keyed pipeline forced onto the hash path: every batch record equals the
    oracle's (the bin path never changes counts, slots or order)."""
    torch = torch_cuda()
    N, B, C, Q = 60000, 4096, 8192, 4
    mset = torch.zeros(2 * B, dtype=torch.int64, device="cuda")
    flow = g.Flow(g.PAYLOAD_KEYED)
    src = flow.source("keys", g.ZipfKeys(1 << 20, 1.1))
    part = flow.partition("bucket", g.KeyBucket(B))
    sink = flow.sink("digest", g.ChunkDigest(0, 0, mset.data_ptr(), B))
    flow.connect(src, part)
    for b in range(B):
        flow.connect(part, sink, b)
    flow.configure(g.Tuning(seed=SEED, batch_chunks=Q, chunk_capacity=C, live_particles=32768, schedule=schedule,
                            trace=True, device_engine=engine, bin_path=g.BIN_PATH_HASH, tile_items=tile))
    st = flow.run(src, N)
    ref = run_oracle(flow, N, C=C, P=32768, Q=Q, schedule=schedule, multiset_bins=B, bin_path=2, hash_tile=tile)
    assert ref["status"] == O.OK, ref["message"]
    assert_same_trace(flow.trace(), ref["trace"])
    torch.cuda.synchronize()
    ms = mset.cpu().numpy().view(np.uint64).reshape(-1, 2)
    assert (ms[:, 0] == ref["multiset"]).all() and int(ms[:, 1].sum()) == N == st["sunk"]


@pytest.mark.parametrize("schedule,engine", [(g.SCHEDULE_HOST, 0), (g.SCHEDULE_DEVICE, 1), (g.SCHEDULE_DEVICE, 2)])
def test_hash_table_exhaustion_aborts_the_batch(schedule, engine):
    """tests/integration/test_hash_bins.cpp:504-546: a table of 8 slots, 16 distinct bins in the only
    source chunk: the batch fails (BatchError), nothing reaches the sink, every chunk returns."""
    N, B = 16, 16
    flow, src, counts = _kernel_axis_flow(N, B, 0)
    kw = dict(seed=1, batch_chunks=1, chunk_capacity=16, live_particles=16, num_chunks=64, schedule=schedule,
              device_engine=engine, bin_path=g.BIN_PATH_HASH)
    flow.configure(g.Tuning(hash_table_entries=8, **kw))
    with pytest.raises(g.BatchError, match="table full"):
        flow.run(src, N)
    assert int(counts.sum().item()) == 0
    ref = O.run(flow.neutral_graph(), seed=1, items=N, capacity=16, live=16, num_chunks=64, batch=1,
                mode=O.MODE_RUNNER if schedule == g.SCHEDULE_HOST else O.MODE_STEP, hist_size=B,
                bin_path=2, hash_tile=256, hash_entries=8)
    assert ref["status"] == O.INVALID and "table full" in ref["message"]
    # sixteen slots hold the tile's sixteen bins
    flow.configure(g.Tuning(hash_table_entries=16, **kw))
    st = flow.run(src, N)
    assert st["sunk"] == N and counts.cpu().tolist() == [1] * B


def test_stage_abi_after_table_full_is_single_use():
    """The stage protocol alone: a full table fails the launch's batch at advance, publishes nothing,
    and the stage refuses further batches (Execution.hpp:2652-2656)."""
    torch_cuda()
    pool = g.Pool(g.PAYLOAD_PARTICLE, 64, 16)
    h = pool.acquire()
    pool.write(0, h.index * 16, np.arange(16, dtype=np.uint32).view(np.float32))
    pool.write(1, h.index * 16, np.arange(16, dtype=np.uint32))
    stage = g.Stage(pool, g.IdMod(16, 0).desc(), 1, SEED, 1, tile_items=256, bin_path=g.BIN_PATH_HASH,
                    hash_table_entries=8)
    stage.launch([(h.index, h.generation, 16, 0)])
    with pytest.raises(g.BatchError, match="table full"):
        stage.advance()
    with pytest.raises(g.StageFailedError):
        stage.launch([(h.index, h.generation, 16, 0)])
    assert pool.counts()["free"] == 63  # only the input is live


def test_dense_path_above_its_limit_is_rejected():
    flow, src, _ = _kernel_axis_flow(1000, 8192, 0)
    flow.configure(g.Tuning(seed=SEED, chunk_capacity=1024, schedule=g.SCHEDULE_DEVICE, bin_path=g.BIN_PATH_DENSE))
    with pytest.raises(ValueError, match="dense bin path"):
        flow.run(src, 1000)
