"""Lanes (Tuning::inflightBatches, flow/Tuning.hpp:34-41; sched/GraphRunner.hpp:766-771) and the
hash path's table-full error (flow/Execution.hpp:1207-1307, graph/BinTable.hpp:101-121) in the
oracle's two schedules."""
import numpy as np
import pytest

import oracle as O

FEEDBACK = [dict(role=0, op="SRC_PARTICLE_UNIFORM", bins=1, routes=[1]),
            dict(role=1, op="CLS_WEIGHT_ROUTE", bins=3, routes=[4, 2, 3], u=[0], f=[1e-3]),
            dict(role=1, op="XF_SCALE_WEIGHT", bins=1, routes=[1], f=[0.5]),
            dict(role=1, op="XF_SCALE_WEIGHT", bins=1, routes=[1], f=[0.75]),
            dict(role=2, op="SNK_RECORD_WEIGHT", bins=0)]


def decode(tr):
    i = 0
    while i < len(tr):
        node, step, n_in = int(tr[i]) & 0xFFFF, int(tr[i + 1]), int(tr[i + 2])
        nb = int(tr[i + 3 + 3 * n_in])
        R = int(tr[i + 4 + 3 * n_in + nb])
        yield node, step, n_in, tr[i + 4 + 3 * n_in:i + 4 + 3 * n_in + nb], R
        i += 5 + 3 * n_in + nb + 4 * R


@pytest.mark.parametrize("mode", [O.MODE_RUNNER, O.MODE_STEP])
@pytest.mark.parametrize("depth", [1, 2, 4])
def test_lanes_keep_every_particle_trajectory(mode, depth):
    n = 20000
    r = O.run(FEEDBACK, seed=42, items=n, capacity=512, live=4096, batch=4, mode=mode, record=True,
              inflight=depth, fuse=mode == O.MODE_STEP, trace_cap=1 << 23)
    assert r["status"] == O.OK, r["message"]
    w, _ = O.feedback_scalar(42, n)
    assert (r["record"] == w).all() and r["sunk"] == n


@pytest.mark.parametrize("depth", [2, 4])
def test_step_lanes_batch_more_per_step(depth):
    n = 40000
    ref = O.run(FEEDBACK, seed=42, items=n, capacity=512, live=8192, batch=4, mode=O.MODE_STEP,
                trace_cap=1 << 23)
    r = O.run(FEEDBACK, seed=42, items=n, capacity=512, live=8192, batch=4, mode=O.MODE_STEP,
              trace_cap=1 << 23, inflight=depth)
    assert r["status"] == O.OK and ref["status"] == O.OK
    per = {}
    for node, step, *_ in decode(r["trace"]):
        per[(node, step)] = per.get((node, step), 0) + 1
    # at most `depth` batches of a node per step, one sink batch, and some step uses the extra lanes
    assert max(v for (k, _), v in per.items() if k != 4) <= depth
    assert max(v for (k, _), v in per.items() if k == 4) == 1
    assert max(v for (k, _), v in per.items() if k == 0) == depth
    assert r["steps"] < ref["steps"]


def test_runner_lanes_overlap_partition_batches():
    # a partition batch spans two runner iterations (classify, then plan + scatter); a second lane
    # lets the node dispatch its next batch in between, which changes the interleaving
    n = 20000
    a = O.run(FEEDBACK, seed=42, items=n, capacity=512, live=4096, batch=2, mode=O.MODE_RUNNER, trace_cap=1 << 23)
    b = O.run(FEEDBACK, seed=42, items=n, capacity=512, live=4096, batch=2, mode=O.MODE_RUNNER, trace_cap=1 << 23,
              inflight=2)
    assert a["status"] == b["status"] == O.OK
    assert b["steps"] < a["steps"]
    assert len(a["trace"]) != len(b["trace"]) or (a["trace"] != b["trace"]).any()


HASH = [dict(role=0, op="SRC_PARTICLE_UNIFORM", bins=1, routes=[1]),
        dict(role=1, op="CLS_ID_MOD", upd="GEN_TO_BIN", bins=16, routes=[2] * 16, u=[0]),
        dict(role=2, op="SNK_HIST", bins=0, u=[2, 0, 16])]


@pytest.mark.parametrize("mode", [O.MODE_RUNNER, O.MODE_STEP])
def test_hash_table_exhaustion_aborts_the_batch(mode):
    # tests/integration/test_hash_bins.cpp:504-546: 8 table slots, 16 distinct bins in a 128-item tile
    kw = dict(seed=1, items=16, capacity=16, live=16, batch=1, mode=mode, hist_size=16, bin_path=2, hash_tile=128)
    r = O.run(HASH, hash_entries=8, **kw)
    assert r["status"] == O.INVALID and "table full" in r["message"]
    assert r["hist"].sum() == 0  # nothing reached the sink
    # 16 slots hold 16 distinct bins; the default table (2 * tile slots) never fills
    for entries in (16, 0):
        r = O.run(HASH, hash_entries=entries, **kw)
        assert r["status"] == O.OK and r["hist"].tolist() == [1] * 16
    # a tile of 8 items holds at most 8 distinct bins
    r = O.run(HASH, hash_entries=8, **{**kw, "hash_tile": 8})
    assert r["status"] == O.OK
    # the dense path has no table
    r = O.run(HASH, hash_entries=8, **{**kw, "bin_path": 1})
    assert r["status"] == O.OK


@pytest.mark.parametrize("mode,depth", [(O.MODE_RUNNER, 1), (O.MODE_STEP, 2)])
def test_keyed_scalar_model_equals_the_scheduled_pipeline(mode, depth):
    g = [dict(role=0, op="SRC_KEYED_ZIPF", bins=1, u=[1 << 20], f=[1.1], routes=[1]),
         dict(role=1, op="CLS_KEY_BUCKET", bins=4096, routes=[2] * 4096),
         dict(role=2, op="SNK_CHUNK_DIGEST", bins=0)]
    n = 50000
    r = O.run(g, seed=9, items=n, capacity=8192, live=32768, batch=4, mode=mode, multiset_bins=4096, inflight=depth)
    assert r["status"] == O.OK, r["message"]
    ms, ct = O.keyed_scalar(9, n, threads=4)
    assert (ms == r["multiset"]).all() and (ct == r["multiset_items"]).all() and int(ct.sum()) == n


def test_batch_loops_do_not_depend_on_threads():
    kw = dict(seed=42, items=60000, capacity=8192, live=16384, batch=8, mode=O.MODE_RUNNER, trace_cap=1 << 23,
              record=True)
    a = O.run(FEEDBACK, threads=1, **kw)
    b = O.run(FEEDBACK, threads=4, **kw)
    assert a["status"] == b["status"] == O.OK
    assert (a["trace"] == b["trace"]).all() and (a["record"] == b["record"]).all()


@pytest.mark.parametrize("chunks,depth", [(2, 1), (4, 1), (3, 2), (8, 4)])
def test_source_lanes_admit_several_chunks(chunks, depth):
    # step mode: a source lane admits up to `chunks` chunks per step, each by GraphRunner's admission
    # rule (GraphRunner.hpp:902-983) with the earlier ones counted; every trajectory is unchanged
    n, C, P = 30011, 512, 4096
    kw = dict(seed=42, items=n, capacity=C, live=P, batch=4, mode=O.MODE_STEP, record=True, fuse=True,
              trace_cap=1 << 23, inflight=depth)
    ref = O.run(FEEDBACK, **kw)
    r = O.run(FEEDBACK, source_chunks=chunks, **kw)
    assert r["status"] == ref["status"] == O.OK, r["message"]
    w, _ = O.feedback_scalar(42, n)
    assert (r["record"] == w).all() and r["sunk"] == n
    src = [(step, outs) for node, step, n_in, counts, outs in decode_outputs(r["trace"]) if node == 0]
    per_step = {}
    for step, outs in src:
        per_step[step] = per_step.get(step, 0) + len(outs)
        # a lane's chunks: C items each but the last, ids consecutive over the lane's chunks
        assert 1 <= len(outs) <= chunks
        assert all(f == C for f in outs[:-1, 2]) and 0 < outs[-1, 2] <= C
    assert max(per_step.values()) == min(chunks * depth, P // C)
    assert sum(int(o[:, 2].sum()) for _, o in src) == n


def test_source_chunks_do_not_change_the_runner_schedule():
    kw = dict(seed=42, items=20000, capacity=512, live=4096, batch=4, mode=O.MODE_RUNNER, trace_cap=1 << 23)
    a = O.run(FEEDBACK, **kw)
    b = O.run(FEEDBACK, source_chunks=8, **kw)
    assert a["status"] == b["status"] == O.OK and (a["trace"] == b["trace"]).all()


def decode_outputs(tr):
    i = 0
    while i < len(tr):
        node, step, n_in = int(tr[i]) & 0xFFFF, int(tr[i + 1]), int(tr[i + 2])
        nb = int(tr[i + 3 + 3 * n_in])
        R = int(tr[i + 4 + 3 * n_in + nb])
        o0 = i + 5 + 3 * n_in + nb
        yield node, step, n_in, tr[i + 4 + 3 * n_in:o0 - 1], tr[o0:o0 + 4 * R].reshape(-1, 4)
        i = o0 + 4 * R


@pytest.mark.parametrize("n,C,P,chunks", [(5000, 512, 300, 4), (4097, 1024, 8192, 8), (1, 512, 4096, 4)])
def test_source_chunks_edges(n, C, P, chunks):
    # a live bound below C (one partial chunk per admission), N not a multiple of C (a partial tail
    # chunk), a one-item run: trajectories unchanged, every lane's chunks full but the last
    kw = dict(seed=42, items=n, capacity=C, live=P, batch=4, mode=O.MODE_STEP, record=True, fuse=True,
              trace_cap=1 << 23, inflight=2)
    r = O.run(FEEDBACK, source_chunks=chunks, **kw)
    assert r["status"] == O.OK, r["message"]
    w, _ = O.feedback_scalar(42, n)
    assert (r["record"] == w).all() and r["sunk"] == n
    admitted = 0
    for node, step, n_in, counts, outs in decode_outputs(r["trace"]):
        if node != 0:
            continue
        assert all(f == C for f in outs[:-1, 2])
        admitted += int(outs[:, 2].sum())
    assert admitted == n
