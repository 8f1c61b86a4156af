"""Serving: a mixed stream of independent flow jobs, continuously batched. Each topology runs one
device engine over copies of its graph (one per job slot); the jobs admitted to an engine run in
the same step launches.

Each served job's result must equal the same job run alone (exact digests for the particle
feedback and keyed hash jobs; photon tallies equal up to the order of the fp64 atomic additions,
1e-12 relative), and equal the CPU oracle: the feedback digest from the oracle's per-id weights,
the hash digest from the oracle's per-bucket multisets, the photon tally within the per-bin
tolerance of the photon tests (5e-3 relative plus the weight of two histories moved to a
neighbouring bin by fp32 rounding).
"""
import numpy as np
import pytest

import graphflow_amd as g
import oracle as O
from gpu_util import torch_cuda

pytestmark = pytest.mark.gpu

JOBS = [(g.JOB_FEEDBACK, 20000, 1), (g.JOB_HASH, 30000, 2), (g.JOB_PHOTON, 20000, 3),
        (g.JOB_FEEDBACK, 3000, 4), (g.JOB_HASH, 4096, 5), (g.JOB_PHOTON, 5000, 6),
        (g.JOB_FEEDBACK, 120000, 7), (g.JOB_HASH, 100000, 8), (g.JOB_PHOTON, 50000, 9),
        (g.JOB_FEEDBACK, 1, 10), (g.JOB_HASH, 8193, 11), (g.JOB_PHOTON, 1000, 12)]


def serve(jobs, slots):
    torch_cuda()
    sv = g.Server(slots=slots, max_items=200000)
    ids = [sv.submit(t, n, s) for t, n, s in jobs]
    res = {r["job"]: r for r in sv.poll(wait_all=True)}
    out = []
    for jid in ids:
        r = res[jid]
        assert r["status"] == 0, (r, sv.last_error())
        r = dict(r)
        if r["topology"] == g.JOB_PHOTON:
            r["tally"] = sv.tally(jid)
        out.append(r)
    sv.close()
    return out


def hash_graph():
    return [dict(role=O.ROLE_SOURCE, op="SRC_KEYED_ZIPF", bins=1, u=[1 << 20], f=[1.1], routes=[1]),
            dict(role=O.ROLE_TRANSFORM, op="CLS_KEY_BUCKET", bins=4096, routes=[2] * 4096),
            dict(role=O.ROLE_SINK, op="SNK_CHUNK_DIGEST", bins=0, routes=[])]


def test_mixed_stream_equals_alone_and_oracle():
    mixed = serve(JOBS, slots=8)
    alone = serve(JOBS, slots=1)
    assert len({r["slot"] for r in mixed}) > 1
    # continuous batching: jobs of one topology ran in the same engine at the same time
    for topo in (g.JOB_FEEDBACK, g.JOB_HASH, g.JOB_PHOTON):
        rs = [r for r in mixed if r["topology"] == topo]
        assert any(a["start_ms"] < b["end_ms"] and b["start_ms"] < a["end_ms"] for a in rs for b in rs if a is not b)
    # one slot per topology: the jobs of a topology ran one after another
    for topo in (g.JOB_FEEDBACK, g.JOB_HASH, g.JOB_PHOTON):
        rs = sorted((r for r in alone if r["topology"] == topo), key=lambda r: r["start_ms"])
        assert all(a["end_ms"] <= b["start_ms"] for a, b in zip(rs, rs[1:]))
    for (topo, n, seed), m, a in zip(JOBS, mixed, alone):
        assert m["items"] == n and a["items"] == n
        if topo == g.JOB_PHOTON:
            assert np.allclose(m["tally"], a["tally"], rtol=1e-12, atol=0.0)
            ref, _ = O.photon_scalar(seed, n, dict(g.PHOTON_DEFAULTS), threads=16)
            assert (np.abs(m["tally"] - ref) <= 5e-3 * ref + 2.0).all()
        else:
            assert m["digest"] == a["digest"], (topo, n, seed)
            if topo == g.JOB_FEEDBACK:
                w, _ = O.feedback_scalar(seed, n, threads=16)
                assert m["digest"] == g.record_digest(w), (n, seed)
            else:
                ref = O.run(hash_graph(), seed=seed, items=n, capacity=8192, live=16384, batch=8, mode=O.MODE_STEP,
                            multiset_bins=4096)
                assert ref["status"] == O.OK
                assert m["digest"] == g.multiset_digest(ref["multiset"], ref["multiset_items"]), (n, seed)


def test_latency_accounting_and_slots():
    jobs = [(g.JOB_FEEDBACK, 10000, s) for s in range(16)]
    res = serve(jobs, slots=4)
    for r in res:
        assert r["submit_ms"] <= r["start_ms"] <= r["end_ms"]
        assert r["engine"] == 3  # the team engine over the topology's slot copies
        assert r["steps"] > 0
    assert {r["slot"] for r in res} <= {0, 1, 2, 3}
    assert len({r["digest"] for r in res}) == 16  # different seeds, different results


def test_small_budget_mixed_topologies_drain():
    """A small memory budget: each engine's pool holds fewer chunks than two jobs' bounds (live bound
    plus the graph's output bounds), so the jobs of a topology are admitted one after another, and a
    job whose bound exceeds the whole pool runs with its live bound lowered to fit. Every job
    finishes with the result it has alone."""
    torch_cuda()
    jobs = [(g.JOB_FEEDBACK, 5000, 1), (g.JOB_PHOTON, 3000, 2), (g.JOB_HASH, 9000, 3),
            (g.JOB_FEEDBACK, 7000, 4), (g.JOB_HASH, 2000, 5), (g.JOB_PHOTON, 4000, 6)]
    sv = g.Server(slots=2, max_items=200000, memory_budget=int(2.6 * 2**30))
    ids = [sv.submit(t, n, s) for t, n, s in jobs]
    res = {r["job"]: r for r in sv.poll(wait_all=True)}
    sv.close()
    assert len(res) == len(jobs)
    for topo in (g.JOB_FEEDBACK, g.JOB_HASH, g.JOB_PHOTON):
        rs = sorted((r for r in res.values() if r["topology"] == topo), key=lambda r: r["start_ms"])
        assert all(a["end_ms"] <= b["start_ms"] for a, b in zip(rs, rs[1:]))
    alone = serve(jobs, slots=1)
    for jid, a in zip(ids, alone):
        r = res[jid]
        assert r["status"] == 0
        if r["topology"] != g.JOB_PHOTON:
            assert r["digest"] == a["digest"]
        else:
            assert abs(r["tally_sum"] - a["tally_sum"]) <= 1e-9 * abs(a["tally_sum"])


def test_reused_slot_jobs_equal_oracle():
    """Jobs of one topology one after another through the same slot (one slot per engine): each job
    starts on an engine that went idle after the previous one, and its result equals the oracle's
    (no tile of the previous job's last steps is run again when the engine wakes)."""
    jobs = [(g.JOB_HASH, 4096, 5), (g.JOB_HASH, 4096, 2), (g.JOB_HASH, 100, 5), (g.JOB_HASH, 30000, 7),
            (g.JOB_FEEDBACK, 5000, 3), (g.JOB_FEEDBACK, 700, 9), (g.JOB_FEEDBACK, 12000, 4)]
    res = serve(jobs, slots=1)
    for (topo, n, seed), r in zip(jobs, res):
        assert r["slot"] == 0 and r["items"] == n
        if topo == g.JOB_FEEDBACK:
            w, _ = O.feedback_scalar(seed, n, threads=16)
            assert r["digest"] == g.record_digest(w), (n, seed)
        else:
            ref = O.run(hash_graph(), seed=seed, items=n, capacity=8192, live=16384, batch=8, mode=O.MODE_STEP,
                        multiset_bins=4096)
            assert r["digest"] == g.multiset_digest(ref["multiset"], ref["multiset_items"]), (n, seed)
