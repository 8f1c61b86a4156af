"""Benchmark of the transform-batch hot path on one MI355X.

Headline workload: the fp32 particle feedback network Generate -> Filter -> {A (x0.5), B (x0.75)}
-> Filter, Filter.sink -> Sink at N=1e8, C=8192, P=16384, Q=8, seed 42 (synthetic inputs generated
on the device by the source node). One "step" of this benchmark is one complete run of the network
over all N particles; `value` is N / seconds per run, measured around `Flow.run` with the pool, the
sink histogram and the engine buffers already resident in HBM.

Beside the headline it reports (each at the size its workload states):
  - the roofline of the dominant kernel, with its launch durations taken from the timed runs
    (HIP events on the engine's stream);
  - the other routing distributions of the feedback network at N=1e8;
  - the same network with 16M particles in flight on the wide engine (kernels at throughput);
  - the keyed hash pipeline at 1e8 keys, Q in {1, 4, 8, 16}, with in-flight batch lanes;
  - photon transport at 1e8 photons (team engine with P photons in flight; wide engine with 16M);
  - the served mixed stream of 4096 jobs (N log-uniform 1e3..1e7, 64 GB budget);
  - the plain one-batch-at-a-time forms (Q=1 at 2e6 particles; the headline's N=1e8, Q=8);
and, in the CPU leg, the oracle's restatement of the reference CPU path (runner schedule, one thread
and all host threads) plus the CPU results every GPU form is checked against.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "particles/s at N=1e8 C=8192 P=16384 Q=8; % HBM peak; served p99 job latency"
GOLDEN = os.path.join(ROOT, "tests", "golden", "feedback_N1e8.json")
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_team_feedback_N1e8_hbm_traffic.txt")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


ENGINES = {0: "plain host-planned", 1: "wide (hipGraph of step kernels)", 2: "micro (one persistent workgroup)",
           3: "team (one persistent launch, workgroups of one XCD, replicated scheduler)",
           4: "team handing over to wide"}


def build_feedback(g, torch, C, P, Q, schedule, seed=42, use_graph=1, fuse=False, engine=0, team=0,
                   distribution="parity", destinations=2, **kw):
    counts = torch.zeros(g.FEEDBACK_HIST_N, dtype=torch.int64, device="cuda")
    flow = g.Flow(g.PAYLOAD_PARTICLE)
    src, *_ = g.feedback_network(flow, g.WeightHistogram(counts.data_ptr(), g.FEEDBACK_HIST_LO, g.FEEDBACK_HIST_N),
                                 1e-3, distribution, destinations)
    flow.configure(g.Tuning(seed=seed, batch_chunks=Q, chunk_capacity=C, live_particles=P, schedule=schedule,
                            use_graph=use_graph, fuse_transforms=fuse, device_engine=engine,
                            max_workgroups=team or None, **kw))
    return flow, src, counts


def run_once(flow, src, buf, N, torch):
    buf.zero_()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st = flow.run(src, N)
    torch.cuda.synchronize()
    return time.perf_counter() - t0, st


def golden_cases():
    try:
        return json.load(open(GOLDEN))
    except OSError:
        return None


def pmc_summary(fetch_csv, write_csv, workload, out, kernel="k_team", command=None):
    """Turn two rocprofv3 --pmc runs (FETCH_SIZE, then WRITE_SIZE; counter_collection.csv, KiB per
    dispatch) of a bench.py command with `--copy-calib 1024` into HBM bytes per launch of the
    kernels whose name contains `kernel`. Each counter is scaled by what it reads for the
    calibration copy of known size (the last device copy of the run: 1 GiB read and 1 GiB written)."""
    import csv

    def per_dispatch(path, counter):
        d = {}
        for r in csv.DictReader(open(path)):
            if r["Counter_Name"] != counter:
                continue
            k = (int(r["Dispatch_Id"]), r["Kernel_Name"])
            d[k] = d.get(k, 0.0) + float(r["Counter_Value"])
        return d

    cmd = command or "--only-headline --no-cpu --no-autotune --steps 1 --warmup 0"
    lines = [f"# HBM traffic of {kernel}, from two rocprofv3 --pmc passes:",
             "#   cd /tmp && rocprofv3 --pmc FETCH_SIZE --output-format csv -d $OUT/pmc_fetch -o run -- "
             f"python $R/bench.py {cmd} --copy-calib 1024",
             "#   (the same with --pmc WRITE_SIZE into $OUT/pmc_write; R = the repository root, OUT = an output directory)",
             "#   python bench.py --pmc-summary $OUT/pmc_fetch/run_counter_collection.csv "
             f"$OUT/pmc_write/run_counter_collection.csv <this file> --pmc-kernel '{kernel}'"
             + (f" --pmc-command '{command}'" if command else ""),
             f"workload={workload}"]
    per = {}
    for counter, path in (("FETCH_SIZE", fetch_csv), ("WRITE_SIZE", write_csv)):
        d = per_dispatch(path, counter)
        copies = [(k, v) for k, v in d.items() if "copyBuffer" in k[1]]
        cal_k, cal_v = max(copies, key=lambda kv: kv[0][0])
        factor = float(1 << 20) / cal_v  # KiB the 1 GiB copy moved / KiB the counter read
        mine = [v for k, v in sorted(d.items()) if kernel in k[1]]
        per[counter] = [v * factor * 1024.0 for v in mine]
        shown = ", ".join(f"{v:.1f}" for v in mine) if len(mine) <= 16 else f"{len(mine)} dispatches, {sum(mine):.1f} in total"
        lines.append(f"{counter}: calibration copy dispatch {cal_k[0]} read {cal_v:.1f} KiB for 1048576 KiB "
                     f"-> factor {factor:.4f}; {kernel} dispatches (KiB): {shown}")
    launches = len(per["FETCH_SIZE"])
    total = sum(per["FETCH_SIZE"]) + sum(per["WRITE_SIZE"])
    lines.append(f"launches={launches}")
    lines.append(f"fetch_bytes_total={sum(per['FETCH_SIZE']):.0f}")
    lines.append(f"write_bytes_total={sum(per['WRITE_SIZE']):.0f}")
    lines.append(f"traffic_bytes_per_launch={total / max(launches, 1):.0f}")
    with open(out, "w") as f:
        f.write("\n".join(lines) + "\n")
    print("\n".join(lines))


def traffic_for(workload):
    """HBM bytes per k_team launch from the committed PMC summary of this workload, or None."""
    import re

    try:
        kv = dict(m.groups() for m in (re.match(r"^([a-z_]+)=(.*)$", l.strip()) for l in open(TRAFFIC_FILE)) if m)
    except OSError:
        return None
    if kv.get("workload") != workload:
        return None
    return float(kv["traffic_bytes_per_launch"]), int(kv["launches"])


# ---- the other routing distributions of config 2 -------------------------------------------------
def distribution_forms(g, torch, N, C, P, Q, golden):
    out = {}
    for name, dest in (("skew_90_10", 2), ("all_one", 2), ("uniform", 16)):
        flow, src, counts = build_feedback(g, torch, C, P, Q, g.SCHEDULE_DEVICE, fuse=True, distribution=name,
                                           destinations=dest)
        run_once(flow, src, counts, min(N, 1_000_000), torch)  # engine build
        dt, st = run_once(flow, src, counts, N, torch)
        rows = g.hist_rows(counts.data_ptr(), g.FEEDBACK_HIST_N, g.FEEDBACK_HIST_LO)
        sha = g.rows_sha256(rows)
        rec = {"value": N / dt, "unit": "particles/s", "items": N, "engine": int(st["engine"]),
               "histogram_rows": int(len(rows)), "histogram_sha256": sha}
        if golden and N == golden["items"]:
            rec["histogram_equals_golden"] = sha == golden["cases"][name]["sha256"]
        out[f"feedback_{name}_{dest}way"] = rec
        log(f"[bench] feedback {name} ({dest} destinations) N={N}: {N / dt:.4g} particles/s, rows {len(rows)}, "
            f"equals golden: {rec.get('histogram_equals_golden')}")
        flow.close()
    return out


# ---- the feedback network with millions of particles in flight -----------------------------------
def wide_feedback_form(g, torch, N, C, live, golden, lanes=31, src_chunks=16, wg=2048, tile=1024):
    """The headline network and seed with a live bound of `live` particles instead of 16384 (wide
    engine, `lanes` batches in flight per node, source lanes of `src_chunks` chunks, Q=16): every
    particle's trajectory is the headline's, so the histogram must equal the headline's fixture.
    A step carries millions of partition passes, so the classify and partition-scatter kernels run
    at their throughput instead of the team engine's step latency; their HIP-event times give
    their bandwidth."""
    Q = 16
    S = live // C * 3 + 2 * lanes * (Q + 4) + 64
    flow, src, counts = build_feedback(g, torch, C, live, Q, g.SCHEDULE_DEVICE, fuse=True, engine=1, team=wg,
                                       tile_items=tile, inflight_batches=lanes, source_chunks=src_chunks,
                                       num_chunks=S)
    run_once(flow, src, counts, 2_000_000, torch)  # engine build, hipGraph capture
    dt, st = run_once(flow, src, counts, N, torch)
    rows = g.hist_rows(counts.data_ptr(), g.FEEDBACK_HIST_N, g.FEEDBACK_HIST_LO)
    sha = g.rows_sha256(rows)
    rec = {"value": N / dt, "unit": "particles/s", "items": N, "live": live, "Q": Q, "inflight_batches": lanes,
           "source_chunks": src_chunks, "workgroups": wg, "engine_steps": int(st["steps"]), "engine": int(st["engine"]),
           "histogram_sha256": sha}
    if golden and N == golden["items"]:
        rec["histogram_equals_golden"] = sha == golden["cases"]["parity"]["sha256"]
    # kernel times (direct launches timed with HIP events)
    flow.configure(g.Tuning(seed=42, batch_chunks=Q, chunk_capacity=C, live_particles=live, schedule=g.SCHEDULE_DEVICE,
                            use_graph=2, fuse_transforms=True, device_engine=1, max_workgroups=wg, tile_items=tile,
                            inflight_batches=lanes, source_chunks=src_chunks, num_chunks=S))
    run_once(flow, src, counts, 2_000_000, torch)
    run_once(flow, src, counts, N, torch)
    kms, kn = flow.kernel_times()
    part_items, _ = flow.item_passes()
    rec["kernel_ms"] = {"classify": kms[0], "plan": kms[1], "partition_scatter": kms[4], "source": kms[5],
                        "sink": kms[6], "retire_select": kms[7]}
    rec["partition_passes"] = int(part_items)
    rec["partition_scatter_gbs"] = 24.0 * part_items / (kms[4] * 1e-3) / 1e9
    rec["classify_scatter_gbs"] = 24.0 * part_items / ((kms[0] + kms[4]) * 1e-3) / 1e9
    rec["classify_scatter_hbm_frac"] = rec["classify_scatter_gbs"] / HBM_PEAK_GBS
    log(f"[bench] feedback network, {live} particles in flight (wide engine) N={N}: {N / dt:.4g} particles/s, "
        f"{int(st['steps'])} steps; partition scatter {rec['partition_scatter_gbs']:.0f} GB/s, classify+scatter "
        f"{rec['classify_scatter_gbs']:.0f} GB/s; histogram equals golden: {rec.get('histogram_equals_golden')}")
    flow.close()
    return rec


# ---- config 3: keyed hash pipeline ---------------------------------------------------------------
def hash_flow(g, ms, B=4096):
    f = g.Flow(g.PAYLOAD_KEYED)
    src = f.source("keys", g.ZipfKeys(1 << 20, 1.1))
    part = f.partition("bucket", g.KeyBucket(B))
    snk = f.sink("digest", g.ChunkDigest(0, 0, ms.data_ptr(), B))
    f.connect(src, part)
    for b in range(B):
        f.connect(part, snk, b)
    return f, src


def hash_forms(g, torch, N, qs, lanes, live, src_chunks=8, tile=1024):
    """Keyed hash pipeline (Zipf(1.1) keys over 2^20 values, murmur3 fmix32 into 4096 buckets,
    C=8192) at each Q: keys/s of the batched wide engine with `lanes` batches in flight per node,
    source lanes of `src_chunks` chunks per step and `live` keys in flight. Per-bucket (hash sum,
    count) pairs are kept for the CPU leg's check. A second run with HIP events around every step
    kernel gives the kernels' time."""
    out, multisets = {}, {}
    B, C = 4096, 8192
    ms = torch.zeros(2 * B, dtype=torch.int64, device="cuda")
    for Q in qs:
        f, src = hash_flow(g, ms, B)
        S = min(524000, live // C + 2 * lanes * (Q + B) + 64)
        tun = dict(seed=42, batch_chunks=Q, chunk_capacity=C, live_particles=live, num_chunks=S,
                   schedule=g.SCHEDULE_DEVICE, device_engine=1, max_workgroups=1024, tile_items=tile,
                   inflight_batches=lanes, source_chunks=src_chunks)
        f.configure(g.Tuning(**tun))
        run_once(f, src, ms, min(N, 2_000_000), torch)  # engine build, hipGraph capture
        dt, st = run_once(f, src, ms, N, torch)
        multisets[Q] = ms.cpu().numpy().view("uint64").reshape(-1, 2).copy()
        rec = {"value": N / dt, "unit": "keys/s", "items": N, "Q": Q, "inflight_batches": lanes, "live": live,
               "source_chunks": src_chunks, "tile_items": tile,
               "num_chunks": S, "engine_steps": int(st["steps"]), "engine": int(st["engine"])}
        # kernel time per step kernel (direct launches timed with HIP events, not the graph replay)
        f.configure(g.Tuning(**{**tun, "use_graph": 2}))
        run_once(f, src, ms, min(N, 2_000_000), torch)
        run_once(f, src, ms, N, torch)
        kms, kn = f.kernel_times()
        part_items, _ = f.item_passes()
        rec["kernel_ms"] = {"classify": kms[0], "plan": kms[1], "partition_scatter": kms[4], "source": kms[5],
                            "sink": kms[6], "retire_select": kms[7]}
        # the partition's classify + scatter kernels: 24 B per partition pass (12 B read, 12 B written)
        kab = kms[0] + kms[4]
        rec["classify_scatter_gbs"] = 24.0 * part_items / (kab * 1e-3) / 1e9 if kab > 0 else None
        rec["classify_scatter_hbm_frac"] = rec["classify_scatter_gbs"] / HBM_PEAK_GBS if kab > 0 else None
        rec["steps_timed"] = int(kn[0])
        out[f"hash_pipeline_B4096_Q{Q}"] = rec
        log(f"[bench] hash pipeline B=4096 Q={Q} lanes={lanes} N={N}: {N / dt:.4g} keys/s ({int(st['steps'])} steps); "
            f"kernels (ms) classify {kms[0]:.1f} plan {kms[1]:.1f} partition scatter {kms[4]:.1f} source {kms[5]:.1f} "
            f"sink {kms[6]:.1f} retire/select {kms[7]:.1f}; classify+scatter {rec['classify_scatter_gbs']:.0f} GB/s")
        f.close()
    return out, multisets


# ---- config 4: photon transport ------------------------------------------------------------------
def photon_forms(g, torch, N, Np, C, P, Q, wide_live=1 << 24, lanes=31, src_chunks=16, wg=2048, tile=1024):
    """Photon transport (10 layers of 1 mfp, albedo 0.9, HG g=0.9, roulette 1e-4 / 0.1, 1024 depth
    bins): the batched persistent form (team engine, P photons in flight), the batched hipGraph form
    with `wide_live` photons in flight (wide engine, `lanes` batches per node, Q=16; 0 = skip) and the
    plain host-planned one-batch form."""
    cfg = dict(g.PHOTON_DEFAULTS)
    out, tallies = {}, {}
    Sw = wide_live // C * 3 + 2 * lanes * (16 + 4) + 64
    forms = [("photon_batched_team", N, P, dict(schedule=g.SCHEDULE_DEVICE, batch_chunks=Q))]
    if wide_live:
        forms.append((f"photon_batched_wide_live_{wide_live}", N, wide_live,
                      dict(schedule=g.SCHEDULE_DEVICE, batch_chunks=16, fuse_transforms=True, device_engine=1,
                           max_workgroups=wg, tile_items=tile, inflight_batches=lanes, source_chunks=src_chunks,
                           num_chunks=Sw)))
    if Np > 0:
        forms.append(("photon_plain_host_planned_Q1", Np, P, dict(schedule=g.SCHEDULE_HOST, batch_chunks=1)))
    for name, n, live, kw in forms:
        tally = torch.zeros(cfg["tally_bins"], dtype=torch.float64, device="cuda")
        flow = g.Flow(g.PAYLOAD_PHOTON)
        src, *_ = g.photon_network(flow, tally.data_ptr())
        flow.configure(g.Tuning(seed=42, chunk_capacity=C, live_particles=live, photon=cfg, **kw))
        flow.run(src, min(n, 100000))  # warm-up: engine build, pool
        dt, st = run_once(flow, src, tally, n, torch)
        part_items, xf_items = flow.item_passes()
        photon_steps = part_items  # one alive-partition pass per photon step (the step is fused into it)
        out[name] = {"value": n / dt, "unit": "photons/s", "items": n, "live": live,
                     "photon_steps_per_s": photon_steps / dt, "photon_steps": int(photon_steps),
                     "engine_steps": int(st["steps"])}
        tallies[name] = tally.cpu().numpy()
        log(f"[bench] {name} N={n}: {n / dt:.4g} photons/s, {photon_steps / dt:.4g} photon steps/s")
        flow.close()
    return out, tallies


# ---- config 5: served mixed stream ---------------------------------------------------------------
def job_stream(jobs, max_items, seed=42):
    import numpy as np

    rng = np.random.default_rng(seed)
    topo = rng.integers(0, 3, jobs)
    items = np.exp(rng.uniform(np.log(1e3), np.log(max_items), jobs)).astype(np.int64)
    gaps = rng.exponential(1.0, jobs)
    seeds = 1000 + np.arange(jobs)
    return topo, items, gaps, seeds


def serve_stream(g, topo, items, arrive, seeds, slots, max_items):
    import threading

    import numpy as np

    sv = g.Server(slots=slots, max_items=int(max_items) + 1, prewarm=True)
    ids = [None] * len(topo)

    def submitter():
        t0 = time.perf_counter()
        for i in range(len(topo)):
            d = arrive[i] - (time.perf_counter() - t0)
            if d > 0:
                time.sleep(d)
            ids[i] = sv.submit(int(topo[i]), int(items[i]), int(seeds[i]))

    th = threading.Thread(target=submitter)
    th.start()
    # collect results as they finish, with a progress line every few seconds
    res, t_last = [], time.perf_counter()
    while True:
        res += sv.poll()
        if not th.is_alive() and len(res) == len(topo):
            break
        if time.perf_counter() - t_last > 5.0:
            t_last = time.perf_counter()
            log(f"[bench] serving: {sum(i is not None for i in ids)} submitted, {len(res)} finished")
        time.sleep(0.05)
    th.join()
    sv.close()
    lat = np.array([r["end_ms"] - r["submit_ms"] for r in res])
    span = (max(r["end_ms"] for r in res) - min(r["submit_ms"] for r in res)) * 1e-3
    return res, lat, span


def served_forms(g, torch, jobs, max_items, rate, slots=16):
    """Config 5: `jobs` independent flow jobs, Poisson arrivals at `rate` per second, N log-uniform over
    [1e3, max_items], topology uniform over {particle feedback, keyed hash pipeline, photon transport},
    64 GB budget, `slots` job slots per topology engine. Aggregate items/s and p50/p99 job latency
    (submit to finish), next to the same stream's jobs run one at a time as single flows on the
    batched device engine (every 16th job) and with the plain host-planned form (the stream's 16
    smallest jobs)."""
    import numpy as np

    topo, items, gaps, seeds = job_stream(jobs, max_items)
    arrive = np.cumsum(gaps / rate)
    res, lat, span = serve_stream(g, topo, items, arrive, seeds, slots, max_items)
    total = int(items.sum())
    bad = [r for r in res if r["status"] != 0]
    out = {"value": total / span, "unit": "items/s", "jobs": jobs, "items": total, "arrival_rate_per_s": rate,
           "n_range": [1000, int(max_items)], "slots": slots, "memory_budget_bytes": 64 << 30,
           "failed_jobs": len(bad), "p50_latency_ms": float(np.percentile(lat, 50)),
           "p99_latency_ms": float(np.percentile(lat, 99)), "makespan_s": span,
           "jobs_per_topology": {g.JOB_NAMES[t]: int((topo == t).sum()) for t in range(3)}}
    by_topo = {}
    for t in range(3):
        lt = np.array([r["end_ms"] - r["submit_ms"] for r in res if r["topology"] == t])
        if len(lt):
            by_topo[g.JOB_NAMES[t]] = {"p50_latency_ms": float(np.percentile(lt, 50)),
                                       "p99_latency_ms": float(np.percentile(lt, 99))}
    out["latency_per_topology"] = by_topo
    # one job at a time: every 16th job as a single flow on the batched device engine (engine picked
    # per topology as for a run alone), then the plain host-planned form, one batch at a time (Q=1),
    # on the stream's 16 smallest jobs
    sample = np.arange(0, jobs, 16)
    small = np.argsort(items)[:16]
    tally = torch.zeros(1024, dtype=torch.float64, device="cuda")
    rec = torch.zeros(int(max_items) + 1, dtype=torch.int32, device="cuda")
    ms = torch.zeros(2 * 4096, dtype=torch.int64, device="cuda")
    flows = {}
    for t in range(3):
        if t == 1:
            f, src = hash_flow(g, ms)
        else:
            f = g.Flow(g.PAYLOAD_PARTICLE if t == 0 else g.PAYLOAD_PHOTON)
            if t == 0:
                src, *_ = g.feedback_network(f, g.RecordWeight(rec.data_ptr(), int(max_items) + 1))
            else:
                src, *_ = g.photon_network(f, tally.data_ptr())
        flows[t] = (f, src)
    one_t = 0.0
    for i in sample:
        f, src = flows[int(topo[i])]
        f.configure(g.Tuning(seed=int(seeds[i]), batch_chunks=8, chunk_capacity=8192,
                             live_particles=32768 if int(topo[i]) == 2 else 65536, schedule=g.SCHEDULE_DEVICE,
                             photon=dict(g.PHOTON_DEFAULTS)))
        f.run(src, min(int(items[i]), 16384))  # engine build (same pool shape) outside the timed run
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        f.run(src, int(items[i]))
        torch.cuda.synchronize()
        one_t += time.perf_counter() - t0
    one = int(items[sample].sum()) / one_t
    out["batched_one_at_a_time_items_per_s"] = one
    out["batched_sample_jobs"] = len(sample)
    out["speedup_vs_batched_one_at_a_time"] = out["value"] / one
    plain_t = 0.0
    for i in small:
        f, src = flows[int(topo[i])]
        f.configure(g.Tuning(seed=int(seeds[i]), batch_chunks=1, chunk_capacity=8192, live_particles=16384,
                             schedule=g.SCHEDULE_HOST, photon=dict(g.PHOTON_DEFAULTS)))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        f.run(src, int(items[i]))
        torch.cuda.synchronize()
        plain_t += time.perf_counter() - t0
    plain = int(items[small].sum()) / plain_t
    out["plain_one_at_a_time_items_per_s"] = plain
    out["plain_sample_jobs"] = len(small)
    out["speedup_vs_plain"] = out["value"] / plain
    log(f"[bench] served {jobs} jobs at {rate:.0f}/s: {out['value']:.4g} items/s, p50 {out['p50_latency_ms']:.1f} ms, "
        f"p99 {out['p99_latency_ms']:.1f} ms; one at a time batched {one:.4g}, plain {plain:.4g} items/s; p50/p99 per "
        f"topology (ms): " + ", ".join(f"{k} {v['p50_latency_ms']:.1f}/{v['p99_latency_ms']:.1f}" for k, v in by_topo.items()))
    for f, _ in flows.values():
        f.close()
    return out


# ---- the CPU leg: reference CPU path (oracle restatement) and every form's CPU check ---------------
def cpu_model():
    try:
        for line in subprocess.run(["lscpu"], capture_output=True, text=True).stdout.splitlines():
            if line.startswith("Model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_leg(g, flow, N, C, P, Q, rows, Nc, forms, hash_sets, photon_tallies, photons):
    import numpy as np

    import oracle as O

    threads = min(os.cpu_count() or 1, 16)
    graph = flow.neutral_graph()
    base = {}
    for th in (1, threads):
        t0 = time.perf_counter()
        r = O.run(graph, seed=42, items=Nc, capacity=C, live=P, batch=Q, mode=O.MODE_RUNNER, threads=th,
                  hist_size=g.FEEDBACK_HIST_N)
        base[th] = (Nc / (time.perf_counter() - t0), r["status"])
        log(f"[bench] cpu: oracle runner schedule N={Nc} threads={th}: {base[th][0]:.4g} particles/s")
    out = {"value": base[1][0], "unit": "particles/s", "cores": 1, "kind": "port",
           "sample": f"oracle restatement of the reference CPU path (GraphRunner schedule, classify / plan / "
                     f"scatter per batch), same network, N={Nc:.0e} C={C} P={P} Q={Q} seed 42, one thread, "
                     f"status={base[1][1]}",
           "all_cores": {"value": base[threads][0], "threads": threads, "status": base[threads][1]},
           "nproc": os.cpu_count(), "cpu_model": cpu_model()}
    checks = {}
    # the headline histogram against the schedule-free model at the full N
    t0 = time.perf_counter()
    w, _ = O.feedback_scalar(42, N, threads=threads)
    v, c = np.unique(w, return_counts=True)
    checks["headline_histogram_equals_oracle"] = bool(len(v) == len(rows) and (rows[:, 0] == v).all()
                                                      and (rows[:, 1] == c).all())
    checks["headline_oracle_seconds"] = time.perf_counter() - t0
    del w
    # config 3: every Q's per-bucket multisets against the schedule-free model
    if hash_sets:
        n = next(iter(forms[k]["items"] for k in forms if k.startswith("hash_pipeline")))
        msum, mcnt = O.keyed_scalar(42, n, threads=threads)
        for q, got in hash_sets.items():
            checks[f"hash_Q{q}_buckets_equal_oracle"] = bool((got[:, 0] == msum).all() and (got[:, 1] == mcnt).all())
    # config 4: the batched photon tallies against the FP64 oracle on the same Philox words
    if photon_tallies:
        t0 = time.perf_counter()
        ref, steps = O.photon_scalar(42, photons, dict(g.PHOTON_DEFAULTS), threads=threads)
        for name, t in photon_tallies.items():
            pre = "photon_fp64" if name == "photon_batched_team" else "photon_wide_fp64"
            checks[f"{pre}_max_bin_rel_err"] = float((np.abs(t - ref) / ref).max())
            checks[f"{pre}_total_rel_err"] = float(abs(t.sum() - ref.sum()) / ref.sum())
        checks["photon_fp64_oracle_steps"] = int(steps)
        checks["photon_oracle_seconds"] = time.perf_counter() - t0
    out["checks"] = checks
    log(f"[bench] cpu checks: {checks}")
    return out


def headline(args, g, torch, golden, N, C, P, Q, fuse, wkey):
    """The headline runs: autotune, warmup, the timed runs, the histogram checks and the roofline."""
    log(f"[bench] batched form N={N} C={C} P={P} Q={Q}")
    flow, src, counts = build_feedback(g, torch, C, P, Q, g.SCHEDULE_DEVICE, fuse=fuse, engine=args.engine,
                                       team=args.team)
    tuned = None
    if not args.no_autotune and args.engine == 0 and not args.team:
        tuned = flow.autotune(src, probe_items=min(N, 10_000_000))
        log(f"[bench] autotuner picked {tuned}")
    eng = tuned["device_engine"] if tuned else args.engine
    wg = tuned["max_workgroups"] if tuned else (args.team or None)
    tile = tuned["tile_items"] if tuned else None
    # the team engine's launches are timed with HIP events on its stream in the timed runs themselves
    # (a handful of launches per run); the wide engine keeps its hipGraph replay
    timed_graph = 2 if eng in (0, 3) else 1
    flow.configure(g.Tuning(seed=42, batch_chunks=Q, chunk_capacity=C, live_particles=P, schedule=g.SCHEDULE_DEVICE,
                            use_graph=timed_graph, fuse_transforms=fuse, device_engine=eng, max_workgroups=wg or None,
                            tile_items=tile or None))
    for i in range(args.warmup):
        dt, st = run_once(flow, src, counts, N, torch)
        log(f"[bench] warmup {i}: {dt:.3f} s, {N / dt:.4g} particles/s, steps={st['steps']}")
    times, k_ms, k_n = [], 0.0, 0
    for i in range(args.steps):
        dt, st = run_once(flow, src, counts, N, torch)
        times.append(dt)
        ms, n = flow.kernel_times()
        k_ms += ms[3]
        k_n += n[3]
        log(f"[bench] step {i}: {dt:.3f} s, {N / dt:.4g} particles/s, steps={st['steps']}")
    tmean = sum(times) / len(times)
    value = N / tmean
    rows = g.hist_rows(counts.data_ptr(), g.FEEDBACK_HIST_N, g.FEEDBACK_HIST_LO)
    row_sum = int(rows[:, 1].sum()) if len(rows) else 0
    sha = g.rows_sha256(rows)
    part_items, xf_items = flow.item_passes()
    log(f"[bench] rows={len(rows)} sum={row_sum} sha256={sha}")

    result = {
        "metric": METRIC, "value": value, "unit": "particles/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": tmean * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+u32", "data": "synthetic (device Philox source, seed 42)",
        "config": {"workload": f"particle_feedback N={N:.0e} C={C} P={P} Q={Q} seed=42",
                   "form": ("batched: device engine " + ENGINES[int(st["engine"])]
                            + (", A/B fused into the filter scatter" if fuse else "")),
                   "engine_steps": int(st["steps"]),
                   "partition_item_passes": int(part_items), "transform_item_passes": int(xf_items),
                   "histogram_rows": int(len(rows)), "histogram_sum": row_sum, "histogram_sha256": sha,
                   "autotuned": tuned},
    }
    if golden and N == golden["items"]:
        result["config"]["histogram_equals_golden"] = sha == golden["cases"]["parity"]["sha256"]
        log(f"[bench] histogram equals the oracle fixture: {result['config']['histogram_equals_golden']}")

    # roofline of the dominant kernel. The team engine's k_team runs whole steps: its bound is the
    # step latency (P=16384 live particles, ~10.5K items a step, a chain of ~152K dependent steps),
    # not HBM bandwidth, so the HBM fraction is reported against the 8 TB/s peak but bound = latency.
    alg_bytes = 24 * part_items + 24 * xf_items + 12 * N + 12 * N  # partition/transform passes, source, sink
    if k_n and int(st["engine"]) == 3:
        per_run = k_n / args.steps
        alg_launch = alg_bytes / per_run
        avg_us = k_ms * 1e3 / k_n
        achieved = alg_launch / (avg_us * 1e-6) / 1e9
        tr = traffic_for(wkey)
        result["roofline"] = {
            "bound": "latency", "kernel": "k_team (whole steps, team of workgroups on one XCD)",
            "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
            "traffic": tr[0] if tr else None,
            "traffic_source": (os.path.relpath(TRAFFIC_FILE, ROOT) + " (committed rocprofv3 --pmc summary of this "
                               "workload, not measured in this run)") if tr else None,
            "algorithmic_bytes_per_launch": alg_launch, "avg_launch_us": avg_us, "launches": k_n,
            "launch_times_from": f"HIP events on the engine stream around every k_team launch of the {args.steps} timed runs",
            "steps_per_launch": int(st["steps"]) / per_run, "avg_step_us": k_ms * 1e3 / (int(st["steps"]) * args.steps),
            "step_floor_note": "a step is two dependent single-wave chains plus one team barrier (dependent L2 "
                               "load ~220 cycles, one-XCD barrier of 32 workgroups 0.5-0.8 us)",
        }
        log(f"[bench] roofline: k_team avg {avg_us:.1f} us over {k_n} launches, {achieved:.1f} GB/s "
            f"({achieved / HBM_PEAK_GBS:.2e} of peak), {result['roofline']['avg_step_us']:.2f} us per step")

    if args.phases:
        flow.configure(g.Tuning(seed=42, batch_chunks=Q, chunk_capacity=C, live_particles=P,
                                schedule=g.SCHEDULE_DEVICE, use_graph=3, fuse_transforms=fuse, device_engine=eng,
                                max_workgroups=wg or None, tile_items=tile or None))
        counts.zero_()
        pst = flow.run(src, N)
        phases, psteps = flow.phase_profile()
        log(f"[bench] phases (us/step over {psteps} steps): { {k: round(v, 3) for k, v in phases.items()} }")
        if args.members and int(pst["engine"]) == 3:
            for m, r in enumerate(flow.member_profile(args.team or 32)):
                log(f"[bench] member {m:2d}: wait {r[0]:.2f} start {r[1]:.2f} scatter {r[2]:.2f} "
                    f"sched {r[3]:.2f} us, segments {r[4]:.2f} sinks {r[5]:.2f} per step, wave work "
                    + " ".join(f"{x:.2f}" for x in r[6:]))
    return result, flow, rows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--items", type=float, default=1e8)
    ap.add_argument("--capacity", type=int, default=8192)
    ap.add_argument("--live", type=int, default=16384)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--plain-items", type=float, default=2e6)
    ap.add_argument("--plain-full-items", type=float, default=None,
                    help="the plain forms at the headline's Q (default 1e8, 0 with --skip-headline; 0 = skip)")
    ap.add_argument("--cpu-items", type=float, default=2e7, help="CPU baseline sample (oracle runner schedule)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--only-headline", action="store_true", help="the headline runs and roofline only")
    ap.add_argument("--skip-headline", action="store_true",
                    help="run only the selected forms (for profiling one form; implies --no-cpu)")
    ap.add_argument("--engine", type=int, default=0, help="device engine: 0 autotuned, 1 wide, 2 micro, 3 team")
    ap.add_argument("--team", type=int, default=0, help="team engine: workgroups in the team (0 = default)")
    ap.add_argument("--no-autotune", action="store_true", help="skip the autotuner's probe runs")
    ap.add_argument("--phases", action="store_true", help="device engines: one more run with the per-phase profile")
    ap.add_argument("--members", action="store_true", help="team engine: with --phases, per-member phase times")
    ap.add_argument("--no-fuse", action="store_true", help="run A/B as separate transform batches")
    ap.add_argument("--photons", type=float, default=1e8, help="photon transport form: photons (0 = skip)")
    ap.add_argument("--plain-photons", type=float, default=1e6)
    ap.add_argument("--photon-live", type=int, default=1 << 24,
                    help="photon transport on the wide engine with this many photons in flight (0 = skip)")
    ap.add_argument("--hash-items", type=float, default=1e8, help="keyed hash pipeline form: keys (0 = skip)")
    ap.add_argument("--hash-q", default="1,4,8,16", help="keyed hash pipeline form: batch sizes Q, comma separated")
    ap.add_argument("--hash-lanes", type=int, default=63, help="keyed hash pipeline: inflight batches per node")
    ap.add_argument("--hash-live", type=int, default=1 << 24, help="keyed hash pipeline: keys in flight (P)")
    ap.add_argument("--hash-source-chunks", type=int, default=16,
                    help="keyed hash pipeline: chunks a source lane admits per step")
    ap.add_argument("--hash-tile", type=int, default=1024, help="keyed hash pipeline: items per workgroup tile")
    ap.add_argument("--serve-jobs", type=int, default=4096, help="served mixed stream: jobs (0 = skip)")
    ap.add_argument("--serve-rate", type=float, default=80.0, help="served mixed stream: Poisson arrivals per second")
    ap.add_argument("--serve-max-items", type=float, default=1e7, help="served mixed stream: N ~ log-uniform[1e3, this]")
    ap.add_argument("--no-distributions", action="store_true")
    ap.add_argument("--wide-live", type=int, default=1 << 24,
                    help="feedback network on the wide engine with this many particles in flight (0 = skip)")
    ap.add_argument("--copy-calib", type=float, default=0.0,
                    help="after the runs, copy this many MiB device to device once (the known-size kernel that "
                         "scales rocprofv3 FETCH_SIZE / WRITE_SIZE readings to bytes)")
    ap.add_argument("--pmc-summary", nargs=3, metavar=("FETCH_CSV", "WRITE_CSV", "OUT"),
                    help="write the HBM traffic summary of two rocprofv3 --pmc runs and exit")
    ap.add_argument("--pmc-kernel", default="k_team", help="--pmc-summary: kernels whose name contains this")
    ap.add_argument("--pmc-command", default=None, help="--pmc-summary: the profiled bench.py arguments")
    ap.add_argument("--pmc-workload", default=None, help="--pmc-summary: workload label (default: the headline's)")
    args = ap.parse_args()

    N, C, P, Q = int(args.items), args.capacity, args.live, args.batch
    fuse = not args.no_fuse
    # the HBM traffic of a launch depends on the workload, not on the team size
    wkey = f"particle_feedback N={N:.0e} C={C} P={P} Q={Q} seed=42 fuse={int(fuse)}"
    if args.pmc_summary:
        pmc_summary(*args.pmc_summary[:2], args.pmc_workload or wkey, args.pmc_summary[2], args.pmc_kernel,
                    args.pmc_command)
        return

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus above 1 is not supported "
                                   "(the path would shard by particle-id range, one flow per GPU)"}))
        sys.exit(2)

    import torch
    import graphflow_amd as g

    golden = golden_cases()
    result, flow, rows = {"metric": METRIC, "n_gpus": 1, "config": {}}, None, None
    if not args.skip_headline:
        result, flow, rows = headline(args, g, torch, golden, N, C, P, Q, fuse, wkey)
    forms, hash_sets, photon_tally = {}, {}, None
    if not args.only_headline:
        Np = int(args.plain_items)
        for name, kw in (("plain_host_planned_Q1", {}), ("plain_fused_batches_Q1", {"fused_batches": True})):
            if Np <= 0:
                break
            pflow, psrc, pcounts = build_feedback(g, torch, C, P, 1, g.SCHEDULE_HOST, **kw)
            dt, _ = run_once(pflow, psrc, pcounts, Np, torch)
            forms[name] = {"value": Np / dt, "unit": "particles/s", "items": Np}
            log(f"[bench] {name} N={Np}: {Np / dt:.4g} particles/s")
            pflow.close()
        # the headline's own configuration (N=1e8, Q=8) in the reference's protocol: the host runner,
        # host-planned batches or one fused batch per partition batch; its histogram must be the
        # headline's
        Nf = int(args.plain_full_items if args.plain_full_items is not None else 0 if args.skip_headline else N)
        for name, kw in (("plain_host_planned_Q8", {}), ("plain_fused_batches_Q8", {"fused_batches": True})):
            if Nf <= 0:
                break
            pflow, psrc, pcounts = build_feedback(g, torch, C, P, Q, g.SCHEDULE_HOST, **kw)
            dt, _ = run_once(pflow, psrc, pcounts, Nf, torch)
            prow = g.hist_rows(pcounts.data_ptr(), g.FEEDBACK_HIST_N, g.FEEDBACK_HIST_LO)
            psha = g.rows_sha256(prow)
            forms[name] = {"value": Nf / dt, "unit": "particles/s", "items": Nf, "Q": Q, "histogram_sha256": psha}
            if golden and Nf == golden["items"]:
                forms[name]["histogram_equals_golden"] = psha == golden["cases"]["parity"]["sha256"]
            log(f"[bench] {name} N={Nf}: {Nf / dt:.4g} particles/s; histogram equals golden: "
                f"{forms[name].get('histogram_equals_golden')}")
            pflow.close()
        if not args.no_distributions:
            forms.update(distribution_forms(g, torch, N, C, P, Q, golden))
        if args.wide_live > 0:
            forms[f"feedback_wide_engine_live_{args.wide_live}"] = wide_feedback_form(g, torch, N, C, args.wide_live,
                                                                                     golden)
        if args.hash_items > 0:
            hf, hash_sets = hash_forms(g, torch, int(args.hash_items), tuple(int(q) for q in args.hash_q.split(",")),
                                       args.hash_lanes, args.hash_live, args.hash_source_chunks, args.hash_tile)
            forms.update(hf)
        if args.photons > 0:
            pf, pt = photon_forms(g, torch, int(args.photons), int(args.plain_photons), C, P, Q, args.photon_live)
            forms.update(pf)
            photon_tally = {k: v for k, v in pt.items() if k.startswith("photon_batched")}
        if args.serve_jobs > 0:
            forms["served_mixed_stream"] = served_forms(g, torch, args.serve_jobs, args.serve_max_items,
                                                        args.serve_rate)
    result["forms"] = forms

    if not args.no_cpu and not args.skip_headline:
        result["cpu_baseline"] = cpu_leg(g, flow, N, C, P, Q, rows, int(args.cpu_items), forms, hash_sets,
                                         photon_tally, int(args.photons))
        ck = result["cpu_baseline"]["checks"]
        result["config"]["histogram_equals_oracle"] = ck["headline_histogram_equals_oracle"]

    if args.copy_calib > 0:
        nbytes = int(args.copy_calib * (1 << 20)) // 4 * 4
        a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        log(f"[bench] copy calibration: {nbytes} bytes read and {nbytes} bytes written by one device copy")

    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
