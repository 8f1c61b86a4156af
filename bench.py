"""Benchmark of the transform-batch hot path on one MI355X.

Workload: the fp32 particle feedback network Generate -> Filter -> {A (x0.5), B (x0.75)} -> Filter,
Filter.sink -> Sink at N=1e8, C=8192, P=16384, Q=8, seed 42 (synthetic inputs generated on the
device by the source node). One "step" of this benchmark is one complete run of the network over
all N particles; `value` is N / seconds per run, measured around `Flow.run` with the pool, the
sink histogram and the engine buffers already resident in HBM.

Beside the headline (batched form: device step engine replayed from a hipGraph) it reports the
plain one-batch-at-a-time form, the roofline of the dominant engine kernel (timed with HIP events
on the engine's stream) and the CPU oracle's restatement of the reference path on a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "particles/s at N=1e8 C=8192 P=16384 Q=8; % HBM peak; served p99 job latency"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


ENGINES = {0: "plain host-planned", 1: "wide (hipGraph of step kernels)", 2: "micro (one persistent workgroup)",
           3: "team (one persistent launch, workgroups of one XCD, replicated scheduler)",
           4: "team handing over to wide"}


def build_feedback(g, torch, N, C, P, Q, schedule, seed=42, use_graph=1, fuse=False, engine=0, tile_items=None,
                   team=0):
    counts = torch.zeros(g.FEEDBACK_HIST_N, dtype=torch.int64, device="cuda")
    flow = g.Flow(g.PAYLOAD_PARTICLE)
    src, *_ = g.feedback_network(flow, g.WeightHistogram(counts.data_ptr(), g.FEEDBACK_HIST_LO, g.FEEDBACK_HIST_N))
    flow.configure(g.Tuning(seed=seed, batch_chunks=Q, chunk_capacity=C, live_particles=P, schedule=schedule,
                            use_graph=use_graph, tile_items=tile_items, fuse_transforms=fuse, device_engine=engine,
                            max_workgroups=team or None))
    return flow, src, counts


def tuning_of(g, N, C, P, Q, fuse, args, tuned, use_graph):
    """The timed configuration again (the autotuner's pick when it ran), with another use_graph."""
    eng = tuned["device_engine"] if tuned else args.engine
    wg = tuned["max_workgroups"] if tuned else (args.team or None)
    tile = tuned["tile_items"] if tuned else None
    return g.Tuning(seed=42, batch_chunks=Q, chunk_capacity=C, live_particles=P, schedule=g.SCHEDULE_DEVICE,
                    use_graph=use_graph, fuse_transforms=fuse, device_engine=eng, max_workgroups=wg or None,
                    tile_items=tile or None)


def run_once(flow, src, counts, N, torch):
    counts.zero_()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st = flow.run(src, N)
    torch.cuda.synchronize()
    return time.perf_counter() - t0, st


TRAFFIC_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "k_team_feedback_N1e8_hbm_traffic.txt")


def pmc_summary(fetch_csv, write_csv, workload, out):
    """Turn two rocprofv3 --pmc runs (FETCH_SIZE, then WRITE_SIZE; counter_collection.csv, KiB per
    dispatch) of `bench.py --copy-calib 1024` into HBM bytes per k_team launch. Each counter is
    scaled by what it reads for the calibration copy of known size (the last device copy of the
    run: 1 GiB read and 1 GiB written)."""
    import csv

    def per_dispatch(path, counter):
        d = {}
        for r in csv.DictReader(open(path)):
            if r["Counter_Name"] != counter:
                continue
            k = (int(r["Dispatch_Id"]), r["Kernel_Name"])
            d[k] = d.get(k, 0.0) + float(r["Counter_Value"])
        return d

    lines = ["# HBM traffic of the team engine kernel (k_team), from two rocprofv3 --pmc passes:",
             "#   cd /tmp && rocprofv3 --pmc FETCH_SIZE --output-format csv -d $OUT/pmc_fetch -o run -- "
             "python $R/bench.py --no-cpu --no-plain --photons 0 --serve-jobs 0 --hash-items 0 --no-autotune --no-roofline "
             "--steps 1 --warmup 0 --items 1e8 --copy-calib 1024",
             "#   (the same with --pmc WRITE_SIZE into $OUT/pmc_write; R = the repository root, OUT = an output directory)",
             "#   python bench.py --pmc-summary $OUT/pmc_fetch/run_counter_collection.csv "
             "$OUT/pmc_write/run_counter_collection.csv <this file>",
             f"workload={workload}"]
    per = {}
    for counter, path in (("FETCH_SIZE", fetch_csv), ("WRITE_SIZE", write_csv)):
        d = per_dispatch(path, counter)
        copies = [(k, v) for k, v in d.items() if "copyBuffer" in k[1]]
        cal_k, cal_v = max(copies, key=lambda kv: kv[0][0])
        factor = float(1 << 20) / cal_v  # KiB the 1 GiB copy moved / KiB the counter read
        team = [v for k, v in sorted(d.items()) if "k_team" in k[1]]
        per[counter] = [v * factor * 1024.0 for v in team]
        lines.append(f"{counter}: calibration copy dispatch {cal_k[0]} read {cal_v:.1f} KiB for 1048576 KiB "
                     f"-> factor {factor:.4f}; k_team dispatches (KiB): {', '.join(f'{v:.1f}' for v in team)}")
    launches = len(per["FETCH_SIZE"])
    total = sum(per["FETCH_SIZE"]) + sum(per["WRITE_SIZE"])
    lines.append(f"launches={launches}")
    lines.append(f"fetch_bytes_total={sum(per['FETCH_SIZE']):.0f}")
    lines.append(f"write_bytes_total={sum(per['WRITE_SIZE']):.0f}")
    lines.append(f"traffic_bytes_per_launch={total / max(launches, 1):.0f}")
    with open(out, "w") as f:
        f.write("\n".join(lines) + "\n")
    print("\n".join(lines))


def traffic_for(workload, launches):
    """HBM bytes per k_team launch from the committed PMC summary of this workload, or None."""
    import re

    try:
        kv = dict(m.groups() for m in (re.match(r"^([a-z_]+)=(.*)$", l.strip()) for l in open(TRAFFIC_FILE)) if m)
    except OSError:
        return None
    if kv.get("workload") != workload or int(kv.get("launches", -1)) != launches:
        return None
    return float(kv["traffic_bytes_per_launch"])


def photon_forms(g, torch, N, Np, C, P, Q, verify):
    """Photon transport (10 layers of 1 mfp, albedo 0.9, HG g=0.9, roulette 1e-4 / 0.1, 1024 depth
    bins): the batched persistent form (team engine) and the plain host-planned one-batch form,
    photons/s and photon steps/s; with verify, the worst per-bin relative difference of the batched
    tallies from the FP64 CPU oracle on the same Philox words."""
    cfg = dict(g.PHOTON_DEFAULTS)
    out = {}
    for name, n, kw in (("photon_batched_team", N, dict(schedule=g.SCHEDULE_DEVICE, batch_chunks=Q)),
                        ("photon_plain_host_planned_Q1", Np, dict(schedule=g.SCHEDULE_HOST, batch_chunks=1))):
        tally = torch.zeros(cfg["tally_bins"], dtype=torch.float64, device="cuda")
        flow = g.Flow(g.PAYLOAD_PHOTON)
        src, *_ = g.photon_network(flow, tally.data_ptr())
        flow.configure(g.Tuning(seed=42, chunk_capacity=C, live_particles=P, photon=cfg, **kw))
        flow.run(src, min(n, 100000))  # warm-up: engine build, pool
        tally.zero_()
        dt, st = run_once(flow, src, tally, n, torch)
        part_items, xf_items = flow.item_passes()
        photon_steps = part_items  # one alive-partition pass per photon step (the step is fused into it)
        rec = {"value": n / dt, "unit": "photons/s", "items": n, "photon_steps_per_s": photon_steps / dt,
               "photon_steps": int(photon_steps), "engine_steps": int(st["steps"])}
        if verify and name == "photon_batched_team":
            import numpy as np
            import oracle as O
            ref, ref_steps = O.photon_scalar(42, n, cfg, threads=min(os.cpu_count() or 1, 16))
            t = tally.cpu().numpy()
            rec["fp64_oracle_max_bin_rel_err"] = float((np.abs(t - ref) / ref).max())
            rec["fp64_oracle_total_rel_err"] = float(abs(t.sum() - ref.sum()) / ref.sum())
            rec["fp64_oracle_photon_steps"] = int(ref_steps)
        out[name] = rec
        log(f"[bench] {name} N={n}: {n / dt:.4g} photons/s, {photon_steps / dt:.4g} photon steps/s "
            + (f"(max bin rel err vs FP64 {rec['fp64_oracle_max_bin_rel_err']:.2e})" if "fp64_oracle_max_bin_rel_err" in rec else ""))
    return out


def hash_form(g, torch, N, qs=(8, 16), autotune=True):
    """Keyed hash pipeline (Zipf(1.1) keys over 2^20 values, murmur3 finalizer into 4096 buckets,
    C=8192) at each Q of `qs`: keys/s on the device engine, one partition pass per key (24 B of
    algorithmic traffic per key). Without autotune: the wide engine, 1024 workgroups, 1024-item
    tiles."""
    out = {}
    bins = 4096
    ms = torch.zeros(2 * bins, dtype=torch.int64, device="cuda")
    for Q in qs:
        f = g.Flow(g.PAYLOAD_KEYED)
        src = f.source("keys", g.ZipfKeys(1 << 20, 1.1))
        part = f.partition("bucket", g.KeyBucket(bins))
        snk = f.sink("digest", g.ChunkDigest(0, 0, ms.data_ptr(), bins))
        f.connect(src, part)
        for b in range(bins):
            f.connect(part, snk, b)
        if autotune:
            f.configure(g.Tuning(seed=42, batch_chunks=Q, chunk_capacity=8192, schedule=g.SCHEDULE_DEVICE))
            tuned = f.autotune(src, probe_items=min(N, 1_000_000))
        else:
            tuned = {"device_engine": 1, "max_workgroups": 1024, "batch_chunks": Q, "tile_items": 1024}
            f.configure(g.Tuning(seed=42, batch_chunks=Q, chunk_capacity=8192, schedule=g.SCHEDULE_DEVICE,
                                 device_engine=1, max_workgroups=1024, tile_items=1024))
        dt, st = run_once(f, src, ms, N, torch)
        out[f"hash_pipeline_B4096_Q{Q}"] = {"value": N / dt, "unit": "keys/s", "items": N, "engine_steps": int(st["steps"]),
                                           "engine": int(st["engine"]), "autotuned": tuned}
        log(f"[bench] hash pipeline B=4096 Q={Q} N={N}: {N / dt:.4g} keys/s ({int(st['steps'])} steps, {tuned})")
    return out


def served_form(g, torch, jobs, rate, max_items, seed=42, plain=True):
    """Served mixed stream: `jobs` independent flow jobs, Poisson arrivals at `rate` per second,
    N ~ log-uniform[1e3, max_items], topology uniform over {particle feedback, keyed hash pipeline,
    photon transport}; aggregate items/s and p50/p99 job latency (submit to finish), next to the
    same jobs run one at a time with the plain host-planned form (Q=1)."""
    import threading
    import numpy as np

    rng = np.random.default_rng(seed)
    topo = rng.integers(0, 3, jobs)
    items = np.exp(rng.uniform(np.log(1e3), np.log(max_items), jobs)).astype(np.int64)
    arrive = np.cumsum(rng.exponential(1.0 / rate, jobs))
    seeds = 1000 + np.arange(jobs)
    sv = g.Server(slots=8, max_items=int(max_items) + 1, prewarm=True)  # every slot builds its flows first
    ids = [None] * jobs

    def submitter():
        t0 = time.perf_counter()
        for i in range(jobs):
            d = arrive[i] - (time.perf_counter() - t0)
            if d > 0:
                time.sleep(d)
            ids[i] = sv.submit(int(topo[i]), int(items[i]), int(seeds[i]))

    th = threading.Thread(target=submitter)
    th.start()
    th.join()
    res = sv.poll(wait_all=True)
    sv.close()
    bad = [r for r in res if r["status"] != 0]
    lat = np.array([r["end_ms"] - r["submit_ms"] for r in res])
    t_first = min(r["submit_ms"] for r in res)
    t_last = max(r["end_ms"] for r in res)
    total = int(items.sum())
    served = total / ((t_last - t_first) * 1e-3)
    per_topo = {}
    for t in range(3):
        rs = [r for r in res if r["topology"] == t]
        if rs:
            per_topo[g.JOB_NAMES[t]] = {
                "jobs": len(rs), "items": int(sum(r["items"] for r in rs)),
                "mean_service_ms": float(np.mean([r["end_ms"] - r["start_ms"] for r in rs])),
                "service_items_per_s": float(sum(r["items"] for r in rs) / max(1e-9, 1e-3 * sum(r["end_ms"] - r["start_ms"] for r in rs)))}
    # plain form: every 4th job of the same stream one at a time, host-planned one batch at a time
    sample = list(range(0, jobs, 4)) if plain else []
    plain_t = 0.0
    tally = torch.zeros(1024, dtype=torch.float64, device="cuda")
    rec = torch.zeros(int(max_items) + 1, dtype=torch.int32, device="cuda")
    ms = torch.zeros(2 * 4096, dtype=torch.int64, device="cuda")
    flows = {}
    for t in range(3):
        f = g.Flow([g.PAYLOAD_PARTICLE, g.PAYLOAD_KEYED, g.PAYLOAD_PHOTON][t])
        if t == 0:
            src, *_ = g.feedback_network(f, g.RecordWeight(rec.data_ptr(), int(max_items) + 1))
        elif t == 1:
            src = f.source("keys", g.ZipfKeys(1 << 20, 1.1))
            part = f.partition("bucket", g.KeyBucket(4096))
            snk = f.sink("digest", g.ChunkDigest(0, 0, ms.data_ptr(), 4096))
            f.connect(src, part)
            for b in range(4096):
                f.connect(part, snk, b)
        else:
            src, *_ = g.photon_network(f, tally.data_ptr())
        flows[t] = (f, src)
    for i in sample:
        f, src = flows[int(topo[i])]
        f.configure(g.Tuning(seed=int(seeds[i]), batch_chunks=1, chunk_capacity=8192, live_particles=16384,
                             schedule=g.SCHEDULE_HOST, photon=dict(g.PHOTON_DEFAULTS)))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        f.run(src, int(items[i]))
        torch.cuda.synchronize()
        plain_t += time.perf_counter() - t0
    plain = int(items[sample].sum()) / plain_t if sample else None
    out = {"value": served, "unit": "items/s", "jobs": jobs, "items": total, "arrival_rate_per_s": rate,
           "per_topology": per_topo, "plain_sample_jobs": len(sample),
           "n_range": [1000, int(max_items)], "slots": 8, "failed_jobs": len(bad),
           "p50_latency_ms": float(np.percentile(lat, 50)), "p99_latency_ms": float(np.percentile(lat, 99)),
           "makespan_s": (t_last - t_first) * 1e-3, "plain_one_at_a_time_items_per_s": plain,
           "speedup_vs_plain": served / plain if plain else None,
           "jobs_per_topology": {g.JOB_NAMES[t]: int((topo == t).sum()) for t in range(3)}}
    log(f"[bench] served mixed stream at {rate:.0f} jobs/s: {jobs} jobs, {total} items, {served:.4g} items/s, "
        f"p50 {out['p50_latency_ms']:.2f} ms, p99 {out['p99_latency_ms']:.2f} ms"
        + (f"; plain one at a time {plain:.4g} items/s ({served / plain:.1f}x)" if plain else ""))
    return out


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
    ap.add_argument("--cpu-items", type=float, default=2e7)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-roofline", action="store_true")
    ap.add_argument("--verify", action="store_true", help="compare the sorted histogram with the oracle")
    ap.add_argument("--engine", type=int, default=0, help="device engine: 0 autotuned, 1 wide, 2 micro, 3 team")
    ap.add_argument("--team", type=int, default=0, help="team engine: workgroups in the team (0 = default)")
    ap.add_argument("--no-autotune", action="store_true", help="skip the autotuner's probe runs")
    ap.add_argument("--phases", action="store_true", help="device engines: one more run with the per-phase profile")
    ap.add_argument("--members", action="store_true", help="team engine: with --phases, per-member phase times")
    ap.add_argument("--no-fuse", action="store_true", help="run A/B as separate transform batches")
    ap.add_argument("--photons", type=float, default=1e7, help="photon transport form: photons (0 = skip)")
    ap.add_argument("--plain-photons", type=float, default=1e6)
    ap.add_argument("--hash-items", type=float, default=4e6, help="keyed hash pipeline form: keys (0 = skip)")
    ap.add_argument("--hash-q", default="8,16", help="keyed hash pipeline form: batch sizes Q, comma separated")
    ap.add_argument("--serve-jobs", type=int, default=256, help="served mixed stream: jobs (0 = skip)")
    ap.add_argument("--serve-rate", type=float, default=1000.0, help="served mixed stream: Poisson arrivals per second")
    ap.add_argument("--serve-max-items", type=float, default=1e6, help="served mixed stream: N ~ log-uniform[1e3, this]")
    ap.add_argument("--copy-calib", type=float, default=0.0,
                    help="after the runs, copy this many MiB device to device once (the known-size kernel that "
                         "scales rocprofv3 FETCH_SIZE / WRITE_SIZE readings to bytes)")
    ap.add_argument("--pmc-summary", nargs=3, metavar=("FETCH_CSV", "WRITE_CSV", "OUT"),
                    help="write the HBM traffic summary of two rocprofv3 --pmc runs and exit")
    args = ap.parse_args()

    N, C, P, Q = int(args.items), args.capacity, args.live, args.batch
    fuse = not args.no_fuse
    # the HBM traffic of a launch depends on the workload, not on the team size
    wkey = f"particle_feedback N={N:.0e} C={C} P={P} Q={Q} seed=42 fuse={int(fuse)}"
    if args.pmc_summary:
        pmc_summary(*args.pmc_summary[:2], wkey, args.pmc_summary[2])
        return

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus above 1 is not supported "
                                   "(the path would shard by particle-id range, one flow per GPU)"}))
        sys.exit(2)

    import torch
    import graphflow_amd as g

    log(f"[bench] batched form N={N} C={C} P={P} Q={Q}")
    flow, src, counts = build_feedback(g, torch, N, C, P, Q, g.SCHEDULE_DEVICE, fuse=fuse, engine=args.engine,
                                       team=args.team)
    tuned = None
    if not args.no_autotune and args.engine == 0 and not args.team:
        tuned = flow.autotune(src, probe_items=min(N, 10_000_000))
        log(f"[bench] autotuner picked {tuned}")
    for i in range(args.warmup):
        dt, st = run_once(flow, src, counts, N, torch)
        log(f"[bench] warmup {i}: {dt:.3f} s, {N / dt:.4g} particles/s, steps={st['steps']}")
    times = []
    for i in range(args.steps):
        dt, st = run_once(flow, src, counts, N, torch)
        times.append(dt)
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

    if args.verify:
        import numpy as np
        import oracle as O
        w, _ = O.feedback_scalar(42, N, threads=min(os.cpu_count() or 1, 16))
        v, c = np.unique(w, return_counts=True)
        ok = len(v) == len(rows) and bool((rows[:, 0] == v).all()) and bool((rows[:, 1] == c).all())
        result["config"]["histogram_equals_oracle"] = ok
        log(f"[bench] histogram equals oracle: {ok}")

    if not args.no_roofline:
        # one extra run with HIP events around every engine kernel, on the engine's own stream
        flow.configure(tuning_of(g, N, C, P, Q, fuse, args, tuned, use_graph=2))
        counts.zero_()
        tst = flow.run(src, N)
        ms, n = flow.kernel_times()
        part_items, xf_items = flow.item_passes()
        # algorithmic bytes: 24 B per partition or transform item pass (one 12 B read and one 12 B
        # write of the payload), 12 B per generated particle (write) and per sunk particle (read)
        alg_bytes = 24 * part_items + 24 * xf_items + 12 * N + 12 * N
        traffic = None
        if n[3] and int(tst["engine"]) == 3:  # team engine: one persistent launch runs many steps
            k_ms, k_n, kname = ms[3], n[3], "k_team (whole steps, team of workgroups on one XCD)"
            others = {"steps_per_launch": int(tst["steps"]) / max(n[3], 1),
                      "avg_step_us": ms[3] * 1e3 / max(int(tst["steps"]), 1)}
            traffic = traffic_for(wkey, int(k_n))
            if traffic is not None:
                others["traffic_source"] = os.path.relpath(TRAFFIC_FILE, os.path.dirname(os.path.abspath(__file__)))
        elif n[3]:  # micro engine: one persistent kernel does every phase
            k_ms, k_n, kname = ms[3], n[3], "k_micro (whole steps in one workgroup)"
            others = {}
        else:
            k_ms, k_n, kname = ms[2], n[2], "k_engine_kb (scatter+transform+source+sink of one step)"
            others = {"ka_avg_launch_us": ms[0] * 1e3 / max(n[0], 1), "kp_avg_launch_us": ms[1] * 1e3 / max(n[1], 1)}
        achieved = alg_bytes / (k_ms * 1e-3) / 1e9 if k_ms > 0 else 0.0
        result["roofline"] = {
            "bound": "hbm", "kernel": kname, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
            "algorithmic_bytes_per_launch": alg_bytes / max(k_n, 1), "avg_launch_us": k_ms * 1e3 / max(k_n, 1),
            "launches": k_n, **others,
        }
        log(f"[bench] roofline: {kname} avg {k_ms * 1e3 / max(k_n, 1):.2f} us over {k_n} launches, "
            f"achieved {achieved:.1f} GB/s ({achieved / HBM_PEAK_GBS:.2e} of peak), traffic/launch {traffic}")

    if args.phases:
        # per-phase device profile (slows the profiled member; not used for any reported rate)
        flow.configure(tuning_of(g, N, C, P, Q, fuse, args, tuned, use_graph=3))
        counts.zero_()
        pst = flow.run(src, N)
        phases, psteps = flow.phase_profile()
        log(f"[bench] phases (us/step over {psteps} steps): { {k: round(v, 3) for k, v in phases.items()} }")
        if args.members and int(pst["engine"]) == 3:
            for m, r in enumerate(flow.member_profile(args.team or 32)):
                log(f"[bench] member {m:2d}: wait {r[0]:.2f} start {r[1]:.2f} scatter {r[2]:.2f} "
                    f"sched {r[3]:.2f} us, segments {r[4]:.2f} sinks {r[5]:.2f} per step, wave work "
                    + " ".join(f"{x:.2f}" for x in r[6:]))

    forms = {}
    if not args.no_plain:
        Np = int(args.plain_items)
        pflow, psrc, pcounts = build_feedback(g, torch, Np, C, P, 1, g.SCHEDULE_HOST)
        dt, _ = run_once(pflow, psrc, pcounts, Np, torch)
        forms["plain_host_planned_Q1"] = {"value": Np / dt, "unit": "particles/s", "items": Np}
        log(f"[bench] plain form (host planned, Q=1) N={Np}: {Np / dt:.4g} particles/s")
    if args.photons > 0:
        forms.update(photon_forms(g, torch, int(args.photons), int(args.plain_photons), C, P, Q,
                                  verify=not args.no_cpu))
    if args.hash_items > 0:
        forms.update(hash_form(g, torch, int(args.hash_items), tuple(int(q) for q in args.hash_q.split(",")),
                               autotune=not args.no_autotune))
    if args.serve_jobs > 0:
        sat = served_form(g, torch, args.serve_jobs, args.serve_rate, args.serve_max_items)
        forms["served_mixed_stream"] = sat
        # the same stream offered at 70 % of the capacity the saturated run measured: latency
        # without a standing queue
        mean_items = sat["items"] / sat["jobs"]
        rate70 = 0.7 * sat["value"] / mean_items
        forms["served_mixed_stream_70pct_load"] = served_form(g, torch, args.serve_jobs, rate70, args.serve_max_items,
                                                             plain=False)
    result["forms"] = forms

    if not args.no_cpu:
        import oracle as O
        Nc = int(args.cpu_items)
        graph = flow.neutral_graph()
        t0 = time.perf_counter()
        r = O.run(graph, seed=42, items=Nc, capacity=C, live=P, batch=Q, mode=O.MODE_STEP,
                  hist_size=g.FEEDBACK_HIST_N)
        dt = time.perf_counter() - t0
        result["cpu_baseline"] = {
            "value": Nc / dt, "unit": "particles/s", "cores": 1, "kind": "port",
            "sample": f"oracle restatement of the reference CPU path, same network, N={Nc:.0e} "
                      f"C={C} P={P} Q={Q} seed 42, single thread, status={r['status']}",
        }
        log(f"[bench] cpu baseline N={Nc}: {Nc / dt:.4g} particles/s ({dt:.1f} s)")

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
