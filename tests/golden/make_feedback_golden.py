"""Writes feedback_N1e8.json: the sorted weight histogram of the fp32 feedback network at
N=1e8, seed 42, for the four routing distributions, from the oracle's schedule-free per-particle
model (or_feedback_scalar). The histogram of a run does not depend on its schedule, so every
engine, C, P and Q must reproduce these rows.

Each entry: rows (distinct weight bit patterns), sum of the counts (= N), SHA-256 of the
'weight_bits,count\\n' rows in ascending order, and the mean number of A/B passes per particle.

    python tests/golden/make_feedback_golden.py [threads]
"""
import hashlib
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import numpy as np  # noqa: E402

import oracle as O  # noqa: E402

N, SEED = 100_000_000, 42
# (name, distribution code of GF_ROUTE_*, destinations)
CASES = [("parity", 0, 2), ("skew_90_10", 1, 2), ("all_one", 2, 2), ("uniform", 3, 16)]


def rows_sha(w):
    v, c = np.unique(w, return_counts=True)
    text = "".join(f"{a},{b}\n" for a, b in zip(v.tolist(), c.tolist())).encode()
    return len(v), int(c.sum()), hashlib.sha256(text).hexdigest()


def main():
    threads = int(sys.argv[1]) if len(sys.argv) > 1 else (os.cpu_count() or 1)
    out = {"items": N, "seed": SEED, "threshold": 1e-3,
           "multipliers": "0.5 for even destinations, 0.75 for odd", "cases": {}}
    for name, dist, nd in CASES:
        t0 = time.time()
        mult = [0.5 if d % 2 == 0 else 0.75 for d in range(nd)]
        w, gens = O.feedback_scalar(SEED, N, 1e-3, dist, nd, mult, threads=threads)
        rows, total, sha = rows_sha(w)
        out["cases"][name] = {"distribution": dist, "destinations": nd, "rows": rows, "sum": total,
                              "sha256": sha, "mean_passes": float(gens.mean())}
        print(name, out["cases"][name], f"{time.time() - t0:.1f} s", flush=True)
        del w, gens
    with open(os.path.join(HERE, "feedback_N1e8.json"), "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main()
