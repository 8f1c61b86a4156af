"""Pinning of the photon transport oracle (CPU only).

The FP64 photon histories (or_photon_scalar) are the reference the GPU tallies are compared with,
so they are pinned here by what the mathematics fixes independently of the restatement:
  * the Henyey-Greenstein sampler by its closed-form moments <ct> = g and <ct^2> = (1 + 2 g^2) / 3;
  * the direction update by |u'| = 1 and <u, u'> = ct;
  * flight + depth tally by the Beer-Lambert profile of a pure absorber (albedo 0), two mu_t;
  * Russian roulette by unbiasedness (the same mean absorbed weight with and without it);
and the fp32 flow restatement that mirrors the device step is held to the FP64 histories within
the per-bin tolerance the GPU tests use.
"""
import numpy as np
import pytest

import oracle as O

PHOTON = {"layers": 10, "tally_bins": 1024, "layer_mfp": 1.0, "mu_t": 1.0, "albedo": 0.9, "hg_g": 0.9,
          "roulette_w": 1e-4, "roulette_p": 0.1}


def photon_graph():
    """Generate -> Step -> Alive; Alive.alive -> Step, Alive.done -> Done (the step is node 1)."""
    return [dict(role=O.ROLE_SOURCE, op="SRC_PHOTON", bins=0, routes=[1]),
            dict(role=O.ROLE_TRANSFORM, op="XF_PHOTON_STEP", bins=1, routes=[2]),
            dict(role=O.ROLE_TRANSFORM, op="CLS_PHOTON_ALIVE", bins=2, routes=[1, 3]),
            dict(role=O.ROLE_SINK, op="SNK_PHOTON_TALLY", bins=0, routes=[])]


@pytest.mark.parametrize("g", [0.9, 0.5, 0.0, -0.3])
def test_hg_moments(g):
    n = 1 << 20
    xi = (np.arange(n) + 0.5) / n  # midpoint rule over the inverse-CDF variable
    ct = O.hg_cos(g, xi)
    assert abs(ct.mean() - g) < 1e-9
    assert abs((ct * ct).mean() - (1.0 + 2.0 * g * g) / 3.0) < 1e-9
    assert ct.min() >= -1.0 and ct.max() <= 1.0


def test_hg_rotation_keeps_unit_length_and_cosine():
    rng = np.random.default_rng(7)
    n = 4096
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    u[:4] = [[0, 0, 1], [0, 0, -1], [0, 0, 1], [0, 0, -1]]  # the lab-frame branch
    ct = rng.uniform(-1, 1, n)
    phi = rng.uniform(0, 2 * np.pi, n)
    out = O.hg_rotate(u, ct, phi)
    assert np.allclose(np.linalg.norm(out, axis=1), 1.0, atol=1e-12)
    assert np.allclose((u * out).sum(axis=1), ct, atol=1e-12)


@pytest.mark.parametrize("mu_t", [1.0, 2.0])
def test_pure_absorber_follows_beer_lambert(mu_t):
    N = 1_000_000
    cfg = dict(PHOTON, albedo=0.0, mu_t=mu_t)
    tally, steps = O.photon_scalar(42, N, cfg, threads=8)
    L = cfg["layers"] * cfg["layer_mfp"] / mu_t
    z = np.arange(cfg["tally_bins"] + 1) * L / cfg["tally_bins"]
    p = np.exp(-mu_t * z[:-1]) - np.exp(-mu_t * z[1:])  # absorbed at the first collision in bin b
    expected = N * p
    score = (tally - expected) / np.sqrt(N * p * (1 - p))
    assert steps == N  # one flight each: absorbed or transmitted
    assert (score * score).mean() < 1.3  # chi-square per degree of freedom
    assert np.abs(score).max() < 5.5
    total_p = 1 - np.exp(-mu_t * L)
    assert abs(tally.sum() - N * total_p) < 5 * np.sqrt(N * total_p * (1 - total_p))


def test_roulette_is_unbiased():
    N = 1_000_000
    base = dict(PHOTON, albedo=0.5)
    off, _ = O.photon_scalar(42, N, dict(base, roulette_w=0.0), threads=8)
    on, _ = O.photon_scalar(43, N, dict(base, roulette_w=0.3), threads=8)  # roulette at the 2nd collision
    # per-photon mean absorbed weight; the roulette estimate has a standard error of ~4e-4 here
    assert abs(on.sum() / N - off.sum() / N) < 2.5e-3


def test_fp32_flow_restatement_matches_fp64_histories():
    N = 100_000
    ref, _ = O.photon_scalar(42, N, PHOTON, threads=8)
    r = O.run(photon_graph(), seed=42, items=N, capacity=8192, live=16384, batch=8, mode=O.MODE_STEP,
              photon=PHOTON, fuse=True)
    assert r["status"] == O.OK
    t = r["tally"]
    # fp32 histories on the same Philox words diverge from the FP64 ones only where rounding moves
    # a photon across a boundary or a roulette threshold: measured 1.4e-6 on the total and 2.4e-3
    # on the worst of the 1024 bins at N=1e5
    assert abs(t.sum() - ref.sum()) <= 1e-5 * ref.sum()
    assert (np.abs(t - ref) <= 5e-3 * ref).all()
