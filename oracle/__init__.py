"""ctypes binding of the CPU oracle (oracle/liboracle.so).

Test infrastructure only: imported by tests/, by ``__graft_entry__.smoke()`` and by the CPU
baseline leg of ``bench.py``. The product package ``graphflow_amd`` never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

OK, INVALID, OVERFLOW, STALL, POOL_EXHAUSTED, BAD_ARGUMENT, PLAN_MISMATCH = range(7)
ROLE_SOURCE, ROLE_TRANSFORM, ROLE_SINK = 0, 1, 2
MODE_RUNNER, MODE_STEP = 0, 1

OPS = {
    "SRC_PARTICLE_UNIFORM": 1, "SRC_KEYED_ZIPF": 2, "SRC_PHOTON": 3,
    "CLS_WEIGHT_ROUTE": 16, "CLS_KEY_BUCKET": 17, "CLS_ID_MOD": 18, "CLS_LUT": 19,
    "CLS_PHOTON_ALIVE": 20,
    "XF_SCALE_WEIGHT": 32, "XF_PHOTON_STEP": 33,
    "SNK_HIST": 48, "SNK_RECORD_WEIGHT": 49, "SNK_CHUNK_DIGEST": 50, "SNK_PHOTON_TALLY": 51,
}
UPDS = {"NONE": 0, "ADD_ON_BIN": 1, "GEN_TO_BIN": 2}


class NodeDesc(C.Structure):
    _fields_ = [("role", C.c_uint32), ("op", C.c_uint32), ("upd", C.c_uint32), ("bins", C.c_uint32),
                ("route_offset", C.c_uint32), ("u", C.c_uint32 * 4), ("f", C.c_float * 4)]


class RunCfg(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("items", C.c_uint64), ("capacity", C.c_uint32),
                ("live", C.c_uint32), ("batch", C.c_uint32), ("num_chunks", C.c_uint32),
                ("mode", C.c_uint32), ("threads", C.c_uint32), ("fuse", C.c_uint32), ("layers", C.c_uint32),
                ("tally_bins", C.c_uint32), ("layer_mfp", C.c_double), ("mu_t", C.c_double),
                ("albedo", C.c_double), ("hg_g", C.c_double), ("roulette_w", C.c_double),
                ("roulette_p", C.c_double), ("inflight", C.c_uint32), ("bin_path", C.c_uint32),
                ("hash_tile", C.c_uint32), ("hash_entries", C.c_uint32), ("fused_stages", C.c_uint32),
                ("source_chunks", C.c_uint32)]


class RunOut(C.Structure):
    _fields_ = [("status", C.c_uint32), ("message", C.c_char * 160), ("pool_chunks", C.c_uint32),
                ("peak_live_chunks", C.c_uint32), ("batches", C.c_uint64), ("steps", C.c_uint64),
                ("generated", C.c_uint64), ("sunk", C.c_uint64),
                ("trace", C.POINTER(C.c_uint32)), ("trace_cap", C.c_uint64), ("trace_len", C.c_uint64),
                ("hist", C.POINTER(C.c_uint64)), ("record", C.POINTER(C.c_uint32)),
                ("digest_log", C.POINTER(C.c_uint32)), ("digest_cap", C.c_uint64),
                ("digest_len", C.c_uint64), ("tally", C.POINTER(C.c_double)),
                ("bin_multiset", C.POINTER(C.c_uint64)), ("bin_multiset_bins", C.c_uint32),
                ("bin_items", C.POINTER(C.c_uint64))]


_lib = None


def build() -> None:
    """Compile liboracle.so from oracle/gf_oracle.c (gcc)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        u32p, u64p = C.POINTER(C.c_uint32), C.POINTER(C.c_uint64)
        _lib.or_philox4x32_10.argtypes = [u32p, u32p, u32p]
        _lib.or_mix64.restype = C.c_uint64
        _lib.or_mix64.argtypes = [C.c_uint64]
        _lib.or_stream_key.restype = C.c_uint64
        _lib.or_stream_key.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint32]
        _lib.or_stream_words.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint32, u32p]
        _lib.or_fmix32.restype = C.c_uint32
        _lib.or_fmix32.argtypes = [C.c_uint32]
        _lib.or_uniform_f32.restype = C.c_float
        _lib.or_uniform_f32.argtypes = [C.c_uint32]
        _lib.or_uniform_f64.restype = C.c_double
        _lib.or_uniform_f64.argtypes = [C.c_uint32, C.c_uint32]
        _lib.or_output_chunk_bound.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, u64p]
        _lib.or_ceil_div.argtypes = [C.c_uint64, C.c_uint64, u64p]
        _lib.or_plan_dense_bins.argtypes = [u32p, C.c_uint32, C.c_uint32, u32p, u32p, C.c_uint32, u32p]
        _lib.or_build_dense_table.argtypes = [u32p, C.c_uint32, u32p, C.c_uint32, u32p]
        _lib.or_tile_count.argtypes = [C.c_uint32, C.c_uint32, u32p]
        _lib.or_flatten_tiles.argtypes = [u32p, u32p, C.c_uint32, C.c_uint32, C.c_uint32, u32p, u32p]
        _lib.or_num_frames_for.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                           C.c_uint32, u32p]
        _lib.or_next_pow2.restype = C.c_uint32
        _lib.or_next_pow2.argtypes = [C.c_uint32]
        _lib.or_hash_bin_slot.restype = C.c_uint32
        _lib.or_hash_bin_slot.argtypes = [C.c_uint32, C.c_uint32]
        _lib.or_hash_table_insert.argtypes = [u32p, u32p, C.c_uint32, C.c_uint32]
        _lib.or_zipf_table.argtypes = [C.c_uint32, C.c_double, C.POINTER(C.c_double)]
        _lib.or_zipf_sample.restype = C.c_uint32
        _lib.or_zipf_sample.argtypes = [C.POINTER(C.c_double), C.c_uint32, C.c_uint32, C.c_uint32]
        _lib.or_derive_pool.argtypes = [C.POINTER(NodeDesc), C.c_uint32, C.c_uint64, C.c_uint32,
                                        C.c_uint32, u32p, u32p]
        _lib.or_run.argtypes = [C.POINTER(NodeDesc), C.c_uint32, C.POINTER(C.c_int32), u32p,
                                C.POINTER(RunCfg), C.POINTER(RunOut)]
        _lib.or_partition_items.argtypes = [C.POINTER(NodeDesc), u32p, C.c_uint64, C.c_uint32, C.c_uint32,
                                            C.POINTER(C.c_float), u32p, u32p, u32p, u32p, u32p, u32p]
        _lib.or_feedback_scalar.argtypes = [C.c_uint64, C.c_uint64, C.c_float, C.c_uint32, C.c_uint32,
                                            C.POINTER(C.c_float), u32p, u32p, C.c_uint32]
        _lib.or_keyed_scalar.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32, C.c_double, C.c_uint32, u64p, u64p,
                                         C.c_uint32]
        _lib.or_photon_scalar.argtypes = [C.c_uint64, C.c_uint64, C.POINTER(RunCfg),
                                          C.POINTER(C.c_double), u64p, C.c_uint32]
        f64p = C.POINTER(C.c_double)
        _lib.or_hg_cos_f64.argtypes = [C.c_double, f64p, f64p, C.c_uint64]
        _lib.or_hg_rotate_f64.argtypes = [f64p, f64p, f64p, f64p, C.c_uint64]
    return _lib


def _p(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype)) if a is not None else None


def philox(ctr, key):
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    o = np.zeros(4, np.uint32)
    lib().or_philox4x32_10(_p(c, C.c_uint32), _p(k, C.c_uint32), _p(o, C.c_uint32))
    return [int(v) for v in o]


def stream_words(seed, id_, event, purpose):
    o = np.zeros(4, np.uint32)
    lib().or_stream_words(seed, id_, event, purpose, _p(o, C.c_uint32))
    return [int(v) for v in o]


def output_chunk_bound(n, c, b):
    o = C.c_uint64(0)
    rc = lib().or_output_chunk_bound(n, c, b, C.byref(o))
    return rc, o.value


def plan_dense_bins(totals, capacity):
    t = np.asarray(totals, dtype=np.uint32)
    base = np.zeros(max(len(t), 1), np.uint32)
    cap = int(sum(-(-int(v) // capacity) for v in t)) + 1 if capacity else 1
    fills = np.zeros(cap, np.uint32)
    r = C.c_uint32(0)
    rc = lib().or_plan_dense_bins(_p(t, C.c_uint32), len(t), capacity, _p(base, C.c_uint32),
                                  _p(fills, C.c_uint32), cap, C.byref(r))
    return rc, base[:len(t)].tolist(), fills[:r.value].tolist(), r.value


def flatten_tiles(chunks, fills, tile):
    ci = np.asarray(chunks, np.uint32)
    fi = np.asarray(fills, np.uint32)
    cap = sum(-(-int(f) // tile) for f in fills) + 1 if tile else 1
    out = np.zeros(3 * cap, np.uint32)
    n = C.c_uint32(0)
    rc = lib().or_flatten_tiles(_p(ci, C.c_uint32), _p(fi, C.c_uint32), len(ci), tile, cap,
                                _p(out, C.c_uint32), C.byref(n))
    return rc, out[:3 * n.value].reshape(-1, 3).tolist()


def num_frames_for(work, tile, frame_cap, max_threads, max_blocks, sms):
    o = C.c_uint32(0)
    rc = lib().or_num_frames_for(work, tile, frame_cap, max_threads, max_blocks, sms, C.byref(o))
    return rc, o.value


def make_nodes(graph):
    """graph: list of dicts {role, op, upd?, bins, routes, u?, f?} -> (NodeDesc[], routes int32[])."""
    n = len(graph)
    arr = (NodeDesc * n)()
    routes = []
    for k, g in enumerate(graph):
        d = arr[k]
        d.role = g["role"]
        d.op = OPS[g["op"]] if isinstance(g["op"], str) else g["op"]
        upd = g.get("upd", 0)
        d.upd = UPDS[upd] if isinstance(upd, str) else upd
        d.bins = g["bins"]
        d.route_offset = len(routes)
        routes.extend(g.get("routes", []))
        for i, v in enumerate(g.get("u", [])):
            d.u[i] = v
        for i, v in enumerate(g.get("f", [])):
            d.f[i] = v
    r = np.asarray(routes if routes else [0], dtype=np.int32)
    return arr, r


def run(graph, *, seed, items, capacity, live=0, batch=1, num_chunks=0, mode=MODE_RUNNER, threads=1, fuse=False,
        trace_cap=0, hist_size=0, record=False, digest_cap=0, multiset_bins=0, lut=None, photon=None,
        inflight=1, bin_path=0, hash_tile=0, hash_entries=0, fused_stages=False, source_chunks=1):
    nodes, routes = make_nodes(graph)
    cfg = RunCfg(seed=seed, items=items, capacity=capacity, live=live, batch=batch,
                 num_chunks=num_chunks, mode=mode, threads=threads, fuse=1 if fuse else 0,
                 inflight=inflight, bin_path=bin_path, hash_tile=hash_tile, hash_entries=hash_entries,
                 fused_stages=1 if fused_stages else 0, source_chunks=source_chunks)
    if photon:
        for k, v in photon.items():
            setattr(cfg, k, v)
    out = RunOut()
    trace = np.zeros(max(trace_cap, 1), np.uint32)
    hist = np.zeros(max(hist_size, 1), np.uint64)
    rec = np.full(max(items, 1) if record else 1, 0xFFFFFFFF, np.uint32)
    dlog = np.zeros(4 * max(digest_cap, 1), np.uint32)
    mset = np.zeros(max(multiset_bins, 1), np.uint64)
    mitems = np.zeros(max(multiset_bins, 1), np.uint64)
    tally = np.zeros(max(photon["tally_bins"], 1) if photon else 1, np.float64)
    out.trace = _p(trace, C.c_uint32) if trace_cap else None
    out.trace_cap = trace_cap
    out.hist = _p(hist, C.c_uint64) if hist_size else None
    out.record = _p(rec, C.c_uint32) if record else None
    out.digest_log = _p(dlog, C.c_uint32) if digest_cap else None
    out.digest_cap = digest_cap
    out.bin_multiset = _p(mset, C.c_uint64) if multiset_bins else None
    out.bin_items = _p(mitems, C.c_uint64) if multiset_bins else None
    out.bin_multiset_bins = multiset_bins
    out.tally = _p(tally, C.c_double) if photon else None
    lut_arr = np.asarray(lut if lut is not None else [0], np.uint32)
    rc = lib().or_run(nodes, len(graph), _p(routes, C.c_int32), _p(lut_arr, C.c_uint32), C.byref(cfg),
                      C.byref(out))
    res = {
        "status": rc, "message": out.message.decode(), "pool_chunks": out.pool_chunks,
        "peak_live_chunks": out.peak_live_chunks, "batches": out.batches, "steps": out.steps,
        "generated": out.generated, "sunk": out.sunk,
    }
    if trace_cap:
        res["trace"] = trace[:min(out.trace_len, trace_cap)].copy()
        res["trace_len"] = out.trace_len
    if hist_size:
        res["hist"] = hist
    if record:
        res["record"] = rec
    if digest_cap:
        res["digest"] = dlog[:4 * min(out.digest_len, digest_cap)].reshape(-1, 4).copy()
        res["digest_len"] = out.digest_len
    if multiset_bins:
        res["multiset"] = mset
        res["multiset_items"] = mitems
    if photon:
        res["tally"] = tally
    return res


def feedback_scalar(seed, items, threshold=1e-3, distribution=0, ndest=2, multipliers=(0.5, 0.75), threads=1):
    m = np.asarray(multipliers, np.float32)
    w = np.zeros(items, np.uint32)
    g = np.zeros(items, np.uint32)
    lib().or_feedback_scalar(seed, items, threshold, distribution, ndest, _p(m, C.c_float),
                             _p(w, C.c_uint32), _p(g, C.c_uint32), threads)
    return w, g


def keyed_scalar(seed, items, values=1 << 20, s=1.1, bins=4096, threads=1):
    ms = np.zeros(bins, np.uint64)
    ct = np.zeros(bins, np.uint64)
    # the source node carries the exponent as fp32 (gf_node_desc.f), so the table uses that value
    rc = lib().or_keyed_scalar(seed, items, values, float(np.float32(s)), bins, _p(ms, C.c_uint64),
                               _p(ct, C.c_uint64), threads)
    assert rc == OK
    return ms, ct


def photon_scalar(seed, photons, photon_cfg, threads=1):
    cfg = RunCfg()
    for k, v in photon_cfg.items():
        setattr(cfg, k, v)
    tally = np.zeros(photon_cfg["tally_bins"], np.float64)
    steps = C.c_uint64(0)
    lib().or_photon_scalar(seed, photons, C.byref(cfg), _p(tally, C.c_double), C.byref(steps), threads)
    return tally, steps.value


def hg_cos(g, xi):
    xi = np.ascontiguousarray(xi, np.float64)
    out = np.zeros_like(xi)
    lib().or_hg_cos_f64(C.c_double(g), _p(xi, C.c_double), _p(out, C.c_double), len(xi))
    return out


def hg_rotate(u, ct, phi):
    u = np.ascontiguousarray(u, np.float64).reshape(-1, 3)
    ct = np.ascontiguousarray(ct, np.float64)
    phi = np.ascontiguousarray(phi, np.float64)
    out = np.zeros_like(u)
    lib().or_hg_rotate_f64(_p(u, C.c_double), _p(ct, C.c_double), _p(phi, C.c_double), _p(out, C.c_double), len(ct))
    return out


def zipf_table(values, s):
    t = np.zeros(values, np.float64)
    lib().or_zipf_table(values, s, _p(t, C.c_double))
    return t


def partition_items(node, n, weight=None, ids=None, gens=None, keys=None, lut=None):
    nodes, _ = make_nodes([node])
    bins = np.zeros(max(n, 1), np.uint32)
    tot = np.zeros(node["bins"], np.uint32)
    rank = np.zeros(max(n, 1), np.uint32)
    w = np.asarray(weight, np.float32) if weight is not None else None
    i = np.asarray(ids, np.uint32) if ids is not None else None
    g = np.asarray(gens, np.uint32) if gens is not None else None
    k = np.asarray(keys, np.uint32) if keys is not None else None
    lu = np.asarray(lut if lut is not None else [0], np.uint32)
    rc = lib().or_partition_items(nodes, _p(lu, C.c_uint32), 0, 0, n, _p(w, C.c_float), _p(i, C.c_uint32),
                                  _p(g, C.c_uint32), _p(k, C.c_uint32), _p(bins, C.c_uint32),
                                  _p(tot, C.c_uint32), _p(rank, C.c_uint32))
    return rc, bins[:n], tot, rank[:n]


def hist_rows(hist, lo):
    """(value, count) rows ascending, from a dense count array whose index 0 is value lo."""
    nz = np.nonzero(hist)[0]
    return np.stack([nz.astype(np.uint64) + np.uint64(lo), hist[nz].astype(np.uint64)], axis=1)
