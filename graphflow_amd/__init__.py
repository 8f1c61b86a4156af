"""graphflow_amd: MI355X-native transform-batch path of graphFlow, Python host binding.

The compute path is libgraphflow_amd.so (hand-written HIP for gfx950) behind the C ABI in
include/graphflow.h. This module binds that ABI with ctypes and mirrors the reference's typed
registration API (include/graphFlow/flow/Flow.hpp:85-412): ``Flow.source / partition /
transform / sink / connect / configure / run``, with the reference's error classes
(``BatchError``, ``SchedulerStallError``, ``ValueError`` for std::invalid_argument).

There is no CPU fallback: every entry point raises ``ExtensionMissing`` when the library is not
built, and ``Flow.run`` needs a visible GPU.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

HERE = os.path.dirname(os.path.abspath(__file__))
# GRAPHFLOW_AMD_LIB selects another build of the same library (for side-by-side timing)
LIB_PATH = os.environ.get("GRAPHFLOW_AMD_LIB") or os.path.join(HERE, "libgraphflow_amd.so")

# ---- status (graphflow.h gf_status) -----------------------------------------------------------
OK, INVALID, OVERFLOW, STALL, POOL_EXHAUSTED, BAD_ARGUMENT, PLAN_MISMATCH, DEVICE_ERROR, STAGE_FAILED = range(9)
PAYLOAD_PARTICLE, PAYLOAD_KEYED, PAYLOAD_PHOTON = 0, 1, 2
ROLE_SOURCE, ROLE_TRANSFORM, ROLE_SINK = 0, 1, 2
SCHEDULE_HOST, SCHEDULE_DEVICE = 0, 1
BIN_PATH_AUTO, BIN_PATH_DENSE, BIN_PATH_HASH = 0, 1, 2

OPS = {
    "SRC_PARTICLE_UNIFORM": 1, "SRC_KEYED_ZIPF": 2, "SRC_PHOTON": 3,
    "CLS_WEIGHT_ROUTE": 16, "CLS_KEY_BUCKET": 17, "CLS_ID_MOD": 18, "CLS_LUT": 19, "CLS_PHOTON_ALIVE": 20,
    "XF_SCALE_WEIGHT": 32, "XF_PHOTON_STEP": 33,
    "SNK_HIST": 48, "SNK_RECORD_WEIGHT": 49, "SNK_CHUNK_DIGEST": 50, "SNK_PHOTON_TALLY": 51,
}
UPDS = {"NONE": 0, "ADD_ON_BIN": 1, "GEN_TO_BIN": 2}
ROUTES = {"parity": 0, "skew_90_10": 1, "all_one": 2, "uniform": 3}


class ExtensionMissing(RuntimeError):
    """libgraphflow_amd.so is not built; there is no fallback path."""


class GraphFlowError(RuntimeError):
    pass


class BatchError(GraphFlowError):
    """Device status or plan mismatch of one batch (sched/GraphRunner.hpp:62-66)."""


class SchedulerStallError(GraphFlowError):
    """No progress possible (sched/GraphRunner.hpp:355-378)."""


class PoolExhaustedError(GraphFlowError):
    pass


class DeviceError(GraphFlowError):
    pass


class StageFailedError(GraphFlowError):
    """A stage is single-use after a failed batch."""


class ChunkHandle(C.Structure):
    _fields_ = [("index", C.c_uint32), ("generation", C.c_uint32)]


class ChunkDesc(C.Structure):
    _fields_ = [("index", C.c_uint32), ("generation", C.c_uint32), ("fill", C.c_uint32), ("bin", C.c_uint32)]

    def tuple(self):
        return (self.index, self.generation, self.fill, self.bin)


class TileDesc(C.Structure):
    _fields_ = [("chunk_index", C.c_uint32), ("tile_offset", C.c_uint32), ("count", C.c_uint32)]


class NodeDesc(C.Structure):
    _fields_ = [("role", C.c_uint32), ("op", C.c_uint32), ("upd", C.c_uint32), ("bins", C.c_uint32),
                ("route_offset", C.c_uint32), ("u", C.c_uint32 * 4), ("f", C.c_float * 4),
                ("ptr", C.c_uint64 * 2), ("size", C.c_uint64 * 2)]


class TuningC(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("batch_chunks", C.c_uint32), ("chunk_capacity", C.c_uint32),
                ("live_particles", C.c_uint32), ("num_chunks", C.c_uint32), ("tile_items", C.c_uint32),
                ("schedule", C.c_uint32), ("steps_per_launch", C.c_uint32), ("trace", C.c_uint32),
                ("use_graph", C.c_uint32), ("fuse_transforms", C.c_uint32), ("max_workgroups", C.c_uint32),
                ("device_engine", C.c_uint32), ("photon_layers", C.c_uint32), ("photon_tally_bins", C.c_uint32),
                ("photon_layer_mfp", C.c_double), ("photon_mu_t", C.c_double), ("photon_albedo", C.c_double),
                ("photon_hg_g", C.c_double), ("photon_roulette_w", C.c_double), ("photon_roulette_p", C.c_double),
                ("team_xcd", C.c_uint32), ("inflight_batches", C.c_uint32), ("bin_path", C.c_uint32),
                ("hash_table_entries", C.c_uint32), ("stall_timeout_ms", C.c_uint32), ("fused_batches", C.c_uint32),
                ("source_chunks", C.c_uint32)]


class Stats(C.Structure):
    _fields_ = [("completed_batches", C.c_uint64), ("steps", C.c_uint64), ("generated", C.c_uint64),
                ("sunk", C.c_uint64), ("pool_chunks", C.c_uint32), ("peak_live_chunks", C.c_uint32),
                ("chunk_capacity", C.c_uint32), ("live_particles", C.c_uint32), ("kernel_launches", C.c_uint64),
                ("wall_ms", C.c_double), ("engine", C.c_uint32), ("handoff_step", C.c_uint32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


# Every entry point of include/graphflow.h with its ctypes signature.
_V, _P = C.c_void_p, C.POINTER
SIGNATURES = {
    "gf_abi_version": (C.c_uint32, []),
    "gf_status_name": (C.c_char_p, [C.c_int]),
    "gf_context_create": (C.c_int, [C.c_int, _P(_V)]),
    "gf_context_destroy": (C.c_int, [_V]),
    "gf_context_last_error": (C.c_int, [_V, C.c_char_p, C.c_uint32]),
    "gf_output_chunk_bound": (C.c_int, [C.c_uint64, C.c_uint64, C.c_uint64, _P(C.c_uint64)]),
    "gf_plan_dense_bins": (C.c_int, [_P(C.c_uint32), C.c_uint32, C.c_uint32, _P(C.c_uint32), _P(C.c_uint32),
                                     C.c_uint32, _P(C.c_uint32)]),
    "gf_pool_create": (C.c_int, [_V, C.c_uint32, C.c_uint32, C.c_uint32, _P(_V)]),
    "gf_pool_destroy": (C.c_int, [_V]),
    "gf_pool_acquire": (C.c_int, [_V, _P(ChunkHandle)]),
    "gf_pool_release": (C.c_int, [_V, ChunkHandle]),
    "gf_pool_acquire_lease": (C.c_int, [_V, ChunkHandle]),
    "gf_pool_release_lease": (C.c_int, [_V, ChunkHandle]),
    "gf_pool_counts": (C.c_int, [_V, _P(C.c_uint32)]),
    "gf_pool_base": (C.c_int, [_V, ChunkHandle, _P(C.c_uint64)]),
    "gf_pool_column": (C.c_int, [_V, C.c_uint32, _P(_V), _P(C.c_uint32)]),
    "gf_pool_write": (C.c_int, [_V, C.c_uint32, C.c_uint64, _V, C.c_uint64]),
    "gf_pool_read": (C.c_int, [_V, C.c_uint32, C.c_uint64, _V, C.c_uint64]),
    "gf_stage_create": (C.c_int, [_V, _P(NodeDesc), C.c_uint32, C.c_uint64, C.c_uint32, C.c_uint32, _P(_V)]),
    "gf_stage_create_tuned": (C.c_int, [_V, _P(NodeDesc), C.c_uint32, _P(TuningC), _P(_V)]),
    "gf_stage_last_error": (C.c_int, [_V, C.c_char_p, C.c_uint32]),
    "gf_transform_batch": (C.c_int, [_V, _P(ChunkDesc), C.c_uint32, C.c_uint64, _V]),
    "gf_transform_batch_noalloc": (C.c_int, [_V, _P(NodeDesc), C.c_uint32, _P(TuningC), _V, C.c_uint32, _V,
                                             C.c_uint32, _V, _V, _V]),
    "gf_pool_lease_count": (C.c_int, [_V, ChunkHandle, _P(C.c_uint32)]),
    "gf_pool_is_live": (C.c_int, [_V, ChunkHandle, _P(C.c_int)]),
    "gf_pool_generation": (C.c_int, [_V, C.c_uint32, _P(C.c_uint32)]),
    "gf_pool_validate_live": (C.c_int, [_V, ChunkHandle]),
    "gf_stage_destroy": (C.c_int, [_V]),
    "gf_stage_launch": (C.c_int, [_V, _P(ChunkDesc), C.c_uint32, C.c_uint64, _V]),
    "gf_stage_advance": (C.c_int, [_V, _V, _P(C.c_uint32), _P(C.c_int)]),
    "gf_stage_finish": (C.c_int, [_V, _P(ChunkDesc), C.c_uint32, _P(C.c_uint32)]),
    "gf_stage_abort": (C.c_int, [_V]),
    "gf_stage_counts": (C.c_int, [_V, _P(C.c_uint32), C.c_uint32]),
    "gf_stage_tiles": (C.c_int, [_V, _P(TileDesc), C.c_uint32, _P(C.c_uint32)]),
    "gf_flow_create": (C.c_int, [_V, C.c_uint32, _P(_V)]),
    "gf_flow_destroy": (C.c_int, [_V]),
    "gf_flow_add_node": (C.c_int, [_V, C.c_char_p, _P(NodeDesc), _P(C.c_uint32)]),
    "gf_flow_connect": (C.c_int, [_V, C.c_uint32, C.c_uint32, C.c_uint32]),
    "gf_flow_configure": (C.c_int, [_V, _P(TuningC)]),
    "gf_flow_run": (C.c_int, [_V, C.c_uint32, C.c_uint64, _P(Stats)]),
    "gf_flow_last_error": (C.c_int, [_V, C.c_char_p, C.c_uint32]),
    "gf_flow_trace": (C.c_int, [_V, _P(C.c_uint32), C.c_uint64, _P(C.c_uint64)]),
    "gf_flow_kernel_times": (C.c_int, [_V, _P(C.c_double), _P(C.c_uint64), C.c_uint32]),
    "gf_flow_item_passes": (C.c_int, [_V, _P(C.c_uint64), _P(C.c_uint64)]),
    "gf_flow_phase_profile": (C.c_int, [_V, _P(C.c_uint64), C.c_uint32]),
    "gf_hist_rows": (C.c_int, [_V, _V, C.c_uint64, C.c_uint32, _P(C.c_uint64), C.c_uint64, _P(C.c_uint64)]),
}

class ServerConfigC(C.Structure):
    _fields_ = [("slots", C.c_uint32), ("team_size", C.c_uint32), ("memory_budget", C.c_uint64),
                ("max_items", C.c_uint64), ("prewarm", C.c_uint32), ("reserved", C.c_uint32),
                ("topology_bytes", C.c_uint64 * 3)]


class JobDescC(C.Structure):
    _fields_ = [("topology", C.c_uint32), ("reserved", C.c_uint32), ("items", C.c_uint64), ("seed", C.c_uint64)]


class JobResultC(C.Structure):
    _fields_ = [("job", C.c_uint64), ("topology", C.c_uint32), ("slot", C.c_uint32), ("status", C.c_uint32),
                ("engine", C.c_uint32), ("items", C.c_uint64), ("steps", C.c_uint64), ("submit_ms", C.c_double),
                ("start_ms", C.c_double), ("end_ms", C.c_double), ("digest", C.c_uint64), ("tally_sum", C.c_double)]


SIGNATURES.update({
    "gf_flow_autotune": (C.c_int, [_V, C.c_uint32, C.c_uint64, _P(TuningC), _P(C.c_double)]),
    "gf_server_create": (C.c_int, [_V, _P(ServerConfigC), _P(_V)]),
    "gf_server_submit": (C.c_int, [_V, _P(JobDescC), _P(C.c_uint64)]),
    "gf_server_poll": (C.c_int, [_V, _P(JobResultC), C.c_uint32, _P(C.c_uint32), C.c_uint32]),
    "gf_server_job_tally": (C.c_int, [_V, C.c_uint64, _P(C.c_double), C.c_uint32]),
    "gf_server_last_error": (C.c_int, [_V, C.c_char_p, C.c_uint32]),
    "gf_server_destroy": (C.c_int, [_V]),
})

_lib = None


def lib():
    """Load libgraphflow_amd.so (raises ExtensionMissing when it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ExtensionMissing(f"{LIB_PATH} is not built; run graphflow_amd/build.py (no CPU fallback exists)")
        h = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _lib = h
    return _lib


def _raise(code: int, message: str = ""):
    if code == OK:
        return
    name = lib().gf_status_name(code).decode()
    msg = f"{message} [{name}]" if message else name
    if code in (INVALID, OVERFLOW, PLAN_MISMATCH):
        raise BatchError(msg)
    if code == STALL:
        raise SchedulerStallError(msg)
    if code == POOL_EXHAUSTED:
        raise PoolExhaustedError(msg)
    if code == BAD_ARGUMENT:
        raise ValueError(msg)
    if code == STAGE_FAILED:
        raise StageFailedError(msg)
    raise DeviceError(msg)


def output_chunk_bound(n: int, capacity: int, max_bins: int) -> int:
    o = C.c_uint64(0)
    _raise(lib().gf_output_chunk_bound(n, capacity, max_bins, C.byref(o)))
    return o.value


# ---- device objects -----------------------------------------------------------------------------
class Context:
    def __init__(self, device: int = 0):
        h = _V()
        _raise(lib().gf_context_create(device, C.byref(h)))
        self.handle = h

    def close(self):
        if self.handle:
            lib().gf_context_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx: Optional[Context] = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


class Pool:
    """ChunkStorage contract (core/ChunkStorage.hpp:35-54) over HBM-resident SoA columns."""

    def __init__(self, payload: int, num_chunks: int, capacity: int, ctx: Optional[Context] = None):
        self.ctx = ctx or default_context()
        h = _V()
        _raise(lib().gf_pool_create(self.ctx.handle, payload, num_chunks, capacity, C.byref(h)))
        self.handle = h
        self.payload, self.num_chunks, self.capacity = payload, num_chunks, capacity

    def acquire(self) -> ChunkHandle:
        out = ChunkHandle()
        _raise(lib().gf_pool_acquire(self.handle, C.byref(out)))
        return out

    def release(self, h: ChunkHandle):
        _raise(lib().gf_pool_release(self.handle, h))

    def acquire_lease(self, h):
        _raise(lib().gf_pool_acquire_lease(self.handle, h))

    def release_lease(self, h):
        _raise(lib().gf_pool_release_lease(self.handle, h))

    def counts(self) -> dict:
        a = (C.c_uint32 * 6)()
        _raise(lib().gf_pool_counts(self.handle, a))
        return dict(zip(["num_chunks", "capacity", "free", "live", "peak_live", "peak_leases"], list(a)))

    def lease_count(self, h) -> int:
        o = C.c_uint32()
        _raise(lib().gf_pool_lease_count(self.handle, h, C.byref(o)))
        return o.value

    def is_live(self, h) -> bool:
        o = C.c_int()
        _raise(lib().gf_pool_is_live(self.handle, h, C.byref(o)))
        return bool(o.value)

    def generation(self, index: int) -> int:
        o = C.c_uint32()
        _raise(lib().gf_pool_generation(self.handle, index, C.byref(o)))
        return o.value

    def validate_live(self, h):
        _raise(lib().gf_pool_validate_live(self.handle, h))

    def base(self, h) -> int:
        o = C.c_uint64()
        _raise(lib().gf_pool_base(self.handle, h, C.byref(o)))
        return o.value

    def write(self, column: int, first_slot: int, array):
        _raise(lib().gf_pool_write(self.handle, column, first_slot, array.ctypes.data, array.size))

    def read(self, column: int, first_slot: int, array):
        _raise(lib().gf_pool_read(self.handle, column, first_slot, array.ctypes.data, array.size))

    def close(self):
        if self.handle:
            lib().gf_pool_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Stage:
    """One node's batch protocol: launch / advance / finish / abort (flow/Execution.hpp:1537-1561)."""

    def __init__(self, pool: Pool, node: NodeDesc, node_key: int = 0, seed: int = 0, batch_chunks: int = 1,
                 tile_items: int = 0, **tuning):
        """tuning: further Tuning fields (bin_path, hash_table_entries, stall_timeout_ms, photon)."""
        self.pool = pool
        self.node = node
        h = _V()
        if tuning:
            t = Tuning(seed=seed, batch_chunks=batch_chunks, tile_items=tile_items or None, **tuning).c()
            _raise(lib().gf_stage_create_tuned(pool.handle, C.byref(node), node_key, C.byref(t), C.byref(h)))
        else:
            _raise(lib().gf_stage_create(pool.handle, C.byref(node), node_key, seed, batch_chunks, tile_items,
                                         C.byref(h)))
        self.handle = h

    def _check(self, code: int):
        if code != OK:
            buf = C.create_string_buffer(512)
            lib().gf_stage_last_error(self.handle, buf, 512)
            _raise(code, buf.value.decode())

    def launch(self, chunks: Sequence[tuple], first_id: int = 0):
        arr = (ChunkDesc * len(chunks))(*[ChunkDesc(*c) for c in chunks])
        self._check(lib().gf_stage_launch(self.handle, arr, len(chunks), first_id, None))

    def transform_batch(self, chunks: Sequence[tuple], first_id: int = 0):
        """One launch for the whole batch (gf_transform_batch); then one advance() and finish()."""
        arr = (ChunkDesc * len(chunks))(*[ChunkDesc(*c) for c in chunks])
        self._check(lib().gf_transform_batch(self.handle, arr, len(chunks), first_id, None))

    def advance(self):
        a, m = C.c_uint32(), C.c_int()
        self._check(lib().gf_stage_advance(self.handle, None, C.byref(a), C.byref(m)))
        return bool(m.value), a.value

    def finish(self, cap: int = 1 << 16):
        arr = (ChunkDesc * cap)()
        n = C.c_uint32()
        _raise(lib().gf_stage_finish(self.handle, arr, cap, C.byref(n)))
        return [arr[i].tuple() for i in range(n.value)]

    def abort(self):
        _raise(lib().gf_stage_abort(self.handle))

    def counts(self, bins: int):
        a = (C.c_uint32 * bins)()
        _raise(lib().gf_stage_counts(self.handle, a, bins))
        return list(a)

    def tiles(self, cap: int = 1 << 16):
        arr = (TileDesc * cap)()
        n = C.c_uint32()
        _raise(lib().gf_stage_tiles(self.handle, arr, cap, C.byref(n)))
        return [(arr[i].chunk_index, arr[i].tile_offset, arr[i].count) for i in range(n.value)]

    def close(self):
        if self.handle:
            lib().gf_stage_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- functor descriptors (the compiled-in functor set) ----------------------------------------------
@dataclass
class Op:
    role: int
    op: str
    bins: int = 1
    upd: str = "NONE"
    u: List[int] = field(default_factory=list)
    f: List[float] = field(default_factory=list)
    ptr: List[int] = field(default_factory=list)
    size: List[int] = field(default_factory=list)

    def desc(self) -> NodeDesc:
        d = NodeDesc()
        d.role, d.op, d.upd, d.bins = self.role, OPS[self.op], UPDS[self.upd], self.bins
        for i, v in enumerate(self.u):
            d.u[i] = v
        for i, v in enumerate(self.f):
            d.f[i] = v
        for i, v in enumerate(self.ptr):
            d.ptr[i] = v
        for i, v in enumerate(self.size):
            d.size[i] = v
        return d

    def neutral(self) -> dict:
        """Backend-neutral description (the same dict drives the oracle)."""
        return dict(role=self.role, op=self.op, upd=self.upd, bins=self.bins, u=list(self.u), f=list(self.f))


def UniformParticles() -> Op:
    """weight = U(0,1] from Philox word 0, id = index, gen = 0."""
    return Op(ROLE_SOURCE, "SRC_PARTICLE_UNIFORM")


def ZipfKeys(values: int = 1 << 20, s: float = 1.1) -> Op:
    return Op(ROLE_SOURCE, "SRC_KEYED_ZIPF", u=[values], f=[s])


def PhotonSource() -> Op:
    return Op(ROLE_SOURCE, "SRC_PHOTON")


def WeightRoute(threshold: float = 1e-3, distribution: str = "parity", destinations: int = 2) -> Op:
    """bin 0 (sink) when weight < threshold, else 1 + route(fmix32(id))."""
    return Op(ROLE_TRANSFORM, "CLS_WEIGHT_ROUTE", bins=1 + destinations, u=[ROUTES[distribution]], f=[threshold])


def KeyBucket(buckets: int = 4096) -> Op:
    return Op(ROLE_TRANSFORM, "CLS_KEY_BUCKET", bins=buckets)


def IdMod(bins: int, distribution: int = 0, update: str = "NONE") -> Op:
    return Op(ROLE_TRANSFORM, "CLS_ID_MOD", bins=bins, upd=update, u=[distribution])


def Lut(bins: int, table_ptr: int, nid: int, ngen: int, offset: int = 0, update: str = "NONE", add_bin: int = 0,
        add: float = 0.0, table_len: Optional[int] = None) -> Op:
    """bin = table[offset + min(id, nid-1) * ngen + min(gen, ngen-1)]; table_len: u32 entries of the
    table (default offset + nid * ngen, the entries this classifier reads)."""
    n = offset + nid * ngen if table_len is None else table_len
    return Op(ROLE_TRANSFORM, "CLS_LUT", bins=bins, upd=update, u=[nid, ngen, add_bin, offset], f=[0.0, add],
              ptr=[table_ptr], size=[n])


def PhotonAlive() -> Op:
    return Op(ROLE_TRANSFORM, "CLS_PHOTON_ALIVE", bins=2)


def ScaleWeight(m: float) -> Op:
    return Op(ROLE_TRANSFORM, "XF_SCALE_WEIGHT", f=[m])


def PhotonStep(tally_ptr: int, tally_bins: int = 1024) -> Op:
    """tally_bins: f64 entries of the tally buffer (at least the run's photon tally bins)."""
    return Op(ROLE_TRANSFORM, "XF_PHOTON_STEP", ptr=[tally_ptr], size=[tally_bins])


def WeightHistogram(counts_ptr: int, lo: int, n: int, field_: int = 0) -> Op:
    return Op(ROLE_SINK, "SNK_HIST", bins=0, u=[field_, lo, n], ptr=[counts_ptr], size=[n])


def RecordWeight(record_ptr: int, n: int) -> Op:
    return Op(ROLE_SINK, "SNK_RECORD_WEIGHT", bins=0, ptr=[record_ptr], size=[n])


def ChunkDigest(log_ptr: int, log_cap: int, multiset_ptr: int = 0, bins: int = 0) -> Op:
    return Op(ROLE_SINK, "SNK_CHUNK_DIGEST", bins=0, ptr=[log_ptr, multiset_ptr], size=[log_cap, bins])


def PhotonDone() -> Op:
    return Op(ROLE_SINK, "SNK_PHOTON_TALLY", bins=0)


# ---- typed flow -------------------------------------------------------------------------------------
@dataclass
class NodeHandle:
    key: int
    bins: int
    name: str

    def out(self, b: int) -> tuple:
        if b >= self.bins:
            raise IndexError("graphFlow: output bin out of range")
        return (self, b)


@dataclass
class Tuning:
    """Optional knobs (flow/Tuning.hpp:27-145)."""
    seed: int = 0
    batch_chunks: Optional[int] = None
    chunk_capacity: Optional[int] = None
    live_particles: Optional[int] = None
    num_chunks: Optional[int] = None
    tile_items: Optional[int] = None
    schedule: int = SCHEDULE_DEVICE
    steps_per_launch: Optional[int] = None
    trace: bool = False
    use_graph: int = 1
    fuse_transforms: bool = False
    max_workgroups: Optional[int] = None
    device_engine: int = 0  # 0 autotuned, 1 wide step kernels, 2 one persistent workgroup, 3 XCD team
    photon: Optional[dict] = None
    team_xcd: Optional[int] = None  # team engine: run the team on this XCD (None: where it first starts)
    inflight_batches: Optional[int] = None  # batches in flight per node (Tuning::inflightBatches)
    bin_path: int = 0  # BIN_PATH_AUTO (hash above 4096 bins), BIN_PATH_DENSE, BIN_PATH_HASH
    hash_table_entries: Optional[int] = None  # hash path: slots of a tile's bin table (None: 2 * tile items)
    stall_timeout_ms: Optional[int] = None  # plain schedule: longest wait for a device stage (Tuning::stallTimeout)
    fused_batches: bool = False  # plain schedule: gf_transform_batch per partition batch (no host round trip)
    source_chunks: Optional[int] = None  # device schedule: chunks a source lane admits per step (None: 1)

    def validate(self):
        for name in ("batch_chunks", "chunk_capacity", "live_particles", "num_chunks", "tile_items",
                     "steps_per_launch", "inflight_batches", "hash_table_entries", "stall_timeout_ms",
                     "source_chunks"):
            v = getattr(self, name)
            if v is not None and v == 0:
                raise ValueError(f"graphFlow::Tuning: {name} must be non-zero")

    def c(self) -> TuningC:
        t = TuningC(seed=self.seed, batch_chunks=self.batch_chunks or 0, chunk_capacity=self.chunk_capacity or 0,
                    live_particles=self.live_particles or 0, num_chunks=self.num_chunks or 0,
                    tile_items=self.tile_items or 0, schedule=self.schedule,
                    steps_per_launch=self.steps_per_launch or 0, trace=1 if self.trace else 0,
                    use_graph=self.use_graph, fuse_transforms=1 if self.fuse_transforms else 0,
                    max_workgroups=self.max_workgroups or 0, device_engine=self.device_engine,
                    team_xcd=0 if self.team_xcd is None else self.team_xcd + 1,
                    inflight_batches=self.inflight_batches or 0, bin_path=self.bin_path,
                    hash_table_entries=self.hash_table_entries or 0, stall_timeout_ms=self.stall_timeout_ms or 0,
                    fused_batches=1 if self.fused_batches else 0, source_chunks=self.source_chunks or 0)
        ph = self.photon or {}
        t.photon_layers = ph.get("layers", 10)
        t.photon_tally_bins = ph.get("tally_bins", 1024)
        t.photon_layer_mfp = ph.get("layer_mfp", 1.0)
        t.photon_mu_t = ph.get("mu_t", 1.0)
        t.photon_albedo = ph.get("albedo", 0.9)
        t.photon_hg_g = ph.get("hg_g", 0.9)
        t.photon_roulette_w = ph.get("roulette_w", 1e-4)
        t.photon_roulette_p = ph.get("roulette_p", 0.1)
        return t


class Flow:
    """Flow<TPayload, TDevice, TStorage> (flow/Flow.hpp:49-754) over the C ABI."""

    def __init__(self, payload: int = PAYLOAD_PARTICLE, ctx: Optional[Context] = None):
        self.ctx = ctx or default_context()
        h = _V()
        _raise(lib().gf_flow_create(self.ctx.handle, payload, C.byref(h)))
        self.handle = h
        self.payload = payload
        self.nodes: List[Op] = []
        self.names: List[str] = []
        self.routes: List[List[int]] = []
        self.tuning = Tuning()

    def _err(self) -> str:
        buf = C.create_string_buffer(512)
        lib().gf_flow_last_error(self.handle, buf, 512)
        return buf.value.decode()

    def _add(self, name: str, op: Op) -> NodeHandle:
        d = op.desc()
        key = C.c_uint32()
        rc = lib().gf_flow_add_node(self.handle, name.encode(), C.byref(d), C.byref(key))
        _raise(rc, self._err() if rc else "")
        self.nodes.append(op)
        self.names.append(name)
        self.routes.append([-1] * (op.bins if op.role != ROLE_SINK else 0))
        return NodeHandle(key.value, op.bins, name)

    def source(self, name: str, op: Op) -> NodeHandle:
        if op.role != ROLE_SOURCE:
            raise TypeError("graphFlow: source() needs a source functor")
        return self._add(name, op)

    def partition(self, name: str, op: Op) -> NodeHandle:
        if op.role != ROLE_TRANSFORM or not op.op.startswith("CLS_"):
            raise TypeError("graphFlow: partition() needs a classifier functor")
        return self._add(name, op)

    def transform(self, name: str, op: Op) -> NodeHandle:
        if op.role != ROLE_TRANSFORM or not op.op.startswith("XF_"):
            raise TypeError("graphFlow: transform() needs a transform functor")
        return self._add(name, op)

    def sink(self, name: str, op: Op) -> NodeHandle:
        if op.role != ROLE_SINK:
            raise TypeError("graphFlow: sink() needs a sink functor")
        return self._add(name, op)

    def connect(self, frm, to: NodeHandle, bin_: Optional[int] = None):
        """connect(node, node) for single-output nodes, connect(node.out(b), node) or connect(node, node, bin)."""
        if isinstance(frm, tuple):
            frm, bin_ = frm
        if bin_ is None:
            if frm.bins != 1:
                raise ValueError("graphFlow: a multi-output node must be connected via out<Bin>(), not connect(node, node)")
            bin_ = 0
        rc = lib().gf_flow_connect(self.handle, frm.key, bin_, to.key)
        _raise(rc, self._err() if rc else "")
        self.routes[frm.key][bin_] = to.key

    def configure(self, tuning: Tuning):
        tuning.validate()
        t = tuning.c()
        _raise(lib().gf_flow_configure(self.handle, C.byref(t)))
        self.tuning = tuning

    def run(self, source: NodeHandle, items: int) -> dict:
        st = Stats()
        rc = lib().gf_flow_run(self.handle, source.key, items, C.byref(st))
        _raise(rc, self._err() if rc else "")
        return st.as_dict()

    def autotune(self, source: NodeHandle, probe_items: int = 0) -> dict:
        """Pick the fastest device-schedule configuration for the knobs left open (device_engine 0,
        batch_chunks unset) by probe runs; configures the flow with it. The probe runs write the
        flow's sinks (clear them before the real run)."""
        t = TuningC()
        ms = C.c_double()
        rc = lib().gf_flow_autotune(self.handle, source.key, probe_items, C.byref(t), C.byref(ms))
        _raise(rc, self._err() if rc else "")
        return {"device_engine": t.device_engine, "max_workgroups": t.max_workgroups, "batch_chunks": t.batch_chunks,
                "tile_items": t.tile_items, "probe_ms": ms.value}

    def trace(self):
        import numpy as np
        n = C.c_uint64()
        lib().gf_flow_trace(self.handle, None, 0, C.byref(n))
        a = np.zeros(max(n.value, 1), np.uint32)
        lib().gf_flow_trace(self.handle, a.ctypes.data_as(C.POINTER(C.c_uint32)), n.value, C.byref(n))
        return a[:n.value]

    def kernel_times(self):
        """Kernel ms and launches per kind of the last timed run (use_graph 2 or 3): classify, plan,
        scatter + retire/select, persistent engine launches, then the wide engine's partition, source
        and sink scatter kernels and its retire/select kernels."""
        ms = (C.c_double * 8)()
        n = (C.c_uint64 * 8)()
        lib().gf_flow_kernel_times(self.handle, ms, n, 8)
        return list(ms), list(n)

    def phase_profile(self):
        """Per-phase wall-clock microseconds per step of the micro or team engine (run with
        use_graph=3; the device wall clock ticks at 100 MHz)."""
        a = (C.c_uint64 * 32)()
        _raise(lib().gf_flow_phase_profile(self.handle, a, 32))
        if int(a[14]):
            steps = int(a[14])
            names = {9: "barrier_wait", 10: "step_start", 12: "step_work", 13: "after_barrier",
                     15: "wave0_segments", 16: "scheduler_retire", 17: "scheduler_select"}
            out = {n: a[i] / 100.0 / steps for i, n in names.items()}
            cyc = {18: "sel_fills", 19: "sel_credit", 20: "sel_check", 21: "sel_lists", 22: "ret_out", 23: "ret_in",
                   24: "plan_nodes", 25: "plan_pops", 27: "cnt_scan", 26: "seg_locate", 7: "seg_place", 28: "seg_rank",
                   29: "seg_store", 30: "seg_agg",
                   31: "step"}
            for i, nm in cyc.items():
                out["cycles_" + nm] = a[i] / steps
            return out, steps
        steps = max(int(a[8]), 1)
        names = ["count", "plan", "source", "partition", "transform", "sink", "retire", "select"]
        return {n: a[i] / 100.0 / steps for i, n in enumerate(names)}, int(a[8])

    def member_profile(self, members=32):
        """Team engine, timed run: per member [barrier wait, step start, scatter phase, scheduler
        wave] in microseconds per step, segments scattered / sink segments per step, then the
        eight waves' work time from the step's start in microseconds per step (run with
        use_graph=3)."""
        stride = 16
        words = 32 + 64 * stride
        a = (C.c_uint64 * words)()
        _raise(lib().gf_flow_phase_profile(self.handle, a, words))
        steps = max(int(a[14]), 1)
        rows = []
        for m in range(members):
            b = 32 + stride * m
            rows.append([a[b] / 100.0 / steps, a[b + 1] / 100.0 / steps, a[b + 2] / 100.0 / steps,
                         a[b + 3] / 100.0 / steps, a[b + 4] / steps, a[b + 5] / steps]
                        + [a[b + 8 + w] / 100.0 / steps for w in range(8)])
        return rows

    def item_passes(self):
        a, b = C.c_uint64(), C.c_uint64()
        lib().gf_flow_item_passes(self.handle, C.byref(a), C.byref(b))
        return a.value, b.value

    def neutral_graph(self) -> list:
        """The registered graph as backend-neutral dicts (role, op, bins, routes, params)."""
        g = []
        for op, r in zip(self.nodes, self.routes):
            d = op.neutral()
            d["routes"] = list(r)
            g.append(d)
        return g

    def close(self):
        if self.handle:
            lib().gf_flow_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def hist_rows(counts_ptr: int, n: int, lo: int, ctx: Optional[Context] = None):
    """Compact a dense device u64 count array into ascending (value, count) rows (numpy Nx2 uint64)."""
    import numpy as np
    ctx = ctx or default_context()
    rows = C.c_uint64()
    _raise(lib().gf_hist_rows(ctx.handle, counts_ptr, n, lo, None, 0, C.byref(rows)))
    out = np.zeros((max(rows.value, 1), 2), np.uint64)
    _raise(lib().gf_hist_rows(ctx.handle, counts_ptr, n, lo, out.ctypes.data_as(C.POINTER(C.c_uint64)), rows.value,
                              C.byref(rows)))
    return out[:rows.value]


def rows_text(rows) -> bytes:
    """The reference's histogram serialisation: one 'value,count\\n' line per row, ascending
    (tools/bench_matrix.sh:717-723)."""
    return b"".join(b"%d,%d\n" % (int(v), int(c)) for v, c in rows)


def rows_sha256(rows) -> str:
    import hashlib
    return hashlib.sha256(rows_text(rows)).hexdigest()


# ---- canonical workloads -----------------------------------------------------------------------------
def feedback_network(flow: Flow, sink_op: Op, threshold: float = 1e-3, distribution: str = "parity",
                     destinations: int = 2, multipliers: Optional[Sequence[float]] = None):
    """Generate -> Filter -> {A, B, ...} -> Filter, Filter.sink -> Sink (fp32 feedback network).

    Destination d multiplies the weight by multipliers[d] (default 0.5 for even d, 0.75 for odd d)."""
    if multipliers is None:
        multipliers = [0.5 if d % 2 == 0 else 0.75 for d in range(destinations)]
    src = flow.source("generate", UniformParticles())
    filt = flow.partition("filter", WeightRoute(threshold, distribution, destinations))
    dests = [flow.transform(f"process {chr(65 + d) if d < 26 else d}", ScaleWeight(multipliers[d]))
             for d in range(destinations)]
    sink = flow.sink("record weight", sink_op)
    flow.connect(src, filt)
    flow.connect(filt.out(0), sink)
    for d, node in enumerate(dests):
        flow.connect(filt.out(1 + d), node)
        flow.connect(node, filt)
    return src, filt, dests, sink


JOB_FEEDBACK, JOB_HASH, JOB_PHOTON = 0, 1, 2
JOB_NAMES = {JOB_FEEDBACK: "feedback", JOB_HASH: "hash", JOB_PHOTON: "photon"}


class Server:
    """Serving loop over the C ABI: independent flow jobs (JOB_FEEDBACK particle feedback network,
    JOB_HASH keyed murmur3 pipeline with 4096 buckets, JOB_PHOTON photon transport) admitted into
    slots that run side by side on the device (gf_server_* in include/graphflow.h)."""

    def __init__(self, slots: int = 8, team_size: int = 0, memory_budget: int = 0, max_items: int = 0,
                 prewarm: bool = False, ctx: Optional[Context] = None):
        self.ctx = ctx or default_context()
        self.cfg = ServerConfigC(slots=slots, team_size=team_size, memory_budget=memory_budget, max_items=max_items,
                                 prewarm=1 if prewarm else 0)
        h = _V()
        rc = lib().gf_server_create(self.ctx.handle, C.byref(self.cfg), C.byref(h))
        if rc != OK:
            buf = C.create_string_buffer(512)
            lib().gf_context_last_error(self.ctx.handle, buf, 512)
            _raise(rc, buf.value.decode(errors="replace"))
        self.handle = h

    def submit(self, topology: int, items: int, seed: int = 42) -> int:
        j = JobDescC(topology=topology, items=items, seed=seed)
        jid = C.c_uint64()
        _raise(lib().gf_server_submit(self.handle, C.byref(j), C.byref(jid)))
        return jid.value

    def poll(self, wait_all: bool = False, cap: int = 8192) -> list:
        arr = (JobResultC * cap)()
        n = C.c_uint32()
        _raise(lib().gf_server_poll(self.handle, arr, cap, C.byref(n), 1 if wait_all else 0))
        out = []
        for r in arr[:n.value]:
            out.append({f: getattr(r, f) for f, _ in JobResultC._fields_})
        return out

    def tally(self, job: int, bins: int = 1024):
        import numpy as np
        t = np.zeros(bins, np.float64)
        _raise(lib().gf_server_job_tally(self.handle, job, t.ctypes.data_as(C.POINTER(C.c_double)), bins))
        return t

    def last_error(self) -> str:
        b = C.create_string_buffer(512)
        lib().gf_server_last_error(self.handle, b, 512)
        return b.value.decode()

    def close(self):
        if getattr(self, "handle", None):
            lib().gf_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def record_digest(weight_bits) -> int:
    """The served feedback result digest of per-id weight bits: sum over ids of mix64(id << 32 | bits)
    modulo 2^64 (numpy; the same value the device computes)."""
    import numpy as np
    w = np.asarray(weight_bits, np.uint64)
    i = np.arange(len(w), dtype=np.uint64)
    with np.errstate(over="ignore"):
        v = (i << np.uint64(32)) | w
        return int(mix64_np(v).sum(dtype=np.uint64))


def multiset_digest(sums, counts) -> int:
    """The served hash result digest of per-bucket (hash sum, count) multiset pairs."""
    import numpy as np
    s = np.asarray(sums, np.uint64)
    c = np.asarray(counts, np.uint64)
    b = np.arange(len(s), dtype=np.uint64)
    with np.errstate(over="ignore"):
        d = mix64_np((b << np.uint64(48)) ^ s) + mix64_np(c * np.uint64(0x9E3779B97F4A7C15) + b)
        return int(d.sum(dtype=np.uint64))


def mix64_np(v):
    """SplitMix64 finalizer over a numpy uint64 array (wrapping arithmetic)."""
    import numpy as np
    v = np.asarray(v, np.uint64).copy()
    with np.errstate(over="ignore"):
        v ^= v >> np.uint64(30)
        v *= np.uint64(0xBF58476D1CE4E5B9)
        v ^= v >> np.uint64(27)
        v *= np.uint64(0x94D049BB133111EB)
        v ^= v >> np.uint64(31)
    return v


def photon_network(flow: Flow, tally_ptr: int):
    """Photon transport as a flow (examples/photon_scattering/PhotonScattering.hpp:84-102, with the
    event routing folded into one alive/done partition): Generate -> Step -> Alive; Alive.alive ->
    Step, Alive.done -> Done. Step is one flight + absorption tally + roulette + Henyey-Greenstein
    scatter; the absorbed weight goes to tally_ptr (double[tally_bins], depth bins over the slab).
    The step node is node 1, so its random stream purpose is 2."""
    src = flow.source("generate", PhotonSource())
    step = flow.transform("step", PhotonStep(tally_ptr))
    alive = flow.partition("alive", PhotonAlive())
    done = flow.sink("done", PhotonDone())
    flow.connect(src, step)
    flow.connect(step, alive)
    flow.connect(alive.out(0), step)
    flow.connect(alive.out(1), done)
    return src, step, alive, done


PHOTON_DEFAULTS = {"layers": 10, "tally_bins": 1024, "layer_mfp": 1.0, "mu_t": 1.0, "albedo": 0.9, "hg_g": 0.9,
                   "roulette_w": 1e-4, "roulette_p": 0.1}


FEEDBACK_HIST_LO = 0x33800000  # bits of 2^-24, the smallest U(0,1] weight
FEEDBACK_HIST_N = 0x3F800000 - 0x33800000 + 1  # up to bits of 1.0
