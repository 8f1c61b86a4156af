// libgraphflow_amd: MI355X transform-batch path behind the graphflow.h C ABI.
//
// Plain form  (GF_SCHEDULE_HOST):   a host GraphRunner (sched/GraphRunner.hpp:716-1257 with every
//   batch completed before the next loop iteration) drives per-node stages; a partition batch is
//   classify kernel -> host plan (DenseBinTable.hpp:39-91) -> stable scatter kernel, like the
//   reference's two host round trips per batch.
// Batched form (GF_SCHEDULE_DEVICE): the device-resident step engine of gf_engine.hpp replayed
//   from a hipGraph; the host never plans.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

#include "gf_team.hpp"
#include "gf_batch.hpp"

using namespace gfd;

namespace {

struct HipError {
    hipError_t e;
    const char* what;
};

#define HIPX(call)                                                                                 \
    do {                                                                                           \
        hipError_t _e = (call);                                                                    \
        if (_e != hipSuccess) throw HipError{_e, #call};                                           \
    } while (0)

struct GfError {
    gf_status s;
    std::string msg;
};

[[noreturn]] void fail(gf_status s, const std::string& m) { throw GfError{s, m}; }

template <class F>
gf_status guard(std::string* err, F&& f)
{
    try {
        f();
        return GF_OK;
    } catch (const GfError& g) {
        if (err) *err = g.msg;
        return g.s;
    } catch (const HipError& h) {
        if (err) *err = std::string(h.what) + ": " + hipGetErrorString(h.e);
        return GF_DEVICE_ERROR;
    } catch (const std::bad_alloc&) {
        if (err) *err = "host allocation failed";
        return GF_BAD_ARGUMENT;
    }
}

// ---- closed-form host math (Reservation.hpp:15-71, DenseBinTable.hpp:39-91) ----------------
uint64_t ceil_div(uint64_t n, uint64_t d) { return n / d + (n % d != 0u ? 1u : 0u); }

uint64_t output_chunk_bound(uint64_t n, uint64_t C, uint64_t B)
{
    if (n == 0u) return 0u;
    if (C == 0u || B == 0u) fail(GF_BAD_ARGUMENT, "outputChunkBound: capacity and maxBins must be non-zero");
    uint64_t per = ceil_div(n, C), bins = B < n ? B : n;
    if (per > UINT64_MAX - bins + 1u) fail(GF_OVERFLOW, "outputChunkBound: chunk count overflow");
    uint64_t r = per + bins - 1u;
    return r < n ? r : n;
}

template <class K>
void set_lds(K kernel, size_t bytes)
{
    if (bytes > 64u * 1024u) HIPX(hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
}

bool is_partition(const gf_node_desc& nd) { return nd.op >= GF_CLS_WEIGHT_ROUTE && nd.op < GF_XF_SCALE_WEIGHT; }

uint32_t reserve_bound_of(const gf_node_desc& nd, uint32_t C, uint32_t Q)
{
    uint64_t b = output_chunk_bound((uint64_t)C * Q, C, is_partition(nd) ? nd.bins : 1u);
    if (b > UINT32_MAX) fail(GF_OVERFLOW, "partition batch output-chunk bound overflow");
    return (uint32_t)b;
}

// Bin path of a partition (core/BinPath.hpp). The dense path keeps 16 bytes of LDS per bin in a
// scatter tile (per-wave running counts): up to 4096 bins that is 64 KB, two scatter workgroups
// per CU of MI355X's 160 KB LDS; automatic selects the hash path above that (the reference's
// limit, Config.hpp:174, is derived from its GPUs' shared memory).
constexpr uint32_t kDenseBinLimit = 4096u;
constexpr uint32_t kHashTileLimit = 1024u;
constexpr uint32_t kKernelKinds = 8;  // kernel-time kinds of gf_flow_kernel_times  // hash path tiles: a tile's table of 2 * 1024 slots fits LDS
constexpr uint32_t kXcds = 8;  // XCDs of an MI355X

uint32_t next_pow2(uint32_t v)
{
    uint32_t r = 1;
    while (r < v) r <<= 1;
    return r;
}

bool on_hash_path(const gf_node_desc& nd, uint32_t bin_path)
{
    const bool part = nd.role == GF_ROLE_TRANSFORM && nd.op >= GF_CLS_WEIGHT_ROUTE && nd.op < GF_XF_SCALE_WEIGHT;
    if (!part) return false;
    if (bin_path == GF_BIN_PATH_DENSE && nd.bins > kDenseBinLimit)
        fail(GF_BAD_ARGUMENT, "graphFlow::Config: the dense bin path supports at most " + std::to_string(kDenseBinLimit) +
                                  " bins; select the hash path");
    return bin_path == GF_BIN_PATH_HASH || (bin_path == GF_BIN_PATH_AUTO && nd.bins > kDenseBinLimit);
}

// Slots of a hash-path tile's bin table (HashBinTable, BinTable.hpp:69-82): nextPow2 of the
// configured entries, or of two per tile item; 0 on the dense path.
uint32_t hash_entries_of(const gf_node_desc& nd, uint32_t bin_path, uint32_t configured, uint32_t tile_items)
{
    if (!on_hash_path(nd, bin_path)) return 0u;
    return next_pow2(configured ? configured : 2u * tile_items);
}

// bytes of the record struct a kernel holds per item (LDS staging), not the payload's HBM bytes
uint32_t record_bytes(uint32_t payload)
{
    return payload == GF_PAYLOAD_PARTICLE ? (uint32_t)sizeof(gfd::Particle)
         : payload == GF_PAYLOAD_KEYED ? (uint32_t)sizeof(gfd::Keyed) : (uint32_t)sizeof(gfd::Photon);
}

// Device buffers a node's functor reads or writes must hold what it indexes: a kernel never checks
// these bounds itself.
void validate_node(const gf_node_desc& nd, uint32_t tally_bins)
{
    if (nd.op == GF_CLS_LUT) {
        if (nd.u[0] == 0u || nd.u[1] == 0u) fail(GF_BAD_ARGUMENT, "graphFlow: a table classifier needs ids and generations");
        if (!nd.ptr[0]) fail(GF_BAD_ARGUMENT, "graphFlow: a table classifier needs its table");
        if ((uint64_t)nd.u[3] + (uint64_t)nd.u[0] * nd.u[1] > nd.size[0])
            fail(GF_BAD_ARGUMENT, "graphFlow: the table classifier reads past its table (size[0])");
    }
    if (nd.op == GF_XF_PHOTON_STEP && (!nd.ptr[0] || nd.size[0] < tally_bins))
        fail(GF_BAD_ARGUMENT, "graphFlow: the photon step's tally buffer (size[0]) holds fewer than the run's tally bins");
}

uint32_t payload_bytes(uint32_t payload)
{
    return payload == GF_PAYLOAD_PARTICLE ? 12u : payload == GF_PAYLOAD_KEYED ? 12u : 36u;
}

std::vector<uint32_t> column_bytes(uint32_t payload)
{
    if (payload == GF_PAYLOAD_PARTICLE) return {4, 4, 4};
    if (payload == GF_PAYLOAD_KEYED) return {4, 8};
    return std::vector<uint32_t>(9, 4);
}

}  // namespace

// ==============================================================================================
// Kernels
// ==============================================================================================
template <class P, int R>
__global__ __launch_bounds__(kThreads) void k_count(const BatchArg, gf_node_desc nd, Cols cols, uint32_t C,
                                                    uint32_t* status, uint32_t hash_entries)
{
    extern __shared__ __align__(16) char smem[];
    const Batch b = batch_arg();
    if (hash_entries) count_tile_hash<P, R>(nd, b, blockIdx.x, cols, C, status, smem, hash_entries);
    else count_tile<P, R>(nd, b, blockIdx.x, cols, C, status, smem);
}

template <class P, int R, bool Stage>
__global__ __launch_bounds__(kThreads) void k_scatter(const BatchArg, gf_node_desc nd, Cols cols, uint32_t C,
                                                      uint32_t* status, uint32_t hash_entries)
{
    extern __shared__ __align__(16) char smem[];
    const Batch b = batch_arg();
    if (hash_entries) scatter_tile_hash<P, R>(nd, b, blockIdx.x, cols, C, status, smem, hash_entries);
    else scatter_tile<P, R, Stage>(nd, b, blockIdx.x, cols, C, status, smem);
}

template <class P, int R>
__global__ __launch_bounds__(kThreads) void k_transform(const BatchArg, gf_node_desc nd, Cols cols, uint32_t C,
                                                        RunParams rp)
{
    const Batch b = batch_arg();
    transform_tile<P, R>(nd, b, blockIdx.x, cols, C, rp);
}

template <class P>
__global__ __launch_bounds__(kThreads) void k_source(const BatchArg, gf_node_desc nd, Cols cols, uint32_t C,
                                                     RunParams rp)
{
    const Batch b = batch_arg();
    source_tile<P>(nd, b, blockIdx.x, cols, C, rp);
}

template <class P>
__global__ __launch_bounds__(kThreads) void k_sink(const BatchArg, gf_node_desc nd, Cols cols, uint32_t C,
                                                   uint32_t* status)
{
    extern __shared__ __align__(16) char smem[];
    const Batch b = batch_arg();
    sink_tile<P>(nd, b, blockIdx.x, cols, C, status, smem);
}

// One partition batch in one launch (gf_batch.hpp). FusedArgs is the first parameter: read in the
// kernel-argument segment (offset 0), where the inputs and reservation of a small batch travel.
template <class P, int R>
__global__ __launch_bounds__(1024) void k_batch_fused(FusedArgs, Cols cols)
{
    extern __shared__ __align__(16) char smem[];
    fused_partition_batch<P, R>(*(const FusedArgs*)__builtin_amdgcn_kernarg_segment_ptr(), cols, smem);
}

// The grid form of the same batch (gf_batch.hpp): count, plan, scatter.
template <class P, int R>
__global__ __launch_bounds__(kThreads) void k_fused_count(FusedArgs, Cols cols)
{
    extern __shared__ __align__(16) char smem[];
    fused_grid_count<P, R>(*(const FusedArgs*)__builtin_amdgcn_kernarg_segment_ptr(), cols, smem);
}

__global__ __launch_bounds__(1024) void k_fused_plan(FusedArgs)
{
    fused_grid_plan(*(const FusedArgs*)__builtin_amdgcn_kernarg_segment_ptr());
}

template <class P, int R>
__global__ __launch_bounds__(kThreads) void k_fused_scatter(FusedArgs, Cols cols)
{
    extern __shared__ __align__(16) char smem[];
    fused_grid_scatter<P, R>(*(const FusedArgs*)__builtin_amdgcn_kernarg_segment_ptr(), cols, smem);
}

// A photon step's private tally copies summed into its tally, copy by copy in a fixed order, and
// cleared for the next run.
__global__ void k_tally_fold(double* copies, uint32_t bins, double* tally)
{
    for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < bins; b += gridDim.x * blockDim.x) {
        double sum = 0.0;
        for (uint32_t k = 0; k < kTallyCopies; ++k) {
            sum += copies[(size_t)k * bins + b];
            copies[(size_t)k * bins + b] = 0.0;
        }
        tally[b] += sum;
    }
}

// Tile descriptors of a batch as the device computes them (TileDescriptor list).
__global__ void k_tiles(const Batch* bp, uint32_t C, gf_tile_desc* out)
{
    const Batch b = *bp;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < b.n_tiles; t += gridDim.x * blockDim.x) {
        TileSpan s = tile_span(b, b.in, b.n_in, t, C);
        out[t] = gf_tile_desc{b.in[s.chunk].index, s.offset, s.count};
    }
}

// ---- step engine ------------------------------------------------------------------------------
__device__ inline bool engine_stopped(const Engine& e, uint32_t* flag)
{
    if (threadIdx.x == 0) *flag = (load_ctl(&e.ctl[CTL_DONE]) | load_ctl(&e.ctl[CTL_STATUS])) != 0u ? 1u : 0u;
    __syncthreads();
    return *flag != 0u;
}

// The step's work prefix (n + 1 words, n <= kMaxNodes) into LDS: a tile's lane is then a binary
// search in LDS instead of a walk over global memory.
__device__ inline void load_prefix(const uint32_t* prefix, uint32_t n, uint32_t* sh)
{
    for (uint32_t k = threadIdx.x; k <= n; k += blockDim.x) sh[k] = prefix[k];
}

template <class P, int R>
__global__ __launch_bounds__(kThreads) void k_engine_ka(Engine e)
{
    extern __shared__ __align__(16) char smem[];
    __shared__ uint32_t flag, pre[kMaxNodes + 1];
    load_prefix(e.ka_prefix, e.n_nodes, pre);
    if (engine_stopped(e, &flag)) return;  // its barrier also publishes `pre`
    const uint32_t total = pre[e.n_nodes];
    for (uint32_t t = blockIdx.x; t < total; t += gridDim.x) {
        const uint32_t k = lane_of(pre, e.n_nodes, t);
        const NodeSt s = e.st[k];
        const Batch b = make_batch(e, k, s);
        const uint32_t E = e.arr[k].hash_entries;
        if (E) count_tile_hash<P, R>(e.nodes[k], b, t - pre[k], e.cols, e.C, &e.ctl[CTL_STATUS], smem, E);
        else count_tile<P, R>(e.nodes[k], b, t - pre[k], e.cols, e.C, &e.ctl[CTL_STATUS], smem);
    }
}

// The plan of every lane of the step over the grid: one workgroup per active lane, pop offsets by
// a decoupled look-back (reads KA's counts across the kernel boundary).
__global__ __launch_bounds__(1024) void k_engine_plan(Engine e)
{
    __shared__ GridPlanLds L;
    __shared__ uint32_t flag;
    if (engine_stopped(e, &flag)) return;
    engine_plan_grid(e, L);
}

// Trace records of the planned step (one workgroup; only when the run records a trace).
__global__ __launch_bounds__(1024) void k_engine_trace(Engine e)
{
    __shared__ unsigned long long len;
    engine_trace_step(e, &len);
}

// Tile kinds one scatter-kernel launch runs. A step launches three instantiations (partition and
// transform tiles, source tiles, sink tiles: disjoint lanes, disjoint chunks), so each holds only its
// own code's registers and occupancy does not follow the largest tile function.
enum { KB_PART = 1u, KB_SRC = 2u, KB_SINK = 4u };

__device__ inline uint32_t kb_group(uint32_t kind)
{
    return kind == KIND_SOURCE ? KB_SRC : kind == KIND_SINK ? KB_SINK : KB_PART;
}

// Tiles round-robin over the workgroups; the tiles of the sinks flagged NODE_WAVE_SINK (sinks fed by
// wide partitions, whose chunks are small) round-robin over the waves instead, several at a time (one
// per row of 4, 8 or 16 lanes).
template <class P, int R, uint32_t Kinds>
__global__ __launch_bounds__(kThreads) void k_engine_kb(Engine e)
{
    extern __shared__ __align__(16) char smem[];
    __shared__ uint32_t flag, pre[kMaxNodes + 1], order[kMaxNodes], wsink[kMaxNodes], scan[kWaves];
    const uint32_t n = e.n_nodes, tid = threadIdx.x;
    // this launch's lanes in node order (partitions first in the partition launch: their tiles are
    // the step's longest); pre = tiles before position i, order[i] = its lane (a thread per lane,
    // n <= kMaxNodes <= the block)
    uint32_t tiles = 0u, part = 0u, other = 0u;
    if (tid < n) {
        const uint32_t kind = e.arr[tid].kind;
        const bool mine = (kb_group(kind) & Kinds) != 0u;
        tiles = mine ? e.kb_prefix[tid + 1u] - e.kb_prefix[tid] : 0u;
        part = mine && kind == KIND_PARTITION ? 1u : 0u;
        other = mine && !part ? 1u : 0u;
        wsink[tid] = mine && (e.arr[tid].flags & NODE_WAVE_SINK) && e.st[tid].active ? 1u : 0u;
    }
    uint32_t n_part, t_part, t_other, n_other;
    const uint32_t i_part = block_exscan(part, scan, &n_part);
    const uint32_t i_other = block_exscan(other, scan, &n_other);
    const uint32_t p_part = block_exscan(part * tiles, scan, &t_part);
    const uint32_t p_other = block_exscan(other * tiles, scan, &t_other);
    const uint32_t nm = n_part + n_other;
    if (tid < n && (part || other)) {
        const uint32_t i = part ? i_part : n_part + i_other;
        order[i] = tid;
        pre[i] = part ? p_part : t_part + p_other;
    }
    if (tid == 0) pre[nm] = t_part + t_other;
    if (engine_stopped(e, &flag)) return;  // its barrier also publishes `pre`, `order` and `wsink`
    const uint32_t total = pre[nm], G = gridDim.x;
    // optional profile (use_graph = 3): wall-clock ticks and tiles per tile kind, summed over workgroups
    unsigned long long tprof = e.prof ? wall_clock64() : 0ull;
    auto prof_tile = [&](uint32_t kind) {
        if (!e.prof) return;
        __syncthreads();
        if (tid == 0) {
            const unsigned long long now = wall_clock64();
            atomicAdd(&e.prof[PROF_KB_TICKS + kind], now - tprof);
            atomicAdd(&e.prof[PROF_KB_TILES + kind], 1ull);
            tprof = now;
        }
    };
    // Partition tiles by XCD: workgroups are dispatched to the 8 XCDs round robin (XCD =
    // blockIdx.x % 8), and each XCD takes a contiguous eighth of the tiles, so the tiles of one batch
    // (whose stores meet in the same output chunks) mostly share one L2, which merges their partial
    // lines before writing them back. Any other launch: the tiles round robin over the workgroups.
    uint32_t t_first = blockIdx.x, t_end = total, t_step = G;
    if constexpr (Kinds == KB_PART) {
        if (G % kXcds == 0u) {
            const uint32_t x = blockIdx.x % kXcds;
            t_first = (uint32_t)((uint64_t)total * x / kXcds) + blockIdx.x / kXcds;
            t_end = (uint32_t)((uint64_t)total * (x + 1u) / kXcds);
            t_step = G / kXcds;
        }
    }
    for (uint32_t t = t_first; t < t_end; t += t_step) {
        const uint32_t i = lane_of(pre, nm, t), k = order[i];
        if (wsink[k]) {
            // skip the lane's tiles: the first tile of this workgroup's round-robin at or after them
            const uint32_t end = pre[i + 1u];
            t += (end - t + G - 1u) / G * G - G;
            continue;
        }
        const gf_node_desc& nd = e.nodes[k];
        const NodeSt s = e.st[k];
        const Batch b = make_batch(e, k, s);
        const uint32_t lt = t - pre[i];
        const uint32_t kind = e.arr[k].kind;
        RunParams rp = e.rp;
        if (e.jobs) rp.seed = e.jobs[e.arr[k].job].seed;  // a served engine: the lane's job seed
        if constexpr ((Kinds & KB_SRC) != 0u) {
            if (kind == KIND_SOURCE) source_tile<P>(nd, b, lt, e.cols, e.C, rp);
        }
        if constexpr ((Kinds & KB_SINK) != 0u) {
            if (kind == KIND_SINK) sink_tile<P>(nd, b, lt, e.cols, e.C, &e.ctl[CTL_STATUS], smem);
        }
        if constexpr ((Kinds & KB_PART) != 0u) {
            if (kind == KIND_PARTITION) {
                const int32_t* fz = e.fuse + nd.route_offset;
                if constexpr (R <= 4) {
                    // hash-path tiles hold at most 1024 items
                    if (e.arr[k].hash_entries) {
                        scatter_tile_hash<P, R>(nd, b, lt, e.cols, e.C, &e.ctl[CTL_STATUS], smem, e.arr[k].hash_entries,
                                                e.real_nodes, fz, &rp, b.totals + nd.bins);
                        prof_tile(kind);
                        continue;
                    }
                }
                if (e.arr[k].flags & NODE_STAGE)
                    scatter_tile<P, R, true>(nd, b, lt, e.cols, e.C, &e.ctl[CTL_STATUS], smem, e.real_nodes, fz, &rp, b.totals + nd.bins);
                else
                    scatter_tile<P, R, false>(nd, b, lt, e.cols, e.C, &e.ctl[CTL_STATUS], smem, e.real_nodes, fz, &rp, b.totals + nd.bins);
            } else if (kind == KIND_TRANSFORM) {
                transform_tile<P, R>(nd, b, lt, e.cols, e.C, rp);
            }
        }
        prof_tile(kind);
    }
    if constexpr ((Kinds & KB_SINK) != 0u) {
        const uint32_t w = threadIdx.x >> 6, W = G * (uint32_t)kWaves;
        unsigned long long wprof = e.prof ? wall_clock64() : 0ull, wtiles = 0ull;
        for (uint32_t i = n_part; i < nm; ++i) {
            const uint32_t k = order[i];
            if (!wsink[k]) continue;
            const uint32_t t0 = pre[i], t1 = pre[i + 1u];
            const gf_node_desc& nd = e.nodes[k];
            const Batch b = make_batch(e, k, e.st[k]);
            // each wave takes 64 / RL consecutive tiles at a time, one per row of RL lanes; rows of 4
            // lanes up to 12 items per tile on average, of 8 up to 24, else of 16
            auto rows = [&](auto rl) {
                constexpr uint32_t NR = 64u / decltype(rl)::value;
                for (uint32_t t = t0 + (blockIdx.x * (uint32_t)kWaves + w) * NR; t < t1; t += NR * W) {
                    sink_tiles_rows<P, decltype(rl)::value>(nd, b, t - t0, t1 - t < NR ? t1 - t : NR, e.cols, e.C,
                                                            &e.ctl[CTL_STATUS]);
                    wtiles += t1 - t < NR ? t1 - t : NR;
                }
            };
            const uint32_t per = t1 > t0 ? b.particles / (t1 - t0) : 0u;
            if (per <= 12u) rows(std::integral_constant<uint32_t, 4u>{});
            else if (per <= 24u) rows(std::integral_constant<uint32_t, 8u>{});
            else rows(std::integral_constant<uint32_t, 16u>{});
        }
        if (e.prof && lane_id() == 0u && wtiles) {
            atomicAdd(&e.prof[PROF_KB_WAVE_TICKS], wall_clock64() - wprof);
            atomicAdd(&e.prof[PROF_KB_WAVE_TILES], wtiles);
        }
    }
}

// Retire of the step over the grid (after every tile of KB): outputs to the destination rings in
// order, inputs back on the free stack.
__global__ __launch_bounds__(1024) void k_engine_route(Engine e)
{
    __shared__ RouteLds L;
    __shared__ uint32_t flag;
    if (engine_stopped(e, &flag)) return;
    engine_route_grid(e, L);
}

// The step's counters and the select of the next step's batches (one workgroup of 1024 threads).
__global__ __launch_bounds__(1024) void k_engine_kr(Engine e)
{
    extern __shared__ __align__(16) char smem[];
    __shared__ uint32_t flag;
    if (engine_stopped(e, &flag)) return;
    SelLds& L = *reinterpret_cast<SelLds*>(smem);
    const unsigned long long t0 = e.prof && threadIdx.x == 0u ? wall_clock64() : 0ull;
    engine_retire_counters(e, L);
    if (e.prof && threadIdx.x == 0u) e.prof[PROF_SEL_RETIRE] += wall_clock64() - t0;
    engine_select(e, L, false);
}

// The selected lanes' input lists, tile / item prefixes and the work prefixes, over the grid.
__global__ __launch_bounds__(1024) void k_engine_lists(Engine e)
{
    __shared__ ListLds L;
    __shared__ uint32_t flag;
    if (engine_stopped(e, &flag)) return;
    engine_lists_grid(e, L);
}

// Micro engine: one persistent workgroup of G tile groups (256 threads each) runs whole steps with
// no kernel boundary: every tile of a node is handed to the groups in rounds, the plan, retire and
// select phases use all threads. The live working set stays in the L2 of one XCD, so each of the
// step's dependent round trips is an L2 hit instead of a cross-XCD miss.
template <class P, int R>
__global__ __launch_bounds__(1024) void k_micro(Engine e, uint32_t max_steps, uint32_t group_lds)
{
    extern __shared__ __align__(16) char smem[];
    __shared__ uint32_t flag;
    const uint32_t G = blockDim.x / kThreads, g = threadIdx.x / kThreads;
    char* gsm = smem + (size_t)g * group_lds;
    unsigned long long t0c = 0;
    auto mark = [&](uint32_t phase) {
        if (e.prof && threadIdx.x == 0) {
            unsigned long long now = wall_clock64();
            e.prof[phase] += now - t0c;
            t0c = now;
        }
    };
    for (uint32_t it = 0; it < max_steps; ++it) {
        if (engine_stopped(e, &flag)) return;
        if (e.prof && threadIdx.x == 0) {
            t0c = wall_clock64();
            e.prof[PROF_STEPS] += 1u;
        }
        for (uint32_t k = 0; k < e.n_nodes; ++k) {
            const NodeSt s = e.st[k];
            if (!s.active || e.arr[k].kind != KIND_PARTITION) continue;
            const Batch b = make_batch(e, k, s);
            for (uint32_t t0 = 0; t0 < s.n_tiles; t0 += G) {
                const uint32_t t = t0 + g;
                const uint32_t E = e.arr[k].hash_entries;
                if (E) count_tile_hash<P, R>(e.nodes[k], b, t < s.n_tiles ? t : kNoTile, e.cols, e.C, &e.ctl[CTL_STATUS], gsm, E);
                else count_tile<P, R>(e.nodes[k], b, t < s.n_tiles ? t : kNoTile, e.cols, e.C, &e.ctl[CTL_STATUS], gsm);
            }
        }
        __syncthreads();
        mark(PROF_COUNT);
        engine_plan(e, *reinterpret_cast<PlanLds*>(smem));
        __syncthreads();
        mark(PROF_PLAN);
        if (engine_stopped(e, &flag)) return;  // a failed plan publishes nothing
        for (uint32_t k = 0; k < e.n_nodes; ++k) {
            const NodeSt s = e.st[k];
            if (!s.active) continue;
            const gf_node_desc& nd = e.nodes[k];
            const Batch b = make_batch(e, k, s);
            const uint32_t kind = e.arr[k].kind;
            const bool stage = (e.arr[k].flags & NODE_STAGE) != 0u;
            for (uint32_t t0 = 0; t0 < s.n_tiles; t0 += G) {
                const uint32_t t = t0 + g < s.n_tiles ? t0 + g : kNoTile;
                if (kind == KIND_SOURCE) {
                    source_tile<P>(nd, b, t, e.cols, e.C, e.rp);
                } else if (kind == KIND_SINK) {
                    sink_tile<P>(nd, b, t, e.cols, e.C, &e.ctl[CTL_STATUS], gsm);
                } else if (kind == KIND_PARTITION) {
                    const int32_t* fz = e.fuse + nd.route_offset;
                    if (e.arr[k].hash_entries)
                        scatter_tile_hash<P, R>(nd, b, t, e.cols, e.C, &e.ctl[CTL_STATUS], gsm, e.arr[k].hash_entries,
                                                e.real_nodes, fz, &e.rp, b.totals + nd.bins);
                    else if (stage)
                        scatter_tile<P, R, true>(nd, b, t, e.cols, e.C, &e.ctl[CTL_STATUS], gsm, e.real_nodes, fz, &e.rp, b.totals + nd.bins);
                    else
                        scatter_tile<P, R, false>(nd, b, t, e.cols, e.C, &e.ctl[CTL_STATUS], gsm, e.real_nodes, fz, &e.rp, b.totals + nd.bins);
                } else {
                    transform_tile<P, R>(nd, b, t, e.cols, e.C, e.rp);
                }
            }
            __syncthreads();
            mark(kind == KIND_SOURCE ? PROF_SOURCE : kind == KIND_SINK ? PROF_SINK : kind == KIND_PARTITION ? PROF_PART : PROF_XF);
        }
        __syncthreads();
        SelLds& L = *reinterpret_cast<SelLds*>(smem);
        engine_retire(e, L);
        mark(PROF_RETIRE);
        engine_select(e, L);
        __syncthreads();
        mark(PROF_SELECT);
    }
}

__global__ __launch_bounds__(kThreads) void k_engine_setup(Engine e, uint32_t do_select)
{
    extern __shared__ __align__(16) char smem[];
    SelLds& L = *reinterpret_cast<SelLds*>(smem);
    const uint32_t tid = threadIdx.x;
    for (uint32_t i = tid; i < e.S; i += kThreads)
        e.free_stack[i] = (unsigned long long)(e.S - 1u - i);  // index 0 first (ChunkLedger.hpp:43-46), generation 0
    for (uint32_t i = tid; i < CNT_WORDS; i += kThreads) e.ctr[i] = 0ull;
    for (uint32_t i = tid; i < CTL_WORDS; i += kThreads) e.ctl[i] = 0u;
    for (uint64_t i = tid; i < (uint64_t)e.route_units_cap * e.n_nodes; i += kThreads) e.rlb[i] = 0ull;
    for (uint64_t i = tid; i < 2ull * e.list_units_cap; i += kThreads) e.llb[i] = 0ull;
    if (e.prof)
        for (uint32_t i = tid; i < PROF_WORDS; i += kThreads) e.prof[i] = 0ull;
    for (uint32_t k = tid; k < e.n_nodes; k += kThreads) {
        NodeSt z{};
        L.st[k] = z;
        e.lb[k] = 0ull;  // epoch 0, no flag: not published
        L.head[k] = 0ull;
        L.tail[k] = 0ull;
    }
    if (tid == 0) {
        L.top = e.S;
        L.generated = L.sunk = L.step = L.batches = 0ull;
        L.part_items = L.xf_items = L.src_items = L.sink_items = 0ull;
    }
    __syncthreads();
    load_lane_tab(e, L);
    load_job_tab(e, L);
    __syncthreads();
    if (do_select) {
        engine_select(e, L);
        return;
    }
    // pre-select state for the team engine: full free stack, empty rings, zero counters
    for (uint32_t k = tid; k < e.n_nodes; k += kThreads) e.st[k] = L.st[k];
    if (tid == 0) e.ctl[CTL_TOP] = e.S;
}

// Continue a run on the wide engine from the pre-select state another engine left behind.
__global__ __launch_bounds__(kThreads) void k_engine_resume(Engine e)
{
    extern __shared__ __align__(16) char smem[];
    SelLds& L = *reinterpret_cast<SelLds*>(smem);
    const uint32_t tid = threadIdx.x;
    load_states(e, L.st);
    if (tid == 0) {
        L.top = e.ctl[CTL_TOP];
        L.generated = e.ctr[CNT_GENERATED];
        L.sunk = e.ctr[CNT_SUNK];
        L.step = e.ctr[CNT_STEP];
        L.batches = e.ctr[CNT_BATCHES];
        L.part_items = e.ctr[CNT_PART_ITEMS];
        L.xf_items = e.ctr[CNT_XF_ITEMS];
        L.src_items = e.ctr[CNT_SRC_ITEMS];
        L.sink_items = e.ctr[CNT_SINK_ITEMS];
        e.ctl[CTL_HANDOFF] = 0u;
    }
    __syncthreads();
    for (uint32_t k = tid; k < e.n_nodes; k += kThreads) {
        L.tail[k] = L.st[k].tail;
        L.head[k] = L.st[k].head;
        L.st[k].active = 0u;
    }
    __syncthreads();
    load_lane_tab(e, L);
    load_job_tab(e, L);
    __syncthreads();
    engine_select(e, L);
}

// ---- dense histogram -> rows ------------------------------------------------------------------
__global__ void k_nz_count(const unsigned long long* c, uint64_t n, uint64_t seg, uint32_t* block_nz)
{
    __shared__ uint32_t red[kWaves];
    uint64_t lo = (uint64_t)blockIdx.x * seg, hi = lo + seg < n ? lo + seg : n;
    uint32_t cnt = 0;
    for (uint64_t i = lo + threadIdx.x; i < hi; i += kThreads) cnt += c[i] != 0ull;
    uint32_t tot;
    block_exscan(cnt, red, &tot);
    if (threadIdx.x == 0) block_nz[blockIdx.x] = tot;
}

__global__ void k_nz_write(const unsigned long long* c, uint64_t n, uint64_t seg, const uint32_t* block_off,
                           uint32_t lo_value, unsigned long long* rows)
{
    __shared__ uint32_t red[kWaves];
    __shared__ uint32_t run;
    uint64_t lo = (uint64_t)blockIdx.x * seg, hi = lo + seg < n ? lo + seg : n;
    if (threadIdx.x == 0) run = block_off[blockIdx.x];
    __syncthreads();
    for (uint64_t i0 = lo; i0 < hi; i0 += kThreads) {
        uint64_t i = i0 + threadIdx.x;
        unsigned long long v = i < hi ? c[i] : 0ull;
        uint32_t tot;
        uint32_t pos = block_exscan(v != 0ull ? 1u : 0u, red, &tot);
        if (v) {
            rows[2ull * (run + pos)] = (unsigned long long)lo_value + i;
            rows[2ull * (run + pos) + 1] = v;
        }
        __syncthreads();
        if (threadIdx.x == 0) run += tot;
        __syncthreads();
    }
}

// ==============================================================================================
// Host objects
// ==============================================================================================
struct gf_context {
    int device = 0;
    hipStream_t stream = nullptr;
    uint32_t cus = 256;
    std::string err;
};

struct gf_pool {
    gf_context* ctx = nullptr;
    uint32_t payload = 0, S = 0, C = 0;
    std::vector<void*> cols;
    std::vector<uint32_t> elem;
    // host ledger (ChunkLedger.hpp:36-134)
    std::vector<uint32_t> free_list, gen, lease;
    uint32_t peak_live = 0, peak_leases = 0;

    Cols dcols() const
    {
        Cols c{};
        for (size_t i = 0; i < cols.size(); ++i) c.c[i] = cols[i];
        return c;
    }
    gf_chunk_handle acquire()
    {
        if (free_list.empty()) fail(GF_POOL_EXHAUSTED, "graphFlow::ChunkPool: pool exhausted");
        uint32_t idx = free_list.back();
        free_list.pop_back();
        peak_live = std::max<uint32_t>(peak_live, S - (uint32_t)free_list.size());
        return gf_chunk_handle{idx, gen[idx]};
    }
    void validate(gf_chunk_handle h) const
    {
        if (h.index >= S) fail(GF_BAD_ARGUMENT, "graphFlow::ChunkPool: chunk index out of range");
        if (h.generation != gen[h.index]) fail(GF_BAD_ARGUMENT, "graphFlow::ChunkPool: stale chunk handle");
    }
    void release(gf_chunk_handle h)
    {
        validate(h);
        if (lease[h.index]) fail(GF_BAD_ARGUMENT, "graphFlow::ChunkPool: release while a reader lease is outstanding");
        if (gen[h.index] == UINT32_MAX) fail(GF_OVERFLOW, "graphFlow::ChunkPool generation overflow");
        gen[h.index] += 1u;
        free_list.push_back(h.index);
    }
    // Return a reserved, never published chunk: back on the free list without a generation bump,
    // so a reservation of k chunks of which the first r were used leaves the ledger as if exactly
    // those r had been acquired (call in reverse acquisition order).
    void cancel(gf_chunk_handle h)
    {
        validate(h);
        if (lease[h.index]) fail(GF_BAD_ARGUMENT, "graphFlow::ChunkPool: cancel while a reader lease is outstanding");
        free_list.push_back(h.index);
    }
    void reset_ledger()
    {
        free_list.resize(S);
        for (uint32_t i = 0; i < S; ++i) free_list[i] = S - 1u - i;
        std::fill(gen.begin(), gen.end(), 0u);
        std::fill(lease.begin(), lease.end(), 0u);
        peak_live = peak_leases = 0;
    }
};

namespace {

template <class T>
T* dalloc(size_t n)
{
    void* p = nullptr;
    HIPX(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
    return (T*)p;
}

template <class T>
T* halloc(size_t n)
{
    void* p = nullptr;
    HIPX(hipHostMalloc(&p, std::max<size_t>(n, 1) * sizeof(T), hipHostMallocDefault));
    return (T*)p;
}

struct DevMem {
    std::vector<void*> dev, host;
    template <class T>
    T* d(size_t n)
    {
        T* p = dalloc<T>(n);
        dev.push_back(p);
        return p;
    }
    template <class T>
    T* h(size_t n)
    {
        T* p = halloc<T>(n);
        host.push_back(p);
        return p;
    }
    void clear()
    {
        for (void* p : dev) (void)hipFree(p);
        for (void* p : host) (void)hipHostFree(p);
        dev.clear();
        host.clear();
    }
    ~DevMem() { clear(); }
};

uint32_t pick_tile_r(uint32_t tile_items)
{
    if (tile_items >= 4096u) return 16u;
    if (tile_items >= 1024u) return 4u;
    return 1u;
}

// Payload dispatch for host launchers.
template <class F>
void with_payload(uint32_t payload, uint32_t r, F&& f)
{
    auto go = [&](auto ptag) {
        using P = typename decltype(ptag)::type;
        if constexpr (std::is_same_v<P, Photon>) {
            if (r >= 4u) f(ptag, std::integral_constant<int, 4>{});
            else f(ptag, std::integral_constant<int, 1>{});
        } else {
            if (r >= 16u) f(ptag, std::integral_constant<int, 16>{});
            else if (r >= 4u) f(ptag, std::integral_constant<int, 4>{});
            else f(ptag, std::integral_constant<int, 1>{});
        }
    };
    if (payload == GF_PAYLOAD_PARTICLE) go(std::type_identity<Particle>{});
    else if (payload == GF_PAYLOAD_KEYED) go(std::type_identity<Keyed>{});
    else go(std::type_identity<Photon>{});
}

uint32_t effective_r(uint32_t payload, uint32_t r)
{
    if (payload == GF_PAYLOAD_PHOTON) return r >= 4u ? 4u : 1u;
    return r >= 16u ? 16u : r >= 4u ? 4u : 1u;
}

// LDS-staged scatter: records of 4-byte fields only (the staged array sits at a 4-byte offset),
// sized by the record struct (Particle 12 B, Photon 36 B)
bool stage_fits(uint32_t bins, uint32_t tile_items, uint32_t rec_bytes)
{
    return rec_bytes % 8u != 0u && bins <= 64u && scatter_smem(bins, tile_items, rec_bytes, true) <= 64u * 1024u;
}

// Guide of the Zipf source's inverse-CDF search (2 words per interval, see RunParams): interval j
// of [0, 1) holds targets in [j, j + 1) / G of the total mass; its search range runs from the first
// rank whose CDF exceeds the lower end less a 1e-12 relative margin to the first rank whose CDF
// exceeds the upper end plus that margin, so the answer of the full search is always inside and the
// range's last rank always exceeds the target (or is the last rank).
std::vector<uint32_t> zipf_guide(const std::vector<double>& cdf)
{
    const uint32_t G = kZipfGuide, n = (uint32_t)cdf.size();
    const double total = cdf[n - 1u];
    std::vector<uint32_t> g(2u * G);
    auto first_above = [&](double v) -> uint32_t {
        return (uint32_t)std::min<size_t>(std::upper_bound(cdf.begin(), cdf.end(), v) - cdf.begin(), n - 1u);
    };
    for (uint32_t j = 0; j < G; ++j) {
        g[2u * j] = first_above((double)j / G * total * (1.0 - 1e-12));
        g[2u * j + 1u] = first_above((double)(j + 1u) / G * total * (1.0 + 1e-12));
    }
    return g;
}

std::vector<double> zipf_table(uint32_t values, double s)
{
    std::vector<double> t(values);
    double acc = 0.0;
    for (uint32_t r = 0; r < values; ++r) {
        acc += std::pow((double)(r + 1u), -s);
        t[r] = acc;
    }
    return t;
}

// Constant A of H(m) ~ A + m^(1-s) / (1-s) + m^-s / 2 (s = 1: A + ln m + 1 / (2m)) fitted to the
// table's last entry, for the device's rank estimate.
double zipf_fit_a(const std::vector<double>& cdf, double s)
{
    const double n = (double)cdf.size();
    if (s == 1.0) return cdf.back() - std::log(n) - 0.5 / n;
    return cdf.back() - std::pow(n, 1.0 - s) / (1.0 - s) - 0.5 * std::pow(n, -s);
}

}  // namespace

// One fused partition batch (gf_batch.hpp) on `st`: one workgroup of as many 256-thread tile groups
// (at most 4) as the LDS holds.
// One launch on one workgroup of G tile groups; with `tiles` > 0 (the batch's tile count, known on the
// host) and more than one round of G tiles, the grid form instead (three launches, a tile per
// workgroup). Both run the same tile functions and plan: equal results.
void launch_fused_kernel(uint32_t payload, uint32_t r, const FusedArgs& a, const Cols& cols, hipStream_t st,
                         uint32_t tiles = 0)
{
    const uint32_t G = std::max<uint32_t>(1u, std::min<uint32_t>(4u, (uint32_t)((160u * 1024u) / a.group_lds)));
    const bool grid = tiles > G && a.n <= kGridIn;
    with_payload(payload, r, [&](auto pt, auto rc) {
        using P = typename decltype(pt)::type;
        constexpr int RR = decltype(rc)::value;
        if (grid) {
            const size_t lds = a.group_lds;
            set_lds(k_fused_count<P, RR>, lds);
            set_lds(k_fused_scatter<P, RR>, lds);
            hipLaunchKernelGGL((k_fused_count<P, RR>), dim3(tiles), dim3(kThreads), lds, st, a, cols);
            hipLaunchKernelGGL(k_fused_plan, dim3(1), dim3(1024), 0, st, a);
            hipLaunchKernelGGL((k_fused_scatter<P, RR>), dim3(tiles), dim3(kThreads), lds, st, a, cols);
        } else {
            const size_t lds = (size_t)G * a.group_lds;
            set_lds(k_batch_fused<P, RR>, lds);
            hipLaunchKernelGGL((k_batch_fused<P, RR>), dim3(1), dim3(kThreads * G), lds, st, a, cols);
        }
    });
    HIPX(hipGetLastError());
}

// ---- no-alloc fused batch ----------------------------------------------------------------------
// The fused kernel's scratch for gf_transform_batch_noalloc: one array per device, held until the
// process ends. Its name is a host stand-in; kernels get the device address that
// hipGetSymbolAddress returns for the current device.
constexpr uint64_t kNoAllocScratchWords = 1ull << 22;  // 16 MiB
__device__ __align__(256) uint32_t scratch[kNoAllocScratchWords];

std::mutex g_scratch_mu;
uint32_t* g_scratch_ptr[64] = {};

uint32_t* scratch_of_device(int device)
{
    if (device < 0 || device >= 64) fail(GF_BAD_ARGUMENT, "graphFlow: device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        int cur = -1;
        HIPX(hipGetDevice(&cur));
        if (cur != device) HIPX(hipSetDevice(device));
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != device) (void)hipSetDevice(cur);
        HIPX(e);
        g_scratch_ptr[device] = (uint32_t*)p;
    }
    return g_scratch_ptr[device];
}

// ---- stage: one node's batch protocol (StageBase, Execution.hpp:1537-1561) ----------------------
struct gf_stage {
    gf_pool* pool = nullptr;
    gf_node_desc nd{};
    uint32_t key = 0, Q = 1, tile_items = 1024, r = 4, bins = 1, cap_in = 1, cap_out = 1, cap_tiles = 1;
    uint32_t reserve_bound = 1;
    uint64_t seed = 0;
    RunParams rp{};
    DevMem mem;
    // One device block per stage with a pinned host mirror of the same layout, so that a launch
    // uploads its batch in one copy and reads its status (and classify counts) back in one:
    //   [Batch | in (cap_in), reserve handles (fused) | tile prefix | item prefix | out (cap_out)
    //    | status (16 B) | counts (cap_tiles x bins)]
    uint8_t* d_blk = nullptr;
    uint8_t* h_blk = nullptr;
    size_t o_in = 0, o_tp = 0, o_out = 0, o_status = 0, o_counts = 0;
    Batch* d_batch = nullptr;
    Batch* h_batch = nullptr;
    uint32_t* d_status = nullptr;
    uint32_t* h_status = nullptr;
    uint32_t* h_counts = nullptr;
    gf_chunk_desc* h_in = nullptr;
    gf_chunk_desc* h_out = nullptr;
    uint32_t* h_prefix = nullptr;
    double* d_zipf = nullptr;
    hipEvent_t ev = nullptr;  // completion of the launch that uses the current mirror
    // The host runner's form (lazy): completion is awaited only where the host needs device results
    // (a classify's counts) or reuses a pinned mirror. Source, transform, sink and scatter launches
    // run ahead on the runner's one stream, where the order of launches keeps every chunk's writer
    // ahead of its readers; sink and scatter failures are read from the stage's status word at the
    // end of the run. Two mirrors alternate, each with its completion event.
    struct Mirror {
        uint8_t* h = nullptr;
        hipEvent_t ev = nullptr;
        bool inflight = false;
    };
    Mirror mir[2];
    uint32_t cur = 0;
    bool lazy = false;
    BatchArg ba{};  // the current launch's batch argument
    bool uploaded = false;
    std::vector<gf_chunk_desc> in, pending;
    std::vector<uint32_t> totals;
    uint32_t stage_no = 0, n_tiles = 0, particles = 0;
    uint64_t first_id = 0;
    bool failed = false, launched = false;
    uint64_t launches = 0;
    uint32_t bin_path = GF_BIN_PATH_AUTO, hash_entries = 0, stall_ms = 0, hash_E = 0;
    size_t lds_count = 0, lds_scatter = 0;
    // fused single-launch form (gf_transform_batch): outputs end where the result words begin, and
    // the scratch (prefixes, totals, bin bases, rows) follows them, so that one copy reads back the
    // used outputs, {R, status}, the prefixes and the bin totals:
    //   [out (cap_out) | result (16 B) | scratch]
    uint8_t* d_fblk = nullptr;
    uint8_t* h_fblk = nullptr;
    size_t o_fres = 0;
    uint32_t* d_result = nullptr;
    uint32_t* d_fscratch = nullptr;
    uint32_t f_n = 0, f_bound = 0;
    std::vector<gf_chunk_handle> reserved;
    bool fused = false, collected = false;
    std::string err;

    ~gf_stage()
    {
        for (auto& m : mir)
            if (m.ev) (void)hipEventDestroy(m.ev);
    }

    int kind() const
    {
        if (nd.role == GF_ROLE_SOURCE) return 0;
        if (nd.role == GF_ROLE_SINK) return 3;
        return is_partition(nd) ? 1 : 2;
    }

    void init(gf_pool* p, const gf_node_desc& d, uint32_t k, uint64_t sd, uint32_t q, uint32_t ti,
              const gf_tuning* t = nullptr)
    {
        pool = p;
        if (t) {
            bin_path = t->bin_path;
            hash_entries = t->hash_table_entries;
            stall_ms = t->stall_timeout_ms;
        }
        nd = d;
        key = k;
        seed = sd;
        Q = q ? q : 1u;
        uint32_t want = ti ? ti : 1024u;
        if (on_hash_path(d, bin_path)) want = std::min(want, kHashTileLimit);
        r = effective_r(p->payload, pick_tile_r(want));
        tile_items = 256u * r;
        hash_E = hash_entries_of(d, bin_path, hash_entries, tile_items);
        const uint32_t rec = record_bytes(p->payload);
        if (kind() == 1) {
            const size_t lc = hash_E ? count_hash_smem(hash_E) : (size_t)std::max<uint32_t>(d.bins, 1u) * 4u;
            const size_t ls = hash_E ? scatter_hash_smem(hash_E)
                                     : scatter_smem(d.bins, tile_items, rec, stage_fits(d.bins, tile_items, rec));
            if (lc > 160u * 1024u || ls > 160u * 1024u)
                fail(GF_BAD_ARGUMENT, "graphFlow: partition width or hash table exceeds the workgroup LDS budget");
            lds_count = lc;
            lds_scatter = ls;
        }
        validate_node(nd, t && t->photon_tally_bins ? t->photon_tally_bins : 1024u);
        bins = nd.role == GF_ROLE_SINK ? 1u : std::max<uint32_t>(nd.bins, 1u);
        if (kind() == 1 && nd.bins == 0) fail(GF_BAD_ARGUMENT, "graphFlow: a partition needs at least one bin");
        cap_in = nd.role == GF_ROLE_SOURCE ? 1u : Q;
        cap_out = nd.role == GF_ROLE_SOURCE ? 1u : nd.role == GF_ROLE_SINK ? 1u : reserve_bound_of(nd, p->C, Q);
        reserve_bound = cap_out;
        cap_tiles = (uint32_t)(ceil_div(p->C, tile_items) * cap_in);
        auto al = [](size_t x) { return (x + 15u) & ~(size_t)15u; };
        const size_t res_bytes = kind() == 1 ? sizeof(gf_chunk_handle) * cap_out : 0u;
        o_in = al(sizeof(Batch));
        o_tp = al(o_in + sizeof(gf_chunk_desc) * cap_in + res_bytes);
        o_out = al(o_tp + 2u * sizeof(uint32_t) * (cap_in + 1u));
        o_status = al(o_out + sizeof(gf_chunk_desc) * cap_out);
        o_counts = o_status + 16u;
        const size_t blk = al(o_counts + sizeof(uint32_t) * (size_t)cap_tiles * bins);
        d_blk = mem.d<uint8_t>(blk);
        for (auto& m : mir) {
            m.h = mem.h<uint8_t>(blk);
            std::memset(m.h, 0, blk);
            HIPX(hipEventCreateWithFlags(&m.ev, hipEventDisableTiming));
        }
        d_batch = (Batch*)d_blk;
        h_batch = (Batch*)mir[0].h;
        h_batch->in = (gf_chunk_desc*)(d_blk + o_in);
        h_batch->tile_prefix = (uint32_t*)(d_blk + o_tp);
        h_batch->item_prefix = h_batch->tile_prefix + (cap_in + 1u);
        h_batch->out = (gf_chunk_desc*)(d_blk + o_out);
        h_batch->counts = (uint32_t*)(d_blk + o_counts);
        h_batch->totals = mem.d<uint32_t>(2u * bins);
        h_batch->tile_items = tile_items;
        h_batch->purpose = 2u * key;
        std::memcpy(mir[1].h, mir[0].h, sizeof(Batch));
        d_status = (uint32_t*)(d_blk + o_status);
        HIPX(hipMemset(d_status, 0, 16u));
        use_mirror(0);
        if (kind() == 1) {
            o_fres = al(sizeof(gf_chunk_desc) * cap_out);
            d_fblk = mem.d<uint8_t>(o_fres + 16u + sizeof(uint32_t) * fused_scratch_words(cap_in, bins, cap_tiles));
            h_fblk = mem.h<uint8_t>(o_fres + 16u + sizeof(uint32_t) * (2u * (cap_in + 1u) + bins));
            d_result = (uint32_t*)(d_fblk + o_fres);
            d_fscratch = (uint32_t*)(d_fblk + o_fres + 16u);
            HIPX(hipMemset(d_result, 0, 16u));  // word 2 of the grid form starts at zero
        }
        rp.seed = seed;
        rp.C = p->C;
        if (nd.op == GF_SRC_KEYED_ZIPF) {
            if (nd.u[0] == 0) fail(GF_BAD_ARGUMENT, "zipf source needs a value count");
            auto t = zipf_table(nd.u[0], (double)nd.f[0]);
            d_zipf = mem.d<double>(t.size());
            HIPX(hipMemcpy(d_zipf, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice));
            auto gd = zipf_guide(t);
            uint32_t* d_guide = mem.d<uint32_t>(gd.size());
            HIPX(hipMemcpy(d_guide, gd.data(), gd.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
            rp.zipf = d_zipf;
            rp.zipf_values = nd.u[0];
            rp.zipf_guide = d_guide;
            rp.zipf_s = (double)nd.f[0];
            rp.zipf_a = zipf_fit_a(t, rp.zipf_s);
        }
    }

    void use_mirror(uint32_t i)
    {
        cur = i;
        h_blk = mir[i].h;
        ev = mir[i].ev;
        h_batch = (Batch*)h_blk;
        h_status = (uint32_t*)(h_blk + o_status);
        h_counts = (uint32_t*)(h_blk + o_counts);
        h_in = (gf_chunk_desc*)(h_blk + o_in);
        h_out = (gf_chunk_desc*)(h_blk + o_out);
        h_prefix = (uint32_t*)(h_blk + o_tp);
    }

    // Mark the end of the current launch on the stream.
    void mark(hipStream_t st)
    {
        HIPX(hipEventRecord(ev, st));
        mir[cur].inflight = true;
    }

    // Wait for the launched device stage. With a stall timeout (Tuning::stallTimeout,
    // GraphRunner.hpp:1742) a stage that has not completed in time fails the run with GF_STALL.
    void wait_stage()
    {
        wait_event(ev);
        mir[cur].inflight = false;
    }

    void wait_event(hipEvent_t e)
    {
        if (!stall_ms) {
            HIPX(hipEventSynchronize(e));
            return;
        }
        const auto t0 = std::chrono::steady_clock::now();
        for (;;) {
            const hipError_t q = hipEventQuery(e);
            if (q == hipSuccess) return;
            if (q != hipErrorNotReady) HIPX(q);
            if (std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(stall_ms))
                fail(GF_STALL, "graphFlow: scheduler stall, a device stage did not complete within the stall timeout");
            std::this_thread::yield();
        }
    }

    void throw_if_failed() const
    {
        if (failed) fail(GF_STAGE_FAILED, "graphFlow: stage is single-use after a failed batch");
    }

    // Complete mirror i's launch.
    void settle(uint32_t i)
    {
        Mirror& m = mir[i];
        if (!m.inflight) return;
        wait_event(m.ev);
        m.inflight = false;
    }

    // Lazy form: before a launch writes a mirror, the launch that used it last is complete.
    void prepare_launch()
    {
        if (!lazy) return;
        settle(cur ^ 1u);
        use_mirror(cur ^ 1u);
    }

    // Lazy form, end of a run: every launch complete, and the status of sink and scatter launches
    // (a classify's is checked when its counts are read) checked.
    void settle_all(hipStream_t st)
    {
        settle(cur ^ 1u);
        settle(cur);
        if (lazy && (kind() == 1 || kind() == 3)) {
            HIPX(hipMemcpyAsync(h_status, d_status, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
            HIPX(hipStreamSynchronize(st));
            check_status(kind() == 3 ? "sink" : "partition scatter");
        }
    }

    // The launch's batch: in the kernel arguments when it fits, else uploaded (record, inputs,
    // prefixes, the first n_out output descriptors, the first count_words planned offsets). The status word is never uploaded: cleared once, it stays zero
    // until a kernel reports a failure, after which the stage takes no further batch. `uploaded`
    // tells whether the launch reads the current mirror.
    void stage_batch(hipStream_t st, uint32_t n_out, size_t count_words = 0)
    {
        uploaded = false;
        const uint32_t n = h_batch->n_in;
        if (n <= kArgIn && n_out <= kArgOut && count_words <= kArgCounts) {
            ba.block = nullptr;
            ba.b = *h_batch;
            ba.inl = ARG_IN | (n_out ? ARG_OUT : 0u) | (count_words ? ARG_COUNTS : 0u);
            std::memcpy(ba.in, h_in, sizeof(gf_chunk_desc) * n);
            std::memcpy(ba.prefix, h_prefix, sizeof(uint32_t) * (n + 1u));
            std::memcpy(ba.prefix + kArgIn + 1u, h_prefix + cap_in + 1u, sizeof(uint32_t) * (n + 1u));
            if (n_out) std::memcpy(ba.out, h_out, sizeof(gf_chunk_desc) * n_out);
            if (count_words) std::memcpy(ba.counts, h_counts, sizeof(uint32_t) * count_words);
            return;
        }
        ba.block = d_batch;
        ba.inl = 0u;
        uploaded = true;
        HIPX(hipMemcpyAsync(d_blk, h_blk, o_out + sizeof(gf_chunk_desc) * n_out, hipMemcpyHostToDevice, st));
        if (count_words)
            HIPX(hipMemcpyAsync(d_blk + o_counts, h_blk + o_counts, sizeof(uint32_t) * count_words,
                                hipMemcpyHostToDevice, st));
    }

    // Read back the status word and the first count_words counts, then mark the stage's event.
    void download_status(hipStream_t st, size_t count_words = 0)
    {
        HIPX(hipMemcpyAsync(h_blk + o_status, d_blk + o_status, 16u + sizeof(uint32_t) * count_words,
                            hipMemcpyDeviceToHost, st));
        mark(st);
    }

    void check_status(const char* what) { check_code(*h_status, what); }

    void check_code(uint32_t s, const char* what)
    {
        if (s == kStatusTableFull) {
            failed = true;
            fail(GF_INVALID, std::string(what) + ": hash bin table full");
        }
        if (s != 0u) {
            failed = true;
            const char* name = s == GF_INVALID ? "invalid" : s == GF_OVERFLOW ? "overflow" : "invalid";
            fail(s == GF_OVERFLOW ? GF_OVERFLOW : GF_INVALID, std::string(what) + ": device status reports " + name);
        }
    }

    void release_pending()
    {
        for (auto& d : pending) {
            try {
                pool->release(gf_chunk_handle{d.index, d.generation});
            } catch (...) {
            }
        }
        pending.clear();
    }

    void cancel_reserved()
    {
        for (size_t i = reserved.size(); i-- > 0;) {
            try {
                pool->cancel(reserved[i]);
            } catch (...) {
            }
        }
        reserved.clear();
    }

    // ---- fused form: one launch per partition batch (gf_batch.hpp); other stages as launch()
    void launch_fused(const gf_chunk_desc* chunks, uint32_t n, uint64_t fid, hipStream_t st)
    {
        if (kind() != 1) {
            launch(chunks, n, fid, st);
            return;
        }
        throw_if_failed();
        if (n == 0) fail(GF_PLAN_MISMATCH, "graphFlow: empty batch");
        if (n > cap_in) fail(GF_PLAN_MISMATCH, "graphFlow: batch exceeds batchChunks");
        prepare_launch();
        const uint32_t C = pool->C;
        in.assign(chunks, chunks + n);
        first_id = fid;
        pending.clear();
        particles = 0;
        n_tiles = 0;
        for (uint32_t c = 0; c < n; ++c) {
            if (chunks[c].fill > C) fail(GF_PLAN_MISMATCH, "graphFlow: chunk fill exceeds the chunk capacity");
            h_in[c] = chunks[c];
            n_tiles += (uint32_t)ceil_div(chunks[c].fill, tile_items);
            particles += chunks[c].fill;
        }
        if (particles == 0) fail(GF_PLAN_MISMATCH, "partition: empty batch");
        // the batch's reservation: outputChunkBound handles, popped in LIFO order
        const uint64_t bound = output_chunk_bound(particles, C, bins);
        if (bound > reserve_bound) fail(GF_PLAN_MISMATCH, "partition plan: reservation exceeds the stage's bound");
        const uint32_t peak0 = pool->peak_live;
        reserved.clear();
        try {
            for (uint64_t r = 0; r < bound; ++r) {
                reserved.push_back(pool->acquire());
                ((gf_chunk_handle*)(h_in + n))[r] = reserved.back();
            }
        } catch (...) {
            cancel_reserved();
            pool->peak_live = peak0;
            throw;
        }
        pool->peak_live = peak0;
        FusedArgs a{};
        if (n <= kArgIn && bound <= kArgRes) {
            // inputs and reservation in the launch's arguments
            a.inl = 1u;
            std::memcpy(a.in_arg, h_in, sizeof(gf_chunk_desc) * n);
            std::memcpy(a.res_arg, h_in + n, sizeof(gf_chunk_handle) * bound);
        } else {
            // inputs and the reservation's handles right after them: one upload
            HIPX(hipMemcpyAsync(d_blk + o_in, h_blk + o_in,
                                sizeof(gf_chunk_desc) * n + sizeof(gf_chunk_handle) * bound, hipMemcpyHostToDevice, st));
        }
        a.nd = nd;
        a.in = h_batch->in;
        a.n = n;
        a.tile_items = tile_items;
        a.C = C;
        a.hash_entries = hash_E;
        a.reserved = (const gf_chunk_handle*)(h_batch->in + n);
        a.n_reserved = (uint32_t)bound;
        a.purpose = 2u * key;
        a.out = (gf_chunk_desc*)(d_fblk + o_fres) - bound;
        a.scratch = d_fscratch;
        a.result = d_result;
        a.group_lds = (uint32_t)((std::max(lds_count, lds_scatter) + 15u) & ~(size_t)15u);
        a.tiles_cap = cap_tiles;
        launch_fused_kernel(pool->payload, r, a, pool->dcols(), st, n_tiles);
        // the used outputs, {R, status}, prefixes and bin totals: one readback
        const size_t o0 = o_fres - sizeof(gf_chunk_desc) * bound;
        HIPX(hipMemcpyAsync(h_fblk + o0, d_fblk + o0,
                            sizeof(gf_chunk_desc) * bound + 16u + sizeof(uint32_t) * (2u * (n + 1u) + bins),
                            hipMemcpyDeviceToHost, st));
        mark(st);
        f_n = n;
        f_bound = (uint32_t)bound;
        fused = true;
        launched = true;
        ++launches;
    }

    // Completion of a fused batch: the outputs it used, the unused reservation returned.
    void fused_collect(uint32_t& allocated)
    {
        wait_stage();
        fused = false;
        const uint32_t* res = (const uint32_t*)(h_fblk + o_fres);
        const gf_chunk_desc* outs = (const gf_chunk_desc*)(h_fblk + o_fres) - f_bound;
        const uint32_t status = res[1], R = res[0];
        if (status) {
            cancel_reserved();
            failed = true;
            if (status == kStatusTableFull) fail(GF_INVALID, "partition classify: hash bin table full");
            if (status == GF_OVERFLOW) fail(GF_OVERFLOW, "partition scatter: device status reports overflow");
            if (status == GF_PLAN_MISMATCH)
                fail(GF_PLAN_MISMATCH, "partition plan: classified items or output chunks do not match the batch");
            fail(GF_INVALID, "partition: device status reports an invalid classifier or updater result");
        }
        if (R > reserved.size()) {
            cancel_reserved();
            failed = true;
            fail(GF_PLAN_MISMATCH, "partition plan: exact chunk count exceeds the reservation bound");
        }
        const uint32_t* tot = res + 4u + 2u * (f_n + 1u);
        totals.assign(tot, tot + bins);
        pending.assign(outs, outs + R);
        for (size_t i = reserved.size(); i-- > R;) pool->cancel(reserved[i]);
        reserved.clear();
        pool->peak_live = std::max<uint32_t>(pool->peak_live, pool->S - (uint32_t)pool->free_list.size());
        allocated = R;
    }

    // ---- launch (Execution.hpp:2363-2430 / 1805-1901 / 1615-1637 / 2809-2828)
    void launch(const gf_chunk_desc* chunks, uint32_t n, uint64_t fid, hipStream_t st)
    {
        throw_if_failed();
        if (n == 0) fail(GF_PLAN_MISMATCH, "graphFlow: empty batch");
        if (n > cap_in) fail(GF_PLAN_MISMATCH, "graphFlow: batch exceeds batchChunks");
        prepare_launch();
        in.assign(chunks, chunks + n);
        first_id = fid;
        stage_no = 0;
        pending.clear();
        particles = 0;
        n_tiles = 0;
        for (uint32_t c = 0; c < n; ++c) {
            h_in[c] = chunks[c];
            h_prefix[c] = n_tiles;
            h_prefix[cap_in + 1u + c] = particles;
            n_tiles += (uint32_t)ceil_div(chunks[c].fill, tile_items);
            particles += chunks[c].fill;
        }
        h_prefix[n] = n_tiles;
        h_prefix[cap_in + 1u + n] = particles;
        h_batch->n_in = n;
        h_batch->n_tiles = n_tiles;
        h_batch->particles = particles;
        h_batch->first_id = fid;
        h_batch->digest_base = 0;
        Cols cols = pool->dcols();
        const uint32_t C = pool->C;
        const int kd = kind();
        if (kd == 0) {
            // the source fills the admitted chunk it is handed (SourceStage::launch)
            h_out[0] = chunks[0];
            h_batch->n_out = 1;
            stage_batch(st, 1);
            uint32_t tiles = (uint32_t)ceil_div(chunks[0].fill, tile_items);
            if (tiles)
                with_payload(pool->payload, r, [&](auto pt, auto) {
                    using P = typename decltype(pt)::type;
                    hipLaunchKernelGGL(k_source<P>, dim3(tiles), dim3(kThreads), 0, st, ba, nd, cols, C, rp);
                });
            pending.assign(chunks, chunks + 1);
        } else if (kd == 3) {
            h_batch->n_out = 0;
            stage_batch(st, 0);
            if (n_tiles)
                with_payload(pool->payload, r, [&](auto pt, auto) {
                    using P = typename decltype(pt)::type;
                    hipLaunchKernelGGL(k_sink<P>, dim3(n_tiles), dim3(kThreads), 64, st, ba, nd, cols, C, d_status);
                });
        } else if (kd == 2) {
            uint64_t R = ceil_div(particles, C);
            if (particles == 0) fail(GF_PLAN_MISMATCH, "transform: empty batch");
            if (R > reserve_bound) fail(GF_PLAN_MISMATCH, "transform: exact chunk count exceeds the reservation bound");
            for (uint32_t k = 0; k < R; ++k) {
                gf_chunk_handle h;
                try {
                    h = pool->acquire();
                } catch (...) {
                    release_pending();
                    failed = true;
                    throw;
                }
                uint64_t rem = particles - (uint64_t)k * C;
                gf_chunk_desc d{h.index, h.generation, (uint32_t)std::min<uint64_t>(rem, C), 0u};
                pending.push_back(d);
                h_out[k] = d;
            }
            h_batch->n_out = (uint32_t)R;
            stage_batch(st, (uint32_t)R);
            with_payload(pool->payload, r, [&](auto pt, auto rc) {
                using P = typename decltype(pt)::type;
                hipLaunchKernelGGL((k_transform<P, decltype(rc)::value>), dim3(n_tiles), dim3(kThreads), 0, st, ba,
                                   nd, cols, C, rp);
            });
        } else {
            h_batch->n_out = 0;
            stage_batch(st, 0);
            with_payload(pool->payload, r, [&](auto pt, auto rc) {
                using P = typename decltype(pt)::type;
                set_lds(k_count<P, decltype(rc)::value>, lds_count);
                hipLaunchKernelGGL((k_count<P, decltype(rc)::value>), dim3(n_tiles), dim3(kThreads), lds_count, st, ba,
                                   nd, cols, C, d_status, hash_E);
            });
        }
        HIPX(hipGetLastError());
        // the status word, and for a classify the [tile][bin] counts right after it; source and
        // transform kernels write no status, so the cleared host word stands
        if (kd == 1 || (kd == 3 && !lazy)) download_status(st, kd == 1 ? (size_t)n_tiles * bins : 0u);
        else if (!lazy || uploaded) mark(st);  // lazy: an event only where the launch reads the mirror
        launched = true;
        ++launches;
    }

    // ---- advance (Execution.hpp:2432-2455, 2520-2642, 2715-2725)
    bool advance(hipStream_t st, uint32_t& allocated)
    {
        throw_if_failed();
        allocated = 0;
        if (fused) {
            fused_collect(allocated);
            return false;
        }
        if (collected) {
            // the fused batch was collected at dispatch (host runner): report its outputs now
            collected = false;
            allocated = (uint32_t)pending.size();
            return false;
        }
        const int kd = kind();
        if (lazy && !(kd == 1 && stage_no == 0)) {
            // outputs known on the host: complete without waiting (failures are read at the end)
            if (kd == 2) allocated = (uint32_t)pending.size();
            if (kd == 1) stage_no = 0;
            return false;
        }
        wait_stage();
        if (kd == 0) {
            check_status("source");
            return false;
        }
        if (kd == 3) {
            check_status("sink");
            return false;
        }
        if (kd == 2) {
            try {
                check_status("transform");
            } catch (...) {
                release_pending();
                throw;
            }
            allocated = (uint32_t)pending.size();
            return false;
        }
        if (stage_no == 0) {
            try {
                plan_and_scatter(st, allocated);
            } catch (...) {
                release_pending();
                failed = true;
                throw;
            }
            stage_no = 1;
            return true;
        }
        try {
            check_status("partition scatter");
        } catch (...) {
            release_pending();
            throw;
        }
        stage_no = 0;
        return false;
    }

    void plan_and_scatter(hipStream_t st, uint32_t& allocated)
    {
        check_status("partition classify");
        const uint32_t B = bins, C = pool->C;
        totals.assign(B, 0u);
        uint64_t sum = 0;
        for (uint32_t t = 0; t < n_tiles; ++t)
            for (uint32_t b = 0; b < B; ++b) totals[b] += h_counts[(size_t)t * B + b];
        for (uint32_t b = 0; b < B; ++b) sum += totals[b];
        if (sum != particles) fail(GF_PLAN_MISMATCH, "partition classify: classified particle count does not match the batch");
        // planDenseBins
        std::vector<uint32_t> base(B);
        uint64_t exact = 0;
        for (uint32_t b = 0; b < B; ++b) {
            base[b] = (uint32_t)exact;
            exact += ceil_div(totals[b], C);
        }
        if (exact > output_chunk_bound(particles, C, B)) fail(GF_PLAN_MISMATCH, "planDenseBins: exact chunk count exceeds the reservation bound");
        if (exact > reserve_bound) fail(GF_PLAN_MISMATCH, "partition plan: exact chunk count exceeds the reservation bound");
        pending.clear();
        for (uint32_t b = 0; b < B; ++b) {
            uint32_t nch = (uint32_t)ceil_div(totals[b], C);
            for (uint32_t k = 0; k < nch; ++k) {
                gf_chunk_handle h = pool->acquire();
                uint32_t rem = totals[b] - k * C;
                pending.push_back(gf_chunk_desc{h.index, h.generation, std::min(rem, C), b});
            }
        }
        for (size_t i = 0; i < pending.size(); ++i) h_out[i] = pending[i];
        // flat output start of every (tile, bin): bin base + items of that bin in earlier tiles
        std::vector<uint32_t> run(B, 0u);
        for (uint32_t t = 0; t < n_tiles; ++t)
            for (uint32_t b = 0; b < B; ++b) {
                uint32_t x = h_counts[(size_t)t * B + b];
                h_counts[(size_t)t * B + b] = base[b] * C + run[b];
                run[b] += x;
            }
        h_batch->n_out = (uint32_t)pending.size();
        stage_batch(st, h_batch->n_out, (size_t)n_tiles * B);
        Cols cols = pool->dcols();
        const bool stage = !hash_E && stage_fits(B, tile_items, record_bytes(pool->payload));
        const size_t lds = lds_scatter;
        with_payload(pool->payload, r, [&](auto pt, auto rc) {
            using P = typename decltype(pt)::type;
            constexpr int RR = decltype(rc)::value;
            if (stage) {
                set_lds(k_scatter<P, RR, true>, lds);
                hipLaunchKernelGGL((k_scatter<P, RR, true>), dim3(n_tiles), dim3(kThreads), lds, st, ba, nd, cols, C,
                                   d_status, hash_E);
            } else {
                set_lds(k_scatter<P, RR, false>, lds);
                hipLaunchKernelGGL((k_scatter<P, RR, false>), dim3(n_tiles), dim3(kThreads), lds, st, ba, nd, cols,
                                   C, d_status, hash_E);
            }
        });
        HIPX(hipGetLastError());
        if (!lazy) download_status(st);
        else if (uploaded) mark(st);
        allocated = (uint32_t)pending.size();
        ++launches;
    }

    void finish(std::vector<gf_chunk_desc>& outputs)
    {
        outputs.insert(outputs.end(), pending.begin(), pending.end());
        pending.clear();
    }

    void abort()
    {
        if (fused) {
            (void)hipEventSynchronize(ev);
            fused = false;
        }
        cancel_reserved();
        collected = false;
        if (kind() != 0) release_pending();
        pending.clear();
        failed = true;
    }
};

// ---- flow ----------------------------------------------------------------------------------------
struct NodeRec {
    std::string name;
    gf_node_desc nd{};
    std::vector<int32_t> routes;
    uint32_t slot = 0, local = 0;  // served flows: the job slot whose graph copy holds the node, its id there
    double* tally_copies = nullptr;  // photon steps: kTallyCopies private tallies of tally_copy_bins
    uint32_t tally_copy_bins = 0;
};

struct EngineBuffers {
    DevMem mem;
    std::string key;
    Engine e{};
    uint32_t stage_mask = 0, wave_sinks = 0, r = 4, grid = 256, route_grid = 1;
    size_t lds_a = 0, lds_b = 0;
    bool micro = false;
    uint32_t group_lds = 0;
    size_t micro_lds = 0;
    // team engine
    bool team = false;
    TeamArgs ta{};
    TeamArgs* ta_dev = nullptr;  // the launch's arguments, read by the kernel from device memory
    size_t cnt_bytes = 0;
    hipGraphExec_t exec = nullptr;
    uint32_t graph_steps = 0;
    ~EngineBuffers()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
    }
};

struct gf_flow {
    gf_context* ctx = nullptr;
    uint32_t payload = 0;
    uint32_t slots = 0;  // served flows: copies of one topology's graph, one per job slot (0: a plain flow)
    std::vector<NodeRec> nodes;
    gf_tuning tuning{};
    std::string err;
    std::vector<uint32_t> trace;
    std::unique_ptr<gf_pool> pool;
    std::unique_ptr<EngineBuffers> eng;
    // kernel time per kind (gf_flow_kernel_times): classify, plan, scatter + retire/select, persistent
    // engines, then the wide engine's partition, source and sink scatter kernels and its retire/select
    double kt_ms[kKernelKinds] = {};
    uint64_t kt_n[kKernelKinds] = {};
    uint64_t part_items = 0, xf_items = 0;
};

namespace {

void check_graph(const gf_flow& f, uint32_t source)
{
    for (const auto& n : f.nodes) validate_node(n.nd, f.tuning.photon_tally_bins ? f.tuning.photon_tally_bins : 1024u);
    if (source >= f.nodes.size() || f.nodes[source].nd.role != GF_ROLE_SOURCE)
        fail(GF_BAD_ARGUMENT, "graphFlow: run() requires the registered source node");
    uint32_t sources = 0;
    for (const auto& n : f.nodes) {
        if (n.nd.role == GF_ROLE_SOURCE) ++sources;
        if (n.nd.role == GF_ROLE_SINK) continue;
        if (n.nd.bins == 0) fail(GF_BAD_ARGUMENT, "graphFlow: a source or transform node needs at least one bin");
        for (size_t b = 0; b < n.routes.size(); ++b) {
            int32_t d = n.routes[b];
            if (d < 0)
                fail(GF_BAD_ARGUMENT, "graphFlow: missing route for node '" + n.name + "' bin " + std::to_string(b));
            if ((size_t)d >= f.nodes.size()) fail(GF_BAD_ARGUMENT, "graphFlow: route destination out of range");
        }
    }
    if (!f.slots && sources != 1) fail(GF_BAD_ARGUMENT, "graphFlow: a graph needs exactly one source node");
    if (f.slots) {
        // a served flow: one source per job slot, and no route leaves its slot's copy of the graph
        std::vector<uint32_t> per(f.slots, 0u);
        for (const auto& n : f.nodes) {
            if (n.slot >= f.slots) fail(GF_BAD_ARGUMENT, "graphFlow: node outside the served flow's slots");
            if (n.nd.role == GF_ROLE_SOURCE) ++per[n.slot];
            for (int32_t d : n.routes)
                if (f.nodes[(size_t)d].slot != n.slot) fail(GF_BAD_ARGUMENT, "graphFlow: a route crosses job slots");
        }
        for (uint32_t c : per)
            if (c != 1u) fail(GF_BAD_ARGUMENT, "graphFlow: every job slot needs exactly one source node");
        if (f.slots > gfd::kMaxJobs) fail(GF_BAD_ARGUMENT, "graphFlow: too many job slots");
    }
}

struct RunShape {
    uint32_t C, P, S, Q, tile_items;
};

RunShape resolve(const gf_flow& f, uint64_t items)
{
    RunShape s{};
    s.C = f.tuning.chunk_capacity ? f.tuning.chunk_capacity : 8192u;
    s.Q = f.tuning.batch_chunks ? f.tuning.batch_chunks : 1u;
    uint64_t P = f.tuning.live_particles ? f.tuning.live_particles : 8ull * s.C;
    if (P > items) P = items;
    if (P == 0) P = 1;
    s.P = (uint32_t)P;
    uint64_t reserve = 0;
    uint32_t maxb = 1;
    for (const auto& n : f.nodes) {
        if (n.nd.role != GF_ROLE_TRANSFORM) continue;
        reserve += output_chunk_bound(s.C, s.C, n.nd.bins);
        maxb = std::max(maxb, n.nd.bins);
    }
    uint64_t S = items == 0 ? 1u : P + reserve;
    if (f.tuning.num_chunks) S = f.tuning.num_chunks;
    if (S > UINT32_MAX || S * (uint64_t)s.C > UINT32_MAX) fail(GF_OVERFLOW, "graphFlow: chunk pool would exceed 2^32 slots");
    s.S = (uint32_t)S;
    // autotuned tile: a workgroup tile must amortise its [tile][bin] metadata for wide partitions
    uint32_t ti = f.tuning.tile_items;
    if (!ti) ti = maxb >= 64u ? 4096u : 1024u;
    for (const auto& n : f.nodes)
        if (on_hash_path(n.nd, f.tuning.bin_path)) ti = std::min(ti, kHashTileLimit);
    s.tile_items = 256u * effective_r(f.payload, pick_tile_r(ti));
    return s;
}

void ensure_pool(gf_flow& f, const RunShape& s)
{
    if (f.pool && f.pool->S == s.S && f.pool->C == s.C && f.pool->payload == f.payload) {
        f.pool->reset_ledger();
        return;
    }
    if (f.pool) {
        for (void* p : f.pool->cols) (void)hipFree(p);
        f.pool.reset();
    }
    auto p = std::make_unique<gf_pool>();
    p->ctx = f.ctx;
    p->payload = f.payload;
    p->S = s.S;
    p->C = s.C;
    p->elem = column_bytes(f.payload);
    for (uint32_t b : p->elem) p->cols.push_back(dalloc<char>((size_t)s.S * s.C * b));
    p->gen.assign(s.S, 0u);
    p->lease.assign(s.S, 0u);
    p->reset_ledger();
    f.pool = std::move(p);
}

void trace_host(std::vector<uint32_t>& tr, bool on, uint32_t node, uint64_t step, const std::vector<gf_chunk_desc>& in,
                uint32_t nbins, const uint32_t* counts, const std::vector<gf_chunk_desc>& out)
{
    if (!on) return;
    tr.push_back(0xB47C0000u | node);
    tr.push_back((uint32_t)step);
    tr.push_back((uint32_t)in.size());
    for (auto& d : in) {
        tr.push_back(d.index);
        tr.push_back(d.generation);
        tr.push_back(d.fill);
    }
    tr.push_back(nbins);
    for (uint32_t i = 0; i < nbins; ++i) tr.push_back(counts[i]);
    tr.push_back((uint32_t)out.size());
    for (auto& d : out) {
        tr.push_back(d.index);
        tr.push_back(d.generation);
        tr.push_back(d.fill);
        tr.push_back(d.bin);
    }
}

// Plain form: GraphRunner::run with every batch completed before the next loop iteration.
// Photon step parameters of a run (the slab of the photon transport flow). Zero layers, tally
// bins, layer thickness or mu_t take the defaults 10, 1024, 1 mfp and 1.
PhotonParams photon_params_of(const gf_tuning& t)
{
    const uint32_t layers = t.photon_layers ? t.photon_layers : 10u;
    const double mfp = t.photon_layer_mfp > 0 ? t.photon_layer_mfp : 1.0, mu_t = t.photon_mu_t > 0 ? t.photon_mu_t : 1.0;
    PhotonParams ph;
    ph.L = (float)(layers * mfp / mu_t);
    ph.mu_t = (float)mu_t;
    ph.albedo = (float)t.photon_albedo;
    ph.g = (float)t.photon_hg_g;
    ph.roulette_w = (float)t.photon_roulette_w;
    ph.roulette_p = (float)t.photon_roulette_p;
    ph.tally_bins = t.photon_tally_bins ? t.photon_tally_bins : 1024u;
    return ph;
}

// The slab of the photon transport workload (10 layers of 1 mfp, mu_t 1, albedo 0.9, g 0.9,
// roulette below 1e-4 with survival 1/10): what a stand-alone photon stage runs with.
gf_tuning default_photon_tuning()
{
    gf_tuning t{};
    t.photon_layers = 10u;
    t.photon_tally_bins = 1024u;
    t.photon_layer_mfp = 1.0;
    t.photon_mu_t = 1.0;
    t.photon_albedo = 0.9;
    t.photon_hg_g = 0.9;
    t.photon_roulette_w = 1e-4;
    t.photon_roulette_p = 0.1;
    return t;
}

void run_host_form(gf_flow& f, uint32_t src, uint64_t N, const RunShape& sh, gf_stats& st)
{
    gf_pool& pool = *f.pool;
    const uint32_t n = (uint32_t)f.nodes.size(), C = sh.C, P = sh.P, Q = sh.Q;
    // lanes (Tuning::inflightBatches, GraphRunner.hpp:766-771): each lane owns its stage, batch,
    // leases and output credit; lane j of node k is lanes[k * D + j]
    const uint32_t D = f.tuning.inflight_batches ? f.tuning.inflight_batches : 1u;
    const bool fused = f.tuning.fused_batches != 0u;
    std::vector<std::unique_ptr<gf_stage>> stages((size_t)n * D);
    for (uint32_t k = 0; k < n; ++k)
        for (uint32_t j = 0; j < D; ++j) {
            auto& sg = stages[(size_t)k * D + j];
            sg = std::make_unique<gf_stage>();
            sg->init(&pool, f.nodes[k].nd, k, f.tuning.seed, Q, sh.tile_items, &f.tuning);
            sg->rp.ph = photon_params_of(f.tuning);
            sg->lazy = true;
        }
    hipStream_t s = f.ctx->stream;
    struct Lane {
        bool armed = false;
        std::vector<gf_chunk_desc> batch;
        uint32_t reserved = 0;
        uint64_t particles = 0;
    };
    std::vector<Lane> lanes((size_t)n * D);
    auto free_lane = [&](uint32_t k) -> int {
        for (uint32_t j = 0; j < D; ++j)
            if (!lanes[(size_t)k * D + j].armed) return (int)j;
        return -1;
    };
    std::vector<std::deque<gf_chunk_desc>> ready(n);
    uint64_t generated = 0, sunk = 0, iter = 0, batches = 0;
    uint32_t reserved = 0, rr = 0;
    const bool tr = f.tuning.trace != 0;
    f.trace.clear();
    f.part_items = f.xf_items = 0;
    auto teardown = [&]() {
        (void)hipStreamSynchronize(s);
        for (size_t i = 0; i < lanes.size(); ++i) {
            stages[i]->abort();
            if (lanes[i].armed)
                for (auto& d : lanes[i].batch) {
                    try {
                        pool.release(gf_chunk_handle{d.index, d.generation});
                    } catch (...) {
                    }
                }
            lanes[i].armed = false;
        }
        for (uint32_t k = 0; k < n; ++k) {
            for (auto& d : ready[k]) {
                try {
                    pool.release(gf_chunk_handle{d.index, d.generation});
                } catch (...) {
                }
            }
            ready[k].clear();
        }
    };
    try {
        for (;;) {
            bool progress = false;
            // 1) retire, every lane of every node
            for (uint32_t k = 0; k < n; ++k)
                for (uint32_t j = 0; j < D; ++j) {
                    Lane& L = lanes[(size_t)k * D + j];
                    if (!L.armed) continue;
                    gf_stage& stg = *stages[(size_t)k * D + j];
                    const gf_node_desc& nd = f.nodes[k].nd;
                    uint32_t allocated = 0;
                    bool more = stg.advance(s, allocated);
                    if (stg.kind() == 1 && more)
                        trace_host(f.trace, tr, k, iter, L.batch, nd.bins, stg.totals.data(), stg.pending);
                    if (allocated) {
                        if (allocated > L.reserved) fail(GF_BAD_ARGUMENT, "graphFlow: node acquired more chunks than reserved");
                        reserved -= L.reserved;
                        L.reserved = 0;
                    } else if (!more && L.reserved) {
                        reserved -= L.reserved;
                        L.reserved = 0;
                    }
                    if (!more) {
                        std::vector<gf_chunk_desc> outs;
                        stg.finish(outs);
                        if (nd.role != GF_ROLE_SINK)
                            for (auto& d : outs) {
                                if (d.bin >= f.nodes[k].routes.size()) fail(GF_BAD_ARGUMENT, "graphFlow: node produced an unroutable bin");
                                ready[(uint32_t)f.nodes[k].routes[d.bin]].push_back(d);
                            }
                        if (nd.role == GF_ROLE_SINK) sunk += L.particles;
                        if (nd.role != GF_ROLE_SOURCE)
                            for (auto& d : L.batch) pool.release(gf_chunk_handle{d.index, d.generation});
                        if (nd.role == GF_ROLE_TRANSFORM) {
                            ++batches;
                            if (is_partition(nd)) f.part_items += L.particles;
                            else f.xf_items += L.particles;
                        }
                        L.armed = false;
                    }
                    progress = true;
                }
            // 2) admission into the source's first free lane
            {
                uint64_t live = generated - sunk, remaining = N - generated;
                uint64_t min_fill = std::min<uint64_t>(std::min<uint64_t>(C, P), remaining);
                bool within = live < P, budget = live < P && (P - live) >= min_fill;
                const int j = free_lane(src);
                if (generated < N && j >= 0 && within && budget && pool.free_list.size() > reserved) {
                    uint64_t fill = std::min<uint64_t>(std::min<uint64_t>(remaining, P - live), C);
                    gf_chunk_handle h = pool.acquire();
                    gf_chunk_desc d{h.index, h.generation, (uint32_t)fill, 0u};
                    Lane& L = lanes[(size_t)src * D + (uint32_t)j];
                    L.batch = {d};
                    L.particles = fill;
                    L.armed = true;
                    stages[(size_t)src * D + (uint32_t)j]->launch(&d, 1, generated, s);
                    trace_host(f.trace, tr, src, iter, {}, 0, nullptr, L.batch);
                    generated += fill;
                    progress = true;
                }
            }
            // 3) dispatch one batch into a free lane: sinks first, then round robin
            {
                bool dispatched = false;
                for (int pass = 0; pass < 2 && !dispatched; ++pass)
                    for (uint32_t kk = 0; kk < n && !dispatched; ++kk) {
                        uint32_t k = pass == 0 ? kk : (rr + kk) % n;
                        const gf_node_desc& nd = f.nodes[k].nd;
                        if (pass == 0 && nd.role != GF_ROLE_SINK) continue;
                        if (ready[k].empty()) continue;
                        const int j = free_lane(k);
                        if (j < 0) continue;
                        Lane& L = lanes[(size_t)k * D + (uint32_t)j];
                        size_t wanted = nd.role == GF_ROLE_TRANSFORM ? std::min<size_t>(Q, ready[k].size()) : 1u;
                        uint32_t reserve = 0;
                        if (nd.role == GF_ROLE_TRANSFORM) {
                            uint64_t parts = 0;
                            for (size_t i = 0; i < wanted; ++i) parts += ready[k][i].fill;
                            while (wanted) {
                                uint64_t bound = output_chunk_bound(parts, C, nd.bins);
                                if (pool.free_list.size() - reserved >= bound) {
                                    reserve = (uint32_t)bound;
                                    break;
                                }
                                --wanted;
                                parts -= ready[k][wanted].fill;
                            }
                            if (!wanted) continue;
                        }
                        L.batch.assign(ready[k].begin(), ready[k].begin() + (long)wanted);
                        ready[k].erase(ready[k].begin(), ready[k].begin() + (long)wanted);
                        L.reserved = reserve;
                        reserved += reserve;
                        L.particles = 0;
                        for (auto& d : L.batch) L.particles += d.fill;
                        L.armed = true;
                        gf_stage& stg = *stages[(size_t)k * D + (uint32_t)j];
                        if (fused && stg.kind() == 1) {
                            // one launch classifies, plans, acquires and scatters; the batch is
                            // complete before the next iteration like every other batch here
                            stg.launch_fused(L.batch.data(), (uint32_t)L.batch.size(), 0, s);
                            uint32_t got = 0;
                            stg.fused_collect(got);
                            stg.collected = true;
                            trace_host(f.trace, tr, k, iter, L.batch, nd.bins, stg.totals.data(), stg.pending);
                        } else {
                            stg.launch(L.batch.data(), (uint32_t)L.batch.size(), 0, s);
                        }
                        if (nd.role == GF_ROLE_SINK) trace_host(f.trace, tr, k, iter, L.batch, 0, nullptr, {});
                        if (stg.kind() == 2) {
                            uint32_t tot = (uint32_t)L.particles;
                            trace_host(f.trace, tr, k, iter, L.batch, 1, &tot, stg.pending);
                        }
                        dispatched = true;
                    }
                if (dispatched) progress = true;
                ++rr;
            }
            bool drained = true;
            for (uint32_t k = 0; k < n; ++k)
                if (!ready[k].empty()) drained = false;
            for (auto& L : lanes)
                if (L.armed) drained = false;
            ++iter;
            if (generated >= N && drained && reserved == 0) break;
            if (!progress) fail(GF_STALL, "graphFlow: scheduler stall, no progress possible");
        }
        // every launch complete (within the stall timeout), then the status of sinks and scatters
        struct EventGuard {
            hipEvent_t e = nullptr;
            ~EventGuard()
            {
                if (e) (void)hipEventDestroy(e);
            }
        } done;
        HIPX(hipEventCreateWithFlags(&done.e, hipEventDisableTiming));
        HIPX(hipEventRecord(done.e, s));
        stages[0]->wait_event(done.e);
        for (auto& sg : stages) sg->settle_all(s);
    } catch (...) {
        teardown();
        throw;
    }
    HIPX(hipStreamSynchronize(s));
    if (pool.free_list.size() != pool.S) fail(GF_BAD_ARGUMENT, "graphFlow: chunk leak after the run");
    st.completed_batches = batches;
    st.steps = iter;
    st.generated = generated;
    st.sunk = sunk;
    st.peak_live_chunks = pool.peak_live;
    uint64_t l = 0;
    for (auto& sg : stages) l += sg->launches;
    st.kernel_launches = l;
}

std::string shape_key(const gf_flow& f, const RunShape& s, uint64_t N)
{
    // N, P and the seed are per-run scalars of the engine (set by run_device), not part of its build;
    // only whether P makes the persistent one-workgroup engine eligible is
    (void)N;
    const bool small_p = (uint64_t)s.P * payload_bytes(f.payload) <= (4ull << 20);
    std::string k = std::to_string(f.payload) + "/" + std::to_string(s.C) + "/" + std::to_string(small_p) + "/" +
                    std::to_string(s.S) + "/" + std::to_string(s.Q) + "/" + std::to_string(s.tile_items) + "/" +
                    std::to_string(f.tuning.trace) + "/" +
                    std::to_string(f.tuning.steps_per_launch) + "/" + std::to_string(f.tuning.fuse_transforms) + "/" +
                    std::to_string(f.tuning.max_workgroups) + "/" + std::to_string(f.tuning.device_engine) + "/" +
                    std::to_string(f.tuning.use_graph) + "/" + std::to_string(f.tuning.team_xcd) + "/" +
                    std::to_string(f.tuning.inflight_batches) + "/" + std::to_string(f.tuning.bin_path) + "/" +
                    std::to_string(f.tuning.hash_table_entries) + "/" + std::to_string(f.slots) + "/" +
                    std::to_string(f.tuning.source_chunks);
    {
        const PhotonParams ph = photon_params_of(f.tuning);
        char buf[160];
        std::snprintf(buf, sizeof buf, "/%a/%a/%a/%a/%a/%a/%u", ph.L, ph.mu_t, ph.albedo, ph.g, ph.roulette_w,
                      ph.roulette_p, ph.tally_bins);
        k += buf;
    }
    for (auto& n : f.nodes) {
        k += "|" + std::to_string(n.nd.role) + "," + std::to_string(n.nd.op) + "," + std::to_string(n.nd.bins);
        for (int32_t r : n.routes) k += "," + std::to_string(r);
        const uint32_t* u = n.nd.u;
        const float* fl = n.nd.f;
        k += "," + std::to_string(u[0]) + "," + std::to_string(u[1]) + "," + std::to_string(u[2]) + "," + std::to_string(u[3]);
        k += "," + std::to_string(fl[0]) + "," + std::to_string(fl[1]) + "," + std::to_string(n.nd.ptr[0]) + "," +
             std::to_string(n.nd.ptr[1]) + "," + std::to_string(n.nd.size[0]);
    }
    return k;
}


// Stream-ordered host<->device copy on the flow's stream, then wait for it: never touches the
// legacy stream, so it is safe while another thread captures a graph on its own stream.
void copy_sync(void* dst, const void* src, size_t bytes, hipMemcpyKind kind, hipStream_t s)
{
    HIPX(hipMemcpyAsync(dst, src, bytes, kind, s));
    HIPX(hipStreamSynchronize(s));
}

void build_engine(gf_flow& f, uint64_t N, const RunShape& sh, uint32_t src)
{
    auto eb = std::make_unique<EngineBuffers>();
    DevMem& m = eb->mem;
    Engine& e = eb->e;
    const uint32_t n = (uint32_t)f.nodes.size();
    const gf_tuning& t = f.tuning;
    // lanes: D per source and transform node, one per sink (Tuning::inflightBatches)
    const uint32_t D = t.inflight_batches ? t.inflight_batches : 1u;
    std::vector<uint32_t> first_lane(n), lanes_of(n);
    uint32_t nv = 0;
    for (uint32_t k = 0; k < n; ++k) {
        first_lane[k] = nv;
        lanes_of[k] = f.nodes[k].nd.role == GF_ROLE_SINK ? 1u : D;
        nv += lanes_of[k];
    }
    if (nv > kMaxNodes)
        fail(GF_BAD_ARGUMENT, "graphFlow: the device engine supports at most " + std::to_string(kMaxNodes) +
                                  " lanes (nodes times inflight batches)");
    gf_pool& pool = *f.pool;
    e.n_nodes = nv;
    e.C = sh.C;
    e.P = sh.P;
    e.Q = sh.Q;
    e.S = sh.S;
    e.ring_cap = sh.S;
    e.N = N;
    e.cols = pool.dcols();
    e.rp.seed = t.seed;
    e.rp.C = sh.C;
    e.rp.ph = photon_params_of(t);
    e.src_chunks = t.source_chunks ? t.source_chunks : 1u;
    std::vector<gf_node_desc> real(n), nds(nv);
    std::vector<int32_t> routes, fuse;
    std::vector<uint32_t> ring_of(nv);
    const uint32_t r = effective_r(f.payload, pick_tile_r(sh.tile_items));
    const uint32_t rec = record_bytes(f.payload);
    // KA and KB hold only their tiles' LDS (the plan, retire and select run in KP and KR)
    size_t lds_a = 256, lds_b = 256, group_lds = 256;
    uint32_t grid_tiles = 0;
    std::vector<NodeArr> arr(nv);
    float zipf_s = 0.0f;
    auto is_plain_xf = [&](int32_t k) {
        return k >= 0 && f.nodes[k].nd.role == GF_ROLE_TRANSFORM && !is_partition(f.nodes[k].nd) && f.nodes[k].nd.bins == 1u;
    };
    // stream purpose of a node: 2 x its id in its job's graph (a served flow's slot copy keeps the
    // purposes of the job run alone)
    auto purpose_of = [&](uint32_t k) { return 2u * (f.slots ? f.nodes[k].local : k); };
    for (uint32_t k = 0; k < n; ++k) {
        real[k] = f.nodes[k].nd;
        real[k].route_offset = (uint32_t)routes.size();
        const bool part = real[k].role == GF_ROLE_TRANSFORM && is_partition(real[k]);
        for (int32_t dst : f.nodes[k].routes) {
            // a plain transform fed by a partition bin runs inside that partition's scatter; routes
            // name the destination's first lane, which owns the destination's ready ring. A fused
            // entry packs the transform's node (low 16 bits) and its stream purpose (high 16 bits).
            if (part && t.fuse_transforms && is_plain_xf(dst)) {
                fuse.push_back((int32_t)((uint32_t)dst | purpose_of((uint32_t)dst) << 16));
                routes.push_back((int32_t)first_lane[f.nodes[dst].routes[0]]);
            } else {
                fuse.push_back(-1);
                routes.push_back((int32_t)first_lane[dst]);
            }
        }
    }
    // sinks fed by a partition of 64 or more bins take many small chunks per step: wave tiles
    std::vector<bool> wave_sink(n, false);
    for (uint32_t k = 0; k < n; ++k) {
        if (!(real[k].role == GF_ROLE_TRANSFORM && is_partition(real[k])) || real[k].bins < 64u) continue;
        for (int32_t dst : f.nodes[k].routes)
            if (dst >= 0 && f.nodes[dst].nd.role == GF_ROLE_SINK) wave_sink[dst] = true;
    }
    uint32_t rings = 0;
    for (uint32_t k = 0; k < n; ++k) {
        const gf_node_desc& nd = real[k];
        const bool part = nd.role == GF_ROLE_TRANSFORM && is_partition(nd);
        const uint32_t B = std::max<uint32_t>(nd.bins, 1u);
        const uint32_t ring = rings++;
        for (uint32_t j = 0; j < lanes_of[k]; ++j) {
            const uint32_t v = first_lane[k] + j;
            nds[v] = nd;
            ring_of[v] = first_lane[k];
            NodeArr& a = arr[v];
            std::memset(&a, 0, sizeof(a));
            a.kind = nd.role == GF_ROLE_SOURCE ? KIND_SOURCE : nd.role == GF_ROLE_SINK ? KIND_SINK : part ? KIND_PARTITION : KIND_TRANSFORM;
            a.cap_in = nd.role == GF_ROLE_SOURCE ? 1u : nd.role == GF_ROLE_SINK ? sh.S : sh.Q;
            a.cap_out = nd.role == GF_ROLE_SOURCE ? e.src_chunks : nd.role == GF_ROLE_SINK ? 1u : reserve_bound_of(nd, sh.C, sh.Q);
            a.reserve_bound = a.cap_out;
            a.real = k;
            a.lanes = j == 0 ? lanes_of[k] : 0u;
            a.ring = ring;
            // a source's tiles hold four items per thread: each item's generation (Philox, a Zipf
            // search) is a chain of dependent steps, and a thread runs its four chains together
            const uint32_t ti = nd.role == GF_ROLE_SOURCE ? 4u * (uint32_t)kThreads : sh.tile_items;
            uint32_t cap_tiles = (uint32_t)(nd.role == GF_ROLE_SOURCE ? ceil_div((uint64_t)sh.C * e.src_chunks, ti)
                                                                        : ceil_div(sh.C, ti) * (nd.role == GF_ROLE_SINK ? 1u : a.cap_in));
            a.in = m.d<gf_chunk_desc>(a.cap_in);
            a.tile_prefix = m.d<uint32_t>(a.cap_in + 1u);
            a.item_prefix = m.d<uint32_t>(a.cap_in + 1u);
            a.out = m.d<gf_chunk_desc>(a.cap_out);
            a.counts = part ? m.d<uint32_t>((size_t)cap_tiles * B) : nullptr;
            a.totals = m.d<uint32_t>(2u * B);
            a.tile_items = ti;
            a.purpose = purpose_of(k);
            a.job = f.nodes[k].slot;
            if (wave_sink[k]) a.flags |= NODE_WAVE_SINK;
            if (part) {
                a.hash_entries = hash_entries_of(nd, t.bin_path, t.hash_table_entries, sh.tile_items);
                if (a.hash_entries) {
                    lds_a = std::max(lds_a, count_hash_smem(a.hash_entries));
                    lds_b = std::max(lds_b, scatter_hash_smem(a.hash_entries));
                    group_lds = std::max(group_lds, std::max(count_hash_smem(a.hash_entries), scatter_hash_smem(a.hash_entries)));
                } else {
                    lds_a = std::max(lds_a, (size_t)B * 4u);
                    bool stg = stage_fits(B, sh.tile_items, rec);
                    if (stg) a.flags |= NODE_STAGE;
                    lds_b = std::max(lds_b, scatter_smem(B, sh.tile_items, rec, stg));
                    group_lds = std::max(group_lds, std::max((size_t)B * 4u, scatter_smem(B, sh.tile_items, rec, stg)));
                }
                grid_tiles += cap_tiles;
            } else if (nd.role != GF_ROLE_SINK) {
                grid_tiles += cap_tiles;
            }
        }
        if (nd.op == GF_SRC_KEYED_ZIPF && !(e.rp.zipf && e.rp.zipf_values == nd.u[0] && zipf_s == nd.f[0])) {
            zipf_s = nd.f[0];  // a served flow's slot copies share one table
            auto z = zipf_table(nd.u[0], (double)nd.f[0]);
            double* dz = m.d<double>(z.size());
            copy_sync(dz, z.data(), z.size() * sizeof(double), hipMemcpyHostToDevice, f.ctx->stream);
            auto gd = zipf_guide(z);
            uint32_t* dg = m.d<uint32_t>(gd.size());
            copy_sync(dg, gd.data(), gd.size() * sizeof(uint32_t), hipMemcpyHostToDevice, f.ctx->stream);
            e.rp.zipf = dz;
            e.rp.zipf_values = nd.u[0];
            e.rp.zipf_guide = dg;
            e.rp.zipf_s = (double)nd.f[0];
            e.rp.zipf_a = zipf_fit_a(z, e.rp.zipf_s);
        }
    }
    if (lds_b > 160u * 1024u || lds_a > 160u * 1024u)
        fail(GF_BAD_ARGUMENT, "graphFlow: partition width exceeds the workgroup LDS budget");
    gf_node_desc* dn = m.d<gf_node_desc>(nv);
    copy_sync(dn, nds.data(), sizeof(gf_node_desc) * nv, hipMemcpyHostToDevice, f.ctx->stream);
    gf_node_desc* drn = m.d<gf_node_desc>(n);
    copy_sync(drn, real.data(), sizeof(gf_node_desc) * n, hipMemcpyHostToDevice, f.ctx->stream);
    int32_t* dr = m.d<int32_t>(std::max<size_t>(routes.size(), 1));
    int32_t* df = m.d<int32_t>(std::max<size_t>(fuse.size(), 1));
    if (!routes.empty()) {
        copy_sync(dr, routes.data(), sizeof(int32_t) * routes.size(), hipMemcpyHostToDevice, f.ctx->stream);
        copy_sync(df, fuse.data(), sizeof(int32_t) * fuse.size(), hipMemcpyHostToDevice, f.ctx->stream);
    }
    uint32_t* dring = m.d<uint32_t>(nv);
    copy_sync(dring, ring_of.data(), sizeof(uint32_t) * nv, hipMemcpyHostToDevice, f.ctx->stream);
    NodeArr* da = m.d<NodeArr>(nv);
    copy_sync(da, arr.data(), sizeof(NodeArr) * nv, hipMemcpyHostToDevice, f.ctx->stream);
    e.nodes = dn;
    e.real_nodes = drn;
    e.routes = dr;
    e.fuse = df;
    e.ring_of = dring;
    e.arr = da;
    e.st = m.d<NodeSt>(nv);
    e.free_stack = m.d<unsigned long long>(sh.S);
    e.rings = m.d<gf_chunk_desc>((size_t)rings * sh.S);
    e.ka_prefix = m.d<uint32_t>(nv + 1u);
    e.kb_prefix = m.d<uint32_t>(nv + 1u);
    e.ctl = m.d<uint32_t>(CTL_WORDS);
    e.ctr = m.d<unsigned long long>(CNT_WORDS);
    e.trace_on = t.trace ? 1u : 0u;
    e.trace_cap = t.trace ? (unsigned long long)1 << 26 : 0ull;
    e.trace = t.trace ? m.d<uint32_t>(e.trace_cap) : nullptr;
    e.prof = t.use_graph == 3u ? m.d<unsigned long long>(PROF_WORDS) : nullptr;
    e.lb = m.d<unsigned long long>(nv);
    // grid retire: a step routes at most S outputs; its input lists hold at most S chunks
    e.route_units_cap = (uint32_t)ceil_div(sh.S, kGridBlock) + 1u;
    e.list_units_cap = (uint32_t)ceil_div(sh.S, kGridBlock) + nv + 1u;
    e.rlb = m.d<unsigned long long>((size_t)e.route_units_cap * nv);
    e.llb = m.d<unsigned long long>(2ull * e.list_units_cap);
    e.n_jobs = f.slots;
    e.jobs = nullptr;
    if (f.slots) {
        // every slot empty (N = 0: done); the host admits jobs between launches
        std::vector<JobSlot> z(f.slots);
        std::memset(z.data(), 0, sizeof(JobSlot) * z.size());
        e.jobs = m.d<JobSlot>(f.slots);
        copy_sync(e.jobs, z.data(), sizeof(JobSlot) * z.size(), hipMemcpyHostToDevice, f.ctx->stream);
    }
    eb->route_grid = std::min<uint32_t>(64u, e.route_units_cap + nv);
    const uint32_t gmax = t.max_workgroups ? t.max_workgroups : 64u;
    eb->grid = std::min<uint32_t>(std::max<uint32_t>(grid_tiles, 1u), gmax);
    eb->r = r;
    eb->lds_a = lds_a;
    eb->lds_b = lds_b;
    eb->key = shape_key(f, sh, N);
    // engine choice: one persistent workgroup when the live items fit the L2 of one XCD (4 MiB)
    // and its LDS holds four tile groups; the multi-workgroup step kernels otherwise
    group_lds = (group_lds + 15u) & ~(size_t)15u;
    eb->group_lds = (uint32_t)group_lds;
    eb->micro_lds = std::max(std::max(sizeof(PlanLds), sizeof(SelLds)), 4u * group_lds);
    const bool micro_fits = eb->micro_lds <= 160u * 1024u;
    if (t.device_engine == 2u && !micro_fits) fail(GF_BAD_ARGUMENT, "graphFlow: micro engine LDS budget exceeded");
    // team engine: every partition narrow enough for a per-segment count row in LDS; its replicated
    // scheduler runs one lane per node
    uint32_t bb = 1;
    bool team_ok = D == 1u && e.src_chunks == 1u && f.slots <= kTeamJobs && n <= kTeamNodes &&
                   routes.size() <= kTeamRoutes && ceil_div(sh.C, kSeg) <= (1u << 17);
    for (uint32_t k = 0; k < n; ++k)
        if (is_partition(real[k]) && real[k].role == GF_ROLE_TRANSFORM) bb = std::max(bb, real[k].bins + 1u);
    if (bb > kTeamBins) team_ok = false;
    bb = (bb + 3u) & ~3u;  // rows of whole 16-byte words: one 128-bit load per four slots
    const uint32_t tpc = (uint32_t)ceil_div(sh.C, kSeg);
    const uint64_t cnt_words = (uint64_t)sh.S * 2u * tpc * bb;
    if (cnt_words * 4u > (1ull << 31)) team_ok = false;
    if (t.device_engine == 3u && !team_ok)
        fail(GF_BAD_ARGUMENT, "graphFlow: the team engine needs at most 32 nodes and partitions of at most 31 bins");
    eb->team = team_ok && (t.device_engine == 3u || t.device_engine == 0u);
    if (eb->team) {
        TeamArgs& ta = eb->ta;
        ta.cnt = m.d<uint32_t>(cnt_words);
        eb->ta_dev = m.d<TeamArgs>(1);
        ta.zero_g = m.d<uint32_t>(1u + kTeamZero);
        ta.tw = m.d<uint32_t>(TW_WORDS);
        // default team: every CU of one XCD (the device's CUs over its 8 XCDs)
        const uint32_t xcd_cus = std::max<uint32_t>(1u, f.ctx->cus / 8u);
        ta.K = std::min<uint32_t>(t.max_workgroups ? t.max_workgroups : xcd_cus, 32u);
        ta.xcd = t.team_xcd;
        ta.tpc = tpc;
        ta.bb = bb;
        ta.cnt_bytes = (uint32_t)(cnt_words * 4u);
        ta.max_steps = t.steps_per_launch ? t.steps_per_launch : 65536u;
        uint32_t rc = 1;
        while (rc * 2u <= kTeamRing / n) rc *= 2u;
        ta.rc = std::min<uint32_t>(rc, 1024u);
        ta.n_routes = (uint32_t)routes.size();
        ta.mprof = t.use_graph == 3u ? m.d<unsigned long long>(64u * kMemberProf) : nullptr;
        ta.c_shift = 32u;
        if ((sh.C & (sh.C - 1u)) == 0u)
            for (uint32_t b = 0; b < 32u; ++b)
                if ((1u << b) == sh.C) ta.c_shift = b;
        eb->cnt_bytes = cnt_words * 4u;
        HIPX(hipFuncSetAttribute((const void*)k_team<Particle>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sizeof(TeamLds)));
        HIPX(hipFuncSetAttribute((const void*)k_team<Keyed>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sizeof(TeamLds)));
        HIPX(hipFuncSetAttribute((const void*)k_team<Photon>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sizeof(TeamLds)));
    }
    if (f.slots && t.device_engine == 2u) fail(GF_BAD_ARGUMENT, "graphFlow: a served flow runs on the wide or the team engine");
    eb->micro = !eb->team && !f.slots && (t.device_engine == 2u ||
                              (t.device_engine == 0u && micro_fits && (uint64_t)sh.P * rec <= (4ull << 20)));
    with_payload(f.payload, r, [&](auto pt, auto rc) {
        using P = typename decltype(pt)::type;
        constexpr int RR = decltype(rc)::value;
        set_lds(k_engine_ka<P, RR>, lds_a);
        set_lds(k_engine_kb<P, RR, KB_PART>, lds_b);
        if (micro_fits) set_lds(k_micro<P, RR>, eb->micro_lds);
    });
    static_assert(sizeof(SelLds) <= 160u * 1024u, "the select's LDS state exceeds a workgroup's LDS");
    set_lds(k_engine_kr, sizeof(SelLds));
    set_lds(k_engine_setup, sizeof(SelLds));
    set_lds(k_engine_resume, sizeof(SelLds));
    (void)src;
    f.eng = std::move(eb);
}

// One step: KA (classify/count tiles), KP (plan, a workgroup per active lane), KB (scatter /
// transform / source / sink tiles), KR (routing and release over the grid, counters and select in
// one workgroup, input lists over the grid). With `ev`, events are recorded around KA, KP and
// KB + KR.
// With events (timed runs): ev[0..6] around classify, plan, the three scatter kernels and the
// retire/select kernels.
void launch_step(const EngineBuffers& eb, uint32_t payload, hipStream_t s, hipEvent_t* ev = nullptr)
{
    with_payload(payload, eb.r, [&](auto pt, auto rc) {
        using P = typename decltype(pt)::type;
        constexpr int RR = decltype(rc)::value;
        if (ev) HIPX(hipEventRecord(ev[0], s));
        hipLaunchKernelGGL((k_engine_ka<P, RR>), dim3(eb.grid), dim3(kThreads), eb.lds_a, s, eb.e);
        if (ev) HIPX(hipEventRecord(ev[1], s));
        hipLaunchKernelGGL(k_engine_plan, dim3(eb.e.n_nodes), dim3(1024), 0, s, eb.e);
        if (eb.e.trace_on) hipLaunchKernelGGL(k_engine_trace, dim3(1), dim3(1024), 0, s, eb.e);
        if (ev) HIPX(hipEventRecord(ev[2], s));
        // partition / transform tiles, then source tiles, then sink tiles (disjoint lanes and chunks;
        // running them on forked streams measured no faster than back to back)
        hipLaunchKernelGGL((k_engine_kb<P, RR, KB_PART>), dim3(eb.grid), dim3(kThreads), eb.lds_b, s, eb.e);
        if (ev) HIPX(hipEventRecord(ev[3], s));
        hipLaunchKernelGGL((k_engine_kb<P, RR, KB_SRC>), dim3(eb.grid), dim3(kThreads), 256, s, eb.e);
        if (ev) HIPX(hipEventRecord(ev[4], s));
        hipLaunchKernelGGL((k_engine_kb<P, RR, KB_SINK>), dim3(eb.grid), dim3(kThreads), 256, s, eb.e);
        if (ev) HIPX(hipEventRecord(ev[5], s));
        hipLaunchKernelGGL(k_engine_route, dim3(eb.route_grid), dim3(1024), 0, s, eb.e);
        hipLaunchKernelGGL(k_engine_kr, dim3(1), dim3(1024), sizeof(SelLds), s, eb.e);
        hipLaunchKernelGGL(k_engine_lists, dim3(eb.route_grid), dim3(1024), 0, s, eb.e);
        if (ev) HIPX(hipEventRecord(ev[6], s));
    });
}

void finish_device(gf_flow& f, EngineBuffers& eb, gf_stats& st, const uint32_t* ctl, uint64_t launches);

// Micro engine host loop: launches of one persistent workgroup, each running up to M steps.
void run_micro(gf_flow& f, EngineBuffers& eb, uint64_t N, gf_stats& st, bool timed)
{
    hipStream_t s = f.ctx->stream;
    const uint32_t M = f.tuning.steps_per_launch ? f.tuning.steps_per_launch : 2048u;
    uint32_t ctl[CTL_WORDS];
    uint64_t launches = 0, steps_launched = 0;
    const uint64_t max_steps = 64ull * (N + 1024u) + 4096u;
    hipEvent_t ev[2] = {nullptr, nullptr};
    if (timed)
        for (auto& x : ev) HIPX(hipEventCreate(&x));
    for (uint32_t i = 0; i < kKernelKinds; ++i) {
        f.kt_ms[i] = 0.0;
        f.kt_n[i] = 0;
    }
    for (;;) {
        if (timed) HIPX(hipEventRecord(ev[0], s));
        with_payload(f.payload, eb.r, [&](auto pt, auto rc) {
            using P = typename decltype(pt)::type;
            constexpr int RR = decltype(rc)::value;
            hipLaunchKernelGGL((k_micro<P, RR>), dim3(1), dim3(1024), eb.micro_lds, s, eb.e, M, eb.group_lds);
        });
        HIPX(hipGetLastError());
        if (timed) HIPX(hipEventRecord(ev[1], s));
        ++launches;
        steps_launched += M;
        HIPX(hipMemcpyAsync(ctl, eb.e.ctl, sizeof(ctl), hipMemcpyDeviceToHost, s));
        HIPX(hipStreamSynchronize(s));
        if (timed) {
            float x = 0;
            HIPX(hipEventElapsedTime(&x, ev[0], ev[1]));
            f.kt_ms[3] += x;
            f.kt_n[3]++;
        }
        if (ctl[CTL_STATUS] || ctl[CTL_DONE]) break;
        if (steps_launched > max_steps) fail(GF_STALL, "graphFlow: device engine exceeded its step bound");
    }
    if (timed)
        for (auto& x : ev) (void)hipEventDestroy(x);
    finish_device(f, eb, st, ctl, launches);
}

void run_wide(gf_flow& f, EngineBuffers& eb, uint64_t N, gf_stats& st, bool timed, uint64_t launches0);

// Team engine host loop: each launch forms the team on one XCD and runs up to max_steps steps; a
// step that exceeds the team's LDS limits hands the rest of the run to the wide engine.
void run_team(gf_flow& f, EngineBuffers& eb, uint64_t N, gf_stats& st, bool timed)
{
    hipStream_t s = f.ctx->stream;
    TeamArgs& ta = eb.ta;
    ta.e = eb.e;
    HIPX(hipMemsetAsync(ta.cnt, 0, eb.cnt_bytes, s));
    if (ta.mprof) HIPX(hipMemsetAsync(ta.mprof, 0, 64u * kMemberProf * sizeof(unsigned long long), s));
    HIPX(hipMemsetAsync(ta.zero_g, 0, sizeof(uint32_t), s));
    uint32_t ctl[CTL_WORDS];
    uint64_t launches = 0, steps_launched = 0;
    const uint64_t max_steps = 64ull * (N + 1024u) + 4096u;
    hipEvent_t ev[2] = {nullptr, nullptr};
    if (timed)
        for (auto& x : ev) HIPX(hipEventCreate(&x));
    for (uint32_t i = 0; i < kKernelKinds; ++i) {
        f.kt_ms[i] = 0.0;
        f.kt_n[i] = 0;
    }
    const uint32_t grid = 8u * ta.K;
    copy_sync(eb.ta_dev, &ta, sizeof(TeamArgs), hipMemcpyHostToDevice, s);
    for (;;) {
        HIPX(hipMemsetAsync(ta.tw, 0, TW_WORDS * sizeof(uint32_t), s));
        if (timed) HIPX(hipEventRecord(ev[0], s));
        with_payload(f.payload, 1u, [&](auto pt, auto) {
            using P = typename decltype(pt)::type;
            hipLaunchKernelGGL(k_team<P>, dim3(grid), dim3(kTeamThreads), sizeof(TeamLds), s, (const TeamArgs*)eb.ta_dev);
        });
        HIPX(hipGetLastError());
        if (timed) HIPX(hipEventRecord(ev[1], s));
        ++launches;
        steps_launched += ta.max_steps;
        HIPX(hipMemcpyAsync(ctl, eb.e.ctl, sizeof(ctl), hipMemcpyDeviceToHost, s));
        HIPX(hipStreamSynchronize(s));
        if (timed) {
            float x = 0;
            HIPX(hipEventElapsedTime(&x, ev[0], ev[1]));
            f.kt_ms[3] += x;
            f.kt_n[3]++;
        }
        if (ctl[CTL_STATUS] || ctl[CTL_DONE]) break;
        if (ctl[CTL_HANDOFF]) {
            if (timed)
                for (auto& x : ev) (void)hipEventDestroy(x);
            unsigned long long step = 0;
            copy_sync(&step, eb.e.ctr + CNT_STEP, sizeof(step), hipMemcpyDeviceToHost, f.ctx->stream);
            st.handoff_step = (uint32_t)step;
            st.engine = 4;
            hipLaunchKernelGGL(k_engine_resume, dim3(1), dim3(kThreads), sizeof(SelLds), s, eb.e);
            HIPX(hipGetLastError());
            run_wide(f, eb, N, st, timed, launches);
            st.engine = 4;
            return;
        }
        if (steps_launched > max_steps) fail(GF_STALL, "graphFlow: device engine exceeded its step bound");
    }
    if (timed)
        for (auto& x : ev) (void)hipEventDestroy(x);
    finish_device(f, eb, st, ctl, launches);
}

void run_device_form(gf_flow& f, uint32_t src, uint64_t N, const RunShape& sh, gf_stats& st, bool timed)
{
    const std::string key = shape_key(f, sh, N);
    if (!f.eng || f.eng->key != key) build_engine(f, N, sh, src);
    EngineBuffers& eb = *f.eng;
    if (eb.e.N != N || eb.e.P != sh.P || eb.e.rp.seed != f.tuning.seed) {
        // same build, new run scalars: every kernel takes the engine by value, so only a captured
        // graph has to be captured again
        eb.e.N = eb.ta.e.N = N;
        eb.e.P = eb.ta.e.P = sh.P;
        eb.e.rp.seed = eb.ta.e.rp.seed = f.tuning.seed;
        if (eb.exec) {
            (void)hipGraphExecDestroy(eb.exec);
            eb.exec = nullptr;
        }
        eb.graph_steps = 0;
    }
    hipStream_t s = f.ctx->stream;
    hipLaunchKernelGGL(k_engine_setup, dim3(1), dim3(kThreads), sizeof(SelLds), s, eb.e, eb.team ? 0u : 1u);
    HIPX(hipGetLastError());
    if (eb.team) {
        st.engine = 3;
        run_team(f, eb, N, st, timed);
        return;
    }
    if (eb.micro) {
        st.engine = 2;
        run_micro(f, eb, N, st, timed);
        return;
    }
    st.engine = 1;
    run_wide(f, eb, N, st, timed, 0);
}

// A run's private tally copies (photon_tally): attached to the flow's photon steps before anything is
// built or launched, folded into each step's tally when the run ends, also when it fails (the
// caller's tally then holds what the launched batches added, as with direct atomic adds).
struct TallyScope {
    gf_flow& f;
    bool folded = false;

    explicit TallyScope(gf_flow& fl) : f(fl)
    {
        const uint32_t tb = f.tuning.photon_tally_bins ? f.tuning.photon_tally_bins : 1024u;
        for (auto& n : f.nodes) {
            if (n.nd.role != GF_ROLE_TRANSFORM || n.nd.op != GF_XF_PHOTON_STEP) continue;
            if (n.tally_copies && n.tally_copy_bins != tb) {
                f.eng.reset();  // its node array names the old copies
                (void)hipFree(n.tally_copies);
                n.tally_copies = nullptr;
            }
            if (!n.tally_copies) {
                n.tally_copies = dalloc<double>((size_t)kTallyCopies * tb);
                HIPX(hipMemset(n.tally_copies, 0, sizeof(double) * kTallyCopies * tb));
                n.tally_copy_bins = tb;
            }
            n.nd.ptr[1] = (uint64_t)(uintptr_t)n.tally_copies;
        }
    }

    void fold()
    {
        folded = true;
        hipStream_t s = f.ctx->stream;
        for (auto& n : f.nodes) {
            if (!n.tally_copies || n.nd.op != GF_XF_PHOTON_STEP) continue;
            const uint32_t tb = n.tally_copy_bins;
            hipLaunchKernelGGL(k_tally_fold, dim3((tb + 255u) / 256u), dim3(256), 0, s, n.tally_copies, tb,
                               (double*)n.nd.ptr[0]);
            HIPX(hipGetLastError());
        }
        HIPX(hipStreamSynchronize(s));
    }

    ~TallyScope()
    {
        if (folded) return;
        try {
            (void)hipStreamSynchronize(f.ctx->stream);
            fold();
        } catch (...) {
        }
    }
};

void run_host(gf_flow& f, uint32_t src, uint64_t N, const RunShape& sh, gf_stats& st)
{
    TallyScope ts(f);
    run_host_form(f, src, N, sh, st);
    ts.fold();
}

void run_device(gf_flow& f, uint32_t src, uint64_t N, const RunShape& sh, gf_stats& st, bool timed)
{
    TallyScope ts(f);
    run_device_form(f, src, N, sh, st, timed);
    ts.fold();
}

void run_wide(gf_flow& f, EngineBuffers& eb, uint64_t N, gf_stats& st, bool timed, uint64_t launches0)
{
    hipStream_t s = f.ctx->stream;
    const uint32_t K = f.tuning.steps_per_launch ? f.tuning.steps_per_launch : 64u;
    const bool use_graph = f.tuning.use_graph == 1u && !timed;
    if (use_graph && (!eb.exec || eb.graph_steps != K)) {
        if (eb.exec) {
            (void)hipGraphExecDestroy(eb.exec);
            eb.exec = nullptr;
        }
        hipGraph_t g;
        HIPX(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        for (uint32_t i = 0; i < K; ++i) launch_step(eb, f.payload, s);
        HIPX(hipStreamEndCapture(s, &g));
        HIPX(hipGraphInstantiate(&eb.exec, g, nullptr, nullptr, 0));
        (void)hipGraphDestroy(g);
        eb.graph_steps = K;
    }
    uint32_t ctl[CTL_WORDS];
    uint64_t launches = launches0;
    hipEvent_t ev[7] = {};
    if (timed)
        for (auto& x : ev) HIPX(hipEventCreate(&x));
    for (uint32_t i = 0; i < kKernelKinds; ++i) {
        f.kt_ms[i] = 0.0;
        f.kt_n[i] = 0;
    }
    // host-side bound: every step admits, plans or retires something, so a correct run needs far
    // fewer steps than this; exceeding it means a scheduling defect, reported as a stall
    const uint64_t max_steps = 64ull * (N + 1024u) + 4096u;
    uint64_t steps_launched = 0;
    for (;;) {
        if (use_graph) {
            HIPX(hipGraphLaunch(eb.exec, s));
            launches += 4ull * K;
        } else {
            for (uint32_t i = 0; i < K; ++i) {
                if (timed) {
                    launch_step(eb, f.payload, s, ev);
                    HIPX(hipEventSynchronize(ev[6]));
                    // kinds 0, 1, 2 (= 4 + 5 + 6 + 7), 4: partition, 5: source, 6: sink, 7: retire/select
                    const int from[7] = {0, 1, 2, 2, 3, 4, 5}, to[7] = {1, 2, 6, 3, 4, 5, 6}, kind[7] = {0, 1, 2, 4, 5, 6, 7};
                    for (int q = 0; q < 7; ++q) {
                        float x = 0;
                        HIPX(hipEventElapsedTime(&x, ev[from[q]], ev[to[q]]));
                        f.kt_ms[kind[q]] += x;
                        f.kt_n[kind[q]]++;
                    }
                } else {
                    launch_step(eb, f.payload, s);
                }
                launches += 4;
            }
        }
        steps_launched += K;
        HIPX(hipGetLastError());
        HIPX(hipMemcpyAsync(ctl, eb.e.ctl, sizeof(ctl), hipMemcpyDeviceToHost, s));
        HIPX(hipStreamSynchronize(s));
        if (ctl[CTL_STATUS] || ctl[CTL_DONE]) break;
        if (steps_launched > max_steps) fail(GF_STALL, "graphFlow: device engine exceeded its step bound");
    }
    if (timed)
        for (auto& x : ev) (void)hipEventDestroy(x);
    finish_device(f, eb, st, ctl, launches);
}

void finish_device(gf_flow& f, EngineBuffers& eb, gf_stats& st, const uint32_t* ctl, uint64_t launches)
{
    unsigned long long ctr[CNT_WORDS];
    copy_sync(ctr, eb.e.ctr, sizeof(ctr), hipMemcpyDeviceToHost, f.ctx->stream);
    if (f.tuning.trace) {
        uint64_t len = std::min<uint64_t>(ctr[CNT_TRACE_LEN], eb.e.trace_cap);
        f.trace.resize(len);
        if (len) copy_sync(f.trace.data(), eb.e.trace, len * sizeof(uint32_t), hipMemcpyDeviceToHost, f.ctx->stream);
    }
    st.completed_batches = ctr[CNT_BATCHES];
    st.steps = ctr[CNT_STEP];
    st.generated = ctr[CNT_GENERATED];
    st.sunk = ctr[CNT_SUNK];
    st.peak_live_chunks = (uint32_t)ctr[CNT_PEAK_LIVE];
    st.kernel_launches = launches;
    f.part_items = ctr[CNT_PART_ITEMS];
    f.xf_items = ctr[CNT_XF_ITEMS];
    if (ctl[CTL_STATUS] == kStatusTableFull) fail(GF_INVALID, "partition classify: hash bin table full");
    if (ctl[CTL_STATUS]) {
        gf_status code = (gf_status)ctl[CTL_STATUS];
        const char* what = code == GF_INVALID ? "partition: device status reports an invalid classifier or updater result"
                         : code == GF_OVERFLOW ? "partition scatter: device status reports a checked updater overflow"
                         : code == GF_POOL_EXHAUSTED ? "graphFlow::ChunkPool: pool exhausted"
                         : code == GF_STALL ? "graphFlow: scheduler stall, no progress possible"
                         : code == GF_BAD_ARGUMENT ? "graphFlow: chunk leak after the run"
                                                   : "partition plan: exact chunk count exceeds the reservation bound";
        fail(code, what);
    }
}

}  // namespace

// ==============================================================================================
// C ABI
// ==============================================================================================
extern "C" {

uint32_t gf_abi_version(void) { return 1u; }

const char* gf_status_name(gf_status s)
{
    switch (s) {
    case GF_OK: return "ok";
    case GF_INVALID: return "invalid";
    case GF_OVERFLOW: return "overflow";
    case GF_STALL: return "stall";
    case GF_POOL_EXHAUSTED: return "pool_exhausted";
    case GF_BAD_ARGUMENT: return "bad_argument";
    case GF_PLAN_MISMATCH: return "plan_mismatch";
    case GF_DEVICE_ERROR: return "device_error";
    case GF_STAGE_FAILED: return "stage_failed";
    }
    return "unknown";
}

gf_status gf_context_create(int device, gf_context** out)
{
    if (!out) return GF_BAD_ARGUMENT;
    return guard(nullptr, [&] {
        auto c = std::make_unique<gf_context>();
        c->device = device;
        HIPX(hipSetDevice(device));
        HIPX(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        hipDeviceProp_t prop;
        HIPX(hipGetDeviceProperties(&prop, device));
        c->cus = (uint32_t)prop.multiProcessorCount;
        *out = c.release();
    });
}

gf_status gf_context_destroy(gf_context* ctx)
{
    if (!ctx) return GF_BAD_ARGUMENT;
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return GF_OK;
}

gf_status gf_context_last_error(gf_context* ctx, char* buffer, uint32_t size)
{
    if (!ctx || !buffer || size == 0) return GF_BAD_ARGUMENT;
    std::snprintf(buffer, size, "%s", ctx->err.c_str());
    return GF_OK;
}

gf_status gf_output_chunk_bound(uint64_t n, uint64_t capacity, uint64_t max_bins, uint64_t* out)
{
    if (!out) return GF_BAD_ARGUMENT;
    return guard(nullptr, [&] { *out = output_chunk_bound(n, capacity, max_bins); });
}

gf_status gf_plan_dense_bins(const uint32_t* totals, uint32_t bins, uint32_t capacity, uint32_t* bin_base,
                             uint32_t* fills, uint32_t fills_cap, uint32_t* num_output_chunks)
{
    if (!totals || !bin_base || !num_output_chunks || bins == 0 || capacity == 0) return GF_BAD_ARGUMENT;
    return guard(nullptr, [&] {
        uint64_t exact = 0, parts = 0;
        for (uint32_t b = 0; b < bins; ++b) {
            bin_base[b] = (uint32_t)exact;
            exact += ceil_div(totals[b], capacity);
            parts += totals[b];
        }
        if (exact > UINT32_MAX) fail(GF_OVERFLOW, "planDenseBins: output chunk count overflow");
        if (exact > output_chunk_bound(parts, capacity, bins)) fail(GF_PLAN_MISMATCH, "planDenseBins: bound");
        if (exact > fills_cap) fail(GF_BAD_ARGUMENT, "planDenseBins: fills buffer too small");
        uint32_t k = 0;
        for (uint32_t b = 0; b < bins; ++b)
            for (uint32_t o = 0; o < totals[b]; o += capacity) fills[k++] = std::min(totals[b] - o, capacity);
        *num_output_chunks = (uint32_t)exact;
    });
}

gf_status gf_pool_create(gf_context* ctx, uint32_t payload, uint32_t num_chunks, uint32_t capacity, gf_pool** out)
{
    if (!ctx || !out || num_chunks == 0 || capacity == 0 || payload > GF_PAYLOAD_PHOTON) return GF_BAD_ARGUMENT;
    if ((uint64_t)num_chunks * capacity > UINT32_MAX) return GF_OVERFLOW;
    return guard(&ctx->err, [&] {
        auto p = std::make_unique<gf_pool>();
        p->ctx = ctx;
        p->payload = payload;
        p->S = num_chunks;
        p->C = capacity;
        p->elem = column_bytes(payload);
        for (uint32_t b : p->elem) p->cols.push_back(dalloc<char>((size_t)num_chunks * capacity * b));
        p->gen.assign(num_chunks, 0u);
        p->lease.assign(num_chunks, 0u);
        p->reset_ledger();
        *out = p.release();
    });
}

gf_status gf_pool_destroy(gf_pool* pool)
{
    if (!pool) return GF_BAD_ARGUMENT;
    for (void* p : pool->cols) (void)hipFree(p);
    delete pool;
    return GF_OK;
}

gf_status gf_pool_acquire(gf_pool* pool, gf_chunk_handle* out)
{
    if (!pool || !out) return GF_BAD_ARGUMENT;
    return guard(&pool->ctx->err, [&] { *out = pool->acquire(); });
}

gf_status gf_pool_release(gf_pool* pool, gf_chunk_handle h)
{
    if (!pool) return GF_BAD_ARGUMENT;
    return guard(&pool->ctx->err, [&] { pool->release(h); });
}

gf_status gf_pool_acquire_lease(gf_pool* pool, gf_chunk_handle h)
{
    if (!pool) return GF_BAD_ARGUMENT;
    return guard(&pool->ctx->err, [&] {
        pool->validate(h);
        if (pool->lease[h.index] == UINT32_MAX) fail(GF_OVERFLOW, "graphFlow::ChunkPool lease overflow");
        pool->peak_leases = std::max(pool->peak_leases, ++pool->lease[h.index]);
    });
}

gf_status gf_pool_release_lease(gf_pool* pool, gf_chunk_handle h)
{
    if (!pool) return GF_BAD_ARGUMENT;
    return guard(&pool->ctx->err, [&] {
        pool->validate(h);
        if (pool->lease[h.index] == 0) fail(GF_BAD_ARGUMENT, "graphFlow::ChunkPool: lease release without a matching acquire");
        --pool->lease[h.index];
    });
}

gf_status gf_pool_counts(gf_pool* pool, uint32_t counts[6])
{
    if (!pool || !counts) return GF_BAD_ARGUMENT;
    counts[0] = pool->S;
    counts[1] = pool->C;
    counts[2] = (uint32_t)pool->free_list.size();
    counts[3] = pool->S - (uint32_t)pool->free_list.size();
    counts[4] = pool->peak_live;
    counts[5] = pool->peak_leases;
    return GF_OK;
}

gf_status gf_pool_lease_count(gf_pool* pool, gf_chunk_handle h, uint32_t* count)
{
    if (!pool || !count) return GF_BAD_ARGUMENT;
    return guard(&pool->ctx->err, [&] {
        if (h.index >= pool->S) fail(GF_BAD_ARGUMENT, "graphFlow::ChunkPool: chunk index out of range");
        *count = pool->lease[h.index];
    });
}

gf_status gf_pool_is_live(gf_pool* pool, gf_chunk_handle h, int* live)
{
    if (!pool || !live) return GF_BAD_ARGUMENT;
    *live = h.index < pool->S && pool->gen[h.index] == h.generation ? 1 : 0;
    return GF_OK;
}

gf_status gf_pool_generation(gf_pool* pool, uint32_t index, uint32_t* generation)
{
    if (!pool || !generation) return GF_BAD_ARGUMENT;
    return guard(&pool->ctx->err, [&] {
        if (index >= pool->S) fail(GF_BAD_ARGUMENT, "graphFlow::ChunkPool: chunk index out of range");
        *generation = pool->gen[index];
    });
}

gf_status gf_pool_validate_live(gf_pool* pool, gf_chunk_handle h)
{
    if (!pool) return GF_BAD_ARGUMENT;
    return guard(&pool->ctx->err, [&] { pool->validate(h); });
}

gf_status gf_pool_base(gf_pool* pool, gf_chunk_handle h, uint64_t* slot)
{
    if (!pool || !slot) return GF_BAD_ARGUMENT;
    if (h.index >= pool->S) return GF_BAD_ARGUMENT;
    *slot = (uint64_t)h.index * pool->C;
    return GF_OK;
}

gf_status gf_pool_column(gf_pool* pool, uint32_t column, void** device_ptr, uint32_t* elem_bytes)
{
    if (!pool || column >= pool->cols.size()) return GF_BAD_ARGUMENT;
    if (device_ptr) *device_ptr = pool->cols[column];
    if (elem_bytes) *elem_bytes = pool->elem[column];
    return GF_OK;
}

gf_status gf_pool_write(gf_pool* pool, uint32_t column, uint64_t first, const void* host, uint64_t count)
{
    if (!pool || column >= pool->cols.size() || first + count > (uint64_t)pool->S * pool->C) return GF_BAD_ARGUMENT;
    return guard(&pool->ctx->err, [&] {
        HIPX(hipMemcpy((char*)pool->cols[column] + first * pool->elem[column], host, count * pool->elem[column],
                       hipMemcpyHostToDevice));
    });
}

gf_status gf_pool_read(gf_pool* pool, uint32_t column, uint64_t first, void* host, uint64_t count)
{
    if (!pool || column >= pool->cols.size() || first + count > (uint64_t)pool->S * pool->C) return GF_BAD_ARGUMENT;
    return guard(&pool->ctx->err, [&] {
        HIPX(hipMemcpy(host, (char*)pool->cols[column] + first * pool->elem[column], count * pool->elem[column],
                       hipMemcpyDeviceToHost));
    });
}

gf_status gf_stage_create_tuned(gf_pool* pool, const gf_node_desc* node, uint32_t node_key, const gf_tuning* tuning,
                                gf_stage** out)
{
    if (!pool || !node || !tuning || !out) return GF_BAD_ARGUMENT;
    return guard(&pool->ctx->err, [&] {
        if (tuning->bin_path > GF_BIN_PATH_HASH) fail(GF_BAD_ARGUMENT, "graphFlow::Tuning: unknown bin path");
        auto s = std::make_unique<gf_stage>();
        s->init(pool, *node, node_key, tuning->seed, tuning->batch_chunks, tuning->tile_items, tuning);
        s->rp.ph = photon_params_of(*tuning);
        *out = s.release();
    });
}

gf_status gf_stage_create(gf_pool* pool, const gf_node_desc* node, uint32_t node_key, uint64_t seed,
                          uint32_t batch_chunks, uint32_t tile_items, gf_stage** out)
{
    if (!pool || !node || !out) return GF_BAD_ARGUMENT;
    return guard(&pool->ctx->err, [&] {
        auto s = std::make_unique<gf_stage>();
        s->init(pool, *node, node_key, seed, batch_chunks, tile_items);
        s->rp.ph = photon_params_of(default_photon_tuning());
        *out = s.release();
    });
}

gf_status gf_stage_destroy(gf_stage* stage)
{
    if (!stage) return GF_BAD_ARGUMENT;
    delete stage;
    return GF_OK;
}

gf_status gf_stage_launch(gf_stage* stage, const gf_chunk_desc* chunks, uint32_t n, uint64_t first_id, void* stream)
{
    if (!stage || !chunks) return GF_BAD_ARGUMENT;
    return guard(&stage->err, [&] {
        stage->launch(chunks, n, first_id, stream ? (hipStream_t)stream : stage->pool->ctx->stream);
    });
}

gf_status gf_transform_batch(gf_stage* stage, const gf_chunk_desc* chunks, uint32_t n, uint64_t first_id, void* stream)
{
    if (!stage || !chunks) return GF_BAD_ARGUMENT;
    return guard(&stage->err, [&] {
        stage->launch_fused(chunks, n, first_id, stream ? (hipStream_t)stream : stage->pool->ctx->stream);
    });
}

gf_status gf_transform_batch_noalloc(gf_pool* pool, const gf_node_desc* node, uint32_t node_key, const gf_tuning* tuning,
                                     const gf_chunk_desc* chunks, uint32_t n, const gf_chunk_handle* reserved,
                                     uint32_t n_reserved, gf_chunk_desc* outputs, uint32_t* result, void* stream)
{
    if (!pool || !node || !tuning || !chunks || !reserved || !outputs || !result) return GF_BAD_ARGUMENT;
    return guard(&pool->ctx->err, [&] {
        const gf_node_desc& nd = *node;
        if (!(nd.role == GF_ROLE_TRANSFORM && is_partition(nd)) || nd.bins == 0u)
            fail(GF_BAD_ARGUMENT, "graphFlow: the no-alloc batch runs partition nodes");
        if (tuning->bin_path > GF_BIN_PATH_HASH) fail(GF_BAD_ARGUMENT, "graphFlow::Tuning: unknown bin path");
        if (n == 0u) fail(GF_PLAN_MISMATCH, "graphFlow: empty batch");
        validate_node(nd, tuning->photon_tally_bins ? tuning->photon_tally_bins : 1024u);
        const uint32_t C = pool->C, B = nd.bins;
        uint32_t want = tuning->tile_items ? tuning->tile_items : 1024u;
        if (on_hash_path(nd, tuning->bin_path)) want = std::min(want, kHashTileLimit);
        const uint32_t r = effective_r(pool->payload, pick_tile_r(want)), T = 256u * r;
        const uint32_t E = hash_entries_of(nd, tuning->bin_path, tuning->hash_table_entries, T);
        const uint32_t rec = record_bytes(pool->payload);
        const size_t lc = E ? count_hash_smem(E) : (size_t)B * 4u;
        const size_t ls = E ? scatter_hash_smem(E) : scatter_smem(B, T, rec, stage_fits(B, T, rec));
        if (lc > 160u * 1024u || ls > 160u * 1024u)
            fail(GF_BAD_ARGUMENT, "graphFlow: partition width or hash table exceeds the workgroup LDS budget");
        // worst case: every input chunk full
        const uint64_t tiles = (uint64_t)n * ceil_div(C, T);
        if (fused_scratch_words(n, B, (uint32_t)std::min<uint64_t>(tiles, UINT32_MAX)) > kNoAllocScratchWords)
            fail(GF_BAD_ARGUMENT, "graphFlow: the batch's tile counts exceed the no-alloc scratch (" +
                                      std::to_string(kNoAllocScratchWords) + " words)");
        FusedArgs a{};
        a.nd = nd;
        a.in = chunks;
        a.n = n;
        a.tile_items = T;
        a.C = C;
        a.hash_entries = E;
        a.reserved = reserved;
        a.n_reserved = n_reserved;
        a.purpose = 2u * node_key;
        a.out = outputs;
        a.scratch = scratch_of_device(pool->ctx->device);
        a.result = result;
        a.group_lds = (uint32_t)((std::max(lc, ls) + 15u) & ~(size_t)15u);
        a.tiles_cap = (uint32_t)((kNoAllocScratchWords - 2ull * (n + 1u) - 2ull * B) / B);
        launch_fused_kernel(pool->payload, r, a, pool->dcols(), stream ? (hipStream_t)stream : pool->ctx->stream);
    });
}

gf_status gf_stage_last_error(gf_stage* stage, char* buffer, uint32_t size)
{
    if (!stage || !buffer || !size) return GF_BAD_ARGUMENT;
    std::snprintf(buffer, size, "%s", stage->err.c_str());
    return GF_OK;
}

gf_status gf_stage_advance(gf_stage* stage, void* stream, uint32_t* allocated, int* more_work)
{
    if (!stage || !allocated || !more_work) return GF_BAD_ARGUMENT;
    return guard(&stage->err, [&] {
        uint32_t a = 0;
        *more_work = stage->advance(stream ? (hipStream_t)stream : stage->pool->ctx->stream, a) ? 1 : 0;
        *allocated = a;
    });
}

gf_status gf_stage_finish(gf_stage* stage, gf_chunk_desc* outputs, uint32_t cap, uint32_t* n_out)
{
    if (!stage || !n_out) return GF_BAD_ARGUMENT;
    if (stage->pending.size() > cap) return GF_BAD_ARGUMENT;
    std::vector<gf_chunk_desc> outs;
    stage->finish(outs);
    for (size_t i = 0; i < outs.size(); ++i) outputs[i] = outs[i];
    *n_out = (uint32_t)outs.size();
    return GF_OK;
}

gf_status gf_stage_abort(gf_stage* stage)
{
    if (!stage) return GF_BAD_ARGUMENT;
    stage->abort();
    return GF_OK;
}

gf_status gf_stage_counts(gf_stage* stage, uint32_t* counts, uint32_t bins)
{
    if (!stage || !counts || bins != stage->totals.size()) return GF_BAD_ARGUMENT;
    std::copy(stage->totals.begin(), stage->totals.end(), counts);
    return GF_OK;
}

gf_status gf_stage_tiles(gf_stage* stage, gf_tile_desc* tiles, uint32_t cap, uint32_t* n)
{
    if (!stage || !n) return GF_BAD_ARGUMENT;
    return guard(&stage->err, [&] {
        *n = stage->n_tiles;
        if (stage->n_tiles > cap) fail(GF_BAD_ARGUMENT, "tile buffer too small");
        if (!stage->n_tiles) return;
        gf_tile_desc* d = dalloc<gf_tile_desc>(stage->n_tiles);
        hipStream_t s = stage->pool->ctx->stream;
        // a batch that travelled in its kernel arguments: its record, inputs and prefixes to the block
        if (!stage->ba.block)
            HIPX(hipMemcpyAsync(stage->d_blk, stage->h_blk, stage->o_out, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_tiles, dim3(std::min<uint32_t>((stage->n_tiles + 255u) / 256u, 1024u)), dim3(256), 0, s,
                           stage->d_batch, stage->pool->C, d);
        HIPX(hipGetLastError());
        HIPX(hipMemcpyAsync(tiles, d, sizeof(gf_tile_desc) * stage->n_tiles, hipMemcpyDeviceToHost, s));
        HIPX(hipStreamSynchronize(s));
        (void)hipFree(d);
    });
}

gf_status gf_flow_create(gf_context* ctx, uint32_t payload, gf_flow** out)
{
    if (!ctx || !out || payload > GF_PAYLOAD_PHOTON) return GF_BAD_ARGUMENT;
    auto f = std::make_unique<gf_flow>();
    f->ctx = ctx;
    f->payload = payload;
    f->tuning.use_graph = 1;
    *out = f.release();
    return GF_OK;
}

gf_status gf_flow_destroy(gf_flow* flow)
{
    if (!flow) return GF_BAD_ARGUMENT;
    flow->eng.reset();
    if (flow->pool)
        for (void* p : flow->pool->cols) (void)hipFree(p);
    for (auto& n : flow->nodes)
        if (n.tally_copies) (void)hipFree(n.tally_copies);
    delete flow;
    return GF_OK;
}

gf_status gf_flow_add_node(gf_flow* flow, const char* name, const gf_node_desc* node, uint32_t* key)
{
    if (!flow || !node || !key) return GF_BAD_ARGUMENT;
    return guard(&flow->err, [&] {
        if (node->role > GF_ROLE_SINK) fail(GF_BAD_ARGUMENT, "graphFlow: unknown node role");
        if (node->role == GF_ROLE_SOURCE && !flow->slots)  // a served flow: one source per job slot
            for (auto& n : flow->nodes)
                if (n.nd.role == GF_ROLE_SOURCE) fail(GF_BAD_ARGUMENT, "graphFlow: a Flow has exactly one source node");
        if (node->role == GF_ROLE_SINK && node->bins != 0) fail(GF_BAD_ARGUMENT, "graphFlow: a sink node has no bins and no routes");
        if (node->role != GF_ROLE_SINK && node->bins == 0)
            fail(GF_BAD_ARGUMENT, "graphFlow: a source or transform node needs at least one bin");
        if (node->bins > 65536u) fail(GF_BAD_ARGUMENT, "graphFlow: partition width exceeds maxSupportedBins");
        if (flow->nodes.size() >= 32767u) fail(GF_OVERFLOW, "graphFlow: graph node count exceeds the RNG stream-purpose range");
        NodeRec r;
        r.name = name ? name : "";
        r.nd = *node;
        r.routes.assign(node->role == GF_ROLE_SINK ? 0u : node->bins, -1);
        *key = (uint32_t)flow->nodes.size();
        flow->nodes.push_back(std::move(r));
    });
}

gf_status gf_flow_connect(gf_flow* flow, uint32_t from, uint32_t bin, uint32_t to)
{
    if (!flow) return GF_BAD_ARGUMENT;
    return guard(&flow->err, [&] {
        if (from >= flow->nodes.size() || to >= flow->nodes.size()) fail(GF_BAD_ARGUMENT, "graphFlow: malformed node handle");
        NodeRec& p = flow->nodes[from];
        if (p.nd.role == GF_ROLE_SINK) fail(GF_BAD_ARGUMENT, "graphFlow: a sink node has no output to connect");
        if (flow->nodes[to].nd.role == GF_ROLE_SOURCE) fail(GF_BAD_ARGUMENT, "graphFlow: a route may not target the source node");
        if (bin >= p.nd.bins) fail(GF_BAD_ARGUMENT, "graphFlow: output bin out of range");
        if (p.routes[bin] != -1) fail(GF_BAD_ARGUMENT, "graphFlow: route for this output is already connected");
        p.routes[bin] = (int32_t)to;
    });
}

gf_status gf_flow_configure(gf_flow* flow, const gf_tuning* t)
{
    if (!flow || !t) return GF_BAD_ARGUMENT;
    return guard(&flow->err, [&] {
        if (t->schedule > GF_SCHEDULE_DEVICE) fail(GF_BAD_ARGUMENT, "graphFlow::Tuning: unknown schedule");
        if (t->bin_path > GF_BIN_PATH_HASH) fail(GF_BAD_ARGUMENT, "graphFlow::Tuning: unknown bin path");
        if (t->inflight_batches > kMaxNodes) fail(GF_BAD_ARGUMENT, "graphFlow::Tuning: inflightBatches exceeds the engine's lanes");
        if (t->inflight_batches > 1u && t->schedule == GF_SCHEDULE_DEVICE && t->device_engine == 3u)
            fail(GF_BAD_ARGUMENT, "graphFlow::Tuning: the team engine runs one batch per node and step (inflightBatches 1)");
        if (t->source_chunks > 64u) fail(GF_BAD_ARGUMENT, "graphFlow::Tuning: sourceChunks is at most 64");
        if (t->source_chunks > 1u && t->schedule == GF_SCHEDULE_DEVICE && t->device_engine == 3u)
            fail(GF_BAD_ARGUMENT, "graphFlow::Tuning: the team engine admits one source chunk per step (sourceChunks 1)");
        flow->tuning = *t;
    });
}

// Autotuner: probe runs over the device-schedule knobs the caller left open (engine 0, Q 0), on
// this flow's own graph, payload and routing; the fastest candidate configures the flow.
// Candidates: team engine with 16/24/32 workgroups (at most the CUs of one XCD), wide engine with
// 1024- or 4096-item tiles and 64, 256 or 1024 workgroups, the one-workgroup engine; Q in {4, 8, 16}
// when unset. None of them changes a result: every engine runs the same step semantics, and Q
// changes batch composition but no particle's trajectory.
gf_status gf_flow_autotune(gf_flow* flow, uint32_t source, uint64_t probe_items, gf_tuning* chosen, double* best_ms)
{
    if (!flow || !chosen) return GF_BAD_ARGUMENT;
    return guard(&flow->err, [&] {
        check_graph(*flow, source);
        const gf_tuning base = flow->tuning;
        if (base.schedule != GF_SCHEDULE_DEVICE) {
            *chosen = base;
            return;
        }
        const uint64_t n = probe_items ? probe_items : (1ull << 21);
        std::vector<gf_tuning> cand;
        std::vector<uint32_t> qs;
        if (base.batch_chunks) qs.push_back(base.batch_chunks);
        else qs = {4u, 8u, 16u};
        const uint32_t xcd_cus = std::max<uint32_t>(1u, std::min<uint32_t>(32u, flow->ctx->cus / 8u));
        for (uint32_t q : qs) {
            gf_tuning t = base;
            t.batch_chunks = q;
            t.use_graph = base.use_graph >= 2u ? 1u : base.use_graph;
            if (base.device_engine == 0u || base.device_engine == 3u)
                for (uint32_t k : {16u, 24u, 32u})
                    if (k <= xcd_cus || k == 16u) {
                        gf_tuning c = t;
                        c.device_engine = 3u;
                        c.max_workgroups = std::min(k, xcd_cus);
                        cand.push_back(c);
                    }
            if (base.device_engine == 0u || base.device_engine == 1u)
                for (uint32_t ti : {1024u, 4096u})
                    for (uint32_t wg : {64u, 256u, 1024u}) {
                        gf_tuning c = t;
                        c.device_engine = 1u;
                        c.tile_items = ti;
                        c.max_workgroups = wg;
                        cand.push_back(c);
                    }
            if (base.device_engine == 0u || base.device_engine == 2u) {
                gf_tuning c = t;
                c.device_engine = 2u;
                cand.push_back(c);
            }
        }
        double best = 1e300;
        gf_tuning pick = base;
        for (const gf_tuning& c : cand) {
            flow->tuning = c;
            gf_stats st{};
            double ms = 1e300;
            const gf_status rc = guard(nullptr, [&] {
                for (int rep = 0; rep < 2; ++rep) {  // the first run builds the engine
                    RunShape sh = resolve(*flow, n);
                    ensure_pool(*flow, sh);
                    auto t0 = std::chrono::steady_clock::now();
                    run_device(*flow, source, n, sh, st, false);
                    HIPX(hipStreamSynchronize(flow->ctx->stream));
                    ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
                }
            });
            if (rc == GF_OK && ms < best) {
                best = ms;
                pick = c;
            }
        }
        if (best >= 1e300) fail(GF_BAD_ARGUMENT, "autotune: no candidate ran");
        flow->tuning = pick;
        flow->tuning.use_graph = base.use_graph;
        *chosen = flow->tuning;
        if (best_ms) *best_ms = best;
    });
}

gf_status gf_flow_run(gf_flow* flow, uint32_t source, uint64_t items, gf_stats* stats)
{
    if (!flow) return GF_BAD_ARGUMENT;
    return guard(&flow->err, [&] {
        check_graph(*flow, source);
        if (items != 0 && items - 1u > 0xFFFFFFFFull) fail(GF_OVERFLOW, "graphFlow: item count exceeds the 32 bit RNG stream-id range");
        RunShape sh = resolve(*flow, items);
        ensure_pool(*flow, sh);
        gf_stats st{};
        st.pool_chunks = sh.S;
        st.chunk_capacity = sh.C;
        st.live_particles = sh.P;
        auto t0 = std::chrono::steady_clock::now();
        if (items == 0) {
            // nothing to admit
        } else if (flow->tuning.schedule == GF_SCHEDULE_DEVICE) {
            run_device(*flow, source, items, sh, st, flow->tuning.use_graph >= 2u);
        } else {
            run_host(*flow, source, items, sh, st);
        }
        st.wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        if (stats) *stats = st;
    });
}

gf_status gf_flow_last_error(gf_flow* flow, char* buffer, uint32_t size)
{
    if (!flow || !buffer || size == 0) return GF_BAD_ARGUMENT;
    std::snprintf(buffer, size, "%s", flow->err.c_str());
    return GF_OK;
}

gf_status gf_flow_trace(gf_flow* flow, uint32_t* host, uint64_t cap, uint64_t* len)
{
    if (!flow || !len) return GF_BAD_ARGUMENT;
    *len = flow->trace.size();
    if (host) std::memcpy(host, flow->trace.data(), sizeof(uint32_t) * std::min<uint64_t>(cap, flow->trace.size()));
    return GF_OK;
}

gf_status gf_flow_kernel_times(gf_flow* flow, double* ms, uint64_t* launches, uint32_t kinds)
{
    if (!flow || !ms || !launches) return GF_BAD_ARGUMENT;
    for (uint32_t i = 0; i < kinds; ++i) {
        ms[i] = i < kKernelKinds ? flow->kt_ms[i] : 0.0;
        launches[i] = i < kKernelKinds ? flow->kt_n[i] : 0u;
    }
    return GF_OK;
}

gf_status gf_flow_phase_profile(gf_flow* flow, uint64_t* ticks, uint32_t words)
{
    if (!flow || !ticks || !flow->eng || !flow->eng->e.prof) return GF_BAD_ARGUMENT;
    return guard(&flow->err, [&] {
        HIPX(hipMemcpy(ticks, flow->eng->e.prof, sizeof(uint64_t) * std::min<uint32_t>(words, PROF_WORDS),
                       hipMemcpyDeviceToHost));
        // words beyond the phase totals: per-member team totals, kMemberProf words per member
        if (words > PROF_WORDS && flow->eng->team && flow->eng->ta.mprof)
            HIPX(hipMemcpy(ticks + PROF_WORDS, flow->eng->ta.mprof,
                           sizeof(uint64_t) * std::min<uint32_t>(words - PROF_WORDS, 64u * kMemberProf), hipMemcpyDeviceToHost));
    });
}

gf_status gf_flow_item_passes(gf_flow* flow, uint64_t* partition_items, uint64_t* transform_items)
{
    if (!flow || !partition_items || !transform_items) return GF_BAD_ARGUMENT;
    *partition_items = flow->part_items;
    *transform_items = flow->xf_items;
    return GF_OK;
}

gf_status gf_hist_rows(gf_context* ctx, const uint64_t* counts, uint64_t n, uint32_t lo, uint64_t* host_rows,
                       uint64_t cap, uint64_t* rows)
{
    if (!ctx || !counts || !rows) return GF_BAD_ARGUMENT;
    return guard(&ctx->err, [&] {
        const uint64_t seg = 1u << 16;
        const uint32_t blocks = (uint32_t)std::max<uint64_t>(ceil_div(n, seg), 1u);
        uint32_t* bnz = dalloc<uint32_t>(blocks);
        hipStream_t s = ctx->stream;
        hipLaunchKernelGGL(k_nz_count, dim3(blocks), dim3(kThreads), 0, s, (const unsigned long long*)counts, n, seg, bnz);
        HIPX(hipGetLastError());
        std::vector<uint32_t> h(blocks);
        HIPX(hipMemcpyAsync(h.data(), bnz, sizeof(uint32_t) * blocks, hipMemcpyDeviceToHost, s));
        HIPX(hipStreamSynchronize(s));
        uint64_t total = 0;
        for (auto& v : h) {
            uint32_t x = v;
            v = (uint32_t)total;
            total += x;
        }
        *rows = total;
        if (host_rows && total <= cap && total) {
            HIPX(hipMemcpyAsync(bnz, h.data(), sizeof(uint32_t) * blocks, hipMemcpyHostToDevice, s));
            unsigned long long* d = dalloc<unsigned long long>(2 * total);
            hipLaunchKernelGGL(k_nz_write, dim3(blocks), dim3(kThreads), 0, s, (const unsigned long long*)counts, n, seg, bnz, lo,
                               d);
            HIPX(hipGetLastError());
            HIPX(hipMemcpyAsync(host_rows, d, sizeof(uint64_t) * 2 * total, hipMemcpyDeviceToHost, s));
            HIPX(hipStreamSynchronize(s));
            (void)hipFree(d);
        }
        (void)hipFree(bnz);
    });
}

}  // extern "C"

#include "gf_server.hpp"
