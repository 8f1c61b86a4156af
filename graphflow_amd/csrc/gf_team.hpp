// Team engine of the batched form: one persistent launch in which a team of K workgroups that
// share one XCD (and therefore one L2) runs whole steps, with one inter-workgroup barrier per step.
//
// Every team member keeps a full copy of the scheduler state in LDS (ready rings, a window of the
// free stack, the selected batches, the plan) and applies the same deterministic updates, so no
// scheduling decision is ever broadcast. The only data the members exchange per step are the
// records themselves and per-segment bin counts:
//
//   * a segment is 256 consecutive slots of one chunk;
//   * whoever writes a record into a chunk that is routed to a partition node also classifies the
//     record with that node's classifier and adds one to cnt[chunk][gen&1][segment][bin] (a
//     wave-aggregated atomic; the last slot of the row counts invalid bins);
//   * when the partition node later takes the chunk, every member reads those counts right
//     after the barrier and derives the exact plan (per-bin totals, output chunk counts, the flat
//     start of every (segment, bin)) before it has loaded a single record.
//
// The stable canonical order (bin-major, then batch input order) follows from the per-segment
// prefixes plus a wave-level rank inside each segment, so a member scatters its segments without
// seeing any other member's records. The scheduling rules are those of engine_select/engine_plan/
// engine_retire (sched/GraphRunner.hpp:902-1133 applied once per node per step).
//
// Coherence: stores are plain vector stores (they write through the CU's L1 into the XCD's L2);
// every load of data written by another member in the same launch is a relaxed agent-scope load,
// which bypasses the L1; before arriving at the barrier every wave waits for its stores and
// atomics. Members are chosen by their hardware XCC id, so they always share the L2 that serves
// those loads. Scheduler state survives between launches in the wide engine's global layout, so
// the team can hand the run over to the wide engine when a step exceeds its LDS limits.
#pragma once

#include "gf_engine.hpp"

namespace gfd {

constexpr uint32_t kTeamNodes = 16;
constexpr uint32_t kTeamBins = 32;  // count row stride: bins of the widest partition + invalid slot
constexpr uint32_t kTeamRoutes = 512;
constexpr uint32_t kSeg = 256;
constexpr uint32_t kTeamRing = 2048;   // ready descriptors over all nodes
constexpr uint32_t kTeamStack = 2048;  // free-stack window
constexpr uint32_t kTeamMargin = 512;  // pops or pushes of one step
constexpr uint32_t kTeamIn = 1024;     // input chunks of one step
constexpr uint32_t kTeamOut = 512;     // output chunks of one step
constexpr uint32_t kTeamSegs = 2048;   // segments of one step
constexpr uint32_t kTeamPrefix = 4096; // [segment][bin] words of the step's partitions
constexpr uint32_t kTeamZero = 1024;   // consumed count rows cleared in the next step
constexpr uint32_t kTeamWaves = 5;     // waves 0-3 scatter segments, wave 4 schedules
constexpr uint32_t kTeamThreads = 64u * kTeamWaves;
constexpr uint32_t kTeamSpin = 1u << 24;

enum { TW_ARRIVED = 0, TW_XCC0 = 1, TW_BAR = 2, TW_ERR = 4, TW_REG = 8, TW_WORDS = 32 };  // TW_ERR + step parity
enum { EXIT_NONE = 0, EXIT_DONE = 1, EXIT_STALL = 2, EXIT_HANDOFF = 3, EXIT_LIMIT = 4, EXIT_STATUS = 5, EXIT_SKIP = 6 };
enum { PROF_T_WAIT = 9, PROF_T_COUNTS = 10, PROF_T_PLAN = 11, PROF_T_SEG = 12, PROF_T_SELECT = 13, PROF_T_STEPS = 14 };

struct TeamArgs {
    Engine e;
    uint32_t* cnt;     // [S][2][tpc][bb] segment bin counts
    uint32_t* zero_g;  // [0] = n, [1..n] = chunk | (gen & 1) << 31 rows still to clear
    uint32_t* tw;      // team words, cleared before every launch
    uint32_t K, tpc, bb, max_steps, rc, n_routes;
};

struct TeamLds {
    gf_node_desc nd[kTeamNodes];
    uint32_t kind[kTeamNodes], bound[kTeamNodes], route_off[kTeamNodes], purpose[kTeamNodes];
    int32_t route[kTeamRoutes], fuse[kTeamRoutes];
    unsigned long long head[kTeamNodes], tail[kTeamNodes];
    gf_chunk_desc ring[kTeamRing];
    unsigned long long fs[kTeamStack];
    // selected batches
    uint32_t active[kTeamNodes], n_in[kTeamNodes], particles[kTeamNodes], n_out[kTeamNodes], pre_off[kTeamNodes];
    uint32_t in_off[kTeamNodes + 1], seg_off[kTeamNodes + 1], out_off[kTeamNodes + 1];
    uint32_t st_node[kTeamNodes];
    unsigned long long first_id[kTeamNodes], digest_base[kTeamNodes];
    gf_chunk_desc in[kTeamIn];
    uint32_t in_items[kTeamIn];
    uint32_t seg_map[kTeamSegs];  // node << 28 | input << 18 | segment of the chunk
    gf_chunk_desc out[kTeamOut];
    uint32_t pre[kTeamPrefix];
    uint32_t tot[kTeamNodes * kTeamBins];
    uint32_t binbase[kTeamNodes * kTeamBins];
    uint32_t zero[kTeamZero];
    unsigned long long generated, sunk, step, batches, part_items, xf_items, src_items, sink_items, peak_live, trace_len;
    unsigned long long digest[kTeamNodes];
    uint32_t fs_base, top, zero_n, status, exit_code, recenter, total_out;
    uint32_t seg_err;  // a segment of this member reported a device status this step
    uint32_t reserved;  // credits the select phase reserved for this step's outputs
    uint32_t shv[8];
};

__device__ inline uint32_t xcc_id() { return __builtin_amdgcn_s_getreg(20 | (3 << 11)) & 0xFu; }  // HW_REG_XCC_ID[3:0]

__device__ inline void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ inline uint32_t seg_pack(uint32_t k, uint32_t i, uint32_t gg) { return k << 28 | i << 18 | gg; }

// Rank of each valid lane among the lanes of its bin, after the `run` earlier items of that bin in
// the same segment (lane b of `run` holds bin b's running count; bins < 64).
__device__ inline uint32_t seg_rank(bool v, uint32_t bin, uint32_t& run)
{
    uint64_t act = __ballot(v);
    uint32_t pos = 0;
    const uint32_t lane = lane_id();
    while (act) {
        const int leader = __ffsll((unsigned long long)act) - 1;
        const uint32_t lb = (uint32_t)__builtin_amdgcn_readlane((int)bin, leader);
        const bool mine = v && bin == lb;
        const uint64_t m = __ballot(mine);
        const uint32_t base = (uint32_t)__builtin_amdgcn_readlane((int)run, (int)lb);
        if (mine) pos = base + (uint32_t)__popcll(m & lanemask_lt());
        if (lane == lb) run += (uint32_t)__popcll(m);
        act &= ~m;
    }
    return pos;
}

// arr[key] += number of lanes with that key, one atomic per distinct key of the wave.
__device__ inline void agg_add(bool v, uint32_t key, uint32_t* arr)
{
    uint64_t act = __ballot(v);
    const uint32_t lane = lane_id();
    while (act) {
        const int leader = __ffsll((unsigned long long)act) - 1;
        const uint32_t lk = (uint32_t)__builtin_amdgcn_readlane((int)key, leader);
        const uint64_t m = __ballot(v && key == lk);
        if (lane == (uint32_t)leader) atomicAdd(&arr[lk], (uint32_t)__popcll(m));
        act &= ~m;
    }
}

// ---- team formation --------------------------------------------------------------------------
// Every workgroup of the grid registers under its XCC id; the members are the first K of the XCC
// that block 0 runs on. Returns the member rank (or ~0u) in sh[0] and the team size in sh[1].
__device__ inline void team_join(const TeamArgs& a, uint32_t* sh)
{
    if (threadIdx.x == 0) {
        uint32_t* tw = a.tw;
        const uint32_t x = xcc_id();
        if (blockIdx.x == 0) atomicExch(&tw[TW_XCC0], x + 1u);
        const uint32_t slot = atomicAdd(&tw[TW_REG + x], 1u);
        __builtin_amdgcn_s_waitcnt(0);
        atomicAdd(&tw[TW_ARRIVED], 1u);
        uint32_t spins = 0, ok = 1u;
        while (ld_u32(tw, TW_ARRIVED) < gridDim.x)
            if (++spins > kTeamSpin) {
                ok = 0u;
                break;
            }
        const uint32_t x0 = ld_u32(tw, TW_XCC0) - 1u;
        const uint32_t reg = ld_u32(tw, TW_REG + (x0 & 15u));
        sh[0] = (ok && x == x0 && slot < a.K) ? slot : ~0u;
        sh[1] = reg < a.K ? reg : a.K;
        if (!ok) set_status(&a.e.ctl[CTL_STATUS], GF_DEVICE_ERROR);
    }
    __syncthreads();
}

// Thread 0: wait until every member has arrived `target` times in total (bounded spin).
__device__ inline bool team_wait(const TeamArgs& a, uint32_t target)
{
    uint32_t spins = 0;
    while (ld_u32(a.tw, TW_BAR) < target)
        if (++spins > kTeamSpin) return false;
    return true;
}

// ---- state load / flush ------------------------------------------------------------------------
__device__ inline void team_load(const TeamArgs& a, TeamLds& L)
{
    const Engine& e = a.e;
    const uint32_t tid = threadIdx.x, n = e.n_nodes, rcm = a.rc - 1u;
    for (uint32_t i = tid; i < n * (uint32_t)(sizeof(gf_node_desc) / 4u); i += kTeamThreads)
        ((uint32_t*)L.nd)[i] = ((const uint32_t*)e.nodes)[i];
    for (uint32_t i = tid; i < a.n_routes; i += kTeamThreads) {
        L.route[i] = e.routes[i];
        L.fuse[i] = e.fuse[i];
    }
    if (tid < n) {
        L.kind[tid] = e.arr[tid].kind;
        L.bound[tid] = e.arr[tid].reserve_bound;
        L.purpose[tid] = e.arr[tid].purpose;
        L.route_off[tid] = e.nodes[tid].route_offset;
        L.head[tid] = e.st[tid].head;
        L.tail[tid] = e.st[tid].tail;
        L.digest[tid] = e.ctr[CNT_DIGEST0 + tid];
        L.active[tid] = 0u;
    }
    if (tid == 0) {
        L.top = e.ctl[CTL_TOP];
        L.fs_base = L.top > kTeamStack / 2u ? L.top - kTeamStack / 2u : 0u;
        L.generated = e.ctr[CNT_GENERATED];
        L.sunk = e.ctr[CNT_SUNK];
        L.step = e.ctr[CNT_STEP];
        L.batches = e.ctr[CNT_BATCHES];
        L.part_items = e.ctr[CNT_PART_ITEMS];
        L.xf_items = e.ctr[CNT_XF_ITEMS];
        L.src_items = e.ctr[CNT_SRC_ITEMS];
        L.sink_items = e.ctr[CNT_SINK_ITEMS];
        L.trace_len = e.ctr[CNT_TRACE_LEN];
        L.peak_live = e.ctr[CNT_PEAK_LIVE];
        L.zero_n = a.zero_g[0];
        L.status = 0u;
        L.exit_code = EXIT_NONE;
        L.recenter = 0u;
        uint32_t big = 0;
        for (uint32_t k = 0; k < n; ++k)
            if (e.st[k].tail - e.st[k].head > a.rc) big = 1u;
        if (big || L.zero_n > kTeamZero) L.exit_code = EXIT_SKIP;  // hand over without touching the state
    }
    __syncthreads();
    if (L.exit_code) return;
    for (uint32_t k = 0; k < n; ++k) {
        const unsigned long long h = L.head[k], cnt = L.tail[k] - L.head[k];
        for (uint32_t i = tid; i < (uint32_t)cnt; i += kTeamThreads)
            L.ring[k * a.rc + (uint32_t)((h + i) & rcm)] = e.rings[(size_t)k * e.ring_cap + (size_t)((h + i) % e.ring_cap)];
    }
    for (uint32_t i = tid; i < L.top - L.fs_base; i += kTeamThreads) L.fs[i] = e.free_stack[L.fs_base + i];
    for (uint32_t i = tid; i < L.zero_n; i += kTeamThreads) L.zero[i] = a.zero_g[1 + i];
    __syncthreads();
}

// Member 0 writes the pre-select state back in the wide engine's layout.
__device__ inline void team_flush(const TeamArgs& a, TeamLds& L)
{
    const Engine& e = a.e;
    const uint32_t tid = threadIdx.x, n = e.n_nodes, rcm = a.rc - 1u;
    for (uint32_t k = 0; k < n; ++k) {
        const unsigned long long h = L.head[k], cnt = L.tail[k] - L.head[k];
        for (uint32_t i = tid; i < (uint32_t)cnt; i += kTeamThreads)
            e.rings[(size_t)k * e.ring_cap + (size_t)((h + i) % e.ring_cap)] = L.ring[k * a.rc + (uint32_t)((h + i) & rcm)];
    }
    if (tid < n) {
        NodeSt s{};
        s.head = L.head[tid];
        s.tail = L.tail[tid];
        e.st[tid] = s;
        e.ctr[CNT_DIGEST0 + tid] = L.digest[tid];
    }
    for (uint32_t i = tid; i < L.top - L.fs_base; i += kTeamThreads) e.free_stack[L.fs_base + i] = L.fs[i];
    for (uint32_t i = tid; i < L.zero_n; i += kTeamThreads) a.zero_g[1 + i] = L.zero[i];
    if (tid == 0) {
        a.zero_g[0] = L.zero_n;
        e.ctl[CTL_TOP] = L.top;
        e.ctr[CNT_GENERATED] = L.generated;
        e.ctr[CNT_SUNK] = L.sunk;
        e.ctr[CNT_STEP] = L.step;
        e.ctr[CNT_BATCHES] = L.batches;
        e.ctr[CNT_PART_ITEMS] = L.part_items;
        e.ctr[CNT_XF_ITEMS] = L.xf_items;
        e.ctr[CNT_SRC_ITEMS] = L.src_items;
        e.ctr[CNT_SINK_ITEMS] = L.sink_items;
        e.ctr[CNT_TRACE_LEN] = L.trace_len;
        e.ctr[CNT_PEAK_LIVE] = L.peak_live;
        if (L.exit_code == EXIT_DONE) {
            e.ctl[CTL_DONE] = 1u;
            if (L.top != e.S) set_status(&e.ctl[CTL_STATUS], GF_BAD_ARGUMENT);
        } else if (L.exit_code == EXIT_STALL) {
            set_status(&e.ctl[CTL_STATUS], GF_STALL);
        } else if (L.exit_code == EXIT_HANDOFF) {
            e.ctl[CTL_HANDOFF] = 1u;
        }
    }
}

// ---- select (wave 4): the next step's batches, or the reason to stop -------------------------
__device__ inline void team_select(const TeamArgs& a, TeamLds& L, bool at_limit)
{
    const Engine& e = a.e;
    const uint32_t lane = lane_id(), n = e.n_nodes, C = e.C, rcm = a.rc - 1u;
    if (lane == 0) {
        const unsigned long long P = e.P, Q = e.Q;
        uint32_t reserved = 0, in_tot = 0, seg_tot = 0, pre_tot = 0, any = 0;
        unsigned long long gen = L.generated;
        for (uint32_t k = 0; k < n; ++k) {
            const uint32_t kind = L.kind[k];
            const unsigned long long avail = L.tail[k] - L.head[k];
            uint32_t nin = 0, parts = 0, nseg = 0, act = 0;
            L.in_off[k] = in_tot;
            L.seg_off[k] = seg_tot;
            if (kind == KIND_SOURCE) {
                const unsigned long long live = gen - L.sunk, remaining = e.N - gen;
                unsigned long long min_fill = C < P ? C : P;
                if (remaining < min_fill) min_fill = remaining;
                if (gen < e.N && live < P && P - live >= min_fill && L.top - reserved >= 1u) {
                    unsigned long long fill = remaining;
                    if (fill > P - live) fill = P - live;
                    if (fill > C) fill = C;
                    parts = (uint32_t)fill;
                    L.first_id[k] = gen;
                    gen += fill;
                    reserved += 1u;
                    act = 1u;
                    nseg = (parts + kSeg - 1u) / kSeg;
                }
            } else if (avail) {
                const unsigned long long h = L.head[k];
                const gf_chunk_desc* rg = L.ring + k * a.rc;
                uint32_t want = kind == KIND_SINK ? (uint32_t)avail : (uint32_t)(avail < Q ? avail : Q);
                for (uint32_t i = 0; i < want; ++i) parts += rg[(uint32_t)((h + i) & rcm)].fill;
                if (kind != KIND_SINK) {
                    const uint32_t B = kind == KIND_PARTITION ? L.nd[k].bins : 1u;
                    while (want) {
                        const uint64_t bound = bound_of(parts, C, B);
                        if ((uint64_t)(L.top - reserved) >= bound) {
                            reserved += (uint32_t)bound;
                            break;
                        }
                        --want;
                        parts -= rg[(uint32_t)((h + want) & rcm)].fill;
                    }
                }
                nin = want;
                act = nin ? 1u : 0u;
                for (uint32_t i = 0; i < nin; ++i) nseg += (rg[(uint32_t)((h + i) & rcm)].fill + kSeg - 1u) / kSeg;
                if (kind == KIND_PARTITION && act) {
                    L.pre_off[k] = pre_tot;
                    pre_tot += nseg * (L.nd[k].bins + 1u);
                }
            }
            L.active[k] = act;
            L.n_in[k] = nin;
            L.particles[k] = parts;
            any |= act;
            in_tot += nin;
            seg_tot += nseg;
        }
        L.in_off[n] = in_tot;
        L.seg_off[n] = seg_tot;
        uint32_t code = EXIT_NONE;
        if (!any) {
            bool drained = true;
            for (uint32_t k = 0; k < n; ++k)
                if (L.tail[k] != L.head[k]) drained = false;
            code = (drained && L.generated >= e.N) ? EXIT_DONE : EXIT_STALL;
        } else if (at_limit) {
            code = EXIT_LIMIT;
        } else {
            bool fits = in_tot <= kTeamIn && seg_tot <= kTeamSegs && pre_tot <= kTeamPrefix && reserved <= kTeamOut &&
                        reserved <= kTeamMargin && in_tot <= kTeamMargin && in_tot <= kTeamZero;
            for (uint32_t k = 0; k < n && fits; ++k)
                if ((L.tail[k] - L.head[k]) - L.n_in[k] + reserved > a.rc) fits = false;
            if (!fits) {
                code = EXIT_HANDOFF;
            } else {
                for (uint32_t k = 0; k < n; ++k) {
                    L.head[k] += L.n_in[k];
                    if (L.kind[k] == KIND_SINK && L.active[k]) {
                        L.digest_base[k] = L.digest[k];
                        L.digest[k] += L.n_in[k];
                    }
                }
                L.generated = gen;
                L.reserved = reserved;
                const uint32_t nb = L.top > kTeamStack / 2u ? L.top - kTeamStack / 2u : 0u;
                L.recenter = (nb != L.fs_base && (L.top < L.fs_base + kTeamMargin || L.top + kTeamMargin > L.fs_base + kTeamStack)) ? 1u : 0u;
            }
        }
        L.exit_code = code;
    }
    wave_sync();
    if (L.exit_code) return;
    // input lists, item prefixes and segment maps
    for (uint32_t k = 0; k < n; ++k) {
        if (!L.active[k]) continue;
        if (L.kind[k] == KIND_SOURCE) {
            const uint32_t ns = L.seg_off[k + 1] - L.seg_off[k];
            for (uint32_t g = lane; g < ns; g += 64u) L.seg_map[L.seg_off[k] + g] = seg_pack(k, 0u, g);
            continue;
        }
        const uint32_t m = L.n_in[k];
        const unsigned long long h0 = L.head[k] - m;
        const gf_chunk_desc* rg = L.ring + k * a.rc;
        uint32_t carry_i = 0, carry_s = 0;
        for (uint32_t i0 = 0; i0 < m; i0 += 64u) {
            const uint32_t i = i0 + lane;
            gf_chunk_desc d{0, 0, 0, 0};
            if (i < m) {
                d = rg[(uint32_t)((h0 + i) & rcm)];
                L.in[L.in_off[k] + i] = d;
            }
            const uint32_t fill = i < m ? d.fill : 0u, ns = (fill + kSeg - 1u) / kSeg;
            const uint32_t inc_i = wave_inscan(fill), inc_s = wave_inscan(ns);
            if (i < m) {
                L.in_items[L.in_off[k] + i] = carry_i + inc_i - fill;
                const uint32_t sp = L.seg_off[k] + carry_s + inc_s - ns;
                for (uint32_t g = 0; g < ns; ++g) L.seg_map[sp + g] = seg_pack(k, i, g);
            }
            carry_i += (uint32_t)__shfl((int)inc_i, 63);
            carry_s += (uint32_t)__shfl((int)inc_s, 63);
        }
    }
    wave_sync();
}

// ---- after the barrier: status, segment counts, clearing, free-stack window ----------------------
__device__ inline void team_counts(const TeamArgs& a, TeamLds& L, uint32_t me, uint32_t T, uint32_t it)
{
    const Engine& e = a.e;
    const uint32_t tid = threadIdx.x, n = e.n_nodes;
    const uint32_t row = a.tpc * a.bb;
    // errors members published before the barrier of the previous step (parity word: a member
    // that is already one step ahead writes the other word)
    if (tid == 0) {
        L.status = it ? ld_u32(a.tw, TW_ERR + ((it - 1u) & 1u)) : 0u;
        L.seg_err = 0u;
    }
    // clear the count rows of the chunks partitions consumed in the previous step
    const uint32_t zw = L.zero_n * row;
    for (uint32_t x = me * kTeamThreads + tid; x < zw; x += T * kTeamThreads) {
        const uint32_t z = L.zero[x / row];
        a.cnt[((size_t)(z & 0x7FFFFFFFu) * 2u + (z >> 31)) * row + x % row] = 0u;
    }
    // segment counts of every partition batch of this step
    for (uint32_t k = 0; k < n; ++k) {
        if (!L.active[k] || L.kind[k] != KIND_PARTITION) continue;
        const uint32_t B1 = L.nd[k].bins + 1u, ns = L.seg_off[k + 1] - L.seg_off[k];
        for (uint32_t x = tid; x < ns * B1; x += kTeamThreads) {
            const uint32_t g = x / B1, b = x - g * B1;
            const uint32_t sm = L.seg_map[L.seg_off[k] + g];
            const gf_chunk_desc d = L.in[L.in_off[k] + ((sm >> 18) & 1023u)];
            L.pre[L.pre_off[k] + x] = ld_u32(a.cnt, (((size_t)d.index * 2u + (d.generation & 1u)) * a.tpc + (sm & 0x3FFFFu)) * a.bb + b);
        }
    }
    // move the free-stack window so that this step's pops and pushes stay inside it
    if (L.recenter) {
        const uint32_t old = L.fs_base, top = L.top;
        const uint32_t nb = top > kTeamStack / 2u ? top - kTeamStack / 2u : 0u;
        constexpr uint32_t per = (kTeamStack + kTeamThreads - 1u) / kTeamThreads;
        unsigned long long keep[per];
        // entries of [max(old, nb), top) stay in LDS at a new offset; member 0 spills [old, nb)
#pragma unroll
        for (uint32_t r = 0; r < per; ++r) {
            const uint32_t gi = (nb > old ? nb : old) + tid + r * kTeamThreads;
            keep[r] = gi < top ? L.fs[gi - old] : 0ull;
        }
        if (me == 0)
            for (uint32_t gi = old + tid; gi < nb && gi < top; gi += kTeamThreads) e.free_stack[gi] = L.fs[gi - old];
        __syncthreads();
#pragma unroll
        for (uint32_t r = 0; r < per; ++r) {
            const uint32_t gi = (nb > old ? nb : old) + tid + r * kTeamThreads;
            if (gi < top) L.fs[gi - nb] = keep[r];
        }
        // entries [nb, old) come back from global memory (written by an earlier spill or launch)
        for (uint32_t gi = nb + tid; gi < old; gi += kTeamThreads) L.fs[gi - nb] = ld_u64(e.free_stack, gi);
        if (tid == 0) {
            L.fs_base = nb;
            L.recenter = 0u;
        }
    }
    __syncthreads();
}

// ---- plan (all threads): totals, exact output counts, pops, flat starts ------------------------
__device__ inline void team_plan(const TeamArgs& a, TeamLds& L)
{
    const Engine& e = a.e;
    const uint32_t tid = threadIdx.x, n = e.n_nodes, C = e.C, w = tid >> 6, lane = lane_id();
    for (uint32_t k = 0; k < n; ++k) {
        if (!L.active[k] || L.kind[k] != KIND_PARTITION) continue;
        const uint32_t B1 = L.nd[k].bins + 1u, ns = L.seg_off[k + 1] - L.seg_off[k];
        uint32_t* m = L.pre + L.pre_off[k];
        for (uint32_t b = w; b < B1; b += kTeamWaves) {
            uint32_t carry = 0;
            for (uint32_t g0 = 0; g0 < ns; g0 += 64u) {
                const uint32_t g = g0 + lane;
                const uint32_t x = g < ns ? m[g * B1 + b] : 0u;
                const uint32_t inc = wave_inscan(x);
                if (g < ns) m[g * B1 + b] = carry + inc - x;
                carry += (uint32_t)__shfl((int)inc, 63);
            }
            if (lane == 0) L.tot[k * kTeamBins + b] = carry;
        }
    }
    __syncthreads();
    if (tid < n) {
        const uint32_t k = tid, kind = L.kind[k];
        uint32_t R = 0, st = 0;
        if (L.active[k]) {
            if (kind == KIND_PARTITION) {
                const uint32_t B = L.nd[k].bins;
                uint32_t sum = 0;
                for (uint32_t b = 0; b < B; ++b) {
                    const uint32_t t = L.tot[k * kTeamBins + b];
                    L.binbase[k * kTeamBins + b] = R;
                    R += (t + C - 1u) / C;
                    sum += t;
                }
                if (L.tot[k * kTeamBins + B] != 0u) st = GF_INVALID;
                else if (sum != L.particles[k] || R > L.bound[k]) st = GF_PLAN_MISMATCH;
            } else if (kind == KIND_TRANSFORM) {
                R = (L.particles[k] + C - 1u) / C;
                if (L.particles[k] == 0u || R > L.bound[k]) st = GF_PLAN_MISMATCH;
            } else if (kind == KIND_SOURCE) {
                R = 1u;
            }
        }
        L.n_out[k] = R;
        L.st_node[k] = st;
    }
    __syncthreads();
    if (tid == 0) {
        uint32_t off = 0, st = 0, inv = 0;
        for (uint32_t k = 0; k < n; ++k) {
            L.out_off[k] = off;
            off += L.n_out[k];
            if (L.st_node[k] == GF_INVALID) inv = 1u;
            else if (!st && L.st_node[k]) st = L.st_node[k];
        }
        L.out_off[n] = off;
        L.total_out = off;
        if (inv) st = GF_INVALID;
        if (!st && off > L.reserved) st = GF_PLAN_MISMATCH;
        if (!st && off > L.top) st = GF_POOL_EXHAUSTED;
        L.status = st;
        if (st) set_status(&e.ctl[CTL_STATUS], st);
    }
    __syncthreads();
    if (L.status) return;
    // pops: output j of the step (node order, bin-major, k order) takes stack[top - 1 - j]
    for (uint32_t j = tid; j < L.total_out; j += kTeamThreads) {
        uint32_t k = 0;
        while (k + 1u < n && L.out_off[k + 1u] <= j) ++k;
        const uint32_t r = j - L.out_off[k], kind = L.kind[k];
        const unsigned long long ent = L.fs[L.top - 1u - j - L.fs_base];
        gf_chunk_desc d;
        d.index = (uint32_t)ent;
        d.generation = (uint32_t)(ent >> 32);
        d.bin = 0u;
        if (kind == KIND_SOURCE) {
            d.fill = L.particles[k];
        } else if (kind == KIND_TRANSFORM) {
            const uint32_t rem = L.particles[k] - r * C;
            d.fill = rem < C ? rem : C;
        } else {
            const uint32_t B = L.nd[k].bins;
            uint32_t b = 0;
            for (uint32_t x = 0; x < B; ++x) {
                const uint32_t t = L.tot[k * kTeamBins + x], bb = L.binbase[k * kTeamBins + x];
                if (t && bb <= r) b = x;
            }
            const uint32_t rem = L.tot[k * kTeamBins + b] - (r - L.binbase[k * kTeamBins + b]) * C;
            d.fill = rem < C ? rem : C;
            d.bin = b;
        }
        L.out[j] = d;
    }
    // flat output start of every (segment, bin)
    for (uint32_t k = 0; k < n; ++k) {
        if (!L.active[k] || L.kind[k] != KIND_PARTITION) continue;
        const uint32_t B = L.nd[k].bins, B1 = B + 1u, ns = L.seg_off[k + 1] - L.seg_off[k];
        uint32_t* m = L.pre + L.pre_off[k];
        for (uint32_t x = tid; x < ns * B1; x += kTeamThreads) {
            const uint32_t b = x % B1;
            if (b < B) m[x] += L.binbase[k * kTeamBins + b] * C;
        }
    }
    __syncthreads();
    if (tid == 0) {
        L.top -= L.total_out;
        const unsigned long long live = e.S - L.top;
        if (live > L.peak_live) L.peak_live = live;
    }
    __syncthreads();
}

// ---- trace (member 0) and retire (wave 4 of every member) --------------------------------------
__device__ inline void team_trace(const TeamArgs& a, TeamLds& L)
{
    const Engine& e = a.e;
    const uint32_t lane = lane_id(), n = e.n_nodes;
    for (uint32_t k = 0; k < n; ++k) {
        if (!L.active[k]) continue;
        const uint32_t kind = L.kind[k];
        const uint32_t nin = kind == KIND_SOURCE ? 0u : L.n_in[k];
        const uint32_t nb = kind == KIND_PARTITION ? L.nd[k].bins : kind == KIND_TRANSFORM ? 1u : 0u;
        const uint32_t R = L.n_out[k];
        const unsigned long long len = 5ull + 3ull * nin + nb + 4ull * R, p0 = L.trace_len;
        auto put = [&](unsigned long long p, uint32_t v) {
            if (p0 + p < e.trace_cap) e.trace[p0 + p] = v;
        };
        if (lane == 0) {
            put(0, 0xB47C0000u | k);
            put(1, (uint32_t)L.step);
            put(2, nin);
            put(3 + 3ull * nin, nb);
            put(4 + 3ull * nin + nb, R);
        }
        for (uint32_t i = lane; i < nin; i += 64u) {
            const gf_chunk_desc d = L.in[L.in_off[k] + i];
            put(3 + 3ull * i, d.index);
            put(4 + 3ull * i, d.generation);
            put(5 + 3ull * i, d.fill);
        }
        for (uint32_t b = lane; b < nb; b += 64u)
            put(4 + 3ull * nin + b, kind == KIND_PARTITION ? L.tot[k * kTeamBins + b] : L.particles[k]);
        const unsigned long long o0 = 5ull + 3ull * nin + nb;
        for (uint32_t i = lane; i < R; i += 64u) {
            const gf_chunk_desc d = L.out[L.out_off[k] + i];
            put(o0 + 4ull * i, d.index);
            put(o0 + 4ull * i + 1, d.generation);
            put(o0 + 4ull * i + 2, d.fill);
            put(o0 + 4ull * i + 3, d.bin);
        }
        wave_sync();
        if (lane == 0) L.trace_len = p0 + len;
        wave_sync();
    }
}

__device__ inline void team_retire(const TeamArgs& a, TeamLds& L)
{
    const Engine& e = a.e;
    const uint32_t n = e.n_nodes, rcm = a.rc - 1u;
    if (lane_id() == 0) {
        // outputs to the destination rings: node order, output order
        for (uint32_t k = 0; k < n; ++k) {
            if (!L.active[k] || L.kind[k] == KIND_SINK) continue;
            const bool part = L.kind[k] == KIND_PARTITION;
            for (uint32_t r = 0; r < L.n_out[k]; ++r) {
                const gf_chunk_desc d = L.out[L.out_off[k] + r];
                const uint32_t dst = (uint32_t)L.route[L.route_off[k] + (part ? d.bin : 0u)];
                const unsigned long long pos = L.tail[dst]++;
                L.ring[dst * a.rc + (uint32_t)(pos & rcm)] = d;
            }
        }
        // consumed inputs back to the free stack (generation + 1); partition inputs' count rows
        // are cleared in the next step
        uint32_t zn = 0;
        for (uint32_t k = 0; k < n; ++k) {
            if (!L.active[k]) continue;
            const uint32_t kind = L.kind[k];
            if (kind == KIND_SINK) {
                L.sunk += L.particles[k];
                L.sink_items += L.particles[k];
            } else if (kind == KIND_SOURCE) {
                L.src_items += L.particles[k];
            } else {
                L.batches += 1u;
                if (kind == KIND_PARTITION) L.part_items += L.particles[k];
                else L.xf_items += L.particles[k];
            }
            if (kind == KIND_SOURCE) continue;
            for (uint32_t i = 0; i < L.n_in[k]; ++i) {
                const gf_chunk_desc d = L.in[L.in_off[k] + i];
                L.fs[L.top - L.fs_base] = (unsigned long long)d.index | ((unsigned long long)(d.generation + 1u) << 32);
                L.top += 1u;
                if (kind == KIND_PARTITION) L.zero[zn++] = d.index | (d.generation & 1u) << 31;
            }
        }
        for (uint32_t k = 0; k < n; ++k) L.active[k] = 0u;
        L.zero_n = zn;
        L.step += 1u;
    }
    wave_sync();
}

// A failed item: the run's status word for the host, this member's flag for the team's stop decision.
__device__ inline void seg_fail(const TeamArgs& a, const TeamLds& L, uint32_t code)
{
    set_status(&a.e.ctl[CTL_STATUS], code);
    const_cast<TeamLds&>(L).seg_err = 1u;
}

// ---- segment work (waves 0-3) ---------------------------------------------------------------------
// Next-node count key of a record stored at (chunk d, offset off), or ~0u when the destination is
// not a partition.
template <class P>
__device__ inline uint32_t next_key(const TeamArgs& a, const TeamLds& L, uint32_t dst, const gf_chunk_desc& d, uint32_t off,
                                    const P& rec)
{
    if (L.kind[dst] != KIND_PARTITION) return ~0u;
    const gf_node_desc& nd = L.nd[dst];
    uint32_t nb = classify(nd, rec);
    if (nb >= nd.bins) nb = nd.bins;
    return (((uint32_t)d.index * 2u + (d.generation & 1u)) * a.tpc + (off >> 8)) * a.bb + nb;
}

template <class P>
__device__ void team_partition_seg(const TeamArgs& a, const TeamLds& L, uint32_t k, uint32_t i, uint32_t gg, uint32_t gs)
{
    const Engine& e = a.e;
    const uint32_t C = e.C, lane = lane_id();
    const gf_node_desc& nd = L.nd[k];
    const uint32_t B = nd.bins, ro = L.route_off[k];
    const gf_chunk_desc d = L.in[L.in_off[k] + i];
    const uint32_t off0 = gg * kSeg, cnt = d.fill - off0 < kSeg ? d.fill - off0 : kSeg;
    const size_t slot0 = (size_t)d.index * C + off0;
    const uint32_t* fst = L.pre + L.pre_off[k] + gs * (B + 1u);
    const gf_chunk_desc* ob = L.out + L.out_off[k];
    P rec[kSeg / 64];
#pragma unroll
    for (uint32_t r = 0; r < kSeg / 64; ++r) {
        const uint32_t it = r * 64u + lane;
        if (it < cnt) rec[r] = P::load_team(e.cols, slot0 + it);
    }
    uint32_t run = 0;
#pragma unroll
    for (uint32_t r = 0; r < kSeg / 64; ++r) {
        const uint32_t it = r * 64u + lane;
        bool v = it < cnt;
        uint32_t bin = ~0u;
        if (v) {
            bin = classify(nd, rec[r]);
            if (bin >= B) {
                seg_fail(a, L, GF_INVALID);
                v = false;
                bin = ~0u;
            }
        }
        const uint32_t pos = seg_rank(v, bin, run);
        uint32_t key = ~0u;
        if (v) {
            if ((fst[bin] + pos) / C >= L.n_out[k]) {
                seg_fail(a, L, GF_PLAN_MISMATCH);
                v = false;
            }
        }
        if (v) {
            const uint32_t flat = fst[bin] + pos;
            const uint32_t q = flat / C, off = flat - q * C;
            const gf_chunk_desc od = ob[q];
            P p = rec[r];
            const uint32_t u = update(nd, bin, p);
            if (u) seg_fail(a, L, u);
            const int32_t fz = L.fuse[ro + bin];
            if (fz >= 0) transform(L.nd[fz], e.rp, 2u * (uint32_t)fz, p);
            p.store(e.cols, (size_t)od.index * C + off);
            key = next_key(a, L, (uint32_t)L.route[ro + bin], od, off, p);
        }
        agg_add(key != ~0u, key, a.cnt);
    }
}

template <class P>
__device__ void team_transform_seg(const TeamArgs& a, const TeamLds& L, uint32_t k, uint32_t i, uint32_t gg)
{
    const Engine& e = a.e;
    const uint32_t C = e.C, lane = lane_id();
    const gf_node_desc& nd = L.nd[k];
    const gf_chunk_desc d = L.in[L.in_off[k] + i];
    const uint32_t off0 = gg * kSeg, cnt = d.fill - off0 < kSeg ? d.fill - off0 : kSeg;
    const size_t slot0 = (size_t)d.index * C + off0;
    const uint32_t pos0 = L.in_items[L.in_off[k] + i] + off0;
    const gf_chunk_desc* ob = L.out + L.out_off[k];
    const uint32_t dst = (uint32_t)L.route[L.route_off[k]];
    P rec[kSeg / 64];
#pragma unroll
    for (uint32_t r = 0; r < kSeg / 64; ++r) {
        const uint32_t it = r * 64u + lane;
        if (it < cnt) rec[r] = P::load_team(e.cols, slot0 + it);
    }
#pragma unroll
    for (uint32_t r = 0; r < kSeg / 64; ++r) {
        const uint32_t it = r * 64u + lane;
        uint32_t key = ~0u;
        if (it < cnt) {
            P p = rec[r];
            transform(nd, e.rp, L.purpose[k], p);
            const uint32_t g = pos0 + it, q = g / C, off = g - q * C;
            if (q >= L.n_out[k]) {
                seg_fail(a, L, GF_PLAN_MISMATCH);
                continue;
            }
            const gf_chunk_desc od = ob[q];
            p.store(e.cols, (size_t)od.index * C + off);
            key = next_key(a, L, dst, od, off, p);
        }
        agg_add(key != ~0u, key, a.cnt);
    }
}

template <class P>
__device__ void team_source_seg(const TeamArgs& a, const TeamLds& L, uint32_t k, uint32_t gg)
{
    const Engine& e = a.e;
    const uint32_t C = e.C, lane = lane_id();
    const gf_node_desc& nd = L.nd[k];
    const gf_chunk_desc od = L.out[L.out_off[k]];
    const uint32_t fill = L.particles[k], off0 = gg * kSeg, cnt = fill - off0 < kSeg ? fill - off0 : kSeg;
    const uint32_t dst = (uint32_t)L.route[L.route_off[k]];
    const unsigned long long id0 = L.first_id[k] + off0;
#pragma unroll
    for (uint32_t r = 0; r < kSeg / 64; ++r) {
        const uint32_t it = r * 64u + lane;
        uint32_t key = ~0u;
        if (it < cnt) {
            P p;
            source(nd, e.rp, L.purpose[k], id0 + it, p);
            p.store(e.cols, (size_t)od.index * C + off0 + it);
            key = next_key(a, L, dst, od, off0 + it, p);
        }
        agg_add(key != ~0u, key, a.cnt);
    }
}

template <class P>
__device__ void team_sink_seg(const TeamArgs& a, const TeamLds& L, uint32_t k, uint32_t i, uint32_t gg)
{
    const Engine& e = a.e;
    const uint32_t C = e.C, lane = lane_id();
    const gf_node_desc& nd = L.nd[k];
    const gf_chunk_desc d = L.in[L.in_off[k] + i];
    const uint32_t off0 = gg * kSeg, cnt = d.fill - off0 < kSeg ? d.fill - off0 : kSeg;
    const size_t slot0 = (size_t)d.index * C + off0;
    P rec[kSeg / 64];
#pragma unroll
    for (uint32_t r = 0; r < kSeg / 64; ++r) {
        const uint32_t it = r * 64u + lane;
        if (it < cnt) rec[r] = P::load_team(e.cols, slot0 + it);
    }
    if constexpr (std::is_same_v<P, Particle>) {
#pragma unroll
        for (uint32_t r = 0; r < kSeg / 64; ++r) {
            if (r * 64u + lane >= cnt) continue;
            const Particle& p = rec[r];
            if (nd.op == GF_SNK_HIST) {
                const uint32_t v = nd.u[0] == 0u ? __float_as_uint(p.w) : p.gen;
                const uint32_t rel = v - nd.u[1];
                if (v < nd.u[1] || rel >= nd.u[2]) seg_fail(a, L, GF_INVALID);
                else atomicAdd((unsigned long long*)nd.ptr[0] + rel, 1ull);
            } else if (nd.op == GF_SNK_RECORD_WEIGHT) {
                if ((uint64_t)p.id < nd.size[0]) ((uint32_t*)nd.ptr[0])[p.id] = __float_as_uint(p.w);
                else seg_fail(a, L, GF_INVALID);
            }
        }
    } else if constexpr (std::is_same_v<P, Keyed>) {
        if (nd.op != GF_SNK_CHUNK_DIGEST) return;
        uint64_t acc = 0, msum = 0;
#pragma unroll
        for (uint32_t r = 0; r < kSeg / 64; ++r) {
            const uint32_t it = r * 64u + lane;
            if (it >= cnt) continue;
            const uint64_t h = keyed_hash(rec[r]);
            acc += mix64(h ^ ((uint64_t)(off0 + it + 1u) * 0xD6E8FEB86659FD93ull));
            msum += h;
        }
        for (int o = 32; o > 0; o >>= 1) {
            acc += __shfl_down(acc, o);
            msum += __shfl_down(msum, o);
        }
        if (lane == 0) {
            const uint64_t entry = L.digest_base[k] + i;
            if (nd.ptr[0] && entry < nd.size[0]) {
                uint32_t* log = (uint32_t*)nd.ptr[0] + 4 * entry;
                if (gg == 0) {
                    log[0] = d.bin;
                    log[1] = d.fill;
                }
                atomicAdd((unsigned long long*)(log + 2), (unsigned long long)acc);
            }
            if (nd.ptr[1] && d.bin < nd.size[1]) {
                unsigned long long* ms = (unsigned long long*)nd.ptr[1] + 2 * d.bin;
                atomicAdd(ms, (unsigned long long)msum);
                atomicAdd(ms + 1, (unsigned long long)cnt);
            }
        }
    }
}

template <class P>
__device__ inline void team_segments(const TeamArgs& a, const TeamLds& L, uint32_t me, uint32_t T, uint32_t w)
{
    const uint32_t G = L.seg_off[a.e.n_nodes];
    const uint32_t lo = (uint32_t)((uint64_t)me * G / T), hi = (uint32_t)((uint64_t)(me + 1u) * G / T);
    for (uint32_t s = lo + w; s < hi; s += 4u) {
        const uint32_t sm = L.seg_map[s];
        const uint32_t k = sm >> 28, i = (sm >> 18) & 1023u, gg = sm & 0x3FFFFu;
        const uint32_t kind = L.kind[k];
        if (kind == KIND_PARTITION) team_partition_seg<P>(a, L, k, i, gg, s - L.seg_off[k]);
        else if (kind == KIND_TRANSFORM) team_transform_seg<P>(a, L, k, i, gg);
        else if (kind == KIND_SOURCE) team_source_seg<P>(a, L, k, gg);
        else team_sink_seg<P>(a, L, k, i, gg);
    }
}

// ---- the persistent team kernel ---------------------------------------------------------------
template <class P>
__global__ __launch_bounds__(kTeamThreads) void k_team(TeamArgs a)
{
    extern __shared__ char smem[];
    TeamLds& L = *reinterpret_cast<TeamLds*>(smem);
    __shared__ uint32_t sh[4];
    team_join(a, sh);
    const uint32_t me = sh[0], T = sh[1];
    if (me == ~0u) return;
    const Engine& e = a.e;
    const uint32_t tid = threadIdx.x, w = tid >> 6;
    const bool prof = e.prof != nullptr && me == 0u && tid == 0u;
    unsigned long long tc = prof ? wall_clock64() : 0ull;
    auto mark = [&](uint32_t ph) {
        if (prof) {
            const unsigned long long now = wall_clock64();
            e.prof[ph] += now - tc;
            tc = now;
        }
    };
    team_load(a, L);
    if (L.exit_code == EXIT_SKIP) {
        if (me == 0u && tid == 0u) e.ctl[CTL_HANDOFF] = 1u;
        return;
    }
    if (w == 4u) team_select(a, L, a.max_steps == 0u);
    __syncthreads();
    uint32_t target = 0;
    for (uint32_t it = 0; L.exit_code == EXIT_NONE; ++it) {
        if (it) {
            if (tid == 0u && !team_wait(a, target)) {
                set_status(&e.ctl[CTL_STATUS], GF_DEVICE_ERROR);
                L.status = GF_DEVICE_ERROR;
            }
            __syncthreads();
            mark(PROF_T_WAIT);
            if (L.status) break;
        }
        team_counts(a, L, me, T, it);
        if (L.status) break;
        mark(PROF_T_COUNTS);
        team_plan(a, L);
        if (L.status) break;
        mark(PROF_T_PLAN);
        if (w < 4u) {
            team_segments<P>(a, L, me, T, w);
        } else {
            if (me == 0u && e.trace_on) team_trace(a, L);
            team_retire(a, L);
        }
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        mark(PROF_T_SEG);
        target += T;
        if (tid == 0u) {
            if (L.seg_err) {
                atomicOr(&a.tw[TW_ERR + (it & 1u)], 1u);
                __builtin_amdgcn_s_waitcnt(0);
            }
            __hip_atomic_fetch_add(&a.tw[TW_BAR], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        if (w == 4u) team_select(a, L, it + 1u >= a.max_steps);
        __syncthreads();
        mark(PROF_T_SELECT);
        if (prof) e.prof[PROF_T_STEPS] += 1u;
    }
    if (me == 0u && !L.status) team_flush(a, L);
}

}  // namespace gfd
