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

constexpr uint32_t kTeamNodes = 32;
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
enum {
    PROF_T_WAIT = 9, PROF_T_COUNTS = 10, PROF_T_PLAN = 11, PROF_T_SEG = 12, PROF_T_SELECT = 13, PROF_T_STEPS = 14,
    PROF_T_SEGW = 15, PROF_T_RETIRE = 16, PROF_T_SELECTW = 17,
    // shader-clock cycles of sub-phases (member 0, lane 0 of the wave doing the work)
    PROF_C_SEL_FILLS = 18, PROF_C_SEL_CREDIT = 19, PROF_C_SEL_CHECK = 20, PROF_C_SEL_LISTS = 21,
    PROF_C_RET_OUT = 22, PROF_C_RET_IN = 23, PROF_C_PLAN_NODES = 24, PROF_C_PLAN_POPS = 25,
    PROF_C_CNT_CLEAR = 26, PROF_C_CNT_SCAN = 27, PROF_C_SEG_RANK = 28, PROF_C_SEG_STORE = 29, PROF_C_SEG_AGG = 30,
    PROF_C_STEP = 31
};

// Sub-phase cycle counter: accumulates the shader-clock delta since the previous mark.
struct Ticker {
    unsigned long long* p;
    unsigned long long t;
    bool on;
    __device__ Ticker(unsigned long long* prof, bool enabled) : p(prof), t(0), on(enabled && prof != nullptr)
    {
        if (on) t = __builtin_readcyclecounter();
    }
    __device__ void operator()(uint32_t slot)
    {
        if (on) {
            const unsigned long long n = __builtin_readcyclecounter();
            p[slot] += n - t;
            t = n;
        }
    }
};

struct TeamArgs {
    Engine e;
    uint32_t* cnt;     // [S][2][tpc][bb] segment bin counts
    uint32_t* zero_g;  // [0] = n, [1..n] = chunk | (gen & 1) << 31 rows still to clear
    uint32_t* tw;      // team words, cleared before every launch
    uint32_t K, tpc, bb, max_steps, rc, n_routes;
    uint32_t c_shift;  // log2(C) when C is a power of two, else 32
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
    uint32_t seg_map[kTeamSegs];  // seg_pack(node, input, segment of the chunk)
    uint32_t seg_row[kTeamSegs];  // partition segments: first word of the segment's count row
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
    uint32_t prof_member;  // this member records sub-phase cycle counts
    uint32_t shv[8];
};

__device__ inline uint32_t xcc_id() { return __builtin_amdgcn_s_getreg(20 | (3 << 11)) & 0xFu; }  // HW_REG_XCC_ID[3:0]

__device__ inline void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Segment map entry: node (5 bits) | input chunk of the node's batch (10 bits) | segment (17 bits).
__device__ inline uint32_t seg_pack(uint32_t k, uint32_t i, uint32_t gg) { return k << 27 | i << 17 | gg; }
__device__ inline uint32_t seg_node(uint32_t sm) { return sm >> 27; }
__device__ inline uint32_t seg_input(uint32_t sm) { return (sm >> 17) & 1023u; }
__device__ inline uint32_t seg_part(uint32_t sm) { return sm & 0x1FFFFu; }

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

// ---- wave-level helpers ------------------------------------------------------------------------
__device__ inline uint32_t rl(uint32_t v, uint32_t k) { return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)k); }
__device__ inline unsigned long long rl64(unsigned long long v, uint32_t k)
{
    return ((unsigned long long)rl((uint32_t)(v >> 32), k) << 32) | rl((uint32_t)v, k);
}
// Value of `v` in lane `src` (all lanes must be active).
__device__ inline uint32_t from_lane(uint32_t v, uint32_t src) { return (uint32_t)__shfl((int)v, (int)src); }
__device__ inline unsigned long long from_lane64(unsigned long long v, uint32_t src)
{
    return ((unsigned long long)from_lane((uint32_t)(v >> 32), src) << 32) | from_lane((uint32_t)v, src);
}
__device__ inline uint32_t wave_sum(uint32_t v) { return rl(wave_inscan(v), 63); }

// Node whose [off, off + cnt) range holds item x: the last node j with cnt_j > 0 and off_j <= x
// (off and cnt are lane-per-node values).
__device__ inline uint32_t node_of(uint32_t x, uint32_t off, uint32_t cnt, uint32_t n)
{
    uint32_t kx = 0;
    for (uint32_t j = 0; j < n; ++j)
        if (rl(cnt, j) && x >= rl(off, j)) kx = j;
    return kx;
}

// ---- select (wave 4): the next step's batches, or the reason to stop -------------------------
// Lane k holds node k. Credits are handed out in node order by a scalar loop over the lanes'
// values; every array write is lane-parallel.
__device__ inline void team_select(const TeamArgs& a, TeamLds& L, bool at_limit)
{
    const Engine& e = a.e;
    const uint32_t lane = lane_id(), n = e.n_nodes, C = e.C, rcm = a.rc - 1u;
    Ticker tk(e.prof, L.prof_member && lane == 0u);
    const bool own = lane < n;
    const uint32_t k = own ? lane : 0u;
    const uint32_t kind = own ? L.kind[k] : 0xFFu;
    const bool xf = kind == KIND_PARTITION || kind == KIND_TRANSFORM;
    const uint32_t B = kind == KIND_PARTITION ? L.nd[k].bins : 1u;
    const unsigned long long head = own ? L.head[k] : 0ull, tail = own ? L.tail[k] : 0ull;
    const uint32_t avail = (uint32_t)(tail - head);
    const gf_chunk_desc* rg = L.ring + k * a.rc;
    uint32_t want = kind == KIND_SINK ? avail : xf ? (avail < e.Q ? avail : e.Q) : 0u;
    uint32_t parts = 0;
    for (uint32_t i = 0; i < want; ++i) parts += rg[(uint32_t)((head + i) & rcm)].fill;
    uint32_t bound = xf ? (uint32_t)bound_of(parts, C, B) : 0u;
    tk(PROF_C_SEL_FILLS);
    // credits in node order
    const uint32_t top = L.top;
    uint32_t credits = top;
    const unsigned long long gen0 = L.generated, sunk = L.sunk;
    unsigned long long gen = gen0;
    uint32_t src_fill = 0, src_node = 0xFFFFFFFFu;
    for (uint32_t j = 0; j < n; ++j) {
        const uint32_t kj = rl(kind, j);
        if (kj == KIND_SOURCE) {
            const unsigned long long P = e.P, live = gen - sunk, remaining = e.N - gen;
            unsigned long long min_fill = C < P ? C : P;
            if (remaining < min_fill) min_fill = remaining;
            if (gen < e.N && live < P && P - live >= min_fill && credits >= 1u) {
                unsigned long long fill = remaining;
                if (fill > P - live) fill = P - live;
                if (fill > C) fill = C;
                src_fill = (uint32_t)fill;
                src_node = j;
                gen += fill;
                credits -= 1u;
            }
        } else if (kj == KIND_PARTITION || kj == KIND_TRANSFORM) {
            if (!rl(want, j)) continue;
            if (rl(bound, j) > credits && lane == j) {
                // largest prefix of the candidates whose bound fits the remaining credits
                uint32_t m = want, p = parts, b = 0;
                while (m) {
                    b = (uint32_t)bound_of(p, C, B);
                    if (b <= credits) break;
                    --m;
                    p -= rg[(uint32_t)((head + m) & rcm)].fill;
                }
                want = m;
                parts = p;
                bound = m ? b : 0u;
            }
            credits -= rl(bound, j);
        }
    }
    tk(PROF_C_SEL_CREDIT);
    uint32_t nin = 0, act = 0, nseg = 0;
    if (kind == KIND_SOURCE) {
        act = lane == src_node ? 1u : 0u;
        parts = act ? src_fill : 0u;
        nseg = (parts + kSeg - 1u) / kSeg;
    } else if (own) {
        nin = want;
        act = nin ? 1u : 0u;
        for (uint32_t i = 0; i < nin; ++i) nseg += (rg[(uint32_t)((head + i) & rcm)].fill + kSeg - 1u) / kSeg;
    }
    const uint32_t prew = (kind == KIND_PARTITION && act) ? nseg * (B + 1u) : 0u;
    const uint32_t i_inc = wave_inscan(nin), s_inc = wave_inscan(nseg), p_inc = wave_inscan(prew);
    const uint32_t in_tot = rl(i_inc, 63), seg_tot = rl(s_inc, 63), pre_tot = rl(p_inc, 63);
    const uint32_t reserved = top - credits;
    const uint64_t actm = __ballot(act != 0u), pend = __ballot(own && tail != head);
    uint32_t code = EXIT_NONE;
    if (!actm) {
        code = (!pend && gen0 >= e.N) ? EXIT_DONE : EXIT_STALL;
    } else if (at_limit) {
        code = EXIT_LIMIT;
    } else {
        const bool fits = in_tot <= kTeamIn && seg_tot <= kTeamSegs && pre_tot <= kTeamPrefix && reserved <= kTeamOut &&
                          reserved <= kTeamMargin && in_tot <= kTeamMargin;
        const bool room = !own || (avail - nin) + reserved <= a.rc;
        if (!fits || __ballot(!room)) code = EXIT_HANDOFF;
    }
    tk(PROF_C_SEL_CHECK);
    if (code) {
        if (lane == 0) L.exit_code = code;
        wave_sync();
        return;
    }
    const uint32_t inoff = i_inc - nin, segoff = s_inc - nseg;
    if (own) {
        L.active[k] = act;
        L.n_in[k] = nin;
        L.particles[k] = parts;
        L.in_off[k] = inoff;
        L.seg_off[k] = segoff;
        L.pre_off[k] = p_inc - prew;
        L.head[k] = head + nin;
        if (kind == KIND_SINK && act) {
            const unsigned long long dg = L.digest[k];
            L.digest_base[k] = dg;
            L.digest[k] = dg + nin;
        }
        if (kind == KIND_SOURCE && act) L.first_id[k] = gen0;
    }
    if (lane == n) {
        L.in_off[n] = in_tot;
        L.seg_off[n] = seg_tot;
    }
    if (lane == 0) {
        L.generated = gen;
        L.reserved = reserved;
        const uint32_t nb = top > kTeamStack / 2u ? top - kTeamStack / 2u : 0u;
        const uint32_t fb = L.fs_base;
        L.recenter = (nb != fb && (top < fb + kTeamMargin || top + kTeamMargin > fb + kTeamStack)) ? 1u : 0u;
        L.exit_code = EXIT_NONE;
    }
    // input list, item prefixes, segment map and count rows, flattened over the step's inputs
    // (node order; the source has no inputs)
    const uint32_t nf = kind == KIND_SOURCE ? 0u : parts, nsf = kind == KIND_SOURCE ? 0u : nseg;
    const uint32_t F = wave_inscan(nf) - nf, S = wave_inscan(nsf) - nsf;
    uint32_t carry_f = 0, carry_s = 0;
    for (uint32_t x0 = 0; x0 < in_tot; x0 += 64u) {
        const uint32_t x = x0 + lane;
        const bool v = x < in_tot;
        const uint32_t kx = node_of(x, inoff, nin, n);
        const uint32_t i = x - from_lane(inoff, kx);
        const unsigned long long hx = from_lane64(head, kx);
        const uint32_t Fx = from_lane(F, kx), Sx = from_lane(S, kx), sox = from_lane(segoff, kx);
        const uint32_t kindx = from_lane(kind, kx);
        gf_chunk_desc d{0, 0, 0, 0};
        if (v) {
            d = L.ring[kx * a.rc + (uint32_t)((hx + i) & rcm)];
            L.in[x] = d;
        }
        const uint32_t fill = v ? d.fill : 0u, ns = (fill + kSeg - 1u) / kSeg;
        const uint32_t fi = wave_inscan(fill), si = wave_inscan(ns);
        if (v) {
            L.in_items[x] = carry_f + fi - fill - Fx;
            const uint32_t sp = sox + carry_s + si - ns - Sx;
            const uint32_t row0 = ((d.index * 2u + (d.generation & 1u)) * a.tpc) * a.bb;
            for (uint32_t g = 0; g < ns; ++g) {
                L.seg_map[sp + g] = seg_pack(kx, i, g);
                if (kindx == KIND_PARTITION) L.seg_row[sp + g] = row0 + g * a.bb;
            }
        }
        carry_f += rl(fi, 63);
        carry_s += rl(si, 63);
    }
    if (src_node != 0xFFFFFFFFu) {
        const uint32_t so = rl(segoff, src_node), ns = rl(nseg, src_node);
        for (uint32_t g = lane; g < ns; g += 64u) L.seg_map[so + g] = seg_pack(src_node, 0u, g);
    }
    wave_sync();
    tk(PROF_C_SEL_LISTS);
}

// ---- after the barrier: status, clears, segment counts and their scan, free-stack window --------
__device__ inline void team_counts(const TeamArgs& a, TeamLds& L, uint32_t me, uint32_t T, uint32_t it)
{
    const Engine& e = a.e;
    const uint32_t tid = threadIdx.x, n = e.n_nodes, w = tid >> 6, lane = lane_id();
    const uint32_t row = a.tpc * a.bb;
    Ticker tk(e.prof, L.prof_member && tid == 0u);
    // errors members published before the previous barrier (parity word: a member already one
    // step ahead writes the other word)
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
    tk(PROF_C_CNT_CLEAR);
    // wave b loads bin b's count of every segment of a partition batch and scans it in registers:
    // L.pre[g][b] = items of bin b in the batch's segments before g, L.tot[b] = all of them
    for (uint32_t k = 0; k < n; ++k) {
        if (!L.active[k] || L.kind[k] != KIND_PARTITION) continue;
        const uint32_t B1 = L.nd[k].bins + 1u, so = L.seg_off[k], ns = L.seg_off[k + 1] - so;
        uint32_t* m = L.pre + L.pre_off[k];
        for (uint32_t b = w; b < B1; b += kTeamWaves) {
            uint32_t carry = 0;
            for (uint32_t g0 = 0; g0 < ns; g0 += 64u) {
                const uint32_t g = g0 + lane;
                const uint32_t x = g < ns ? ld_u32(a.cnt, (size_t)L.seg_row[so + g] + b) : 0u;
                const uint32_t inc = wave_inscan(x);
                if (g < ns) m[g * B1 + b] = carry + inc - x;
                carry += rl(inc, 63);
            }
            if (lane == 0) L.tot[k * kTeamBins + b] = carry;
        }
    }
    tk(PROF_C_CNT_SCAN);
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

// ---- plan (wave 0): exact output counts, bin bases, checks, pops --------------------------------
// Lane k holds node k; a partition's bins are handled one per lane.
__device__ inline void team_plan(const TeamArgs& a, TeamLds& L)
{
    const Engine& e = a.e;
    const uint32_t tid = threadIdx.x, n = e.n_nodes, C = e.C, lane = lane_id();
    if (tid < 64u) {
        Ticker tk(e.prof, L.prof_member && lane == 0u);
        const bool own = lane < n;
        const uint32_t k = own ? lane : 0u;
        const uint32_t act = own ? L.active[k] : 0u, kind = own ? L.kind[k] : 0xFFu;
        const uint32_t parts = own ? L.particles[k] : 0u, bound = own ? L.bound[k] : 0u;
        uint32_t R = 0, st = 0;
        if (act) {
            if (kind == KIND_TRANSFORM) {
                R = (parts + C - 1u) / C;
                if (parts == 0u || R > bound) st = GF_PLAN_MISMATCH;
            } else if (kind == KIND_SOURCE) {
                R = 1u;
            }
        }
        uint64_t pm = __ballot(act && kind == KIND_PARTITION);
        uint32_t inv = 0, pk = ~0u, pB = 0, bbl = 0, totl = 0;  // bins of the last partition node, lane b
        while (pm) {
            const uint32_t kp = (uint32_t)(__ffsll((unsigned long long)pm) - 1);
            pm &= pm - 1u;
            const uint32_t B = L.nd[kp].bins;
            const uint32_t t = lane <= B ? L.tot[kp * kTeamBins + lane] : 0u;
            const uint32_t tv = lane < B ? t : 0u, r = (tv + C - 1u) / C;
            const uint32_t ri = wave_inscan(r), ti = wave_inscan(tv);
            if (lane < B) L.binbase[kp * kTeamBins + lane] = ri - r;
            const uint32_t Rk = rl(ri, 63), sum = rl(ti, 63);
            if (rl(t, B)) inv = 1u;
            if (lane == kp) {
                R = Rk;
                if (sum != parts || Rk > bound) st = GF_PLAN_MISMATCH;
            }
            pk = kp;
            pB = B;
            bbl = ri - r;
            totl = tv;
        }
        const uint64_t bad = __ballot(st != 0u);
        uint32_t status = inv ? (uint32_t)GF_INVALID : bad ? rl(st, (uint32_t)(__ffsll((unsigned long long)bad) - 1)) : 0u;
        const uint32_t oi = wave_inscan(R), off = oi - R, total = rl(oi, 63);
        if (!status && total > L.reserved) status = GF_PLAN_MISMATCH;
        if (!status && total > L.top) status = GF_POOL_EXHAUSTED;
        if (own) {
            L.n_out[k] = R;
            L.out_off[k] = off;
        }
        if (lane == 0) {
            L.out_off[n] = total;
            L.total_out = total;
            L.status = status;
            if (status) set_status(&e.ctl[CTL_STATUS], status);
        }
        tk(PROF_C_PLAN_NODES);
        if (!status) {
            // output j of the step (node order, bin-major, k order) takes stack[top - 1 - j]
            const uint32_t top = L.top, fbase = L.fs_base;
            for (uint32_t j0 = 0; j0 < total; j0 += 64u) {
                const uint32_t j = j0 + lane;
                const uint32_t kx = node_of(j, off, R, n);
                const uint32_t r = j - from_lane(off, kx), kk = from_lane(kind, kx), pp = from_lane(parts, kx);
                if (j < total) {
                    const unsigned long long ent = L.fs[top - 1u - j - fbase];
                    gf_chunk_desc d;
                    d.index = (uint32_t)ent;
                    d.generation = (uint32_t)(ent >> 32);
                    d.bin = 0u;
                    if (kk == KIND_SOURCE) {
                        d.fill = pp;
                    } else if (kk == KIND_TRANSFORM) {
                        const uint32_t rem = pp - r * C;
                        d.fill = rem < C ? rem : C;
                    } else {
                        // the non-empty bin whose output range holds r
                        uint32_t b = 0, tb = 0, bb = 0;
                        if (kx == pk) {
                            for (uint32_t x = 0; x < pB; ++x) {
                                const uint32_t tx = rl(totl, x), bx = rl(bbl, x);
                                if (tx && bx <= r) {
                                    b = x;
                                    tb = tx;
                                    bb = bx;
                                }
                            }
                        } else {
                            const uint32_t B = L.nd[kx].bins;
                            for (uint32_t x = 0; x < B; ++x) {
                                const uint32_t tx = L.tot[kx * kTeamBins + x], bx = L.binbase[kx * kTeamBins + x];
                                if (tx && bx <= r) {
                                    b = x;
                                    tb = tx;
                                    bb = bx;
                                }
                            }
                        }
                        const uint32_t rem = tb - (r - bb) * C;
                        d.fill = rem < C ? rem : C;
                        d.bin = b;
                    }
                    L.out[j] = d;
                }
            }
            if (lane == 0) {
                L.top = top - total;
                const unsigned long long live = e.S - (top - total);
                if (live > L.peak_live) L.peak_live = live;
            }
        }
        tk(PROF_C_PLAN_POPS);
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

// Retire (wave 4 of every member): outputs to the destination rings in node and output order,
// consumed inputs back to the free stack (generation + 1), partition inputs' count rows queued for
// clearing in the next step, counters.
__device__ inline void team_retire(const TeamArgs& a, TeamLds& L)
{
    const Engine& e = a.e;
    const uint32_t lane = lane_id(), n = e.n_nodes, rcm = a.rc - 1u;
    Ticker tk(e.prof, L.prof_member && lane == 0u);
    const bool own = lane < n;
    const uint32_t k = own ? lane : 0u;
    const uint32_t kind = own ? L.kind[k] : 0xFFu, act = own ? L.active[k] : 0u;
    const uint32_t nout = own ? L.n_out[k] : 0u, ooff = own ? L.out_off[k] : 0u, ro = own ? L.route_off[k] : 0u;
    const uint32_t nin = own ? L.n_in[k] : 0u, inoff = own ? L.in_off[k] : 0u, parts = own ? L.particles[k] : 0u;
    unsigned long long tail = own ? L.tail[k] : 0ull;
    const uint32_t total_out = L.total_out, in_tot = L.in_off[n], top = L.top, fb = L.fs_base;
    for (uint32_t x0 = 0; x0 < total_out; x0 += 64u) {
        const uint32_t x = x0 + lane;
        const uint32_t kx = node_of(x, ooff, nout, n);
        const uint32_t kk = from_lane(kind, kx), rox = from_lane(ro, kx);
        const bool v = x < total_out && kk != KIND_SINK;
        gf_chunk_desc d{0, 0, 0, 0};
        uint32_t dst = 0xFFFFFFFFu;
        if (v) {
            d = L.out[x];
            dst = (uint32_t)L.route[rox + (kk == KIND_PARTITION ? d.bin : 0u)];
        }
        uint64_t pend = __ballot(v);
        unsigned long long pos = 0;
        while (pend) {
            const int leader = __ffsll((unsigned long long)pend) - 1;
            const uint32_t lb = rl(dst, (uint32_t)leader);
            const bool mine = v && dst == lb;
            const uint64_t m = __ballot(mine);
            const unsigned long long tl = rl64(tail, lb);
            if (mine) pos = tl + (uint32_t)__popcll(m & lanemask_lt());
            if (lane == lb) tail += (uint32_t)__popcll(m);
            pend &= ~m;
        }
        if (v) L.ring[dst * a.rc + (uint32_t)(pos & rcm)] = d;
    }
    tk(PROF_C_RET_OUT);
    uint32_t zn = 0;
    for (uint32_t x0 = 0; x0 < in_tot; x0 += 64u) {
        const uint32_t x = x0 + lane;
        const bool v = x < in_tot;
        const uint32_t kk = from_lane(kind, node_of(x, inoff, nin, n));
        gf_chunk_desc d{0, 0, 0, 0};
        if (v) {
            d = L.in[x];
            L.fs[top + x - fb] = (unsigned long long)d.index | ((unsigned long long)(d.generation + 1u) << 32);
        }
        const bool z = v && kk == KIND_PARTITION;
        const uint64_t zm = __ballot(z);
        if (z) L.zero[zn + (uint32_t)__popcll(zm & lanemask_lt())] = d.index | (d.generation & 1u) << 31;
        zn += (uint32_t)__popcll(zm);
    }
    const bool a1 = act != 0u;
    const uint32_t sink_p = wave_sum(a1 && kind == KIND_SINK ? parts : 0u);
    const uint32_t src_p = wave_sum(a1 && kind == KIND_SOURCE ? parts : 0u);
    const uint32_t part_p = wave_sum(a1 && kind == KIND_PARTITION ? parts : 0u);
    const uint32_t xf_p = wave_sum(a1 && kind == KIND_TRANSFORM ? parts : 0u);
    const uint32_t nb = (uint32_t)__popcll(__ballot(a1 && (kind == KIND_PARTITION || kind == KIND_TRANSFORM)));
    if (own) {
        L.tail[k] = tail;
        L.active[k] = 0u;
    }
    if (lane == 0) {
        L.top = top + in_tot;
        L.zero_n = zn;
        L.sunk += sink_p;
        L.sink_items += sink_p;
        L.src_items += src_p;
        L.part_items += part_p;
        L.xf_items += xf_p;
        L.batches += nb;
        L.step += 1u;
    }
    wave_sync();
    tk(PROF_C_RET_IN);
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

constexpr uint32_t kSegRounds = kSeg / 64u;

__device__ inline uint32_t div_c(uint32_t x, uint32_t C, uint32_t shift) { return shift < 32u ? x >> shift : x / C; }

// Records of one input segment, prefetched right after the barrier by the wave that will scatter it.
template <class P>
struct SegRecs {
    P rec[kSegRounds];
    uint32_t s;  // segment index, or ~0u
};

template <class P>
__device__ inline void load_segment(const TeamArgs& a, const TeamLds& L, uint32_t s, SegRecs<P>& R)
{
    const uint32_t sm = L.seg_map[s];
    const uint32_t k = seg_node(sm), i = seg_input(sm), gg = seg_part(sm);
    R.s = ~0u;
    if (L.kind[k] == KIND_SOURCE) return;
    const gf_chunk_desc d = L.in[L.in_off[k] + i];
    const uint32_t off0 = gg * kSeg, cnt = d.fill - off0 < kSeg ? d.fill - off0 : kSeg;
    const size_t slot0 = (size_t)d.index * a.e.C + off0;
    const uint32_t lane = lane_id();
#pragma unroll
    for (uint32_t r = 0; r < kSegRounds; ++r) {
        const uint32_t it = r * 64u + lane;
        if (it < cnt) R.rec[r] = P::load_team(a.e.cols, slot0 + it);
    }
    R.s = s;
}

// Per-bin constants of one partition node, bin b in lane b of the wave, loaded once per launch:
// the fused transform (node, op, f[0]), the destination after fusion and, when that destination
// is a partition, its classifier parameters (dop = 0 otherwise).
struct BinTab {
    uint32_t k;  // node described (wave-uniform), or ~0u
    int32_t fz;
    uint32_t dst, xop, dop, dbins, du0;
    float xf0, df0;
};

__device__ inline void load_bintab(const TeamLds& L, uint32_t k, BinTab& t)
{
    const uint32_t lane = lane_id(), B = L.nd[k].bins, ro = L.route_off[k];
    t.k = k;
    t.fz = -1;
    t.dst = 0u;
    t.xop = 0u;
    t.xf0 = 1.0f;
    t.dop = 0u;
    t.dbins = 1u;
    t.du0 = 0u;
    t.df0 = 0.0f;
    if (lane < B) {
        t.fz = L.fuse[ro + lane];
        t.dst = (uint32_t)L.route[ro + lane];
        if (t.fz >= 0) {
            t.xop = L.nd[t.fz].op;
            t.xf0 = L.nd[t.fz].f[0];
        }
        if (L.kind[t.dst] == KIND_PARTITION) {
            const gf_node_desc& dn = L.nd[t.dst];
            t.dop = dn.op;
            t.dbins = dn.bins;
            t.du0 = dn.u[0];
            t.df0 = dn.f[0];
        }
    }
}

// One lane's bin constants and flat output position.
struct ItemBin {
    uint32_t flat, dst, xop, dop, dbins, du0;
    int32_t fz;
    float xf0, df0;
};

// One pass over the distinct bins of a round (the bin id is wave-uniform inside the pass): the
// rank of every lane inside the segment after `run` (lane b holds bin b's running count), its
// flat output position from fb (lane b holds bin b's flat start for this segment), and the bin's
// table entries, all through readlane.
__device__ inline void seg_group(bool v, uint32_t bin, uint32_t& run, uint32_t fb, const BinTab& t, ItemBin& o)
{
    uint64_t act = __ballot(v);
    const uint32_t lane = lane_id();
    while (act) {
        const uint32_t lb = rl(bin, (uint32_t)(__ffsll((unsigned long long)act) - 1));
        const bool mine = v && bin == lb;
        const uint64_t m = __ballot(mine);
        const uint32_t base = rl(run, lb) + rl(fb, lb);
        const uint32_t dst = rl(t.dst, lb), xop = rl(t.xop, lb), dop = rl(t.dop, lb), dbins = rl(t.dbins, lb),
                       du0 = rl(t.du0, lb), fz = rl((uint32_t)t.fz, lb), xf0 = rl(__float_as_uint(t.xf0), lb),
                       df0 = rl(__float_as_uint(t.df0), lb);
        if (mine) {
            o.flat = base + (uint32_t)__popcll(m & lanemask_lt());
            o.dst = dst;
            o.xop = xop;
            o.dop = dop;
            o.dbins = dbins;
            o.du0 = du0;
            o.fz = (int32_t)fz;
            o.xf0 = __uint_as_float(xf0);
            o.df0 = __uint_as_float(df0);
        }
        if (lane == lb) run += (uint32_t)__popcll(m);
        act &= ~m;
    }
}

template <class P>
__device__ void team_partition_seg(const TeamArgs& a, const TeamLds& L, uint32_t k, uint32_t i, uint32_t gg, uint32_t gs,
                                   const SegRecs<P>& R, BinTab& t)
{
    const Engine& e = a.e;
    const uint32_t C = e.C, lane = lane_id();
    if (t.k != k) load_bintab(L, k, t);
    const gf_node_desc& nd = L.nd[k];
    const uint32_t B = nd.bins;
    const gf_chunk_desc d = L.in[L.in_off[k] + i];
    const uint32_t off0 = gg * kSeg, cnt = d.fill - off0 < kSeg ? d.fill - off0 : kSeg;
    const uint32_t* pre = L.pre + L.pre_off[k] + gs * (B + 1u);
    const uint32_t fb = lane < B ? L.binbase[k * kTeamBins + lane] * C + pre[lane] : 0u;
    const gf_chunk_desc* ob = L.out + L.out_off[k];
    const uint32_t nout = L.n_out[k];
    Ticker tk(e.prof, L.prof_member && threadIdx.x == 0u);
    bool v[kSegRounds];
    uint32_t bin[kSegRounds], key[kSegRounds];
    ItemBin ib[kSegRounds];
#pragma unroll
    for (uint32_t r = 0; r < kSegRounds; ++r) {
        v[r] = r * 64u + lane < cnt;
        bin[r] = ~0u;
        if (v[r]) {
            bin[r] = classify(nd, R.rec[r]);
            if (bin[r] >= B) {
                seg_fail(a, L, GF_INVALID);
                v[r] = false;
                bin[r] = ~0u;
            }
        }
    }
    uint32_t run = 0;
#pragma unroll
    for (uint32_t r = 0; r < kSegRounds; ++r) seg_group(v[r], bin[r], run, fb, t, ib[r]);
    tk(PROF_C_SEG_RANK);
#pragma unroll
    for (uint32_t r = 0; r < kSegRounds; ++r) {
        key[r] = ~0u;
        if (!v[r]) continue;
        const uint32_t q = div_c(ib[r].flat, C, a.c_shift), off = ib[r].flat - q * C;
        if (q >= nout) {
            seg_fail(a, L, GF_PLAN_MISMATCH);
            continue;
        }
        const gf_chunk_desc od = ob[q];
        P p = R.rec[r];
        const uint32_t u = update(nd, bin[r], p);
        if (u) seg_fail(a, L, u);
        if (!transform_p(ib[r].xop, ib[r].xf0, p) && ib[r].fz >= 0) transform(L.nd[ib[r].fz], e.rp, 2u * (uint32_t)ib[r].fz, p);
        p.store(e.cols, (size_t)od.index * C + off);
        if (ib[r].dop) {
            uint32_t nb = classify_p(ib[r].dop, ib[r].dbins, ib[r].du0, ib[r].df0, p);
            if (nb == kNeedDesc) nb = classify(L.nd[ib[r].dst], p);
            if (nb >= ib[r].dbins) nb = ib[r].dbins;
            key[r] = (((uint32_t)od.index * 2u + (od.generation & 1u)) * a.tpc + (off >> 8)) * a.bb + nb;
        }
    }
    tk(PROF_C_SEG_STORE);
#pragma unroll
    for (uint32_t r = 0; r < kSegRounds; ++r) agg_add(key[r] != ~0u, key[r], a.cnt);
    tk(PROF_C_SEG_AGG);
}

template <class P>
__device__ void team_transform_seg(const TeamArgs& a, const TeamLds& L, uint32_t k, uint32_t i, uint32_t gg,
                                   const SegRecs<P>& R)
{
    const Engine& e = a.e;
    const uint32_t C = e.C, lane = lane_id();
    const gf_node_desc& nd = L.nd[k];
    const gf_chunk_desc d = L.in[L.in_off[k] + i];
    const uint32_t off0 = gg * kSeg, cnt = d.fill - off0 < kSeg ? d.fill - off0 : kSeg;
    const uint32_t pos0 = L.in_items[L.in_off[k] + i] + off0;
    const gf_chunk_desc* ob = L.out + L.out_off[k];
    const uint32_t nout = L.n_out[k], dst = (uint32_t)L.route[L.route_off[k]];
#pragma unroll
    for (uint32_t r = 0; r < kSegRounds; ++r) {
        const uint32_t it = r * 64u + lane;
        uint32_t key = ~0u;
        if (it < cnt) {
            const uint32_t g = pos0 + it, q = div_c(g, C, a.c_shift), off = g - q * C;
            if (q < nout) {
                P p = R.rec[r];
                transform(nd, e.rp, L.purpose[k], p);
                const gf_chunk_desc od = ob[q];
                p.store(e.cols, (size_t)od.index * C + off);
                key = next_key(a, L, dst, od, off, p);
            } else {
                seg_fail(a, L, GF_PLAN_MISMATCH);
            }
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
    for (uint32_t r = 0; r < kSegRounds; ++r) {
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
__device__ void team_sink_seg(const TeamArgs& a, const TeamLds& L, uint32_t k, uint32_t i, uint32_t gg, const SegRecs<P>& R)
{
    const uint32_t lane = lane_id();
    const gf_node_desc& nd = L.nd[k];
    const gf_chunk_desc d = L.in[L.in_off[k] + i];
    const uint32_t off0 = gg * kSeg, cnt = d.fill - off0 < kSeg ? d.fill - off0 : kSeg;
    if constexpr (std::is_same_v<P, Particle>) {
#pragma unroll
        for (uint32_t r = 0; r < kSegRounds; ++r) {
            if (r * 64u + lane >= cnt) continue;
            const Particle& p = R.rec[r];
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
        for (uint32_t r = 0; r < kSegRounds; ++r) {
            const uint32_t it = r * 64u + lane;
            if (it >= cnt) continue;
            const uint64_t h = keyed_hash(R.rec[r]);
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

// Segments [lo, hi) of this member, wave w takes lo + w, lo + w + 4, ...
__device__ inline void member_range(const TeamLds& L, uint32_t n, uint32_t me, uint32_t T, uint32_t* lo, uint32_t* hi)
{
    const uint32_t G = L.seg_off[n];
    *lo = (uint32_t)((uint64_t)me * G / T);
    *hi = (uint32_t)((uint64_t)(me + 1u) * G / T);
}

template <class P>
__device__ inline void team_prefetch(const TeamArgs& a, const TeamLds& L, uint32_t me, uint32_t T, uint32_t w, SegRecs<P>& R)
{
    uint32_t lo, hi;
    member_range(L, a.e.n_nodes, me, T, &lo, &hi);
    R.s = ~0u;
    if (lo + w < hi) load_segment<P>(a, L, lo + w, R);
}

template <class P>
__device__ inline void team_segments(const TeamArgs& a, const TeamLds& L, uint32_t me, uint32_t T, uint32_t w, SegRecs<P>& R,
                                     BinTab& t)
{
    uint32_t lo, hi;
    member_range(L, a.e.n_nodes, me, T, &lo, &hi);
    for (uint32_t s = lo + w; s < hi; s += 4u) {
        const uint32_t sm = L.seg_map[s];
        const uint32_t k = seg_node(sm), i = seg_input(sm), gg = seg_part(sm);
        const uint32_t kind = L.kind[k];
        if (kind == KIND_SOURCE) {
            team_source_seg<P>(a, L, k, gg);
            continue;
        }
        if (R.s != s) load_segment<P>(a, L, s, R);
        if (kind == KIND_PARTITION) team_partition_seg<P>(a, L, k, i, gg, s - L.seg_off[k], R, t);
        else if (kind == KIND_TRANSFORM) team_transform_seg<P>(a, L, k, i, gg, R);
        else team_sink_seg<P>(a, L, k, i, gg, R);
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
    const bool wprof = e.prof != nullptr && me == 0u && lane_id() == 0u;
    unsigned long long tc = prof ? wall_clock64() : 0ull;
    auto mark = [&](uint32_t ph) {
        if (prof) {
            const unsigned long long now = wall_clock64();
            e.prof[ph] += now - tc;
            tc = now;
        }
    };
    if (tid == 0u) L.prof_member = (e.prof != nullptr && me == 0u) ? 1u : 0u;
    team_load(a, L);
    if (L.exit_code == EXIT_SKIP) {
        if (me == 0u && tid == 0u) e.ctl[CTL_HANDOFF] = 1u;
        return;
    }
    if (w == 4u) team_select(a, L, a.max_steps == 0u);
    __syncthreads();
    uint32_t target = 0;
    unsigned long long cyc = 0;
    SegRecs<P> R;
    BinTab tab;
    tab.k = ~0u;
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
        if (w < 4u) team_prefetch<P>(a, L, me, T, w, R);
        team_counts(a, L, me, T, it);
        if (L.status) break;
        mark(PROF_T_COUNTS);
        team_plan(a, L);
        if (L.status) break;
        mark(PROF_T_PLAN);
        const unsigned long long w0 = wprof ? wall_clock64() : 0ull;
        if (w < 4u) {
            team_segments<P>(a, L, me, T, w, R, tab);
        } else {
            if (me == 0u && e.trace_on) team_trace(a, L);
            team_retire(a, L);
        }
        if (wprof && (w == 0u || w == 4u)) e.prof[w == 0u ? PROF_T_SEGW : PROF_T_RETIRE] += wall_clock64() - w0;
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
        if (w == 4u) {
            const unsigned long long s0 = wprof ? wall_clock64() : 0ull;
            team_select(a, L, it + 1u >= a.max_steps);
            if (wprof) e.prof[PROF_T_SELECTW] += wall_clock64() - s0;
        }
        __syncthreads();
        mark(PROF_T_SELECT);
        if (prof) e.prof[PROF_T_STEPS] += 1u;
        if (prof) {
            const unsigned long long c = __builtin_readcyclecounter();
            if (it) e.prof[PROF_C_STEP] += c - cyc;
            cyc = c;
        }
    }
    if (me == 0u && !L.status) team_flush(a, L);
}

}  // namespace gfd
