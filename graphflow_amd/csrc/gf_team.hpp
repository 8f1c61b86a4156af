// Team engine of the batched form: one persistent launch in which a team of K workgroups that
// share one XCD (and therefore one L2) runs whole steps, with one inter-workgroup barrier per step.
//
// Every team member keeps a full copy of the scheduler state in LDS (ready rings, a window of the
// free stack, the selected batches, the plan) and applies the same deterministic updates, so no
// scheduling decision is ever broadcast. The only data the members exchange per step are the
// records themselves and per-segment bin counts:
//
//   * a segment is kSeg consecutive slots of one chunk;
//   * whoever writes a record into a chunk that is routed to a partition node also classifies the
//     record with that node's classifier and adds one to cnt[chunk][gen&1][segment][bin] (a
//     wave-aggregated atomic; the last slot of the row counts invalid bins);
//   * when the partition node later takes the chunk, the counts give the exact plan: per-bin
//     totals, output chunk counts and the flat start of every (segment, bin).
//
// The stable canonical order (bin-major, then batch input order) follows from the per-segment
// prefixes plus a wave-level rank inside each segment, so a member scatters its segments without
// seeing any other member's records. The scheduling rules are those of engine_select/engine_plan/
// engine_retire (sched/GraphRunner.hpp:902-1133 applied once per node per step).
//
// Per step, after the barrier, with no workgroup barrier until the step's end:
//   * waves 0-5 each load their segment's records and the count rows of the segment's node, derive
//     from them alone where the segment's items go (bin totals -> output chunk of each bin, the
//     prefix of the segment's earlier items -> slot) and scatter;
//   * wave 7 loads the same count rows, plans the whole step (pops the output chunks), retires it
//     and selects the next step into the other half of the double-buffered batch state;
//   * wave 6 runs the sink segments.
// Then every wave waits for its stores, the member arrives, and wave 7 pushes the step's inputs on
// the free stack (after the workgroup barrier, because the scatter waves read the popped entries).
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
constexpr uint32_t kTeamJobs = 8;    // job slots of a served engine on the team engine
constexpr uint32_t kTeamBins = 32;  // count row stride: bins of the widest partition + invalid slot
constexpr uint32_t kTeamRoutes = 512;
constexpr uint32_t kSegShift = 7;
constexpr uint32_t kSeg = 1u << kSegShift;  // slots per segment
constexpr uint32_t kSegRounds = kSeg / 64u;
static_assert(kSegRounds == 2u, "segments of two 64-lane rounds");
constexpr uint32_t kTeamRing = 2048;   // ready descriptors over all nodes
constexpr uint32_t kTeamStack = 2048;  // free-stack window
constexpr uint32_t kTeamMargin = 512;  // pops or pushes of one step
constexpr uint32_t kTeamIn = 512;      // input chunks of one step
constexpr uint32_t kTeamOut = 512;     // output chunks of one step
constexpr uint32_t kTeamSegs = 1024;   // segments of one step
constexpr uint32_t kTeamZero = 512;    // count rows consumed by one step's partitions, cleared in the next
constexpr uint32_t kTeamWaves = 8;     // waves 0-5 scatter segments, wave 6 sinks, wave 7 schedules
constexpr uint32_t kWorkers = 6;
constexpr uint32_t kSchedWave = 7;
constexpr uint32_t kSinkWave = 6;
// Order in which the worker waves take a member's segments. Waves w and w + 4 share a SIMD; the
// wave beside the scheduler (wave 3) comes last, so the scheduler rarely competes for issue slots.
__device__ constexpr uint32_t worker_rank(uint32_t w) { return w < 3u ? w : w == 3u ? 5u : w - 1u; }
constexpr uint32_t kClrThreads = 64u * 4u;  // threads of member 0 that spill the free-stack window
constexpr uint32_t kTeamThreads = 64u * kTeamWaves;
constexpr uint32_t kTeamSpin = 1u << 24;

// Team words: the team's XCC (+1), the step barrier, the closed team size (bit 31 set once the
// leader closed registration) and one registration counter per XCC. The barrier is one 64-bit
// word: arrivals in bits 0-31, members that failed an item in an even step in bits 32-47 and in
// an odd step in bits 48-63, so a member learns of the previous step's failures in the same load
// that ends its wait.
enum { TW_XCC0 = 1, TW_BAR = 2, TW_TEAM = 6, TW_REG = 8, TW_WORDS = 32 };
constexpr uint32_t kBarWord = TW_BAR / 2u;  // index of the barrier in 64-bit words
constexpr uint32_t kJoinSpin = 1u << 12;  // leader's wait for the rest of the team (polls of L2)
enum { EXIT_NONE = 0, EXIT_DONE = 1, EXIT_STALL = 2, EXIT_HANDOFF = 3, EXIT_LIMIT = 4, EXIT_STATUS = 5, EXIT_SKIP = 6 };
enum {
    PROF_T_WAIT = 9, PROF_T_COUNTS = 10, PROF_T_PLAN = 11, PROF_T_SEG = 12, PROF_T_SELECT = 13, PROF_T_STEPS = 14,
    PROF_T_SEGW = 15, PROF_T_RETIRE = 16, PROF_T_SELECTW = 17,
    // shader-clock cycles of sub-phases (member 0, lane 0 of the wave doing the work)
    PROF_C_SEL_FILLS = 18, PROF_C_SEL_CREDIT = 19, PROF_C_SEL_CHECK = 20, PROF_C_SEL_LISTS = 21,
    PROF_C_RET_OUT = 22, PROF_C_RET_IN = 23, PROF_C_PLAN_NODES = 24, PROF_C_PLAN_POPS = 25,
    PROF_C_SEG_LOCATE = 26, PROF_C_CNT_SCAN = 27, PROF_C_SEG_RANK = 28, PROF_C_SEG_STORE = 29, PROF_C_SEG_AGG = 30,
    PROF_C_STEP = 31, PROF_C_SEG_PLACE = 7  // slot 7 is free in the team engine
};

// Sub-phase cycle counter: adds the shader-clock delta since the previous mark to an LDS
// accumulator (a fire-and-forget LDS atomic, so the measurement does not stall the wave); member 0
// writes the accumulators to the profile buffer when the launch ends.
struct Ticker {
    unsigned long long* p;  // LDS accumulators, or nullptr
    unsigned long long t;
    __device__ Ticker(unsigned long long* acc, bool enabled) : p(enabled ? acc : nullptr), t(0)
    {
        if (p) t = __builtin_readcyclecounter();
    }
    __device__ void operator()(uint32_t slot)
    {
        if (p) {
            const unsigned long long n = __builtin_readcyclecounter();
            atomicAdd(&p[slot], n - t);
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
    uint32_t cnt_bytes;  // size of cnt (below 2 GiB)
    uint32_t c_shift;  // log2(C) when C is a power of two, else 32
    uint32_t xcd;      // XCC the team must use, plus one; 0: the XCC of the first workgroup to run
    unsigned long long* mprof;  // optional per-member wall-clock totals, kMemberProf words per member
};

// Per-member profile words: barrier wait, step start (window move), scatter phase, scheduler wave,
// segments scattered by the worker waves, sink segments, and from MP_WAVE on the time from the
// step's start until each wave has finished its work.
enum { MP_WAIT = 0, MP_COUNTPLAN = 1, MP_SEG = 2, MP_SCHED = 3, MP_NSEG = 4, MP_NSINK = 5, MP_WAVE = 8, kMemberProf = 16 };

// One step's selected batches; two of them alternate so that the scheduler wave can select step
// s+1 while the other waves still scatter step s.
struct StepBuf {
    uint32_t active[kTeamNodes], n_in[kTeamNodes], particles[kTeamNodes];
    uint32_t ob[kTeamNodes];  // outputs of the sources and transforms before node k (node order)
    uint32_t in_off[kTeamNodes + 1], seg_off[kTeamNodes + 1], nseg[kTeamNodes];
    uint32_t top0;      // free-stack top when the step starts: output j takes entry top0 - 1 - j
    uint32_t g_main;    // segments [0, g_main) belong to non-sink nodes, [g_main, seg_off[n]) to sinks
    unsigned long long first_id[kTeamNodes], digest_base[kTeamNodes];
    uint32_t reserved;  // credits reserved for the step's outputs
    uint32_t recenter;  // move the free-stack window before the step's pops
    uint32_t sum_sink;  // items the step's sinks take
    uint32_t part_mask; // active partition nodes of the step (bit k = node k)
    gf_chunk_desc in[kTeamIn];
    uint32_t in_items[kTeamIn];  // items of the node's batch before input i
    uint32_t in_seg[kTeamIn];    // first segment (step numbering) of input i
};

struct TeamLds {
    gf_node_desc nd[kTeamNodes];
    uint32_t kind[kTeamNodes], bound[kTeamNodes], route_off[kTeamNodes], purpose[kTeamNodes];
    int32_t route[kTeamRoutes], fuse[kTeamRoutes];
    unsigned long long head[kTeamNodes], tail[kTeamNodes];
    gf_chunk_desc ring[kTeamRing];
    unsigned long long fs[kTeamStack];
    StepBuf sb[2];
    // plan of the current step
    uint32_t n_out[kTeamNodes], out_off[kTeamNodes + 1];
    gf_chunk_desc out[kTeamOut];
    uint32_t out_ridx[kTeamOut];  // route-table entry of output j
    uint32_t tot[kTeamNodes * kTeamBins];
    uint32_t binbase[kTeamNodes * kTeamBins];
    uint32_t zero[kTeamZero];  // count rows to clear in the next step: chunk | (generation & 1) << 31
    unsigned long long generated, sunk, step, batches, part_items, xf_items, src_items, sink_items, peak_live, trace_len;
    unsigned long long digest[kTeamNodes];
    // jobs: the node's job slot and seed; per job its items, live bound, admitted and sunk items (an
    // engine that is not served: one job, the run's N, P and seed)
    unsigned long long seed[kTeamNodes];
    uint32_t job[kTeamNodes];
    unsigned long long jN[kTeamJobs], jP[kTeamJobs], jgen[kTeamJobs], jsunk[kTeamJobs];
    uint32_t njobs;
    uint32_t fs_base, top, zero_n, status, exit_code, total_out, err_prev;
    uint32_t seg_err;      // a segment of this member reported a device status this step
    uint32_t prof_member;  // this member records sub-phase timings
    unsigned long long prof_acc[PROF_WORDS];
    unsigned long long wave_acc[kTeamWaves];  // member profile: per-wave work time
};
static_assert(sizeof(TeamLds) + 64u <= 160u * 1024u, "team scheduler state must fit one CU's LDS");

__device__ inline uint32_t xcc_id() { return __builtin_amdgcn_s_getreg(20 | (3 << 11)) & 0xFu; }  // HW_REG_XCC_ID[3:0]

__device__ inline void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---- wave-level helpers ------------------------------------------------------------------------
__device__ inline uint32_t rl(uint32_t v, uint32_t k) { return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)k); }
__device__ inline unsigned long long rl64(unsigned long long v, uint32_t k)
{
    return ((unsigned long long)rl((uint32_t)(v >> 32), k) << 32) | rl((uint32_t)v, k);
}
// Value of `v` in lane `src` (all lanes must be active).
__device__ inline uint32_t from_lane(uint32_t v, uint32_t src) { return (uint32_t)__builtin_amdgcn_ds_bpermute((int)(src << 2), (int)v); }
__device__ inline unsigned long long from_lane64(unsigned long long v, uint32_t src)
{
    return ((unsigned long long)from_lane((uint32_t)(v >> 32), src) << 32) | from_lane((uint32_t)v, src);
}
__device__ inline uint32_t wave_sum(uint32_t v) { return rl(wave_inscan(v), 63); }
// Value of `v` in lane y (per-lane y < m, m small): a scalar loop over the candidate lanes, usable
// in divergent code where ds_bpermute would read inactive lanes.
__device__ inline uint32_t from_lane_nd(uint32_t v, uint32_t y, uint32_t m)
{
    uint32_t r = 0;
    for (uint32_t j = 0; j < m; ++j)
        if (y == j) r = rl(v, j);
    return r;
}

// Node whose [off, off + cnt) range holds item x: the last node j with cnt_j > 0 and off_j <= x
// (off and cnt are lane-per-node values).
__device__ inline uint32_t node_of(uint32_t x, uint32_t off, uint32_t cnt, uint32_t n)
{
    uint32_t kx = 0;
    for (uint32_t j = 0; j < n; ++j)
        if (rl(cnt, j) && x >= rl(off, j)) kx = j;
    return kx;
}

// arr[key] += number of a segment's items with that key, over both 64-lane rounds (key ~0u: none):
// one atomic per distinct key of the segment, so the items of one destination segment in the two
// rounds share it.
__device__ inline void seg_agg_add(const uint32_t (&key)[2], uint32_t* arr)
{
    uint64_t a0 = __ballot(key[0] != ~0u), a1 = __ballot(key[1] != ~0u);
    const uint32_t lane = lane_id();
    while (a0 | a1) {
        const uint32_t lk = a0 ? rl(key[0], (uint32_t)(__ffsll((unsigned long long)a0) - 1))
                               : rl(key[1], (uint32_t)(__ffsll((unsigned long long)a1) - 1));
        const uint64_t m0 = __ballot(key[0] == lk), m1 = __ballot(key[1] == lk);
        if (lane == 0u) atomicAdd(&arr[lk], (uint32_t)(__popcll(m0) + __popcll(m1)));
        a0 &= ~m0;
        a1 &= ~m1;
    }
}

__device__ inline uint32_t div_c(uint32_t x, uint32_t C, uint32_t shift) { return shift < 32u ? x >> shift : x / C; }
// ceil(x / C), with a shift when C is a power of two
__device__ inline uint32_t ceil_c(uint32_t x, uint32_t C, uint32_t shift)
{
    const uint32_t q = div_c(x, C, shift);
    return q + (x - q * C != 0u ? 1u : 0u);
}

// Output-chunk bound of a batch of n items into B bins: min(ceil(n / C) + min(B, n) - 1, n), in
// 32-bit arithmetic
__device__ inline uint32_t team_bound(uint32_t n, uint32_t C, uint32_t shift, uint32_t B)
{
    if (n == 0u) return 0u;
    const uint32_t per = ceil_c(n, C, shift), extra = (B < n ? B : n) - 1u;
    return per > n - extra ? n : per + extra;
}

// ---- team formation --------------------------------------------------------------------------
// The team lives on one XCC: the requested one (a.xcd = XCC + 1), or the XCC of the first
// workgroup that runs. Workgroups elsewhere leave at once, so the launch never waits for CUs that
// other kernels hold. On the team's XCC the first K to register are candidates; the first of them
// (the leader) waits a bounded time for the other K - 1, then closes the team at the count it saw,
// and candidates that registered after that leave too. Returns the member rank (or ~0u) in sh[0]
// and the team size in sh[1].
__device__ inline void team_join(const TeamArgs& a, uint32_t* sh)
{
    if (threadIdx.x == 0) {
        uint32_t* tw = a.tw;
        const uint32_t x = xcc_id();
        uint32_t x0;
        if (a.xcd) {
            x0 = a.xcd - 1u;
        } else {
            const uint32_t prev = atomicCAS(&tw[TW_XCC0], 0u, x + 1u);
            x0 = prev ? prev - 1u : x;
        }
        uint32_t mine = ~0u, T = 0u, ok = 1u;
        if (x == x0) {
            const uint32_t slot = atomicAdd(&tw[TW_REG + x], 1u);
            if (slot < a.K) {
                if (slot == 0u) {
                    uint32_t spins = 0, reg = ld_u32(tw, TW_REG + x);
                    while (reg < a.K && ++spins < kJoinSpin) reg = ld_u32(tw, TW_REG + x);
                    T = reg < a.K ? reg : a.K;
                    atomicExch(&tw[TW_TEAM], 0x80000000u | T);
                } else {
                    uint32_t spins = 0, w = ld_u32(tw, TW_TEAM);
                    while (!(w & 0x80000000u)) {
                        if (++spins > kTeamSpin) {
                            ok = 0u;
                            break;
                        }
                        w = ld_u32(tw, TW_TEAM);
                    }
                    T = w & 0x7FFFFFFFu;
                }
                if (ok && slot < T) mine = slot;
            }
        }
        sh[0] = mine;
        sh[1] = T;
        if (!ok) set_status(&a.e.ctl[CTL_STATUS], GF_DEVICE_ERROR);
    }
    __syncthreads();
}

// Thread 0: wait until every member has arrived `target` times in total (bounded spin); `word` is
// the barrier word that ended the wait.
__device__ inline bool team_wait(const TeamArgs& a, uint32_t target, unsigned long long& word)
{
    uint32_t spins = 0;
    for (;;) {
        word = ld_u64(a.tw, kBarWord);
        if ((uint32_t)word >= target) return true;
        if (++spins > kTeamSpin) return false;
    }
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
        L.fuse[i] = e.fuse[i] < 0 ? -1 : (e.fuse[i] & 0xFFFF);  // the fused node; its stream purpose is 2 x node here
    }
    if (tid < n) {
        L.kind[tid] = e.arr[tid].kind;
        L.bound[tid] = e.arr[tid].reserve_bound;
        L.purpose[tid] = e.arr[tid].purpose;
        L.route_off[tid] = e.nodes[tid].route_offset;
        L.head[tid] = e.st[tid].head;
        L.tail[tid] = e.st[tid].tail;
        L.digest[tid] = e.ctr[CNT_DIGEST0 + tid];
        L.sb[0].active[tid] = 0u;
        L.sb[1].active[tid] = 0u;
        L.job[tid] = e.arr[tid].job;
        L.seed[tid] = e.jobs ? e.jobs[e.arr[tid].job].seed : e.rp.seed;
    }
    if (e.jobs) {
        if (tid < e.n_jobs) {
            L.jN[tid] = e.jobs[tid].N;
            L.jP[tid] = e.jobs[tid].P;
            L.jgen[tid] = e.jobs[tid].gen;
            L.jsunk[tid] = e.jobs[tid].sunk;
        }
    } else if (tid == 0) {
        L.jN[0] = e.N;
        L.jP[0] = e.P;
        L.jgen[0] = e.ctr[CNT_GENERATED];
        L.jsunk[0] = e.ctr[CNT_SUNK];
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
        L.njobs = e.jobs ? e.n_jobs : 1u;
        L.zero_n = a.zero_g[0];
        L.status = 0u;
        L.err_prev = 0u;
        L.seg_err = 0u;
        L.exit_code = EXIT_NONE;
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

// Member 0 writes the pre-select state back in the wide engine's layout, with the count rows the
// last executed step consumed (cleared by the next launch's first step), or the rows this launch was
// handed when it executed no step.
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
    if (e.jobs && tid < L.njobs) {
        e.jobs[tid].gen = L.jgen[tid];
        e.jobs[tid].sunk = L.jsunk[tid];
    }
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

// ---- scheduler registers (wave 7) ---------------------------------------------------------------
// Lane k = node k, kept in registers for the whole launch: ring cursors, the fill and segment sums
// of the node's ready chunks, the sink digest count and the node's static kind and bins.
struct SchedRegs {
    unsigned long long head, tail, digest;
    unsigned long long items;  // items of the node's batches in this launch
    uint32_t rsum, rseg, kind, bins;
    uint32_t nbat;             // batches of the node in this launch
};

__device__ inline void sched_load(const TeamArgs& a, const TeamLds& L, SchedRegs& g)
{
    const uint32_t lane = lane_id(), rcm = a.rc - 1u;
    const bool own = lane < a.e.n_nodes;
    const uint32_t k = own ? lane : 0u;
    g.kind = own ? L.kind[k] : 0xFFu;
    g.bins = g.kind == KIND_PARTITION ? L.nd[k].bins : 1u;
    g.head = own ? L.head[k] : 0ull;
    g.tail = own ? L.tail[k] : 0ull;
    g.digest = own ? L.digest[k] : 0ull;
    g.rsum = 0u;
    g.rseg = 0u;
    g.items = 0ull;
    g.nbat = 0u;
    for (unsigned long long p = g.head; p < g.tail; ++p) {
        const uint32_t f = L.ring[k * a.rc + (uint32_t)(p & rcm)].fill;
        g.rsum += f;
        g.rseg += (f + kSeg - 1u) >> kSegShift;
    }
}

__device__ inline void sched_store(const TeamArgs& a, TeamLds& L, const SchedRegs& g)
{
    const uint32_t lane = lane_id();
    if (lane < a.e.n_nodes) {
        L.head[lane] = g.head;
        L.tail[lane] = g.tail;
        L.digest[lane] = g.digest;
    }
    // launch totals per kind
    unsigned long long src = 0, part = 0, xf = 0, nb = 0;
    for (uint32_t j = 0; j < a.e.n_nodes; ++j) {
        const uint32_t kj = rl(g.kind, j);
        const unsigned long long it = rl64(g.items, j);
        if (kj == KIND_SOURCE) src += it;
        else if (kj == KIND_PARTITION) part += it;
        else if (kj == KIND_TRANSFORM) xf += it;
        nb += rl(g.nbat, j);
    }
    if (lane == 0) {
        L.src_items += src;
        L.part_items += part;
        L.xf_items += xf;
        L.batches += nb;
    }
    wave_sync();
}

// ---- select (wave 7): the next step's batches into Bn, or the reason to stop ------------------
// Credits are handed out in node order by a scalar loop over the nodes that may take a batch; a
// node that takes its whole ring needs no ring reads (the ring sums are in registers).
__device__ inline void team_select(const TeamArgs& a, TeamLds& L, StepBuf& Bn, bool at_limit, SchedRegs& g)
{
    const Engine& e = a.e;
    const uint32_t lane = lane_id(), n = e.n_nodes, C = e.C, rcm = a.rc - 1u;
    Ticker tk(L.prof_acc, L.prof_member && lane == 0u);
    const bool own = lane < n;
    const uint32_t k = own ? lane : 0u;
    const uint32_t kind = g.kind, B = g.bins;
    const bool xf = kind == KIND_PARTITION || kind == KIND_TRANSFORM;
    const unsigned long long head = g.head, tail = g.tail;
    const uint32_t avail = (uint32_t)(tail - head);
    const gf_chunk_desc* rg = L.ring + k * a.rc;
    uint32_t want = 0, parts = 0, nseg = 0;
    if (kind == KIND_SINK || (xf && avail <= e.Q)) {
        want = avail;
        parts = g.rsum;
        nseg = g.rseg;
    } else if (xf) {
        want = e.Q;
        for (uint32_t i = 0; i < want; ++i) {
            const uint32_t f = rg[(uint32_t)((head + i) & rcm)].fill;
            parts += f;
            nseg += (f + kSeg - 1u) >> kSegShift;
        }
    }
    uint32_t bound = xf ? team_bound(parts, C, a.c_shift, B) : 0u;
    tk(PROF_C_SEL_FILLS);
    // credits in node order, over the nodes that may take a batch
    const uint32_t top = L.top;
    uint32_t credits = top;
    unsigned long long gen = L.generated;
    uint32_t src_fill = 0;
    // a source's admission (live < P, room for min(C, P, remaining); one source per job, against its
    // job's counters) does not depend on credits beyond the one chunk it needs; each source lane
    // evaluates it up front
    uint32_t adm = 0;
    const uint32_t jb = own ? L.job[k] : 0u;
    const unsigned long long gen0 = kind == KIND_SOURCE ? L.jgen[jb] : 0ull;
    if (kind == KIND_SOURCE) {
        const unsigned long long P = L.jP[jb], N = L.jN[jb], live = gen0 - L.jsunk[jb], remaining = N - gen0;
        unsigned long long min_fill = C < P ? C : P;
        if (remaining < min_fill) min_fill = remaining;
        if (gen0 < N && live < P && P - live >= min_fill) {
            unsigned long long fill = remaining;
            if (fill > P - live) fill = P - live;
            if (fill > C) fill = C;
            adm = (uint32_t)fill;
        }
    }
    uint64_t cand = __ballot(adm != 0u || (xf && want));
    while (cand) {
        const uint32_t j = (uint32_t)(__ffsll((unsigned long long)cand) - 1);
        cand &= cand - 1u;
        if (rl(kind, j) == KIND_SOURCE) {
            if (credits >= 1u) {
                if (lane == j) src_fill = adm;
                gen += rl(adm, j);
                credits -= 1u;
            }
        } else {
            if (rl(bound, j) > credits && lane == j) {
                // largest prefix of the candidates whose bound fits the remaining credits
                uint32_t m = want, p = parts, sg = nseg, b = 0;
                while (m) {
                    b = team_bound(p, C, a.c_shift, B);
                    if (b <= credits) break;
                    --m;
                    const uint32_t f = rg[(uint32_t)((head + m) & rcm)].fill;
                    p -= f;
                    sg -= (f + kSeg - 1u) >> kSegShift;
                }
                want = m;
                parts = p;
                nseg = sg;
                bound = m ? b : 0u;
            }
            credits -= rl(bound, j);
        }
    }
    tk(PROF_C_SEL_CREDIT);
    uint32_t nin = 0, act = 0;
    if (kind == KIND_SOURCE) {
        act = src_fill ? 1u : 0u;
        parts = src_fill;
        nseg = (parts + kSeg - 1u) >> kSegShift;
    } else if (own && want) {
        nin = want;
        act = 1u;
    } else {
        parts = 0u;
        nseg = 0u;
    }
    // outputs of the sources and transforms (the partitions' follow from the counts)
    const uint32_t rnp = !act ? 0u : kind == KIND_SOURCE ? 1u : kind == KIND_TRANSFORM ? ceil_c(parts, C, a.c_shift) : 0u;
    // segments are numbered in node order, the sinks' after everybody else's
    const uint32_t nsm = kind == KIND_SINK ? 0u : nseg, nss = kind == KIND_SINK ? nseg : 0u;
    // two scans of 16-bit field pairs: inputs and outputs (each below kTeamRing + kTeamNodes over
    // all nodes), and the segment counts clamped to kTeamSegs + 1 (a clamped count fails the fit
    // check below either way, and 32 of them stay below 2^16)
    static_assert(kTeamRing + kTeamNodes < 65536u && kTeamNodes * (kTeamSegs + 1u) < 65536u, "16-bit scan fields");
    const uint32_t ksm = nsm <= kTeamSegs ? nsm : kTeamSegs + 1u, kss = nss <= kTeamSegs ? nss : kTeamSegs + 1u;
    const uint32_t io_inc = wave_inscan(nin | rnp << 16), s_inc = wave_inscan(ksm | kss << 16);
    const uint32_t i_inc = io_inc & 0xFFFFu, o_inc = io_inc >> 16, sm_inc = s_inc & 0xFFFFu, ss_inc = s_inc >> 16;
    const uint32_t g_main = rl(sm_inc, 63);
    const uint32_t in_tot = rl(i_inc, 63), seg_tot = g_main + rl(ss_inc, 63);
    const uint32_t reserved = top - credits;
    const uint64_t actm = __ballot(act != 0u), pend = __ballot(own && tail != head);
    uint32_t code = EXIT_NONE;
    // every job fully admitted (lane j: job j)
    const bool admitted = __ballot(lane < L.njobs && L.jgen[lane < kTeamJobs ? lane : 0u] < L.jN[lane < kTeamJobs ? lane : 0u]) == 0ull;
    if (!actm) {
        code = (!pend && admitted) ? EXIT_DONE : EXIT_STALL;
    } else if (at_limit) {
        code = EXIT_LIMIT;
    } else {
        const bool fits = in_tot <= kTeamIn && seg_tot <= kTeamSegs && reserved <= kTeamOut && reserved <= kTeamMargin &&
                          in_tot <= kTeamMargin;
        const bool room = !own || (avail - nin) + reserved <= a.rc;
        if (!fits || __ballot(!room)) code = EXIT_HANDOFF;
    }
    tk(PROF_C_SEL_CHECK);
    if (code) {
        if (lane == 0) L.exit_code = code;
        wave_sync();
        return;
    }
    const uint32_t inoff = i_inc - nin, segoff = kind == KIND_SINK ? g_main + ss_inc - nss : sm_inc - nsm;
    if (own) {
        Bn.active[k] = act;
        Bn.n_in[k] = nin;
        Bn.particles[k] = parts;
        Bn.in_off[k] = inoff;
        Bn.seg_off[k] = segoff;
        Bn.nseg[k] = nseg;
        Bn.ob[k] = o_inc - rnp;
        if (kind == KIND_SINK && act) {
            Bn.digest_base[k] = g.digest;
            g.digest += nin;
        }
        if (kind == KIND_SOURCE && act) {
            Bn.first_id[k] = gen0;
            L.jgen[jb] = gen0 + parts;  // one source per job
        }
        if (act && kind != KIND_SOURCE) {
            g.head = head + nin;
            g.rsum -= parts;
            g.rseg -= nseg;
        }
    }
    if (act) {
        g.items += parts;
        if (xf) g.nbat += 1u;
    }
    uint32_t sum_sink = 0;
    for (uint64_t sm = __ballot(act && kind == KIND_SINK); sm; sm &= sm - 1u)
        sum_sink += rl(parts, (uint32_t)(__ffsll((unsigned long long)sm) - 1));
    if (lane == n) {
        Bn.in_off[n] = in_tot;
        Bn.seg_off[n] = seg_tot;
    }
    const uint32_t pmask = (uint32_t)__ballot(act && kind == KIND_PARTITION);  // nodes < 32
    if (lane == 0) {
        L.generated = gen;
        Bn.g_main = g_main;
        Bn.reserved = reserved;
        Bn.sum_sink = sum_sink;
        Bn.part_mask = pmask;
        Bn.top0 = top;
        const uint32_t nb = top > kTeamStack / 2u ? top - kTeamStack / 2u : 0u;
        const uint32_t fb = L.fs_base;
        Bn.recenter = (nb != fb && (top < fb + kTeamMargin || top + kTeamMargin > fb + kTeamStack)) ? 1u : 0u;
        L.exit_code = EXIT_NONE;
    }
    // per active node (node order): its inputs' descriptors and first segments, and for plain
    // transforms the items before each input
    for (uint64_t actn = __ballot(act && kind != KIND_SOURCE); actn; actn &= actn - 1u) {
        const uint32_t kk = (uint32_t)(__ffsll((unsigned long long)actn) - 1);
        const uint32_t m = rl(nin, kk), io = rl(inoff, kk), so = rl(segoff, kk);
        const unsigned long long h = rl64(head, kk);
        const bool plain = rl(kind, kk) == KIND_TRANSFORM;
        const gf_chunk_desc* rk = L.ring + kk * a.rc;
        uint32_t carry_f = 0, carry_s = 0;
        for (uint32_t y0 = 0; y0 < m; y0 += 64u) {
            const uint32_t y = y0 + lane;
            const bool v = y < m;
            gf_chunk_desc d{0, 0, 0, 0};
            if (v) d = rk[(uint32_t)((h + y) & rcm)];
            const uint32_t ns = (d.fill + kSeg - 1u) >> kSegShift, si = wave_inscan(ns);
            const uint32_t first = so + carry_s + si - ns;
            if (v) {
                Bn.in[io + y] = d;
                Bn.in_seg[io + y] = first;
            }
            carry_s += rl(si, 63);
            if (plain) {
                const uint32_t fi = wave_inscan(d.fill);
                if (v) Bn.in_items[io + y] = carry_f + fi - d.fill;
                carry_f += rl(fi, 63);
            }
        }
    }
    wave_sync();
    tk(PROF_C_SEL_LISTS);
}

// Move the free-stack window so that the step's pops and pushes stay inside it (all threads).
__device__ inline void team_recenter(const TeamArgs& a, TeamLds& L, uint32_t me)
{
    const Engine& e = a.e;
    const uint32_t tid = threadIdx.x;
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
    if (me == 0 && tid < kClrThreads)
        for (uint32_t gi = old + tid; gi < nb && gi < top; gi += kClrThreads) e.free_stack[gi] = L.fs[gi - old];
    __syncthreads();
#pragma unroll
    for (uint32_t r = 0; r < per; ++r) {
        const uint32_t gi = (nb > old ? nb : old) + tid + r * kTeamThreads;
        if (gi < top) L.fs[gi - nb] = keep[r];
    }
    // entries [nb, old) come back from global memory (written by an earlier spill or launch)
    for (uint32_t gi = nb + tid; gi < old; gi += kTeamThreads) L.fs[gi - nb] = ld_u64(e.free_stack, gi);
    __syncthreads();
    if (tid == 0) L.fs_base = nb;
    __syncthreads();
}

// ---- count rows ----------------------------------------------------------------------------------
// Sink wave: clear this member's share of the count rows the previous step's partitions consumed (or,
// in a launch's first step, the rows the previous launch left). A chunk's rows are one contiguous
// block of tpc * bb words; its next use as a partition input has the other generation parity, so
// the rows are zero again before any producer counts into them.
__device__ inline void team_clear_rows(const TeamArgs& a, const TeamLds& L, uint32_t me, uint32_t T)
{
    const uint32_t lane = lane_id(), blk = a.tpc * a.bb, zw = L.zero_n * blk;
    for (uint32_t x = me * 64u + lane; x < zw; x += T * 64u) {
        const uint32_t z = L.zero[x / blk];
        a.cnt[((size_t)(z & 0x7FFFFFFFu) * 2u + (z >> 31)) * blk + x % blk] = 0u;
    }
}

// Per-bin sums of the count rows of partition node k's segments [g0, g1) (segments numbered within
// the node's batch, in batch input order): lane lofs + b receives in `all` the sum of slot b over
// [g0, g1) and in `lo` the sum over [g0, split) (b < B1; slot B1 - 1 counts invalid bins). Rows are
// bb words (a multiple of four); each lane takes two segments and loads their rows four slots at
// a time with 128-bit agent-coherent (sc1) buffer loads, which read zero past the array's end, so
// lanes without a segment load from there. A row entry counts at most kSeg items, so two slots
// share one word as 16-bit fields whose sums over the 128 rows of a pass stay below 2^15: one
// wave reduction per slot pair for the totals and one for the prefix. `tsum` (wave-uniform)
// receives the sum of `all` over the slots.
__device__ inline void seg_sums(const TeamArgs& a, const StepBuf& Bc, uint32_t k, uint32_t B1, uint32_t g0, uint32_t split,
                                uint32_t g1, uint32_t lofs, uint32_t& lo, uint32_t& all, uint32_t& tsum)
{
    static_assert(kSeg * 128u < 65536u, "16-bit count fields");
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    constexpr int kSc1 = 16;  // cache policy: agent scope (bypasses the CU's L1)
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc((void*)a.cnt, (short)0, (int)a.cnt_bytes, 0x00020000);
    const uint32_t lane = lane_id(), so = Bc.seg_off[k], io = Bc.in_off[k], mi = Bc.n_in[k];
    // first segment (step numbering) and count-row base of input `lane`; batches of more than 64
    // inputs search LDS
    const gf_chunk_desc dl = lane < mi ? Bc.in[io + lane] : gf_chunk_desc{0, 0, 0, 0};
    const uint32_t fl = lane < mi ? Bc.in_seg[io + lane] : 0xFFFFFFFFu;
    const uint32_t rowl = ((dl.index * 2u + (dl.generation & 1u)) * a.tpc) * a.bb;
    lo = 0u;
    all = 0u;
    tsum = 0u;
    for (uint32_t h = g0; h < g1; h += 128u) {
        // the input holding segment h + lane (a) and h + 64 + lane (b): the last one starting at or before it
        const uint32_t ga = so + h + lane, gb = ga + 64u;
        uint32_t fa = rl(fl, 0), ra = rl(rowl, 0), fb = fa, rb = ra;
        if (mi <= 64u) {
            for (uint32_t y = 1; y < mi; ++y) {
                const uint32_t fy = rl(fl, y), ry = rl(rowl, y);
                const bool ca = fy <= ga, cb = fy <= gb;
                fa = ca ? fy : fa;
                ra = ca ? ry : ra;
                fb = cb ? fy : fb;
                rb = cb ? ry : rb;
            }
        } else {
            for (uint32_t y = 1; y < mi; ++y) {
                const uint32_t fy = Bc.in_seg[io + y];
                if (fy <= gb) {
                    const gf_chunk_desc d = Bc.in[io + y];
                    const uint32_t ry = ((d.index * 2u + (d.generation & 1u)) * a.tpc) * a.bb;
                    if (fy <= ga) {
                        fa = fy;
                        ra = ry;
                    }
                    fb = fy;
                    rb = ry;
                }
            }
        }
        // byte offsets of the two rows, one past the array for segments >= g1
        const uint32_t oa = h + lane < g1 ? (ra + (ga - fa) * a.bb) * 4u : a.cnt_bytes;
        const uint32_t ob = h + 64u + lane < g1 ? (rb + (gb - fb) * a.bb) * 4u : a.cnt_bytes;
        // rows before `split` in this pass: none, all, or those of the lanes below it
        const bool none = split <= h, whole = split >= h + 128u;
        const bool pa = h + lane < split, pb = h + 64u + lane < split;
        for (uint32_t c0 = 0; c0 < B1; c0 += 4u) {
            const u32x4 va = __builtin_amdgcn_raw_buffer_load_b128(rs, oa + c0 * 4u, 0, kSc1);
            const u32x4 vb = __builtin_amdgcn_raw_buffer_load_b128(rs, ob + c0 * 4u, 0, kSc1);
            const uint32_t a01 = va[0] | va[1] << 16, a23 = va[2] | va[3] << 16;
            const uint32_t b01 = vb[0] | vb[1] << 16, b23 = vb[2] | vb[3] << 16;
            const uint32_t t01 = rl(wave_inscan(a01 + b01), 63), t23 = rl(wave_inscan(a23 + b23), 63);
            const uint32_t nq = B1 - c0;  // slots of this group
            tsum += (t01 & 0xFFFFu) + (nq > 1u ? t01 >> 16 : 0u) + (nq > 2u ? t23 & 0xFFFFu : 0u) + (nq > 3u ? t23 >> 16 : 0u);
            uint32_t p01 = 0u, p23 = 0u;
            if (whole) {
                p01 = t01;
                p23 = t23;
            } else if (!none) {
                p01 = rl(wave_inscan((pa ? a01 : 0u) + (pb ? b01 : 0u)), 63);
                p23 = rl(wave_inscan((pa ? a23 : 0u) + (pb ? b23 : 0u)), 63);
            }
            // slot c0 + q goes to lane lofs + c0 + q
            const uint32_t q = lane - lofs - c0, sh = (q & 1u) * 16u;
            if (q < 4u && c0 + q < B1) {
                all += ((q < 2u ? t01 : t23) >> sh) & 0xFFFFu;
                lo += ((q < 2u ? p01 : p23) >> sh) & 0xFFFFu;
            }
        }
    }
}

// ---- plan (wave 7) ------------------------------------------------------------------------------
// Every node's exact output count from the count rows (partitions) or the batch (sources,
// transforms), the checks of the batch, and the step's output chunks popped off the free stack.
__device__ inline void team_plan(const TeamArgs& a, TeamLds& L, const StepBuf& Bc, const SchedRegs& g)
{
    const Engine& e = a.e;
    const uint32_t lane = lane_id(), n = e.n_nodes, C = e.C;
    Ticker tk(L.prof_acc, L.prof_member && lane == 0u);
    // lanes 0..31: nodes; lanes 32..63: bins of the first partition node of the step
    const bool own = lane < n;
    const uint32_t k = own ? lane : 0u;
    const uint32_t act = own ? Bc.active[k] : 0u, kind = g.kind;
    const uint32_t parts = own ? Bc.particles[k] : 0u, bound = own ? L.bound[k] : 0u, ro = own ? L.route_off[k] : 0u;
    const uint32_t pm = Bc.part_mask, reserved = Bc.reserved, top = L.top, fbase = L.fs_base;
    const uint32_t kp = pm ? (uint32_t)(__ffs(pm) - 1) : 0u;
    const uint32_t B = pm ? rl(g.bins, kp) : 0u;
    uint32_t t = 0u, tall = 0u;  // tall: items over all slots (equals the valid bins' sum unless slot B counts)
    if (pm) {
        uint32_t unused;
        seg_sums(a, Bc, kp, B + 1u, 0u, 0u, Bc.nseg[kp], 32u, unused, t, tall);
    }
    tk(PROF_C_CNT_SCAN);
    const uint32_t bl = lane - 32u;  // bin of this lane (lanes >= 32)
    const uint32_t tv = (lane >= 32u && bl < B) ? t : 0u, r = ceil_c(tv, C, a.c_shift);
    const uint32_t ri = wave_inscan(r);
    const uint32_t bbase = ri - r;  // first output of bin bl within the partition's outputs
    uint32_t R = 0, st = 0, inv = 0;
    if (act) {
        if (kind == KIND_TRANSFORM) {
            R = ceil_c(parts, C, a.c_shift);
            if (parts == 0u || R > bound) st = GF_PLAN_MISMATCH;
        } else if (kind == KIND_SOURCE) {
            R = 1u;
        }
    }
    if (pm) {
        const uint32_t Rk = rl(ri, 63);
        if (rl(t, 32u + B)) inv = 1u;
        if (lane == kp) {
            R = Rk;
            if (tall != parts || Rk > bound) st = GF_PLAN_MISMATCH;
        }
        if (lane >= 32u && bl < B) L.binbase[kp * kTeamBins + bl] = bbase;
        if (lane >= 32u && bl <= B) L.tot[kp * kTeamBins + bl] = t;
    }
    // further partition nodes of the same step (rare): bins in lanes 0.., one pass each
    for (uint32_t pmk = pm & (pm - 1u); pmk; pmk &= pmk - 1u) {
        const uint32_t kq = (uint32_t)(__ffs(pmk) - 1);
        const uint32_t Bq = rl(g.bins, kq);
        uint32_t unused, tq, sq;
        seg_sums(a, Bc, kq, Bq + 1u, 0u, 0u, Bc.nseg[kq], 0u, unused, tq, sq);
        const uint32_t tqv = lane < Bq ? tq : 0u, rq = ceil_c(tqv, C, a.c_shift);
        const uint32_t rqi = wave_inscan(rq);
        if (lane < Bq) L.binbase[kq * kTeamBins + lane] = rqi - rq;
        if (lane <= Bq) L.tot[kq * kTeamBins + lane] = tq;
        // wave-wide values first: a readlane inside the lane branch would let the scans move
        // under that branch's exec mask
        const uint32_t Rq = rl(rqi, 63);
        if (rl(tq, Bq)) inv = 1u;
        if (lane == kq) {
            R = Rq;
            if (sq != parts || Rq > bound) st = GF_PLAN_MISMATCH;
        }
    }
    wave_sync();
    const uint64_t bad = __ballot(st != 0u);
    uint32_t status = inv ? (uint32_t)GF_INVALID : bad ? rl(st, (uint32_t)(__ffsll((unsigned long long)bad) - 1)) : 0u;
    const uint32_t oi = wave_inscan(own ? R : 0u), off = oi - R, total = rl(oi, 63);
    if (!status && total > reserved) status = GF_PLAN_MISMATCH;
    if (!status && total > top) status = GF_POOL_EXHAUSTED;
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
        auto emit = [&](uint32_t j0, uint32_t cnt, uint32_t fill0, uint32_t rem0, uint32_t bin, uint32_t ridx) {
            // outputs j0 .. j0+cnt-1: fills C, C, ..., the rest of rem0 (fill0 = C or the whole)
            for (uint32_t q = lane; q < cnt; q += 64u) {
                const uint32_t j = j0 + q;
                const unsigned long long ent = L.fs[top - 1u - j - fbase];
                const uint32_t rem = rem0 - q * fill0;
                gf_chunk_desc d;
                d.index = (uint32_t)ent;
                d.generation = (uint32_t)(ent >> 32);
                d.fill = rem < fill0 ? rem : fill0;
                d.bin = bin;
                L.out[j] = d;
                L.out_ridx[j] = ridx;
            }
        };
        // node lanes emit their source / transform outputs and bin lanes the first partition's,
        // all lanes at once
        const uint32_t kpo = rl(off, kp), kpr = rl(ro, kp);
        uint32_t cnt = 0, j0 = 0, fill0 = C, rem0 = 0, bin = 0, ridx = 0;
        if (own && R && (kind == KIND_SOURCE || kind == KIND_TRANSFORM)) {
            cnt = R;
            j0 = off;
            fill0 = kind == KIND_SOURCE ? parts : C;
            rem0 = parts;
            ridx = ro;
        } else if (lane >= 32u && r) {
            cnt = r;
            j0 = kpo + bbase;
            rem0 = tv;
            bin = bl;
            ridx = kpr + bl;
        }
        for (uint32_t q = 0; q < cnt; ++q) {
            const uint32_t j = j0 + q;
            const unsigned long long ent = L.fs[top - 1u - j - fbase];
            const uint32_t rem = rem0 - q * fill0;
            gf_chunk_desc d;
            d.index = (uint32_t)ent;
            d.generation = (uint32_t)(ent >> 32);
            d.fill = rem < fill0 ? rem : fill0;
            d.bin = bin;
            L.out[j] = d;
            L.out_ridx[j] = ridx;
        }
        // further partition nodes of the step, one bin at a time
        for (uint32_t pmk = pm & (pm - 1u); pmk; pmk &= pmk - 1u) {
            const uint32_t kq = (uint32_t)(__ffs(pmk) - 1);
            const uint32_t Bq = rl(g.bins, kq), base = rl(off, kq), rok = rl(ro, kq);
            for (uint32_t x = 0; x < Bq; ++x) {
                const uint32_t tx = L.tot[kq * kTeamBins + x];
                if (tx) emit(base + L.binbase[kq * kTeamBins + x], ceil_c(tx, C, a.c_shift), C, tx, x, rok + x);
            }
        }
        if (lane == 0) {
            L.top = top - total;
            const unsigned long long live = e.S - (top - total);
            if (live > L.peak_live) L.peak_live = live;
        }
    }
    wave_sync();
    tk(PROF_C_PLAN_POPS);
}

// ---- trace (member 0) and retire (wave 7 of every member) --------------------------------------
__device__ inline void team_trace(const TeamArgs& a, TeamLds& L, const StepBuf& Bc)
{
    const Engine& e = a.e;
    const uint32_t lane = lane_id(), n = e.n_nodes;
    for (uint32_t k = 0; k < n; ++k) {
        if (!Bc.active[k]) continue;
        const uint32_t kind = L.kind[k];
        const uint32_t nin = kind == KIND_SOURCE ? 0u : Bc.n_in[k];
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
            const gf_chunk_desc d = Bc.in[Bc.in_off[k] + i];
            put(3 + 3ull * i, d.index);
            put(4 + 3ull * i, d.generation);
            put(5 + 3ull * i, d.fill);
        }
        for (uint32_t b = lane; b < nb; b += 64u)
            put(4 + 3ull * nin + b, kind == KIND_PARTITION ? L.tot[k * kTeamBins + b] : Bc.particles[k]);
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

// Retire (wave 7 of every member): outputs to the destination rings in node and output order,
// the free-stack top and the counters.
__device__ inline void team_retire(const TeamArgs& a, TeamLds& L, StepBuf& Bc, SchedRegs& g)
{
    const Engine& e = a.e;
    const uint32_t lane = lane_id(), n = e.n_nodes, rcm = a.rc - 1u;
    Ticker tk(L.prof_acc, L.prof_member && lane == 0u);
    const uint32_t total_out = L.total_out, in_tot = Bc.in_off[n], top = L.top;
    for (uint32_t x0 = 0; x0 < total_out; x0 += 64u) {
        const uint32_t x = x0 + lane;
        const bool v = x < total_out;
        gf_chunk_desc d{0, 0, 0, 0};
        uint32_t dst = 0xFFFFFFFFu;
        if (v) {
            d = L.out[x];
            dst = (uint32_t)L.route[L.out_ridx[x]];
        }
        const uint32_t ns = (d.fill + kSeg - 1u) >> kSegShift;
        uint64_t pend = __ballot(v);
        unsigned long long pos = 0;
        while (pend) {
            const uint32_t lb = rl(dst, (uint32_t)(__ffsll((unsigned long long)pend) - 1));
            const bool mine = v && dst == lb;
            const uint64_t m = __ballot(mine);
            const unsigned long long tl = rl64(g.tail, lb);
            if (mine) pos = tl + (uint32_t)__popcll(m & lanemask_lt());
            // fills and segment counts as one sum: the step's outputs hold at most kTeamSegs * kSeg
            // items (18 bits) in at most kTeamSegs + kTeamOut segments
            const uint32_t fss = wave_sum(mine ? d.fill | ns << 18 : 0u);
            if (lane == lb) {
                g.tail += (uint32_t)__popcll(m);
                g.rsum += fss & 0x3FFFFu;
                g.rseg += fss >> 18;
            }
            pend &= ~m;
        }
        if (v) L.ring[dst * a.rc + (uint32_t)(pos & rcm)] = d;
    }
    tk(PROF_C_RET_OUT);
    // the consumed inputs go back on the free stack in team_push, once the scatter waves no longer
    // read the popped entries; the top moves now, for the next step's credits
    // sunk items per job (lane = node; an engine that is not served: job 0 = L.sunk)
    if (lane < e.n_nodes && g.kind == KIND_SINK && Bc.active[lane])
        atomicAdd(&L.jsunk[L.job[lane]], (unsigned long long)Bc.particles[lane]);
    if (lane == 0) {
        L.top = top + in_tot;
        L.sunk += Bc.sum_sink;
        L.sink_items += Bc.sum_sink;
        L.step += 1u;
    }
    wave_sync();
    tk(PROF_C_RET_IN);
}

// Wave 6, after the step's workgroup barrier: the step's consumed inputs back on the free stack
// (generation + 1), where its pops were taken from, and the partition inputs' count rows on the
// list the sink wave clears in the next step.
__device__ inline void team_push(const TeamArgs& a, TeamLds& L, const StepBuf& Bc)
{
    const uint32_t lane = lane_id(), in_tot = Bc.in_off[a.e.n_nodes];
    const uint32_t ptop = Bc.top0 - L.total_out, fb = L.fs_base;
    for (uint32_t x = lane; x < in_tot; x += 64u) {
        const gf_chunk_desc d = Bc.in[x];
        L.fs[ptop + x - fb] = (unsigned long long)d.index | ((unsigned long long)(d.generation + 1u) << 32);
    }
    uint32_t zn = 0;
    for (uint32_t pmk = Bc.part_mask; pmk; pmk &= pmk - 1u) {
        const uint32_t k = (uint32_t)(__ffs(pmk) - 1);
        const uint32_t io = Bc.in_off[k], m = Bc.n_in[k];
        for (uint32_t i = lane; i < m; i += 64u) {
            const gf_chunk_desc d = Bc.in[io + i];
            L.zero[zn + i] = d.index | (d.generation & 1u) << 31;
        }
        zn += m;
    }
    if (lane == 0) L.zero_n = zn;
    wave_sync();
}

// A failed item: the run's status word for the host, this member's flag for the team's stop decision.
__device__ inline void seg_fail(const TeamArgs& a, const TeamLds& L, uint32_t code)
{
    set_status(&a.e.ctl[CTL_STATUS], code);
    const_cast<TeamLds&>(L).seg_err = 1u;
}

// ---- segment work (waves 0-5 and the sink wave) ---------------------------------------------------------------------
// Next-node count key of a record stored at (chunk d, offset off), or ~0u when the destination is
// not a partition.
template <class P>
__device__ inline uint32_t next_key(const TeamArgs& a, uint32_t dkind, const gf_node_desc& nd, const gf_chunk_desc& d, uint32_t off,
                                    const P& rec)
{
    if (dkind != KIND_PARTITION) return ~0u;
    uint32_t nb = classify(nd, rec);
    if (nb >= nd.bins) nb = nd.bins;
    return (((uint32_t)d.index * 2u + (d.generation & 1u)) * a.tpc + (off >> kSegShift)) * a.bb + nb;
}

// Records of one input segment, requested by the wave that scatters it before it reads the count rows.
template <class P>
struct SegRecs {
    P rec[kSegRounds];
    uint32_t s;  // segment index, or ~0u
};

// Node, input chunk of the node's batch and segment of that chunk holding step segment s
// (wave-uniform): a ballot over the nodes' segment ranges, then over the node's inputs.
struct SegLoc {
    uint32_t k, i, gg, kind;
};

__device__ inline SegLoc locate(const TeamLds& L, const StepBuf& Bc, uint32_t n, uint32_t s)
{
    const uint32_t lane = lane_id();
    const uint32_t lo = lane < n ? Bc.seg_off[lane] : 0xFFFFFFFFu, hi = lane < n ? lo + Bc.nseg[lane] : 0u;
    SegLoc r;
    r.k = (uint32_t)(__ffsll((unsigned long long)__ballot(lane < n && lo <= s && s < hi)) - 1);
    const uint32_t so = rl(lo, r.k);
    r.i = 0u;
    r.gg = s - so;
    r.kind = L.kind[r.k];
    if (r.kind == KIND_SOURCE) return r;
    const uint32_t io = Bc.in_off[r.k], m = Bc.n_in[r.k];
    for (uint32_t y0 = 0; y0 < m; y0 += 64u) {
        const uint32_t f = y0 + lane < m ? Bc.in_seg[io + y0 + lane] : 0xFFFFFFFFu;
        const uint64_t ym = __ballot(f <= s);
        if (ym) {
            const uint32_t hb = 63u - (uint32_t)__clzll((unsigned long long)ym);
            r.i = y0 + hb;
            r.gg = s - rl(f, hb);
        }
    }
    return r;
}

template <class P>
__device__ inline void load_segment(const TeamArgs& a, const StepBuf& Bc, const SegLoc& loc, uint32_t s, SegRecs<P>& R)
{
    const uint32_t k = loc.k, i = loc.i, gg = loc.gg;
    R.s = ~0u;
    if (loc.kind == KIND_SOURCE) return;
    const gf_chunk_desc d = Bc.in[Bc.in_off[k] + i];
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

// Per-bin constants of one partition node, bin b in lane b of the wave, loaded once per launch
// and packed for one ds_bpermute per value: w0 = destination after fusion | (fused node + 1) << 8
// | fused op << 16 | destination classifier op << 24 (0 when the destination is not a
// partition), w1 = destination bins | destination u[0] << 16, and the two f[0] parameters.
struct BinTab {
    uint32_t k;  // node described (wave-uniform), or ~0u
    uint32_t w0, w1;
    float xf0, df0;
};

__device__ inline void load_bintab(const TeamLds& L, uint32_t k, BinTab& t)
{
    const uint32_t lane = lane_id(), B = L.nd[k].bins, ro = L.route_off[k];
    t.k = k;
    t.w0 = 0u;
    t.w1 = 1u;
    t.xf0 = 1.0f;
    t.df0 = 0.0f;
    if (lane < B) {
        const int32_t fz = L.fuse[ro + lane];
        const uint32_t dst = (uint32_t)L.route[ro + lane];
        uint32_t xop = 0, dop = 0;
        if (fz >= 0) {
            xop = L.nd[fz].op;
            t.xf0 = L.nd[fz].f[0];
        }
        if (L.kind[dst] == KIND_PARTITION) {
            const gf_node_desc& dn = L.nd[dst];
            dop = dn.op;
            t.w1 = dn.bins | (dn.u[0] & 0xFFFFu) << 16;
            t.df0 = dn.f[0];
        }
        t.w0 = dst | (uint32_t)(fz + 1) << 8 | xop << 16 | dop << 24;
    }
}

// Rank of every valid lane inside the segment after `run` earlier items of its bin (lane b of run
// holds bin b's running count; bins < 64).
__device__ inline uint32_t seg_rank(bool v, uint32_t bin, uint32_t& run)
{
    uint64_t act = __ballot(v);
    uint32_t pos = 0;
    const uint32_t lane = lane_id();
    while (act) {
        const uint32_t lb = rl(bin, (uint32_t)(__ffsll((unsigned long long)act) - 1));
        const bool mine = v && bin == lb;
        const uint64_t m = __ballot(mine);
        const uint32_t base = rl(run, lb);
        if (mine) pos = base + (uint32_t)__popcll(m & lanemask_lt());
        if (lane == lb) run += (uint32_t)__popcll(m);
        act &= ~m;
    }
    return pos;
}

// Where one node's outputs go, as a scatter wave derives it from the batch and the count rows
// alone: the node's outputs are entries base .. base + R - 1 of the step's pops, and (partitions)
// the items of bin b in segment gs start at flat slot bb_b + pre_b of the node's outputs.
struct Place {
    uint32_t k;    // node described (wave-uniform), or ~0u
    uint32_t ok;   // the node's batch agrees with its counts (otherwise wave 7 fails the step)
    uint32_t base, R;
    uint32_t gs;   // node-local segment `pre` belongs to
    uint32_t bb;   // lane b: binbase of bin b times C
    uint32_t pre;  // lane b: items of bin b in the node's segments before gs
    uint32_t fs0;  // LDS index of the step's pop 0 (the free stack's top entry when the step starts)
    uint32_t room; // pops the LDS window holds
};

// Output count of partition node k from its count rows.
__device__ inline uint32_t part_outputs(const TeamArgs& a, const TeamLds& L, const StepBuf& Bc, uint32_t k)
{
    const uint32_t B = L.nd[k].bins, C = a.e.C;
    uint32_t unused, tot, ts;
    seg_sums(a, Bc, k, B + 1u, 0u, 0u, Bc.nseg[k], 0u, unused, tot, ts);
    const uint32_t tv = lane_id() < B ? tot : 0u;
    return wave_sum(ceil_c(tv, C, a.c_shift));
}

// Placement of node k's segment gs (node-local numbering). A wave's segments come in increasing
// order, so a later segment of the same partition only adds the rows in between to the prefix.
__device__ inline void place_node(const TeamArgs& a, const TeamLds& L, const StepBuf& Bc, uint32_t k, uint32_t kind, uint32_t gs,
                                  Place& pl)
{
    const uint32_t C = a.e.C, lane = lane_id();
    if (pl.k == k) {
        if (kind != KIND_PARTITION || gs == pl.gs) return;
        if (gs > pl.gs) {
            uint32_t add, unused, us;
            seg_sums(a, Bc, k, L.nd[k].bins + 1u, pl.gs, gs, gs, 0u, add, unused, us);
            pl.pre += add;
            pl.gs = gs;
            return;
        }
    }
    pl.k = k;
    pl.gs = gs;
    pl.room = Bc.top0 - L.fs_base;
    pl.fs0 = pl.room - 1u;
    uint32_t base = Bc.ob[k];
    for (uint32_t pmk = Bc.part_mask & ((1u << k) - 1u); pmk; pmk &= pmk - 1u)
        base += part_outputs(a, L, Bc, (uint32_t)(__ffs(pmk) - 1));
    pl.base = base;
    pl.ok = 1u;
    pl.bb = 0u;
    pl.pre = 0u;
    const uint32_t parts = Bc.particles[k];
    if (kind == KIND_SOURCE) {
        pl.R = 1u;
    } else if (kind == KIND_TRANSFORM) {
        pl.R = ceil_c(parts, C, a.c_shift);
    } else {
        const uint32_t B = L.nd[k].bins;
        uint32_t pre, tot, sum;  // sum: items over all slots, the valid bins' sum when slot B is 0
        seg_sums(a, Bc, k, B + 1u, 0u, gs, Bc.nseg[k], 0u, pre, tot, sum);
        const uint32_t tv = lane < B ? tot : 0u, r = ceil_c(tv, C, a.c_shift), ri = wave_inscan(r);
        const uint32_t Rk = rl(ri, 63), inv = rl(tot, B);
        pl.R = Rk;
        pl.ok = (inv == 0u && sum == parts && Rk <= L.bound[k]) ? 1u : 0u;
        pl.bb = lane < B ? (ri - r) * C : 0u;
        pl.pre = pre;
    }
}

// Output chunk of pop j of the step (entry top0 - 1 - j of the free stack), or false when the
// entry is not in the LDS window (only a failed plan gets there).
__device__ inline bool pop_desc(const TeamLds& L, const Place& pl, uint32_t j, gf_chunk_desc& od)
{
    if (j >= pl.room) return false;
    const unsigned long long ent = L.fs[pl.fs0 - j];
    od.index = (uint32_t)ent;
    od.generation = (uint32_t)(ent >> 32);
    od.fill = 0u;
    od.bin = 0u;
    return true;
}

template <class P>
__device__ void team_partition_seg(const TeamArgs& a, const TeamLds& L, const StepBuf& Bc, uint32_t k, uint32_t i, uint32_t gg,
                                   const SegRecs<P>& R, BinTab& t, const Place& pl)
{
    const Engine& e = a.e;
    const uint32_t C = e.C, lane = lane_id();
    if (t.k != k) load_bintab(L, k, t);
    const gf_node_desc nd = L.nd[k];
    const uint32_t B = nd.bins;
    const gf_chunk_desc d = Bc.in[Bc.in_off[k] + i];
    const uint32_t off0 = gg * kSeg, cnt = d.fill - off0 < kSeg ? d.fill - off0 : kSeg;
    const uint32_t fb = lane < B ? pl.bb + pl.pre : 0u;  // flat start of bin `lane` in this segment
    Ticker tk(const_cast<unsigned long long*>(L.prof_acc), L.prof_member && threadIdx.x == 0u);
    bool v[kSegRounds];
    uint32_t bin[kSegRounds], pos[kSegRounds], key[kSegRounds];
#pragma unroll
    for (uint32_t r = 0; r < kSegRounds; ++r) {
        v[r] = r * 64u + lane < cnt;
        bin[r] = 0u;
        if (v[r]) {
            bin[r] = classify(nd, R.rec[r]);
            if (bin[r] >= B) {
                seg_fail(a, L, GF_INVALID);
                v[r] = false;
                bin[r] = 0u;
            }
        }
    }
    uint32_t run = 0;
#pragma unroll
    for (uint32_t r = 0; r < kSegRounds; ++r) pos[r] = seg_rank(v[r], bin[r], run);
    tk(PROF_C_SEG_RANK);
#pragma unroll
    for (uint32_t r = 0; r < kSegRounds; ++r) {
        // the bin's constants for every lane at once
        const uint32_t flat = from_lane(fb, bin[r]) + pos[r];
        const uint32_t w0 = from_lane(t.w0, bin[r]), w1 = from_lane(t.w1, bin[r]);
        const float xf0 = __uint_as_float(from_lane(__float_as_uint(t.xf0), bin[r]));
        const float df0 = __uint_as_float(from_lane(__float_as_uint(t.df0), bin[r]));
        key[r] = ~0u;
        if (!v[r]) continue;
        const uint32_t q = div_c(flat, C, a.c_shift), off = flat - q * C;
        gf_chunk_desc od;
        if (q >= pl.R || !pop_desc(L, pl, pl.base + q, od)) {
            seg_fail(a, L, GF_PLAN_MISMATCH);
            continue;
        }
        P p = R.rec[r];
        const uint32_t u = update(nd, bin[r], p);
        if (u) seg_fail(a, L, u);
        const int32_t fz = (int32_t)((w0 >> 8) & 0xFFu) - 1;
        if (!transform_p((w0 >> 16) & 0xFFu, xf0, p) && fz >= 0) {
            RunParams rp = e.rp;
            rp.seed = L.seed[fz];
            transform(L.nd[fz], rp, L.purpose[fz], p);
        }
        p.store(e.cols, (size_t)od.index * C + off);
        const uint32_t dop = w0 >> 24;
        if (dop) {
            const uint32_t dbins = w1 & 0xFFFFu;
            uint32_t nb = classify_p(dop, dbins, w1 >> 16, df0, p);
            if (nb == kNeedDesc) nb = classify(L.nd[w0 & 0xFFu], p);
            if (nb >= dbins) nb = dbins;
            key[r] = (((uint32_t)od.index * 2u + (od.generation & 1u)) * a.tpc + (off >> kSegShift)) * a.bb + nb;
        }
    }
    tk(PROF_C_SEG_STORE);
    seg_agg_add(key, a.cnt);
    tk(PROF_C_SEG_AGG);
}

template <class P>
__device__ void team_transform_seg(const TeamArgs& a, const TeamLds& L, const StepBuf& Bc, uint32_t k, uint32_t i, uint32_t gg,
                                   const SegRecs<P>& R, const Place& pl)
{
    const Engine& e = a.e;
    const uint32_t C = e.C, lane = lane_id();
    const gf_node_desc nd = L.nd[k];
    const gf_chunk_desc d = Bc.in[Bc.in_off[k] + i];
    const uint32_t off0 = gg * kSeg, cnt = d.fill - off0 < kSeg ? d.fill - off0 : kSeg;
    const uint32_t pos0 = Bc.in_items[Bc.in_off[k] + i] + off0;
    const uint32_t dst = (uint32_t)L.route[L.route_off[k]], dkind = L.kind[dst], purpose = L.purpose[k];
    const gf_node_desc dnd = L.nd[dst];
    RunParams rp = e.rp;
    rp.seed = L.seed[k];  // the node's job seed
    uint32_t key[kSegRounds];
#pragma unroll
    for (uint32_t r = 0; r < kSegRounds; ++r) {
        const uint32_t it = r * 64u + lane;
        key[r] = ~0u;
        if (it < cnt) {
            const uint32_t g = pos0 + it, q = div_c(g, C, a.c_shift), off = g - q * C;
            gf_chunk_desc od;
            if (q < pl.R && pop_desc(L, pl, pl.base + q, od)) {
                P p = R.rec[r];
                transform(nd, rp, purpose, p);
                p.store(e.cols, (size_t)od.index * C + off);
                key[r] = next_key(a, dkind, dnd, od, off, p);
            } else {
                seg_fail(a, L, GF_PLAN_MISMATCH);
            }
        }
    }
    seg_agg_add(key, a.cnt);
}

template <class P>
__device__ void team_source_seg(const TeamArgs& a, const TeamLds& L, const StepBuf& Bc, uint32_t k, uint32_t gg, const Place& pl)
{
    const Engine& e = a.e;
    const uint32_t C = e.C, lane = lane_id();
    const gf_node_desc nd = L.nd[k];
    gf_chunk_desc od;
    if (!pop_desc(L, pl, pl.base, od)) {
        seg_fail(a, L, GF_PLAN_MISMATCH);
        return;
    }
    const uint32_t fill = Bc.particles[k], off0 = gg * kSeg, cnt = fill - off0 < kSeg ? fill - off0 : kSeg;
    const uint32_t dst = (uint32_t)L.route[L.route_off[k]], dkind = L.kind[dst], purpose = L.purpose[k];
    const gf_node_desc dnd = L.nd[dst];
    const unsigned long long id0 = Bc.first_id[k] + off0;
    RunParams rp = e.rp;
    rp.seed = L.seed[k];  // the node's job seed
    uint32_t key[kSegRounds];
#pragma unroll
    for (uint32_t r = 0; r < kSegRounds; ++r) {
        const uint32_t it = r * 64u + lane;
        key[r] = ~0u;
        if (it < cnt) {
            P p;
            source(nd, rp, purpose, id0 + it, p);
            p.store(e.cols, (size_t)od.index * C + off0 + it);
            key[r] = next_key(a, dkind, dnd, od, off0 + it, p);
        }
    }
    seg_agg_add(key, a.cnt);
}

template <class P>
__device__ void team_sink_seg(const TeamArgs& a, const TeamLds& L, const StepBuf& Bc, uint32_t k, uint32_t i, uint32_t gg,
                              const SegRecs<P>& R)
{
    const uint32_t lane = lane_id();
    const gf_node_desc nd = L.nd[k];
    const gf_chunk_desc d = Bc.in[Bc.in_off[k] + i];
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
            const uint64_t entry = Bc.digest_base[k] + i;
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

// Segments [lo, hi) of this member (the step's non-sink segments split evenly over the members):
// worker wave w takes lo + r, lo + r + kWorkers, ... with r = worker_rank(w).
__device__ inline void member_range(const StepBuf& Bc, uint32_t n, uint32_t me, uint32_t T, uint32_t* lo, uint32_t* hi)
{
    (void)n;
    const uint32_t G = Bc.g_main;
    // me < T and G <= kTeamSegs: the products fit 32 bits
    *lo = me * G / T;
    *hi = (me + 1u) * G / T;
}

// Sink segments of this member, taken by its sink wave (after it cleared its share of the count
// rows).
template <class P>
__device__ inline void team_sinks(const TeamArgs& a, const TeamLds& L, const StepBuf& Bc, uint32_t me, uint32_t T)
{
    const uint32_t G0 = Bc.g_main, Gs = Bc.seg_off[a.e.n_nodes] - G0;
    const uint32_t lo = G0 + me * Gs / T, hi = G0 + (me + 1u) * Gs / T;
    SegRecs<P> R;
    for (uint32_t s = lo; s < hi; ++s) {
        const SegLoc loc = locate(L, Bc, a.e.n_nodes, s);
        load_segment<P>(a, Bc, loc, s, R);
        team_sink_seg<P>(a, L, Bc, loc.k, loc.i, loc.gg, R);
    }
}

// Worker wave w: segments lo + r, lo + r + kWorkers, ... of this member (r = worker_rank(w)). A
// segment's records are requested before its count rows, so both loads are in flight together.
template <class P>
__device__ inline void team_segments(const TeamArgs& a, const TeamLds& L, const StepBuf& Bc, uint32_t me, uint32_t T, uint32_t w,
                                     BinTab& t)
{
    uint32_t lo, hi;
    member_range(Bc, a.e.n_nodes, me, T, &lo, &hi);
    Place pl;
    pl.k = ~0u;
    SegRecs<P> R;
    Ticker tk(const_cast<unsigned long long*>(L.prof_acc), L.prof_member && threadIdx.x == 0u);
    for (uint32_t s = lo + worker_rank(w); s < hi; s += kWorkers) {
        const SegLoc loc = locate(L, Bc, a.e.n_nodes, s);
        load_segment<P>(a, Bc, loc, s, R);
        tk(PROF_C_SEG_LOCATE);
        const uint32_t k = loc.k, kind = loc.kind;
        place_node(a, L, Bc, k, kind, s - Bc.seg_off[k], pl);
        tk(PROF_C_SEG_PLACE);
        if (!pl.ok) continue;
        if (kind == KIND_SOURCE) team_source_seg<P>(a, L, Bc, k, loc.gg, pl);
        else if (kind == KIND_PARTITION) team_partition_seg<P>(a, L, Bc, k, loc.i, loc.gg, R, t, pl);
        else team_transform_seg<P>(a, L, Bc, k, loc.i, loc.gg, R, pl);
    }
}

// ---- the persistent team kernel ---------------------------------------------------------------
// A copy of `p` the compiler cannot see through: loads of the arguments through it stay inside the
// step loop, where the scalar cache serves them, instead of holding every argument any wave role
// uses in scalar registers for the whole launch (which spills them into vector-register lanes).
typedef const TeamArgs __attribute__((address_space(4))) ConstTeamArgs;  // constant address space: scalar loads
__device__ inline const TeamArgs& fresh_args(const TeamArgs* ga)
{
    ConstTeamArgs* p = (ConstTeamArgs*)ga;
    asm volatile("" : "+s"(p));
    return *(const TeamArgs*)p;
}

// The launch's arguments are read from device memory (`ga`).
template <class P>
__global__ __launch_bounds__(kTeamThreads) void k_team(const TeamArgs* __restrict__ ga)
{
    const TeamArgs& a = *ga;
    extern __shared__ __align__(16) char smem[];
    TeamLds& L = *reinterpret_cast<TeamLds*>(smem);
    __shared__ uint32_t sh[4];
    team_join(a, sh);
    const uint32_t me = sh[0], T = sh[1];
    if (me == ~0u) return;
    const Engine& e = a.e;
    const uint32_t tid = threadIdx.x, w = tid >> 6;
    const bool prof = e.prof != nullptr && me == 0u && tid == 0u;
    unsigned long long* mp = a.mprof ? a.mprof + (size_t)me * kMemberProf : nullptr;
    unsigned long long mt = mp ? wall_clock64() : 0ull;
    auto mmark = [&](uint32_t slot) {
        if (mp && tid == 0u) {
            const unsigned long long now = wall_clock64();
            atomicAdd(&mp[slot], now - mt);
            mt = now;
        }
    };
    const bool wprof = e.prof != nullptr && me == 0u && lane_id() == 0u;
    unsigned long long tc = prof ? wall_clock64() : 0ull;
    auto mark = [&](uint32_t ph) {
        if (prof) {
            const unsigned long long now = wall_clock64();
            atomicAdd(&L.prof_acc[ph], now - tc);
            tc = now;
        }
    };
    if (tid == 0u) L.prof_member = (e.prof != nullptr && me == 0u) ? 1u : 0u;
    if (tid < PROF_WORDS) L.prof_acc[tid] = 0ull;
    if (tid < kTeamWaves) L.wave_acc[tid] = 0ull;
    team_load(a, L);
    if (L.exit_code == EXIT_SKIP) {
        if (me == 0u && tid == 0u) e.ctl[CTL_HANDOFF] = 1u;
        return;
    }
    SchedRegs sg;
    if (w == kSchedWave) {
        sched_load(a, L, sg);
        team_select(a, L, L.sb[0], a.max_steps == 0u, sg);
    }
    __syncthreads();
    uint32_t target = 0;
    unsigned long long cyc = 0;
    BinTab tab;
    tab.k = ~0u;
    for (uint32_t it = 0;; ++it) {
        const TeamArgs& a = fresh_args(ga);
        const Engine& e = a.e;
        StepBuf& Bc = L.sb[it & 1u];
        StepBuf& Bn = L.sb[(it + 1u) & 1u];  // the previous step's batches until select overwrites them
        if (it) {
            // the previous step's records and counts of every member are in memory
            if (tid == 0u) {
                unsigned long long word = 0ull;
                if (!team_wait(a, target, word)) {
                    set_status(&e.ctl[CTL_STATUS], GF_DEVICE_ERROR);
                    L.status = GF_DEVICE_ERROR;
                }
                L.err_prev = (uint32_t)(word >> (32u + 16u * ((it - 1u) & 1u))) & 0xFFFFu;
            }
            __syncthreads();
            mark(PROF_T_WAIT);
            mmark(MP_WAIT);
            if (L.status || L.err_prev) break;
        }
        if (L.exit_code != EXIT_NONE) break;
        if (Bc.recenter) team_recenter(a, L, me);
        mark(PROF_T_COUNTS);
        mmark(MP_COUNTPLAN);
        const unsigned long long ts = mp ? wall_clock64() : 0ull;
        const unsigned long long w0 = wprof ? wall_clock64() : 0ull;
        const unsigned long long w0m = mp ? wall_clock64() : 0ull;
        if (mp && tid == 0u) {
            uint32_t lo, hi;
            member_range(Bc, e.n_nodes, me, T, &lo, &hi);
            const uint32_t Gs = Bc.seg_off[e.n_nodes] - Bc.g_main;
            atomicAdd(&mp[MP_NSEG], (unsigned long long)(hi - lo));
            atomicAdd(&mp[MP_NSINK], (unsigned long long)((me + 1u) * Gs / T - me * Gs / T));
        }
        if (w < kWorkers) {
            team_segments<P>(a, L, Bc, me, T, w, tab);
            if (wprof && w == 0u) atomicAdd(&L.prof_acc[PROF_T_SEGW], wall_clock64() - w0);
        } else if (w == kSinkWave) {
            team_clear_rows(a, L, me, T);
            team_sinks<P>(a, L, Bc, me, T);
        } else {
            // the scheduler wave plans this step, retires it and selects the next one
            team_plan(a, L, Bc, sg);
            if (!L.status) {
                if (me == 0u && e.trace_on) team_trace(a, L, Bc);
                const unsigned long long w1 = wprof ? wall_clock64() : 0ull;
                team_retire(a, L, Bc, sg);
                const unsigned long long w2 = wprof ? wall_clock64() : 0ull;
                if (wprof) atomicAdd(&L.prof_acc[PROF_T_RETIRE], w2 - w1);
                team_select(a, L, Bn, it + 1u >= a.max_steps, sg);
                if (wprof) atomicAdd(&L.prof_acc[PROF_T_SELECTW], wall_clock64() - w2);
            }
            if (mp && lane_id() == 0u) atomicAdd(&mp[MP_SCHED], wall_clock64() - w0m);
        }
        if (mp && lane_id() == 0u) atomicAdd(&L.wave_acc[w], wall_clock64() - ts);
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        target += T;
        if (tid == 0u) {
            // arrive, with this member's item failures of the step in the step parity's field
            const unsigned long long add = 1ull + (L.seg_err ? 1ull << (32u + 16u * (it & 1u)) : 0ull);
            __hip_atomic_fetch_add((unsigned long long*)a.tw + kBarWord, add, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            L.seg_err = 0u;
        }
        mark(PROF_T_SEG);
        mmark(MP_SEG);
        if (L.status) break;
        if (w == kSchedWave) team_push(a, L, Bc);
        mark(PROF_T_SELECT);
        if (prof) {
            L.prof_acc[PROF_T_STEPS] += 1u;
            const unsigned long long c = __builtin_readcyclecounter();
            if (it) L.prof_acc[PROF_C_STEP] += c - cyc;
            cyc = c;
        }
    }
    if (w == kSchedWave) sched_store(a, L, sg);
    __syncthreads();
    if (mp && tid < kTeamWaves) mp[MP_WAVE + tid] = L.wave_acc[tid];
    if (me == 0u && !L.status && !L.err_prev) team_flush(a, L);
    if (e.prof != nullptr && me == 0u) {
        __syncthreads();
        if (tid < PROF_WORDS) e.prof[tid] += L.prof_acc[tid];
    }
}

}  // namespace gfd
