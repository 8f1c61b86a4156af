// Device-resident step engine of the batched form.
//
// One step = four kernels, no host work in between:
//   KA: classify+count every partition tile of the step;
//   KP: one workgroup plans every node in node order: per-tile flat offsets, exact output chunk
//       counts, pops from the device free stack, output descriptors, trace;
//   KB: scatter / transform / source / sink tiles of every active node;
//   KR: one workgroup retires the step (routes outputs to the destination ready rings, releases
//       inputs to the free stack) and selects the next step's batches (source admission against
//       P, largest credit-fitting prefix of <= Q ready chunks per transform node, all ready chunks
//       per sink).
// The scheduling rules are those of GraphRunner::run (sched/GraphRunner.hpp:902-1133) applied
// once per lane per step; the host only replays a hipGraph of K steps and polls a done flag.
//
// Lanes (Tuning::inflightBatches, flow/Tuning.hpp:34-41; GraphRunner.hpp:766-771): every source
// and transform node runs up to D batches per step. The engine's nodes are the lanes, in (node,
// lane) order; the lanes of a node share its ready ring, which belongs to its first lane
// (ring_of), and routes name that first lane. A sink keeps one lane: it takes all ready chunks.
//
// The one-workgroup phases are latency bound: they keep the whole control state in LDS and
// touch global memory in a constant number of parallel round trips per step.
#pragma once

#include "gf_device.hpp"

namespace gfd {

constexpr uint32_t kMaxNodes = 128;  // engine nodes (lanes)
constexpr uint32_t kListCap = 2048;   // chunk descriptors staged in LDS by the select phase
constexpr uint32_t kPlanWords = 8192; // [tile][bin] words of one partition planned in LDS

// Dynamic per-node state (64 bytes), rewritten by the last workgroups.
struct NodeSt {
    uint32_t active, n_in, n_tiles, particles;
    uint32_t n_out, in_slot, pad1, pad2;  // in_slot: ring slot of the lane's first input
    unsigned long long first_id, digest_base, head, tail;
};

// Static per-node arrays and shape.
struct NodeArr {
    gf_chunk_desc* in;
    uint32_t* tile_prefix;
    uint32_t* item_prefix;
    gf_chunk_desc* out;
    uint32_t* counts;
    uint32_t* totals;
    uint32_t tile_items, purpose, cap_in, cap_out, reserve_bound, kind;  // kind: 0 src, 1 part, 2 xf, 3 sink
    uint32_t flags;  // NODE_STAGE: scatter stages its tile in LDS; NODE_WAVE_SINK: one tile per wave
    uint32_t real;   // the graph node this lane belongs to (trace records, stream purpose = 2 * real)
    uint32_t lanes;  // first lane of a node: the node's lanes (it stages candidates for all of them)
    uint32_t ring;   // first lane of a node: index of the node's ready ring in `rings`
    uint32_t hash_entries;  // partitions on the hash bin path: slots of a tile's LDS bin table; 0: dense path
    uint32_t job;           // served engines: the job slot whose copy of the graph the lane belongs to
};

enum { NODE_STAGE = 1u, NODE_WAVE_SINK = 2u };

enum { KIND_SOURCE = 0, KIND_PARTITION = 1, KIND_TRANSFORM = 2, KIND_SINK = 3 };

// One job slot of a served engine: the engine runs the union of M copies of a topology's graph, one
// per slot, and the host admits a job into a free slot between launches (items, live bound, seed;
// admitted and sunk items start at 0). The select keeps admitted and sunk up to date; a job is done
// when sunk == N.
struct JobSlot {
    unsigned long long N, P, seed, gen, sunk;
};
constexpr uint32_t kMaxJobs = 64;

struct Engine {
    uint32_t n_nodes, C, P, Q, S, ring_cap, trace_on, n_jobs;  // n_jobs: slots of a served engine, else 0
    uint32_t src_chunks, pad0;  // chunks a source lane admits per step
    uint64_t N;
    RunParams rp;
    Cols cols;
    const gf_node_desc* nodes;       // per lane (a copy of its node's descriptor)
    const gf_node_desc* real_nodes;  // per graph node (fused transforms are looked up here)
    const int32_t* routes;  // effective route per (node, bin) to the destination's first lane, fused transforms bypassed
    const int32_t* fuse;    // per (node, bin): graph node of the transform applied inside the scatter, or -1
    const uint32_t* ring_of;  // per lane: the lane whose ring holds its node's ready chunks
    const NodeArr* arr;
    NodeSt* st;
    unsigned long long* free_stack;  // entry = index | generation << 32
    gf_chunk_desc* rings;            // n_nodes * ring_cap
    uint32_t* ka_prefix;             // n_nodes + 1
    uint32_t* kb_prefix;             // n_nodes + 1
    uint32_t* ctl;                   // control words, CTL_*
    unsigned long long* ctr;         // counters, CNT_*
    uint32_t* trace;
    unsigned long long trace_cap;
    unsigned long long* prof;  // optional per-phase wall-clock totals (micro engine), PROF_*
    unsigned long long* lb;    // look-back words of the grid plan, one per lane
    unsigned long long* rlb;   // look-back words of the grid routing, [unit][lane]
    unsigned long long* llb;   // look-back words of the grid input lists, [unit][2] (tiles, items)
    uint32_t route_units_cap, list_units_cap;
    JobSlot* jobs;             // served engines: n_jobs slots; null otherwise (one job: N, P, the run's seed)
};

enum { PROF_COUNT = 0, PROF_PLAN = 1, PROF_SOURCE = 2, PROF_PART = 3, PROF_XF = 4, PROF_SINK = 5, PROF_RETIRE = 6,
       PROF_SELECT = 7, PROF_STEPS = 8,
       // the select's phases (staging, decisions, input lists, prefixes and termination) and the
       // retire counters before it
       PROF_SEL_STAGE = 9, PROF_SEL_DECIDE = 10, PROF_SEL_LISTS = 11, PROF_SEL_END = 12, PROF_SEL_RETIRE = 13,
       // wide engine scatter kernel: workgroup wall-clock ticks and tiles per KIND_*, then the
       // one-wave sink tiles (ticks summed over waves)
       PROF_KB_TICKS = 16, PROF_KB_TILES = 22, PROF_KB_WAVE_TICKS = 28, PROF_KB_WAVE_TILES = 29,
       PROF_WORDS = 32 };

enum { CTL_STATUS = 0, CTL_DONE = 1, CTL_TOP = 2, CTL_TICKET_A = 3, CTL_TICKET_B = 4, CTL_HANDOFF = 5,
       CTL_TICKET_C = 6, CTL_LIST_DONE = 7, CTL_WORDS = 8 };
enum {
    CNT_GENERATED = 0, CNT_SUNK = 1, CNT_STEP = 2, CNT_BATCHES = 3, CNT_PART_ITEMS = 4, CNT_XF_ITEMS = 5,
    CNT_SRC_ITEMS = 6, CNT_SINK_ITEMS = 7, CNT_TRACE_LEN = 8, CNT_PEAK_LIVE = 9, CNT_FUSED_ITEMS = 10,
    CNT_DIGEST0 = 16, CNT_WORDS = 16 + kMaxNodes
};

constexpr uint32_t kMaxWaves = 16;  // last-workgroup phases run with 256 or 1024 threads
__device__ inline uint32_t bsz() { return blockDim.x; }
__device__ inline uint32_t bwaves() { return blockDim.x >> 6; }

__device__ inline bool is_partition_op(uint32_t op) { return op >= GF_CLS_WEIGHT_ROUTE && op < GF_XF_SCALE_WEIGHT; }
__device__ inline uint32_t load_ctl(uint32_t* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

// Block-wide exclusive scan of one value per thread; returns the exclusive prefix, *total gets the sum.
__device__ inline uint32_t block_exscan(uint32_t v, uint32_t* sh, uint32_t* total)
{
    const uint32_t tid = threadIdx.x, lane = lane_id(), w = tid >> 6;
    uint32_t x = v;
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = (uint32_t)__shfl_up((int)x, o);
        if (lane >= (uint32_t)o) x += y;
    }
    if (lane == 63) sh[w] = x;
    __syncthreads();
    uint32_t off = 0, tot = 0;
    for (uint32_t k = 0; k < bwaves(); ++k) {
        if (k < w) off += sh[k];
        tot += sh[k];
    }
    __syncthreads();
    *total = tot;
    return off + x - v;
}

// Exclusive scan of n values a[0..n) in place (one workgroup, any memory), returns the total.
__device__ inline uint32_t block_exscan_array(uint32_t* a, uint32_t n, uint32_t* sh)
{
    const uint32_t tid = threadIdx.x;
    const uint32_t per = (n + bsz() - 1u) / bsz();
    const uint32_t lo = tid * per, hi = lo + per < n ? lo + per : n;
    uint32_t s = 0;
    for (uint32_t i = lo; i < hi; ++i) s += a[i];
    uint32_t total;
    uint32_t run = block_exscan(s, sh, &total);
    for (uint32_t i = lo; i < hi; ++i) {
        uint32_t x = a[i];
        a[i] = run;
        run += x;
    }
    __syncthreads();
    return total;
}

// Wave-wide inclusive scan on the DPP network (no LDS round trip): Hillis-Steele inside each
// 16-lane row with row_shr:1,2,4,8, then row_bcast:15 into rows 1 and 3 and row_bcast:31 into
// rows 2 and 3. Lanes whose DPP source is outside the row read 0.
__device__ inline uint32_t wave_inscan(uint32_t x)
{
    int v = (int)x;
    v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xF, 0xF, false);  // row_shr:1
    v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xF, 0xF, false);  // row_shr:2
    v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xF, 0xF, false);  // row_shr:4
    v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xF, 0xF, false);  // row_shr:8
    v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xA, 0xF, false);  // row_bcast:15 -> rows 1, 3
    v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xC, 0xF, false);  // row_bcast:31 -> rows 2, 3
    return (uint32_t)v;
}

__device__ inline Batch make_batch(const Engine& e, uint32_t k, const NodeSt& s)
{
    const NodeArr& a = e.arr[k];
    Batch b;
    b.active = s.active;
    b.n_in = s.n_in;
    b.n_tiles = s.n_tiles;
    b.n_out = s.n_out;
    b.particles = s.particles;
    b.reserved = 0;
    b.tile_items = a.tile_items;
    b.purpose = a.purpose;
    b.first_id = s.first_id;
    b.digest_base = s.digest_base;
    b.in = a.in;
    b.tile_prefix = a.tile_prefix;
    b.item_prefix = a.item_prefix;
    b.out = a.out;
    b.counts = a.counts;
    b.totals = a.totals;
    return b;
}

__device__ inline void load_states(const Engine& e, NodeSt* sh)
{
    const uint32_t words = e.n_nodes * (uint32_t)(sizeof(NodeSt) / 4u);
    const uint32_t* src = (const uint32_t*)e.st;
    uint32_t* dst = (uint32_t*)sh;
    for (uint32_t i = threadIdx.x; i < words; i += bsz()) dst[i] = src[i];
}

__device__ inline void store_states(const Engine& e, const NodeSt* sh)
{
    const uint32_t words = e.n_nodes * (uint32_t)(sizeof(NodeSt) / 4u);
    uint32_t* dst = (uint32_t*)e.st;
    const uint32_t* src = (const uint32_t*)sh;
    for (uint32_t i = threadIdx.x; i < words; i += bsz()) dst[i] = src[i];
}

// Trace record (layout identical to the oracle): 0xB47C0000|node, step, n_in, (index,gen,fill)*n_in,
// nbins, counts, R, (index,gen,fill,bin)*R.
__device__ inline void trace_record(const Engine& e, uint32_t k, const Batch& b, uint32_t nbins,
                                    const uint32_t* counts, unsigned long long* len_sh, uint32_t step)
{
    const uint32_t tid = threadIdx.x;
    unsigned long long len = (unsigned long long)5u + 3ull * b.n_in + nbins + 4ull * b.n_out;
    __syncthreads();
    if (tid == 0) {
        *len_sh = e.ctr[CNT_TRACE_LEN];
        e.ctr[CNT_TRACE_LEN] = *len_sh + len;
    }
    __syncthreads();
    unsigned long long p0 = *len_sh;
    auto put = [&](unsigned long long p, uint32_t v) {
        if (p0 + p < e.trace_cap) e.trace[p0 + p] = v;
    };
    if (tid == 0) {
        put(0, 0xB47C0000u | e.arr[k].real);
        put(1, step);
        put(2, b.n_in);
        put(3 + 3ull * b.n_in, nbins);
        put(4 + 3ull * b.n_in + nbins, b.n_out);
    }
    for (uint32_t i = tid; i < b.n_in; i += bsz()) {
        put(3 + 3ull * i, b.in[i].index);
        put(4 + 3ull * i, b.in[i].generation);
        put(5 + 3ull * i, b.in[i].fill);
    }
    for (uint32_t i = tid; i < nbins; i += bsz()) put(4 + 3ull * b.n_in + i, counts[i]);
    unsigned long long o0 = 5 + 3ull * b.n_in + nbins;
    for (uint32_t i = tid; i < b.n_out; i += bsz()) {
        put(o0 + 4ull * i, b.out[i].index);
        put(o0 + 4ull * i + 1, b.out[i].generation);
        put(o0 + 4ull * i + 2, b.out[i].fill);
        put(o0 + 4ull * i + 3, b.out[i].bin);
    }
    __syncthreads();
}

// ------------------------------------------------------------------------------------------
// KA last workgroup: plan
// ------------------------------------------------------------------------------------------
struct PlanLds {
    NodeSt st[kMaxNodes];
    uint32_t R[kMaxNodes];
    uint32_t off[kMaxNodes];
    uint32_t tb_off[kMaxNodes];  // offset of node k's [tot(B), base(B)] in `words`, or ~0u when in global
    uint32_t scan[kMaxWaves];
    uint32_t top, status, fail, total_R, words_used;
    unsigned long long len;
    unsigned long long digest[kMaxNodes];
    uint32_t words[kPlanWords + 1024];
};

// Per-(tile, bin) exclusive prefix over tiles, totals and bin bases of one partition, in LDS when
// the [tile][bin] matrix fits, else in global memory.
__device__ inline void plan_partition_counts(const Engine& e, uint32_t k, PlanLds& L)
{
    const gf_node_desc& nd = e.nodes[k];
    const NodeArr& a = e.arr[k];
    const uint32_t B = nd.bins, nt = L.st[k].n_tiles, tid = threadIdx.x, C = e.C;
    const uint32_t w = tid >> 6, lane = lane_id();
    const uint32_t need = nt * B;
    uint32_t* tot;
    uint32_t* base;
    if (need <= kPlanWords && L.words_used + 2u * B <= 1024u) {
        uint32_t* m = L.words;  // [tile][bin]
        for (uint32_t i = tid; i < need; i += bsz()) m[i] = a.counts[i];
        tot = L.words + kPlanWords + L.words_used;
        base = tot + B;
        __syncthreads();
        // one wave per bin, 64 tiles per iteration with a carried prefix
        for (uint32_t bin = w; bin < B; bin += bwaves()) {
            uint32_t carry = 0;
            for (uint32_t t0 = 0; t0 < nt; t0 += 64u) {
                uint32_t t = t0 + lane;
                uint32_t x = t < nt ? m[t * B + bin] : 0u;
                uint32_t inc = wave_inscan(x);
                if (t < nt) m[t * B + bin] = carry + inc - x;
                carry += (uint32_t)__shfl((int)inc, 63);
            }
            if (lane == 0) tot[bin] = carry;
        }
        __syncthreads();
        if (tid == 0) {
            L.tb_off[k] = kPlanWords + L.words_used;
            L.words_used += 2u * B;
        }
    } else {
        // thread per bin (consecutive bins: coalesced), 16 tiles loaded before any is rewritten
        tot = a.totals;
        base = a.totals + B;
        uint32_t* cn = a.counts;
        for (uint32_t bin = tid; bin < B; bin += bsz()) {
            uint32_t acc = 0;
            for (uint32_t t0 = 0; t0 < nt; t0 += 16u) {
                uint32_t x[16];
#pragma unroll
                for (uint32_t j = 0; j < 16u; ++j) x[j] = t0 + j < nt ? cn[(size_t)(t0 + j) * B + bin] : 0u;
#pragma unroll
                for (uint32_t j = 0; j < 16u; ++j) {
                    if (t0 + j < nt) cn[(size_t)(t0 + j) * B + bin] = acc;
                    acc += x[j];
                }
            }
            tot[bin] = acc;
        }
        if (tid == 0) L.tb_off[k] = ~0u;
        __syncthreads();
    }
    uint32_t s = 0;
    for (uint32_t bin = tid; bin < B; bin += bsz()) {
        s += tot[bin];
        base[bin] = (tot[bin] + C - 1u) / C;
    }
    uint32_t sum;
    block_exscan(s, L.scan, &sum);
    uint32_t R = block_exscan_array(base, B, L.scan);
    if (tid == 0) {
        L.R[k] = R;
        if (sum != L.st[k].particles || R > a.reserve_bound) {
            set_status(&e.ctl[CTL_STATUS], GF_PLAN_MISMATCH);
            L.fail = 1u;
        }
    }
    // every (tile, bin)'s items of the bin in earlier tiles, and beside the totals (for the trace)
    // each bin's first output chunk: a scatter tile adds that chunk's flat start itself, so the
    // [tile][bin] matrix is not rewritten a second time
    if (need <= kPlanWords && L.tb_off[k] != ~0u) {
        const uint32_t* m = L.words;
        for (uint32_t i = tid; i < need; i += bsz()) a.counts[i] = m[i];
        for (uint32_t bin = tid; bin < B; bin += bsz()) {
            a.totals[bin] = tot[bin];
            a.totals[B + bin] = base[bin];
        }
    }
    __syncthreads();
}

__device__ inline void engine_plan(const Engine& e, PlanLds& L)
{
    const uint32_t tid = threadIdx.x, n = e.n_nodes, C = e.C;
    load_states(e, L.st);
    if (tid == 0) {
        L.top = e.ctl[CTL_TOP];
        L.status = load_ctl(&e.ctl[CTL_STATUS]);
        L.fail = 0u;
        L.words_used = 0u;
    }
    for (uint32_t k = tid; k < n; k += bsz()) L.digest[k] = e.ctr[CNT_DIGEST0 + k];
    __syncthreads();
    if (L.status) return;
    // pass 1: exact output chunk count of every node, in node order
    for (uint32_t k = 0; k < n; ++k) {
        if (!L.st[k].active) {
            if (tid == 0) L.R[k] = 0u;
            continue;
        }
        const uint32_t kind = e.arr[k].kind;
        if (kind == KIND_PARTITION) {
            plan_partition_counts(e, k, L);
        } else if (tid == 0) {
            if (kind == KIND_SOURCE) {
                L.R[k] = (L.st[k].particles + C - 1u) / C;  // the admitted chunks
            } else if (kind == KIND_SINK) {
                L.R[k] = 0u;
                L.st[k].digest_base = L.digest[k];
                L.digest[k] += L.st[k].n_in;
            } else {
                uint32_t R = (L.st[k].particles + C - 1u) / C;
                L.R[k] = R;
                if (L.st[k].particles == 0u || R > e.arr[k].reserve_bound) {
                    set_status(&e.ctl[CTL_STATUS], GF_PLAN_MISMATCH);
                    L.fail = 1u;
                }
            }
        }
        __syncthreads();
    }
    if (tid == 0) {
        uint32_t acc = 0;
        for (uint32_t k = 0; k < n; ++k) {
            L.off[k] = acc;
            acc += L.R[k];
        }
        L.total_R = acc;
        if (!L.fail && acc > L.top) {
            set_status(&e.ctl[CTL_STATUS], GF_POOL_EXHAUSTED);
            L.fail = 1u;
        }
    }
    __syncthreads();
    if (L.fail) return;
    // pass 2: every pop of the step (node order, bin-major, k order): output j takes
    // stack[top - 1 - j]; a partition's outputs are written per bin (thread per bin), the others
    // per output
    const uint32_t top = L.top;
    for (uint32_t k = 0; k < n; ++k) {
        if (!L.st[k].active || !L.R[k]) continue;
        const uint32_t kind = e.arr[k].kind, o = L.off[k];
        gf_chunk_desc* out = e.arr[k].out;
        if (kind == KIND_PARTITION) {
            const uint32_t B = e.nodes[k].bins;
            const uint32_t* tot = L.tb_off[k] != ~0u ? L.words + L.tb_off[k] : e.arr[k].totals;
            const uint32_t* base = tot + B;
            for (uint32_t bin = tid; bin < B; bin += bsz()) {
                const uint32_t tb = tot[bin], b0 = base[bin];
                for (uint32_t q = 0, rem = tb; rem; ++q) {
                    const unsigned long long ent = e.free_stack[top - 1u - (o + b0 + q)];
                    gf_chunk_desc d;
                    d.index = (uint32_t)ent;
                    d.generation = (uint32_t)(ent >> 32);
                    d.fill = rem < C ? rem : C;
                    d.bin = bin;
                    out[b0 + q] = d;
                    rem -= d.fill;
                }
            }
        } else {
            for (uint32_t r = tid; r < L.R[k]; r += bsz()) {
                const unsigned long long ent = e.free_stack[top - 1u - (o + r)];
                gf_chunk_desc d;
                d.index = (uint32_t)ent;
                d.generation = (uint32_t)(ent >> 32);
                const uint32_t rem = L.st[k].particles - r * C;  // source and transform: C, ..., C, rest
                d.fill = rem < C ? rem : C;
                d.bin = 0u;
                out[r] = d;
            }
        }
    }
    for (uint32_t k = tid; k < n; k += bsz()) {
        L.st[k].n_out = L.R[k];
        if (e.arr[k].kind == KIND_TRANSFORM) e.arr[k].totals[0] = L.st[k].particles;
        e.ctr[CNT_DIGEST0 + k] = L.digest[k];
    }
    __syncthreads();
    store_states(e, L.st);
    if (tid == 0) {
        uint32_t nt = top - L.total_R;
        e.ctl[CTL_TOP] = nt;
        unsigned long long live = e.S - nt;
        if (live > e.ctr[CNT_PEAK_LIVE]) e.ctr[CNT_PEAK_LIVE] = live;
    }
    if (e.trace_on) {
        __syncthreads();
        const uint32_t step = (uint32_t)e.ctr[CNT_STEP];
        for (uint32_t k = 0; k < n; ++k) {
            if (!L.st[k].active) continue;
            const uint32_t kind = e.arr[k].kind;
            Batch b = make_batch(e, k, L.st[k]);
            if (kind == KIND_PARTITION) trace_record(e, k, b, e.nodes[k].bins, e.arr[k].totals, &L.len, step);
            else if (kind == KIND_TRANSFORM) trace_record(e, k, b, 1u, e.arr[k].totals, &L.len, step);
            else trace_record(e, k, b, 0u, nullptr, &L.len, step);
        }
    }
    __syncthreads();
}

// ------------------------------------------------------------------------------------------
// KB last workgroup: retire + select
// ------------------------------------------------------------------------------------------
struct SelLds {
    NodeSt st[kMaxNodes];
    unsigned long long tail[kMaxNodes];
    unsigned long long head[kMaxNodes];
    uint32_t ring_at[kMaxNodes];  // ring slot of node k's tail (retire) / head (select): cursor % ring_cap
    uint32_t off[kMaxNodes + 1];
    uint32_t lst_off[kMaxNodes + 1];  // staged ready descriptors of node k: lst[lst_off[k] ..)
    uint32_t wanted[kMaxNodes];
    uint32_t taken[kMaxNodes];  // first lane of a node: ready chunks its lanes took this step
    uint32_t in_at[kMaxNodes];  // lane: position of its first input in its node's ready chunks
    uint32_t kind[kMaxNodes], ring_of[kMaxNodes], lanes[kMaxNodes], bins[kMaxNodes], job[kMaxNodes];  // load_lane_tab
    unsigned long long jN[kMaxJobs], jP[kMaxJobs], jgen[kMaxJobs], jsunk[kMaxJobs];  // load_job_tab
    uint32_t njobs;
    uint32_t wcnt[kMaxWaves][kMaxNodes];
    uint32_t scan[kMaxWaves];
    uint32_t top, total, any, staged, overflow;
    unsigned long long generated, sunk, step, batches, part_items, xf_items, src_items, sink_items;
    gf_chunk_desc lst[kListCap];
    uint32_t fpre[kListCap + 1];  // exclusive prefix of the staged descriptors' fills
    uint32_t dfill[kMaxNodes], dtiles[kMaxNodes];  // source lanes: admitted items and tiles, first id
    unsigned long long dfirst[kMaxNodes];
    unsigned long long red[kMaxWaves * 8];  // block_sum64
};
// the one-workgroup phases run a thread per lane: every block has at least kMaxNodes threads
static_assert(kMaxNodes <= (uint32_t)kThreads, "a block must hold a thread per lane");

// Block-wide sums of V 64-bit values per thread; every thread gets the sums (sh: kMaxWaves * V words).
template <int V>
__device__ inline void block_sum64(unsigned long long (&v)[V], unsigned long long* sh)
{
    const uint32_t lane = lane_id(), w = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < V; ++i) {
        for (int o = 32; o > 0; o >>= 1) {
            const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)v[i], o);
            const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(v[i] >> 32), o);
            v[i] += ((unsigned long long)hi << 32) | lo;
        }
        if (lane == 0) sh[w * V + i] = v[i];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < V; ++i) {
        unsigned long long t = 0ull;
        for (uint32_t ww = 0; ww < bwaves(); ++ww) t += sh[ww * V + i];
        v[i] = t;
    }
    __syncthreads();
}

// Lane owning flat item j: the k with prefix[k] <= j < prefix[k+1] (prefix has n+1 entries).
__device__ inline uint32_t lane_of(const uint32_t* prefix, uint32_t n, uint32_t j)
{
    uint32_t lo = 0, hi = n;  // first index in [0, n] with prefix > j, minus one
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (prefix[mid] <= j) lo = mid + 1u;
        else hi = mid;
    }
    return lo - 1u;
}

__device__ inline uint64_t bound_of(uint64_t n, uint64_t C, uint64_t B)
{
    if (n == 0u) return 0u;
    // 32-bit division whenever the operands fit (a batch's items, the chunk capacity)
    uint64_t per = n + C - 1u <= 0xFFFFFFFFull ? (uint64_t)((uint32_t)(n + C - 1u) / (uint32_t)C) : (n + C - 1u) / C;
    uint64_t bins = B < n ? B : n;
    uint64_t r = per + bins - 1u;
    return r < n ? r : n;
}

constexpr uint32_t kAhead = 4;  // block-sized rounds of descriptor loads issued together

// The static per-lane fields the one-workgroup phases consult, staged in LDS by one parallel load
// (a thread per lane): their walks over the lanes in node order then read LDS only, instead of
// waiting on one global round trip per lane.
__device__ inline void load_lane_tab(const Engine& e, SelLds& L)
{
    for (uint32_t k = threadIdx.x; k < e.n_nodes; k += bsz()) {
        L.kind[k] = e.arr[k].kind;
        L.ring_of[k] = e.ring_of[k];
        L.lanes[k] = e.arr[k].lanes;
        L.bins[k] = e.nodes[k].bins;
        L.job[k] = e.arr[k].job;
    }
}

// The jobs' items, live bounds and admitted / sunk counts into LDS (a thread per job). An engine
// that is not served runs one job: the run's N and P, the engine's admitted and sunk totals
// (L.generated and L.sunk must be current).
__device__ inline void load_job_tab(const Engine& e, SelLds& L)
{
    const uint32_t tid = threadIdx.x;
    if (e.jobs) {
        if (tid < e.n_jobs) {
            const JobSlot& j = e.jobs[tid];
            L.jN[tid] = j.N;
            L.jP[tid] = j.P;
            L.jgen[tid] = j.gen;
            L.jsunk[tid] = j.sunk;
        }
        if (tid == 0) L.njobs = e.n_jobs;
    } else if (tid == 0) {
        L.jN[0] = e.N;
        L.jP[0] = e.P;
        L.jgen[0] = L.generated;
        L.jsunk[0] = L.sunk;
        L.njobs = 1u;
    }
}

__device__ inline void engine_retire(const Engine& e, SelLds& L)
{
    const uint32_t tid = threadIdx.x, n = e.n_nodes, w = tid >> 6;
    load_states(e, L.st);
    load_lane_tab(e, L);
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
    }
    __syncthreads();
    for (uint32_t k = tid; k < n; k += bsz()) {
        L.tail[k] = L.st[k].tail;
        L.head[k] = L.st[k].head;
        L.ring_at[k] = (uint32_t)(L.st[k].tail % e.ring_cap);  // the one 64-bit remainder per node
    }
    // outputs to route: node order, output order
    if (tid == 0) {
        uint32_t acc = 0;
        for (uint32_t k = 0; k < n; ++k) {
            L.off[k] = acc;
            if (L.st[k].active && L.kind[k] != KIND_SINK) acc += L.st[k].n_out;
        }
        L.off[n] = acc;
        L.total = acc;
    }
    __syncthreads();
    // route one block-sized round at a time with an order-preserving per-destination rank; the
    // descriptors and destinations of kAhead rounds are loaded together
    for (uint32_t g0 = 0; g0 < L.total; g0 += kAhead * bsz()) {
        gf_chunk_desc dd[kAhead];
        uint32_t dk[kAhead], ds[kAhead];
#pragma unroll
        for (uint32_t u = 0; u < kAhead; ++u) {
            const uint32_t j = g0 + u * bsz() + tid;
            dd[u] = gf_chunk_desc{0, 0, 0, 0};
            dk[u] = 0u;
            if (j < L.total) {
                uint32_t k = 0;
                while (k + 1u < n && L.off[k + 1u] <= j) ++k;
                dk[u] = k;
                dd[u] = e.arr[k].out[j - L.off[k]];
            }
        }
#pragma unroll
        for (uint32_t u = 0; u < kAhead; ++u)
            ds[u] = g0 + u * bsz() + tid < L.total ? (uint32_t)e.routes[e.nodes[dk[u]].route_offset + dd[u].bin] : 0xFFFFFFFFu;
#pragma unroll
        for (uint32_t u = 0; u < kAhead; ++u) {
            if (g0 + u * bsz() >= L.total) break;
            const bool v = g0 + u * bsz() + tid < L.total;
            const gf_chunk_desc d = dd[u];
            const uint32_t dst = ds[u];
            for (uint32_t i = tid; i < kMaxWaves * kMaxNodes; i += bsz()) (&L.wcnt[0][0])[i] = 0u;
            __syncthreads();
            uint64_t peers = peer_mask(v, dst);
            const uint32_t rank = (uint32_t)__popcll(peers & lanemask_lt());
            if (v && (peers & lanemask_lt()) == 0ull) L.wcnt[w][dst] = (uint32_t)__popcll(peers);
            __syncthreads();
            if (v) {
                uint32_t pre = 0;
                for (uint32_t ww = 0; ww < w; ++ww) pre += L.wcnt[ww][dst];
                // slot of tail + pre + rank: a ring never holds more than ring_cap (= S) chunks, so
                // pre + rank < ring_cap and one subtraction wraps it
                uint32_t slot = L.ring_at[dst] + pre + rank;
                if (slot >= e.ring_cap) slot -= e.ring_cap;
                e.rings[(size_t)e.arr[dst].ring * e.ring_cap + slot] = d;
            }
            __syncthreads();
            for (uint32_t k = tid; k < n; k += bsz()) {
                uint32_t sum = 0;
                for (uint32_t ww = 0; ww < bwaves(); ++ww) sum += L.wcnt[ww][k];
                L.tail[k] += sum;
                uint32_t at = L.ring_at[k] + sum;
                if (at >= e.ring_cap) at -= e.ring_cap;
                L.ring_at[k] = at;
            }
            __syncthreads();
        }
    }
    // release every consumed input to the free stack (generation + 1)
    if (tid == 0) {
        uint32_t acc = 0;
        for (uint32_t k = 0; k < n; ++k) {
            L.off[k] = acc;
            if (L.st[k].active && L.kind[k] != KIND_SOURCE) acc += L.st[k].n_in;
        }
        L.off[n] = acc;
        L.total = acc;
    }
    __syncthreads();
    for (uint32_t g0 = tid; g0 < L.total; g0 += kAhead * bsz()) {
        gf_chunk_desc dd[kAhead];
#pragma unroll
        for (uint32_t u = 0; u < kAhead; ++u) {
            const uint32_t j = g0 + u * bsz();
            dd[u] = gf_chunk_desc{0, 0, 0, 0};
            if (j < L.total) {
                uint32_t k = 0;
                while (k + 1u < n && L.off[k + 1u] <= j) ++k;
                dd[u] = e.arr[k].in[j - L.off[k]];
            }
        }
#pragma unroll
        for (uint32_t u = 0; u < kAhead; ++u) {
            const uint32_t j = g0 + u * bsz();
            if (j < L.total)
                e.free_stack[L.top + j] = (unsigned long long)dd[u].index | ((unsigned long long)(dd[u].generation + 1u) << 32);
        }
    }
    __syncthreads();
    if (tid == 0) {
        L.top += L.total;
        for (uint32_t k = 0; k < n; ++k) {
            if (!L.st[k].active) continue;
            const uint32_t kind = L.kind[k];
            if (kind == KIND_SINK) {
                L.sunk += L.st[k].particles;
                L.sink_items += L.st[k].particles;
            } else if (kind == KIND_SOURCE) {
                L.src_items += L.st[k].particles;
            } else {
                L.batches += 1u;
                if (kind == KIND_PARTITION) L.part_items += L.st[k].particles;
                else L.xf_items += L.st[k].particles;
            }
            L.st[k].active = 0u;
        }
        L.step += 1u;
    }
    __syncthreads();
    load_job_tab(e, L);
    __syncthreads();
}

// Select of the next step. With lists = false the input lists, tile / item prefixes and work
// prefixes are left to the grid list kernel (engine_lists_grid), which reads each lane's inputs
// from its ring at in_slot. The caller has loaded the lane table, the rings' head and tail and the
// job table into L.
__device__ inline void engine_select(const Engine& e, SelLds& L, bool lists = true)
{
    const uint32_t tid = threadIdx.x, n = e.n_nodes, C = e.C, Q = e.Q;
    // optional profile (use_graph = 3): wall-clock ticks of the select's phases
    unsigned long long pt = e.prof && tid == 0u ? wall_clock64() : 0ull;
    auto pmark = [&](uint32_t slot) {
        if (!e.prof || tid != 0u) return;
        const unsigned long long now = wall_clock64();
        e.prof[slot] += now - pt;
        pt = now;
    };
    // the lane table (load_lane_tab) and the rings' head / tail are in LDS: every caller loads them
    __syncthreads();
    // stage the candidate ready descriptors of every transform's ring in LDS: <= Q per lane (the
    // first lane of a node stages for all its lanes). A sink takes all its ready chunks, which its
    // input list reads straight from the ring. A thread per lane, one block scan for the offsets.
    {
        uint32_t staged_w = 0u;
        if (tid < n) {
            const uint32_t kind = L.kind[tid];
            const unsigned long long avail = L.tail[tid] - L.head[tid];
            const unsigned long long cap = (unsigned long long)Q * L.lanes[tid];
            const uint32_t want = kind == KIND_SOURCE || L.ring_of[tid] != tid ? 0u
                                : kind == KIND_SINK ? (uint32_t)avail : (avail < cap ? (uint32_t)avail : (uint32_t)cap);
            L.wanted[tid] = want;
            L.taken[tid] = 0u;
            L.ring_at[tid] = (uint32_t)(L.head[tid] % e.ring_cap);
            if (kind != KIND_SINK) staged_w = want;
        }
        uint32_t staged;
        const uint32_t off = block_exscan(staged_w, L.scan, &staged);
        if (tid < n) L.lst_off[tid] = off;
        if (tid == 0) {
            L.lst_off[n] = staged;
            L.staged = staged;
            L.overflow = staged > kListCap ? 1u : 0u;
        }
    }
    __syncthreads();
    // descriptor i of node k's ready chunks (i < the node's ring occupancy <= ring_cap)
    auto ring_entry = [&](uint32_t k, uint32_t i) -> const gf_chunk_desc& {
        uint32_t slot = L.ring_at[k] + i;
        if (slot >= e.ring_cap) slot -= e.ring_cap;
        return e.rings[(size_t)e.arr[k].ring * e.ring_cap + slot];
    };
    if (!L.overflow) {
        for (uint32_t j = tid; j < L.staged; j += bsz()) {
            const uint32_t k = lane_of(L.lst_off, n, j);
            L.lst[j] = ring_entry(k, j - L.lst_off[k]);
            L.fpre[j] = L.lst[j].fill;
        }
        __syncthreads();
        // the items of any run of a node's candidates are a difference of two prefix entries
        // (staged <= kListCap descriptors of <= C items: 32-bit sums)
        const uint32_t sum = block_exscan_array(L.fpre, L.staged, L.scan);
        if (tid == 0) L.fpre[L.staged] = sum;
    }
    __syncthreads();
    pmark(PROF_SEL_STAGE);
    // decisions (GraphRunner's order: node order, every lane after the earlier ones). A lane depends
    // on the earlier lanes only through the credits they reserved, the candidates its node's earlier
    // lanes took and, at the source, the items admitted before it. When the step's whole demand fits
    // the free stack no lane shrinks: lane i of a node takes min(Q, wanted - i Q) candidates and the
    // source's lanes admit in turn as if alone. A thread per lane decides that case (one block sum
    // checks the demand against the free stack); otherwise one thread walks the lanes in node order.
    bool serial = L.overflow != 0u;
    if (!serial) {
        if (tid < n && L.kind[tid] == KIND_SOURCE) {
            // the source of one job: its lanes admit in turn, up to src_chunks chunks each, against
            // the job's items and live bound. The node's admissions in closed form: while at least
            // C items remain and the live bound leaves room for C, every chunk is full; then at most
            // one partial chunk (the tail of N, or a live bound below C). Lane j takes chunks
            // [j K, (j + 1) K) of that sequence; a thread per lane.
            const uint32_t l = L.ring_of[tid], j = tid - l, jb = L.job[tid];
            const unsigned long long N = L.jN[jb], Pj = L.jP[jb], gen = L.jgen[jb];
            const unsigned long long live = gen - L.jsunk[jb], rem = N - gen;
            const unsigned long long cap = (unsigned long long)L.lanes[l] * e.src_chunks;
            unsigned long long full = 0ull, tail = 0ull;  // full chunks, then the items of a partial one
            if (gen < N && live < Pj) {
                if (Pj >= C) {
                    full = (Pj - live) / C;
                    if (rem / C < full) full = rem / C;
                    if (cap < full) full = cap;
                    const unsigned long long rem1 = rem - full * C, live1 = live + full * C;
                    if (full < cap && rem1 > 0ull && rem1 < C && live1 < Pj && Pj - live1 >= rem1) tail = rem1;
                } else if (rem >= Pj) {
                    if (live == 0ull) tail = Pj;  // one chunk of P items
                } else if (Pj - live >= rem) {
                    tail = rem;
                }
            }
            const unsigned long long total = full + (tail ? 1ull : 0ull);
            const unsigned long long c0 = (unsigned long long)j * e.src_chunks;
            unsigned long long c1 = c0 + e.src_chunks;
            if (c1 > total) c1 = total;
            unsigned long long fill = 0ull;
            if (c1 > c0) {
                const unsigned long long fc = (c1 < full ? c1 : full) - (c0 < full ? c0 : full);  // full chunks of the lane
                fill = fc * C + (tail && c1 == total ? tail : 0ull);
            }
            const uint32_t T = e.arr[tid].tile_items;
            L.dfill[tid] = (uint32_t)fill;
            L.dtiles[tid] = (uint32_t)((fill + T - 1u) / T);
            // an empty lane's first id is the end of the node's admissions (the job's next id)
            L.dfirst[tid] = c1 > c0 ? gen + (c0 < full ? c0 : full) * C : gen + full * C + tail;
        }
        __syncthreads();
        unsigned long long dem[1] = {0ull};
        if (tid < n) {
            const uint32_t kind = L.kind[tid];
            if (kind == KIND_SOURCE) {
                dem[0] = (L.dfill[tid] + C - 1u) / C;  // chunks admitted
            } else if (kind != KIND_SINK) {
                const uint32_t l = L.ring_of[tid], w = L.wanted[l], i = tid - l;
                const uint32_t at = i * Q < w ? i * Q : w, m = w - at < Q ? w - at : Q;
                const unsigned long long parts = L.fpre[L.lst_off[l] + at + m] - L.fpre[L.lst_off[l] + at];
                dem[0] = bound_of(parts, C, kind == KIND_PARTITION ? L.bins[tid] : 1u);
            }
        }
        block_sum64(dem, L.red);
        serial = dem[0] > (unsigned long long)L.top;
    }
    if (!serial) {
        unsigned long long act[2] = {0ull, 0ull};  // active lanes, admitted items
        if (tid < n) {
            NodeSt& s = L.st[tid];
            const uint32_t kind = L.kind[tid];
            s.active = 0u;
            s.n_in = 0u;
            s.n_out = 0u;
            s.n_tiles = 0u;
            if (kind == KIND_SOURCE) {
                if (L.dfill[tid]) {
                    s.particles = L.dfill[tid];
                    s.first_id = L.dfirst[tid];
                    s.n_tiles = L.dtiles[tid];
                    s.active = 1u;
                    act[1] = L.dfill[tid];
                }
                if (L.ring_of[tid] == tid) {
                    const uint32_t last = tid + L.lanes[tid] - 1u;
                    L.jgen[L.job[tid]] = L.dfirst[last] + L.dfill[last];
                }
            } else if (kind == KIND_SINK) {
                s.n_in = L.wanted[tid];
                s.active = s.n_in ? 1u : 0u;
                L.in_at[tid] = 0u;
                s.in_slot = L.ring_at[tid];
            } else {
                const uint32_t l = L.ring_of[tid], w = L.wanted[l], i = tid - l;
                const uint32_t at = i * Q < w ? i * Q : w, m = w - at < Q ? w - at : Q;
                s.n_in = m;
                s.active = m ? 1u : 0u;
                L.in_at[tid] = at;
                uint32_t slot = L.ring_at[l] + at;  // at < ring occupancy <= ring_cap
                if (slot >= e.ring_cap) slot -= e.ring_cap;
                s.in_slot = slot;
            }
            // a node's lanes take all its candidates together: wanted <= Q * lanes
            if (kind != KIND_SOURCE && L.ring_of[tid] == tid) L.head[tid] += L.wanted[tid];
            act[0] = s.active;
        }
        block_sum64(act, L.red);
        if (tid == 0) {
            L.any = act[0] ? 1u : 0u;
            L.generated += act[1];
        }
    }
    if (serial && tid == 0) {
        uint32_t reserved = 0;
        L.any = 0u;
        for (uint32_t k = 0; k < n; ++k) {
            NodeSt& s = L.st[k];
            const uint32_t kind = L.kind[k];
            s.active = 0u;
            s.n_in = 0u;
            s.n_out = 0u;
            s.n_tiles = 0u;
            if (kind == KIND_SOURCE) {
                const uint32_t jb = L.job[k];
                const unsigned long long first = L.jgen[jb], N = L.jN[jb], Pj = L.jP[jb];
                unsigned long long gen = first;
                for (uint32_t c = 0; c < e.src_chunks; ++c) {
                    const unsigned long long live = gen - L.jsunk[jb], remaining = N - gen;
                    unsigned long long min_fill = C < Pj ? C : Pj;
                    if (remaining < min_fill) min_fill = remaining;
                    if (!(gen < N && live < Pj && (Pj - live) >= min_fill && L.top - reserved >= 1u)) break;
                    unsigned long long fill = remaining;
                    if (fill > Pj - live) fill = Pj - live;
                    if (fill > C) fill = C;
                    gen += fill;
                    reserved += 1u;
                }
                if (gen > first) {
                    const uint32_t fill = (uint32_t)(gen - first);
                    s.particles = fill;
                    s.first_id = first;
                    s.n_tiles = (fill + e.arr[k].tile_items - 1u) / e.arr[k].tile_items;
                    L.jgen[jb] = gen;
                    L.generated += fill;
                    s.active = 1u;
                }
            } else if (kind == KIND_SINK) {
                s.n_in = L.wanted[k];
                s.active = s.n_in ? 1u : 0u;
                L.in_at[k] = 0u;
                s.in_slot = L.ring_at[k];
                L.head[k] += s.n_in;
            } else {
                // this lane's candidates follow the chunks the node's earlier lanes took
                const uint32_t l = L.ring_of[k], at = L.taken[l];
                uint32_t wanted = L.wanted[l] - at < Q ? L.wanted[l] - at : Q;
                unsigned long long parts = 0;
                if (L.overflow) {
                    for (uint32_t i = 0; i < wanted; ++i) parts += ring_entry(l, at + i).fill;
                } else {
                    parts = L.fpre[L.lst_off[l] + at + wanted] - L.fpre[L.lst_off[l] + at];
                }
                const uint32_t B = kind == KIND_PARTITION ? L.bins[k] : 1u;
                while (wanted) {
                    uint64_t bound = bound_of(parts, C, B);
                    if ((uint64_t)(L.top - reserved) >= bound) {
                        reserved += (uint32_t)bound;
                        break;
                    }
                    --wanted;
                    parts -= L.overflow ? ring_entry(l, at + wanted).fill : L.lst[L.lst_off[l] + at + wanted].fill;
                }
                s.n_in = wanted;
                s.active = wanted ? 1u : 0u;
                L.in_at[k] = at;
                L.taken[l] = at + wanted;
                uint32_t slot = L.ring_at[l] + at;  // at < ring occupancy <= ring_cap
                if (slot >= e.ring_cap) slot -= e.ring_cap;
                s.in_slot = slot;
                L.head[l] += wanted;
            }
            L.any |= s.active;
        }
    }
    __syncthreads();
    pmark(PROF_SEL_DECIDE);
    // input lists and tile / item prefixes of every active non-source node
    for (uint32_t k = 0; k < n && lists; ++k) {
        NodeSt& s = L.st[k];
        if (!s.active || L.kind[k] == KIND_SOURCE) continue;
        const NodeArr& a = e.arr[k];
        const uint32_t m = s.n_in, T = a.tile_items, l = L.ring_of[k], at = L.in_at[k];
        uint32_t carry_t = 0, carry_i = 0;
        for (uint32_t g0 = 0; g0 < m; g0 += kAhead * bsz()) {
            // kAhead consecutive inputs per thread: their loads together, one pair of block scans
            // over the threads' sums
            gf_chunk_desc dd[kAhead];
            uint32_t tl[kAhead], st = 0u, sf = 0u;
#pragma unroll
            for (uint32_t u = 0; u < kAhead; ++u) {
                const uint32_t i = g0 + tid * kAhead + u;
                dd[u] = gf_chunk_desc{0, 0, 0, 0};
                if (i < m)
                    dd[u] = L.overflow || L.kind[k] == KIND_SINK ? ring_entry(l, at + i) : L.lst[L.lst_off[l] + at + i];
            }
#pragma unroll
            for (uint32_t u = 0; u < kAhead; ++u) {
                tl[u] = (dd[u].fill + T - 1u) / T;  // 0 for the empty entries past m
                st += tl[u];
                sf += dd[u].fill;
            }
            uint32_t tt, ti;
            uint32_t pt = block_exscan(st, L.scan, &tt);
            uint32_t pi = block_exscan(sf, L.scan, &ti);
#pragma unroll
            for (uint32_t u = 0; u < kAhead; ++u) {
                const uint32_t i = g0 + tid * kAhead + u;
                if (i < m) {
                    a.in[i] = dd[u];
                    a.tile_prefix[i] = carry_t + pt;
                    a.item_prefix[i] = carry_i + pi;
                }
                pt += tl[u];
                pi += dd[u].fill;
            }
            carry_t += tt;
            carry_i += ti;
        }
        if (tid == 0) {
            a.tile_prefix[m] = carry_t;
            a.item_prefix[m] = carry_i;
            s.n_tiles = carry_t;
            s.particles = carry_i;
        }
        __syncthreads();
    }
    pmark(PROF_SEL_LISTS);
    // work prefixes, ring cursors, termination: a thread per lane
    __syncthreads();
    unsigned long long busy[2] = {0ull, 0ull};  // lanes with ready chunks, jobs not fully admitted
    {
        uint32_t ta = 0u, tb = 0u;
        if (tid < n) {
            NodeSt& s = L.st[tid];
            s.head = L.head[tid];
            s.tail = L.tail[tid];
            if (s.active) {
                tb = s.n_tiles;
                if (L.kind[tid] == KIND_PARTITION) ta = s.n_tiles;
            }
            busy[0] = L.tail[tid] != L.head[tid] ? 1ull : 0ull;
        }
        if (tid < L.njobs) busy[1] = L.jgen[tid] < L.jN[tid] ? 1ull : 0ull;
        // a served engine keeps its jobs' counters in the job table
        if (e.jobs && tid < L.njobs) {
            e.jobs[tid].gen = L.jgen[tid];
            e.jobs[tid].sunk = L.jsunk[tid];
        }
        if (lists) {
            uint32_t sa, sb;
            const uint32_t pa = block_exscan(ta, L.scan, &sa);
            const uint32_t pb = block_exscan(tb, L.scan, &sb);
            if (tid < n) {
                e.ka_prefix[tid] = pa;
                e.kb_prefix[tid] = pb;
            }
            if (tid == 0) {
                e.ka_prefix[n] = sa;
                e.kb_prefix[n] = sb;
            }
        }
        block_sum64(busy, L.red);
        // done (every lane idle, nothing in flight, every job admitted): no work prefix may keep an
        // earlier step's tiles, since the grid list kernel that rewrites them stops with the engine
        // and a served engine is woken for its next jobs without a setup
        if (!L.any && busy[0] == 0ull && busy[1] == 0ull)
            for (uint32_t k = tid; k <= n; k += bsz()) {
                e.ka_prefix[k] = 0u;
                e.kb_prefix[k] = 0u;
            }
    }
    if (tid == 0) {
        e.ctl[CTL_TOP] = L.top;
        e.ctl[CTL_TICKET_A] = 0u;  // the next step's grid plan takes its lanes from ticket 0
        e.ctl[CTL_TICKET_B] = 0u;  // and its grid routing its units
        e.ctr[CNT_GENERATED] = L.generated;
        e.ctr[CNT_SUNK] = L.sunk;
        e.ctr[CNT_STEP] = L.step;
        e.ctr[CNT_BATCHES] = L.batches;
        e.ctr[CNT_PART_ITEMS] = L.part_items;
        e.ctr[CNT_XF_ITEMS] = L.xf_items;
        e.ctr[CNT_SRC_ITEMS] = L.src_items;
        e.ctr[CNT_SINK_ITEMS] = L.sink_items;
        if (!L.any) {
            // every job admitted and nothing in flight: done (a served engine: idle until the host
            // admits more jobs)
            const bool drained = busy[0] == 0ull;
            if (drained && busy[1] == 0ull) {
                e.ctl[CTL_DONE] = 1u;
                if (L.top != e.S) set_status(&e.ctl[CTL_STATUS], GF_BAD_ARGUMENT);
            } else {
                set_status(&e.ctl[CTL_STATUS], GF_STALL);
            }
        }
    }
    __syncthreads();
    store_states(e, L.st);
    __syncthreads();
    pmark(PROF_SEL_END);
}

// Node owning work item t: largest k < n with prefix[k] <= t (prefix has n+1 entries).
__device__ inline uint32_t find_node(const uint32_t* prefix, uint32_t n, uint32_t t)
{
    uint32_t k = 0;
    while (k + 1u < n && prefix[k + 1u] <= t) ++k;
    return k;
}

// ------------------------------------------------------------------------------------------
// KP over the grid: one workgroup per active lane, pop offsets by a decoupled look-back
// ------------------------------------------------------------------------------------------
// Workgroups take the step's active lanes in node order by ticket (CTL_TICKET_A, cleared by the
// select of the previous step). A lane's workgroup plans its batch alone: per-(tile, bin) prefix
// over tiles, bin totals, exact output chunk count R and bin bases. Its output chunks start at the
// step's pop position O = sum of R over the earlier lanes, found by a decoupled look-back over the
// lanes' published words (Merrill & Garland, "Single-pass Parallel Prefix Scan with Decoupled
// Look-back", 2016): word = epoch << 36 | flag << 34 | value, flag 1 = the lane's R, flag 2 = the
// inclusive prefix through the lane. The epoch is the step number (28 bits; a word of a step 2^28
// steps earlier is never read), so no word is ever cleared. A
// workgroup publishes its R before it waits, and waits only on lanes with earlier tickets, which
// are held by running workgroups: every wait ends.
// A word carries its value itself and no other data is published through it, so its stores and
// loads are relaxed agent-scope atomics: an acquire / release here would write back and invalidate
// the XCD's L2 (buffer_wbl2 / buffer_inv on gfx950) at every publish and every spin.
enum { LB_AGG = 1ull, LB_INC = 2ull };

__device__ inline unsigned long long lb_word(uint32_t epoch, unsigned long long flag, unsigned long long v)
{
    return ((unsigned long long)(epoch & 0xFFFFFFFu) << 36) | (flag << 34) | (v & ((1ull << 34) - 1ull));
}

__device__ inline void lb_publish(unsigned long long* w, unsigned long long v)
{
    __hip_atomic_store(w, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Exclusive prefix of the published values before unit u (thread 0 only).
__device__ inline unsigned long long lb_lookback(unsigned long long* words, uint32_t u, uint32_t epoch)
{
    unsigned long long acc = 0;
    for (int32_t j = (int32_t)u - 1; j >= 0; --j) {
        unsigned long long w;
        for (;;) {
            w = __hip_atomic_load(&words[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if ((uint32_t)(w >> 36) == (epoch & 0xFFFFFFFu) && ((w >> 34) & 3ull) != 0ull) break;
            __builtin_amdgcn_s_sleep(1);
        }
        acc += w & ((1ull << 34) - 1ull);
        if (((w >> 34) & 3ull) == LB_INC) break;
    }
    return acc;
}

struct GridPlanLds {
    uint32_t unit, units, v, status, fail;
    uint32_t scan[kMaxWaves];
    unsigned long long off, top;
    uint32_t act[kMaxNodes];  // active lanes in node order
};

// Plan of one partition lane: per (tile, bin) exclusive prefix over tiles in place, totals[bin],
// totals[B + bin] = the bin's first output chunk; returns R, *items = the classified items.
__device__ inline uint32_t plan_lane_counts(const Engine& e, uint32_t k, uint32_t nt, uint32_t* scan, uint32_t* items)
{
    const NodeArr& a = e.arr[k];
    const uint32_t B = e.nodes[k].bins, C = e.C, tid = threadIdx.x;
    uint32_t carry_r = 0, carry_n = 0;
    for (uint32_t b0 = 0; b0 < B; b0 += bsz()) {
        const uint32_t bin = b0 + tid;
        uint32_t tot = 0;
        if (bin < B) {
            for (uint32_t t0 = 0; t0 < nt; t0 += 8u) {
                uint32_t x[8];
#pragma unroll
                for (uint32_t j = 0; j < 8u; ++j) x[j] = t0 + j < nt ? a.counts[(size_t)(t0 + j) * B + bin] : 0u;
#pragma unroll
                for (uint32_t j = 0; j < 8u; ++j) {
                    if (t0 + j < nt) a.counts[(size_t)(t0 + j) * B + bin] = tot;
                    tot += x[j];
                }
            }
            a.totals[bin] = tot;
        }
        const uint32_t ch = bin < B ? (tot + C - 1u) / C : 0u;
        uint32_t sum_r, sum_n;
        const uint32_t ex = block_exscan(ch, scan, &sum_r);
        block_exscan(tot, scan, &sum_n);
        if (bin < B) a.totals[B + bin] = carry_r + ex;
        carry_r += sum_r;
        carry_n += sum_n;
    }
    *items = carry_n;
    return carry_r;
}

__device__ inline void engine_plan_grid(const Engine& e, GridPlanLds& L)
{
    const uint32_t tid = threadIdx.x, n = e.n_nodes, C = e.C;
    // the active lanes in node order: a thread per lane (n <= kMaxNodes <= the block), one scan
    const uint32_t on = tid < n && e.st[tid].active ? 1u : 0u;
    if (tid == 0) L.status = load_ctl(&e.ctl[CTL_STATUS]);
    uint32_t units;
    const uint32_t pos = block_exscan(on, L.scan, &units);
    if (on) L.act[pos] = tid;
    if (tid == 0) L.units = units;
    __syncthreads();
    if (L.status) return;
    const uint32_t epoch = (uint32_t)e.ctr[CNT_STEP];
    for (;;) {
        if (tid == 0) {
            L.unit = atomicAdd(&e.ctl[CTL_TICKET_A], 1u);
            L.fail = 0u;
            L.top = e.ctl[CTL_TOP];  // read before this unit publishes (the last unit rewrites it)
        }
        __syncthreads();
        const uint32_t u = L.unit;
        if (u >= L.units) return;
        const uint32_t k = L.act[u];
        NodeSt s = e.st[k];
        const NodeArr& a = e.arr[k];
        const uint32_t kind = a.kind;
        uint32_t R = 0;
        if (kind == KIND_PARTITION) {
            uint32_t items = 0;
            R = plan_lane_counts(e, k, s.n_tiles, L.scan, &items);
            if (tid == 0 && (items != s.particles || R > a.reserve_bound)) {
                set_status(&e.ctl[CTL_STATUS], GF_PLAN_MISMATCH);
                L.fail = 1u;
            }
        } else if (kind == KIND_SOURCE) {
            R = (s.particles + C - 1u) / C;  // the admitted chunks
        } else if (kind == KIND_TRANSFORM) {
            R = (s.particles + C - 1u) / C;
            if (tid == 0 && (s.particles == 0u || R > a.reserve_bound)) {
                set_status(&e.ctl[CTL_STATUS], GF_PLAN_MISMATCH);
                L.fail = 1u;
            }
        }
        __syncthreads();
        if (tid == 0) {
            lb_publish(&e.lb[u], lb_word(epoch, u == 0u ? LB_INC : LB_AGG, R));
            const unsigned long long off = lb_lookback(e.lb, u, epoch);
            if (u != 0u) lb_publish(&e.lb[u], lb_word(epoch, LB_INC, off + R));
            L.off = off;
            if (off + R > L.top && !L.fail) {
                set_status(&e.ctl[CTL_STATUS], GF_POOL_EXHAUSTED);
                L.fail = 1u;
            }
            if (kind == KIND_SINK) {
                s.digest_base = e.ctr[CNT_DIGEST0 + k];
                e.ctr[CNT_DIGEST0 + k] = s.digest_base + s.n_in;
                e.st[k].digest_base = s.digest_base;
            }
            e.st[k].n_out = R;
            if (kind == KIND_TRANSFORM) a.totals[0] = s.particles;
            if (u + 1u == L.units) {
                const uint32_t nt = (uint32_t)(L.top - (off + R));
                e.ctl[CTL_TOP] = nt;
                const unsigned long long live = e.S - nt;
                if (live > e.ctr[CNT_PEAK_LIVE]) e.ctr[CNT_PEAK_LIVE] = live;
            }
        }
        __syncthreads();
        if (L.fail || !R) continue;
        // pops: output j of the step takes stack[top - 1 - j]; bin-major, then chunk order
        const uint32_t top = (uint32_t)L.top, o = (uint32_t)L.off;
        gf_chunk_desc* out = a.out;
        if (kind == KIND_PARTITION) {
            const uint32_t B = e.nodes[k].bins;
            for (uint32_t bin = tid; bin < B; bin += bsz()) {
                const uint32_t tb = a.totals[bin], b0 = a.totals[B + bin];
                for (uint32_t q = 0, rem = tb; rem; ++q) {
                    const unsigned long long ent = e.free_stack[top - 1u - (o + b0 + q)];
                    gf_chunk_desc d;
                    d.index = (uint32_t)ent;
                    d.generation = (uint32_t)(ent >> 32);
                    d.fill = rem < C ? rem : C;
                    d.bin = bin;
                    out[b0 + q] = d;
                    rem -= d.fill;
                }
            }
        } else {
            for (uint32_t r = tid; r < R; r += bsz()) {
                const unsigned long long ent = e.free_stack[top - 1u - (o + r)];
                gf_chunk_desc d;
                d.index = (uint32_t)ent;
                d.generation = (uint32_t)(ent >> 32);
                const uint32_t rem = s.particles - r * C;  // source and transform: C, ..., C, rest
                d.fill = rem < C ? rem : C;
                d.bin = 0u;
                out[r] = d;
            }
        }
        __syncthreads();
    }
}

// Trace records of a step planned over the grid (one workgroup, node order).
__device__ inline void engine_trace_step(const Engine& e, unsigned long long* len_sh)
{
    if (load_ctl(&e.ctl[CTL_STATUS])) return;
    const uint32_t step = (uint32_t)e.ctr[CNT_STEP];
    for (uint32_t k = 0; k < e.n_nodes; ++k) {
        const NodeSt s = e.st[k];
        if (!s.active) continue;
        const uint32_t kind = e.arr[k].kind;
        Batch b = make_batch(e, k, s);
        if (kind == KIND_PARTITION) trace_record(e, k, b, e.nodes[k].bins, e.arr[k].totals, len_sh, step);
        else if (kind == KIND_TRANSFORM) trace_record(e, k, b, 1u, e.arr[k].totals, len_sh, step);
        else trace_record(e, k, b, 0u, nullptr, len_sh, step);
    }
}

// ------------------------------------------------------------------------------------------
// KR over the grid: routing + release (engine_route_grid), the step's counters and the select in
// one workgroup (engine_retire_counters + engine_select without lists), input lists over the grid
// (engine_lists_grid). Same results as engine_retire + engine_select.
// ------------------------------------------------------------------------------------------
constexpr uint32_t kGridBlock = 4096;  // outputs (routing) or inputs (lists) per unit: 1024 threads x 4


// Exclusive prefix of the published values before unit u over words[stride * j + col] (thread-local).
__device__ inline unsigned long long lb_lookback_col(unsigned long long* words, uint32_t u, uint32_t stride,
                                                     uint32_t col, uint32_t first, uint32_t epoch)
{
    unsigned long long acc = 0;
    for (int32_t j = (int32_t)u - 1; j >= (int32_t)first; --j) {
        unsigned long long w;
        for (;;) {
            w = __hip_atomic_load(&words[(size_t)j * stride + col], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if ((uint32_t)(w >> 36) == (epoch & 0xFFFFFFFu) && ((w >> 34) & 3ull) != 0ull) break;
            __builtin_amdgcn_s_sleep(1);
        }
        acc += w & ((1ull << 34) - 1ull);
        if (((w >> 34) & 3ull) == LB_INC) break;
    }
    return acc;
}

struct RouteLds {
    uint32_t off[kMaxNodes + 1];   // outputs before lane k (node order; active non-sink lanes)
    uint32_t ioff[kMaxNodes + 1];  // inputs before lane k (active non-source lanes)
    const gf_chunk_desc* out[kMaxNodes];  // per lane: its output / input descriptors, route offset,
    const gf_chunk_desc* in[kMaxNodes];   // ready ring
    uint32_t route_off[kMaxNodes], ring[kMaxNodes];
    uint32_t scan[kMaxWaves];
    uint32_t cnt[kMaxWaves][kMaxNodes];
    uint32_t run[kMaxNodes];       // per destination ring: this unit's outputs so far
    uint32_t base[kMaxNodes];      // per ring: outputs of the earlier units
    uint32_t ring_at[kMaxNodes];   // per ring: slot of its tail at the step's start
    uint32_t unit, units, total, total_in, top;
};

// Routing of the step's outputs to the destination rings, in order, and the release of its inputs.
// Units of kGridBlock outputs are taken by ticket; a unit ranks its outputs per destination ring,
// and the ring offsets of the earlier units come from one decoupled look-back per ring.
__device__ inline void engine_route_grid(const Engine& e, RouteLds& L)
{
    const uint32_t tid = threadIdx.x, n = e.n_nodes, w = tid >> 6;
    // a thread per lane (n <= kMaxNodes <= the block): the lane's fields into LDS, its output and
    // input counts through two block scans
    uint32_t no = 0u, ni = 0u;
    if (tid < n) {
        const NodeSt& st = e.st[tid];
        const NodeArr& a = e.arr[tid];
        const uint32_t kind = a.kind;
        if (st.active && kind != KIND_SINK) no = st.n_out;
        if (st.active && kind != KIND_SOURCE) ni = st.n_in;
        L.out[tid] = a.out;
        L.in[tid] = a.in;
        L.ring[tid] = a.ring;
        L.route_off[tid] = e.nodes[tid].route_offset;
        L.ring_at[tid] = (uint32_t)(st.tail % e.ring_cap);
    }
    if (tid == 0) L.top = e.ctl[CTL_TOP];
    uint32_t to, ti;
    const uint32_t po = block_exscan(no, L.scan, &to);
    const uint32_t pi = block_exscan(ni, L.scan, &ti);
    if (tid < n) {
        L.off[tid] = po;
        L.ioff[tid] = pi;
    }
    if (tid == 0) {
        L.off[n] = to;
        L.ioff[n] = ti;
        L.total = to;
        L.total_in = ti;
        L.units = (to + kGridBlock - 1u) / kGridBlock;
    }
    __syncthreads();
    const uint32_t epoch = (uint32_t)e.ctr[CNT_STEP];
    // every consumed input back on the free stack (generation + 1), above the step's pops
    for (uint32_t j = blockIdx.x * bsz() + tid; j < L.total_in; j += gridDim.x * bsz()) {
        const uint32_t k = lane_of(L.ioff, n, j);
        const gf_chunk_desc d = L.in[k][j - L.ioff[k]];
        e.free_stack[L.top + j] = (unsigned long long)d.index | ((unsigned long long)(d.generation + 1u) << 32);
    }
    for (;;) {
        if (tid == 0) L.unit = atomicAdd(&e.ctl[CTL_TICKET_B], 1u);
        for (uint32_t k = tid; k < n; k += bsz()) L.run[k] = 0u;
        __syncthreads();
        const uint32_t u = L.unit;
        if (u >= L.units) return;
        gf_chunk_desc dd[kAhead];
        uint32_t ds[kAhead], rk[kAhead], dk[kAhead];
        // the kAhead rounds' descriptors, then their destinations, each as one batch of loads
#pragma unroll
        for (uint32_t r = 0; r < kAhead; ++r) {
            const uint32_t j = u * kGridBlock + r * bsz() + tid;
            dd[r] = gf_chunk_desc{0, 0, 0, 0};
            dk[r] = 0u;
            if (j < L.total) {
                dk[r] = lane_of(L.off, n, j);
                dd[r] = L.out[dk[r]][j - L.off[dk[r]]];
            }
        }
#pragma unroll
        for (uint32_t r = 0; r < kAhead; ++r)
            ds[r] = u * kGridBlock + r * bsz() + tid < L.total ? (uint32_t)e.routes[L.route_off[dk[r]] + dd[r].bin]
                                                               : 0xFFFFFFFFu;
#pragma unroll
        for (uint32_t r = 0; r < kAhead; ++r) {
            const bool v = u * kGridBlock + r * bsz() + tid < L.total;
            for (uint32_t i = tid; i < kMaxWaves * n; i += bsz()) (&L.cnt[0][0])[(i / n) * kMaxNodes + i % n] = 0u;
            __syncthreads();
            const uint64_t peers = peer_mask(v, ds[r]);
            if (v && (peers & lanemask_lt()) == 0ull) L.cnt[w][ds[r]] = (uint32_t)__popcll(peers);
            __syncthreads();
            if (v) {
                uint32_t pre = L.run[ds[r]];
                for (uint32_t ww = 0; ww < w; ++ww) pre += L.cnt[ww][ds[r]];
                rk[r] = pre + (uint32_t)__popcll(peers & lanemask_lt());
            }
            __syncthreads();
            for (uint32_t k = tid; k < n; k += bsz()) {
                uint32_t sum = 0;
                for (uint32_t ww = 0; ww < bwaves(); ++ww) sum += L.cnt[ww][k];
                L.run[k] += sum;
            }
            __syncthreads();
        }
        // per ring: publish this unit's count, look back for the earlier units', publish the prefix
        if (tid < n) {
            unsigned long long* word = &e.rlb[(size_t)u * n + tid];
            lb_publish(word, lb_word(epoch, u == 0u ? LB_INC : LB_AGG, L.run[tid]));
            const unsigned long long b = lb_lookback_col(e.rlb, u, n, tid, 0u, epoch);
            if (u != 0u) lb_publish(word, lb_word(epoch, LB_INC, b + L.run[tid]));
            L.base[tid] = (uint32_t)b;
        }
        __syncthreads();
#pragma unroll
        for (uint32_t r = 0; r < kAhead; ++r) {
            if (ds[r] == 0xFFFFFFFFu) continue;
            // a ring never holds more than ring_cap (= S) chunks: one subtraction wraps the slot
            uint32_t slot = L.ring_at[ds[r]] + L.base[ds[r]] + rk[r];
            if (slot >= e.ring_cap) slot -= e.ring_cap;
            e.rings[(size_t)L.ring[ds[r]] * e.ring_cap + slot] = dd[r];
        }
        if (u + 1u == L.units)
            for (uint32_t k = tid; k < n; k += bsz()) e.st[k].tail += L.base[k] + L.run[k];
        __syncthreads();
    }
}

// The rest of the retire (one workgroup, after the grid routing): counters, the inputs' release
// offset on the free stack, every lane inactive; then the caller runs engine_select(e, L, false).
__device__ inline void engine_retire_counters(const Engine& e, SelLds& L)
{
    const uint32_t tid = threadIdx.x, n = e.n_nodes;
    load_states(e, L.st);
    load_lane_tab(e, L);
    __syncthreads();
    // the retired lanes' counts: a thread per lane, one block sum (released inputs, sunk items,
    // admitted items, transform batches, partition items, transform items)
    unsigned long long v[6] = {0ull, 0ull, 0ull, 0ull, 0ull, 0ull};
    if (tid < n) {
        const NodeSt& s = L.st[tid];
        L.tail[tid] = s.tail;
        L.head[tid] = s.head;
        if (s.active) {
            const uint32_t kind = L.kind[tid];
            if (kind != KIND_SOURCE) v[0] = s.n_in;
            if (kind == KIND_SINK) {
                v[1] = s.particles;
            } else if (kind == KIND_SOURCE) {
                v[2] = s.particles;
            } else {
                v[3] = 1ull;
                if (kind == KIND_PARTITION) v[4] = s.particles;
                else v[5] = s.particles;
            }
        }
    }
    const unsigned long long sunk_here = v[1];  // this lane's sunk items (a sink lane)
    block_sum64(v, L.red);
    if (tid < n) L.st[tid].active = 0u;
    if (tid == 0) {
        L.top = e.ctl[CTL_TOP];
        L.generated = e.ctr[CNT_GENERATED];
        L.sunk = e.ctr[CNT_SUNK];
        L.step = e.ctr[CNT_STEP];
        L.batches = e.ctr[CNT_BATCHES];
        L.part_items = e.ctr[CNT_PART_ITEMS];
        L.xf_items = e.ctr[CNT_XF_ITEMS];
        L.src_items = e.ctr[CNT_SRC_ITEMS];
        L.sink_items = e.ctr[CNT_SINK_ITEMS] + v[1];
        L.top += (uint32_t)v[0];
        L.sunk += v[1];
        L.src_items += v[2];
        L.batches += v[3];
        L.part_items += v[4];
        L.xf_items += v[5];
        L.step += 1u;
    }
    __syncthreads();
    load_job_tab(e, L);
    __syncthreads();
    // a served engine: each sink lane's items to its job (an engine that is not served has them in
    // L.sunk already)
    if (e.jobs && sunk_here) atomicAdd(&L.jsunk[L.job[tid]], sunk_here);
    __syncthreads();
}

struct ListLds {
    uint32_t uoff[kMaxNodes + 1];  // list units before lane k (active non-source lanes)
    uint32_t scan[kMaxWaves];
    uint32_t unit, units, k, first, last, ex_t, ex_i, tot_t, tot_i;
};

// Input lists and tile / item prefixes of every active non-source lane: units of kGridBlock inputs
// in (lane, position) order by ticket; a lane's prefixes continue across its units through a
// decoupled look-back over the lane's earlier units. The workgroup that completes the last unit
// writes the step's work prefixes.
__device__ inline void engine_lists_grid(const Engine& e, ListLds& L)
{
    const uint32_t tid = threadIdx.x, n = e.n_nodes;
    // list units of every lane: a thread per lane (n <= kMaxNodes <= the block), one block scan
    uint32_t nu = 0u;
    if (tid < n) {
        const NodeSt& st = e.st[tid];
        if (st.active && e.arr[tid].kind != KIND_SOURCE) nu = (st.n_in + kGridBlock - 1u) / kGridBlock;
    }
    uint32_t units;
    const uint32_t pu = block_exscan(nu, L.scan, &units);
    if (tid < n) L.uoff[tid] = pu;
    if (tid == 0) {
        L.uoff[n] = units;
        L.units = units;
    }
    __syncthreads();
    const uint32_t epoch = (uint32_t)e.ctr[CNT_STEP];
    bool finisher = false;
    if (L.units == 0u) {
        finisher = blockIdx.x == 0;
    } else {
        for (;;) {
            if (tid == 0) L.unit = atomicAdd(&e.ctl[CTL_TICKET_C], 1u);
            __syncthreads();
            const uint32_t u = L.unit;
            if (u >= L.units) break;
            if (tid == 0) {
                L.k = lane_of(L.uoff, n, u);
                L.first = L.uoff[L.k];
                L.last = L.uoff[L.k + 1u] - 1u;
            }
            __syncthreads();
            const uint32_t k = L.k;
            const NodeSt st = e.st[k];
            const NodeArr& a = e.arr[k];
            const uint32_t l = e.ring_of[k], T = a.tile_items;
            const uint32_t i0 = (u - L.first) * kGridBlock;
            gf_chunk_desc dd[kAhead];
            uint32_t tl[kAhead], sum_t = 0, sum_i = 0;
#pragma unroll
            for (uint32_t r = 0; r < kAhead; ++r) {
                const uint32_t i = i0 + tid * kAhead + r;
                dd[r] = gf_chunk_desc{0, 0, 0, 0};
                if (i < st.n_in) {
                    uint32_t slot = st.in_slot + i;
                    if (slot >= e.ring_cap) slot -= e.ring_cap;
                    dd[r] = e.rings[(size_t)e.arr[l].ring * e.ring_cap + slot];
                }
                tl[r] = (dd[r].fill + T - 1u) / T;
                sum_t += tl[r];
                sum_i += dd[r].fill;
            }
            uint32_t tt, ti;
            uint32_t pt = block_exscan(sum_t, L.scan, &tt);
            uint32_t pi = block_exscan(sum_i, L.scan, &ti);
            if (tid == 0) {
                unsigned long long* wt = &e.llb[2ull * u];
                const bool head = u == L.first;
                lb_publish(wt, lb_word(epoch, head ? LB_INC : LB_AGG, tt));
                lb_publish(wt + 1, lb_word(epoch, head ? LB_INC : LB_AGG, ti));
                const unsigned long long bt = head ? 0ull : lb_lookback_col(e.llb, u, 2u, 0u, L.first, epoch);
                const unsigned long long bi = head ? 0ull : lb_lookback_col(e.llb, u, 2u, 1u, L.first, epoch);
                if (!head) {
                    lb_publish(wt, lb_word(epoch, LB_INC, bt + tt));
                    lb_publish(wt + 1, lb_word(epoch, LB_INC, bi + ti));
                }
                L.ex_t = (uint32_t)bt;
                L.ex_i = (uint32_t)bi;
                L.tot_t = (uint32_t)bt + tt;
                L.tot_i = (uint32_t)bi + ti;
            }
            __syncthreads();
            pt += L.ex_t;
            pi += L.ex_i;
#pragma unroll
            for (uint32_t r = 0; r < kAhead; ++r) {
                const uint32_t i = i0 + tid * kAhead + r;
                if (i < st.n_in) {
                    a.in[i] = dd[r];
                    a.tile_prefix[i] = pt;
                    a.item_prefix[i] = pi;
                }
                pt += tl[r];
                pi += dd[r].fill;
            }
            if (tid == 0 && u == L.last) {
                a.tile_prefix[st.n_in] = L.tot_t;
                a.item_prefix[st.n_in] = L.tot_i;
                e.st[k].n_tiles = L.tot_t;
                e.st[k].particles = L.tot_i;
            }
            __syncthreads();
            if (tid == 0) {
                // the unit's stores are ordered before its completion count; the finisher reads
                // every lane's n_tiles after an acquire
                const uint32_t done = __hip_atomic_fetch_add(&e.ctl[CTL_LIST_DONE], 1u, __ATOMIC_ACQ_REL,
                                                             __HIP_MEMORY_SCOPE_AGENT);
                L.k = done + 1u == L.units ? 1u : 0u;
            }
            __syncthreads();
            if (L.k) finisher = true;
        }
    }
    // the finisher (one workgroup; `finisher` is uniform in it) writes the work prefixes: a thread
    // per lane, two block scans. Its acquire (thread 0, on the done count) ordered the other
    // workgroups' n_tiles before this fence; each thread's fence orders its own loads after it.
    if (finisher) {
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        uint32_t ta = 0u, tb = 0u;
        if (tid < n) {
            const NodeSt& st = e.st[tid];
            if (st.active) {
                tb = st.n_tiles;
                if (e.arr[tid].kind == KIND_PARTITION) ta = st.n_tiles;
            }
        }
        uint32_t sa, sb;
        const uint32_t pa = block_exscan(ta, L.scan, &sa);
        const uint32_t pb = block_exscan(tb, L.scan, &sb);
        if (tid < n) {
            e.ka_prefix[tid] = pa;
            e.kb_prefix[tid] = pb;
        }
        if (tid == 0) {
            e.ka_prefix[n] = sa;
            e.kb_prefix[n] = sb;
            e.ctl[CTL_LIST_DONE] = 0u;
            e.ctl[CTL_TICKET_C] = 0u;
        }
    }
}

}  // namespace gfd
