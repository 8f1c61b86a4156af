// Device-resident step engine of the batched form.
//
// One step = two kernels:
//   KA: classify+count every partition tile of the step; the last workgroup to finish (atomic
//       ticket, no spinning) plans every node in node order: per-tile flat offsets, exact output
//       chunk counts, pops from the device free stack, output descriptors, trace.
//   KB: scatter / transform / source / sink tiles of every active node; the last workgroup
//       retires the step (routes outputs to the destination ready rings, releases inputs to the
//       free stack) and selects the next step's batches (source admission against P, largest
//       credit-fitting prefix of <= Q ready chunks per transform node, all ready chunks per sink).
// The scheduling rules are those of GraphRunner::run (sched/GraphRunner.hpp:902-1133) applied
// once per node per step; the host only replays a hipGraph of K steps and polls a done flag.
#pragma once

#include "gf_device.hpp"

namespace gfd {

struct Engine {
    uint32_t n_nodes, C, P, Q, S, ring_cap, grid, pad;
    uint64_t N;
    RunParams rp;
    Cols cols;
    const gf_node_desc* nodes;
    const int32_t* routes;
    const uint32_t* reserve_bound; // per node PartitionStage/TransformStage reserve bound
    Batch* batches;                // per node
    uint32_t* free_stack;          // S
    uint32_t* gens;                // S
    gf_chunk_desc* rings;          // n_nodes * ring_cap
    unsigned long long* ring_head; // n_nodes
    unsigned long long* ring_tail; // n_nodes
    uint32_t* ka_prefix;           // n_nodes + 1
    uint32_t* kb_prefix;           // n_nodes + 1
    uint32_t* ctl;                 // control words, see CTL_*
    unsigned long long* ctr;       // counters, see CNT_*
    uint32_t* trace;
    unsigned long long trace_cap;
};

enum { CTL_STATUS = 0, CTL_DONE = 1, CTL_TOP = 2, CTL_TICKET_A = 3, CTL_TICKET_B = 4, CTL_WORDS = 8 };
enum {
    CNT_GENERATED = 0, CNT_SUNK = 1, CNT_STEP = 2, CNT_BATCHES = 3, CNT_PART_ITEMS = 4, CNT_XF_ITEMS = 5,
    CNT_SRC_ITEMS = 6, CNT_SINK_ITEMS = 7, CNT_TRACE_LEN = 8, CNT_PEAK_LIVE = 9, CNT_DIGEST0 = 16,
    CNT_WORDS = 16 + 64
};

__device__ inline bool is_partition_op(uint32_t op) { return op >= GF_CLS_WEIGHT_ROUTE && op < GF_XF_SCALE_WEIGHT; }

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
    for (int k = 0; k < kWaves; ++k) {
        if ((uint32_t)k < w) off += sh[k];
        tot += sh[k];
    }
    __syncthreads();
    *total = tot;
    return off + x - v;
}

// Exclusive scan of n values at a[0..n) in place (one workgroup), returns the total.
__device__ inline uint32_t block_exscan_array(uint32_t* a, uint32_t n, uint32_t* sh)
{
    const uint32_t tid = threadIdx.x;
    const uint32_t per = (n + kThreads - 1u) / kThreads;
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

// Trace record (layout identical to the oracle): 0xB47C0000|node, step, n_in, (index,gen,fill)*n_in,
// nbins, counts, R, (index,gen,fill,bin)*R.
__device__ inline void trace_record(const Engine& e, uint32_t k, const Batch& b, uint32_t nbins,
                                    const uint32_t* counts, unsigned long long* len_sh)
{
    if (!e.trace) return;
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
        put(0, 0xB47C0000u | k);
        put(1, (uint32_t)e.ctr[CNT_STEP]);
        put(2, b.n_in);
        put(3 + 3ull * b.n_in, nbins);
        put(4 + 3ull * b.n_in + nbins, b.n_out);
    }
    for (uint32_t i = tid; i < b.n_in; i += kThreads) {
        put(3 + 3ull * i, b.in[i].index);
        put(4 + 3ull * i, b.in[i].generation);
        put(5 + 3ull * i, b.in[i].fill);
    }
    for (uint32_t i = tid; i < nbins; i += kThreads) put(4 + 3ull * b.n_in + i, counts[i]);
    unsigned long long o0 = 5 + 3ull * b.n_in + nbins;
    for (uint32_t i = tid; i < b.n_out; i += kThreads) {
        put(o0 + 4ull * i, b.out[i].index);
        put(o0 + 4ull * i + 1, b.out[i].generation);
        put(o0 + 4ull * i + 2, b.out[i].fill);
        put(o0 + 4ull * i + 3, b.out[i].bin);
    }
    __syncthreads();
}

__device__ inline uint32_t load_ctl(uint32_t* p) { return atomicAdd(p, 0u); }

struct PlanShared {
    uint32_t scan[kWaves];
    uint32_t top;
    uint32_t ok;
    uint32_t stop;
    uint32_t R;
    uint32_t sum;
    unsigned long long len;
};

// Plan of one partition batch (Execution.hpp:2520-2642 on the device): per (tile, bin) flat
// output start, per bin totals (b.totals[0..B)), bin bases (b.totals[B..2B)), exact chunk count R,
// pops from the free stack in (bin, k) order, output descriptors.
__device__ inline void plan_partition(const Engine& e, uint32_t k, Batch& b, PlanShared& ps)
{
    const gf_node_desc& nd = e.nodes[k];
    const uint32_t B = nd.bins, nt = b.n_tiles, tid = threadIdx.x, C = e.C;
    uint32_t* cnt = b.counts;
    uint32_t* tot = b.totals;
    uint32_t* base = b.totals + B;
    if (B < 64u) {
        for (uint32_t bin = 0; bin < B; ++bin) {
            const uint32_t per = (nt + kThreads - 1u) / kThreads;
            const uint32_t lo = tid * per, hi = lo + per < nt ? lo + per : nt;
            uint32_t s = 0;
            for (uint32_t t = lo; t < hi; ++t) s += cnt[(size_t)t * B + bin];
            uint32_t total;
            uint32_t run = block_exscan(s, ps.scan, &total);
            for (uint32_t t = lo; t < hi; ++t) {
                uint32_t x = cnt[(size_t)t * B + bin];
                cnt[(size_t)t * B + bin] = run;
                run += x;
            }
            if (tid == 0) tot[bin] = total;
            __syncthreads();
        }
    } else {
        for (uint32_t bin = tid; bin < B; bin += kThreads) {
            uint32_t acc = 0;
            for (uint32_t t = 0; t < nt; ++t) {
                uint32_t x = cnt[(size_t)t * B + bin];
                cnt[(size_t)t * B + bin] = acc;
                acc += x;
            }
            tot[bin] = acc;
        }
        __syncthreads();
    }
    // checked totals and exact chunk counts
    uint32_t s = 0;
    for (uint32_t bin = tid; bin < B; bin += kThreads) {
        s += tot[bin];
        base[bin] = (tot[bin] + C - 1u) / C;
    }
    uint32_t sum;
    block_exscan(s, ps.scan, &sum);
    uint32_t R = block_exscan_array(base, B, ps.scan);
    if (tid == 0) {
        ps.ok = 1u;
        if (sum != b.particles) {
            set_status(&e.ctl[CTL_STATUS], GF_PLAN_MISMATCH);
            ps.ok = 0u;
        } else if (R > e.reserve_bound[k]) {
            set_status(&e.ctl[CTL_STATUS], GF_PLAN_MISMATCH);
            ps.ok = 0u;
        } else if (R > ps.top) {
            set_status(&e.ctl[CTL_STATUS], GF_POOL_EXHAUSTED);
            ps.ok = 0u;
        }
        ps.R = R;
    }
    __syncthreads();
    if (!ps.ok) return;
    const uint32_t top = ps.top;
    for (uint32_t bin = tid; bin < B; bin += kThreads) {
        uint32_t nch = (tot[bin] + C - 1u) / C;
        for (uint32_t j = 0; j < nch; ++j) {
            uint32_t r = base[bin] + j;
            uint32_t idx = e.free_stack[top - 1u - r];
            uint32_t rem = tot[bin] - j * C;
            b.out[r] = gf_chunk_desc{idx, e.gens[idx], rem < C ? rem : C, bin};
        }
    }
    const size_t ntb = (size_t)nt * B;
    for (size_t i = tid; i < ntb; i += kThreads) cnt[i] += base[i % B] * C;
    __syncthreads();
    if (tid == 0) {
        ps.top = top - R;
        b.n_out = R;
    }
    __syncthreads();
}

__device__ inline void plan_transform(const Engine& e, uint32_t k, Batch& b, PlanShared& ps)
{
    const uint32_t tid = threadIdx.x, C = e.C;
    const uint32_t R = (b.particles + C - 1u) / C;
    if (tid == 0) {
        ps.ok = 1u;
        if (b.particles == 0u || R > e.reserve_bound[k]) {
            set_status(&e.ctl[CTL_STATUS], GF_PLAN_MISMATCH);
            ps.ok = 0u;
        } else if (R > ps.top) {
            set_status(&e.ctl[CTL_STATUS], GF_POOL_EXHAUSTED);
            ps.ok = 0u;
        }
    }
    __syncthreads();
    if (!ps.ok) return;
    const uint32_t top = ps.top;
    for (uint32_t r = tid; r < R; r += kThreads) {
        uint32_t idx = e.free_stack[top - 1u - r];
        uint32_t rem = b.particles - r * C;
        b.out[r] = gf_chunk_desc{idx, e.gens[idx], rem < C ? rem : C, 0u};
    }
    if (tid == 0) {
        b.totals[0] = b.particles;
        ps.top = top - R;
        b.n_out = R;
    }
    __syncthreads();
}

// Last workgroup of KA.
__device__ inline void engine_plan(const Engine& e, PlanShared& ps)
{
    const uint32_t tid = threadIdx.x;
    if (tid == 0) ps.top = e.ctl[CTL_TOP];
    __syncthreads();
    for (uint32_t k = 0; k < e.n_nodes; ++k) {
        Batch& b = e.batches[k];
        if (!b.active) continue;
        if (tid == 0) ps.stop = load_ctl(&e.ctl[CTL_STATUS]) != 0u ? 1u : 0u;
        __syncthreads();
        if (ps.stop) break;
        const gf_node_desc& nd = e.nodes[k];
        if (nd.role == GF_ROLE_SOURCE) {
            if (tid == 0) {
                if (ps.top == 0u) {
                    set_status(&e.ctl[CTL_STATUS], GF_POOL_EXHAUSTED);
                } else {
                    uint32_t idx = e.free_stack[--ps.top];
                    b.out[0] = gf_chunk_desc{idx, e.gens[idx], b.particles, 0u};
                    b.n_out = 1u;
                }
            }
            __syncthreads();
            trace_record(e, k, b, 0u, nullptr, &ps.len);
        } else if (nd.role == GF_ROLE_SINK) {
            if (tid == 0) {
                b.n_out = 0u;
                b.digest_base = e.ctr[CNT_DIGEST0 + (k & 63u)];
                e.ctr[CNT_DIGEST0 + (k & 63u)] += b.n_in;
            }
            __syncthreads();
            trace_record(e, k, b, 0u, nullptr, &ps.len);
        } else if (is_partition_op(nd.op)) {
            plan_partition(e, k, b, ps);
            if (ps.ok) trace_record(e, k, b, nd.bins, b.totals, &ps.len);
        } else {
            plan_transform(e, k, b, ps);
            if (ps.ok) trace_record(e, k, b, 1u, b.totals, &ps.len);
        }
        __syncthreads();
    }
    if (tid == 0) {
        e.ctl[CTL_TOP] = ps.top;
        uint32_t live = e.S - ps.top;
        if (live > e.ctr[CNT_PEAK_LIVE]) e.ctr[CNT_PEAK_LIVE] = live;
    }
    __syncthreads();
}

// Tile prefix / item prefix of a node's input list (one workgroup).
__device__ inline void build_prefixes(Batch& b, uint32_t tile_items, uint32_t* sh)
{
    const uint32_t tid = threadIdx.x, n = b.n_in;
    for (uint32_t i = tid; i < n; i += kThreads) {
        b.tile_prefix[i] = (b.in[i].fill + tile_items - 1u) / tile_items;
        b.item_prefix[i] = b.in[i].fill;
    }
    __syncthreads();
    uint32_t nt = block_exscan_array(b.tile_prefix, n, sh);
    uint32_t ni = block_exscan_array(b.item_prefix, n, sh);
    if (tid == 0) {
        b.tile_prefix[n] = nt;
        b.item_prefix[n] = ni;
        b.n_tiles = nt;
        b.particles = ni;
    }
    __syncthreads();
}

struct SelectShared {
    uint32_t scan[kWaves];
    uint32_t reserved;
    uint32_t any;
    uint32_t wanted;
    unsigned long long head;
};

__device__ inline uint64_t bound_of(uint64_t n, uint64_t C, uint64_t B)
{
    if (n == 0u) return 0u;
    uint64_t per = (n + C - 1u) / C;
    uint64_t bins = B < n ? B : n;
    uint64_t r = per + bins - 1u;
    return r < n ? r : n;
}

// Last workgroup of KB, part 1: retire the current step in node order.
__device__ inline void engine_retire(const Engine& e)
{
    const uint32_t tid = threadIdx.x, lane = lane_id();
    // ---- retire in node order
    for (uint32_t k = 0; k < e.n_nodes; ++k) {
        Batch& b = e.batches[k];
        if (!b.active) continue;
        const gf_node_desc& nd = e.nodes[k];
        if (nd.role != GF_ROLE_SINK && tid < 64u) {
            for (uint32_t o0 = 0; o0 < b.n_out; o0 += 64u) {
                uint32_t o = o0 + lane;
                bool v = o < b.n_out;
                gf_chunk_desc d = v ? b.out[o] : gf_chunk_desc{0, 0, 0, 0};
                uint32_t dst = v ? (uint32_t)e.routes[nd.route_offset + d.bin] : 0xFFFFFFFFu;
                uint64_t peers = peer_mask(v, dst);
                unsigned long long cur = v ? e.ring_tail[dst] : 0ull;
                unsigned long long pos = cur + (unsigned long long)__popcll(peers & lanemask_lt());
                if (v) e.rings[(size_t)dst * e.ring_cap + (size_t)(pos % e.ring_cap)] = d;
                if (v && (peers & lanemask_lt()) == 0ull) e.ring_tail[dst] = cur + (unsigned long long)__popcll(peers);
            }
        }
        if (nd.role != GF_ROLE_SOURCE) {
            const uint32_t top = e.ctl[CTL_TOP];
            for (uint32_t i = tid; i < b.n_in; i += kThreads) {
                uint32_t idx = b.in[i].index;
                e.gens[idx] += 1u;
                e.free_stack[top + i] = idx;
            }
            __syncthreads();
            if (tid == 0) e.ctl[CTL_TOP] = top + b.n_in;
        }
        if (tid == 0) {
            if (nd.role == GF_ROLE_SINK) {
                e.ctr[CNT_SUNK] += b.particles;
                e.ctr[CNT_SINK_ITEMS] += b.particles;
            } else if (nd.role == GF_ROLE_SOURCE) {
                e.ctr[CNT_SRC_ITEMS] += b.particles;
            } else {
                e.ctr[CNT_BATCHES] += 1u;
                if (is_partition_op(nd.op)) e.ctr[CNT_PART_ITEMS] += b.particles;
                else e.ctr[CNT_XF_ITEMS] += b.particles;
            }
            b.active = 0u;
        }
        __syncthreads();
    }
    if (tid == 0) e.ctr[CNT_STEP] += 1u;
    __syncthreads();
}

// Last workgroup of KB, part 2 (and the setup kernel): select the next step in node order.
__device__ inline void engine_select(const Engine& e, SelectShared& ss)
{
    const uint32_t tid = threadIdx.x;
    const uint32_t C = e.C;
    if (tid == 0) {
        ss.reserved = 0u;
        ss.any = 0u;
    }
    __syncthreads();
    const uint32_t P = e.P, Q = e.Q;
    for (uint32_t k = 0; k < e.n_nodes; ++k) {
        Batch& b = e.batches[k];
        const gf_node_desc& nd = e.nodes[k];
        if (tid == 0) {
            ss.wanted = 0u;
            b.active = 0u;
            b.n_in = 0u;
            b.n_out = 0u;
            const uint32_t freec = e.ctl[CTL_TOP];
            const unsigned long long head = e.ring_head[k], avail_ready = e.ring_tail[k] - head;
            ss.head = head;
            if (nd.role == GF_ROLE_SOURCE) {
                unsigned long long gen = e.ctr[CNT_GENERATED], sunk = e.ctr[CNT_SUNK];
                unsigned long long live = gen - sunk, remaining = e.N - gen;
                unsigned long long min_fill = C < P ? C : P;
                if (remaining < min_fill) min_fill = remaining;
                if (gen < e.N && live < P && (P - live) >= min_fill && freec - ss.reserved >= 1u) {
                    unsigned long long fill = remaining;
                    if (fill > P - live) fill = P - live;
                    if (fill > C) fill = C;
                    b.particles = (uint32_t)fill;
                    b.first_id = gen;
                    b.n_tiles = (uint32_t)((fill + b.tile_items - 1u) / b.tile_items);
                    e.ctr[CNT_GENERATED] = gen + fill;
                    ss.reserved += 1u;
                    b.active = 1u;
                }
            } else if (nd.role == GF_ROLE_TRANSFORM) {
                uint32_t wanted = avail_ready < Q ? (uint32_t)avail_ready : Q;
                unsigned long long parts = 0;
                for (uint32_t i = 0; i < wanted; ++i) parts += e.rings[(size_t)k * e.ring_cap + (head + i) % e.ring_cap].fill;
                while (wanted) {
                    uint64_t bound = bound_of(parts, C, nd.bins);
                    if ((uint64_t)(freec - ss.reserved) >= bound) {
                        ss.reserved += (uint32_t)bound;
                        break;
                    }
                    --wanted;
                    parts -= e.rings[(size_t)k * e.ring_cap + (head + wanted) % e.ring_cap].fill;
                }
                ss.wanted = wanted;
            } else {
                ss.wanted = (uint32_t)avail_ready;
            }
        }
        __syncthreads();
        const uint32_t wanted = ss.wanted;
        if (nd.role != GF_ROLE_SOURCE && wanted) {
            const unsigned long long head = ss.head;
            for (uint32_t i = tid; i < wanted; i += kThreads)
                b.in[i] = e.rings[(size_t)k * e.ring_cap + (head + i) % e.ring_cap];
            __syncthreads();
            if (tid == 0) {
                b.n_in = wanted;
                e.ring_head[k] = head + wanted;
                b.active = 1u;
            }
            __syncthreads();
            build_prefixes(b, b.tile_items, ss.scan);
        }
        if (tid == 0) ss.any |= b.active;
        __syncthreads();
    }
    if (tid == 0) {
        uint32_t ka = 0, kb = 0;
        for (uint32_t k = 0; k < e.n_nodes; ++k) {
            e.ka_prefix[k] = ka;
            e.kb_prefix[k] = kb;
            const Batch& b = e.batches[k];
            if (b.active) {
                if (e.nodes[k].role == GF_ROLE_TRANSFORM && is_partition_op(e.nodes[k].op)) ka += b.n_tiles;
                kb += b.n_tiles;
            }
        }
        e.ka_prefix[e.n_nodes] = ka;
        e.kb_prefix[e.n_nodes] = kb;
        if (!ss.any) {
            bool drained = true;
            for (uint32_t k = 0; k < e.n_nodes; ++k)
                if (e.ring_tail[k] != e.ring_head[k]) drained = false;
            if (drained && e.ctr[CNT_GENERATED] >= e.N) {
                e.ctl[CTL_DONE] = 1u;
                if (e.ctl[CTL_TOP] != e.S) set_status(&e.ctl[CTL_STATUS], GF_BAD_ARGUMENT);
            } else {
                set_status(&e.ctl[CTL_STATUS], GF_STALL);
            }
        }
    }
    __syncthreads();
}

// Node owning work item t: largest k < n with prefix[k] <= t (prefix has n+1 entries).
__device__ inline uint32_t find_node(const uint32_t* prefix, uint32_t n, uint32_t t)
{
    uint32_t k = 0;
    while (k + 1u < n && prefix[k + 1u] <= t) ++k;
    return k;
}

// Ticket: returns true in the last workgroup to arrive (all its global writes are fenced).
__device__ inline bool last_block(uint32_t* ticket, uint32_t* flag_sh)
{
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = atomicAdd(ticket, 1u);
        *flag_sh = (t == gridDim.x - 1u) ? 1u : 0u;
    }
    __syncthreads();
    if (*flag_sh) {
        __threadfence();
        return true;
    }
    return false;
}

}  // namespace gfd
