// One partition batch in one launch: classify, exact plan, output acquisition and stable scatter
// of a ragged list of <= Q input chunks, with no host round trip between them.
//
// The reference runs a partition batch as classify kernel -> counts to the host -> host plan
// (planDenseBins) -> output chunks acquired on the host -> table to the device -> scatter kernel
// (flow/Execution.hpp:2363-2725), two host round trips per batch. Here one workgroup of G tile
// groups (256 threads each) runs all of it:
//   1. tile / item prefixes of the input list (TileDescriptor flattening, Execution.hpp:2377-2389);
//   2. per-tile bin counts into the [tile][bin] rows (dense LDS counters or the hash path's table);
//   3. the exact plan: per (tile, bin) exclusive prefix over tiles, bin totals, R output chunks,
//      bin bases (DenseBinTable.hpp:39-77), with the reference's checks (classified items equal the
//      batch, R within the reservation);
//   4. output acquisition: output r takes reserved handle r, in bin-major, chunk order. The caller
//      hands in the batch's reservation (outputChunkBound handles, popped from its free list in
//      LIFO order) and takes back the unused tail, so the pool ends as if exactly R chunks were
//      acquired in that order;
//   5. the stable scatter (canonical order: bin-major, then input order) with the updater.
// The result words are {R, status}. Scratch (the [tile][bin] rows, totals and bin bases, tile and
// item prefixes) is caller-provided device memory.
#pragma once

#include "gf_engine.hpp"

namespace gfd {

// Scratch words of one fused batch: prefixes 2 (n + 1), totals + bin bases 2B, [tile][bin] rows.
__host__ __device__ inline uint64_t fused_scratch_words(uint32_t n, uint32_t bins, uint32_t tiles)
{
    return 2ull * (n + 1u) + 2ull * bins + (uint64_t)tiles * bins;
}

constexpr uint32_t kArgRes = 64u;

struct FusedArgs {
    gf_node_desc nd;
    const gf_chunk_desc* in;          // n input chunks (device)
    uint32_t n, tile_items, C, hash_entries;
    const gf_chunk_handle* reserved;  // n_reserved output handles (device)
    uint32_t n_reserved, purpose;
    gf_chunk_desc* out;               // R output descriptors (device, n_reserved entries)
    uint32_t* scratch;                // fused_scratch_words(n, bins, tiles) words (device)
    uint32_t* result;                 // [0] R, [1] status
    uint32_t group_lds, tiles_cap;    // LDS bytes per tile group; most tiles the scratch holds
    // inl = 1: the inputs and the reservation travel in the launch itself (n <= kArgIn,
    // n_reserved <= kArgRes), read in the kernel-argument segment like a plain stage's batch
    uint32_t inl, pad;
    gf_chunk_desc in_arg[kArgIn];
    gf_chunk_handle res_arg[kArgRes];
};

// Tile and item prefixes of the input list (TileDescriptor flattening, Execution.hpp:2377-2389),
// by one thread; bad = a fill above C or more tiles than the scratch holds.
__device__ inline void fused_prefixes(const FusedArgs& a, const gf_chunk_desc* in, uint32_t* tile_prefix,
                                      uint32_t* item_prefix, uint32_t& tiles, uint32_t& items, uint32_t& bad)
{
    uint32_t t = 0, p = 0;
    bad = 0u;
    for (uint32_t c = 0; c < a.n; ++c) {
        tile_prefix[c] = t;
        item_prefix[c] = p;
        const uint32_t f = in[c].fill;
        if (f > a.C) bad = 1u;
        t += (f + a.tile_items - 1u) / a.tile_items;
        p += f;
    }
    tile_prefix[a.n] = t;
    item_prefix[a.n] = p;
    if (t > a.tiles_cap) bad = 1u;
    tiles = t;
    items = p;
}

__device__ inline Batch fused_batch(const FusedArgs& a, const gf_chunk_desc* in, uint32_t* tile_prefix,
                                    uint32_t* item_prefix, uint32_t tiles, uint32_t items)
{
    const uint32_t B = a.nd.bins;
    Batch b;
    b.active = 1u;
    b.n_in = a.n;
    b.n_tiles = tiles;
    b.n_out = 0u;
    b.particles = items;
    b.reserved = 0u;
    b.tile_items = a.tile_items;
    b.purpose = a.purpose;
    b.first_id = 0ull;
    b.digest_base = 0ull;
    b.in = const_cast<gf_chunk_desc*>(in);
    b.tile_prefix = tile_prefix;
    b.item_prefix = item_prefix;
    b.out = a.out;
    b.totals = a.scratch + 2u * (a.n + 1u);  // [B] totals, then [B] bin bases
    b.counts = b.totals + 2u * B;            // [tile][bin]
    return b;
}

template <class P, int R>
__device__ inline void fused_count_tile(const FusedArgs& a, const Batch& b, uint32_t t, const Cols& cols,
                                        uint32_t* status, char* gsm)
{
    if (a.hash_entries) count_tile_hash<P, R>(a.nd, b, t, cols, a.C, status, gsm, a.hash_entries);
    else count_tile<P, R>(a.nd, b, t, cols, a.C, status, gsm);
}

template <class P, int R>
__device__ inline void fused_scatter_tile(const FusedArgs& a, const Batch& b, uint32_t t, const Cols& cols,
                                          uint32_t* status, char* gsm)
{
    const uint32_t B = a.nd.bins;
    const bool stage = !a.hash_entries && B <= 64u &&
                       scatter_smem(B, a.tile_items, sizeof(P), true) <= (size_t)a.group_lds;
    if (a.hash_entries)
        scatter_tile_hash<P, R>(a.nd, b, t, cols, a.C, status, gsm, a.hash_entries, nullptr, nullptr, nullptr,
                                b.totals + B);
    else if (stage)
        scatter_tile<P, R, true>(a.nd, b, t, cols, a.C, status, gsm, nullptr, nullptr, nullptr, b.totals + B);
    else
        scatter_tile<P, R, false>(a.nd, b, t, cols, a.C, status, gsm, nullptr, nullptr, nullptr, b.totals + B);
}

// 3. the exact plan: per (tile, bin) exclusive prefix over tiles in place, bin totals, R, bin bases
// (DenseBinTable.hpp:39-77) with the checks (classified items equal the batch, R within the
// reservation); 4. output r takes reserved handle r (bin-major, chunk order). The whole block;
// returns 0 or GF_PLAN_MISMATCH, and R.
__device__ inline uint32_t fused_plan(const FusedArgs& a, const Batch& b, const gf_chunk_handle* reserved,
                                      uint32_t* sh_scan, uint32_t* sh_word, uint32_t& R)
{
    const uint32_t tid = threadIdx.x, B = a.nd.bins, C = a.C, nt = b.n_tiles;
    uint32_t* counts = b.counts;
    uint32_t* totals = b.totals;
    uint32_t carry_r = 0, carry_n = 0;
    for (uint32_t b0 = 0; b0 < B; b0 += blockDim.x) {
        const uint32_t bin = b0 + tid;
        uint32_t tot = 0;
        if (bin < B) {
            for (uint32_t t = 0; t < nt; ++t) {
                const uint32_t x = counts[(size_t)t * B + bin];
                counts[(size_t)t * B + bin] = tot;
                tot += x;
            }
            totals[bin] = tot;
        }
        const uint32_t ch = bin < B ? (tot + C - 1u) / C : 0u;
        uint32_t sum_r, sum_n;
        const uint32_t ex = block_exscan(ch, sh_scan, &sum_r);
        block_exscan(tot, sh_scan, &sum_n);
        if (bin < B) totals[B + bin] = carry_r + ex;
        carry_r += sum_r;
        carry_n += sum_n;
    }
    if (tid == 0) *sh_word = carry_n != b.particles || carry_r > a.n_reserved ? (uint32_t)GF_PLAN_MISMATCH : 0u;
    __syncthreads();
    const uint32_t st = *sh_word;
    __syncthreads();
    R = carry_r;
    if (st) return st;
    for (uint32_t bin = tid; bin < B; bin += blockDim.x) {
        const uint32_t b0 = totals[B + bin];
        for (uint32_t q = 0, rem = totals[bin]; rem; ++q) {
            const gf_chunk_handle h = reserved[b0 + q];
            gf_chunk_desc d;
            d.index = h.index;
            d.generation = h.generation;
            d.fill = rem < C ? rem : C;
            d.bin = bin;
            a.out[b0 + q] = d;
            rem -= d.fill;
        }
    }
    __syncthreads();
    return 0u;
}

// The one-workgroup form: G tile groups of 256 threads run every phase, G tiles at a time.
template <class P, int R>
__device__ void fused_partition_batch(const FusedArgs& a, const Cols& cols, char* smem)
{
    __shared__ uint32_t sh_n_tiles, sh_particles, sh_status, sh_word, sh_scan[kMaxWaves];
    const uint32_t tid = threadIdx.x, G = blockDim.x / kThreads, g = tid / kThreads;
    const uint32_t n = a.n;
    uint32_t* tile_prefix = a.scratch;
    uint32_t* item_prefix = tile_prefix + (n + 1u);
    char* gsm = smem + (size_t)g * a.group_lds;
    const gf_chunk_desc* in = a.inl ? a.in_arg : a.in;
    const gf_chunk_handle* reserved = a.inl ? a.res_arg : a.reserved;
    // 1. prefixes (n <= Q: one thread)
    if (tid == 0) {
        uint32_t t, p, bad;
        fused_prefixes(a, in, tile_prefix, item_prefix, t, p, bad);
        sh_n_tiles = t;
        sh_particles = p;
        sh_status = bad ? (uint32_t)GF_PLAN_MISMATCH : 0u;
        a.result[0] = 0u;
        a.result[1] = 0u;
    }
    __syncthreads();
    if (sh_status) {
        if (tid == 0) a.result[1] = sh_status;
        return;
    }
    const Batch b = fused_batch(a, in, tile_prefix, item_prefix, sh_n_tiles, sh_particles);
    const uint32_t nt = b.n_tiles;
    // 2. per-tile counts, G tiles at a time
    for (uint32_t t0 = 0; t0 < nt; t0 += G)
        fused_count_tile<P, R>(a, b, t0 + g < nt ? t0 + g : kNoTile, cols, &a.result[1], gsm);
    __syncthreads();
    if (tid == 0) sh_status = __hip_atomic_load(&a.result[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __syncthreads();
    if (sh_status) return;
    // 3. + 4. plan and outputs
    uint32_t Rout;
    const uint32_t st = fused_plan(a, b, reserved, sh_scan, &sh_word, Rout);
    if (st) {
        if (tid == 0) a.result[1] = st;
        return;
    }
    // 5. stable scatter, G tiles at a time
    for (uint32_t t0 = 0; t0 < nt; t0 += G)
        fused_scatter_tile<P, R>(a, b, t0 + g < nt ? t0 + g : kNoTile, cols, &a.result[1], gsm);
    __syncthreads();
    if (tid == 0) a.result[0] = Rout;
}

// The grid form, for batches of many tiles: three launches in stream order, no host round trip.
// Count: a tile per workgroup, errors into result word 2 (a word no other phase writes before the
// plan reads it). Plan: one workgroup; moves that error to the status word, or plans, writes R and
// the outputs, and clears word 2 for the next batch. Scatter: a tile per workgroup, unless the
// status word is set. Every workgroup takes the prefixes of the <= kGridIn input chunks into LDS.
constexpr uint32_t kGridIn = 256u;

template <class P, int R>
__device__ void fused_grid_count(const FusedArgs& a, const Cols& cols, char* smem)
{
    __shared__ uint32_t sh_tp[kGridIn + 1u], sh_ip[kGridIn + 1u], sh_t, sh_p, sh_bad;
    const gf_chunk_desc* in = a.inl ? a.in_arg : a.in;
    if (threadIdx.x == 0) fused_prefixes(a, in, sh_tp, sh_ip, sh_t, sh_p, sh_bad);
    __syncthreads();
    if (sh_bad) return;  // the plan reports it
    const Batch b = fused_batch(a, in, sh_tp, sh_ip, sh_t, sh_p);
    fused_count_tile<P, R>(a, b, blockIdx.x < b.n_tiles ? blockIdx.x : kNoTile, cols, &a.result[2], smem);
}

__device__ inline void fused_grid_plan(const FusedArgs& a)
{
    __shared__ uint32_t sh_t, sh_p, sh_err, sh_word, sh_scan[kMaxWaves];
    const uint32_t tid = threadIdx.x, n = a.n;
    uint32_t* tile_prefix = a.scratch;
    uint32_t* item_prefix = tile_prefix + (n + 1u);
    const gf_chunk_desc* in = a.inl ? a.in_arg : a.in;
    const gf_chunk_handle* reserved = a.inl ? a.res_arg : a.reserved;
    if (tid == 0) {
        uint32_t bad;
        fused_prefixes(a, in, tile_prefix, item_prefix, sh_t, sh_p, bad);
        const uint32_t e = a.result[2];
        sh_err = bad ? (uint32_t)GF_PLAN_MISMATCH : e;
        a.result[0] = 0u;
        a.result[1] = sh_err;
        a.result[2] = 0u;
    }
    __syncthreads();
    if (sh_err) return;
    const Batch b = fused_batch(a, in, tile_prefix, item_prefix, sh_t, sh_p);
    uint32_t Rout;
    const uint32_t st = fused_plan(a, b, reserved, sh_scan, &sh_word, Rout);
    if (tid == 0) {
        if (st) a.result[1] = st;
        else a.result[0] = Rout;
    }
}

template <class P, int R>
__device__ void fused_grid_scatter(const FusedArgs& a, const Cols& cols, char* smem)
{
    __shared__ uint32_t sh_tp[kGridIn + 1u], sh_ip[kGridIn + 1u], sh_t, sh_p, sh_st;
    const gf_chunk_desc* in = a.inl ? a.in_arg : a.in;
    if (threadIdx.x == 0) {
        uint32_t bad;
        fused_prefixes(a, in, sh_tp, sh_ip, sh_t, sh_p, bad);
        sh_st = a.result[1];
    }
    __syncthreads();
    if (sh_st) return;
    const Batch b = fused_batch(a, in, sh_tp, sh_ip, sh_t, sh_p);
    fused_scatter_tile<P, R>(a, b, blockIdx.x < b.n_tiles ? blockIdx.x : kNoTile, cols, &a.result[1], smem);
}

}  // namespace gfd
