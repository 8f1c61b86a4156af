// Device-side building blocks of the MI355X transform-batch path (gfx950, wave64).
//
// Payload layouts (SoA columns in HBM), the keyed Philox stream, the compiled-in functor set
// and the per-tile work functions shared by the plain per-stage kernels and the device-resident
// step engine. A tile is a workgroup of 4 wave64s over T = 256*R consecutive items of one chunk;
// wave w owns the contiguous item range [w*64R, (w+1)*64R) of the tile so that per-wave counters
// give the canonical stable order (bin-major, then batch input order) without a tile-wide sort.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "../../include/graphflow.h"

namespace gfd {

constexpr int kWaves = 4;
constexpr int kThreads = 64 * kWaves;

// ------------------------------------------------------------------------------------------
// RNG: Philox4x32-10 (Salmon et al., SC'11) keyed by the stream key of
// rng/ParticleRng.hpp:57-94 (SplitMix64 finalizer over the packed (id, event, purpose)).
// ------------------------------------------------------------------------------------------
__host__ __device__ inline uint64_t mix64(uint64_t v)
{
    v ^= v >> 30;
    v *= 0xBF58476D1CE4E5B9ull;
    v ^= v >> 27;
    v *= 0x94D049BB133111EBull;
    v ^= v >> 31;
    return v;
}

__host__ __device__ inline uint64_t stream_key(uint64_t seed, uint64_t id, uint64_t event, uint32_t purpose)
{
    uint64_t packed = (id & 0xFFFFFFFFull) | ((event & 0xFFFFull) << 32) | ((uint64_t)(purpose & 0xFFFFu) << 48);
    return mix64(packed ^ mix64(seed));
}

__device__ inline void philox10(uint32_t c[4], uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        uint32_t hi0 = __umulhi(0xD2511F53u, c[0]), lo0 = 0xD2511F53u * c[0];
        uint32_t hi1 = __umulhi(0xCD9E8D57u, c[2]), lo1 = 0xCD9E8D57u * c[2];
        uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = lo1;
        c[2] = n2;
        c[3] = lo0;
    }
}

__device__ inline void stream_words(uint64_t seed, uint64_t id, uint64_t event, uint32_t purpose, uint32_t out[4])
{
    uint64_t k = stream_key(seed, id, event, purpose);
    out[0] = out[1] = out[2] = out[3] = 0u;
    philox10(out, (uint32_t)k, (uint32_t)(k >> 32));
}

__host__ __device__ inline uint32_t fmix32(uint32_t h)
{
    h ^= h >> 16;
    h *= 0x85EBCA6Bu;
    h ^= h >> 13;
    h *= 0xC2B2AE35u;
    h ^= h >> 16;
    return h;
}

__device__ inline float uniform_oc(uint32_t w) { return (float)((w >> 8) + 1u) * (1.0f / 16777216.0f); }
__device__ inline float uniform_co(uint32_t w) { return (float)(w >> 8) * (1.0f / 16777216.0f); }
__device__ inline double uniform_f64(uint32_t w0, uint32_t w1)
{
    return ((double)(w0 >> 5) * 67108864.0 + (double)(w1 >> 6)) * (1.0 / 9007199254740992.0);
}

// ------------------------------------------------------------------------------------------
// Payload layouts. Column i of the pool is cols.c[i]; slot s lives at index s of every column.
// ------------------------------------------------------------------------------------------
struct Cols {
    void* c[9];
};

// Relaxed agent-scope loads: they bypass the per-CU L1, so a workgroup reads what another
// workgroup of the same XCD stored before a barrier, never a stale L1 line.
__device__ inline uint32_t ld_u32(const void* p, size_t i)
{
    return __hip_atomic_load((const uint32_t*)p + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline uint64_t ld_u64(const void* p, size_t i)
{
    return __hip_atomic_load((const unsigned long long*)p + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

struct Particle {
    float w;
    uint32_t id, gen;
    __device__ static Particle load(const Cols& k, size_t s)
    {
        Particle p;
        p.w = ((const float*)k.c[0])[s];
        p.id = ((const uint32_t*)k.c[1])[s];
        p.gen = ((const uint32_t*)k.c[2])[s];
        return p;
    }
    __device__ static Particle load_team(const Cols& k, size_t s)
    {
        Particle p;
        p.w = __uint_as_float(ld_u32(k.c[0], s));
        p.id = ld_u32(k.c[1], s);
        p.gen = ld_u32(k.c[2], s);
        return p;
    }
    __device__ void store(const Cols& k, size_t s) const
    {
        ((float*)k.c[0])[s] = w;
        ((uint32_t*)k.c[1])[s] = id;
        ((uint32_t*)k.c[2])[s] = gen;
    }
};

struct Keyed {
    uint32_t key;
    uint64_t val;
    __device__ static Keyed load(const Cols& k, size_t s)
    {
        Keyed p;
        p.key = ((const uint32_t*)k.c[0])[s];
        p.val = ((const uint64_t*)k.c[1])[s];
        return p;
    }
    __device__ static Keyed load_team(const Cols& k, size_t s)
    {
        Keyed p;
        p.key = ld_u32(k.c[0], s);
        p.val = ld_u64(k.c[1], s);
        return p;
    }
    __device__ void store(const Cols& k, size_t s) const
    {
        ((uint32_t*)k.c[0])[s] = key;
        ((uint64_t*)k.c[1])[s] = val;
    }
};

struct Photon {
    float x, y, z, ux, uy, uz, w;
    uint32_t id, step;
    __device__ static Photon load(const Cols& k, size_t s)
    {
        Photon p;
        p.x = ((const float*)k.c[0])[s];
        p.y = ((const float*)k.c[1])[s];
        p.z = ((const float*)k.c[2])[s];
        p.ux = ((const float*)k.c[3])[s];
        p.uy = ((const float*)k.c[4])[s];
        p.uz = ((const float*)k.c[5])[s];
        p.w = ((const float*)k.c[6])[s];
        p.id = ((const uint32_t*)k.c[7])[s];
        p.step = ((const uint32_t*)k.c[8])[s];
        return p;
    }
    __device__ static Photon load_team(const Cols& k, size_t s)
    {
        Photon p;
        p.x = __uint_as_float(ld_u32(k.c[0], s));
        p.y = __uint_as_float(ld_u32(k.c[1], s));
        p.z = __uint_as_float(ld_u32(k.c[2], s));
        p.ux = __uint_as_float(ld_u32(k.c[3], s));
        p.uy = __uint_as_float(ld_u32(k.c[4], s));
        p.uz = __uint_as_float(ld_u32(k.c[5], s));
        p.w = __uint_as_float(ld_u32(k.c[6], s));
        p.id = ld_u32(k.c[7], s);
        p.step = ld_u32(k.c[8], s);
        return p;
    }
    __device__ void store(const Cols& k, size_t s) const
    {
        ((float*)k.c[0])[s] = x;
        ((float*)k.c[1])[s] = y;
        ((float*)k.c[2])[s] = z;
        ((float*)k.c[3])[s] = ux;
        ((float*)k.c[4])[s] = uy;
        ((float*)k.c[5])[s] = uz;
        ((float*)k.c[6])[s] = w;
        ((uint32_t*)k.c[7])[s] = id;
        ((uint32_t*)k.c[8])[s] = step;
    }
};

// ------------------------------------------------------------------------------------------
// Node, batch and run parameters as the kernels see them.
// ------------------------------------------------------------------------------------------
struct PhotonParams {
    float L, mu_t, albedo, g, roulette_w, roulette_p;
    uint32_t tally_bins;
};

constexpr uint32_t kZipfGuide = 1u << 16;  // guide table intervals of the Zipf source

struct RunParams {
    uint64_t seed;
    uint32_t C;
    const double* zipf;
    uint32_t zipf_values;
    // Zipf guide table: for u in [j, j+1) / kZipfGuide the rank lies in [guide[2j], guide[2j+1]]
    // (bounds taken with a 1e-12 relative margin, so the search below returns the rank of the
    // full-range search); null: search the whole table
    const uint32_t* zipf_guide;
    // rank estimate inside a wide guide interval: the partial sums H(m) = sum_{i<=m} i^-s follow
    // H(m) ~ A + m^(1-s) / (1-s) + m^-s / 2 (Euler-Maclaurin; s = 1: A + ln m + 1 / (2m)), with A
    // fitted to the table's last entry. Only a starting point: the search verifies it.
    double zipf_a, zipf_s;
    PhotonParams ph;
};

// One node's current batch. The plain stages and the step engine fill the same structure.
struct Batch {
    uint32_t active, n_in, n_tiles, n_out;
    uint32_t particles, reserved, tile_items, purpose;
    uint64_t first_id;
    uint64_t digest_base;
    gf_chunk_desc* in;     // cap_in
    uint32_t* tile_prefix; // cap_in + 1: tiles before input chunk c
    uint32_t* item_prefix; // cap_in + 1: items before input chunk c
    gf_chunk_desc* out;    // cap_out
    uint32_t* counts;      // [tile][bin] counts, after planning the flat output start of (tile, bin)
    uint32_t* totals;      // [bins]
};

// A plain stage's batch as a kernel argument. A small batch travels in the launch itself: the
// record and copies of the arrays the kernel only reads (inputs and prefixes, output descriptors,
// planned offsets), so the launch needs no upload. Otherwise `block` names the uploaded record.
constexpr uint32_t kArgIn = 16u, kArgOut = 32u, kArgCounts = 128u;
enum : uint32_t { ARG_IN = 1u, ARG_OUT = 2u, ARG_COUNTS = 4u };

struct BatchArg {
    const Batch* block;  // the record in device memory, or null: `b` and the arrays flagged in `inl`
    Batch b;
    uint32_t inl, pad[3];
    gf_chunk_desc in[kArgIn];
    gf_chunk_desc out[kArgOut];
    uint32_t prefix[2u * (kArgIn + 1u)];  // tile prefix, then item prefix
    uint32_t counts[kArgCounts];
};

// A plain-stage kernel takes its BatchArg as the first parameter, which the code-object ABI places at
// offset 0 of the kernel-argument segment; the batch is read there, so that the arrays it carries are
// addressed in the argument segment itself (a pointer to the parameter object would have the compiler
// copy the whole argument into per-lane scratch memory).
__device__ inline Batch batch_arg()
{
    const BatchArg& a = *(const BatchArg*)__builtin_amdgcn_kernarg_segment_ptr();
    if (a.block) return *a.block;
    Batch b = a.b;
    if (a.inl & ARG_IN) {
        b.in = const_cast<gf_chunk_desc*>(a.in);
        b.tile_prefix = const_cast<uint32_t*>(a.prefix);
        b.item_prefix = b.tile_prefix + (kArgIn + 1u);
    }
    if (a.inl & ARG_OUT) b.out = const_cast<gf_chunk_desc*>(a.out);
    if (a.inl & ARG_COUNTS) b.counts = const_cast<uint32_t*>(a.counts);
    return b;
}

__device__ inline void set_status(uint32_t* status, uint32_t code) { atomicCAS(status, 0u, code); }

__device__ inline uint32_t route_of(uint32_t dist, uint32_t ndest, uint32_t h)
{
    switch (dist) {
    case GF_ROUTE_PARITY: return h & 1u;
    case GF_ROUTE_SKEW_90_10: return (h % 10u == 0u) ? 1u : 0u;
    case GF_ROUTE_ALL_ONE: return 0u;
    default: return h % ndest;
    }
}

// Classifier functors (pure functions of the payload, Concepts.hpp:134-145). classify_p takes the
// scalar parameters a classifier reads (op, bins, u[0], f[0]); the table classifier also needs the
// descriptor and reports kNeedDesc there.
constexpr uint32_t kNeedDesc = 0xFFFFFFFEu;

__device__ inline uint32_t classify_p(uint32_t op, uint32_t bins, uint32_t u0, float f0, const Particle& p)
{
    switch (op) {
    case GF_CLS_WEIGHT_ROUTE:
        if (p.w < f0) return 0u;
        return 1u + route_of(u0, bins - 1u, fmix32(p.id));
    case GF_CLS_ID_MOD:
        if (u0 == 1u) return 0u;
        if (u0 == 2u) return (p.id % 8u == 0u) ? p.id % bins : 0u;
        return p.id % bins;
    case GF_CLS_LUT: return kNeedDesc;
    default: return 0xFFFFFFFFu;
    }
}
__device__ inline uint32_t classify_p(uint32_t op, uint32_t bins, uint32_t, float, const Keyed& p)
{
    if (op == GF_CLS_KEY_BUCKET) return fmix32(p.key) % bins;
    return 0xFFFFFFFFu;
}
__device__ inline uint32_t classify_p(uint32_t op, uint32_t, uint32_t, float, const Photon& p)
{
    if (op == GF_CLS_PHOTON_ALIVE) return p.w > 0.0f ? 0u : 1u;
    return 0xFFFFFFFFu;
}

__device__ inline uint32_t classify(const gf_node_desc& nd, const Particle& p)
{
    const uint32_t c = classify_p(nd.op, nd.bins, nd.u[0], nd.f[0], p);
    if (c != kNeedDesc) return c;
    const uint32_t a = p.id < nd.u[0] ? p.id : nd.u[0] - 1u;
    const uint32_t g = p.gen < nd.u[1] ? p.gen : nd.u[1] - 1u;
    return ((const uint32_t*)nd.ptr[0])[nd.u[3] + a * nd.u[1] + g];
}
__device__ inline uint32_t classify(const gf_node_desc& nd, const Keyed& p) { return classify_p(nd.op, nd.bins, nd.u[0], nd.f[0], p); }
__device__ inline uint32_t classify(const gf_node_desc& nd, const Photon& p) { return classify_p(nd.op, nd.bins, nd.u[0], nd.f[0], p); }

// Partition updaters (UpdateResult, Concepts.hpp:33-39): 0 success, GF_INVALID, GF_OVERFLOW.
__device__ inline uint32_t update(const gf_node_desc& nd, uint32_t bin, Particle& p)
{
    if (nd.upd == GF_UPD_ADD_ON_BIN) {
        if (bin == nd.u[2]) {
            float w = p.w + nd.f[1];
            if (isinf(w)) return GF_OVERFLOW;
            p.w = w;
            p.gen += 1u;
        }
    } else if (nd.upd == GF_UPD_GEN_TO_BIN) {
        p.gen = bin;
    }
    return 0u;
}
__device__ inline uint32_t update(const gf_node_desc&, uint32_t, Keyed&) { return 0u; }
__device__ inline uint32_t update(const gf_node_desc&, uint32_t, Photon&) { return 0u; }

// One photon flight + interaction + roulette + Henyey-Greenstein scatter (fp32), tallying the
// absorbed weight into depth bins.
__device__ inline void photon_step(const RunParams& rp, uint32_t purpose, Photon& p, double* tally)
{
    if (!(p.w > 0.0f)) return;
    uint32_t r[4];
    stream_words(rp.seed, p.id, p.step, purpose, r);
    const PhotonParams& c = rp.ph;
    float s = -logf(uniform_oc(r[0])) / c.mu_t;
    float nx = p.x + s * p.ux, ny = p.y + s * p.uy, nz = p.z + s * p.uz;
    p.step += 1u;
    if (nz < 0.0f || nz > c.L) {
        p.w = 0.0f;
        p.z = nz;
        return;
    }
    p.x = nx;
    p.y = ny;
    p.z = nz;
    uint32_t bin = (uint32_t)(nz / c.L * (float)c.tally_bins);
    if (bin >= c.tally_bins) bin = c.tally_bins - 1u;
    atomicAdd(&tally[bin], (double)(p.w * (1.0f - c.albedo)));
    p.w = p.w * c.albedo;
    if (p.w < c.roulette_w) {
        if (uniform_co(r[3]) < c.roulette_p) {
            p.w = p.w / c.roulette_p;
        } else {
            p.w = 0.0f;
            return;
        }
    }
    float g = c.g;
    float xi = uniform_co(r[1]);
    float ct;
    if (g != 0.0f) {
        const float tmp = (1.0f - g * g) / (1.0f - g + 2.0f * g * xi);
        ct = (1.0f + g * g - tmp * tmp) / (2.0f * g);
    } else {
        ct = 2.0f * xi - 1.0f;  // isotropic
    }
    ct = fminf(1.0f, fmaxf(-1.0f, ct));
    float st = sqrtf(1.0f - ct * ct);
    float phi = 6.2831853071795864f * uniform_co(r[2]);
    float cp = cosf(phi), sp = sinf(phi);
    float ux = p.ux, uy = p.uy, uz = p.uz;
    if (fabsf(uz) > 0.99999f) {
        p.ux = st * cp;
        p.uy = st * sp;
        p.uz = uz > 0.0f ? ct : -ct;
    } else {
        float t = sqrtf(1.0f - uz * uz);
        p.ux = st * (ux * uz * cp - uy * sp) / t + ux * ct;
        p.uy = st * (uy * uz * cp + ux * sp) / t + uy * ct;
        p.uz = -st * cp * t + uz * ct;
    }
}

// Plain transform functors. transform_p applies the transforms that need only (op, f[0]) and
// reports whether it did.
__device__ inline bool transform_p(uint32_t op, float f0, Particle& p)
{
    if (op == GF_XF_SCALE_WEIGHT) {
        p.w = p.w * f0;
        p.gen += 1u;
        return true;
    }
    return false;
}
__device__ inline bool transform_p(uint32_t, float, Keyed&) { return false; }
__device__ inline bool transform_p(uint32_t, float, Photon&) { return false; }

__device__ inline void transform(const gf_node_desc& nd, const RunParams&, uint32_t, Particle& p)
{
    transform_p(nd.op, nd.f[0], p);
}
__device__ inline void transform(const gf_node_desc&, const RunParams&, uint32_t, Keyed&) {}
// The tally a photon step adds into. A flow's runs spread it over kTallyCopies private copies
// (ptr[1], tally_bins doubles each, folded into ptr[0] when the run ends): the 1024 double bins
// are 64 cache lines, and every step's atomic add would otherwise queue behind the others on
// them. Waves take the copies round robin. Without copies (a stage made through the C ABI),
// the step adds into ptr[0] itself.
constexpr uint32_t kTallyCopies = 128u;

__device__ inline double* photon_tally(const gf_node_desc& nd, uint32_t bins)
{
    if (!nd.ptr[1]) return (double*)nd.ptr[0];
    const uint32_t w = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    return (double*)nd.ptr[1] + (size_t)(w & (kTallyCopies - 1u)) * bins;
}

__device__ inline void transform(const gf_node_desc& nd, const RunParams& rp, uint32_t purpose, Photon& p)
{
    if (nd.op == GF_XF_PHOTON_STEP) photon_step(rp, purpose, p, photon_tally(nd, rp.ph.tally_bins));
}

// Zipf rank of u: the smallest r with cdf[r] > u * cdf[n - 1] (n - 1 if none), the full-table
// binary search's answer. The guide interval [lo, hi] holds it; a wide interval (the tail, where
// a guide slot spans many ranks) is first probed at the analytic estimate with one window of four
// consecutive entries, which settles the rank in one round trip when the estimate is within one
// rank, and narrows the interval otherwise.
__device__ inline uint32_t zipf_rank(const RunParams& rp, double u)
{
    const double* cdf = rp.zipf;
    const double target = u * cdf[rp.zipf_values - 1u];
    uint32_t lo = 0, hi = rp.zipf_values - 1u;
    if (rp.zipf_guide) {
        const uint32_t j = (uint32_t)(u * (double)kZipfGuide);  // u < 1: j < kZipfGuide
        lo = rp.zipf_guide[2u * j];
        hi = rp.zipf_guide[2u * j + 1u];
    }
    if (hi - lo >= 4u) {
        // continuous solution m* of H(m) = target (two fixed-point steps); rank floor(m*) (0-based)
        const double s = rp.zipf_s;
        double m;
        if (s == 1.0) {
            m = exp(target - rp.zipf_a);
            m = exp(target - rp.zipf_a - 0.5 / m);
        } else {
            const double e = 1.0 / (1.0 - s);
            m = pow((1.0 - s) * (target - rp.zipf_a), e);
            m = pow((1.0 - s) * (target - rp.zipf_a - 0.5 * pow(m, -s)), e);
        }
        // window [g - 1, g + 2] inside [lo, hi]: g in [lo + 1, hi - 2] (a NaN estimate compares false)
        double gd = floor(m);
        if (!(gd >= (double)(lo + 1u))) gd = (double)(lo + 1u);
        if (!(gd <= (double)(hi - 2u))) gd = (double)(hi - 2u);
        const uint32_t g = (uint32_t)gd;
        const double c0 = cdf[g - 1u], c1 = cdf[g], c2 = cdf[g + 1u], c3 = cdf[g + 2u];
        if (c0 > target) hi = g - 1u;  // rank <= g - 1
        else if (c1 > target) return g;
        else if (c2 > target) return g + 1u;
        else if (c3 > target) return g + 2u;
        else lo = g + 3u;  // rank >= g + 3
    }
    while (lo < hi) {
        uint32_t mid = lo + (hi - lo) / 2u;
        if (cdf[mid] > target) hi = mid;
        else lo = mid + 1u;
    }
    return lo;
}

// Zipf ranks of NQ uniforms with their loads batched: the NQ guide intervals in one round trip,
// then one window of four consecutive entries per item in a second (at the analytic estimate in a
// wide interval, at the interval itself in a narrow one), and the rare miss finishes with a binary
// search. Same answers as zipf_rank.
template <int NQ>
__device__ inline void zipf_ranks(const RunParams& rp, const double (&u)[NQ], uint32_t (&rank)[NQ])
{
    const double* cdf = rp.zipf;
    const uint32_t n = rp.zipf_values;
    if (n < 4u || !rp.zipf_guide) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) rank[q] = zipf_rank(rp, u[q]);
        return;
    }
    const double total = cdf[n - 1u];
    double target[NQ];
    uint32_t lo[NQ], hi[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        target[q] = u[q] * total;
        const uint32_t j = (uint32_t)(u[q] * (double)kZipfGuide);  // u < 1: j < kZipfGuide
        lo[q] = rp.zipf_guide[2u * j];
        hi[q] = rp.zipf_guide[2u * j + 1u];
    }
    uint32_t base[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        if (hi[q] - lo[q] >= 4u) {
            const double s = rp.zipf_s;
            double m;
            if (s == 1.0) {
                m = exp(target[q] - rp.zipf_a);
                m = exp(target[q] - rp.zipf_a - 0.5 / m);
            } else {
                const double e = 1.0 / (1.0 - s);
                m = pow((1.0 - s) * (target[q] - rp.zipf_a), e);
                m = pow((1.0 - s) * (target[q] - rp.zipf_a - 0.5 * pow(m, -s)), e);
            }
            double gd = floor(m);
            if (!(gd >= (double)(lo[q] + 1u))) gd = (double)(lo[q] + 1u);
            if (!(gd <= (double)(hi[q] - 2u))) gd = (double)(hi[q] - 2u);
            base[q] = (uint32_t)gd - 1u;  // window [g - 1, g + 2] inside [lo, hi]
        } else {
            base[q] = lo[q] < n - 4u ? lo[q] : n - 4u;  // window holds [lo, hi]
        }
    }
    double c[NQ][4];
#pragma unroll
    for (int q = 0; q < NQ; ++q)
#pragma unroll
        for (int x = 0; x < 4; ++x) c[q][x] = cdf[base[q] + (uint32_t)x];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        // every entry before the rank is <= target and the rank's entry (but the last) is above it
        uint32_t k = 0;
#pragma unroll
        for (int x = 0; x < 4; ++x) k += c[q][x] <= target[q] ? 1u : 0u;
        uint32_t a = lo[q], h = hi[q];
        if (k > 0u && k < 4u) {
            a = h = base[q] + k;
        } else if (k == 0u) {
            if (base[q] <= a) h = a;  // the rank is at most base <= lo
            else h = base[q];
        } else if (base[q] + 4u > a) {
            a = base[q] + 4u < h ? base[q] + 4u : h;
        }
        while (a < h) {
            const uint32_t mid = a + (h - a) / 2u;
            if (cdf[mid] > target[q]) h = mid;
            else a = mid + 1u;
        }
        rank[q] = a;
    }
}

// Source functors.
__device__ inline void source(const gf_node_desc&, const RunParams& rp, uint32_t purpose, uint64_t id, Particle& p)
{
    uint32_t r[4];
    stream_words(rp.seed, id, 0, purpose, r);
    p.w = uniform_oc(r[0]);
    p.id = (uint32_t)id;
    p.gen = 0u;
}
__device__ inline void source(const gf_node_desc&, const RunParams& rp, uint32_t purpose, uint64_t id, Keyed& p)
{
    uint32_t r[4];
    stream_words(rp.seed, id, 0, purpose, r);
    const double u = uniform_f64(r[0], r[1]);
    p.key = zipf_rank(rp, u);
    p.val = id;
}
__device__ inline void source(const gf_node_desc&, const RunParams&, uint32_t, uint64_t id, Photon& p)
{
    p.x = p.y = p.z = 0.0f;
    p.ux = p.uy = 0.0f;
    p.uz = 1.0f;
    p.w = 1.0f;
    p.id = (uint32_t)id;
    p.step = 0u;
}

__device__ inline uint64_t keyed_hash(const Keyed& p)
{
    return mix64(((uint64_t)p.key << 32) ^ mix64(p.val + 0x9E3779B97F4A7C15ull));
}

// ------------------------------------------------------------------------------------------
// Wave helpers
// ------------------------------------------------------------------------------------------
__device__ inline uint32_t lane_id() { return threadIdx.x & 63u; }
// Thread index inside the 256-thread tile group (a workgroup may hold several tile groups).
__device__ inline uint32_t ltid() { return threadIdx.x & (uint32_t)(kThreads - 1); }
// Tile index meaning "this group has no tile in this round": every __syncthreads is still executed.
constexpr uint32_t kNoTile = 0xFFFFFFFFu;
__device__ inline uint64_t lanemask_lt() { return (1ull << lane_id()) - 1ull; }

// Flat output position -> (chunk, slot) with a shift and a mask when the chunk capacity C is a power
// of two (every workload's C), a division otherwise; `sh` = cap_shift(C), uniform per launch.
__host__ __device__ inline uint32_t cap_shift(uint32_t C) { return (C & (C - 1u)) == 0u ? (uint32_t)__builtin_ctz(C) : 32u; }
__device__ inline uint32_t cdiv(uint32_t f, uint32_t C, uint32_t sh) { return sh < 32u ? f >> sh : f / C; }
__device__ inline uint32_t cmod(uint32_t f, uint32_t C, uint32_t sh) { return sh < 32u ? f & (C - 1u) : f % C; }

// Lanes with the same bin (the "peers" of each lane); iterations = distinct bins in the wave.
__device__ inline uint64_t peer_mask(bool valid, uint32_t bin)
{
    uint64_t active = __ballot(valid);
    uint64_t peers = 0;
    while (active) {
        int leader = __ffsll((unsigned long long)active) - 1;
        uint32_t lb = (uint32_t)__shfl((int)bin, leader);
        uint64_t m = __ballot(valid && bin == lb);
        if (valid && bin == lb) peers = m;
        active &= ~m;
    }
    return peers;
}

// The same set from one ballot per bin bit (bins < 2^nbits): a lane's peers are the valid lanes
// that agree with it on every bit. Costs nbits ballots however many bins the wave holds, so wide
// partitions use it.
__device__ inline uint64_t peer_mask_bits(bool valid, uint32_t bin, uint32_t nbits)
{
    uint64_t peers = __ballot(valid);
    for (uint32_t j = 0; j < nbits; ++j) {
        const bool bit = ((bin >> j) & 1u) != 0u;
        const uint64_t m = __ballot(valid && bit);
        peers &= bit ? m : ~m;
    }
    return valid ? peers : 0ull;
}

// Chunk of tile t: largest c with prefix[c] <= t (prefix has n+1 entries).
__device__ inline uint32_t find_chunk(const uint32_t* prefix, uint32_t n, uint32_t t)
{
    // Chunk c holds tile t iff prefix[c] <= t < prefix[c + 1]. Every chunk of an input list holds at
    // least one item, so at least one tile: the chunks before c hold c tiles plus at most
    // X = prefix[n] - n more, hence t - X <= c <= t. Common case (X tiles all after t, e.g. the many
    // small per-bucket chunks a sink takes): c = t, from loads issued together with prefix[n].
    const uint32_t tot = prefix[n];
    const bool in = t < n;
    const uint32_t pt = in ? prefix[t] : 0u, pt1 = in ? prefix[t + 1u] : 0u;
    if (in && pt == t && pt1 > t) return t;
    const uint32_t X = tot - n;
    uint32_t lo = t > X ? t - X : 0u, hi = t + 1u < n ? t + 1u : n;  // invariant: prefix[lo] <= t < prefix[hi]
    if (lo >= n || prefix[lo] > t || (hi < n && prefix[hi] <= t)) {
        lo = 0u;  // the window's ends do not hold the invariant: search the whole list
        hi = n;
    }
    while (hi - lo > 1u) {
        uint32_t mid = (lo + hi) >> 1;
        if (prefix[mid] <= t) lo = mid;
        else hi = mid;
    }
    return lo;
}

struct TileSpan {
    uint32_t chunk;  // position in the batch's input list
    uint32_t offset; // first item of the tile inside the chunk
    uint32_t count;
    size_t base;     // first slot
};

__device__ inline TileSpan tile_span(const Batch& b, const gf_chunk_desc* list, uint32_t n, uint32_t t, uint32_t C)
{
    TileSpan s;
    if (t == kNoTile) {
        s.chunk = s.offset = s.count = 0u;
        s.base = 0;
        return s;
    }
    s.chunk = find_chunk(b.tile_prefix, n, t);
    gf_chunk_desc d = list[s.chunk];
    s.offset = (t - b.tile_prefix[s.chunk]) * b.tile_items;
    uint32_t rem = d.fill - s.offset;
    s.count = rem < b.tile_items ? rem : b.tile_items;
    s.base = (size_t)d.index * C + s.offset;
    return s;
}

// ------------------------------------------------------------------------------------------
// Tile functions
// ------------------------------------------------------------------------------------------

// Partition classify of one tile: per-tile bin counts (Execution.hpp:869-954 semantics; the
// tile-local histogram lives in LDS and is fed by wave-aggregated increments).
template <class P, int R>
__device__ void count_tile(const gf_node_desc& nd, const Batch& b, uint32_t t, const Cols& cols, uint32_t C,
                           uint32_t* status, char* smem)
{
    const uint32_t B = nd.bins, tid = ltid(), w = tid >> 6, lane = lane_id();
    uint32_t* tcnt = (uint32_t*)smem;
    for (uint32_t i = tid; i < B; i += kThreads) tcnt[i] = 0u;
    __syncthreads();
    TileSpan s = tile_span(b, b.in, b.n_in, t, C);
#pragma unroll
    for (int r = 0; r < R; ++r) {
        uint32_t i = w * 64u * R + (uint32_t)r * 64u + lane;
        bool v = i < s.count;
        uint32_t bin = 0u;
        if (v) {
            P p = P::load(cols, s.base + i);
            bin = classify(nd, p);
            if (bin >= B) {
                set_status(status, GF_INVALID);
                v = false;
            }
        }
        if (B >= 64u) {
            // wide partitions: a wave's items mostly fall into distinct bins, so one LDS atomic
            // per item beats aggregating per distinct bin
            if (v) atomicAdd(&tcnt[bin], 1u);
        } else {
            uint64_t peers = peer_mask(v, bin);
            if (v && (peers & lanemask_lt()) == 0ull) atomicAdd(&tcnt[bin], (uint32_t)__popcll(peers));
        }
    }
    __syncthreads();
    if (t != kNoTile) {
        uint32_t* dst = b.counts + (size_t)t * B;
        for (uint32_t i = tid; i < B; i += kThreads) dst[i] = tcnt[i];
    }
    __syncthreads();
}

// ------------------------------------------------------------------------------------------
// Hash bin path (BinPath::hash, core/BinPath.hpp): partitions wider than the dense LDS counters
// count and rank a tile's items through an open-addressing table of E = 2^k slots in LDS
// (graph/BinTable.hpp:69-137: home slot (bin * 0x9E3779B1) & (E - 1), linear probing, a bin's
// key set once by compare-and-swap). LDS grows with the tile, not with the partition width, so a
// partition may have up to 65536 bins. A tile with more distinct bins than slots fails its batch
// before any output is planned (Execution.hpp:1207-1307).
// ------------------------------------------------------------------------------------------
constexpr uint32_t kEmptyKey = 0xFFFFFFFFu;
constexpr uint32_t kNoSlot = 0xFFFFFFFFu;
constexpr uint32_t kStatusTableFull = 17u;  // device status word: a tile's bin table is full (-> GF_INVALID)

__device__ inline uint32_t hash_home(uint32_t bin, uint32_t mask) { return (bin * 0x9E3779B1u) & mask; }

// Slot of `bin`, inserting it; kNoSlot when every slot holds another bin.
__device__ inline uint32_t hash_insert(uint32_t* keys, uint32_t mask, uint32_t bin)
{
    uint32_t slot = hash_home(bin, mask);
    for (uint32_t p = 0; p <= mask; ++p) {
        const uint32_t old = atomicCAS(&keys[slot], kEmptyKey, bin);
        if (old == kEmptyKey || old == bin) return slot;
        slot = (slot + 1u) & mask;
    }
    return kNoSlot;
}

// Slot of `bin` in a built table, kNoSlot when absent.
__device__ inline uint32_t hash_find(const uint32_t* keys, uint32_t mask, uint32_t bin)
{
    uint32_t slot = hash_home(bin, mask);
    for (uint32_t p = 0; p <= mask; ++p) {
        const uint32_t k = keys[slot];
        if (k == bin) return slot;
        if (k == kEmptyKey) return kNoSlot;
        slot = (slot + 1u) & mask;
    }
    return kNoSlot;
}

// LDS bytes of count_tile_hash / scatter_tile_hash with E slots.
__host__ __device__ inline size_t count_hash_smem(uint32_t E) { return (size_t)E * 8u + 16u; }
__host__ __device__ inline size_t scatter_hash_smem(uint32_t E) { return (size_t)E * 4u * (kWaves + 2u) + 16u; }

// Hash-path classify of one tile: the tile's bins counted in the LDS table, then the tile's dense
// count row written from it (the plan scans [tile][bin] rows whatever the bin path).
template <class P, int R>
__device__ void count_tile_hash(const gf_node_desc& nd, const Batch& b, uint32_t t, const Cols& cols, uint32_t C,
                                uint32_t* status, char* smem, uint32_t E)
{
    const uint32_t B = nd.bins, tid = ltid(), w = tid >> 6, lane = lane_id(), mask = E - 1u;
    uint32_t* keys = (uint32_t*)smem;
    uint32_t* cnt = keys + E;
    uint32_t* full = cnt + E;
    for (uint32_t i = tid; i < E; i += kThreads) {
        keys[i] = kEmptyKey;
        cnt[i] = 0u;
    }
    if (tid == 0) *full = 0u;
    __syncthreads();
    TileSpan s = tile_span(b, b.in, b.n_in, t, C);
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const uint32_t i = w * 64u * R + (uint32_t)r * 64u + lane;
        if (i < s.count) {
            P p = P::load(cols, s.base + i);
            const uint32_t bin = classify(nd, p);
            if (bin >= B) {
                set_status(status, GF_INVALID);
            } else {
                const uint32_t slot = hash_insert(keys, mask, bin);
                if (slot == kNoSlot) *full = 1u;
                else atomicAdd(&cnt[slot], 1u);
            }
        }
    }
    __syncthreads();
    if (tid == 0 && *full) set_status(status, kStatusTableFull);
    if (t != kNoTile) {
        uint32_t* dst = b.counts + (size_t)t * B;
        for (uint32_t bin = tid; bin < B; bin += kThreads) {
            const uint32_t slot = hash_find(keys, mask, bin);
            dst[bin] = slot == kNoSlot ? 0u : cnt[slot];
        }
    }
    __syncthreads();
}

// Hash-path stable scatter of one tile: scatter_tile's ranking with the per-wave running counts
// kept per table slot instead of per bin.
template <class P, int R>
__device__ void scatter_tile_hash(const gf_node_desc& nd, const Batch& b, uint32_t t, const Cols& cols, uint32_t C,
                                  uint32_t* status, char* smem, uint32_t E, const gf_node_desc* all = nullptr,
                                  const int32_t* fuse = nullptr, const RunParams* rp = nullptr,
                                  const uint32_t* bin_base = nullptr)
{
    const uint32_t csh = cap_shift(C);
    const uint32_t B = nd.bins, tid = ltid(), w = tid >> 6, lane = lane_id(), mask = E - 1u;
    uint32_t* keys = (uint32_t*)smem;
    uint32_t* wcnt = keys + E;            // [kWaves][E] running, then exclusive wave prefix
    uint32_t* full = wcnt + kWaves * E;
    for (uint32_t i = tid; i < E; i += kThreads) keys[i] = kEmptyKey;
    for (uint32_t i = tid; i < kWaves * E; i += kThreads) wcnt[i] = 0u;
    if (tid == 0) *full = 0u;
    __syncthreads();
    TileSpan s = tile_span(b, b.in, b.n_in, t, C);
    P rec[R];
    uint32_t bin[R], slot[R], rank[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const uint32_t i = w * 64u * R + (uint32_t)r * 64u + lane;
        bool v = i < s.count;
        bin[r] = 0xFFFFFFFFu;
        slot[r] = kNoSlot;
        if (v) {
            rec[r] = P::load(cols, s.base + i);
            bin[r] = classify(nd, rec[r]);
            if (bin[r] >= B) {
                set_status(status, GF_INVALID);
                v = false;
            } else {
                slot[r] = hash_insert(keys, mask, bin[r]);
                if (slot[r] == kNoSlot) {
                    *full = 1u;
                    v = false;
                }
            }
            if (!v) bin[r] = 0xFFFFFFFFu;
        }
        const uint64_t peers = peer_mask_bits(v, slot[r], 32u - (uint32_t)__clz((int)(E - 1u)));
        const uint32_t cur = v ? wcnt[w * E + slot[r]] : 0u;
        rank[r] = cur + (uint32_t)__popcll(peers & lanemask_lt());
        if (v && (peers & lanemask_lt()) == 0ull) wcnt[w * E + slot[r]] = cur + (uint32_t)__popcll(peers);
    }
    __syncthreads();
    if (tid == 0 && *full) set_status(status, kStatusTableFull);
    for (uint32_t k = tid; k < E; k += kThreads) {
        uint32_t acc = 0;
        for (int ww = 0; ww < kWaves; ++ww) {
            const uint32_t x = wcnt[ww * E + k];
            wcnt[ww * E + k] = acc;
            acc += x;
        }
    }
    __syncthreads();
    const uint32_t* toff = b.counts + (size_t)(t == kNoTile ? 0u : t) * B;
    // flat positions, then output chunks, then the stores: every round's loads are issued before the
    // first store (the compiler cannot move a load of the plan past a store to the pool)
    uint32_t f[R], ci[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const uint32_t k = bin[r] == 0xFFFFFFFFu ? 0u : bin[r];
        f[r] = toff[k] + (bin_base ? bin_base[k] * C : 0u) + (bin[r] == 0xFFFFFFFFu ? 0u : wcnt[w * E + slot[r]]) + rank[r];
    }
#pragma unroll
    for (int r = 0; r < R; ++r) ci[r] = bin[r] == 0xFFFFFFFFu ? 0u : b.out[cdiv(f[r], C, csh)].index;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        if (bin[r] == 0xFFFFFFFFu) continue;
        const uint32_t k = bin[r];
        P p = rec[r];
        const uint32_t u = update(nd, k, p);
        if (u) set_status(status, u);
        if (fuse && fuse[k] >= 0) transform(all[fuse[k] & 0xFFFF], *rp, (uint32_t)fuse[k] >> 16, p);
        p.store(cols, (size_t)ci[r] * C + cmod(f[r], C, csh));
    }
    __syncthreads();
}

// LDS bytes of scatter_tile: the per-wave running counts (16 bits: a wave ranks at most 64 R <= 1024
// items of a tile of at most 4096), padded to 16 bytes; with staging also the tile totals, their
// prefix and the staged records.
__host__ __device__ inline size_t scatter_wcnt_bytes(uint32_t bins) { return ((size_t)kWaves * bins * 2u + 15u) & ~(size_t)15u; }
__host__ __device__ inline size_t scatter_smem(uint32_t bins, uint32_t tile_items, uint32_t rec_bytes, bool stage)
{
    if (!stage) return scatter_wcnt_bytes(bins);
    return scatter_wcnt_bytes(bins) + (size_t)bins * 20u + (size_t)tile_items * (rec_bytes + 4u);
}

// Stable scatter of one tile (Execution.hpp:1096-1187 / :1327-1502 semantics with the canonical
// order): the records are held in registers between the ranking and the write; with staging the
// tile is reordered in LDS so consecutive lanes write consecutive output slots of one bin run.
// b.counts[t*B + bin] holds the flat output position of the tile's first item of `bin`, less
// bin_base[bin] * C when bin_base is given (the device engines keep each bin's first output chunk
// beside the bin totals instead of adding it to every tile's entry); flat position f lives at
// slot out[f / C].index * C + f % C.
template <class P, int R>
__device__ void scatter_tile_2p(const gf_node_desc& nd, const Batch& b, uint32_t t, const Cols& cols, uint32_t C,
                                uint32_t* status, char* smem, const gf_node_desc* all, const int32_t* fuse,
                                const RunParams* rp, const uint32_t* bin_base);

template <class P, int R, bool Stage>
__device__ void scatter_tile(const gf_node_desc& nd, const Batch& b, uint32_t t, const Cols& cols, uint32_t C,
                             uint32_t* status, char* smem, const gf_node_desc* all = nullptr,
                             const int32_t* fuse = nullptr, const RunParams* rp = nullptr,
                             const uint32_t* bin_base = nullptr)
{
    const uint32_t csh = cap_shift(C);
    if constexpr (R > 4) {
        // tiles of more than 1024 items: never staged (the staged records would exceed 64 KB of LDS)
        scatter_tile_2p<P, R>(nd, b, t, cols, C, status, smem, all, fuse, rp, bin_base);
        return;
    }
    const uint32_t B = nd.bins, tid = ltid(), w = tid >> 6, lane = lane_id();
    uint16_t* wcnt = (uint16_t*)smem;                                    // [kWaves][B] running, then exclusive wave prefix
    uint32_t* ttot = (uint32_t*)(smem + scatter_wcnt_bytes(B));          // Stage: [B] tile totals
    uint32_t* tstart = ttot + B;                                         // Stage: [B] exclusive prefix of ttot over bins
    P* stage = (P*)(tstart + B);                                         // [T]
    uint32_t* flat = (uint32_t*)(stage + (Stage ? b.tile_items : 0u));
    for (uint32_t i = tid; i < (uint32_t)(scatter_wcnt_bytes(B) / 4u); i += kThreads) ((uint32_t*)smem)[i] = 0u;
    __syncthreads();
    TileSpan s = tile_span(b, b.in, b.n_in, t, C);
    P rec[R];
    uint32_t bin[R], rank[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        uint32_t i = w * 64u * R + (uint32_t)r * 64u + lane;
        bool v = i < s.count;
        bin[r] = 0xFFFFFFFFu;
        if (v) {
            rec[r] = P::load(cols, s.base + i);
            bin[r] = classify(nd, rec[r]);
            if (bin[r] >= B) {
                set_status(status, GF_INVALID);
                v = false;
                bin[r] = 0xFFFFFFFFu;
            }
        }
        uint64_t peers = B > 64u ? peer_mask_bits(v, bin[r], 32u - (uint32_t)__clz((int)(B - 1u))) : peer_mask(v, bin[r]);
        uint32_t cur = v ? (uint32_t)wcnt[w * B + bin[r]] : 0u;
        rank[r] = cur + (uint32_t)__popcll(peers & lanemask_lt());
        // every lane of the wave has read `cur` before a leader stores the new running count
        if (v && (peers & lanemask_lt()) == 0ull) wcnt[w * B + bin[r]] = (uint16_t)(cur + (uint32_t)__popcll(peers));
    }
    __syncthreads();
    for (uint32_t k = tid; k < B; k += kThreads) {
        uint32_t acc = 0;
        for (int ww = 0; ww < kWaves; ++ww) {
            uint32_t x = wcnt[ww * B + k];
            wcnt[ww * B + k] = (uint16_t)acc;
            acc += x;
        }
        if constexpr (Stage) ttot[k] = acc;
    }
    __syncthreads();
    const uint32_t* toff = b.counts + (size_t)(t == kNoTile ? 0u : t) * B;
    if constexpr (Stage) {
        // With C >= the tile, a bin's run of the tile covers at most two output chunks: their
        // indices (and the run's flat start) are loaded once per bin into LDS, and each staged item
        // keeps its absolute slot, so no item waits on a load of the plan or of a descriptor.
        uint32_t* tb = flat + b.tile_items;  // [B] flat start of the tile's run of each bin
        uint32_t* cx = tb + B;               // [B][2] output chunk indices the run covers
        const bool direct = C >= b.tile_items;
        if (tid == 0) {
            uint32_t acc = 0;
            for (uint32_t k = 0; k < B; ++k) {
                tstart[k] = acc;
                acc += ttot[k];
            }
        }
        if (direct)
            for (uint32_t k = tid; k < B; k += kThreads) {
                if (!ttot[k]) continue;
                const uint32_t f0 = toff[k] + (bin_base ? bin_base[k] * C : 0u), c0 = cdiv(f0, C, csh);
                tb[k] = f0;
                cx[2u * k] = b.out[c0].index;
                cx[2u * k + 1u] = cdiv(f0 + ttot[k] - 1u, C, csh) > c0 ? b.out[c0 + 1u].index : 0u;
            }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (bin[r] == 0xFFFFFFFFu) continue;
            uint32_t k = bin[r], within = wcnt[w * B + k] + rank[r];
            uint32_t u = update(nd, k, rec[r]);
            if (u) set_status(status, u);
            if (fuse && fuse[k] >= 0) transform(all[fuse[k] & 0xFFFF], *rp, (uint32_t)fuse[k] >> 16, rec[r]);
            uint32_t p = tstart[k] + within;
            stage[p] = rec[r];
            if (direct) {
                const uint32_t f = tb[k] + within;
                flat[p] = cx[2u * k + (cdiv(f, C, csh) - cdiv(tb[k], C, csh))] * C + cmod(f, C, csh);  // slots stay below 2^32 (S * C)
            } else {
                flat[p] = toff[k] + (bin_base ? bin_base[k] * C : 0u) + within;
            }
        }
        __syncthreads();
        if (direct) {
            for (uint32_t j = tid; j < s.count; j += kThreads) stage[j].store(cols, flat[j]);
        } else {
            for (uint32_t j = tid; j < s.count; j += kThreads) {
                uint32_t f = flat[j];
                size_t slot = (size_t)b.out[cdiv(f, C, csh)].index * C + cmod(f, C, csh);
                stage[j].store(cols, slot);
            }
        }
        __syncthreads();
    } else {
        // flat positions, then output chunks, then the stores (all loads before the first store)
        uint32_t f[R], ci[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const uint32_t k = bin[r] == 0xFFFFFFFFu ? 0u : bin[r];
            f[r] = toff[k] + (bin_base ? bin_base[k] * C : 0u) + (uint32_t)wcnt[w * B + k] + rank[r];
        }
#pragma unroll
        for (int r = 0; r < R; ++r) ci[r] = bin[r] == 0xFFFFFFFFu ? 0u : b.out[cdiv(f[r], C, csh)].index;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (bin[r] == 0xFFFFFFFFu) continue;
            uint32_t k = bin[r];
            P p = rec[r];
            uint32_t u = update(nd, k, p);
            if (u) set_status(status, u);
            if (fuse && fuse[k] >= 0) transform(all[fuse[k] & 0xFFFF], *rp, (uint32_t)fuse[k] >> 16, p);
            p.store(cols, (size_t)ci[r] * C + cmod(f[r], C, csh));
        }
        __syncthreads();
    }
}

// scatter_tile for tiles of more than 4 rounds per thread (unstaged), in two passes over the tile:
// the ranking pass keeps each item's bin and in-wave rank in one register (16 bits each: dense
// partitions have at most 4096 bins, a wave ranks at most 64 R items), the placing pass loads the
// records again (L2 hits) four rounds at a time. Registers do not grow with the records of R rounds.
template <class P, int R>
__device__ void scatter_tile_2p(const gf_node_desc& nd, const Batch& b, uint32_t t, const Cols& cols, uint32_t C,
                                uint32_t* status, char* smem, const gf_node_desc* all, const int32_t* fuse,
                                const RunParams* rp, const uint32_t* bin_base)
{
    const uint32_t csh = cap_shift(C);
    static_assert(R % 4 == 0, "rounds in groups of four");
    uint32_t tid = ltid();
    // the round offsets derive from the thread index: made opaque here, so they are recomputed per
    // tile instead of being hoisted out of the caller's tile loop into R live 64-bit registers
    asm volatile("" : "+v"(tid));
    const uint32_t B = nd.bins, w = tid >> 6, lane = tid & 63u;
    uint16_t* wcnt = (uint16_t*)smem;  // [kWaves][B] running, then exclusive wave prefix
    for (uint32_t i = tid; i < (uint32_t)(scatter_wcnt_bytes(B) / 4u); i += kThreads) ((uint32_t*)smem)[i] = 0u;
    __syncthreads();
    TileSpan s = tile_span(b, b.in, b.n_in, t, C);
    uint32_t br[R];  // bin | rank << 16; 0xFFFFFFFF: no item
    const uint32_t nbits = 32u - (uint32_t)__clz((int)(B > 1u ? B - 1u : 1u));
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const uint32_t i = w * 64u * R + (uint32_t)r * 64u + lane;
        bool v = i < s.count;
        uint32_t bin = 0u;
        if (v) {
            bin = classify(nd, P::load(cols, s.base + i));
            if (bin >= B) {
                set_status(status, GF_INVALID);
                v = false;
            }
        }
        const uint64_t peers = B > 64u ? peer_mask_bits(v, bin, nbits) : peer_mask(v, bin);
        const uint32_t cur = v ? (uint32_t)wcnt[w * B + bin] : 0u;
        const uint32_t rank = cur + (uint32_t)__popcll(peers & lanemask_lt());
        // every lane of the wave has read `cur` before a leader stores the new running count
        if (v && (peers & lanemask_lt()) == 0ull) wcnt[w * B + bin] = (uint16_t)(cur + (uint32_t)__popcll(peers));
        br[r] = v ? (bin | rank << 16) : 0xFFFFFFFFu;
    }
    __syncthreads();
    for (uint32_t k = tid; k < B; k += kThreads) {
        uint32_t acc = 0;
        for (int ww = 0; ww < kWaves; ++ww) {
            const uint32_t x = wcnt[ww * B + k];
            wcnt[ww * B + k] = (uint16_t)acc;
            acc += x;
        }
    }
    __syncthreads();
    const uint32_t* toff = b.counts + (size_t)(t == kNoTile ? 0u : t) * B;
    // the placing pass with and without the bins' first output chunks (a uniform choice made once,
    // so no per-item branch waits on a load)
    auto place = [&](auto has_base) {
        constexpr bool HB = decltype(has_base)::value;
#pragma unroll
        for (int r0 = 0; r0 < R; r0 += 4) {
            uint32_t f[4], ci[4];
            P rec[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const uint32_t x = br[r0 + u], k = x == 0xFFFFFFFFu ? 0u : x & 0xFFFFu;
                f[u] = toff[k] + (uint32_t)wcnt[w * B + k] + (x >> 16);
                if constexpr (HB) f[u] += bin_base[k] * C;
                if (x != 0xFFFFFFFFu) rec[u] = P::load(cols, s.base + w * 64u * R + (uint32_t)(r0 + u) * 64u + lane);
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) ci[u] = br[r0 + u] == 0xFFFFFFFFu ? 0u : b.out[cdiv(f[u], C, csh)].index;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const uint32_t x = br[r0 + u];
                if (x == 0xFFFFFFFFu) continue;
                const uint32_t k = x & 0xFFFFu;
                P p = rec[u];
                const uint32_t st = update(nd, k, p);
                if (st) set_status(status, st);
                if (fuse && fuse[k] >= 0) transform(all[fuse[k] & 0xFFFF], *rp, (uint32_t)fuse[k] >> 16, p);
                p.store(cols, (size_t)ci[u] * C + cmod(f[u], C, csh));
            }
            __builtin_amdgcn_sched_barrier(0);  // one group's records live at a time
        }
    };
    if (bin_base) place(std::true_type{});
    else place(std::false_type{});
    __syncthreads();
}

// Out-of-place transform of one tile (TransformKernel, Execution.hpp:708-776): dense
// concatenation, global position = items before the chunk + offset.
template <class P, int R>
__device__ void transform_tile(const gf_node_desc& nd, const Batch& b, uint32_t t, const Cols& cols, uint32_t C,
                               const RunParams& rp)
{
    const uint32_t csh = cap_shift(C);
    TileSpan s = tile_span(b, b.in, b.n_in, t, C);
    const uint32_t pos0 = s.count ? b.item_prefix[s.chunk] + s.offset : 0u;
    for (uint32_t i = ltid(); i < s.count; i += kThreads) {
        P p = P::load(cols, s.base + i);
        transform(nd, rp, b.purpose, p);
        uint32_t g = pos0 + i;
        p.store(cols, (size_t)b.out[cdiv(g, C, csh)].index * C + cmod(g, C, csh));
    }
}

// Source of one tile of the admitted chunks (SourceKernel, Execution.hpp:664-692): item g of the
// lane's batch (chunks of C items but the last) goes to chunk out[g / C], slot g % C.
template <class P>
__device__ void source_tile(const gf_node_desc& nd, const Batch& b, uint32_t t, const Cols& cols, uint32_t C,
                            const RunParams& rp)
{
    const uint32_t csh = cap_shift(C);
    if (t == kNoTile) return;
    uint32_t off = t * b.tile_items;
    uint32_t count = b.particles - off < b.tile_items ? b.particles - off : b.tile_items;
    if constexpr (std::is_same_v<P, Keyed>) {
        // four items per thread: their Zipf searches share two round trips
        for (uint32_t i0 = ltid(); i0 < count; i0 += 4u * kThreads) {
            double u[4];
            uint32_t key[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const uint32_t g = off + i0 + (uint32_t)q * kThreads;
                uint32_t r[4];
                stream_words(rp.seed, b.first_id + g, 0, b.purpose, r);
                u[q] = i0 + (uint32_t)q * kThreads < count ? uniform_f64(r[0], r[1]) : 0.0;
            }
            zipf_ranks<4>(rp, u, key);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const uint32_t i = i0 + (uint32_t)q * kThreads;
                if (i >= count) continue;
                const uint32_t g = off + i;
                Keyed p;
                p.key = key[q];
                p.val = b.first_id + g;
                p.store(cols, (size_t)b.out[cdiv(g, C, csh)].index * C + cmod(g, C, csh));
            }
        }
    } else {
        for (uint32_t i = ltid(); i < count; i += kThreads) {
            P p;
            const uint32_t g = off + i;
            source(nd, rp, b.purpose, b.first_id + g, p);
            p.store(cols, (size_t)b.out[cdiv(g, C, csh)].index * C + cmod(g, C, csh));
        }
    }
}

// Sink of one tile of one input chunk (SinkKernel, Execution.hpp:1508-1534).
template <class P>
__device__ void sink_tile(const gf_node_desc& nd, const Batch& b, uint32_t t, const Cols& cols, uint32_t C,
                          uint32_t* status, char* smem)
{
    TileSpan s = tile_span(b, b.in, b.n_in, t, C);
    if constexpr (std::is_same_v<P, Particle>) {
        for (uint32_t i = ltid(); i < s.count; i += kThreads) {
            Particle p = Particle::load(cols, s.base + i);
            if (nd.op == GF_SNK_HIST) {
                uint32_t v = nd.u[0] == 0u ? __float_as_uint(p.w) : p.gen;
                uint32_t rel = v - nd.u[1];
                if (v < nd.u[1] || rel >= nd.u[2]) set_status(status, GF_INVALID);
                else atomicAdd((unsigned long long*)nd.ptr[0] + rel, 1ull);
            } else if (nd.op == GF_SNK_RECORD_WEIGHT) {
                if ((uint64_t)p.id < nd.size[0]) ((uint32_t*)nd.ptr[0])[p.id] = __float_as_uint(p.w);
                else set_status(status, GF_INVALID);
            }
        }
    } else if constexpr (std::is_same_v<P, Keyed>) {
        if (nd.op != GF_SNK_CHUNK_DIGEST) return;
        const gf_chunk_desc d = s.count ? b.in[s.chunk] : gf_chunk_desc{0, 0, 0, 0};
        uint64_t acc = 0, msum = 0;
        for (uint32_t i = ltid(); i < s.count; i += kThreads) {
            Keyed p = Keyed::load(cols, s.base + i);
            uint64_t h = keyed_hash(p);
            acc += mix64(h ^ ((uint64_t)(s.offset + i + 1u) * 0xD6E8FEB86659FD93ull));
            msum += h;
        }
        // workgroup reduction of the two sums
        uint64_t* red = (uint64_t*)smem;
        for (int o = 32; o > 0; o >>= 1) {
            acc += __shfl_down(acc, o);
            msum += __shfl_down(msum, o);
        }
        if (lane_id() == 0) {
            red[2 * (ltid() >> 6)] = acc;
            red[2 * (ltid() >> 6) + 1] = msum;
        }
        __syncthreads();
        if (ltid() == 0 && s.count) {
            uint64_t a = 0, m = 0;
            for (int ww = 0; ww < kWaves; ++ww) {
                a += red[2 * ww];
                m += red[2 * ww + 1];
            }
            uint64_t entry = b.digest_base + s.chunk;
            if (nd.ptr[0] && entry < nd.size[0]) {
                uint32_t* log = (uint32_t*)nd.ptr[0] + 4 * entry;
                if (s.offset == 0) {
                    log[0] = d.bin;
                    log[1] = d.fill;
                }
                atomicAdd((unsigned long long*)(log + 2), (unsigned long long)a);
            }
            if (nd.ptr[1] && d.bin < nd.size[1]) {
                unsigned long long* ms = (unsigned long long*)nd.ptr[1] + 2 * d.bin;
                atomicAdd(ms, (unsigned long long)m);
                atomicAdd(ms + 1, (unsigned long long)s.count);
            }
        }
        __syncthreads();
    }
}

// Sink of NT consecutive tiles t .. t + nt - 1 (nt <= NT) by one wave (same results as sink_tile):
// for sinks that take many small chunks per step. The tiles' spans, their first 64 items and the
// digests' reductions are computed for all NT tiles together, so one wave keeps NT tiles' loads in
// flight instead of one tile's dependent chain at a time; larger tiles continue item by item.
template <class P, uint32_t NT>
__device__ void sink_tiles_wave(const gf_node_desc& nd, const Batch& b, uint32_t t, uint32_t nt, const Cols& cols,
                                uint32_t C, uint32_t* status)
{
    const uint32_t lane = lane_id();
    TileSpan sp[NT];
#pragma unroll
    for (uint32_t q = 0; q < NT; ++q) sp[q] = tile_span(b, b.in, b.n_in, q < nt ? t + q : kNoTile, C);
    if constexpr (std::is_same_v<P, Particle>) {
        auto item = [&](const Particle& p) {
            if (nd.op == GF_SNK_HIST) {
                uint32_t v = nd.u[0] == 0u ? __float_as_uint(p.w) : p.gen;
                uint32_t rel = v - nd.u[1];
                if (v < nd.u[1] || rel >= nd.u[2]) set_status(status, GF_INVALID);
                else atomicAdd((unsigned long long*)nd.ptr[0] + rel, 1ull);
            } else if (nd.op == GF_SNK_RECORD_WEIGHT) {
                if ((uint64_t)p.id < nd.size[0]) ((uint32_t*)nd.ptr[0])[p.id] = __float_as_uint(p.w);
                else set_status(status, GF_INVALID);
            }
        };
        Particle pk[NT];
#pragma unroll
        for (uint32_t q = 0; q < NT; ++q)
            if (lane < sp[q].count) pk[q] = Particle::load(cols, sp[q].base + lane);
#pragma unroll
        for (uint32_t q = 0; q < NT; ++q)
            if (lane < sp[q].count) item(pk[q]);
#pragma unroll
        for (uint32_t q = 0; q < NT; ++q)
            for (uint32_t i = lane + 64u; i < sp[q].count; i += 64u) item(Particle::load(cols, sp[q].base + i));
    } else if constexpr (std::is_same_v<P, Keyed>) {
        if (nd.op != GF_SNK_CHUNK_DIGEST) return;
        uint64_t acc[NT], msum[NT];
        Keyed pk[NT];
#pragma unroll
        for (uint32_t q = 0; q < NT; ++q) {
            acc[q] = msum[q] = 0ull;
            if (lane < sp[q].count) pk[q] = Keyed::load(cols, sp[q].base + lane);
        }
        auto item = [&](uint32_t q, const Keyed& p, uint32_t i) {
            const uint64_t h = keyed_hash(p);
            acc[q] += mix64(h ^ ((uint64_t)(sp[q].offset + i + 1u) * 0xD6E8FEB86659FD93ull));
            msum[q] += h;
        };
#pragma unroll
        for (uint32_t q = 0; q < NT; ++q)
            if (lane < sp[q].count) item(q, pk[q], lane);
#pragma unroll
        for (uint32_t q = 0; q < NT; ++q)
            for (uint32_t i = lane + 64u; i < sp[q].count; i += 64u) item(q, Keyed::load(cols, sp[q].base + i), i);
#pragma unroll
        for (uint32_t q = 0; q < NT; ++q)
            for (int o = 32; o > 0; o >>= 1) {
                acc[q] += __shfl_down(acc[q], o);
                msum[q] += __shfl_down(msum[q], o);
            }
        if (lane == 0) {
#pragma unroll
            for (uint32_t q = 0; q < NT; ++q) {
                if (!sp[q].count) continue;
                const gf_chunk_desc d = b.in[sp[q].chunk];
                const uint64_t entry = b.digest_base + sp[q].chunk;
                if (nd.ptr[0] && entry < nd.size[0]) {
                    uint32_t* log = (uint32_t*)nd.ptr[0] + 4 * entry;
                    if (sp[q].offset == 0) {
                        log[0] = d.bin;
                        log[1] = d.fill;
                    }
                    atomicAdd((unsigned long long*)(log + 2), (unsigned long long)acc[q]);
                }
                if (nd.ptr[1] && d.bin < nd.size[1]) {
                    unsigned long long* ms = (unsigned long long*)nd.ptr[1] + 2 * d.bin;
                    atomicAdd(ms, (unsigned long long)msum[q]);
                    atomicAdd(ms + 1, (unsigned long long)sp[q].count);
                }
            }
        }
    }
}

// Sink of 64 / RL consecutive tiles t .. t + nt - 1 by one wave, one tile per row of RL lanes (same
// results as sink_tile): the small per-bucket chunks a sink of a wide partition takes need a few row
// iterations each, the rows' loads are in flight together, and a row's digest sums are reduced
// within the row. Tiles of more than 64 items go to the whole wave afterwards, one at a time. The
// caller picks RL from the batch's items per tile (narrow rows for the tiny chunks of a batch of
// few keys per bucket: more chunks per round trip).
template <class P, uint32_t RL>
__device__ void sink_tiles_rows(const gf_node_desc& nd, const Batch& b, uint32_t t, uint32_t nt, const Cols& cols,
                                uint32_t C, uint32_t* status)
{
    const uint32_t lane = lane_id(), row = lane / RL, j = lane % RL;
    const TileSpan sp = tile_span(b, b.in, b.n_in, row < nt ? t + row : kNoTile, C);
    const bool small = sp.count <= 64u;
    if constexpr (std::is_same_v<P, Particle>) {
        if (small)
            for (uint32_t i = j; i < sp.count; i += RL) {
                const Particle p = Particle::load(cols, sp.base + i);
                if (nd.op == GF_SNK_HIST) {
                    uint32_t v = nd.u[0] == 0u ? __float_as_uint(p.w) : p.gen;
                    uint32_t rel = v - nd.u[1];
                    if (v < nd.u[1] || rel >= nd.u[2]) set_status(status, GF_INVALID);
                    else atomicAdd((unsigned long long*)nd.ptr[0] + rel, 1ull);
                } else if (nd.op == GF_SNK_RECORD_WEIGHT) {
                    if ((uint64_t)p.id < nd.size[0]) ((uint32_t*)nd.ptr[0])[p.id] = __float_as_uint(p.w);
                    else set_status(status, GF_INVALID);
                }
            }
    } else if constexpr (std::is_same_v<P, Keyed>) {
        if (nd.op != GF_SNK_CHUNK_DIGEST) return;
        uint64_t acc = 0ull, msum = 0ull;
        if (small)
            for (uint32_t i = j; i < sp.count; i += RL) {
                const Keyed p = Keyed::load(cols, sp.base + i);
                const uint64_t h = keyed_hash(p);
                acc += mix64(h ^ ((uint64_t)(sp.offset + i + 1u) * 0xD6E8FEB86659FD93ull));
                msum += h;
            }
        for (int o = (int)RL / 2; o > 0; o >>= 1) {
            acc += __shfl_down(acc, o, (int)RL);
            msum += __shfl_down(msum, o, (int)RL);
        }
        if (j == 0u && small && sp.count) {
            const gf_chunk_desc d = b.in[sp.chunk];
            const uint64_t entry = b.digest_base + sp.chunk;
            if (nd.ptr[0] && entry < nd.size[0]) {
                uint32_t* log = (uint32_t*)nd.ptr[0] + 4 * entry;
                if (sp.offset == 0) {
                    log[0] = d.bin;
                    log[1] = d.fill;
                }
                atomicAdd((unsigned long long*)(log + 2), (unsigned long long)acc);
            }
            if (nd.ptr[1] && d.bin < nd.size[1]) {
                unsigned long long* ms = (unsigned long long*)nd.ptr[1] + 2 * d.bin;
                atomicAdd(ms, (unsigned long long)msum);
                atomicAdd(ms + 1, (unsigned long long)sp.count);
            }
        }
    }
    // the larger tiles, each by the whole wave
    for (uint64_t big = __ballot(j == 0u && !small); big; big &= big - 1ull) {
        const uint32_t r = (uint32_t)__ffsll((unsigned long long)big) - 1u;
        sink_tiles_wave<P, 1>(nd, b, t + r / RL, 1u, cols, C, status);
    }
}

}  // namespace gfd
