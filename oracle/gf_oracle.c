/* CPU oracle of graphFlow's transform-batch path — TEST INFRASTRUCTURE only: it checks results.
 *
 * Nothing in the product (graphflow_amd) links or loads this file. Only tests/, the smoke check
 * and the CPU-baseline leg of bench.py use it, always as the reference the GPU results are
 * compared with, never as the measured GPU path.
 *
 * Pinning of each function (what fixes its output independently of this restatement):
 *   or_ceil_div, or_output_chunk_bound   literals tests/unit/test_reservation.cpp:49-61, exhaustive
 *                                        small-case property :74-101 (tests/test_oracle_primitives.py)
 *   or_plan_dense_bins/build_dense_table literals tests/unit/test_dense_bin_table.cpp:18-59
 *   or_tile_count/or_flatten_tiles       C=1000,T=128 tail 104 (tests/integration/test_dense_bins.cpp:195-218)
 *                                        and the coverage/ordering invariants
 *   or_num_frames_for                    literals tests/unit/test_launch_policy.cpp:27-57
 *   or_next_pow2/or_hash_bin_slot        literals tests/integration/test_hash_bins.cpp:296-305
 *   or_philox4x32_10                     Random123 known-answer vectors (Salmon et al., SC'11)
 *   or_mix64/or_stream_key               closed form + bijection/uniqueness property
 *                                        (tests/unit/test_philox.cpp:48-61 pattern)
 *   or_fmix32                            murmur3 finalizer known values (smhasher fmix32)
 *   or_partition_items / batch protocol  stable counting sort == numpy stable argsort by bin;
 *                                        kernel-axis histograms + SHA-256 known answers
 *                                        (tools/BENCH_TABLE.md:215-222); scripted routing fixture
 *                                        (tests/integration/test_scripted_routing.cpp:172-183)
 *   or_run (runner / step schedules)     conservation (sum rows == N), equality with the
 *                                        schedule-free or_feedback_scalar model, S=P+K*R
 *                                        (tests/integration/test_batching.cpp:577-602),
 *                                        credit-shrunk prefix (test_batching.cpp:550-575)
 *   or_zipf_table/or_zipf_sample         closed-form Zipf probabilities (top-1 / top-10 mass)
 *   or_keyed_scalar                      equals or_run's per-bucket multisets and counts (step and
 *                                        runner schedules) on the same seed
 *   or_hg_cos_f64                        closed-form HG moments <ct> = g, <ct^2> = (1 + 2 g^2) / 3
 *   or_hg_rotate_f64                     unit length and <u, u'> = ct for any u, both frames
 *   or_photon_scalar                     albedo 0 -> Beer-Lambert absorption profile N (e^-mu z0 - e^-mu z1)
 *                                        per depth bin (two mu_t), roulette unbiasedness (same mean
 *                                        absorbed weight with and without roulette), fp32 flow
 *                                        restatement within the stated per-bin tolerance
 */
#include "gf_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------------------------
 * Closed-form primitives
 * ---------------------------------------------------------------------------------------- */

/* Reservation.hpp:15-20 */
int or_ceil_div(uint64_t n, uint64_t d, uint64_t* out)
{
    if (d == 0) return OR_BAD_ARGUMENT;
    *out = n / d + (n % d != 0 ? 1u : 0u);
    return OR_OK;
}

/* Reservation.hpp:35-50: R = min(n, ceil(n/C) + min(B, n) - 1) */
int or_output_chunk_bound(uint64_t n, uint64_t capacity, uint64_t max_bins, uint64_t* out)
{
    if (n == 0) { *out = 0; return OR_OK; }
    if (capacity == 0 || max_bins == 0) return OR_BAD_ARGUMENT;
    uint64_t per_chunk;
    or_ceil_div(n, capacity, &per_chunk);
    uint64_t bins = max_bins < n ? max_bins : n;
    if (per_chunk > UINT64_MAX - bins + 1u) return OR_OVERFLOW;
    uint64_t bound = per_chunk + bins - 1u;
    *out = bound < n ? bound : n;
    return OR_OK;
}

/* DenseBinTable.hpp:39-77 */
int or_plan_dense_bins(const uint32_t* totals, uint32_t bins, uint32_t capacity, uint32_t* bin_base,
                       uint32_t* fills, uint32_t fills_cap, uint32_t* num_output_chunks)
{
    if (bins == 0 || capacity == 0) return OR_BAD_ARGUMENT;
    uint64_t exact = 0, particles = 0;
    for (uint32_t b = 0; b < bins; ++b) {
        bin_base[b] = (uint32_t)exact;
        uint64_t c;
        or_ceil_div(totals[b], capacity, &c);
        exact += c;
        particles += totals[b];
    }
    if (exact > UINT32_MAX) return OR_OVERFLOW;
    uint64_t bound;
    int rc = or_output_chunk_bound(particles, capacity, bins, &bound);
    if (rc) return rc;
    if (exact > bound) return OR_PLAN_MISMATCH;
    if (exact > fills_cap) return OR_BAD_ARGUMENT;
    uint32_t k = 0;
    for (uint32_t b = 0; b < bins; ++b)
        for (uint32_t off = 0; off < totals[b]; off += capacity) {
            uint32_t rem = totals[b] - off;
            fills[k++] = rem < capacity ? rem : capacity;
        }
    *num_output_chunks = (uint32_t)exact;
    return OR_OK;
}

/* DenseBinTable.hpp:80-91: [binBase[B], chunkIndex[R]] */
int or_build_dense_table(const uint32_t* bin_base, uint32_t bins, const uint32_t* chunk_indices,
                         uint32_t num_output_chunks, uint32_t* table)
{
    memcpy(table, bin_base, sizeof(uint32_t) * bins);
    memcpy(table + bins, chunk_indices, sizeof(uint32_t) * num_output_chunks);
    return OR_OK;
}

/* Execution.hpp:113-124 */
int or_tile_count(uint32_t fill, uint32_t tile_size, uint32_t* out)
{
    if (tile_size == 0) return OR_BAD_ARGUMENT;
    *out = fill == 0 ? 0u : (uint32_t)(((uint64_t)fill + tile_size - 1u) / tile_size);
    return OR_OK;
}

/* Execution.hpp:2377-2389: per chunk ceil(fill/T) tiles {chunkIndex, t*T, min(T, fill - t*T)} */
int or_flatten_tiles(const uint32_t* chunk_index, const uint32_t* fill, uint32_t n, uint32_t tile_size,
                     uint32_t cap, uint32_t* tiles, uint32_t* num_tiles)
{
    uint32_t k = 0;
    for (uint32_t c = 0; c < n; ++c) {
        uint32_t nt;
        int rc = or_tile_count(fill[c], tile_size, &nt);
        if (rc) return rc;
        for (uint32_t t = 0; t < nt; ++t) {
            if (k >= cap) return OR_OVERFLOW;
            uint32_t off = t * tile_size, rem = fill[c] - off;
            tiles[3 * k + 0] = chunk_index[c];
            tiles[3 * k + 1] = off;
            tiles[3 * k + 2] = rem < tile_size ? rem : tile_size;
            ++k;
        }
    }
    *num_tiles = k;
    return OR_OK;
}

/* LaunchPolicy.hpp:46-76 */
int or_num_frames_for(uint64_t work, uint32_t tile_size, uint32_t frame_cap, uint32_t max_threads,
                      uint32_t max_blocks, uint32_t sms, uint32_t* out)
{
    if (tile_size == 0 || frame_cap == 0 || max_threads == 0 || max_blocks == 0) return OR_BAD_ARGUMENT;
    if (tile_size > max_threads) return OR_BAD_ARGUMENT;
    if (work == 0) { *out = 0; return OR_OK; }
    uint64_t tiles = (work + tile_size - 1u) / tile_size;
    uint64_t s = sms == 0 ? 1u : sms;
    uint64_t capped = (uint64_t)frame_cap * s;
    uint64_t f = tiles;
    if ((uint64_t)max_blocks < f) f = max_blocks;
    if (capped < f) f = capped;
    *out = (uint32_t)f;
    return OR_OK;
}

/* BinTable.hpp:49-55 */
uint32_t or_next_pow2(uint32_t v)
{
    uint32_t r = 1;
    while (r < v) r <<= 1;
    return r;
}

/* BinTable.hpp:90-93 */
uint32_t or_hash_bin_slot(uint32_t bin, uint32_t mask) { return (bin * 0x9E3779B1u) & mask; }

/* BinTable.hpp:101-121, sequential form: insert or bump; 0 when the table is full. */
int or_hash_table_insert(uint32_t* keys, uint32_t* counts, uint32_t entries, uint32_t bin)
{
    uint32_t mask = entries - 1u, slot = or_hash_bin_slot(bin, mask);
    for (uint32_t p = 0; p < entries; ++p) {
        if (keys[slot] == 0xFFFFFFFFu || keys[slot] == bin) {
            keys[slot] = bin;
            counts[slot] += 1u;
            return 1;
        }
        slot = (slot + 1u) & mask;
    }
    return 0;
}

/* ------------------------------------------------------------------------------------------
 * RNG: Philox4x32-10 from the published algorithm (Salmon, Moraes, Dror, Shaw, SC'11), the
 * stream key mapping of rng/ParticleRng.hpp:57-94, and the murmur3 finalizer.
 * ---------------------------------------------------------------------------------------- */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* ParticleRng.hpp:69-77 (SplitMix64 finalizer) */
uint64_t or_mix64(uint64_t v)
{
    v ^= v >> 30;
    v *= 0xBF58476D1CE4E5B9ull;
    v ^= v >> 27;
    v *= 0x94D049BB133111EBull;
    v ^= v >> 31;
    return v;
}

/* ParticleRng.hpp:57-61, 86-94 */
uint64_t or_stream_key(uint64_t seed, uint64_t id, uint64_t event, uint32_t purpose)
{
    uint64_t packed = (id & 0xFFFFFFFFull) | ((event & 0xFFFFull) << 32) | ((uint64_t)(purpose & 0xFFFFu) << 48);
    return or_mix64(packed ^ or_mix64(seed));
}

/* The first Philox block of one stream position: key = stream key (lo, hi), counter 0. */
void or_stream_words(uint64_t seed, uint64_t id, uint64_t event, uint32_t purpose, uint32_t out[4])
{
    uint64_t k = or_stream_key(seed, id, event, purpose);
    uint32_t key[2] = {(uint32_t)k, (uint32_t)(k >> 32)};
    uint32_t ctr[4] = {0, 0, 0, 0};
    or_philox4x32_10(ctr, key, out);
}

uint32_t or_fmix32(uint32_t h)
{
    h ^= h >> 16;
    h *= 0x85EBCA6Bu;
    h ^= h >> 13;
    h *= 0xC2B2AE35u;
    h ^= h >> 16;
    return h;
}

/* U(0,1] with 24 bits: ((x >> 8) + 1) * 2^-24, exact in fp32. */
float or_uniform_f32(uint32_t word) { return (float)((word >> 8) + 1u) * (1.0f / 16777216.0f); }

/* U[0,1) with 53 bits from two words, exact in fp64. */
double or_uniform_f64(uint32_t w0, uint32_t w1)
{
    return ((double)(w0 >> 5) * 67108864.0 + (double)(w1 >> 6)) * (1.0 / 9007199254740992.0);
}

void or_zipf_table(uint32_t values, double s, double* cdf)
{
    double acc = 0.0;
    for (uint32_t r = 0; r < values; ++r) {
        acc += pow((double)(r + 1u), -s);
        cdf[r] = acc;
    }
}

/* smallest r with cdf[r] > u * cdf[values-1] */
uint32_t or_zipf_sample(const double* cdf, uint32_t values, uint32_t w0, uint32_t w1)
{
    double target = or_uniform_f64(w0, w1) * cdf[values - 1u];
    uint32_t lo = 0, hi = values - 1u;
    while (lo < hi) {
        uint32_t mid = lo + (hi - lo) / 2u;
        if (cdf[mid] > target) hi = mid; else lo = mid + 1u;
    }
    return lo;
}

/* ------------------------------------------------------------------------------------------
 * Items, ops and the photon step
 * ---------------------------------------------------------------------------------------- */
typedef struct or_item {
    float w;       /* particle weight / photon weight */
    uint32_t id;
    uint32_t gen;  /* particle generation / photon step count */
    uint32_t key;  /* hash pipeline key */
    uint64_t val;  /* hash pipeline payload */
    float px, py, pz, ux, uy, uz;
} or_item;

typedef struct or_ctx {
    const or_node_desc* nodes;
    uint32_t n_nodes;
    const int32_t* routes;
    const uint32_t* lut;
    const or_run_cfg* cfg;
    double* zipf;
    uint32_t zipf_values;
    or_run_out* out;
} or_ctx;

static uint32_t route_of(uint32_t dist, uint32_t ndest, uint32_t h)
{
    switch (dist) {
    case OR_ROUTE_PARITY: return h & 1u;
    case OR_ROUTE_SKEW_90_10: return (h % 10u == 0u) ? 1u : 0u;
    case OR_ROUTE_ALL_ONE: return 0u;
    default: return h % ndest;
    }
}

/* classifier of a partition node (Concepts.hpp:134-145 shape: pure function of the payload) */
static uint32_t or_classify(const or_ctx* x, const or_node_desc* nd, const or_item* it)
{
    switch (nd->op) {
    case OR_CLS_WEIGHT_ROUTE:
        if (it->w < nd->f[0]) return 0u;
        return 1u + route_of(nd->u[0], nd->bins - 1u, or_fmix32(it->id));
    case OR_CLS_KEY_BUCKET: return or_fmix32(it->key) % nd->bins;
    case OR_CLS_ID_MOD:
        if (nd->u[0] == 1u) return 0u;
        if (nd->u[0] == 2u) return (it->id % 8u == 0u) ? it->id % nd->bins : 0u;
        return it->id % nd->bins;
    case OR_CLS_LUT: {
        uint32_t a = it->id < nd->u[0] ? it->id : nd->u[0] - 1u;
        uint32_t g = it->gen < nd->u[1] ? it->gen : nd->u[1] - 1u;
        return x->lut[nd->u[3] + a * nd->u[1] + g];
    }
    case OR_CLS_PHOTON_ALIVE: return it->w > 0.0f ? 0u : 1u;
    default: return 0xFFFFFFFFu;
    }
}

/* partition updater; returns OR_OK / OR_INVALID / OR_OVERFLOW (UpdateResult, Concepts.hpp:33-39) */
static int or_update(const or_node_desc* nd, uint32_t bin, or_item* it)
{
    switch (nd->upd) {
    case OR_UPD_ADD_ON_BIN:
        if (bin == nd->u[2]) {
            float w = it->w + nd->f[1];
            if (isinf(w)) return OR_OVERFLOW;
            it->w = w;
            it->gen += 1u;
        }
        return OR_OK;
    case OR_UPD_GEN_TO_BIN: it->gen = bin; return OR_OK;
    default: return OR_OK;
    }
}

static void or_photon_step_f32(const or_run_cfg* c, uint64_t seed, uint32_t purpose, or_item* p, double* tally)
{
    if (!(p->w > 0.0f)) return;
    uint32_t r[4];
    or_stream_words(seed, p->id, p->gen, purpose, r);
    float L = (float)(c->layers * c->layer_mfp / c->mu_t);
    float s = -logf(or_uniform_f32(r[0])) / (float)c->mu_t;
    float nx = p->px + s * p->ux, ny = p->py + s * p->uy, nz = p->pz + s * p->uz;
    p->gen += 1u;
    if (nz < 0.0f || nz > L) { p->w = 0.0f; p->pz = nz; return; }
    p->px = nx; p->py = ny; p->pz = nz;
    uint32_t bin = (uint32_t)(nz / L * (float)c->tally_bins);
    if (bin >= c->tally_bins) bin = c->tally_bins - 1u;
    float a = (float)c->albedo;
    tally[bin] += (double)(p->w * (1.0f - a));
    p->w = p->w * a;
    if (p->w < (float)c->roulette_w) {
        float u = (float)(r[3] >> 8) * (1.0f / 16777216.0f);
        if (u < (float)c->roulette_p) p->w = p->w / (float)c->roulette_p;
        else { p->w = 0.0f; return; }
    }
    float g = (float)c->hg_g;
    float xi = (float)(r[1] >> 8) * (1.0f / 16777216.0f);
    float ct;
    if (g != 0.0f) {
        float tmp = (1.0f - g * g) / (1.0f - g + 2.0f * g * xi);
        ct = (1.0f + g * g - tmp * tmp) / (2.0f * g);
    } else {
        ct = 2.0f * xi - 1.0f;
    }
    if (ct > 1.0f) ct = 1.0f;
    if (ct < -1.0f) ct = -1.0f;
    float st = sqrtf(1.0f - ct * ct);
    float phi = 6.2831853071795864f * ((float)(r[2] >> 8) * (1.0f / 16777216.0f));
    float cp = cosf(phi), sp = sinf(phi);
    float ux = p->ux, uy = p->uy, uz = p->uz;
    if (fabsf(uz) > 0.99999f) {
        p->ux = st * cp;
        p->uy = st * sp;
        p->uz = uz > 0.0f ? ct : -ct;
    } else {
        float t = sqrtf(1.0f - uz * uz);
        p->ux = st * (ux * uz * cp - uy * sp) / t + ux * ct;
        p->uy = st * (uy * uz * cp + ux * sp) / t + uy * ct;
        p->uz = -st * cp * t + uz * ct;
    }
}

static void or_source_item(const or_ctx* x, const or_node_desc* nd, uint32_t purpose, uint64_t id, or_item* it)
{
    memset(it, 0, sizeof(*it));
    it->id = (uint32_t)id;
    uint32_t r[4];
    switch (nd->op) {
    case OR_SRC_PARTICLE_UNIFORM:
        or_stream_words(x->cfg->seed, id, 0, purpose, r);
        it->w = or_uniform_f32(r[0]);
        break;
    case OR_SRC_KEYED_ZIPF:
        or_stream_words(x->cfg->seed, id, 0, purpose, r);
        it->key = or_zipf_sample(x->zipf, x->zipf_values, r[0], r[1]);
        it->val = id;
        break;
    case OR_SRC_PHOTON:
        it->w = 1.0f;
        it->uz = 1.0f;
        break;
    default: break;
    }
}

static void or_transform_item(const or_ctx* x, const or_node_desc* nd, uint32_t purpose, or_item* it)
{
    switch (nd->op) {
    case OR_XF_SCALE_WEIGHT:
        it->w = it->w * nd->f[0];
        it->gen += 1u;
        break;
    case OR_XF_PHOTON_STEP: or_photon_step_f32(x->cfg, x->cfg->seed, purpose, it, x->out->tally); break;
    default: break;
    }
}

static uint64_t item_hash(const or_item* it)
{
    return or_mix64(((uint64_t)it->key << 32) ^ or_mix64(it->val + 0x9E3779B97F4A7C15ull));
}

/* ------------------------------------------------------------------------------------------
 * Pool: fixed SoA chunks and the LIFO free list of ChunkLedger.hpp:36-134
 * ---------------------------------------------------------------------------------------- */
typedef struct or_pool {
    uint32_t S, C;
    or_item* slots;
    uint32_t* free_stack;
    uint32_t top;
    uint32_t* gen;
    uint32_t peak_live;
} or_pool;

static int pool_init(or_pool* p, uint32_t S, uint32_t C)
{
    p->S = S; p->C = C;
    p->slots = (or_item*)calloc((size_t)S * C, sizeof(or_item));
    p->free_stack = (uint32_t*)malloc(sizeof(uint32_t) * S);
    p->gen = (uint32_t*)calloc(S, sizeof(uint32_t));
    if (!p->slots || !p->free_stack || !p->gen) return OR_BAD_ARGUMENT;
    /* initial order S-1..0: index 0 is handed out first (ChunkLedger.hpp:43-46) */
    for (uint32_t i = 0; i < S; ++i) p->free_stack[i] = S - 1u - i;
    p->top = S;
    p->peak_live = 0;
    return OR_OK;
}

static void pool_free(or_pool* p)
{
    free(p->slots); free(p->free_stack); free(p->gen);
}

static int pool_acquire(or_pool* p, uint32_t* idx)
{
    if (p->top == 0) return OR_POOL_EXHAUSTED;
    *idx = p->free_stack[--p->top];
    uint32_t live = p->S - p->top;
    if (live > p->peak_live) p->peak_live = live;
    return OR_OK;
}

static void pool_release(or_pool* p, uint32_t idx)
{
    p->gen[idx] += 1u;
    p->free_stack[p->top++] = idx;
}

/* ------------------------------------------------------------------------------------------
 * Queues, batches and the trace
 * ---------------------------------------------------------------------------------------- */
typedef struct desc { uint32_t index, gen, fill, bin; } desc;

typedef struct fifo { desc* d; uint64_t head, tail, cap; } fifo;

static void fifo_push(fifo* f, desc v)
{
    if (f->tail - f->head == f->cap) {
        uint64_t n = f->tail - f->head, nc = f->cap ? f->cap * 2 : 64;
        desc* nd = (desc*)malloc(sizeof(desc) * nc);
        for (uint64_t i = 0; i < n; ++i) nd[i] = f->d[(f->head + i) % f->cap];
        free(f->d);
        f->d = nd; f->head = 0; f->tail = n; f->cap = nc;
    }
    f->d[f->tail % f->cap] = v;
    f->tail++;
}
static uint64_t fifo_size(const fifo* f) { return f->tail - f->head; }
static desc fifo_at(const fifo* f, uint64_t i) { return f->d[(f->head + i) % f->cap]; }

typedef struct batch {
    uint32_t n_in;
    desc* in;        /* input descriptors (cap Q or S) */
    uint32_t n_out;
    desc* out;       /* output descriptors */
    uint32_t* counts;/* per bin totals (partition) */
    or_item* items;  /* gathered items in batch order */
    uint32_t* bins;  /* classified bin per item */
    uint64_t particles;
    uint64_t first_id; /* source lanes: id of the chunk's first item */
    uint32_t reserved;
    int stage;       /* runner: 0 launched, 1 planned */
    int armed;
} batch;

static void trace_put(or_run_out* o, uint32_t v)
{
    if (o->trace && o->trace_len < o->trace_cap) o->trace[o->trace_len] = v;
    o->trace_len++;
}

/* record layout: 0xB47C0000|node, step, n_in, (index,gen,fill)*n_in, nbins, counts*nbins, R,
 * (index,gen,fill,bin)*R */
static void trace_batch(or_run_out* o, uint32_t node, uint64_t step, const batch* b, uint32_t nbins)
{
    if (!o->trace) return;
    trace_put(o, 0xB47C0000u | node);
    trace_put(o, (uint32_t)step);
    trace_put(o, b->n_in);
    for (uint32_t i = 0; i < b->n_in; ++i) {
        trace_put(o, b->in[i].index); trace_put(o, b->in[i].gen); trace_put(o, b->in[i].fill);
    }
    trace_put(o, nbins);
    for (uint32_t k = 0; k < nbins; ++k) trace_put(o, b->counts[k]);
    trace_put(o, b->n_out);
    for (uint32_t i = 0; i < b->n_out; ++i) {
        trace_put(o, b->out[i].index); trace_put(o, b->out[i].gen);
        trace_put(o, b->out[i].fill); trace_put(o, b->out[i].bin);
    }
}

static void fail(or_run_out* o, int status, const char* msg)
{
    if (o->status == OR_OK) {
        o->status = (uint32_t)status;
        snprintf(o->message, sizeof(o->message), "%s", msg);
    }
}

/* OpenMP threads for a batch loop of n items (the reference's CPU backend runs a batch's blocks on
 * the host cores): every result is independent of the thread count. */
static int batch_threads(const or_ctx* x, uint64_t n)
{
    int t = x->cfg && x->cfg->threads ? (int)x->cfg->threads : 1;
    return n >= 4096u ? t : 1;
}

/* gather the batch's items in batch order: chunk order, then slot order (the chunks in parallel) */
static void gather(const or_ctx* x, const or_pool* p, batch* b)
{
    uint64_t k = 0;
    for (uint32_t c = 0; c < b->n_in; ++c) k += b->in[c].fill;
    b->particles = k;
    uint64_t* at = (uint64_t*)malloc(sizeof(uint64_t) * (b->n_in + 1u));
    at[0] = 0;
    for (uint32_t c = 0; c < b->n_in; ++c) at[c + 1] = at[c] + b->in[c].fill;
    const int64_t nc = b->n_in;
#pragma omp parallel for schedule(dynamic, 1) num_threads(batch_threads(x, k))
    for (int64_t c = 0; c < nc; ++c) {
        const or_item* src = p->slots + (size_t)b->in[c].index * p->C;
        memcpy(b->items + at[c], src, sizeof(or_item) * b->in[c].fill);
    }
    free(at);
}

/* Partition classify (Execution.hpp:869-954 semantics): bins per item, totals per bin. */
static int partition_classify(const or_ctx* x, const or_node_desc* nd, const or_pool* p, batch* b)
{
    gather(x, p, b);
    memset(b->counts, 0, sizeof(uint32_t) * nd->bins);
    const int64_t n = (int64_t)b->particles;
    const int T = batch_threads(x, b->particles);
    const uint32_t B = nd->bins;
    /* per-thread counts over contiguous item ranges, summed afterwards (integer sums: the totals
     * do not depend on the thread count) */
    uint32_t* tc = (uint32_t*)calloc((size_t)T * B, sizeof(uint32_t));
    int bad = 0;
#pragma omp parallel num_threads(T) reduction(| : bad)
    {
#ifdef _OPENMP
        const int t = omp_get_thread_num(), nt = omp_get_num_threads();
#else
        const int t = 0, nt = 1;
#endif
        const int64_t lo = n * t / nt, hi = n * (t + 1) / nt;
        uint32_t* c = tc + (size_t)t * B;
        for (int64_t i = lo; i < hi; ++i) {
            uint32_t bin = or_classify(x, nd, &b->items[i]);
            if (bin >= B) { bad = 1; bin = 0; b->bins[i] = 0xFFFFFFFFu; continue; }
            b->bins[i] = bin;
            c[bin]++;
        }
    }
    for (int t = 0; t < T; ++t)
        for (uint32_t k = 0; k < B; ++k) b->counts[k] += tc[(size_t)t * B + k];
    free(tc);
    if (bad) return OR_INVALID;
    return OR_OK;
}

/* Exact plan + output acquisition + stable scatter (Execution.hpp:2520-2642 and the scatter
 * kernels :1096-1187 / :1327-1502). Slot order inside a bin is the canonical stable order:
 * bin-major, then batch input order. */
/* the plain transform node fused into bin `bin` of partition `nd`, or -1 */
static int fused_of(const or_ctx* x, const or_node_desc* nd, uint32_t bin)
{
    if (!x->cfg->fuse || x->cfg->mode != OR_MODE_STEP) return -1;
    int32_t d = x->routes[nd->route_offset + bin];
    const or_node_desc* t = &x->nodes[d];
    if (t->role == OR_ROLE_TRANSFORM && t->op >= OR_XF_SCALE_WEIGHT && t->op < OR_SNK_HIST && t->bins == 1u) return d;
    return -1;
}

static int partition_plan_scatter(const or_ctx* x, const or_node_desc* nd, uint32_t purpose, or_pool* p,
                                  batch* b, uint32_t reserve_bound)
{
    uint64_t sum = 0;
    for (uint32_t k = 0; k < nd->bins; ++k) sum += b->counts[k];
    if (sum != b->particles) return OR_PLAN_MISMATCH;
    uint32_t* base = (uint32_t*)malloc(sizeof(uint32_t) * nd->bins);
    uint32_t* fills = (uint32_t*)malloc(sizeof(uint32_t) * (reserve_bound + 1u));
    uint32_t R = 0;
    int rc = or_plan_dense_bins(b->counts, nd->bins, p->C, base, fills, reserve_bound + 1u, &R);
    if (rc == OR_OK && R > reserve_bound) rc = OR_PLAN_MISMATCH;
    if (rc == OR_OK && R > p->top) rc = OR_POOL_EXHAUSTED;
    if (rc) { free(base); free(fills); return rc; }
    /* acquire in bin order, k order */
    uint32_t r = 0;
    for (uint32_t bin = 0; bin < nd->bins; ++bin) {
        uint64_t nch;
        or_ceil_div(b->counts[bin], p->C, &nch);
        for (uint32_t k = 0; k < nch; ++k, ++r) {
            uint32_t idx = 0;
            pool_acquire(p, &idx);
            b->out[r].index = idx;
            b->out[r].gen = p->gen[idx];
            b->out[r].fill = fills[r];
            b->out[r].bin = bin;
        }
    }
    b->n_out = R;
    /* stable scatter with the updater (purpose + 1): T contiguous item ranges, each range's start
     * per bin is the count of the bin in the earlier ranges, so every item gets its sequential rank */
    (void)purpose;
    const int T = batch_threads(x, b->particles);
    const uint64_t n = b->particles, B = nd->bins;
    uint32_t* cursor = (uint32_t*)calloc((size_t)T * B, sizeof(uint32_t));
    int* urc_t = (int*)calloc((size_t)T, sizeof(int));
#pragma omp parallel num_threads(T)
    {
#ifdef _OPENMP
        const int t = omp_get_thread_num(), nt = omp_get_num_threads();
#else
        const int t = 0, nt = 1;
#endif
        const uint64_t lo = n * (uint64_t)t / (uint64_t)nt, hi = n * (uint64_t)(t + 1) / (uint64_t)nt;
        uint32_t* cur = cursor + (size_t)t * B;
        for (uint64_t i = lo; i < hi; ++i) cur[b->bins[i]]++;
#pragma omp barrier
#pragma omp single
        {
            for (uint64_t k = 0; k < B; ++k) {
                uint32_t acc = 0;
                for (int tt = 0; tt < nt; ++tt) {
                    uint32_t c = cursor[(size_t)tt * B + k];
                    cursor[(size_t)tt * B + k] = acc;
                    acc += c;
                }
            }
        }
        for (uint64_t i = lo; i < hi; ++i) {
            uint32_t bin = b->bins[i];
            uint32_t j = cur[bin]++;
            or_item it = b->items[i];
            int u = or_update(nd, bin, &it);
            if (u != OR_OK && urc_t[t] == OR_OK) urc_t[t] = u;
            int fz = fused_of(x, nd, bin);
            if (fz >= 0) or_transform_item(x, &x->nodes[fz], 2u * (uint32_t)fz, &it);
            uint32_t chunk = b->out[base[bin] + j / p->C].index;
            p->slots[(size_t)chunk * p->C + j % p->C] = it;
        }
    }
    int urc = OR_OK;
    for (int t = 0; t < T && urc == OR_OK; ++t) urc = urc_t[t];
    free(urc_t); free(cursor); free(base); free(fills);
    return urc;
}

/* TransformStage::launch + TransformKernel (Execution.hpp:708-776, 1805-1901): dense
 * concatenation into R = ceil(sum(fill)/C) chunks. */
static int transform_batch(const or_ctx* x, const or_node_desc* nd, uint32_t purpose, or_pool* p, batch* b,
                           uint32_t reserve_bound)
{
    gather(x, p, b);
    uint64_t R;
    or_ceil_div(b->particles, p->C, &R);
    if (b->particles == 0) return OR_PLAN_MISMATCH;
    if (R > reserve_bound) return OR_PLAN_MISMATCH;
    if (R > p->top) return OR_POOL_EXHAUSTED;
    for (uint32_t k = 0; k < R; ++k) {
        uint32_t idx = 0;
        pool_acquire(p, &idx);
        uint64_t rem = b->particles - (uint64_t)k * p->C;
        b->out[k].index = idx;
        b->out[k].gen = p->gen[idx];
        b->out[k].fill = (uint32_t)(rem < p->C ? rem : p->C);
        b->out[k].bin = 0;
    }
    b->n_out = (uint32_t)R;
    b->counts[0] = (uint32_t)b->particles;
    const int64_t n = (int64_t)b->particles;
#pragma omp parallel for schedule(static) num_threads(batch_threads(x, b->particles))
    for (int64_t i = 0; i < n; ++i) {
        or_item it = b->items[i];
        or_transform_item(x, nd, purpose, &it);
        p->slots[(size_t)b->out[i / p->C].index * p->C + i % p->C] = it;
    }
    return OR_OK;
}

static void sink_batch(const or_ctx* x, const or_node_desc* nd, const or_pool* p, batch* b)
{
    or_run_out* o = x->out;
    if (nd->op == OR_SNK_HIST && batch_threads(x, b->particles) > 1) {
        /* histogram increments commute: the chunks' items in parallel */
        int bad = 0;
        for (uint32_t c = 0; c < b->n_in; ++c) {
            const or_item* src = p->slots + (size_t)b->in[c].index * p->C;
            const int64_t n = b->in[c].fill;
#pragma omp parallel for schedule(static) num_threads(batch_threads(x, b->particles)) reduction(| : bad)
            for (int64_t i = 0; i < n; ++i) {
                uint32_t v = nd->u[0] == 0u ? 0u : src[i].gen;
                if (nd->u[0] == 0u) memcpy(&v, &src[i].w, 4);
                uint32_t rel = v - nd->u[1];
                if (v < nd->u[1] || rel >= nd->u[2]) bad = 1;
                else if (o->hist) {
#pragma omp atomic
                    o->hist[rel]++;
                }
            }
        }
        if (bad) fail(o, OR_INVALID, "sink: histogram value out of range");
        return;
    }
    for (uint32_t c = 0; c < b->n_in; ++c) {
        const or_item* src = p->slots + (size_t)b->in[c].index * p->C;
        uint64_t h0 = 0;
        for (uint32_t i = 0; i < b->in[c].fill; ++i) {
            const or_item* it = &src[i];
            switch (nd->op) {
            case OR_SNK_HIST: {
                uint32_t v = nd->u[0] == 0u ? 0u : it->gen;
                if (nd->u[0] == 0u) memcpy(&v, &it->w, 4);
                uint32_t rel = v - nd->u[1];
                if (v < nd->u[1] || rel >= nd->u[2]) fail(o, OR_INVALID, "sink: histogram value out of range");
                else if (o->hist) o->hist[rel]++;
                break;
            }
            case OR_SNK_RECORD_WEIGHT: {
                uint32_t v;
                memcpy(&v, &it->w, 4);
                if (o->record) o->record[it->id] = v;
                break;
            }
            case OR_SNK_CHUNK_DIGEST: {
                uint64_t h = item_hash(it);
                h0 += or_mix64(h ^ ((uint64_t)(i + 1u) * 0xD6E8FEB86659FD93ull));
                if (o->bin_multiset && b->in[c].bin < o->bin_multiset_bins) {
                    o->bin_multiset[b->in[c].bin] += h;
                    o->bin_items[b->in[c].bin] += 1u;
                }
                break;
            }
            default: break;
            }
        }
        if (nd->op == OR_SNK_CHUNK_DIGEST) {
            if (o->digest_log && o->digest_len < o->digest_cap) {
                uint32_t* d = o->digest_log + 4 * o->digest_len;
                d[0] = b->in[c].bin; d[1] = b->in[c].fill; d[2] = (uint32_t)h0; d[3] = (uint32_t)(h0 >> 32);
            }
            o->digest_len++;
        }
    }
}

static void source_batch(const or_ctx* x, const or_node_desc* nd, uint32_t purpose, or_pool* p, const desc* chunk,
                         uint64_t first_id)
{
    or_item* dst = p->slots + (size_t)chunk->index * p->C;
    const int64_t n = chunk->fill;
#pragma omp parallel for schedule(static) num_threads(batch_threads(x, (uint64_t)n))
    for (int64_t i = 0; i < n; ++i) or_source_item(x, nd, purpose, first_id + (uint64_t)i, &dst[i]);
}

/* ------------------------------------------------------------------------------------------
 * Pool derivation (Flow.hpp:580-626) and the two schedules
 * ---------------------------------------------------------------------------------------- */
int or_derive_pool(const or_node_desc* nodes, uint32_t n_nodes, uint64_t items, uint32_t capacity,
                   uint32_t live, uint32_t* live_out, uint32_t* pool_out)
{
    if (capacity == 0) return OR_BAD_ARGUMENT;
    uint64_t reserve = 0;
    for (uint32_t k = 0; k < n_nodes; ++k) {
        if (nodes[k].role != OR_ROLE_TRANSFORM) continue;
        uint64_t b;
        int rc = or_output_chunk_bound(capacity, capacity, nodes[k].bins, &b);
        if (rc) return rc;
        reserve += b;
    }
    uint64_t P = live ? live : (uint64_t)8u * capacity; /* accelerator default P = 8*C */
    if (P > items) P = items;
    if (P == 0) P = 1;
    *live_out = (uint32_t)P;
    if (items == 0) { *pool_out = 1; return OR_OK; }
    if (P + reserve > UINT32_MAX) return OR_OVERFLOW;
    *pool_out = (uint32_t)(P + reserve);
    return OR_OK;
}

static uint32_t reserve_bound_of(const or_node_desc* nd, uint32_t C, uint32_t Q)
{
    uint64_t b;
    /* PartitionStage::checkedReserveBound(C, Q, B) and TransformStage's (C, Q) form */
    or_output_chunk_bound((uint64_t)C * Q, C, nd->op >= OR_XF_SCALE_WEIGHT && nd->op < OR_SNK_HIST ? 1u : nd->bins, &b);
    return (uint32_t)b;
}

static int is_partition(const or_node_desc* nd) { return nd->op >= OR_CLS_WEIGHT_ROUTE && nd->op < OR_XF_SCALE_WEIGHT; }

/* Bin path of a partition (BinPath.hpp): hash above 4096 bins unless the run fixes one (the
 * build's dense limit: 24 bytes of LDS per bin in a scatter tile of MI355X's 160 KB). */
static int uses_hash_path(const or_run_cfg* cfg, const or_node_desc* nd)
{
    return cfg->bin_path == 2u || (cfg->bin_path == 0u && nd->bins > 4096u);
}

/* Hash-path classify check (Execution.hpp:1207-1307): every tile of hash_tile items of every input
 * chunk inserts its bins into a table of hash_entries slots (BinTable.hpp:101-121); a tile with
 * more distinct bins than slots fails the batch before any output is planned. With hash_entries 0
 * the table has nextPow2(2 * tile) slots and can never fill. */
static int hash_tables_fit(const or_run_cfg* cfg, const or_node_desc* nd, const batch* b)
{
    const uint32_t T = cfg->hash_tile, E = cfg->hash_entries;
    if (!uses_hash_path(cfg, nd) || E == 0u || T == 0u) return 1;
    uint32_t* keys = (uint32_t*)malloc(sizeof(uint32_t) * E);
    uint32_t* cnt = (uint32_t*)malloc(sizeof(uint32_t) * E);
    int ok = 1;
    uint64_t item = 0;
    for (uint32_t c = 0; c < b->n_in && ok; ++c) {
        for (uint32_t off = 0; off < b->in[c].fill && ok; off += T) {
            const uint32_t n = b->in[c].fill - off < T ? b->in[c].fill - off : T;
            for (uint32_t e = 0; e < E; ++e) { keys[e] = 0xFFFFFFFFu; cnt[e] = 0u; }
            for (uint32_t i = 0; i < n; ++i)
                if (!or_hash_table_insert(keys, cnt, E, b->bins[item + off + i])) { ok = 0; break; }
        }
        item += b->in[c].fill;
    }
    free(keys); free(cnt);
    return ok;
}

int or_run(const or_node_desc* nodes, uint32_t n_nodes, const int32_t* routes, const uint32_t* lut,
           const or_run_cfg* cfg, or_run_out* out)
{
    out->status = OR_OK;
    out->message[0] = 0;
    out->trace_len = 0;
    out->digest_len = 0;
    out->batches = out->steps = out->generated = out->sunk = 0;
    or_ctx x = {nodes, n_nodes, routes, lut, cfg, NULL, 0, out};
    uint32_t C = cfg->capacity, Q = cfg->batch ? cfg->batch : 1u, P = 0, S = 0;
    /* Tuning::inflightBatches (Tuning.hpp:34-41): lanes per node (GraphRunner.hpp:766-771) */
    const uint32_t D = cfg->inflight ? cfg->inflight : 1u;
    int rc = or_derive_pool(nodes, n_nodes, cfg->items, C, cfg->live, &P, &S);
    if (rc) { fail(out, rc, "pool derivation"); return rc; }
    if (cfg->num_chunks) S = cfg->num_chunks;
    out->pool_chunks = S;
    int src = -1;
    for (uint32_t k = 0; k < n_nodes; ++k) {
        if (nodes[k].role == OR_ROLE_SOURCE) {
            if (src >= 0) { fail(out, OR_BAD_ARGUMENT, "a graph needs exactly one source node"); return OR_BAD_ARGUMENT; }
            src = (int)k;
        }
        if (nodes[k].role != OR_ROLE_SINK)
            for (uint32_t b = 0; b < nodes[k].bins; ++b) {
                int32_t d = routes[nodes[k].route_offset + b];
                if (d < 0 || (uint32_t)d >= n_nodes || nodes[d].role == OR_ROLE_SOURCE) {
                    fail(out, OR_BAD_ARGUMENT, "missing or invalid route");
                    return OR_BAD_ARGUMENT;
                }
            }
        if (nodes[k].op == OR_SRC_KEYED_ZIPF && !x.zipf) {
            x.zipf_values = nodes[k].u[0];
            x.zipf = (double*)malloc(sizeof(double) * x.zipf_values);
            or_zipf_table(x.zipf_values, (double)nodes[k].f[0], x.zipf);
        }
    }
    if (src < 0) { fail(out, OR_BAD_ARGUMENT, "no source node"); return OR_BAD_ARGUMENT; }

    or_pool pool;
    if (pool_init(&pool, S, C)) { fail(out, OR_BAD_ARGUMENT, "pool allocation"); return OR_BAD_ARGUMENT; }

    fifo* ready = (fifo*)calloc(n_nodes, sizeof(fifo));
    /* lane j of node k is lane[k * D + j]; every lane owns its batch, leases and credit */
    batch* lane = (batch*)calloc((size_t)n_nodes * D, sizeof(batch));
    /* step mode: a source lane admits up to SC chunks per step (runner mode: one per dispatch) */
    const uint32_t SC = cfg->mode == OR_MODE_STEP && cfg->source_chunks ? cfg->source_chunks : 1u;
    uint64_t item_cap = (uint64_t)Q * C;
    for (uint32_t k = 0; k < n_nodes; ++k)
        for (uint32_t j = 0; j < D; ++j) {
            batch* b = &lane[k * D + j];
            uint32_t cap_in = nodes[k].role == OR_ROLE_SINK ? S : Q;
            b->in = (desc*)malloc(sizeof(desc) * (cap_in + 1u));
            uint32_t cap_out = nodes[k].role == OR_ROLE_SOURCE ? SC : reserve_bound_of(&nodes[k], C, Q) + 1u;
            b->out = (desc*)malloc(sizeof(desc) * cap_out);
            b->counts = (uint32_t*)malloc(sizeof(uint32_t) * (nodes[k].bins + 1u));
            if (nodes[k].role == OR_ROLE_TRANSFORM) {
                b->items = (or_item*)malloc(sizeof(or_item) * item_cap);
                b->bins = (uint32_t*)malloc(sizeof(uint32_t) * item_cap);
            }
        }
    uint64_t N = cfg->items, generated = 0, sunk = 0;
    uint32_t reserved = 0, rr = 0;
    uint64_t iter = 0;

#define FAIL_RUN(code, msg) do { fail(out, code, msg); goto done; } while (0)

    if (cfg->mode == OR_MODE_RUNNER) {
        /* GraphRunner::run (GraphRunner.hpp:716-1257) with every launched device stage complete by
         * the next loop iteration: one legal interleaving of the reference's event polling. A
         * partition batch spans two iterations (classify, then plan + scatter), so with several
         * lanes a node can dispatch its next batch while the previous one is still in flight. */
        for (;;) {
            int progress = 0;
            /* 1) retire, every lane of every node */
            for (uint32_t k = 0; k < n_nodes; ++k)
                for (uint32_t j = 0; j < D; ++j) {
                    batch* b = &lane[k * D + j];
                    if (!b->armed) continue;
                    const or_node_desc* nd = &nodes[k];
                    uint32_t purpose = 2u * k;
                    int more = 0;
                    uint32_t allocated = 0;
                    if (nd->role == OR_ROLE_TRANSFORM && is_partition(nd) && b->stage == 2) {
                        allocated = b->n_out;  /* fused batch: planned and scattered at dispatch */
                    } else if (nd->role == OR_ROLE_TRANSFORM && is_partition(nd)) {
                        if (b->stage == 0) {
                            int r2 = partition_plan_scatter(&x, nd, purpose + 1u, &pool, b, reserve_bound_of(nd, C, Q));
                            if (r2 == OR_OVERFLOW) FAIL_RUN(OR_OVERFLOW, "partition scatter: checked updater overflow");
                            if (r2 == OR_INVALID) FAIL_RUN(OR_INVALID, "partition scatter: invalid updater result");
                            if (r2) FAIL_RUN(r2, "partition plan failed");
                            trace_batch(out, k, iter, b, nd->bins);
                            allocated = b->n_out;
                            b->stage = 1;
                            more = 1;
                        }
                    } else if (nd->role == OR_ROLE_TRANSFORM) {
                        allocated = b->n_out;
                    }
                    if (allocated) {
                        if (allocated > b->reserved) FAIL_RUN(OR_BAD_ARGUMENT, "node acquired more chunks than reserved");
                        reserved -= b->reserved;
                        b->reserved = 0;
                    } else if (!more && b->reserved) {
                        reserved -= b->reserved;
                        b->reserved = 0;
                    }
                    if (!more) {
                        /* finish: route outputs (GraphRunner.hpp:808-852) */
                        if (nd->role != OR_ROLE_SINK)
                            for (uint32_t o = 0; o < b->n_out; ++o) {
                                int32_t d = routes[nd->route_offset + b->out[o].bin];
                                fifo_push(&ready[d], b->out[o]);
                            }
                        if (nd->role == OR_ROLE_SINK) sunk += b->particles;
                        if (nd->role != OR_ROLE_SOURCE)
                            for (uint32_t i = 0; i < b->n_in; ++i) pool_release(&pool, b->in[i].index);
                        if (nd->role == OR_ROLE_TRANSFORM) out->batches++;
                        b->armed = 0;
                    }
                    progress = 1;
                }
            /* 2) admission into the source's first free lane (GraphRunner.hpp:902-983) */
            {
                uint64_t live = generated - sunk, remaining = N - generated;
                uint64_t min_fill = C < P ? C : P;
                if (remaining < min_fill) min_fill = remaining;
                int within = live < P;
                int budget = live < P && (P - live) >= min_fill;
                batch* b = NULL;
                for (uint32_t j = 0; j < D && !b; ++j)
                    if (!lane[(uint32_t)src * D + j].armed) b = &lane[(uint32_t)src * D + j];
                if (generated < N && b && within && budget && pool.top > reserved) {
                    uint64_t fill = remaining;
                    if (fill > P - live) fill = P - live;
                    if (fill > C) fill = C;
                    uint32_t idx = 0;
                    pool_acquire(&pool, &idx);
                    b->n_in = 0;
                    b->n_out = 1;
                    b->out[0].index = idx; b->out[0].gen = pool.gen[idx];
                    b->out[0].fill = (uint32_t)fill; b->out[0].bin = 0;
                    b->particles = fill;
                    source_batch(&x, &nodes[src], 2u * (uint32_t)src, &pool, &b->out[0], generated);
                    trace_batch(out, (uint32_t)src, iter, b, 0);
                    generated += fill;
                    b->armed = 1;
                    progress = 1;
                }
            }
            /* 3) dispatch one batch into a free lane: sinks first, then round robin
             *    (GraphRunner.hpp:985-1133) */
            {
                int dispatched = 0;
                for (int pass = 0; pass < 2 && !dispatched; ++pass) {
                    for (uint32_t kk = 0; kk < n_nodes && !dispatched; ++kk) {
                        uint32_t k = pass == 0 ? kk : (rr + kk) % n_nodes;
                        const or_node_desc* nd = &nodes[k];
                        if (pass == 0 && nd->role != OR_ROLE_SINK) continue;
                        if (fifo_size(&ready[k]) == 0) continue;
                        batch* b = NULL;
                        for (uint32_t j = 0; j < D && !b; ++j)
                            if (!lane[k * D + j].armed) b = &lane[k * D + j];
                        if (!b) continue;
                        uint64_t wanted = nd->role == OR_ROLE_TRANSFORM ? (fifo_size(&ready[k]) < Q ? fifo_size(&ready[k]) : Q) : 1u;
                        uint32_t reserve = 0;
                        if (nd->role == OR_ROLE_TRANSFORM) {
                            uint64_t parts = 0;
                            for (uint64_t i = 0; i < wanted; ++i) parts += fifo_at(&ready[k], i).fill;
                            while (wanted) {
                                uint64_t bound;
                                or_output_chunk_bound(parts, C, nd->bins, &bound);
                                if (pool.top - reserved >= bound) { reserve = (uint32_t)bound; break; }
                                --wanted;
                                parts -= fifo_at(&ready[k], wanted).fill;
                            }
                            if (!wanted) continue;
                        }
                        b->n_in = (uint32_t)wanted;
                        for (uint64_t i = 0; i < wanted; ++i) b->in[i] = fifo_at(&ready[k], i);
                        ready[k].head += wanted;
                        b->reserved = reserve;
                        reserved += reserve;
                        b->n_out = 0;
                        b->stage = 0;
                        b->armed = 1;
                        uint32_t purpose = 2u * k;
                        if (nd->role == OR_ROLE_SINK) {
                            b->particles = 0;
                            for (uint32_t i = 0; i < b->n_in; ++i) b->particles += b->in[i].fill;
                            sink_batch(&x, nd, &pool, b);
                            trace_batch(out, k, iter, b, 0);
                        } else if (is_partition(nd)) {
                            int r2 = partition_classify(&x, nd, &pool, b);
                            if (r2) FAIL_RUN(r2, "partition classify: device status reports an invalid classifier result");
                            if (!hash_tables_fit(cfg, nd, b)) FAIL_RUN(OR_INVALID, "partition classify: hash bin table full");
                            if (cfg->fused_stages) {
                                r2 = partition_plan_scatter(&x, nd, purpose + 1u, &pool, b, reserve_bound_of(nd, C, Q));
                                if (r2 == OR_OVERFLOW) FAIL_RUN(OR_OVERFLOW, "partition scatter: checked updater overflow");
                                if (r2 == OR_INVALID) FAIL_RUN(OR_INVALID, "partition scatter: invalid updater result");
                                if (r2) FAIL_RUN(r2, "partition plan failed");
                                trace_batch(out, k, iter, b, nd->bins);
                                b->stage = 2;
                            }
                        } else {
                            int r2 = transform_batch(&x, nd, purpose, &pool, b, reserve_bound_of(nd, C, Q));
                            if (r2) FAIL_RUN(r2, "transform batch failed");
                            trace_batch(out, k, iter, b, 1);
                        }
                        dispatched = 1;
                    }
                }
                if (dispatched) progress = 1;
                rr++;
            }
            if (out->status) goto done;
            /* 4) termination */
            int drained = 1;
            for (uint32_t k = 0; k < n_nodes && drained; ++k) {
                if (fifo_size(&ready[k])) drained = 0;
                for (uint32_t j = 0; j < D; ++j) if (lane[k * D + j].armed) drained = 0;
            }
            iter++;
            if (generated >= N && drained && reserved == 0) break;
            if (!progress) FAIL_RUN(OR_STALL, "scheduler stall: no progress possible");
        }
    } else {
        /* Step schedule of the device-resident batched form. Per step, in node order, every lane of
         * a node takes at most one batch: a source lane admits one chunk (the admission rule of
         * GraphRunner.hpp:902-983 with the step's earlier admissions counted), a transform lane the
         * largest credit-fitting prefix of <= Q ready chunks after the earlier lanes' batches, and
         * a sink all ready chunks in one batch. Execution pops output chunks in (node, lane)
         * order; retirement routes outputs and releases inputs in the same order. */
        for (;;) {
            uint32_t res = 0;
            int any = 0;
            for (uint32_t k = 0; k < n_nodes; ++k) {
                const or_node_desc* nd = &nodes[k];
                int stop = 0;
                for (uint32_t j = 0; j < D; ++j) {
                    batch* b = &lane[k * D + j];
                    b->armed = 0;
                    b->n_in = 0;
                    b->n_out = 0;
                    if (stop) continue;
                    if (nd->role == OR_ROLE_SOURCE) {
                        /* up to SC chunks, each admitted by GraphRunner's rule with the earlier
                         * admissions counted; a refusal holds for the rest of the step, so the
                         * node's later lanes stop too. Only a batch's last chunk can be partial
                         * (a fill below C needs the tail of N or a live bound below C, and either
                         * refuses the next chunk), so its items are a dense concatenation. */
                        b->particles = 0;
                        b->first_id = generated;
                        for (uint32_t c = 0; c < SC; ++c) {
                            uint64_t live = generated - sunk, remaining = N - generated;
                            uint64_t min_fill = C < P ? C : P;
                            if (remaining < min_fill) min_fill = remaining;
                            if (!(generated < N && live < P && (P - live) >= min_fill && pool.top - res >= 1u)) {
                                stop = 1;
                                break;
                            }
                            uint64_t fill = remaining;
                            if (fill > P - live) fill = P - live;
                            if (fill > C) fill = C;
                            b->out[c].fill = (uint32_t)fill;
                            b->n_out = c + 1u;
                            b->particles += fill;
                            generated += fill;
                            res += 1;
                        }
                        b->armed = b->n_out != 0;
                    } else if (nd->role == OR_ROLE_TRANSFORM) {
                        uint64_t wanted = fifo_size(&ready[k]) < Q ? fifo_size(&ready[k]) : Q, parts = 0;
                        for (uint64_t i = 0; i < wanted; ++i) parts += fifo_at(&ready[k], i).fill;
                        uint32_t reserve = 0;
                        while (wanted) {
                            uint64_t bound;
                            or_output_chunk_bound(parts, C, nd->bins, &bound);
                            if (pool.top - res >= bound) { reserve = (uint32_t)bound; break; }
                            --wanted;
                            parts -= fifo_at(&ready[k], wanted).fill;
                        }
                        if (wanted) {
                            b->n_in = (uint32_t)wanted;
                            for (uint64_t i = 0; i < wanted; ++i) b->in[i] = fifo_at(&ready[k], i);
                            ready[k].head += wanted;
                            res += reserve;
                            b->armed = 1;
                        } else {
                            stop = 1;
                        }
                    } else {
                        uint64_t n = fifo_size(&ready[k]);
                        if (n) {
                            b->n_in = (uint32_t)n;
                            b->particles = 0;
                            for (uint64_t i = 0; i < n; ++i) { b->in[i] = fifo_at(&ready[k], i); b->particles += b->in[i].fill; }
                            ready[k].head += n;
                            b->armed = 1;
                        }
                        stop = 1; /* one sink batch per step */
                    }
                    any |= b->armed;
                }
            }
            if (!any) {
                int drained = 1;
                for (uint32_t k = 0; k < n_nodes; ++k) if (fifo_size(&ready[k])) drained = 0;
                if (generated >= N && drained) break;
                FAIL_RUN(OR_STALL, "scheduler stall: no progress possible");
            }
            /* execute in (node, lane) order */
            for (uint32_t k = 0; k < n_nodes; ++k)
                for (uint32_t j = 0; j < D; ++j) {
                    const or_node_desc* nd = &nodes[k];
                    batch* b = &lane[k * D + j];
                    if (!b->armed) continue;
                    uint32_t purpose = 2u * k;
                    if (nd->role == OR_ROLE_SOURCE) {
                        uint64_t first = b->first_id;
                        for (uint32_t c = 0; c < b->n_out; ++c) {
                            uint32_t idx = 0;
                            pool_acquire(&pool, &idx);
                            b->out[c].index = idx; b->out[c].gen = pool.gen[idx]; b->out[c].bin = 0;
                            source_batch(&x, nd, purpose, &pool, &b->out[c], first);
                            first += b->out[c].fill;
                        }
                        trace_batch(out, k, iter, b, 0);
                    } else if (nd->role == OR_ROLE_SINK) {
                        sink_batch(&x, nd, &pool, b);
                        trace_batch(out, k, iter, b, 0);
                    } else if (is_partition(nd)) {
                        int r2 = partition_classify(&x, nd, &pool, b);
                        if (r2) FAIL_RUN(r2, "partition classify: device status reports an invalid classifier result");
                        if (!hash_tables_fit(cfg, nd, b)) FAIL_RUN(OR_INVALID, "partition classify: hash bin table full");
                        r2 = partition_plan_scatter(&x, nd, purpose + 1u, &pool, b, reserve_bound_of(nd, C, Q));
                        if (r2 == OR_OVERFLOW) FAIL_RUN(OR_OVERFLOW, "partition scatter: checked updater overflow");
                        if (r2) FAIL_RUN(r2, "partition plan failed");
                        trace_batch(out, k, iter, b, nd->bins);
                    } else {
                        int r2 = transform_batch(&x, nd, purpose, &pool, b, reserve_bound_of(nd, C, Q));
                        if (r2) FAIL_RUN(r2, "transform batch failed");
                        trace_batch(out, k, iter, b, 1);
                    }
                }
            /* retire in (node, lane) order */
            for (uint32_t k = 0; k < n_nodes; ++k)
                for (uint32_t j = 0; j < D; ++j) {
                    const or_node_desc* nd = &nodes[k];
                    batch* b = &lane[k * D + j];
                    if (!b->armed) continue;
                    if (nd->role != OR_ROLE_SINK)
                        for (uint32_t o = 0; o < b->n_out; ++o) {
                            int32_t d = routes[nd->route_offset + b->out[o].bin];
                            if (is_partition(nd) && fused_of(&x, nd, b->out[o].bin) >= 0) d = routes[nodes[d].route_offset];
                            fifo_push(&ready[d], b->out[o]);
                        }
                    if (nd->role == OR_ROLE_SINK) sunk += b->particles;
                    if (nd->role != OR_ROLE_SOURCE)
                        for (uint32_t i = 0; i < b->n_in; ++i) pool_release(&pool, b->in[i].index);
                    if (nd->role == OR_ROLE_TRANSFORM) out->batches++;
                    b->armed = 0;
                }
            iter++;
        }
    }
    if (pool.top != pool.S) fail(out, OR_BAD_ARGUMENT, "chunk leak after the run");
done:
    out->steps = iter;
    out->generated = generated;
    out->sunk = sunk;
    out->peak_live_chunks = pool.peak_live;
    for (uint32_t k = 0; k < n_nodes; ++k) free(ready[k].d);
    for (uint64_t i = 0; i < (uint64_t)n_nodes * D; ++i) {
        free(lane[i].in); free(lane[i].out); free(lane[i].counts);
        free(lane[i].items); free(lane[i].bins);
    }
    free(ready); free(lane);
    free(x.zipf);
    pool_free(&pool);
    return (int)out->status;
}


/* One partition batch over explicit arrays: bins, totals, and the stable rank j of every item
 * inside its bin (the slot is then table[B + binBase[bin] + j/C], j % C). */
int or_partition_items(const or_node_desc* node, const uint32_t* lut, uint64_t seed, uint32_t purpose,
                       uint32_t n_items, const float* weight, const uint32_t* id, const uint32_t* gen,
                       const uint32_t* key, uint32_t* bins_out, uint32_t* totals, uint32_t* rank_out)
{
    (void)seed; (void)purpose;
    or_ctx x;
    memset(&x, 0, sizeof(x));
    x.lut = lut;
    memset(totals, 0, sizeof(uint32_t) * node->bins);
    for (uint32_t i = 0; i < n_items; ++i) {
        or_item it;
        memset(&it, 0, sizeof(it));
        if (weight) it.w = weight[i];
        if (id) it.id = id[i];
        if (gen) it.gen = gen[i];
        if (key) it.key = key[i];
        uint32_t b = or_classify(&x, node, &it);
        if (b >= node->bins) return OR_INVALID;
        bins_out[i] = b;
        rank_out[i] = totals[b]++;
    }
    return OR_OK;
}

/* Schedule-free feedback model (tests/reference/ScalarModel.hpp:84-133 pattern, feedback-network functors). */
/* Schedule-free model of the keyed hash pipeline: key i = Zipf(s) rank from the source's stream
 * (seed, i, 0, purpose 0), value i, bucket fmix32(key) % bins; per bucket the sum of the item hashes
 * and the item count. Every schedule delivers each key to its bucket exactly once, so these equal a
 * run's per-bucket multiset digests whatever C, P, Q and lanes. */
int or_keyed_scalar(uint64_t seed, uint64_t items, uint32_t values, double s, uint32_t bins, uint64_t* multiset,
                    uint64_t* counts, uint32_t threads)
{
    if (!values || !bins) return OR_BAD_ARGUMENT;
    double* cdf = (double*)malloc(sizeof(double) * values);
    if (!cdf) return OR_BAD_ARGUMENT;
    or_zipf_table(values, s, cdf);
    memset(multiset, 0, sizeof(uint64_t) * bins);
    memset(counts, 0, sizeof(uint64_t) * bins);
    const int64_t n = (int64_t)items;
#pragma omp parallel num_threads(threads ? threads : 1)
    {
        uint64_t* ms = (uint64_t*)calloc(bins, sizeof(uint64_t));
        uint64_t* ct = (uint64_t*)calloc(bins, sizeof(uint64_t));
#pragma omp for schedule(static)
        for (int64_t i = 0; i < n; ++i) {
            uint32_t r[4];
            or_stream_words(seed, (uint64_t)i, 0, 0u, r);
            or_item it;
            memset(&it, 0, sizeof(it));
            it.key = or_zipf_sample(cdf, values, r[0], r[1]);
            it.val = (uint64_t)i;
            const uint32_t b = or_fmix32(it.key) % bins;
            ms[b] += item_hash(&it);
            ct[b] += 1u;
        }
#pragma omp critical
        for (uint32_t b = 0; b < bins; ++b) {
            multiset[b] += ms[b];
            counts[b] += ct[b];
        }
        free(ms);
        free(ct);
    }
    free(cdf);
    return OR_OK;
}

int or_feedback_scalar(uint64_t seed, uint64_t items, float threshold, uint32_t distribution, uint32_t ndest,
                       const float* multipliers, uint32_t* weight_bits, uint32_t* gens, uint32_t threads)
{
    (void)threads;
#pragma omp parallel for schedule(static) num_threads(threads ? threads : 1)
    for (int64_t i = 0; i < (int64_t)items; ++i) {
        uint32_t r[4];
        or_stream_words(seed, (uint64_t)i, 0, 0u, r);
        float w = or_uniform_f32(r[0]);
        uint32_t g = 0;
        uint32_t d = route_of(distribution, ndest, or_fmix32((uint32_t)i));
        while (!(w < threshold)) {
            w = w * multipliers[d];
            g += 1u;
        }
        uint32_t bits;
        memcpy(&bits, &w, 4);
        weight_bits[i] = bits;
        if (gens) gens[i] = g;
    }
    return OR_OK;
}

/* Henyey-Greenstein polar cosine by inversion of the HG cumulative distribution (Henyey &
 * Greenstein 1941; inversion as published for MCML, Wang, Jacques & Zheng 1995):
 * ct = (1 + g^2 - ((1 - g^2) / (1 - g + 2 g xi))^2) / (2 g), isotropic 2 xi - 1 when g = 0.
 * (The reference example scatters isotropically, examples/photon_scattering/PhotonFunctors.hpp:172-190;
 * the anisotropic phase function comes from the photon transport workload definition.) */
static double or_hg_cos(double g, double xi)
{
    double ct;
    if (g != 0.0) {
        double tmp = (1.0 - g * g) / (1.0 - g + 2.0 * g * xi);
        ct = (1.0 + g * g - tmp * tmp) / (2.0 * g);
    } else {
        ct = 2.0 * xi - 1.0;
    }
    if (ct > 1.0) ct = 1.0;
    if (ct < -1.0) ct = -1.0;
    return ct;
}

/* New direction after scattering by polar cosine ct and azimuth phi about the old direction u
 * (the standard MCML rotation; near the poles the frame is the lab frame). */
static void or_hg_rotate(double* ux, double* uy, double* uz, double ct, double phi)
{
    double st = sqrt(1.0 - ct * ct), cp = cos(phi), sp = sin(phi);
    if (fabs(*uz) > 0.99999) {
        *ux = st * cp;
        *uy = st * sp;
        *uz = *uz > 0.0 ? ct : -ct;
    } else {
        double tt = sqrt(1.0 - *uz * *uz);
        double nux = st * (*ux * *uz * cp - *uy * sp) / tt + *ux * ct;
        double nuy = st * (*uy * *uz * cp + *ux * sp) / tt + *uy * ct;
        double nuz = -st * cp * tt + *uz * ct;
        *ux = nux;
        *uy = nuy;
        *uz = nuz;
    }
}

int or_hg_cos_f64(double g, const double* xi, double* ct, uint64_t n)
{
    for (uint64_t i = 0; i < n; ++i) ct[i] = or_hg_cos(g, xi[i]);
    return OR_OK;
}

int or_hg_rotate_f64(const double* u, const double* ct, const double* phi, double* out, uint64_t n)
{
    for (uint64_t i = 0; i < n; ++i) {
        double x = u[3 * i], y = u[3 * i + 1], z = u[3 * i + 2];
        or_hg_rotate(&x, &y, &z, ct[i], phi[i]);
        out[3 * i] = x;
        out[3 * i + 1] = y;
        out[3 * i + 2] = z;
    }
    return OR_OK;
}

/* FP64 photon histories on the same Philox words as the fp32 device step. */
int or_photon_scalar(uint64_t seed, uint64_t photons, const or_run_cfg* c, double* tally, uint64_t* steps_out,
                     uint32_t threads)
{
    const uint32_t purpose = 2u; /* step node key 1 */
    const double L = c->layers * c->layer_mfp / c->mu_t;
    uint64_t steps = 0;
    int nt = threads ? (int)threads : 1;
    double* part = (double*)calloc((size_t)nt * c->tally_bins, sizeof(double));
#pragma omp parallel for schedule(dynamic, 1024) num_threads(nt) reduction(+ : steps)
    for (int64_t id = 0; id < (int64_t)photons; ++id) {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        double* t = part + (size_t)tid * c->tally_bins;
        double x = 0, y = 0, z = 0, ux = 0, uy = 0, uz = 1, w = 1;
        uint32_t step = 0;
        while (w > 0.0) {
            uint32_t r[4];
            or_stream_words(seed, (uint64_t)id, step, purpose, r);
            double s = -log((double)((r[0] >> 8) + 1u) * (1.0 / 16777216.0)) / c->mu_t;
            double nx = x + s * ux, ny = y + s * uy, nz = z + s * uz;
            step++;
            steps++;
            if (nz < 0.0 || nz > L) break;
            x = nx; y = ny; z = nz;
            uint32_t bin = (uint32_t)(nz / L * (double)c->tally_bins);
            if (bin >= c->tally_bins) bin = c->tally_bins - 1u;
            t[bin] += w * (1.0 - c->albedo);
            w *= c->albedo;
            if (w < c->roulette_w) {
                double u = (double)(r[3] >> 8) * (1.0 / 16777216.0);
                if (u < c->roulette_p) w /= c->roulette_p;
                else break;
            }
            const double ct = or_hg_cos(c->hg_g, (double)(r[1] >> 8) * (1.0 / 16777216.0));
            const double phi = 6.283185307179586 * ((double)(r[2] >> 8) * (1.0 / 16777216.0));
            or_hg_rotate(&ux, &uy, &uz, ct, phi);
        }
    }
    for (int k = 0; k < nt; ++k)
        for (uint32_t b = 0; b < c->tally_bins; ++b) tally[b] += part[(size_t)k * c->tally_bins + b];
    free(part);
    if (steps_out) *steps_out = steps;
    return OR_OK;
}
