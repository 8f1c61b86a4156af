/* CPU oracle of graphFlow's transform-batch path.
 *
 * TEST INFRASTRUCTURE ONLY. This library checks results: only tests/, the smoke check and the
 * CPU-baseline leg of bench.py load it. The product library (graphflow_amd) never links,
 * loads or calls anything here, and nothing here is ever the thing being measured as GPU
 * throughput.
 *
 * It restates, in plain C loops over host arrays, the semantics of the reference's
 *   - reservation math           include/graphFlow/core/Reservation.hpp:15-71
 *   - dense output plan          include/graphFlow/graph/DenseBinTable.hpp:39-91
 *   - tile flattening            include/graphFlow/flow/Execution.hpp:113-124, 2377-2389
 *   - frame count                include/graphFlow/core/LaunchPolicy.hpp:46-76
 *   - hash table geometry        include/graphFlow/graph/BinTable.hpp:49-93
 *   - stream key mapping         include/graphFlow/rng/ParticleRng.hpp:51-129
 *   - chunk free list            include/graphFlow/core/ChunkLedger.hpp:36-134
 *   - partition batch protocol   include/graphFlow/flow/Execution.hpp:2363-2725
 *   - transform batch            include/graphFlow/flow/Execution.hpp:708-776, 1805-1901
 *   - scheduler loop             include/graphFlow/sched/GraphRunner.hpp:716-1257
 *   - pool derivation            include/graphFlow/flow/Flow.hpp:580-626
 * with the benchmark functors (fp32 feedback network, murmur3/Zipf hash pipeline, layered-slab
 * photon transport in FP64).
 *
 * Pinning status of every function is listed in the header comment of gf_oracle.c.
 */
#ifndef GF_ORACLE_H
#define GF_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    OR_OK = 0,
    OR_INVALID = 1,      /* invalid classifier / updater result (StatusCode::invalid)     */
    OR_OVERFLOW = 2,     /* checked overflow (StatusCode::overflow)                       */
    OR_STALL = 3,        /* scheduler cannot make progress (SchedulerStallError)           */
    OR_POOL_EXHAUSTED = 4,
    OR_BAD_ARGUMENT = 5, /* std::invalid_argument / logic_error on configuration          */
    OR_PLAN_MISMATCH = 6 /* BatchError: plan / count mismatch                              */
};

/* ---- closed-form primitives ---------------------------------------------------------- */
int or_ceil_div(uint64_t n, uint64_t d, uint64_t* out);
int or_output_chunk_bound(uint64_t n, uint64_t capacity, uint64_t max_bins, uint64_t* out);
int or_plan_dense_bins(const uint32_t* totals, uint32_t bins, uint32_t capacity, uint32_t* bin_base,
                       uint32_t* fills, uint32_t fills_cap, uint32_t* num_output_chunks);
int or_build_dense_table(const uint32_t* bin_base, uint32_t bins, const uint32_t* chunk_indices,
                         uint32_t num_output_chunks, uint32_t* table);
int or_tile_count(uint32_t fill, uint32_t tile_size, uint32_t* out);
int or_flatten_tiles(const uint32_t* chunk_index, const uint32_t* fill, uint32_t n, uint32_t tile_size,
                     uint32_t cap, uint32_t* tiles /* 3 words per tile */, uint32_t* num_tiles);
int or_num_frames_for(uint64_t work, uint32_t tile_size, uint32_t frame_cap, uint32_t max_threads,
                      uint32_t max_blocks, uint32_t sms, uint32_t* out);
uint32_t or_next_pow2(uint32_t v);
uint32_t or_hash_bin_slot(uint32_t bin, uint32_t mask);
int or_hash_table_insert(uint32_t* keys, uint32_t* counts, uint32_t entries, uint32_t bin);

/* ---- RNG ----------------------------------------------------------------------------- */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
uint64_t or_mix64(uint64_t v);
uint64_t or_stream_key(uint64_t seed, uint64_t id, uint64_t event, uint32_t purpose);
void or_stream_words(uint64_t seed, uint64_t id, uint64_t event, uint32_t purpose, uint32_t out[4]);
uint32_t or_fmix32(uint32_t h);
float or_uniform_f32(uint32_t word);
double or_uniform_f64(uint32_t w0, uint32_t w1);

/* Zipf inverse-CDF table: cdf[r] = sum_{i<=r} (i+1)^-s, r < values. */
void or_zipf_table(uint32_t values, double s, double* cdf);
uint32_t or_zipf_sample(const double* cdf, uint32_t values, uint32_t w0, uint32_t w1);

/* ---- graph description ---------------------------------------------------------------- */
enum { OR_ROLE_SOURCE = 0, OR_ROLE_TRANSFORM = 1, OR_ROLE_SINK = 2 };

enum {
    /* sources */
    OR_SRC_PARTICLE_UNIFORM = 1, /* weight=U(0,1], id=index, gen=0                                  */
    OR_SRC_KEYED_ZIPF = 2,       /* key=Zipf rank (u[0] values, f[0] exponent), value=index        */
    OR_SRC_PHOTON = 3,
    /* partition classifiers */
    OR_CLS_WEIGHT_ROUTE = 16, /* bin = w<f[0] ? 0 : 1+route(fmix32(id)); u[0]=distribution           */
    OR_CLS_KEY_BUCKET = 17,   /* bin = fmix32(key) % bins                                          */
    OR_CLS_ID_MOD = 18,       /* u[0]: 0 balanced id%B, 1 one-hot 0, 2 skewed (id%8==0 ? id%B : 0)  */
    OR_CLS_LUT = 19,          /* bin = lut[min(id,u[0]-1)*u[1] + min(gen,u[1]-1)]                   */
    OR_CLS_PHOTON_ALIVE = 20, /* bin = alive ? 0 : 1                                                */
    /* plain transforms */
    OR_XF_SCALE_WEIGHT = 32, /* w *= f[0]; gen++                                                    */
    OR_XF_PHOTON_STEP = 33,
    /* sinks */
    OR_SNK_HIST = 48,          /* count of field u[0] (0 weight bits, 2 gen) in [u[1], u[1]+u[2]) */
    OR_SNK_RECORD_WEIGHT = 49, /* record[id] = weight bits                                        */
    OR_SNK_CHUNK_DIGEST = 50,  /* per sunk chunk: {bin, fill, order hash lo, hi}                   */
    OR_SNK_PHOTON_TALLY = 51
};

enum { OR_UPD_NONE = 0, OR_UPD_ADD_ON_BIN = 1 /* bin==u[2]: w += f[1]; gen++ (overflow if inf) */,
       OR_UPD_GEN_TO_BIN = 2 /* gen = bin */ };

enum { OR_ROUTE_PARITY = 0, OR_ROUTE_SKEW_90_10 = 1, OR_ROUTE_ALL_ONE = 2, OR_ROUTE_UNIFORM = 3 };

typedef struct or_node_desc {
    uint32_t role;
    uint32_t op;
    uint32_t upd;
    uint32_t bins;         /* partition width; 1 for plain transform/source; 0 for sinks */
    uint32_t route_offset; /* first entry of this node's bins-sized route list          */
    uint32_t u[4];
    float f[4];
} or_node_desc;

enum { OR_MODE_RUNNER = 0, OR_MODE_STEP = 1 };

typedef struct or_run_cfg {
    uint64_t seed;
    uint64_t items;     /* N */
    uint32_t capacity;  /* C */
    uint32_t live;      /* P; 0 derives min(N, 8*C) */
    uint32_t batch;     /* Q */
    uint32_t num_chunks;/* S; 0 derives P + K*R */
    uint32_t mode;      /* OR_MODE_RUNNER / OR_MODE_STEP */
    uint32_t threads;   /* OpenMP threads for the batch loops (results do not depend on it) */
    uint32_t fuse;      /* step mode: a plain transform fed by a partition bin runs inside that
                           partition's scatter; its items go to the transform's destination */
    /* photon slab (OR_XF_PHOTON_STEP / OR_SNK_PHOTON_TALLY) */
    uint32_t layers;
    uint32_t tally_bins;
    double layer_mfp, mu_t, albedo, hg_g, roulette_w, roulette_p;
    uint32_t inflight;     /* Tuning::inflightBatches: batches in flight per node (0 -> 1)          */
    uint32_t bin_path;     /* BinPath: 0 automatic (hash above 4096 bins), 1 dense, 2 hash           */
    uint32_t hash_tile;    /* hash path: items per classify tile (the device's tile size)            */
    uint32_t hash_entries; /* hash path: slots of a tile's bin table (0 -> nextPow2(2 * tile), never full) */
    uint32_t fused_stages; /* runner mode: a partition batch classifies, plans, acquires and scatters at
                              dispatch (one launch) and retires at the next iteration             */
    uint32_t source_chunks;/* step mode: chunks a source lane admits per step (0 -> 1), each by the
                              admission rule with the step's earlier admissions counted          */
} or_run_cfg;

typedef struct or_run_out {
    uint32_t status;
    char message[160];
    uint32_t pool_chunks;
    uint32_t peak_live_chunks;
    uint64_t batches;        /* completed non-source, non-sink transform batches */
    uint64_t steps;          /* loop iterations (runner) or steps (step mode)     */
    uint64_t generated;
    uint64_t sunk;
    /* trace of every planned batch (optional): see or_run in gf_oracle.c for the layout */
    uint32_t* trace;
    uint64_t trace_cap;
    uint64_t trace_len;
    /* sink outputs (caller allocated, may be null when unused) */
    uint64_t* hist;          /* OR_SNK_HIST counts, size u[2]                     */
    uint32_t* record;        /* OR_SNK_RECORD_WEIGHT, size N                      */
    uint32_t* digest_log;    /* OR_SNK_CHUNK_DIGEST, 4 words per sunk chunk       */
    uint64_t digest_cap;
    uint64_t digest_len;
    double* tally;           /* OR_SNK_PHOTON_TALLY absorbed weight per depth bin */
    uint64_t* bin_multiset;  /* OR_SNK_CHUNK_DIGEST: per bin sum of item hashes, size bins */
    uint32_t bin_multiset_bins;
    uint64_t* bin_items;     /* per bin item count */
} or_run_out;

int or_derive_pool(const or_node_desc* nodes, uint32_t n_nodes, uint64_t items, uint32_t capacity,
                   uint32_t live, uint32_t* live_out, uint32_t* pool_out);
int or_run(const or_node_desc* nodes, uint32_t n_nodes, const int32_t* routes, const uint32_t* lut,
           const or_run_cfg* cfg, or_run_out* out);

/* One partition batch on explicit host arrays (the hot path in isolation).
 * Input: n_items records (SoA) in batch order; output: per bin totals, the stable slot of every
 * item (flat position bin_base*C + j), and the payload after the updater. */
int or_partition_items(const or_node_desc* node, const uint32_t* lut, uint64_t seed, uint32_t purpose,
                       uint32_t n_items, const float* weight, const uint32_t* id, const uint32_t* gen,
                       const uint32_t* key, uint32_t* bins_out, uint32_t* totals, uint32_t* rank_out);

/* Schedule-free per-particle model of the fp32 feedback network (ScalarModel.hpp:84-133 pattern):
 * returns the final weight bits per id. */
int or_feedback_scalar(uint64_t seed, uint64_t items, float threshold, uint32_t distribution, uint32_t ndest,
                       const float* multipliers, uint32_t* weight_bits, uint32_t* gens, uint32_t threads);

/* Schedule-free model of the keyed hash pipeline: per-bucket item-hash sums and counts. */
int or_keyed_scalar(uint64_t seed, uint64_t items, uint32_t values, double s, uint32_t bins, uint64_t* multiset,
                    uint64_t* counts, uint32_t threads);

/* FP64 photon histories, schedule free, same RNG words as the device (ScalarTransport.hpp:87-211 pattern). */
int or_hg_cos_f64(double g, const double* xi, double* ct, uint64_t n);
int or_hg_rotate_f64(const double* u, const double* ct, const double* phi, double* out, uint64_t n);
int or_photon_scalar(uint64_t seed, uint64_t photons, const or_run_cfg* cfg, double* tally, uint64_t* steps_out,
                     uint32_t threads);

#ifdef __cplusplus
}
#endif
#endif
