/* graphflow.h — C ABI of the MI355X transform-batch library (libgraphflow_amd.so).
 *
 * This is the drop-in boundary for graphFlow's transform-batch path. Every entry point takes
 * plain pointers and sizes, never throws, and returns a gf_status. Entry points are called from
 * one host thread per context (the reference's single scheduler thread,
 * docs/architecture.md:89-92); device work is stream ordered.
 *
 * Reference interfaces replaced (paths relative to the reference repository):
 *   gf_pool_*        ChunkStorage concept      include/graphFlow/core/ChunkStorage.hpp:35-54
 *                    (leaseCount/isLive/generation/validateLive: core/ChunkLedger.hpp:152-187),
 *                    ChunkPool / ChunkLedger   include/graphFlow/core/ChunkPool.hpp:155-269,
 *                                              include/graphFlow/core/ChunkLedger.hpp:83-188
 *   gf_stage_*       StageBase hooks           include/graphFlow/flow/Execution.hpp:1537-1561
 *                    (= GraphRunner::Hooks     include/graphFlow/sched/GraphRunner.hpp:451-461)
 *   gf_flow_*        Flow<TPayload,TDevice,TStorage> registration/run
 *                                              include/graphFlow/flow/Flow.hpp:85-412
 *   gf_tuning        Tuning                    include/graphFlow/flow/Tuning.hpp:27-145
 *   gf_stats         SchedulerStats            include/graphFlow/sched/GraphRunner.hpp:154-210
 *   gf_status        StatusCode + BatchError / SchedulerStallError / invalid_argument
 *                                              include/graphFlow/flow/Status.hpp:21-96,
 *                                              include/graphFlow/sched/GraphRunner.hpp:62-66,355-378
 *   gf_output_chunk_bound, gf_plan_dense_bins
 *                                              include/graphFlow/core/Reservation.hpp:35-50,
 *                                              include/graphFlow/graph/DenseBinTable.hpp:39-91
 *
 * Arbitrary C++ device functors cannot cross a C ABI, so nodes carry op descriptors: the
 * functor set is compiled into the library (GF_OP_*).
 */
#ifndef GRAPHFLOW_H
#define GRAPHFLOW_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status --------------------------------------------------------------------------- */
typedef enum gf_status {
    GF_OK = 0,
    GF_INVALID = 1,        /* StatusCode::invalid: classifier/updater rejected an item      */
    GF_OVERFLOW = 2,       /* StatusCode::overflow: checked counter / updater overflow      */
    GF_STALL = 3,          /* SchedulerStallError                                            */
    GF_POOL_EXHAUSTED = 4, /* ChunkPool "pool exhausted"                                     */
    GF_BAD_ARGUMENT = 5,   /* std::invalid_argument / std::logic_error                      */
    GF_PLAN_MISMATCH = 6,  /* BatchError from planning / finalise                            */
    GF_DEVICE_ERROR = 7,   /* HIP runtime error                                              */
    GF_STAGE_FAILED = 8    /* a stage is single-use after a failed batch                      */
} gf_status;

/* ---- chunk model (core/Chunk.hpp:19-37, graph/TileDescriptor.hpp:25-30) ----------------- */
typedef struct gf_chunk_handle { uint32_t index, generation; } gf_chunk_handle;
typedef struct gf_chunk_desc { uint32_t index, generation, fill, bin; } gf_chunk_desc;
typedef struct gf_tile_desc { uint32_t chunk_index, tile_offset, count; } gf_tile_desc;

/* ---- payloads --------------------------------------------------------------------------- */
enum {
    GF_PAYLOAD_PARTICLE = 0, /* weight f32, id u32, gen u32                      (12 B) */
    GF_PAYLOAD_KEYED = 1,    /* key u32, value u64                               (12 B) */
    GF_PAYLOAD_PHOTON = 2    /* x,y,z,ux,uy,uz,w f32, id u32, step u32           (36 B) */
};

/* ---- node roles and ops ------------------------------------------------------------------ */
enum { GF_ROLE_SOURCE = 0, GF_ROLE_TRANSFORM = 1, GF_ROLE_SINK = 2 };

enum {
    GF_SRC_PARTICLE_UNIFORM = 1, /* weight = U(0,1] from Philox word 0, id = index, gen = 0       */
    GF_SRC_KEYED_ZIPF = 2,       /* key = Zipf(f[0]) rank over u[0] values, value = index        */
    GF_SRC_PHOTON = 3,           /* origin, +z direction, weight 1                                */
    GF_CLS_WEIGHT_ROUTE = 16,    /* bin = w < f[0] ? 0 : 1 + route(fmix32(id)), u[0] distribution */
    GF_CLS_KEY_BUCKET = 17,      /* bin = fmix32(key) % bins                                      */
    GF_CLS_ID_MOD = 18,          /* u[0]: 0 balanced, 1 one-hot, 2 skewed                          */
    GF_CLS_LUT = 19,             /* bin = lut[u[3] + min(id,u[0]-1)*u[1] + min(gen,u[1]-1)], ptr[0] */
    GF_CLS_PHOTON_ALIVE = 20,    /* bin = weight > 0 ? 0 : 1                                       */
    GF_XF_SCALE_WEIGHT = 32,     /* weight *= f[0]; gen += 1                                       */
    GF_XF_PHOTON_STEP = 33,      /* one flight + interaction + roulette + HG scatter; tally ptr[0] */
    GF_SNK_HIST = 48,            /* ptr[0] u64 counts of field u[0] (0 weight bits, 2 gen) - u[1], size u[2] */
    GF_SNK_RECORD_WEIGHT = 49,   /* ptr[0] u32 record[id] = weight bits                            */
    GF_SNK_CHUNK_DIGEST = 50,    /* ptr[0] u32x4 log per sunk chunk, size[0] entries; ptr[1] u64x2 per bin */
    GF_SNK_PHOTON_TALLY = 51     /* no-op sink for finished photons                                */
};
enum { GF_UPD_NONE = 0, GF_UPD_ADD_ON_BIN = 1, GF_UPD_GEN_TO_BIN = 2 };
enum { GF_ROUTE_PARITY = 0, GF_ROUTE_SKEW_90_10 = 1, GF_ROUTE_ALL_ONE = 2, GF_ROUTE_UNIFORM = 3 };

typedef struct gf_node_desc {
    uint32_t role, op, upd, bins;
    uint32_t route_offset; /* filled by gf_flow_add_node */
    uint32_t u[4];
    float f[4];
    uint64_t ptr[2];  /* device buffers of sinks / LUT classifiers */
    uint64_t size[2]; /* element counts of ptr[] */
} gf_node_desc;

/* ---- tuning / config (flow/Tuning.hpp) ---------------------------------------------------- */
enum { GF_SCHEDULE_HOST = 0 /* plain: host GraphRunner, host planning */,
       GF_SCHEDULE_DEVICE = 1 /* batched: device-resident step engine  */ };

typedef struct gf_tuning {
    uint64_t seed;
    uint32_t batch_chunks;   /* Q (0 -> 1)                              */
    uint32_t chunk_capacity; /* C (0 -> 8192)                           */
    uint32_t live_particles; /* P (0 -> min(N, 8*C))                    */
    uint32_t num_chunks;     /* S (0 -> P + K*R)                        */
    uint32_t tile_items;     /* items per workgroup tile (0 -> autotune) */
    uint32_t schedule;       /* GF_SCHEDULE_*                           */
    uint32_t steps_per_launch; /* device engine steps per graph launch (0 -> 64) */
    uint32_t trace;          /* record a batch trace                    */
    uint32_t use_graph;      /* 0 direct launches, 1 hipGraph of K steps, 2 direct launches timed with HIP events,
                                3 as 2 plus the device engines' per-phase wall-clock profile */
    uint32_t fuse_transforms;/* device schedule: apply a plain transform fed by a partition bin inside
                                that partition's scatter and route its items to the transform's destination */
    uint32_t max_workgroups; /* device schedule: workgroups per step kernel (0 -> 64); team size of
                                the team engine (0 -> the CUs of one XCD, at most 32) */
    uint32_t device_engine;  /* device schedule: 0 autotuned, 1 wide (multi-workgroup step kernels in a
                                hipGraph), 2 micro (one persistent 1024-thread workgroup runs the steps),
                                3 team (one persistent launch, a team of workgroups on one XCD with a
                                replicated scheduler and one barrier per step) */
    uint32_t photon_layers, photon_tally_bins;
    double photon_layer_mfp, photon_mu_t, photon_albedo, photon_hg_g, photon_roulette_w, photon_roulette_p;
    uint32_t team_xcd;       /* team engine: XCD the team runs on, plus one (0 -> the XCD of the first
                                workgroup to start); flows that run side by side use different XCDs */
    uint32_t inflight_batches; /* Tuning::inflightBatches: batches in flight per source and transform node
                                  (0 -> 1); plain schedule: lanes of the host runner; device schedule: a
                                  node runs up to this many batches per step (wide and micro engines) */
    uint32_t bin_path;       /* Tuning::binPath: GF_BIN_PATH_* (hash above 4096 bins when automatic)      */
    uint32_t hash_table_entries; /* hash path: slots of a tile's bin table (0 -> nextPow2(2 * tile items), which
                                    never fills); a tile with more distinct bins fails its batch (GF_INVALID) */
    uint32_t stall_timeout_ms; /* Tuning::stallTimeout: plain schedule, longest wait for a device stage
                                  before the run fails with GF_STALL (0 -> no limit)                       */
    uint32_t fused_batches;  /* plain schedule: each partition batch as one fused batch (gf_transform_batch)
                                instead of classify -> host plan -> scatter                                */
    uint32_t source_chunks;  /* device schedule (wide and micro engines): chunks a source lane admits per
                                step, each by the admission rule with the earlier ones counted (0 -> 1,
                                at most 64); the plain schedule admits one chunk per dispatch            */
} gf_tuning;

enum { GF_BIN_PATH_AUTO = 0, GF_BIN_PATH_DENSE = 1, GF_BIN_PATH_HASH = 2 };

typedef struct gf_stats {
    uint64_t completed_batches, steps, generated, sunk;
    uint32_t pool_chunks, peak_live_chunks, chunk_capacity, live_particles;
    uint64_t kernel_launches;
    double wall_ms;
    uint32_t engine;         /* engine that ran: 0 plain host-planned, 1 wide, 2 micro, 3 team,
                                4 team handing over to wide */
    uint32_t handoff_step;   /* step at which the team handed over (engine 4)              */
} gf_stats;

typedef struct gf_context gf_context;
typedef struct gf_pool gf_pool;
typedef struct gf_stage gf_stage;
typedef struct gf_flow gf_flow;

/* ---- context ------------------------------------------------------------------------------ */
gf_status gf_context_create(int device, gf_context** out);
gf_status gf_context_destroy(gf_context* ctx);
/* message of the last failed call that reported through the context (gf_server_create) */
gf_status gf_context_last_error(gf_context* ctx, char* buffer, uint32_t size);
const char* gf_status_name(gf_status s);
uint32_t gf_abi_version(void);

/* ---- closed-form plan helpers (host) ------------------------------------------------------- */
gf_status gf_output_chunk_bound(uint64_t n, uint64_t capacity, uint64_t max_bins, uint64_t* out);
gf_status gf_plan_dense_bins(const uint32_t* totals, uint32_t bins, uint32_t capacity, uint32_t* bin_base,
                             uint32_t* fills, uint32_t fills_cap, uint32_t* num_output_chunks);

/* ---- pool: ChunkStorage contract (ChunkStorage.hpp:35-54) ----------------------------------- */
gf_status gf_pool_create(gf_context* ctx, uint32_t payload, uint32_t num_chunks, uint32_t capacity, gf_pool** out);
gf_status gf_pool_destroy(gf_pool* pool);
gf_status gf_pool_acquire(gf_pool* pool, gf_chunk_handle* out);
gf_status gf_pool_release(gf_pool* pool, gf_chunk_handle handle);
gf_status gf_pool_acquire_lease(gf_pool* pool, gf_chunk_handle handle);
gf_status gf_pool_release_lease(gf_pool* pool, gf_chunk_handle handle);
/* counts[0..5] = numChunks, chunkCapacity, freeCount, liveCount, peakLive, peakLeases */
gf_status gf_pool_counts(gf_pool* pool, uint32_t counts[6]);
gf_status gf_pool_base(gf_pool* pool, gf_chunk_handle handle, uint64_t* slot);
/* leaseCount(handle): reader leases of the chunk; GF_BAD_ARGUMENT for an index out of range */
gf_status gf_pool_lease_count(gf_pool* pool, gf_chunk_handle handle, uint32_t* count);
/* isLive(handle): 1 when the index is in range and the generation matches, else 0 */
gf_status gf_pool_is_live(gf_pool* pool, gf_chunk_handle handle, int* live);
/* generation(chunkIndex): the chunk's current generation; GF_BAD_ARGUMENT out of range */
gf_status gf_pool_generation(gf_pool* pool, uint32_t index, uint32_t* generation);
/* validateLive(handle): GF_OK, or GF_BAD_ARGUMENT ("chunk index out of range" / "stale chunk handle") */
gf_status gf_pool_validate_live(gf_pool* pool, gf_chunk_handle handle);
gf_status gf_pool_column(gf_pool* pool, uint32_t column, void** device_ptr, uint32_t* elem_bytes);
gf_status gf_pool_write(gf_pool* pool, uint32_t column, uint64_t first_slot, const void* host, uint64_t count);
gf_status gf_pool_read(gf_pool* pool, uint32_t column, uint64_t first_slot, void* host, uint64_t count);

/* ---- stage: one node's batch protocol (Execution.hpp:1537-1561) -----------------------------
 * launch:  enqueue the first device stage of a batch (classify for a partition, the whole
 *          transform / source / sink kernel otherwise).
 * advance: called after the previous device stage completed; *more_work = 1 when a further
 *          device stage was enqueued; *allocated = output chunks acquired by this call.
 * finish:  copy the published output descriptors (bin-major, then chunk order).
 * abort:   release unpublished outputs; the stage is single-use afterwards. */
gf_status gf_stage_create(gf_pool* pool, const gf_node_desc* node, uint32_t node_key, uint64_t seed,
                          uint32_t batch_chunks, uint32_t tile_items, gf_stage** out);
/* The same with the run's Tuning: seed, batch_chunks (Q), tile_items, bin_path, hash_table_entries,
 * stall_timeout_ms and the photon slab (a reference stage is built from the resolved Config,
 * flow/Flow.hpp:328-412). */
gf_status gf_stage_create_tuned(gf_pool* pool, const gf_node_desc* node, uint32_t node_key, const gf_tuning* tuning,
                                gf_stage** out);
gf_status gf_stage_destroy(gf_stage* stage);
gf_status gf_stage_launch(gf_stage* stage, const gf_chunk_desc* chunks, uint32_t n, uint64_t first_id, void* stream);
gf_status gf_stage_advance(gf_stage* stage, void* stream, uint32_t* allocated, int* more_work);
gf_status gf_stage_finish(gf_stage* stage, gf_chunk_desc* outputs, uint32_t cap, uint32_t* n_out);
gf_status gf_stage_abort(gf_stage* stage);
/* One partition batch with no host round trip (no counterpart kernel in the reference, which plans on
 * the host between its classify and scatter kernels, flow/Execution.hpp:2363-2725): classify, device
 * plan, device output acquisition and stable scatter of a ragged list of n <= Q chunks, in one launch
 * on one workgroup, or for batches of more than one round of that workgroup's tiles (and n <= 256)
 * in three stream-ordered launches over the grid (count, plan, scatter), same result. The stage
 * reserves outputChunkBound(items, C, bins) chunks from its pool before the launch; the kernel
 * uses the first R in bin-major order. gf_stage_advance (one call, more_work = 0) returns the
 * unused reservation and reports allocated = R; gf_stage_finish copies the R outputs. Plain
 * transforms, sources and sinks run their one kernel as with gf_stage_launch. */
gf_status gf_transform_batch(gf_stage* stage, const gf_chunk_desc* chunks, uint32_t n, uint64_t first_id, void* stream);
/* gf_transform_batch without allocation. No counterpart in the reference: for callers that cannot
 * allocate device memory (code under hipGraph capture, tight serving loops). Every argument but
 * pool, node and tuning is device memory the caller owns: chunks (n input descriptors), reserved
 * (n_reserved output handles, normally outputChunkBound(items, C, bins) popped from the pool's free
 * list), outputs (n_reserved descriptors; the first R are written), result ({R, status}; status 0,
 * GF_INVALID, GF_OVERFLOW or GF_PLAN_MISMATCH; 17 = a hash bin table was full). The kernel's
 * scratch is a static device array of 4 Mi words per device (16 MiB, held until the process
 * ends): GF_BAD_ARGUMENT when n full chunks' tile counts would not fit it. Calls on one device
 * share that array, so calls that overlap on one device race; the caller returns the unused
 * reservation (handles R .. n_reserved-1). Stream ordered, asynchronous, no host synchronisation. */
gf_status gf_transform_batch_noalloc(gf_pool* pool, const gf_node_desc* node, uint32_t node_key, const gf_tuning* tuning,
                                     const gf_chunk_desc* chunks, uint32_t n, const gf_chunk_handle* reserved,
                                     uint32_t n_reserved, gf_chunk_desc* outputs, uint32_t* result, void* stream);
/* message of the stage's last failed call */
gf_status gf_stage_last_error(gf_stage* stage, char* buffer, uint32_t size);
/* per-bin totals of the last classify (bins entries) */
gf_status gf_stage_counts(gf_stage* stage, uint32_t* counts, uint32_t bins);
/* tile descriptors of the last launch, flattened on the device (TileDescriptor list) */
gf_status gf_stage_tiles(gf_stage* stage, gf_tile_desc* tiles, uint32_t cap, uint32_t* n);

/* ---- flow: typed graph registration and run (Flow.hpp:85-412) ------------------------------- */
gf_status gf_flow_create(gf_context* ctx, uint32_t payload, gf_flow** out);
gf_status gf_flow_destroy(gf_flow* flow);
gf_status gf_flow_add_node(gf_flow* flow, const char* name, const gf_node_desc* node, uint32_t* key);
gf_status gf_flow_connect(gf_flow* flow, uint32_t from, uint32_t bin, uint32_t to);
gf_status gf_flow_configure(gf_flow* flow, const gf_tuning* tuning);
gf_status gf_flow_run(gf_flow* flow, uint32_t source, uint64_t items, gf_stats* stats);
/* autotuner (no counterpart in the reference): probe runs of probe_items (0 -> 2^21) over the
 * device-schedule knobs left open in the flow's tuning (device_engine 0: team sizes, wide-engine
 * tile sizes and grids of 64/256/1024 workgroups, the one-workgroup engine; batch_chunks 0: Q in {4, 8, 16}); the fastest
 * configures the flow and is returned with its probe time. The probe runs write the flow's sinks. */
gf_status gf_flow_autotune(gf_flow* flow, uint32_t source, uint64_t probe_items, gf_tuning* chosen, double* best_ms);
gf_status gf_flow_last_error(gf_flow* flow, char* buffer, uint32_t size);
gf_status gf_flow_trace(gf_flow* flow, uint32_t* host, uint64_t cap, uint64_t* len);
/* kernel timing of the last device-engine run with use_graph 2 or 3: total ms and launches per kind
 * (0: classify, 1: plan, 2: scatter + retire/select of the wide engine, 3: team / one-workgroup engine
 * launches, 4: partition and transform scatter, 5: source tiles, 6: sink tiles, 7: retire/select;
 * kinds beyond 7 read 0) */
gf_status gf_flow_kernel_times(gf_flow* flow, double* ms, uint64_t* launches, uint32_t kinds);
/* micro or team engine run with use_graph = 3: wall-clock ticks (100 MHz) per step phase and the
 * number of steps (micro: count, plan, source, partition scatter, transform, sink, retire, select,
 * steps; team: words 9-17 phase ticks, word 14 steps, words 18-31 shader cycles of sub-phases, then
 * 8 words per team member: barrier wait, counts+plan, scatter phase, scheduler wave, segments,
 * sink segments) */
gf_status gf_flow_phase_profile(gf_flow* flow, uint64_t* ticks, uint32_t words);
/* algorithmic bytes moved by the last run's partition/transform/source/sink item passes */
gf_status gf_flow_item_passes(gf_flow* flow, uint64_t* partition_items, uint64_t* transform_items);

/* ---- serving: independent flow jobs admitted into slots as they arrive ------------------------
 * No counterpart in the reference (its runner executes one flow per run). Jobs are one of three
 * topologies: the particle feedback network (result: digest of the per-id weight bits), the keyed
 * murmur3 hash pipeline with 4096 buckets (result: digest of the per-bucket multisets) and photon
 * transport (result: the 1024-bin absorbed-weight tally). A topology's `slots` job slots live in
 * its device engines, each over copies of the topology's graph (one per slot) whose every step
 * runs the chunks of all its admitted jobs together: team engines of up to 6 (feedback) or 8
 * (photon) slots on distinct XCDs, one wide engine for the hash pipeline. One driver thread admits
 * jobs first come, first served within a topology into a free slot whose engine's unreserved
 * chunks hold the job's pool bound (live bound + the graph's output bounds), launches the engines
 * side by side and finishes jobs. Submission, polling and destruction may be called from one host
 * thread while the engines run. */
enum { GF_JOB_FEEDBACK = 0, GF_JOB_HASH = 1, GF_JOB_PHOTON = 2 };

typedef struct gf_server_config {
    uint32_t slots;             /* job slots per topology (0 -> 16, at most 42)                     */
    uint32_t team_size;         /* team engine workgroups (0 -> CUs of one XCD minus 4)             */
    uint64_t memory_budget;     /* device bytes of the engines' pools and result buffers (0 -> 64 GiB) */
    uint64_t max_items;         /* largest job (0 -> 1e7; sizes the feedback result buffers)        */
    uint32_t prewarm;           /* unused: the engines are built before create returns              */
    uint32_t reserved;
    uint64_t topology_bytes[3]; /* out: device bytes each topology's engine holds                   */
} gf_server_config;

typedef struct gf_job_desc {
    uint32_t topology;          /* GF_JOB_* */
    uint32_t reserved;
    uint64_t items;             /* N: particles, keys or photons */
    uint64_t seed;
} gf_job_desc;

typedef struct gf_job_result {
    uint64_t job;
    uint32_t topology, slot;
    uint32_t status;            /* gf_status of the job's run */
    uint32_t engine;            /* gf_stats.engine of the run */
    uint64_t items, steps;
    double submit_ms, start_ms, end_ms; /* server clock: submitted, started on its slot, finished */
    uint64_t digest;            /* feedback / hash result digest (order independent) */
    double tally_sum;           /* photon: total absorbed weight */
} gf_job_result;

typedef struct gf_server gf_server;
gf_status gf_server_create(gf_context* ctx, const gf_server_config* cfg, gf_server** out);
gf_status gf_server_submit(gf_server* server, const gf_job_desc* job, uint64_t* job_id);
/* completed jobs since the last poll (up to cap); wait_all = 1 first waits until every submitted
 * job has finished */
gf_status gf_server_poll(gf_server* server, gf_job_result* out, uint32_t cap, uint32_t* n, uint32_t wait_all);
gf_status gf_server_job_tally(gf_server* server, uint64_t job, double* tally, uint32_t bins);
gf_status gf_server_last_error(gf_server* server, char* buffer, uint32_t size);
gf_status gf_server_destroy(gf_server* server);

/* ---- device histogram rows: compact a dense u64 count array into ascending (value,count) -- */
gf_status gf_hist_rows(gf_context* ctx, const uint64_t* device_counts, uint64_t n, uint32_t lo,
                       uint64_t* host_rows /* 2 per row */, uint64_t cap, uint64_t* rows);

#ifdef __cplusplus
}
#endif
#endif
