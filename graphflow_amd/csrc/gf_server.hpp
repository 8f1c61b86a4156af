// Serving loop: independent flow jobs (particle feedback, keyed hash pipeline, photon transport)
// arrive over time and run continuously batched on the device.
//
// A served engine runs the union of M copies of a topology's graph, one copy per job slot (a
// served flow). Every step of the engine selects, classifies, plans and scatters the ready chunks of
// all its admitted jobs together (one team step, or one set of wide-engine launches over every
// slot's lanes), so jobs share the step and its fixed cost; a node's batch (<= Q chunks) stays
// within one job, as in the reference's per-flow batching, which keeps each job's result equal to
// its result alone. A slot's copy keeps the job's stream purposes (2 x node id in the job's own
// graph) and takes the job's seed from the engine's job table. The particle feedback and photon
// topologies run on team engines (one XCD each, at most 8 slots of at most 32 nodes); a team step
// that does not fit the team's limits hands the engine to the wide engine until it is idle. The
// hash pipeline's 4096-bucket partition runs on one wide engine.
//
// One host thread (the driver) runs the loop: it admits queued jobs into free slots of engines with no
// launch in flight, writes their items, live bounds and seeds into the job tables between launches,
// and keeps every busy engine launching on its own stream (a team launch of up to kServeTeamSteps
// steps, or a hipGraph of kServeSteps wide steps; the engines run side by side on the device, each
// at its own pace). A launch is followed by copies of the engine's job table into pinned host memory
// and an event; when the event completes the driver finishes the jobs whose items are all sunk.
//
// Admission (first come, first served within a topology): a job needs a free slot on one of its
// topology's engines whose unreserved chunks hold the job's pool bound S_j = P_j + the graph's
// transform output bounds (flow/Flow.hpp:580-626; P_j = the job's live bound = min(N, P)). An
// engine's pool is its share of the topology's third of the memory budget; a job whose bound
// exceeds a whole pool runs with its live bound lowered to fit.
//
// There is no serving loop in the reference; the flows are the reference's own networks
// (examples/particle_feedback, the hash classify pipeline of flow/Execution.hpp:1207-1502 and
// examples/photon_scattering), and each job's result is the result of its flow run alone.
//
// Included at the end of gf_runtime.hip (uses the flow and engine internals defined there).

#include <condition_variable>
#include <mutex>
#include <thread>
#include <unordered_map>

namespace {

// Order-independent digest of a record-weight result: sum over ids of mix64(id << 32 | bits).
__global__ void k_record_digest(const uint32_t* rec, uint64_t n, unsigned long long* out)
{
    unsigned long long acc = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        acc += gfd::mix64(i << 32 | rec[i]);
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o);
    if ((threadIdx.x & 63u) == 0u) atomicAdd(out, acc);
}

__global__ void k_fill_u32(uint32_t* p, uint64_t n, uint32_t v)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) p[i] = v;
}

__global__ void k_zero_u64(unsigned long long* p, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) p[i] = 0ull;
}

constexpr uint32_t kServeLive = 65536;      // live bound of a served feedback job
constexpr uint32_t kServeHashLive = 65536;  // keys in flight per served hash-pipeline job
constexpr uint32_t kServePhotonLive = 32768; // photons in flight per served photon job (fewer dependent steps)
constexpr uint32_t kServeCap = 8192;        // chunk capacity
constexpr uint32_t kServeBatch = 8;         // Q
constexpr uint32_t kServeSourceChunks = 8;  // chunks a wide engine's source lane admits per step
constexpr uint32_t kServeBuckets = 4096;
constexpr uint32_t kServeTallyBins = 1024;
constexpr uint32_t kServeSlots = 16;        // job slots per topology (default)
constexpr uint32_t kServeSteps = 16;        // wide-engine steps per launch between two admissions
constexpr uint32_t kServeTeamSteps = 256;   // team-engine steps per launch between two admissions
constexpr uint32_t kTeamSlotsFeedback = 6;  // slot copies of a team engine: 5 nodes each, <= 32 nodes
constexpr uint32_t kTeamSlotsPhoton = 8;    // 4 nodes each

uint32_t payload_of_topology(uint32_t topo)
{
    return topo == GF_JOB_FEEDBACK ? GF_PAYLOAD_PARTICLE : topo == GF_JOB_HASH ? GF_PAYLOAD_KEYED : GF_PAYLOAD_PHOTON;
}

// One job slot of an engine: its result buffers and the job it runs.
struct ServeSlot {
    bool busy = false;
    uint64_t job = 0, items = 0, seed = 0, bound = 0, step0 = 0;
    double submit_ms = 0, start_ms = 0;
    uint32_t* rec = nullptr;           // feedback: weight bits by particle id
    unsigned long long* ms = nullptr;  // hash: (hash sum, count) per bucket
    double* tally = nullptr;           // photon: absorbed weight per depth bin
};

// A served engine: one served flow of one topology.
struct ServeEngine {
    uint32_t topo = 0, xcd = 0;
    gf_context* ctx = nullptr;
    gf_flow* flow = nullptr;
    std::vector<ServeSlot> slots;
    std::vector<gfd::JobSlot> table;  // host image of the engine's job table
    uint64_t pool = 0, reserved = 0, reserve_graph = 0, live_max = 0, bytes = 0;
    unsigned long long* dig = nullptr;
    uint64_t step = 0;   // the engine's step counter at its last poll
    uint32_t busy = 0;
    bool team = false;   // a team engine (runs on the wide engine while `wide` is set)
    bool wide = false;
    // a launch in flight: its job table, control words and step counter are copied into pinned host
    // memory behind it, and `done` records its completion (the driver polls it)
    bool inflight = false;
    hipEvent_t done = nullptr;
    gfd::JobSlot* jt_h = nullptr;
    uint32_t* ctl_h = nullptr;
    unsigned long long* step_h = nullptr;
};

}  // namespace

struct ServerJob {
    uint64_t id;
    gf_job_desc d;
    double t_submit;
};

struct gf_server {
    gf_context* ctx = nullptr;  // the creating context: its device
    gf_server_config cfg{};
    std::mutex mu;
    std::condition_variable cv;
    std::deque<ServerJob> queue;
    std::vector<gf_job_result> done;
    std::unordered_map<uint64_t, std::vector<double>> tallies;
    uint64_t next_id = 1, submitted = 0, finished = 0;
    bool stop = false;
    std::thread driver;
    std::vector<ServeEngine> eng;
    std::chrono::steady_clock::time_point t0;
    std::string err;

    double now_ms() const { return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count(); }
};

namespace {

// Add one slot's copy of the topology's graph to the served flow; returns the copy's source node.
uint32_t add_slot_graph(gf_flow* f, uint32_t topo, uint32_t slot, ServeSlot& ss, uint64_t max_items)
{
    auto ck = [](gf_status s, const char* what) {
        if (s != GF_OK) fail(s, std::string("serving: ") + what);
    };
    auto node = [](uint32_t role, uint32_t op, uint32_t bins) {
        gf_node_desc d{};
        d.role = role;
        d.op = op;
        d.bins = bins;
        return d;
    };
    uint32_t local = 0;
    auto add = [&](const char* name, const gf_node_desc& d) {
        uint32_t k;
        ck(gf_flow_add_node(f, (std::string("slot ") + std::to_string(slot) + " " + name).c_str(), &d, &k), "node");
        f->nodes[k].slot = slot;
        f->nodes[k].local = local++;
        return k;
    };
    if (topo == GF_JOB_FEEDBACK) {
        // particle_feedback: Generate -> Filter -> {A x0.5, B x0.75} -> Filter, Filter.sink -> Sink
        HIPX(hipMalloc(&ss.rec, sizeof(uint32_t) * max_items));
        gf_node_desc src = node(GF_ROLE_SOURCE, GF_SRC_PARTICLE_UNIFORM, 1);
        gf_node_desc filt = node(GF_ROLE_TRANSFORM, GF_CLS_WEIGHT_ROUTE, 3);
        filt.u[0] = GF_ROUTE_PARITY;
        filt.f[0] = 1e-3f;
        gf_node_desc a = node(GF_ROLE_TRANSFORM, GF_XF_SCALE_WEIGHT, 1), b = a;
        a.f[0] = 0.5f;
        b.f[0] = 0.75f;
        gf_node_desc snk = node(GF_ROLE_SINK, GF_SNK_RECORD_WEIGHT, 0);
        snk.ptr[0] = (uint64_t)(uintptr_t)ss.rec;
        snk.size[0] = max_items;
        const uint32_t ks = add("generate", src), kf = add("filter", filt), ka = add("process A", a);
        const uint32_t kb = add("process B", b), kk = add("record weight", snk);
        ck(gf_flow_connect(f, ks, 0, kf), "connect");
        ck(gf_flow_connect(f, kf, 0, kk), "connect");
        ck(gf_flow_connect(f, kf, 1, ka), "connect");
        ck(gf_flow_connect(f, kf, 2, kb), "connect");
        ck(gf_flow_connect(f, ka, 0, kf), "connect");
        ck(gf_flow_connect(f, kb, 0, kf), "connect");
        return ks;
    }
    if (topo == GF_JOB_HASH) {
        // keyed hash pipeline: Zipf keys -> murmur3 bucket partition (4096) -> multiset digest sink
        HIPX(hipMalloc(&ss.ms, sizeof(unsigned long long) * 2u * kServeBuckets));
        gf_node_desc src = node(GF_ROLE_SOURCE, GF_SRC_KEYED_ZIPF, 1);
        src.u[0] = 1u << 20;
        src.f[0] = 1.1f;
        gf_node_desc part = node(GF_ROLE_TRANSFORM, GF_CLS_KEY_BUCKET, kServeBuckets);
        gf_node_desc snk = node(GF_ROLE_SINK, GF_SNK_CHUNK_DIGEST, 0);
        snk.ptr[1] = (uint64_t)(uintptr_t)ss.ms;
        snk.size[1] = kServeBuckets;
        const uint32_t ks = add("keys", src), kp = add("bucket", part), kk = add("digest", snk);
        ck(gf_flow_connect(f, ks, 0, kp), "connect");
        for (uint32_t bb = 0; bb < kServeBuckets; ++bb) ck(gf_flow_connect(f, kp, bb, kk), "connect");
        return ks;
    }
    // photon transport: Generate -> Step -> Alive; Alive.alive -> Step, Alive.done -> Done
    HIPX(hipMalloc(&ss.tally, sizeof(double) * kServeTallyBins));
    gf_node_desc src = node(GF_ROLE_SOURCE, GF_SRC_PHOTON, 1);
    gf_node_desc step = node(GF_ROLE_TRANSFORM, GF_XF_PHOTON_STEP, 1);
    step.ptr[0] = (uint64_t)(uintptr_t)ss.tally;
    step.size[0] = kServeTallyBins;
    gf_node_desc alive = node(GF_ROLE_TRANSFORM, GF_CLS_PHOTON_ALIVE, 2);
    gf_node_desc done = node(GF_ROLE_SINK, GF_SNK_PHOTON_TALLY, 0);
    const uint32_t ks = add("generate", src), kst = add("step", step), ka = add("alive", alive), kd = add("done", done);
    ck(gf_flow_connect(f, ks, 0, kst), "connect");
    ck(gf_flow_connect(f, kst, 0, ka), "connect");
    ck(gf_flow_connect(f, ka, 0, kst), "connect");
    ck(gf_flow_connect(f, ka, 1, kd), "connect");
    return ks;
}

void destroy_engine(ServeEngine& te)
{
    if (te.ctx && te.inflight) (void)hipStreamSynchronize(te.ctx->stream);
    if (te.done) (void)hipEventDestroy(te.done);
    if (te.jt_h) (void)hipHostFree(te.jt_h);
    if (te.ctl_h) (void)hipHostFree(te.ctl_h);
    if (te.step_h) (void)hipHostFree(te.step_h);
    if (te.flow) gf_flow_destroy(te.flow);
    for (auto& ss : te.slots) {
        if (ss.rec) (void)hipFree(ss.rec);
        if (ss.ms) (void)hipFree(ss.ms);
        if (ss.tally) (void)hipFree(ss.tally);
    }
    if (te.dig) (void)hipFree(te.dig);
    if (te.ctx) gf_context_destroy(te.ctx);
    te = ServeEngine{};
}

// Reset an engine's device state to no work in flight: the team engine's pre-select state (its
// count rows clear), or the wide engine's select of empty slots (idle).
void reset_engine(ServeEngine& te)
{
    EngineBuffers& eb = *te.flow->eng;
    hipStream_t st = te.ctx->stream;
    if (te.team) {
        TeamArgs& ta = eb.ta;
        ta.e = eb.e;
        HIPX(hipMemsetAsync(ta.cnt, 0, eb.cnt_bytes, st));
        HIPX(hipMemsetAsync(ta.zero_g, 0, sizeof(uint32_t), st));
        HIPX(hipMemcpyAsync(eb.ta_dev, &ta, sizeof(TeamArgs), hipMemcpyHostToDevice, st));
    }
    hipLaunchKernelGGL(k_engine_setup, dim3(1), dim3(kThreads), sizeof(SelLds), st, eb.e, te.team ? 0u : 1u);
    HIPX(hipGetLastError());
    HIPX(hipStreamSynchronize(st));
    te.wide = false;
}

// Build a served engine of `topo`: M slot copies of its graph, a pool of `pool_bytes`, the job table
// empty (every slot idle), on the team engine (XCD `xcd`) when `team` and the union fits it, the
// wide engine's K-step graph captured either way (a team engine's hand-off runs on it).
void build_serve_engine(gf_server& sv, ServeEngine& te, uint32_t topo, uint32_t M, uint64_t pool_bytes, bool team, uint32_t xcd)
{
    te.topo = topo;
    te.xcd = xcd;
    if (gf_context_create(sv.ctx->device, &te.ctx) != GF_OK) fail(GF_DEVICE_ERROR, "serving: context creation failed");
    const uint32_t payload = payload_of_topology(topo);
    if (gf_flow_create(te.ctx, payload, &te.flow) != GF_OK) fail(GF_DEVICE_ERROR, "serving: flow creation failed");
    gf_flow& f = *te.flow;
    f.slots = M;
    te.slots.resize(M);
    uint32_t src0 = 0;
    for (uint32_t s = 0; s < M; ++s) {
        const uint32_t k = add_slot_graph(te.flow, topo, s, te.slots[s], sv.cfg.max_items);
        if (s == 0) src0 = k;
    }
    HIPX(hipMalloc(&te.dig, sizeof(unsigned long long)));
    // the graph's transform output bounds (Flow::transformReserve)
    uint64_t reserve = 0;
    for (uint32_t k = 0; k < (uint32_t)f.nodes.size(); ++k)
        if (f.nodes[k].slot == 0 && f.nodes[k].nd.role == GF_ROLE_TRANSFORM)
            reserve += output_chunk_bound(kServeCap, kServeCap, f.nodes[k].nd.bins);
    te.reserve_graph = reserve;
    te.live_max = topo == GF_JOB_PHOTON ? kServePhotonLive : topo == GF_JOB_HASH ? kServeHashLive : kServeLive;
    const uint64_t chunk_bytes = (uint64_t)kServeCap * payload_bytes(payload);
    uint64_t pool = pool_bytes / chunk_bytes;
    pool = std::min<uint64_t>(pool, (uint64_t)M * (te.live_max + reserve));  // more than M full jobs is never used
    pool = std::min<uint64_t>(pool, 0xFFFFFFFFull / kServeCap);              // slot addresses stay 32 bit
    if (pool < reserve + 2u) fail(GF_BAD_ARGUMENT, "serving: the memory budget cannot hold one job of this topology");
    te.pool = pool;
    gf_tuning t = default_photon_tuning();
    t.batch_chunks = kServeBatch;
    t.chunk_capacity = kServeCap;
    t.live_particles = (uint32_t)std::min<uint64_t>((uint64_t)M * te.live_max, 0xFFFFFFFFull);
    t.num_chunks = (uint32_t)pool;
    t.schedule = GF_SCHEDULE_DEVICE;
    t.fuse_transforms = 1;
    t.use_graph = 1;
    t.device_engine = team ? 0u : 1u;  // 0: the team engine when the union fits it
    t.inflight_batches = 1;
    t.source_chunks = team ? 1u : kServeSourceChunks;  // the team engine admits one chunk per lane and step
    t.steps_per_launch = team ? kServeTeamSteps : kServeSteps;
    t.max_workgroups = team ? sv.cfg.team_size : 512u;
    t.team_xcd = team ? xcd : 0u;
    if (gf_flow_configure(te.flow, &t) != GF_OK) fail(GF_BAD_ARGUMENT, "serving: " + f.err);
    check_graph(f, src0);
    const RunShape sh = resolve(f, 0xFFFFFFFFull);
    ensure_pool(f, sh);
    build_engine(f, 0, sh, src0);
    EngineBuffers& eb = *f.eng;
    te.team = eb.team;
    hipStream_t s = te.ctx->stream;
    hipGraph_t g;
    HIPX(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    for (uint32_t i = 0; i < kServeSteps; ++i) launch_step(eb, f.payload, s);
    HIPX(hipStreamEndCapture(s, &g));
    HIPX(hipGraphInstantiate(&eb.exec, g, nullptr, nullptr, 0));
    (void)hipGraphDestroy(g);
    eb.graph_steps = kServeSteps;
    reset_engine(te);
    te.table.assign(M, gfd::JobSlot{0, 0, 0, 0, 0});
    HIPX(hipEventCreateWithFlags(&te.done, hipEventDisableTiming));
    HIPX(hipHostMalloc((void**)&te.jt_h, sizeof(gfd::JobSlot) * M, hipHostMallocDefault));
    HIPX(hipHostMalloc((void**)&te.ctl_h, sizeof(uint32_t) * CTL_WORDS, hipHostMallocDefault));
    HIPX(hipHostMalloc((void**)&te.step_h, sizeof(unsigned long long), hipHostMallocDefault));
    te.bytes = pool * chunk_bytes + (uint64_t)M * (topo == GF_JOB_FEEDBACK ? 4u * sv.cfg.max_items : 65536u) + (64ull << 20);
}

// One launch on an engine's stream (a team launch, or the wide engine's K-step graph), followed by
// the copies of its job table, control words and step counter into pinned host memory and the
// completion event the driver polls.
void launch_engine(ServeEngine& te)
{
    EngineBuffers& eb = *te.flow->eng;
    hipStream_t st = te.ctx->stream;
    HIPX(hipMemsetAsync(eb.e.ctl + CTL_DONE, 0, sizeof(uint32_t), st));  // wake an idle engine
    if (te.team && !te.wide) {
        HIPX(hipMemsetAsync(eb.ta.tw, 0, TW_WORDS * sizeof(uint32_t), st));
        with_payload(te.flow->payload, 1u, [&](auto pt, auto) {
            using P = typename decltype(pt)::type;
            hipLaunchKernelGGL(k_team<P>, dim3(8u * eb.ta.K), dim3(kTeamThreads), sizeof(TeamLds), st, (const TeamArgs*)eb.ta_dev);
        });
        HIPX(hipGetLastError());
    } else {
        HIPX(hipGraphLaunch(eb.exec, st));
    }
    HIPX(hipMemcpyAsync(te.jt_h, eb.e.jobs, sizeof(gfd::JobSlot) * te.table.size(), hipMemcpyDeviceToHost, st));
    HIPX(hipMemcpyAsync(te.ctl_h, eb.e.ctl, sizeof(uint32_t) * CTL_WORDS, hipMemcpyDeviceToHost, st));
    HIPX(hipMemcpyAsync(te.step_h, eb.e.ctr + CNT_STEP, sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
    HIPX(hipEventRecord(te.done, st));
    te.inflight = true;
}

// The job's pool bound and live bound on an engine (P_j = min(N, P), lowered to fit the pool).
void job_bounds(const ServeEngine& te, uint64_t items, uint64_t& live, uint64_t& bound)
{
    live = std::min<uint64_t>(items, te.live_max);
    if (live + te.reserve_graph > te.pool) live = te.pool - te.reserve_graph;
    if (live == 0) live = 1;
    bound = live + te.reserve_graph;
}

// Start a job on a free slot: clear its result buffers, write its job table entry.
void start_job(gf_server& sv, ServeEngine& te, uint32_t s, const ServerJob& j)
{
    ServeSlot& ss = te.slots[s];
    hipStream_t st = te.ctx->stream;
    if (te.topo == GF_JOB_FEEDBACK) {
        hipLaunchKernelGGL(k_fill_u32, dim3(256), dim3(256), 0, st, ss.rec, j.d.items, 0xFFFFFFFFu);
    } else if (te.topo == GF_JOB_HASH) {
        hipLaunchKernelGGL(k_zero_u64, dim3(32), dim3(256), 0, st, ss.ms, (uint64_t)2u * kServeBuckets);
    } else {
        hipLaunchKernelGGL(k_zero_u64, dim3(4), dim3(256), 0, st, (unsigned long long*)ss.tally, (uint64_t)kServeTallyBins);
    }
    HIPX(hipGetLastError());
    uint64_t live, bound;
    job_bounds(te, j.d.items, live, bound);
    te.table[s] = gfd::JobSlot{j.d.items, live, j.d.seed, 0ull, 0ull};
    HIPX(hipMemcpyAsync(te.flow->eng->e.jobs + s, &te.table[s], sizeof(gfd::JobSlot), hipMemcpyHostToDevice, st));
    ss.busy = true;
    ss.job = j.id;
    ss.items = j.d.items;
    ss.seed = j.d.seed;
    ss.bound = bound;
    ss.submit_ms = j.t_submit;
    ss.start_ms = sv.now_ms();
    ss.step0 = te.step;
    te.reserved += bound;
    te.busy++;
}

// The finished job's result from its slot's buffers (the engine's stream is idle).
void finish_job(gf_server& sv, ServeEngine& te, uint32_t s, uint64_t step, gf_job_result& r, std::vector<double>& tally)
{
    ServeSlot& ss = te.slots[s];
    hipStream_t st = te.ctx->stream;
    r.job = ss.job;
    r.topology = te.topo;
    r.slot = s;
    r.status = GF_OK;
    r.engine = te.team && !te.wide ? 3u : 1u;  // the engine that ran the job's last steps
    r.items = ss.items;
    r.steps = step - ss.step0;
    r.submit_ms = ss.submit_ms;
    r.start_ms = ss.start_ms;
    if (te.topo == GF_JOB_FEEDBACK) {
        HIPX(hipMemsetAsync(te.dig, 0, sizeof(unsigned long long), st));
        hipLaunchKernelGGL(k_record_digest, dim3(256), dim3(256), 0, st, ss.rec, ss.items, te.dig);
        HIPX(hipGetLastError());
        unsigned long long d = 0;
        HIPX(hipMemcpyAsync(&d, te.dig, sizeof d, hipMemcpyDeviceToHost, st));
        HIPX(hipStreamSynchronize(st));
        r.digest = d;
    } else if (te.topo == GF_JOB_HASH) {
        std::vector<unsigned long long> ms(2u * kServeBuckets);
        HIPX(hipMemcpyAsync(ms.data(), ss.ms, sizeof(unsigned long long) * ms.size(), hipMemcpyDeviceToHost, st));
        HIPX(hipStreamSynchronize(st));
        uint64_t d = 0;
        for (uint32_t b = 0; b < kServeBuckets; ++b)
            d += gfd::mix64(((uint64_t)b << 48) ^ ms[2u * b]) + gfd::mix64(ms[2u * b + 1u] * 0x9E3779B97F4A7C15ull + b);
        r.digest = d;
    } else {
        tally.assign(kServeTallyBins, 0.0);
        HIPX(hipMemcpyAsync(tally.data(), ss.tally, sizeof(double) * kServeTallyBins, hipMemcpyDeviceToHost, st));
        HIPX(hipStreamSynchronize(st));
        double sum = 0;
        for (double v : tally) sum += v;
        r.tally_sum = sum;
    }
    r.end_ms = sv.now_ms();
    // the slot is empty again (N = 0) before the engine's next launch
    te.table[s] = gfd::JobSlot{0, 0, 0, 0, 0};
    HIPX(hipMemcpyAsync(te.flow->eng->e.jobs + s, &te.table[s], sizeof(gfd::JobSlot), hipMemcpyHostToDevice, st));
    HIPX(hipStreamSynchronize(st));
    te.reserved -= ss.bound;
    te.busy--;
    ss.busy = false;
}

// Fail every job of an engine whose step reported an error, and reset the engine.
void fail_engine(gf_server& sv, ServeEngine& te, uint32_t status, std::vector<gf_job_result>& out)
{
    for (uint32_t s = 0; s < (uint32_t)te.slots.size(); ++s) {
        ServeSlot& ss = te.slots[s];
        if (!ss.busy) continue;
        gf_job_result r{};
        r.job = ss.job;
        r.topology = te.topo;
        r.slot = s;
        r.status = status;
        r.engine = te.team ? 3u : 1u;
        r.items = ss.items;
        r.submit_ms = ss.submit_ms;
        r.start_ms = ss.start_ms;
        r.end_ms = sv.now_ms();
        out.push_back(r);
        ss.busy = false;
        te.table[s] = gfd::JobSlot{0, 0, 0, 0, 0};
    }
    te.busy = 0;
    te.reserved = 0;
    EngineBuffers& eb = *te.flow->eng;
    HIPX(hipMemcpyAsync(eb.e.jobs, te.table.data(), sizeof(gfd::JobSlot) * te.table.size(), hipMemcpyHostToDevice,
                        te.ctx->stream));
    HIPX(hipMemsetAsync(eb.e.ctl, 0, CTL_WORDS * sizeof(uint32_t), te.ctx->stream));
    reset_engine(te);
}

// Fail the admitted jobs of every engine after an error outside the engines' own reports.
void fail_all_jobs(gf_server& sv, gf_status rc)
{
    for (ServeEngine& te : sv.eng) {
        if (te.inflight) {
            (void)hipStreamSynchronize(te.ctx->stream);
            te.inflight = false;
        }
        for (uint32_t s = 0; s < (uint32_t)te.slots.size(); ++s) {
            ServeSlot& ss = te.slots[s];
            if (!ss.busy) continue;
            gf_job_result r{};
            r.job = ss.job;
            r.topology = te.topo;
            r.slot = s;
            r.status = rc;
            r.items = ss.items;
            r.submit_ms = ss.submit_ms;
            r.start_ms = ss.start_ms;
            r.end_ms = sv.now_ms();
            sv.done.push_back(r);
            sv.finished++;
            ss.busy = false;
            te.reserved -= ss.bound;
            te.busy--;
        }
    }
}

void server_driver(gf_server* svp)
{
    gf_server& sv = *svp;
    for (;;) {
        bool progressed = false;
        // 1. finished launches: finish the jobs whose items are all sunk
        {
            std::vector<gf_job_result> results;
            std::vector<std::pair<uint64_t, std::vector<double>>> tallies;
            std::string msg;
            const gf_status rc = guard(&msg, [&] {
                for (ServeEngine& te : sv.eng) {
                    if (!te.inflight) continue;
                    const hipError_t q = hipEventQuery(te.done);
                    if (q == hipErrorNotReady) continue;
                    HIPX(q);
                    te.inflight = false;
                    progressed = true;
                    EngineBuffers& eb = *te.flow->eng;
                    hipStream_t st = te.ctx->stream;
                    const uint32_t* ctl = te.ctl_h;
                    const unsigned long long step = *te.step_h;
                    te.step = step;
                    if (ctl[CTL_STATUS]) {
                        const uint32_t code = ctl[CTL_STATUS];
                        fail_engine(sv, te, code, results);
                        fail((gf_status)code, "serving: an engine of topology " + std::to_string(te.topo) +
                                                  " reported an error");
                    }
                    for (uint32_t s = 0; s < (uint32_t)te.slots.size(); ++s) {
                        ServeSlot& ss = te.slots[s];
                        if (!ss.busy || te.jt_h[s].sunk < ss.items) continue;
                        gf_job_result r{};
                        std::vector<double> tally;
                        finish_job(sv, te, s, step, r, tally);
                        if (te.topo == GF_JOB_PHOTON) tallies.emplace_back(r.job, std::move(tally));
                        results.push_back(r);
                    }
                    if (te.team && !te.wide && ctl[CTL_HANDOFF]) {
                        // a team step that does not fit the team's limits: the engine continues on the
                        // wide engine from the team's pre-select state until it is idle
                        hipLaunchKernelGGL(k_engine_resume, dim3(1), dim3(kThreads), sizeof(SelLds), st, eb.e);
                        HIPX(hipGetLastError());
                        te.wide = true;
                    } else if (te.team && te.wide && ctl[CTL_DONE] && !te.busy) {
                        reset_engine(te);  // idle: back to the team engine
                    }
                }
            });
            std::lock_guard<std::mutex> lk(sv.mu);
            if (rc != GF_OK) sv.err = msg;
            for (auto& tl : tallies) sv.tallies[tl.first] = std::move(tl.second);
            for (const auto& r : results) sv.done.push_back(r);
            sv.finished += results.size();
            if (rc != GF_OK && results.empty()) fail_all_jobs(sv, rc);
            if (!results.empty() || rc != GF_OK) sv.cv.notify_all();
        }
        // 2. admission: first come, first served within each topology, into the first engine of the
        //    job's topology with no launch in flight, a free slot and room for the job's pool bound
        std::vector<std::pair<uint32_t, ServerJob>> admit;  // (engine, job)
        {
            std::unique_lock<std::mutex> lk(sv.mu);
            auto quiet = [&] {  // nothing running and nothing admitted
                for (const ServeEngine& te : sv.eng)
                    if (te.busy || te.inflight) return false;
                return true;
            };
            // with launches in flight the driver polls their events; otherwise it sleeps until a
            // submission or the server's end
            if (quiet()) sv.cv.wait(lk, [&] { return sv.stop || !sv.queue.empty(); });
            if (sv.stop && sv.queue.empty() && quiet()) break;
            bool blocked[3] = {false, false, false};
            std::vector<uint32_t> free_slots(sv.eng.size());
            std::vector<uint64_t> room(sv.eng.size());
            for (size_t i = 0; i < sv.eng.size(); ++i) {
                free_slots[i] = sv.eng[i].inflight ? 0u : (uint32_t)sv.eng[i].slots.size() - sv.eng[i].busy;
                room[i] = sv.eng[i].pool - sv.eng[i].reserved;
            }
            for (auto it = sv.queue.begin(); it != sv.queue.end();) {
                const uint32_t t = it->d.topology;
                int32_t pick = -1;
                if (!blocked[t])
                    for (size_t i = 0; i < sv.eng.size() && pick < 0; ++i) {
                        if (sv.eng[i].topo != t || !free_slots[i]) continue;
                        uint64_t live, bound;
                        job_bounds(sv.eng[i], it->d.items, live, bound);
                        if (bound <= room[i]) {
                            pick = (int32_t)i;
                            free_slots[i]--;
                            room[i] -= bound;
                        }
                    }
                if (pick >= 0) {
                    admit.emplace_back((uint32_t)pick, *it);
                    it = sv.queue.erase(it);
                } else {
                    blocked[t] = true;  // first come, first served within the topology
                    ++it;
                }
            }
        }
        // 3. start the admitted jobs; the next launch of every busy engine with none in flight
        //    (engines advance independently, side by side on their own streams)
        {
            std::string msg;
            const gf_status rc = guard(&msg, [&] {
                for (const auto& aj : admit) {
                    ServeEngine& te = sv.eng[aj.first];
                    uint32_t s = 0;
                    while (te.slots[s].busy) ++s;
                    start_job(sv, te, s, aj.second);
                    progressed = true;
                }
                for (ServeEngine& te : sv.eng)
                    if (te.busy && !te.inflight) {
                        launch_engine(te);
                        progressed = true;
                    }
            });
            if (rc != GF_OK) {
                std::lock_guard<std::mutex> lk(sv.mu);
                sv.err = msg;
                fail_all_jobs(sv, rc);
                sv.cv.notify_all();
            }
        }
        // launches in flight and nothing else to do: poll again shortly
        if (!progressed) std::this_thread::sleep_for(std::chrono::microseconds(50));
    }
}

}  // namespace

extern "C" {

gf_status gf_server_create(gf_context* ctx, const gf_server_config* cfg, gf_server** out)
{
    if (!ctx || !out) return GF_BAD_ARGUMENT;
    return guard(&ctx->err, [&] {
        auto sv = std::make_unique<gf_server>();
        sv->ctx = ctx;
        if (cfg) sv->cfg = *cfg;
        gf_server_config& c = sv->cfg;
        if (!c.slots) c.slots = kServeSlots;
        // the hash pipeline's wide engine: 3 lanes per slot copy
        if (3u * c.slots > gfd::kMaxNodes || c.slots > gfd::kMaxJobs) fail(GF_BAD_ARGUMENT, "serving: at most 42 job slots per topology");
        if (!c.team_size)
            c.team_size = std::max<uint32_t>(1u, std::min<uint32_t>(32u, ctx->cus / 8u) - std::min<uint32_t>(4u, ctx->cus / 16u));
        if (!c.memory_budget) c.memory_budget = 64ull << 30;
        if (!c.max_items) c.max_items = 10000000ull;
        if (c.max_items > (1ull << 32)) fail(GF_BAD_ARGUMENT, "serving: max_items exceeds the 32 bit stream ids");
        sv->t0 = std::chrono::steady_clock::now();
        // the result buffers first, then a third of the rest of the budget per topology, split over
        // the topology's engines: team engines of at most 6 (feedback) or 8 (photon) slots on
        // distinct XCDs, one wide engine for the hash pipeline
        const uint64_t buffers = (uint64_t)c.slots * (4ull * c.max_items + 65536ull + 8192ull) + (256ull << 20);
        if (buffers >= c.memory_budget) fail(GF_BAD_ARGUMENT, "serving: the memory budget cannot hold the result buffers");
        const uint64_t share = (c.memory_budget - buffers) / 3u;
        struct Plan { uint32_t topo, slots; bool team; };
        std::vector<Plan> plan;
        for (uint32_t t : {(uint32_t)GF_JOB_FEEDBACK, (uint32_t)GF_JOB_PHOTON}) {
            const uint32_t cap = t == GF_JOB_FEEDBACK ? kTeamSlotsFeedback : kTeamSlotsPhoton;
            const uint32_t ne = (c.slots + cap - 1u) / cap;
            for (uint32_t i = 0; i < ne; ++i) plan.push_back({t, c.slots / ne + (i < c.slots % ne ? 1u : 0u), true});
        }
        plan.push_back({(uint32_t)GF_JOB_HASH, c.slots, false});
        uint32_t engines_of[3] = {0, 0, 0};
        for (const Plan& pl : plan) engines_of[pl.topo]++;
        sv->eng.resize(plan.size());
        uint32_t xcd = 0;
        for (size_t i = 0; i < plan.size(); ++i) {
            const Plan& pl = plan[i];
            try {
                build_serve_engine(*sv, sv->eng[i], pl.topo, pl.slots, share / engines_of[pl.topo], pl.team,
                                   pl.team ? 1u + (xcd++ % 8u) : 0u);
            } catch (...) {
                for (ServeEngine& te : sv->eng) destroy_engine(te);
                throw;
            }
        }
        for (uint32_t t = 0; t < 3u; ++t) c.topology_bytes[t] = 0;
        for (const ServeEngine& te : sv->eng) c.topology_bytes[te.topo] += te.bytes;
        gf_server* p = sv.get();
        p->driver = std::thread(server_driver, p);
        *out = sv.release();
    });
}

gf_status gf_server_submit(gf_server* sv, const gf_job_desc* job, uint64_t* id)
{
    if (!sv || !job || !id) return GF_BAD_ARGUMENT;
    if (job->topology > GF_JOB_PHOTON || job->items == 0) return GF_BAD_ARGUMENT;
    if (job->items > sv->cfg.max_items) return GF_BAD_ARGUMENT;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        if (sv->stop) return GF_BAD_ARGUMENT;
        ServerJob j{sv->next_id++, *job, sv->now_ms()};
        *id = j.id;
        sv->queue.push_back(j);
        sv->submitted++;
    }
    sv->cv.notify_all();
    return GF_OK;
}

gf_status gf_server_poll(gf_server* sv, gf_job_result* out, uint32_t cap, uint32_t* n, uint32_t wait_all)
{
    if (!sv || !n || (cap && !out)) return GF_BAD_ARGUMENT;
    std::unique_lock<std::mutex> lk(sv->mu);
    if (wait_all) sv->cv.wait(lk, [&] { return sv->finished == sv->submitted; });
    const uint32_t k = (uint32_t)std::min<size_t>(cap, sv->done.size());
    for (uint32_t i = 0; i < k; ++i) out[i] = sv->done[i];
    sv->done.erase(sv->done.begin(), sv->done.begin() + k);
    *n = k;
    return GF_OK;
}

gf_status gf_server_job_tally(gf_server* sv, uint64_t job, double* tally, uint32_t bins)
{
    if (!sv || !tally) return GF_BAD_ARGUMENT;
    std::lock_guard<std::mutex> lk(sv->mu);
    auto it = sv->tallies.find(job);
    if (it == sv->tallies.end() || bins > it->second.size()) return GF_BAD_ARGUMENT;
    std::memcpy(tally, it->second.data(), sizeof(double) * bins);
    return GF_OK;
}

gf_status gf_server_last_error(gf_server* sv, char* buffer, uint32_t size)
{
    if (!sv || !buffer || !size) return GF_BAD_ARGUMENT;
    std::lock_guard<std::mutex> lk(sv->mu);
    std::snprintf(buffer, size, "%s", sv->err.c_str());
    return GF_OK;
}

gf_status gf_server_destroy(gf_server* sv)
{
    if (!sv) return GF_BAD_ARGUMENT;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->stop = true;
    }
    sv->cv.notify_all();
    if (sv->driver.joinable()) sv->driver.join();
    for (ServeEngine& te : sv->eng) destroy_engine(te);
    delete sv;
    return GF_OK;
}

}  // extern "C"
