// Serving loop: independent flow jobs (particle feedback, keyed hash pipeline, photon transport)
// arrive over time, are admitted into slots as slots and device memory allow, and run side by side
// on the device. Each slot owns a context (its own stream) and keeps one flow per topology, so a
// slot's pools and ready lists stay resident in HBM from job to job and no device memory is freed
// while other jobs run. The team engine of slot i runs on XCD i mod 8 with a team of the XCD's CUs
// minus a few (left for the transient workgroups of the other slots' launches), so up to eight jobs
// hold an XCD each; the hash pipeline's 4096-bucket partition runs on the wide engine.
//
// There is no serving loop in the reference; the flows it runs are the reference's own networks
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

constexpr uint32_t kServeLive = 16384;      // live particles of a served job (the headline P)
constexpr uint32_t kServePhotonLive = 8192; // photons in flight per served photon job
constexpr uint32_t kServeCap = 8192;        // chunk capacity
constexpr uint32_t kServeBatch = 8;         // Q
constexpr uint32_t kServeBuckets = 4096;
constexpr uint32_t kServeTallyBins = 1024;

// One topology's flow in one slot, with its result buffers.
struct SlotFlow {
    gf_flow* flow = nullptr;
    uint32_t source = 0;
    uint64_t bytes = 0;          // device memory charged to the budget
    uint32_t* rec = nullptr;     // feedback: weight bits by particle id
    unsigned long long* ms = nullptr;  // hash: (hash sum, count) per bucket; digest word after them
    double* tally = nullptr;     // photon: absorbed weight per depth bin
    unsigned long long* dig = nullptr;
    uint64_t rec_cap = 0;
};

}  // namespace

struct ServerJob {
    uint64_t id;
    gf_job_desc d;
    double t_submit;
};

struct gf_server {
    gf_context* ctx = nullptr;  // the creating context: device and its CU count
    gf_server_config cfg{};
    std::mutex mu;
    std::condition_variable cv;
    std::deque<ServerJob> queue;
    std::vector<gf_job_result> done;
    std::unordered_map<uint64_t, std::vector<double>> tallies;
    uint64_t next_id = 1, submitted = 0, finished = 0;
    uint64_t mem_in_use = 0;
    uint32_t warm = 0;  // slots done with their prewarm
    bool stop = false;
    std::vector<std::thread> workers;
    std::chrono::steady_clock::time_point t0;
    std::string err;

    double now_ms() const { return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count(); }
};

namespace {

uint64_t pool_bytes_of(uint32_t payload, uint32_t S, uint32_t C) { return (uint64_t)S * C * payload_bytes(payload); }

void destroy_slot_flow(SlotFlow& sf);
void build_slot_flow(SlotFlow& sf, gf_server& sv, gf_context* ctx, uint32_t topo, uint32_t slot);

// Build the topology's flow for a slot (caller holds no lock; the budget was reserved by the caller).
// A failed build frees what it had created.
SlotFlow make_slot_flow(gf_server& sv, gf_context* ctx, uint32_t topo, uint32_t slot)
{
    SlotFlow sf;
    try {
        build_slot_flow(sf, sv, ctx, topo, slot);
    } catch (...) {
        destroy_slot_flow(sf);
        throw;
    }
    return sf;
}

void build_slot_flow(SlotFlow& sf, gf_server& sv, gf_context* ctx, uint32_t topo, uint32_t slot)
{
    const uint32_t payload = topo == GF_JOB_FEEDBACK ? GF_PAYLOAD_PARTICLE : topo == GF_JOB_HASH ? GF_PAYLOAD_KEYED : GF_PAYLOAD_PHOTON;
    auto ck = [](gf_status s, const char* what) {
        if (s != GF_OK) fail(s, std::string("serving: ") + what);
    };
    ck(gf_flow_create(ctx, payload, &sf.flow), "flow create");
    gf_flow* f = sf.flow;
    auto node = [](uint32_t role, uint32_t op, uint32_t bins) {
        gf_node_desc d{};
        d.role = role;
        d.op = op;
        d.bins = bins;
        return d;
    };
    uint32_t k = 0;
    if (topo == GF_JOB_FEEDBACK) {
        // particle_feedback: Generate -> Filter -> {A x0.5, B x0.75} -> Filter, Filter.sink -> Sink
        sf.rec_cap = sv.cfg.max_items;
        HIPX(hipMalloc(&sf.rec, sizeof(uint32_t) * sf.rec_cap));
        gf_node_desc src = node(GF_ROLE_SOURCE, GF_SRC_PARTICLE_UNIFORM, 1);
        gf_node_desc filt = node(GF_ROLE_TRANSFORM, GF_CLS_WEIGHT_ROUTE, 3);
        filt.u[0] = GF_ROUTE_PARITY;
        filt.f[0] = 1e-3f;
        gf_node_desc a = node(GF_ROLE_TRANSFORM, GF_XF_SCALE_WEIGHT, 1), b = a;
        a.f[0] = 0.5f;
        b.f[0] = 0.75f;
        gf_node_desc snk = node(GF_ROLE_SINK, GF_SNK_RECORD_WEIGHT, 0);
        snk.ptr[0] = (uint64_t)(uintptr_t)sf.rec;
        snk.size[0] = sf.rec_cap;
        uint32_t ks, kf, ka, kb, kk;
        ck(gf_flow_add_node(f, "generate", &src, &ks), "node");
        ck(gf_flow_add_node(f, "filter", &filt, &kf), "node");
        ck(gf_flow_add_node(f, "process A", &a, &ka), "node");
        ck(gf_flow_add_node(f, "process B", &b, &kb), "node");
        ck(gf_flow_add_node(f, "record weight", &snk, &kk), "node");
        ck(gf_flow_connect(f, ks, 0, kf), "connect");
        ck(gf_flow_connect(f, kf, 0, kk), "connect");
        ck(gf_flow_connect(f, kf, 1, ka), "connect");
        ck(gf_flow_connect(f, kf, 2, kb), "connect");
        ck(gf_flow_connect(f, ka, 0, kf), "connect");
        ck(gf_flow_connect(f, kb, 0, kf), "connect");
        k = ks;
    } else if (topo == GF_JOB_HASH) {
        // keyed hash pipeline: Zipf keys -> murmur3 bucket partition (4096) -> multiset digest sink
        HIPX(hipMalloc(&sf.ms, sizeof(unsigned long long) * (2u * kServeBuckets + 1u)));
        gf_node_desc src = node(GF_ROLE_SOURCE, GF_SRC_KEYED_ZIPF, 1);
        src.u[0] = 1u << 20;
        src.f[0] = 1.1f;
        gf_node_desc part = node(GF_ROLE_TRANSFORM, GF_CLS_KEY_BUCKET, kServeBuckets);
        gf_node_desc snk = node(GF_ROLE_SINK, GF_SNK_CHUNK_DIGEST, 0);
        snk.ptr[1] = (uint64_t)(uintptr_t)sf.ms;
        snk.size[1] = kServeBuckets;
        uint32_t ks, kp, kk;
        ck(gf_flow_add_node(f, "keys", &src, &ks), "node");
        ck(gf_flow_add_node(f, "bucket", &part, &kp), "node");
        ck(gf_flow_add_node(f, "digest", &snk, &kk), "node");
        ck(gf_flow_connect(f, ks, 0, kp), "connect");
        for (uint32_t bb = 0; bb < kServeBuckets; ++bb) ck(gf_flow_connect(f, kp, bb, kk), "connect");
        k = ks;
    } else {
        // photon transport: Generate -> Step -> Alive; Alive.alive -> Step, Alive.done -> Done
        HIPX(hipMalloc(&sf.tally, sizeof(double) * kServeTallyBins));
        gf_node_desc src = node(GF_ROLE_SOURCE, GF_SRC_PHOTON, 1);
        gf_node_desc step = node(GF_ROLE_TRANSFORM, GF_XF_PHOTON_STEP, 1);
        step.ptr[0] = (uint64_t)(uintptr_t)sf.tally;
        step.size[0] = kServeTallyBins;
        gf_node_desc alive = node(GF_ROLE_TRANSFORM, GF_CLS_PHOTON_ALIVE, 2);
        gf_node_desc done = node(GF_ROLE_SINK, GF_SNK_PHOTON_TALLY, 0);
        uint32_t ks, kst, ka, kd;
        ck(gf_flow_add_node(f, "generate", &src, &ks), "node");
        ck(gf_flow_add_node(f, "step", &step, &kst), "node");
        ck(gf_flow_add_node(f, "alive", &alive, &ka), "node");
        ck(gf_flow_add_node(f, "done", &done, &kd), "node");
        ck(gf_flow_connect(f, ks, 0, kst), "connect");
        ck(gf_flow_connect(f, kst, 0, ka), "connect");
        ck(gf_flow_connect(f, ka, 0, kst), "connect");
        ck(gf_flow_connect(f, ka, 1, kd), "connect");
        k = ks;
    }
    HIPX(hipMalloc(&sf.dig, sizeof(unsigned long long)));
    sf.source = k;
    // fixed pool: S = P + K*R for the topology at its full live bound, whatever N a job has
    gf_tuning t = default_photon_tuning();
    t.batch_chunks = kServeBatch;
    t.chunk_capacity = kServeCap;
    t.live_particles = topo == GF_JOB_PHOTON ? kServePhotonLive : kServeLive;
    t.schedule = GF_SCHEDULE_DEVICE;
    t.fuse_transforms = 1;
    t.use_graph = 1;
    const bool team = topo != GF_JOB_HASH;
    t.device_engine = team ? 3u : 1u;
    // the hash pipeline's wide engine: a wide grid spreads the sink's many small per-bucket chunks
    t.max_workgroups = team ? sv.cfg.team_size : 256u;
    t.team_xcd = team ? 1u + slot % 8u : 0u;
    ck(gf_flow_configure(f, &t), "configure");
    const RunShape sh = resolve(*f, t.live_particles);
    t.num_chunks = sh.S;
    ck(gf_flow_configure(f, &t), "configure");
    sf.bytes = pool_bytes_of(payload, sh.S, kServeCap) + sizeof(uint32_t) * sf.rec_cap + (64ull << 20);
}

void destroy_slot_flow(SlotFlow& sf)
{
    if (sf.flow) gf_flow_destroy(sf.flow);
    if (sf.rec) (void)hipFree(sf.rec);
    if (sf.ms) (void)hipFree(sf.ms);
    if (sf.tally) (void)hipFree(sf.tally);
    if (sf.dig) (void)hipFree(sf.dig);
    sf = SlotFlow{};
}

uint64_t host_mix64(uint64_t v) { return gfd::mix64(v); }

// Run one job on a slot's flow; fills the result (and the photon tally).
void run_job(gf_server& sv, gf_context* ctx, SlotFlow& sf, const ServerJob& j, gf_job_result& r,
             std::vector<double>* tally)
{
    hipStream_t s = ctx->stream;
    const uint64_t N = j.d.items;
    gf_flow* f = sf.flow;
    f->tuning.seed = j.d.seed;
    if (j.d.topology == GF_JOB_FEEDBACK) {
        if (N > sf.rec_cap) fail(GF_BAD_ARGUMENT, "serving: job larger than the slot's result buffer");
        hipLaunchKernelGGL(k_fill_u32, dim3(256), dim3(256), 0, s, sf.rec, N, 0xFFFFFFFFu);
    } else if (j.d.topology == GF_JOB_HASH) {
        hipLaunchKernelGGL(k_zero_u64, dim3(32), dim3(256), 0, s, sf.ms, (uint64_t)2u * kServeBuckets);
    } else {
        hipLaunchKernelGGL(k_zero_u64, dim3(4), dim3(256), 0, s, (unsigned long long*)sf.tally, (uint64_t)kServeTallyBins);
    }
    HIPX(hipGetLastError());
    gf_stats st{};
    r.start_ms = sv.now_ms();
    const gf_status rc = gf_flow_run(f, sf.source, N, &st);
    r.status = rc;
    r.steps = st.steps;
    r.engine = st.engine;
    if (rc != GF_OK) fail(rc, f->err);
    if (j.d.topology == GF_JOB_FEEDBACK) {
        HIPX(hipMemsetAsync(sf.dig, 0, sizeof(unsigned long long), s));
        hipLaunchKernelGGL(k_record_digest, dim3(256), dim3(256), 0, s, sf.rec, N, sf.dig);
        HIPX(hipGetLastError());
        unsigned long long d = 0;
        HIPX(hipMemcpyAsync(&d, sf.dig, sizeof d, hipMemcpyDeviceToHost, s));
        HIPX(hipStreamSynchronize(s));
        r.digest = d;
    } else if (j.d.topology == GF_JOB_HASH) {
        std::vector<unsigned long long> ms(2u * kServeBuckets);
        HIPX(hipMemcpyAsync(ms.data(), sf.ms, sizeof(unsigned long long) * ms.size(), hipMemcpyDeviceToHost, s));
        HIPX(hipStreamSynchronize(s));
        uint64_t d = 0;
        for (uint32_t b = 0; b < kServeBuckets; ++b) d += host_mix64(((uint64_t)b << 48) ^ ms[2u * b]) + host_mix64(ms[2u * b + 1u] * 0x9E3779B97F4A7C15ull + b);
        r.digest = d;
    } else {
        tally->assign(kServeTallyBins, 0.0);
        HIPX(hipMemcpyAsync(tally->data(), sf.tally, sizeof(double) * kServeTallyBins, hipMemcpyDeviceToHost, s));
        HIPX(hipStreamSynchronize(s));
        double sum = 0;
        for (double v : *tally) sum += v;
        r.tally_sum = sum;
    }
}

void server_worker(gf_server* sv, uint32_t slot)
{
    gf_context* ctx = nullptr;
    if (gf_context_create(sv->ctx->device, &ctx) != GF_OK) {
        {
            std::lock_guard<std::mutex> lk(sv->mu);
            sv->err = "serving: slot context creation failed";
            sv->warm++;
        }
        sv->cv.notify_all();
        return;
    }
    SlotFlow flows[3];
    if (sv->cfg.prewarm) {
        // build every topology's flow and engine up front (one tiny job each), within the budget
        for (uint32_t tp = 0; tp < 3u; ++tp) {
            {
                std::lock_guard<std::mutex> lk(sv->mu);
                if (sv->mem_in_use + sv->cfg.topology_bytes[tp] > sv->cfg.memory_budget) continue;
                sv->mem_in_use += sv->cfg.topology_bytes[tp];
            }
            std::string msg;
            gf_job_result r{};
            std::vector<double> tally;
            ServerJob w{0, gf_job_desc{tp, 0u, 64u, 1u}, 0.0};
            if (guard(&msg, [&] {
                    flows[tp] = make_slot_flow(*sv, ctx, tp, slot);
                    run_job(*sv, ctx, flows[tp], w, r, &tally);
                }) != GF_OK) {
                std::lock_guard<std::mutex> lk(sv->mu);
                sv->err = msg;
                if (!flows[tp].flow) sv->mem_in_use -= sv->cfg.topology_bytes[tp];  // never built: refund
            }
        }
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->warm++;
    }
    sv->cv.notify_all();
    for (;;) {
        ServerJob j;
        uint64_t need = 0;
        {
            std::unique_lock<std::mutex> lk(sv->mu);
            // first come, first served: the oldest job runs in the next free slot once its
            // topology's flow is resident here or fits the free budget. A slot waiting for memory
            // first frees its own resident flows of other topologies (it is idle, so they are too):
            // every slot that waits does the same, so a head job whose topology fits the budget
            // alone is always admitted eventually.
            for (;;) {
                if (sv->stop && sv->queue.empty()) break;
                if (!sv->queue.empty()) {
                    const ServerJob& h = sv->queue.front();
                    const uint32_t tp = h.d.topology;
                    need = flows[tp].flow ? 0u : sv->cfg.topology_bytes[tp];
                    if (sv->mem_in_use + need <= sv->cfg.memory_budget) break;
                    bool evicted = false;
                    for (uint32_t o = 0; o < 3u; ++o)
                        if (o != tp && flows[o].flow) {
                            destroy_slot_flow(flows[o]);
                            sv->mem_in_use -= sv->cfg.topology_bytes[o];
                            evicted = true;
                        }
                    if (evicted) {
                        sv->cv.notify_all();
                        continue;
                    }
                }
                sv->cv.wait(lk);
            }
            if (sv->stop && sv->queue.empty()) break;
            j = sv->queue.front();
            sv->queue.pop_front();
            sv->mem_in_use += need;
        }
        gf_job_result r{};
        r.job = j.id;
        r.topology = j.d.topology;
        r.items = j.d.items;
        r.slot = slot;
        r.submit_ms = j.t_submit;
        std::vector<double> tally;
        std::string msg;
        const gf_status rc = guard(&msg, [&] {
            SlotFlow& sf = flows[j.d.topology];
            if (!sf.flow) {
                try {
                    sf = make_slot_flow(*sv, ctx, j.d.topology, slot);
                } catch (...) {
                    destroy_slot_flow(sf);
                    {
                        std::lock_guard<std::mutex> lk(sv->mu);
                        sv->mem_in_use -= need;  // the flow was never built: refund its charge
                    }
                    sv->cv.notify_all();
                    throw;
                }
            }
            run_job(*sv, ctx, sf, j, r, &tally);
        });
        if (rc != GF_OK) r.status = rc;
        r.end_ms = sv->now_ms();
        {
            std::lock_guard<std::mutex> lk(sv->mu);
            if (j.d.topology == GF_JOB_PHOTON && r.status == GF_OK) sv->tallies[j.id] = std::move(tally);
            if (rc != GF_OK) sv->err = msg;
            sv->done.push_back(r);
            sv->finished++;
        }
        sv->cv.notify_all();
    }
    for (auto& sf : flows) destroy_slot_flow(sf);
    gf_context_destroy(ctx);
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
        if (!c.slots) c.slots = 8u;
        if (c.slots > 64u) fail(GF_BAD_ARGUMENT, "serving: at most 64 slots");
        if (!c.team_size) c.team_size = std::max<uint32_t>(1u, std::min<uint32_t>(32u, ctx->cus / 8u) - std::min<uint32_t>(4u, ctx->cus / 16u));
        if (!c.memory_budget) c.memory_budget = 64ull << 30;
        if (!c.max_items) c.max_items = 10000000ull;
        // device bytes a slot holds for each topology once it has run one such job
        c.topology_bytes[GF_JOB_FEEDBACK] = pool_bytes_of(GF_PAYLOAD_PARTICLE, kServeLive + 5u, kServeCap) + 4ull * c.max_items + (64ull << 20);
        c.topology_bytes[GF_JOB_HASH] = pool_bytes_of(GF_PAYLOAD_KEYED, kServeLive + kServeBuckets, kServeCap) + (64ull << 20);
        c.topology_bytes[GF_JOB_PHOTON] = pool_bytes_of(GF_PAYLOAD_PHOTON, kServePhotonLive + 3u, kServeCap) + (64ull << 20);
        for (uint32_t t = 0; t < 3u; ++t)
            if (c.topology_bytes[t] > c.memory_budget) fail(GF_BAD_ARGUMENT, "serving: the memory budget cannot hold one flow");
        sv->t0 = std::chrono::steady_clock::now();
        gf_server* p = sv.get();
        for (uint32_t i = 0; i < c.slots; ++i) p->workers.emplace_back(server_worker, p, i);
        if (c.prewarm) {
            std::unique_lock<std::mutex> lk(p->mu);
            p->cv.wait(lk, [&] { return p->warm == c.slots || !p->err.empty(); });
        }
        *out = sv.release();
    });
}

gf_status gf_server_submit(gf_server* sv, const gf_job_desc* job, uint64_t* id)
{
    if (!sv || !job || !id) return GF_BAD_ARGUMENT;
    if (job->topology > GF_JOB_PHOTON || job->items == 0) return GF_BAD_ARGUMENT;
    if (job->topology == GF_JOB_FEEDBACK && job->items > sv->cfg.max_items) return GF_BAD_ARGUMENT;
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
    for (auto& t : sv->workers) t.join();
    delete sv;
    return GF_OK;
}

}  // extern "C"
