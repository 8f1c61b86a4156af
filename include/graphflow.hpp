// graphflow.hpp — typed C++ Flow API over the C ABI of libgraphflow_amd.so.
//
// Mirrors graphFlow's user-facing surface (include/graphFlow/flow/Flow.hpp:85-412):
// source / partition / transform / sink registration, connect(node, node),
// connect(node.out(bin), node), connect(node, bin, node), configure(Tuning) and
// run(source, N) -> SchedulerStats, with the reference's exception types
// (sched/GraphRunner.hpp:62-66,355-378; flow/Flow.hpp:488-570). The payload is a template
// parameter as in Flow<TPayload, TDevice, TStorage>; the device is a Context (one HIP device
// and its stream) and the storage is the library's SoA chunk pool. Functors are the op
// descriptors compiled into the library (see graphflow.h), built by the factory functions below.
//
// Header only; link with -lgraphflow_amd. No allocation or exception crosses the C ABI: every
// call's gf_status is turned into an exception here, on the caller's side.
#pragma once

#include <cstdint>
#include <chrono>
#include <optional>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "graphflow.h"

namespace graphFlow::amd {

// ---- errors (same names as graphFlow's) -------------------------------------------------------
struct BatchError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct SchedulerStallError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct PoolExhaustedError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct DeviceError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

inline void check(gf_status s, const std::string& what)
{
    if (s == GF_OK) return;
    const std::string msg = what.empty() ? std::string(gf_status_name(s)) : what + " [" + gf_status_name(s) + "]";
    switch (s) {
    case GF_INVALID:
    case GF_OVERFLOW:
    case GF_PLAN_MISMATCH: throw BatchError(msg);
    case GF_STALL: throw SchedulerStallError(msg);
    case GF_POOL_EXHAUSTED: throw PoolExhaustedError(msg);
    case GF_BAD_ARGUMENT: throw std::invalid_argument(msg);
    case GF_STAGE_FAILED: throw std::logic_error(msg);
    default: throw DeviceError(msg);
    }
}

// outputChunkBound (core/Reservation.hpp:35-50)
inline std::uint64_t outputChunkBound(std::uint64_t n, std::uint64_t capacity, std::uint64_t maxBins)
{
    std::uint64_t r = 0;
    check(gf_output_chunk_bound(n, capacity, maxBins, &r), "outputChunkBound");
    return r;
}

// ---- device ----------------------------------------------------------------------------------
class Context {
public:
    explicit Context(int device = 0) { check(gf_context_create(device, &h_), "gf_context_create"); }
    ~Context()
    {
        if (h_) gf_context_destroy(h_);
    }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    gf_context* handle() const { return h_; }

private:
    gf_context* h_ = nullptr;
};

// ---- functors: op descriptors of the compiled-in set --------------------------------------------
struct Op {
    gf_node_desc d{};
};

inline Op op(std::uint32_t role, std::uint32_t code, std::uint32_t bins = 1)
{
    Op o;
    o.d.role = role;
    o.d.op = code;
    o.d.bins = bins;
    return o;
}
inline Op UniformParticles() { return op(GF_ROLE_SOURCE, GF_SRC_PARTICLE_UNIFORM); }
inline Op ZipfKeys(std::uint32_t values = 1u << 20, float s = 1.1f)
{
    Op o = op(GF_ROLE_SOURCE, GF_SRC_KEYED_ZIPF);
    o.d.u[0] = values;
    o.d.f[0] = s;
    return o;
}
inline Op PhotonSource() { return op(GF_ROLE_SOURCE, GF_SRC_PHOTON); }
// bin 0 when weight < threshold, else 1 + route(fmix32(id)) over `destinations`
inline Op WeightRoute(float threshold = 1e-3f, std::uint32_t distribution = GF_ROUTE_PARITY, std::uint32_t destinations = 2)
{
    Op o = op(GF_ROLE_TRANSFORM, GF_CLS_WEIGHT_ROUTE, 1u + destinations);
    o.d.u[0] = distribution;
    o.d.f[0] = threshold;
    return o;
}
inline Op KeyBucket(std::uint32_t buckets = 4096) { return op(GF_ROLE_TRANSFORM, GF_CLS_KEY_BUCKET, buckets); }
inline Op PhotonAlive() { return op(GF_ROLE_TRANSFORM, GF_CLS_PHOTON_ALIVE, 2); }
inline Op ScaleWeight(float m)
{
    Op o = op(GF_ROLE_TRANSFORM, GF_XF_SCALE_WEIGHT);
    o.d.f[0] = m;
    return o;
}
inline Op PhotonStep(double* tally, std::uint64_t bins = 1024)
{
    Op o = op(GF_ROLE_TRANSFORM, GF_XF_PHOTON_STEP);
    o.d.ptr[0] = (std::uint64_t)(uintptr_t)tally;
    o.d.size[0] = bins;
    return o;
}
// dense u64 histogram over weight bit patterns [lo, lo + n)
inline Op WeightHistogram(std::uint64_t* counts, std::uint32_t lo, std::uint32_t n)
{
    Op o = op(GF_ROLE_SINK, GF_SNK_HIST, 0);
    o.d.u[0] = 0;
    o.d.u[1] = lo;
    o.d.u[2] = n;
    o.d.ptr[0] = (std::uint64_t)(uintptr_t)counts;
    o.d.size[0] = n;
    return o;
}
inline Op RecordWeight(std::uint32_t* record, std::uint64_t n)
{
    Op o = op(GF_ROLE_SINK, GF_SNK_RECORD_WEIGHT, 0);
    o.d.ptr[0] = (std::uint64_t)(uintptr_t)record;
    o.d.size[0] = n;
    return o;
}
inline Op ChunkDigest(std::uint32_t* log, std::uint64_t logCap, std::uint64_t* multiset, std::uint32_t bins)
{
    Op o = op(GF_ROLE_SINK, GF_SNK_CHUNK_DIGEST, 0);
    o.d.ptr[0] = (std::uint64_t)(uintptr_t)log;
    o.d.size[0] = logCap;
    o.d.ptr[1] = (std::uint64_t)(uintptr_t)multiset;
    o.d.size[1] = bins;
    return o;
}
inline Op PhotonDone() { return op(GF_ROLE_SINK, GF_SNK_PHOTON_TALLY, 0); }

// ---- tuning and stats (flow/Tuning.hpp:27-145, sched/GraphRunner.hpp:154-210) --------------------
enum class Schedule : std::uint32_t { plain = GF_SCHEDULE_HOST, device = GF_SCHEDULE_DEVICE };
// BinPath (core/BinPath.hpp): dense LDS counters or per-tile hash tables; automatic = hash above 4096 bins
enum class BinPath : std::uint32_t { automatic = GF_BIN_PATH_AUTO, dense = GF_BIN_PATH_DENSE, hash = GF_BIN_PATH_HASH };

struct Tuning {
    std::uint64_t seed = 0;
    std::optional<std::uint32_t> batchChunks, chunkCapacity, liveParticles, numChunks, tileSize, stepsPerLaunch;
    std::optional<std::uint32_t> inflightBatches;           // Tuning::inflightBatches
    BinPath binPath = BinPath::automatic;                   // Tuning::binPath
    std::optional<std::uint32_t> hashTableEntries;          // hash path: slots of a tile's bin table
    std::optional<std::chrono::milliseconds> stallTimeout;  // Tuning::stallTimeout
    bool fusedBatches = false;                              // plain schedule: gf_transform_batch per partition batch
    std::optional<std::uint32_t> sourceChunks;              // device schedule: chunks a source lane admits per step
    Schedule schedule = Schedule::device;
    bool trace = false;
    bool fuseTransforms = false;
    std::uint32_t deviceEngine = 0;  // 0 automatic, 1 wide, 2 one workgroup, 3 team
    std::optional<std::uint32_t> teamSize, teamXcd;

    gf_tuning c() const
    {
        gf_tuning t{};
        t.seed = seed;
        t.batch_chunks = batchChunks.value_or(0);
        t.chunk_capacity = chunkCapacity.value_or(0);
        t.live_particles = liveParticles.value_or(0);
        t.num_chunks = numChunks.value_or(0);
        t.tile_items = tileSize.value_or(0);
        t.schedule = (std::uint32_t)schedule;
        t.steps_per_launch = stepsPerLaunch.value_or(0);
        t.trace = trace ? 1u : 0u;
        t.use_graph = 1;
        t.fuse_transforms = fuseTransforms ? 1u : 0u;
        t.max_workgroups = teamSize.value_or(0);
        t.device_engine = deviceEngine;
        t.team_xcd = teamXcd ? *teamXcd + 1u : 0u;
        t.inflight_batches = inflightBatches.value_or(0);
        t.bin_path = (std::uint32_t)binPath;
        t.hash_table_entries = hashTableEntries.value_or(0);
        t.stall_timeout_ms = stallTimeout ? (std::uint32_t)stallTimeout->count() : 0u;
        t.fused_batches = fusedBatches ? 1u : 0u;
        t.source_chunks = sourceChunks.value_or(0);
        t.photon_layers = 10;
        t.photon_tally_bins = 1024;
        t.photon_layer_mfp = 1.0;
        t.photon_mu_t = 1.0;
        t.photon_albedo = 0.9;
        t.photon_hg_g = 0.9;
        t.photon_roulette_w = 1e-4;
        t.photon_roulette_p = 0.1;
        return t;
    }
};

struct SchedulerStats {
    std::uint64_t completedBatches = 0, steps = 0, generated = 0, sunk = 0;
    std::uint32_t poolChunks = 0, peakLiveChunks = 0, chunkCapacity = 0, liveParticles = 0;
    std::uint64_t kernelLaunches = 0;
    double wallMs = 0;
    std::uint32_t engine = 0;
};

// ---- typed flow ----------------------------------------------------------------------------------
struct NodeHandle {
    std::uint32_t key = 0;
    struct Out {
        std::uint32_t key, bin;
    };
    Out out(std::uint32_t bin) const { return Out{key, bin}; }
    template <std::uint32_t Bin>
    Out out() const
    {
        return Out{key, Bin};
    }
};

template <std::uint32_t Payload>
class Flow {
public:
    explicit Flow(Context& ctx) { check(gf_flow_create(ctx.handle(), Payload, &h_), "gf_flow_create"); }
    ~Flow()
    {
        if (h_) gf_flow_destroy(h_);
    }
    Flow(const Flow&) = delete;
    Flow& operator=(const Flow&) = delete;

    NodeHandle source(const std::string& name, const Op& f) { return add(name, f, GF_ROLE_SOURCE); }
    NodeHandle partition(const std::string& name, const Op& classifier, std::uint32_t update = GF_UPD_NONE)
    {
        Op o = classifier;
        o.d.upd = update;
        return add(name, o, GF_ROLE_TRANSFORM);
    }
    NodeHandle transform(const std::string& name, const Op& f) { return add(name, f, GF_ROLE_TRANSFORM); }
    NodeHandle sink(const std::string& name, const Op& f) { return add(name, f, GF_ROLE_SINK); }

    void connect(NodeHandle from, NodeHandle to) { call(gf_flow_connect(h_, from.key, 0, to.key), "connect"); }
    void connect(NodeHandle::Out from, NodeHandle to) { call(gf_flow_connect(h_, from.key, from.bin, to.key), "connect"); }
    void connect(NodeHandle from, std::uint32_t bin, NodeHandle to) { call(gf_flow_connect(h_, from.key, bin, to.key), "connect"); }

    void configure(const Tuning& t)
    {
        const gf_tuning c = t.c();
        call(gf_flow_configure(h_, &c), "configure");
    }

    SchedulerStats run(NodeHandle source, std::uint64_t items)
    {
        gf_stats st{};
        call(gf_flow_run(h_, source.key, items, &st), "run");
        SchedulerStats r;
        r.completedBatches = st.completed_batches;
        r.steps = st.steps;
        r.generated = st.generated;
        r.sunk = st.sunk;
        r.poolChunks = st.pool_chunks;
        r.peakLiveChunks = st.peak_live_chunks;
        r.chunkCapacity = st.chunk_capacity;
        r.liveParticles = st.live_particles;
        r.kernelLaunches = st.kernel_launches;
        r.wallMs = st.wall_ms;
        r.engine = st.engine;
        return r;
    }

    // the autotuner's pick for the knobs left open (no counterpart in graphFlow)
    gf_tuning autotune(NodeHandle source, std::uint64_t probeItems = 0)
    {
        gf_tuning t{};
        double ms = 0;
        call(gf_flow_autotune(h_, source.key, probeItems, &t, &ms), "autotune");
        return t;
    }

    gf_flow* handle() const { return h_; }

private:
    NodeHandle add(const std::string& name, Op o, std::uint32_t role)
    {
        o.d.role = role;
        NodeHandle n;
        call(gf_flow_add_node(h_, name.c_str(), &o.d, &n.key), "add node '" + name + "'");
        return n;
    }
    void call(gf_status s, const std::string& what)
    {
        if (s == GF_OK) return;
        char buf[512] = {0};
        gf_flow_last_error(h_, buf, sizeof buf);
        check(s, what + ": " + buf);
    }
    gf_flow* h_ = nullptr;
};

using ParticleFlow = Flow<GF_PAYLOAD_PARTICLE>;
using KeyedFlow = Flow<GF_PAYLOAD_KEYED>;
using PhotonFlow = Flow<GF_PAYLOAD_PHOTON>;

}  // namespace graphFlow::amd
