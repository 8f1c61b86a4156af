// The particle feedback network through the typed C++ Flow API (include/graphflow.hpp), written
// the way graphFlow user code is (examples/particle_feedback, flow/Flow.hpp:85-412).
//   flow_example cpu        host-only checks: outputChunkBound literals and the status mapping
//   flow_example gpu N      runs N particles on the device and prints the record digest
#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "graphflow.hpp"

namespace gfa = graphFlow::amd;

static uint64_t mix64(uint64_t v)
{
    v ^= v >> 30;
    v *= 0xBF58476D1CE4E5B9ull;
    v ^= v >> 27;
    v *= 0x94D049BB133111EBull;
    v ^= v >> 31;
    return v;
}

template <class E>
static bool throws(gf_status s)
{
    try {
        gfa::check(s, "probe");
    } catch (const E&) {
        return true;
    } catch (...) {
    }
    return false;
}

int main(int argc, char** argv)
{
    const std::string mode = argc > 1 ? argv[1] : "cpu";
    if (mode == "cpu") {
        // tests/unit/test_reservation.cpp:49-61
        bool ok = gfa::outputChunkBound(4097, 128, 2) == 34 && gfa::outputChunkBound(17, 127, 2) == 2 &&
                  gfa::outputChunkBound(64, 128, 64) == 64;
        ok = ok && throws<gfa::BatchError>(GF_INVALID) && throws<gfa::BatchError>(GF_PLAN_MISMATCH) &&
             throws<gfa::SchedulerStallError>(GF_STALL) && throws<gfa::PoolExhaustedError>(GF_POOL_EXHAUSTED) &&
             throws<std::invalid_argument>(GF_BAD_ARGUMENT) && throws<std::logic_error>(GF_STAGE_FAILED) &&
             throws<gfa::DeviceError>(GF_DEVICE_ERROR);
        bool bad_arg = false;
        try {
            gfa::outputChunkBound(5, 0, 1);
        } catch (const std::invalid_argument&) {
            bad_arg = true;
        }
        std::printf("cpu checks %s\n", ok && bad_arg ? "ok" : "FAILED");
        return ok && bad_arg ? 0 : 1;
    }
    const uint64_t N = argc > 2 ? std::stoull(argv[2]) : 1000000ull;
    try {
        gfa::Context ctx(0);
        uint32_t* rec = nullptr;
        if (hipMalloc((void**)&rec, sizeof(uint32_t) * N) != hipSuccess) return 2;
        gfa::ParticleFlow flow(ctx);
        auto gen = flow.source("generate", gfa::UniformParticles());
        auto filter = flow.partition("filter", gfa::WeightRoute(1e-3f, GF_ROUTE_PARITY, 2));
        auto a = flow.transform("process A", gfa::ScaleWeight(0.5f));
        auto b = flow.transform("process B", gfa::ScaleWeight(0.75f));
        auto sink = flow.sink("record weight", gfa::RecordWeight(rec, N));
        flow.connect(gen, filter);
        flow.connect(filter.out<0>(), sink);
        flow.connect(filter.out<1>(), a);
        flow.connect(filter.out<2>(), b);
        flow.connect(a, filter);
        flow.connect(b, filter);
        gfa::Tuning t;
        t.seed = 42;
        t.batchChunks = 8;
        t.chunkCapacity = 8192;
        t.liveParticles = 16384;
        t.fuseTransforms = true;
        flow.configure(t);
        const gfa::SchedulerStats st = flow.run(gen, N);
        std::vector<uint32_t> h(N);
        if (hipMemcpy(h.data(), rec, sizeof(uint32_t) * N, hipMemcpyDeviceToHost) != hipSuccess) return 2;
        uint64_t d = 0;
        for (uint64_t i = 0; i < N; ++i) d += mix64(i << 32 | h[i]);
        std::printf("sunk %llu steps %llu engine %u digest %llu\n", (unsigned long long)st.sunk,
                    (unsigned long long)st.steps, st.engine, (unsigned long long)d);
        // a misuse surfaces as the reference's exception type
        bool caught = false;
        try {
            flow.connect(gen, 7, sink);
        } catch (const std::invalid_argument&) {
            caught = true;
        }
        std::printf("bad connect -> invalid_argument: %s\n", caught ? "yes" : "no");
        (void)hipFree(rec);
        return caught ? 0 : 1;
    } catch (const std::exception& e) {
        std::printf("error: %s\n", e.what());
        return 1;
    }
}
