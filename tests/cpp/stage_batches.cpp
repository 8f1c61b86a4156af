// One partition batch at a time through the stage C ABI, the way a graphFlow GraphRunner with a
// HipStage binding drives it (flow/Execution.hpp:1537-1561): launch -> advance until done -> finish
// (classify, host plan, scatter), and the fused gf_transform_batch -> advance -> finish. Checks every
// batch (three outputs of one chunk of C items into 3 balanced bins, fills summing to C, the same
// outputs from both protocols, the pool whole afterwards) and prints the time per batch.
#include "graphflow.h"

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <vector>

#define CK(x)                                                                 \
    do {                                                                      \
        gf_status s_ = (x);                                                   \
        if (s_ != GF_OK) {                                                    \
            std::printf("FAIL status %d at line %d\n", (int)s_, __LINE__);    \
            return 1;                                                         \
        }                                                                     \
    } while (0)

int main(int argc, char** argv)
{
    const int n = argc > 1 ? std::atoi(argv[1]) : 2000;
    const uint32_t C = 8192, S = 64;
    gf_context* ctx;
    CK(gf_context_create(0, &ctx));
    gf_pool* pool;
    CK(gf_pool_create(ctx, GF_PAYLOAD_PARTICLE, S, C, &pool));
    gf_chunk_handle h;
    CK(gf_pool_acquire(pool, &h));
    std::vector<float> w(C, 1.0f);
    std::vector<uint32_t> id(C), gen(C, 0u);
    for (uint32_t i = 0; i < C; ++i) id[i] = i;
    CK(gf_pool_write(pool, 0, (uint64_t)h.index * C, w.data(), C));
    CK(gf_pool_write(pool, 1, (uint64_t)h.index * C, id.data(), C));
    CK(gf_pool_write(pool, 2, (uint64_t)h.index * C, gen.data(), C));
    gf_node_desc nd = {};
    nd.role = GF_ROLE_TRANSFORM;
    nd.op = GF_CLS_ID_MOD;  // balanced: id % 3
    nd.bins = 3;
    const gf_chunk_desc in = {h.index, h.generation, C, 0u};
    gf_chunk_desc out[64], first[2][64];
    uint32_t first_n[2] = {0u, 0u};
    double us[2] = {0.0, 0.0};
    for (int mode = 0; mode < 2; ++mode) {
        gf_stage* st;
        CK(gf_stage_create(pool, &nd, 1, 42, 8, 1024, &st));
        const int warm = 100;
        std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
        for (int it = 0; it < warm + n; ++it) {
            if (it == warm) t0 = std::chrono::steady_clock::now();
            // allocated: the outputs an advance acquired (the planning advance of the three-step
            // protocol, the one advance of the fused batch), summed over the batch's advances
            uint32_t alloc = 0, a = 0, nout = 0;
            int more = 1;
            if (mode == 0) {
                CK(gf_stage_launch(st, &in, 1, 0, nullptr));
                while (more) {
                    CK(gf_stage_advance(st, nullptr, &a, &more));
                    alloc += a;
                }
            } else {
                CK(gf_transform_batch(st, &in, 1, 0, nullptr));
                CK(gf_stage_advance(st, nullptr, &alloc, &more));
            }
            CK(gf_stage_finish(st, out, 64, &nout));
            uint32_t items = 0;
            for (uint32_t k = 0; k < nout; ++k) items += out[k].fill;
            if (nout != 3u || items != C || alloc != 3u) {
                std::printf("FAIL batch %d: %u outputs, %u items, allocated %u\n", it, nout, items, alloc);
                return 1;
            }
            if (it == 0) {
                first_n[mode] = nout;
                for (uint32_t k = 0; k < nout; ++k) first[mode][k] = out[k];
            }
            for (uint32_t k = 0; k < nout; ++k) CK(gf_pool_release(pool, gf_chunk_handle{out[k].index, out[k].generation}));
        }
        us[mode] = std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count() / n;
        CK(gf_stage_destroy(st));
    }
    for (uint32_t k = 0; k < first_n[0]; ++k)
        if (first[0][k].fill != first[1][k].fill || first[0][k].bin != first[1][k].bin) {
            std::printf("FAIL: the protocols' outputs differ at %u\n", k);
            return 1;
        }
    uint32_t counts[6];
    CK(gf_pool_counts(pool, counts));
    if (counts[2] != S - 1u) {
        std::printf("FAIL: %u free chunks after the batches, expected %u\n", counts[2], S - 1u);
        return 1;
    }
    std::printf("ok batches %d three_step_us %.1f transform_batch_us %.1f\n", n, us[0], us[1]);
    return 0;
}
