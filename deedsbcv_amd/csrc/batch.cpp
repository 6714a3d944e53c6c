// Multi-atlas batch registration: many independent pairs, registered in
// groups of up to `group` pairs.  The pairs of a group run level-synchronously
// on one context (register_group in register.hip): all their pre-BP kernels
// of a level, then the BP of all their trees in one launch sequence -- one
// launch per round for every tree, so the latency-bound rounds are paid once
// per level instead of once per pair -- then their post-BP kernels.  Every
// kernel sees the inputs it sees in a registration of the pair alone, so each
// output is bit-identical to deeds_register on that pair.
#include <algorithm>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "common.h"

namespace deeds {

namespace {

void check_pairs(const deeds_pair* pairs, int count, const deeds_schedule* S, int group) {
    DEEDS_REQUIRE(count >= 0 && (count == 0 || pairs), "pairs is NULL");
    DEEDS_REQUIRE(group >= 1, "concurrency (pairs per group) must be >= 1");
    for (int i = 0; i < count; i++) {
        const deeds_pair& P = pairs[i];
        check_schedule(S, P.m, P.n, P.o);
        DEEDS_REQUIRE(P.fixed && P.moving && P.flow, "a pair has a null pointer");
    }
}

// Idle contexts per device for device-pointer batches, kept for reuse (their
// workspaces stay allocated); like the default contexts they live until exit.
std::mutex g_pool_mu;
std::map<int, std::vector<deeds_ctx*>> g_pool;

deeds_ctx* acquire_ctx(int device) {
    {
        std::lock_guard<std::mutex> lk(g_pool_mu);
        auto& v = g_pool[device];
        if (!v.empty()) {
            deeds_ctx* c = v.back();
            v.pop_back();
            return c;
        }
    }
    deeds_ctx* c = nullptr;
    const int rc = deeds_ctx_create(device, &c);
    if (rc != DEEDS_OK) fail(rc, std::string("batch: cannot create a context: ") + deeds_last_error());
    return c;
}

void release_ctx(int device, deeds_ctx* c) {
    std::lock_guard<std::mutex> lk(g_pool_mu);
    g_pool[device].push_back(c);
}

size_t flow_count(const deeds_pair& P, const deeds_schedule& S) {
    const int g = S.grid[S.levels - 1];
    return 3 * (size_t)(P.m / g) * (P.n / g) * (P.o / g);
}

}  // namespace

}  // namespace deeds

using namespace deeds;

extern "C" {

int deeds_register_batch(const deeds_pair* pairs, int count, const deeds_schedule* sched, int concurrency,
                         deeds_timing* timings) {
    return guarded(nullptr, [&] {
        check_pairs(pairs, count, sched, concurrency);
        deeds_ctx* c = default_ctx();
        DEEDS_HIP(hipSetDevice(c->device));
        for (int g0 = 0; g0 < count; g0 += concurrency) {
            const int gn = std::min(concurrency, count - g0);
            // device copies of the group's inputs and outputs
            std::vector<deeds_pair> dev(gn);
            for (int i = 0; i < gn; i++) {
                const deeds_pair& P = pairs[g0 + i];
                const size_t sz = (size_t)P.m * P.n * P.o, nf = flow_count(P, *sched);
                float* din = c->get<float>("h.batch.in." + std::to_string(i), 2 * sz);
                float* dout = c->get<float>("h.batch.out." + std::to_string(i), nf + (P.warped ? sz : 0));
                h2d(din, P.fixed, sz, c->stream);
                h2d(din + sz, P.moving, sz, c->stream);
                dev[i] = P;
                dev[i].fixed = din;
                dev[i].moving = din + sz;
                dev[i].flow = dout;
                dev[i].warped = P.warped ? dout + nf : nullptr;
            }
            register_group(c, dev.data(), gn, *sched, timings ? timings + g0 : nullptr);
            for (int i = 0; i < gn; i++) {
                const deeds_pair& P = pairs[g0 + i];
                d2h(P.flow, dev[i].flow, flow_count(P, *sched), c->stream);
                if (P.warped) d2h(P.warped, dev[i].warped, (size_t)P.m * P.n * P.o, c->stream);
            }
            DEEDS_HIP(hipStreamSynchronize(c->stream));
        }
    });
}

int deeds_register_batch_device(int device, const deeds_pair* pairs, int count, const deeds_schedule* sched,
                                int concurrency, deeds_timing* timings) {
    return guarded(nullptr, [&] {
        int n = 0;
        DEEDS_HIP(hipGetDeviceCount(&n));
        DEEDS_REQUIRE(device >= 0 && device < n, "device index out of range");
        check_pairs(pairs, count, sched, concurrency);
        deeds_ctx* c = acquire_ctx(device);
        try {
            DEEDS_HIP(hipSetDevice(device));
            for (int g0 = 0; g0 < count; g0 += concurrency)
                register_group(c, pairs + g0, std::min(concurrency, count - g0), *sched,
                               timings ? timings + g0 : nullptr);
        } catch (...) {
            release_ctx(device, c);
            throw;
        }
        release_ctx(device, c);
    });
}

}  // extern "C"
