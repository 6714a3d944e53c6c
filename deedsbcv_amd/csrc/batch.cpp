// Multi-atlas batch registration: many independent pairs, several at once.
//
// A single registration leaves the GPU idle while the host builds MST trees
// and while its latency-bound BP rounds run with a few workgroups each.  The
// batch runs up to `concurrency` pairs simultaneously, each on its own
// context (own streams, own workspace, own host tree threads), so the kernels
// of one pair fill the gaps of another.  Every pair goes through exactly the
// same code as deeds_ctx_register, so its outputs are bit-identical to a
// registration of that pair alone.
#include <atomic>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "common.h"

namespace deeds {

namespace {

// Idle contexts per device, kept for reuse across batches (their workspaces
// stay allocated).  Like the default contexts, they live until process exit.
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
    int rc = deeds_ctx_create(device, &c);
    if (rc != DEEDS_OK) fail(rc, std::string("batch: cannot create a context: ") + deeds_last_error());
    return c;
}

void release_ctx(int device, deeds_ctx* c) {
    std::lock_guard<std::mutex> lk(g_pool_mu);
    g_pool[device].push_back(c);
}

void run_batch(int device, const deeds_pair* pairs, int count, const deeds_schedule* S, int concurrency,
               deeds_timing* timings, bool host_pointers) {
    DEEDS_REQUIRE(count >= 0 && (count == 0 || pairs), "pairs is NULL");
    DEEDS_REQUIRE(concurrency >= 1, "concurrency must be >= 1");
    for (int i = 0; i < count; i++) {
        const deeds_pair& P = pairs[i];
        check_schedule(S, P.m, P.n, P.o);
        DEEDS_REQUIRE(P.fixed && P.moving && P.flow, "a pair has a null pointer");
    }
    std::atomic<int> next{0};
    std::mutex err_mu;
    int status = DEEDS_OK;
    std::string message;
    auto worker = [&]() {
        deeds_ctx* ctx = nullptr;
        try {
            ctx = acquire_ctx(device);
            DEEDS_HIP(hipSetDevice(device));
            for (;;) {
                {
                    std::lock_guard<std::mutex> lk(err_mu);
                    if (status != DEEDS_OK) break;
                }
                const int i = next.fetch_add(1);
                if (i >= count) break;
                const deeds_pair& P = pairs[i];
                deeds_timing* t = timings ? &timings[i] : nullptr;
                if (!host_pointers) {
                    register_pair(ctx, P.fixed, P.moving, P.X, P.m, P.n, P.o, *S, P.flow, P.warped, t);
                    continue;
                }
                const size_t sz = (size_t)P.m * P.n * P.o;
                const int g = S->grid[S->levels - 1];
                const size_t sz1 = (size_t)(P.m / g) * (P.n / g) * (P.o / g);
                float* din = ctx->get<float>("b.in", 2 * sz);
                float* dout = ctx->get<float>("b.out", sz + 3 * sz1);
                h2d(din, P.fixed, sz, ctx->stream);
                h2d(din + sz, P.moving, sz, ctx->stream);
                register_pair(ctx, din, din + sz, P.X, P.m, P.n, P.o, *S, dout + sz, P.warped ? dout : nullptr, t);
                d2h(P.flow, dout + sz, 3 * sz1, ctx->stream);
                if (P.warped) d2h(P.warped, dout, sz, ctx->stream);
                DEEDS_HIP(hipStreamSynchronize(ctx->stream));
            }
        } catch (const Error& e) {
            std::lock_guard<std::mutex> lk(err_mu);
            if (status == DEEDS_OK) {
                status = e.code;
                message = e.msg;
            }
        } catch (const std::exception& e) {
            std::lock_guard<std::mutex> lk(err_mu);
            if (status == DEEDS_OK) {
                status = DEEDS_ERR_INTERNAL;
                message = e.what();
            }
        }
        if (ctx) release_ctx(device, ctx);
    };
    const int nw = std::min(concurrency, std::max(count, 1));
    std::vector<std::thread> threads;
    for (int w = 0; w < nw; w++) threads.emplace_back(worker);
    for (auto& th : threads) th.join();
    if (status != DEEDS_OK) fail(status, "batch: " + message);
}

}  // namespace

}  // namespace deeds

using namespace deeds;

extern "C" {

int deeds_register_batch(const deeds_pair* pairs, int count, const deeds_schedule* sched, int concurrency,
                         deeds_timing* timings) {
    return guarded(nullptr, [&] {
        int device = default_ctx()->device;
        run_batch(device, pairs, count, sched, concurrency, timings, true);
    });
}

int deeds_register_batch_device(int device, const deeds_pair* pairs, int count, const deeds_schedule* sched,
                                int concurrency, deeds_timing* timings) {
    return guarded(nullptr, [&] {
        int n = 0;
        DEEDS_HIP(hipGetDeviceCount(&n));
        DEEDS_REQUIRE(device >= 0 && device < n, "device index out of range");
        run_batch(device, pairs, count, sched, concurrency, timings, false);
    });
}

}  // extern "C"
