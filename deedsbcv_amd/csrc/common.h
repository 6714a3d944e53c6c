// Internal shared definitions of the deeds_hip library: error handling, the
// context (device, streams, cached device workspace) and small helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/deeds_hip.h"

namespace deeds {

struct Error {
    int code;
    std::string msg;
};

[[noreturn]] inline void fail(int code, const std::string& msg) { throw Error{code, msg}; }

#define DEEDS_HIP(expr)                                                                       \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
            ::deeds::fail(DEEDS_ERR_HIP, std::string(#expr) + " failed: " + hipGetErrorString(e_)); \
    } while (0)

#define DEEDS_REQUIRE(cond, msg)                                     \
    do {                                                             \
        if (!(cond)) ::deeds::fail(DEEDS_ERR_ARG, std::string(msg)); \
    } while (0)

// Fault injection (deeds_inject_fault, or DEEDS_FAULT_INJECT=alloc:N / launch:N
// in the environment): true when an injected failure of `kind` is due.  Kinds:
// 1 device or pinned allocation, 2 kernel launch.
bool fault_now(int kind);

#define DEEDS_LAUNCH_CHECK()                                                                 \
    do {                                                                                     \
        if (::deeds::fault_now(2)) ::deeds::fail(DEEDS_ERR_HIP, "injected kernel launch failure"); \
        DEEDS_HIP(hipGetLastError());                                                        \
    } while (0)

// A device allocation that grows on demand and is kept for reuse.
struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
};

}  // namespace deeds

struct deeds_ctx {
    int device = 0;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;     // active stream (own or caller's)
    hipStream_t aux_stream = nullptr;  // second registration direction
    hipStream_t side_stream[2] = {nullptr, nullptr};  // small work issued from host helper threads (image stats)
    std::string err;
    std::map<std::string, deeds::DevBuf> ws;
    std::map<std::string, deeds::DevBuf> pinned_ws;
    std::shared_ptr<void> soft_cache;  // last staged soft-variant tree (soft_api.cpp)
    std::shared_ptr<void> graph_cache;  // a batch's captured level graphs (register.hip)

    // Device buffer `name` of at least `bytes` (contents unspecified).
    void* buf(const std::string& name, size_t bytes);
    // Pinned host buffer `name` of at least `bytes`, kept for reuse.  Request
    // each name once per call: growing it waits for the context's streams.
    void* pinned(const std::string& name, size_t bytes);
    template <class T>
    T* get(const std::string& name, size_t count) {
        return static_cast<T*>(buf(name, count * sizeof(T)));
    }
    void release_all();
    // Waits for the context's streams, ignoring errors (after a failed call).
    void drain();
};

namespace deeds {

// Thread-local default context used by the host-pointer entry points.
deeds_ctx* default_ctx();
void set_thread_error(const std::string& msg);

// Runs f, converting exceptions to a status and recording the message.
template <class F>
int guarded(deeds_ctx* ctx, F&& f) {
    try {
        if (ctx) DEEDS_HIP(hipSetDevice(ctx->device));
        f();
        return DEEDS_OK;
    } catch (const Error& e) {
        if (ctx) {
            ctx->err = e.msg;
            ctx->drain();  // nothing of the failed call is still running when the caller regains control
        }
        set_thread_error(e.msg);
        return e.code;
    } catch (const std::bad_alloc&) {
        if (ctx) ctx->err = "host allocation failed";
        set_thread_error("host allocation failed");
        return DEEDS_ERR_ALLOC;
    } catch (const std::exception& e) {
        if (ctx) ctx->err = e.what();
        set_thread_error(e.what());
        return DEEDS_ERR_INTERNAL;
    }
}

inline unsigned ceil_div(long long a, long long b) { return (unsigned)((a + b - 1) / b); }

// Validates a registration schedule for an m x n x o pair (throws Error).
void check_schedule(const deeds_schedule* S, int m, int n, int o);
// One symmetric registration on ctx, device pointers (register.hip).
void register_pair(deeds_ctx* ctx, const float* d_fixed, const float* d_moving, const float* X, int m, int n, int o,
                   const deeds_schedule& S, float* d_flow, float* d_warped, deeds_timing* timing);
// linearBCV's multi-level affine / rigid estimation on ctx (linear_host.cpp).
void linear_register(deeds_ctx* ctx, const float* d_fixed, const float* d_moving, int m, int n, int o,
                     const deeds_schedule& S, bool rigid, float* X_out, int* points_used);
// Numbers of batch level-graph captures and replays so far (JSON object).
std::string graph_report();
// Turns the batch level graphs on or off for later batches.
void set_batch_graphs(bool on);
// Level-synchronous registration of `count` pairs (device pointers) on ctx;
// timings is optional (count entries).  Pairs are validated by the caller.
void register_group(deeds_ctx* ctx, const deeds_pair* pairs, int count, const deeds_schedule& S,
                    deeds_timing* timings);

// Upload / download helpers on a stream.
template <class T>
void h2d(T* dst, const T* src, size_t count, hipStream_t s) {
    DEEDS_HIP(hipMemcpyAsync(dst, src, count * sizeof(T), hipMemcpyHostToDevice, s));
}
template <class T>
void d2h(T* dst, const T* src, size_t count, hipStream_t s) {
    DEEDS_HIP(hipMemcpyAsync(dst, src, count * sizeof(T), hipMemcpyDeviceToHost, s));
}

}  // namespace deeds
