// C-ABI entry points of deeds_hip: context management and the host-pointer
// forms, one per reference function (see include/deeds_hip.h for the
// reference file:line each replaces).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <cstring>
#include <chrono>
#include <future>
#include <memory>
#include <type_traits>

#include "bp_host.h"
#include "common.h"
#include "kernels.h"
#include "mst_host.h"
#include "stats.h"

using namespace deeds;

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------

void* deeds_ctx::buf(const std::string& name, size_t bytes) {
    DevBuf& b = ws[name];
    if (b.bytes >= bytes && b.ptr) return b.ptr;
    if (b.ptr) {
        DEEDS_HIP(hipStreamSynchronize(stream));
        if (aux_stream) DEEDS_HIP(hipStreamSynchronize(aux_stream));
        for (hipStream_t ss : side_stream)
            if (ss) DEEDS_HIP(hipStreamSynchronize(ss));
        DEEDS_HIP(hipFree(b.ptr));
        b.ptr = nullptr;
        b.bytes = 0;
    }
    size_t want = std::max<size_t>(bytes, 256);
    if (fault_now(1)) fail(DEEDS_ERR_ALLOC, "injected device allocation failure for '" + name + "'");
    hipError_t e = hipMalloc(&b.ptr, want);
    if (e != hipSuccess) {
        b.ptr = nullptr;
        (void)hipGetLastError();
        fail(DEEDS_ERR_ALLOC, "hipMalloc of " + std::to_string(want) + " bytes for '" + name + "' failed: " +
                                  hipGetErrorString(e));
    }
    b.bytes = want;
    return b.ptr;
}

void* deeds_ctx::pinned(const std::string& name, size_t bytes) {
    DevBuf& b = pinned_ws[name];
    if (b.bytes >= bytes && b.ptr) return b.ptr;
    if (b.ptr) {
        // copies from the old block may still be queued
        DEEDS_HIP(hipStreamSynchronize(stream));
        if (aux_stream) DEEDS_HIP(hipStreamSynchronize(aux_stream));
        for (hipStream_t ss : side_stream)
            if (ss) DEEDS_HIP(hipStreamSynchronize(ss));
        DEEDS_HIP(hipHostFree(b.ptr));
        b.ptr = nullptr;
        b.bytes = 0;
    }
    size_t want = std::max<size_t>(bytes, 256);
    if (fault_now(1)) fail(DEEDS_ERR_ALLOC, "injected pinned allocation failure for '" + name + "'");
    hipError_t e = hipHostMalloc(&b.ptr, want, hipHostMallocDefault);
    if (e != hipSuccess) {
        b.ptr = nullptr;
        (void)hipGetLastError();
        fail(DEEDS_ERR_ALLOC, "hipHostMalloc of " + std::to_string(want) + " bytes for '" + name + "' failed: " +
                                  hipGetErrorString(e));
    }
    b.bytes = want;
    return b.ptr;
}

void deeds_ctx::drain() {
    if (stream) (void)hipStreamSynchronize(stream);
    if (aux_stream) (void)hipStreamSynchronize(aux_stream);
    for (hipStream_t ss : side_stream)
        if (ss) (void)hipStreamSynchronize(ss);
    (void)hipGetLastError();
}

void deeds_ctx::release_all() {
    if (stream) (void)hipStreamSynchronize(stream);
    if (aux_stream) (void)hipStreamSynchronize(aux_stream);
    for (hipStream_t ss : side_stream)
        if (ss) (void)hipStreamSynchronize(ss);
    for (auto& kv : ws)
        if (kv.second.ptr) (void)hipFree(kv.second.ptr);
    ws.clear();
    for (auto& kv : pinned_ws)
        if (kv.second.ptr) (void)hipHostFree(kv.second.ptr);
    pinned_ws.clear();
    soft_cache.reset();  // it points into the released buffers
    graph_cache.reset();  // so do the captured graphs
}

namespace deeds {

namespace {
// Injected faults: per kind, the number of events still to pass before one
// fails (-1: none armed).  One-shot: the failing event disarms its kind.
std::atomic<long long> g_fault[3] = {{-1}, {-1}, {-1}};
std::once_flag g_fault_env;

void fault_from_env() {
    const char* e = std::getenv("DEEDS_FAULT_INJECT");
    if (!e) return;
    long long n = 0;
    if (std::sscanf(e, "alloc:%lld", &n) == 1 && n >= 0) g_fault[1] = n;
    else if (std::sscanf(e, "launch:%lld", &n) == 1 && n >= 0) g_fault[2] = n;
}
}  // namespace

bool fault_now(int kind) {
    std::call_once(g_fault_env, fault_from_env);
    if (kind < 1 || kind > 2) return false;
    long long v = g_fault[kind].load(std::memory_order_relaxed);
    while (v >= 0) {
        if (g_fault[kind].compare_exchange_weak(v, v - 1, std::memory_order_relaxed)) return v == 0;
    }
    return false;
}

namespace {
thread_local std::string t_error;
// The per-thread default context is intentionally never destroyed at thread or
// process exit (the HIP runtime may already be shut down by then); it is
// replaced only by deeds_set_default_device.
thread_local deeds_ctx* t_ctx = nullptr;
int g_default_device = 0;
}  // namespace

void set_thread_error(const std::string& msg) { t_error = msg; }

deeds_ctx* default_ctx() {
    if (!t_ctx) {
        deeds_ctx* c = nullptr;
        int rc = deeds_ctx_create(g_default_device, &c);
        if (rc != DEEDS_OK) fail(rc, "cannot create the default context: " + t_error);
        t_ctx = c;
    }
    return t_ctx;
}

}  // namespace deeds

extern "C" {

const char* deeds_version(void) { return "deeds_hip 0.1 (gfx950)"; }

const char* deeds_last_error(void) { return t_error.c_str(); }

int deeds_set_batch_graphs(int enabled) {
    set_batch_graphs(enabled != 0);
    return DEEDS_OK;
}

int deeds_set_datacost_tiles(int planes_per_workgroup, int pipelined) {
    if (planes_per_workgroup < 0 || pipelined < 0 || pipelined > 3) return DEEDS_ERR_ARG;
    pin_datacost_tiles(planes_per_workgroup, pipelined);
    return DEEDS_OK;
}

size_t deeds_tuning_report(char* buf, size_t cap) {
    const std::string r = "{\"datacost_rows\": " + datacost_tuning_report() + ", \"graphs\": " + graph_report() + "}";
    if (buf && cap > 0) {
        const size_t k = std::min(cap - 1, r.size());
        std::memcpy(buf, r.data(), k);
        buf[k] = 0;
    }
    return r.size();
}

int deeds_inject_fault(int kind, long long after) {
    if (kind != DEEDS_FAULT_ALLOC && kind != DEEDS_FAULT_LAUNCH) return DEEDS_ERR_ARG;
    std::call_once(g_fault_env, fault_from_env);
    g_fault[kind] = after < 0 ? -1 : after;
    return DEEDS_OK;
}

int deeds_set_default_device(int device) {
    return guarded(nullptr, [&] {
        int count = 0;
        DEEDS_HIP(hipGetDeviceCount(&count));
        DEEDS_REQUIRE(device >= 0 && device < count, "device index out of range");
        g_default_device = device;
        if (t_ctx) {
            deeds_ctx_destroy(t_ctx);
            t_ctx = nullptr;
        }
    });
}

int deeds_ctx_create(int device, deeds_ctx** out) {
    if (!out) {
        set_thread_error("deeds_ctx_create: out is NULL");
        return DEEDS_ERR_ARG;
    }
    *out = nullptr;
    auto ctx = new (std::nothrow) deeds_ctx();
    if (!ctx) return DEEDS_ERR_ALLOC;
    int rc = guarded(nullptr, [&] {
        int count = 0;
        DEEDS_HIP(hipGetDeviceCount(&count));
        DEEDS_REQUIRE(device >= 0 && device < count, "device index out of range (no GPU visible?)");
        ctx->device = device;
        DEEDS_HIP(hipSetDevice(device));
        DEEDS_HIP(hipStreamCreateWithFlags(&ctx->own_stream, hipStreamNonBlocking));
        DEEDS_HIP(hipStreamCreateWithFlags(&ctx->aux_stream, hipStreamNonBlocking));
        for (hipStream_t& ss : ctx->side_stream) DEEDS_HIP(hipStreamCreateWithFlags(&ss, hipStreamNonBlocking));
        ctx->stream = ctx->own_stream;
    });
    if (rc != DEEDS_OK) {
        if (ctx->own_stream) (void)hipStreamDestroy(ctx->own_stream);
        if (ctx->aux_stream) (void)hipStreamDestroy(ctx->aux_stream);
        for (hipStream_t ss : ctx->side_stream)
            if (ss) (void)hipStreamDestroy(ss);
        delete ctx;
        return rc;
    }
    *out = ctx;
    return DEEDS_OK;
}

int deeds_ctx_destroy(deeds_ctx* ctx) {
    if (!ctx) return DEEDS_OK;
    (void)hipSetDevice(ctx->device);
    ctx->release_all();
    if (ctx->own_stream) (void)hipStreamDestroy(ctx->own_stream);
    if (ctx->aux_stream) (void)hipStreamDestroy(ctx->aux_stream);
    for (hipStream_t ss : ctx->side_stream)
        if (ss) (void)hipStreamDestroy(ss);
    delete ctx;
    return DEEDS_OK;
}

const char* deeds_ctx_last_error(const deeds_ctx* ctx) { return ctx ? ctx->err.c_str() : t_error.c_str(); }

int deeds_ctx_set_stream(deeds_ctx* ctx, void* hip_stream) {
    if (!ctx) return DEEDS_ERR_ARG;
    return guarded(ctx, [&] {
        DEEDS_HIP(hipStreamSynchronize(ctx->stream));
        ctx->stream = hip_stream ? (hipStream_t)hip_stream : ctx->own_stream;
    });
}

int deeds_ctx_synchronize(deeds_ctx* ctx) {
    if (!ctx) return DEEDS_ERR_ARG;
    return guarded(ctx, [&] {
        DEEDS_HIP(hipStreamSynchronize(ctx->stream));
        DEEDS_HIP(hipStreamSynchronize(ctx->aux_stream));
    });
}

int deeds_ctx_trim(deeds_ctx* ctx) {
    if (!ctx) return DEEDS_ERR_ARG;
    return guarded(ctx, [&] { ctx->release_all(); });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// argument checks shared by both forms
// ---------------------------------------------------------------------------
namespace {

void check_dims(int m, int n, int o) {
    DEEDS_REQUIRE(m > 0 && n > 0 && o > 0, "volume dimensions must be positive");
    DEEDS_REQUIRE((long long)m * n * o < (1LL << 31), "volume too large (more than 2^31 voxels)");
}

void check_mind(int m, int n, int o, int qs) {
    check_dims(m, n, o);
    DEEDS_REQUIRE(qs >= 1, "MIND patch radius qs must be >= 1");
    DEEDS_REQUIRE(m >= 2 * qs + 1 && n >= 2 * qs + 1 && o >= 2 * qs + 1,
                  "each dimension must be >= 2*qs+1 (box filter window)");
    DEEDS_REQUIRE(m <= mind_max_dim() && n <= mind_max_dim() && o <= mind_max_dim(),
                  "a dimension exceeds the box filter's LDS tile limit");
}

void check_grid(int m, int n, int o, int step) {
    check_dims(m, n, o);
    DEEDS_REQUIRE(step >= 1, "grid step must be >= 1");
    DEEDS_REQUIRE(m / step >= 1 && n / step >= 1 && o / step >= 1, "grid step larger than a dimension");
}

void check_labels(int hw, float quant) {
    DEEDS_REQUIRE(hw >= 0 && hw <= kMaxBPRadius, "search radius hw must be in [0, 10]");
    DEEDS_REQUIRE(quant >= 1.0f && quant == std::floor(quant), "quantisation must be an integer >= 1");
}

}  // namespace

// ---------------------------------------------------------------------------
// device-pointer forms
// ---------------------------------------------------------------------------
extern "C" {

int deeds_ctx_descriptor(deeds_ctx* ctx, uint64_t* d_mindq, const float* d_im, int m, int n, int o, int qs) {
    if (!ctx) return DEEDS_ERR_ARG;
    return guarded(ctx, [&] {
        check_mind(m, n, o, qs);
        DEEDS_REQUIRE(d_mindq && d_im, "null pointer");
        size_t sz6 = (size_t)m * n * o * 6;
        float* w0 = ctx->get<float>("mind.w0", sz6);
        float* w1 = ctx->get<float>("mind.w1", sz6);
        launch_descriptor(d_mindq, d_im, m, n, o, qs, w0, w1, ctx->stream);
    });
}

int deeds_ctx_datacost(deeds_ctx* ctx, const uint64_t* d_F, const uint64_t* d_M, float* d_cost, int m, int n, int o,
                       int step, int hw, float quant, float alpha, int randnum) {
    if (!ctx) return DEEDS_ERR_ARG;
    return guarded(ctx, [&] {
        check_grid(m, n, o, step);
        check_labels(hw, quant);
        DEEDS_REQUIRE(d_F && d_M && d_cost, "null pointer");
        CostPlan plan = datacost_plan(m, n, o, step, hw, quant, alpha, randnum);
        uint64_t* Mp = ctx->get<uint64_t>("cost.mpad", (size_t)plan.mp * plan.np * plan.op);
        launch_pad_codes(Mp, d_M, m, n, o, plan.pad, ctx->stream);
        launch_datacost(d_cost, d_F, Mp, m, n, o, step, hw, (int)quant, plan, ctx->stream);
    });
}

}  // extern "C"

namespace deeds {

size_t tree_ints_bound(int N) { return tree_device_ints(N); }

// Bytes of a tree's staging area: the schedule ints, then device scratch
// that is not copied: 3*N floats for the BP parent offsets and 2*N ints for
// the label selection of trees with a depth too wide for LDS.
size_t tree_stage_bytes(int N) {
    return tree_ints_bound(N) * sizeof(int) + 3 * (size_t)N * sizeof(float) + 2 * (size_t)N * sizeof(int);
}

// Packs a host tree's BP schedule (ordered, parents, depth_start, prog_start,
// prog, leaves, leaf_w, sel_info) into the pinned staging area and enqueues its copy into the device
// area d_stage of the same layout (tree_stage_bytes(N)).  The pinned memory
// must stay untouched until the copy has executed on `s`.
BPDevTree pack_tree(const Tree& t, void* pinned, void* d_stage, size_t* bytes) {
    size_t ints = t.ordered.size() + t.parents.size() + t.depth_start.size() + t.prog_start.size() + t.prog.size() +
                  t.leaves.size() + t.leaf_w.size() + t.sel_info.size() + t.sel_jump.size();
    size_t bound = tree_ints_bound(t.N);
    if (ints > bound) fail(DEEDS_ERR_INTERNAL, "tree schedule larger than its staging bound");
    int* hi = static_cast<int*>(pinned);
    int* di = static_cast<int*>(d_stage);
    size_t at = 0;
    BPDevTree out{};
    auto put = [&](const std::vector<int>& v) {
        int* dst = di + at;
        std::copy(v.begin(), v.end(), hi + at);
        at += v.size();
        return dst;
    };
    out.ordered = put(t.ordered);
    out.parents = put(t.parents);
    out.depth_start = put(t.depth_start);
    out.prog_start = put(t.prog_start);
    out.prog = put(t.prog);
    out.n_segments = (int)t.heads.size();
    out.leaves = put(t.leaves);
    out.leaf_w = put(t.leaf_w);
    out.n_leaves = (int)t.leaves.size();
    out.sel_info = put(t.sel_info);
    out.sel_jump = put(t.sel_jump);
    out.max_width = t.max_width;
    out.root = t.N > 0 ? t.ordered[0] : 0;
    out.off3 = reinterpret_cast<float*>(di + bound);
    out.sel_labs = di + bound + 3 * (size_t)t.N;
    *bytes = at * sizeof(int);
    return out;
}

BPDevTree stage_tree(const Tree& t, void* pinned, void* d_stage, hipStream_t s) {
    size_t bytes = 0;
    const BPDevTree out = pack_tree(t, pinned, d_stage, &bytes);
    DEEDS_HIP(hipMemcpyAsync(d_stage, pinned, bytes, hipMemcpyHostToDevice, s));
    return out;
}

// Bottom-up rounds, then the top-down selection.
void run_bp(const Tree& t, const BPDevTree& dt, float* d_cost, int16_t* d_inds, int16_t* d_comp, const float* u0,
            const float* v0, const float* w0, float* u1, float* v1, float* w1, int hw, float quant, hipStream_t s) {
    launch_bp_rounds(d_cost, d_inds, dt.off3, dt, t.N, t.round_start.data(), t.rounds, t.tail_round, u0, v0, w0, hw,
                     quant, s);
    const int first_anchor = t.maxd >= kSelJump ? t.depth_start[kSelJump] : t.N;
    launch_bp_select(d_cost, d_inds, d_comp, first_anchor, dt, t.N, t.maxd, u0, v0, w0, u1, v1, w1, hw, quant, s);
}

namespace {
size_t batch_desc_bytes(int trees) { return ((size_t)trees * sizeof(BPBatchTree) + 255) & ~(size_t)255; }
}  // namespace

size_t bp_batch_bytes(int trees, long long nodes) {
    // leaves and segment heads are disjoint sets of nodes: at most N work items
    // per tree, plus two items per tree for its spine
    return batch_desc_bytes(trees) + ((size_t)nodes + 2 * (size_t)trees) * sizeof(int2);
}

void run_bp_batch(const std::vector<BPBatchItem>& items, int hw, float quant, void* pinned, void* d_stage,
                  size_t bytes, hipStream_t s) {
    const int T = (int)items.size();
    if (T == 0) return;
    if (T == 1) {
        // one tree: the single-tree kernels take their arguments by value (no descriptor loads)
        const BPBatchItem& it = items[0];
        run_bp(*it.tree, it.dt, it.cost, it.inds, it.comp, it.u0, it.v0, it.w0, it.u1, it.v1, it.w1, hw, quant, s);
        return;
    }
    long long nodes = 0;
    int max_n = 0, width = 1, rounds = 0;
    for (const BPBatchItem& it : items) {
        nodes += it.tree->N;
        max_n = std::max(max_n, it.tree->N);
        width = std::max(width, it.tree->max_width);
        rounds = std::max(rounds, it.tree->tail_round);  // the rounds launched one by one
    }
    if (bp_batch_bytes(T, nodes) > bytes) fail(DEEDS_ERR_INTERNAL, "BP batch larger than its staging area");
    const size_t desc = batch_desc_bytes(T);
    BPBatchTree* hd = static_cast<BPBatchTree*>(pinned);
    int2* hw2 = reinterpret_cast<int2*>(static_cast<char*>(pinned) + desc);
    for (int t = 0; t < T; t++) {
        const BPBatchItem& it = items[t];
        const BPDevTree& d = it.dt;
        hd[t] = BPBatchTree{it.cost,      it.inds,    d.prog_start, d.prog,  d.leaves, d.leaf_w, d.off3,
                            d.parents,    d.ordered,  d.depth_start, d.sel_info, it.u0,  it.v0,    it.w0,
                            it.u1,        it.v1,      it.w1,        d.sel_labs, d.root, it.tree->N,
                            it.tree->maxd, d.n_segments};
    }
    // work list: the leaves of every tree, then round by round every tree's segments of that round
    std::vector<int> round_off(1, 0);
    long long at = 0;
    for (int t = 0; t < T; t++)
        for (int i = 0; i < items[t].dt.n_leaves; i++) hw2[at++] = make_int2(t, i);
    round_off.push_back((int)at);
    for (int r = 0; r < rounds; r++) {
        for (int t = 0; t < T; t++) {
            const Tree& tr = *items[t].tree;
            if (r >= tr.tail_round) continue;
            for (int h = tr.round_start[r]; h < tr.round_start[r + 1]; h++)
                hw2[at++] = make_int2(t, tr.prog_start[h]);
        }
        round_off.push_back((int)at);
    }
    // spines: ((tree, first segment), (end segment, 0)) per tree with one
    for (int t = 0; t < T; t++) {
        const Tree& tr = *items[t].tree;
        if (tr.tail_round >= tr.rounds) continue;
        hw2[at++] = make_int2(t, tr.round_start[tr.tail_round]);
        hw2[at++] = make_int2(tr.round_start[tr.rounds], 0);
    }
    round_off.push_back((int)at);
    if (at > nodes + 2LL * T) fail(DEEDS_ERR_INTERNAL, "BP work list exceeds its staging bound");
    DEEDS_HIP(hipMemcpyAsync(d_stage, pinned, desc + (size_t)at * sizeof(int2), hipMemcpyHostToDevice, s));
    launch_bp_batch(static_cast<const BPBatchTree*>(d_stage), T, max_n,
                    reinterpret_cast<const int2*>(static_cast<const char*>(d_stage) + desc), round_off, width, hw,
                    quant, s);
}

}  // namespace deeds

extern "C" {

int deeds_ctx_regularisation(deeds_ctx* ctx, float* d_cost, const float* d_u0, const float* d_v0, const float* d_w0,
                             float* d_u1, float* d_v1, float* d_w1, int hw, int step, float quant, const int* ordered,
                             const int* parents, const float* edgemst, int m, int n, int o, int16_t* d_allinds) {
    if (!ctx) return DEEDS_ERR_ARG;
    return guarded(ctx, [&] {
        check_grid(m, n, o, step);
        check_labels(hw, quant);
        DEEDS_REQUIRE(d_cost && d_u0 && d_v0 && d_w0 && d_u1 && d_v1 && d_w1 && ordered && parents && edgemst,
                      "null pointer");
        int m1 = m / step, n1 = n / step, o1 = o / step;
        Tree t;
        t.m1 = m1; t.n1 = n1; t.o1 = o1; t.N = m1 * n1 * o1;
        t.ordered.assign(ordered, ordered + t.N);
        t.parents.assign(parents, parents + t.N);
        t.edgemst.assign(edgemst, edgemst + t.N);
        for (int v : t.ordered) DEEDS_REQUIRE(v >= 0 && v < t.N, "ordered holds an out-of-range node");
        try {
            build_schedule(t);
        } catch (const std::runtime_error& e) {
            fail(DEEDS_ERR_ARG, e.what());
        }
        int len = 2 * hw + 1;
        size_t len3 = (size_t)len * len * len;
        size_t stage_bytes = tree_stage_bytes(t.N);
        void* pinned = nullptr;
        DEEDS_HIP(hipHostMalloc(&pinned, stage_bytes, hipHostMallocDefault));
        std::unique_ptr<void, hipError_t (*)(void*)> hold(pinned, hipHostFree);
        void* d_stage = ctx->buf("bp.tree", stage_bytes);
        BPDevTree dt = stage_tree(t, pinned, d_stage, ctx->stream);
        int16_t* inds = d_allinds ? d_allinds : ctx->get<int16_t>("bp.inds", (size_t)t.N * len3);
        int16_t* comp = ctx->get<int16_t>("bp.comp", (size_t)t.N * len3);
        run_bp(t, dt, d_cost, inds, comp, d_u0, d_v0, d_w0, d_u1, d_v1, d_w1, hw, quant, ctx->stream);
        // the caller's table gets the root's row as well (regularisation.h:273-274)
        if (d_allinds) launch_root_identity(d_allinds, t.ordered[0], (int)len3, ctx->stream);
        // the staging copy must complete before the pinned block is released
        DEEDS_HIP(hipStreamSynchronize(ctx->stream));
    });
}

// ---------------------------------------------------------------------------
// host-pointer forms
// ---------------------------------------------------------------------------

int deeds_descriptor(uint64_t* mindq, const float* im, int m, int n, int o, int qs) {
    return guarded(nullptr, [&] {
        DEEDS_REQUIRE(mindq && im, "null pointer");
        check_mind(m, n, o, qs);
        deeds_ctx* c = default_ctx();
        DEEDS_HIP(hipSetDevice(c->device));
        size_t sz = (size_t)m * n * o;
        float* d_im = c->get<float>("h.im", sz);
        uint64_t* d_out = c->get<uint64_t>("h.mind", sz);
        h2d(d_im, im, sz, c->stream);
        int rc = deeds_ctx_descriptor(c, d_out, d_im, m, n, o, qs);
        if (rc) fail(rc, c->err);
        d2h(mindq, d_out, sz, c->stream);
        DEEDS_HIP(hipStreamSynchronize(c->stream));
    });
}

int deeds_datacost(const uint64_t* F, const uint64_t* M, float* cost, int m, int n, int o, int step, int hw,
                   float quant, float alpha, int randnum) {
    return guarded(nullptr, [&] {
        DEEDS_REQUIRE(F && M && cost, "null pointer");
        check_grid(m, n, o, step);
        check_labels(hw, quant);
        deeds_ctx* c = default_ctx();
        DEEDS_HIP(hipSetDevice(c->device));
        size_t sz = (size_t)m * n * o;
        int len = 2 * hw + 1;
        size_t ncost = (size_t)(m / step) * (n / step) * (o / step) * len * len * len;
        uint64_t* dF = c->get<uint64_t>("h.F", sz);
        uint64_t* dM = c->get<uint64_t>("h.M", sz);
        float* dc = c->get<float>("h.cost", ncost);
        h2d(dF, F, sz, c->stream);
        h2d(dM, M, sz, c->stream);
        int rc = deeds_ctx_datacost(c, dF, dM, dc, m, n, o, step, hw, quant, alpha, randnum);
        if (rc) fail(rc, c->err);
        d2h(cost, dc, ncost, c->stream);
        DEEDS_HIP(hipStreamSynchronize(c->stream));
    });
}

int deeds_prims_graph(const float* im, int* ordered, int* parents, float* edgemst, int step, int m, int n, int o) {
    return guarded(nullptr, [&] {
        DEEDS_REQUIRE(im && ordered && parents && edgemst, "null pointer");
        check_grid(m, n, o, step);
        deeds_ctx* c = default_ctx();
        DEEDS_HIP(hipSetDevice(c->device));
        size_t sz = (size_t)m * n * o;
        int m1 = m / step, n1 = n / step, o1 = o / step;
        size_t N = (size_t)m1 * n1 * o1;
        float* d_im = c->get<float>("h.im", sz);
        float* d_ec = c->get<float>("h.ec", 6 * N);
        h2d(d_im, im, sz, c->stream);
        launch_edge_sad(d_ec, d_im, step, m, n, o, c->stream);
        std::vector<float> ec(6 * N);
        d2h(ec.data(), d_ec, 6 * N, c->stream);
        float mean = 0, sd = 0;
        image_stats(im, (long long)sz, &mean, &sd);
        DEEDS_HIP(hipStreamSynchronize(c->stream));
        Tree t;
        build_tree(ec.data(), sd, step, m1, n1, o1, t);
        std::copy(t.ordered.begin(), t.ordered.end(), ordered);
        std::copy(t.parents.begin(), t.parents.end(), parents);
        std::copy(t.edgemst.begin(), t.edgemst.end(), edgemst);
    });
}

int deeds_ctx_prims_graph(deeds_ctx* ctx, const float* d_im, int* ordered, int* parents, float* edgemst, int step,
                          int m, int n, int o) {
    if (!ctx) return DEEDS_ERR_ARG;
    return guarded(ctx, [&] {
        DEEDS_REQUIRE(d_im && ordered && parents && edgemst, "null pointer");
        check_grid(m, n, o, step);
        const long long sz = (long long)m * n * o;
        const int m1 = m / step, n1 = n / step, o1 = o / step;
        const size_t N = (size_t)m1 * n1 * o1;
        float* d_ec = ctx->get<float>("prims.ec", 6 * N);
        float* h_ec = static_cast<float*>(ctx->pinned("prims.ec", 6 * N * sizeof(float)));
        launch_edge_sad(d_ec, d_im, step, m, n, o, ctx->stream);
        d2h(h_ec, d_ec, 6 * N, ctx->stream);
        hipEvent_t ev;
        DEEDS_HIP(hipEventCreate(&ev));
        std::unique_ptr<std::remove_pointer<hipEvent_t>::type, hipError_t (*)(hipEvent_t)> hold(ev, hipEventDestroy);
        DEEDS_HIP(hipEventRecord(ev, ctx->stream));
        // Prim's on the edge costs (integer SADs only) while the GPU computes the statistics, as the
        // registration driver does; the caller gets no BP schedule
        Tree t;
        auto shape = std::async(std::launch::async, [&] {
            DEEDS_HIP(hipEventSynchronize(ev));
            return build_tree_shape(h_ec, step, m1, n1, o1, t);
        });
        float mean = 0, sd = 0;
        StatsScratch sc;
        const size_t bytes = StatsScratch::bytes(sz);
        sc.dev = ctx->buf("prims.stats", bytes);
        sc.host = ctx->pinned("prims.stats", bytes);
        bool shaped = false, same_order = false;
        try {
            image_stats_device(d_im, sz, sc, ctx->stream, &mean, &sd);
            if (shape.wait_for(std::chrono::seconds(0)) == std::future_status::ready) {
                shaped = shape.get();  // done, or declined (non-integer SADs)
                if (shaped) same_order = weights_distinguish_costs(h_ec, step, m1, n1, o1, sd);
            } else {
                // Prim's on the costs is running: the order check runs beside it
                same_order = weights_distinguish_costs(h_ec, step, m1, n1, o1, sd);
                shaped = shape.get();
            }
        } catch (...) {
            if (shape.valid()) shape.wait();
            throw;
        }
        if (shaped && same_order)
            finish_tree(h_ec, sd, step, t, false);
        else
            build_tree(h_ec, sd, step, m1, n1, o1, t, false);
        std::copy(t.ordered.begin(), t.ordered.end(), ordered);
        std::copy(t.parents.begin(), t.parents.end(), parents);
        std::copy(t.edgemst.begin(), t.edgemst.end(), edgemst);
    });
}

int deeds_ctx_image_stats(deeds_ctx* ctx, const float* d_im, long long count, float* mean, float* sd) {
    if (!ctx) return DEEDS_ERR_ARG;
    return guarded(ctx, [&] {
        DEEDS_REQUIRE(d_im && mean && sd, "null pointer");
        DEEDS_REQUIRE(count > 0, "count must be positive");
        StatsScratch sc;
        const size_t bytes = StatsScratch::bytes(count);
        sc.dev = ctx->buf("stats.dev", bytes);
        DEEDS_HIP(hipHostMalloc(&sc.host, bytes, hipHostMallocDefault));
        std::unique_ptr<void, hipError_t (*)(void*)> hold(sc.host, hipHostFree);
        image_stats_device(d_im, count, sc, ctx->stream, mean, sd);
    });
}

int deeds_image_stats(const float* im, long long count, float* mean, float* sd) {
    return guarded(nullptr, [&] {
        DEEDS_REQUIRE(im && mean && sd, "null pointer");
        DEEDS_REQUIRE(count > 0, "count must be positive");
        image_stats(im, count, mean, sd);
    });
}

int deeds_regularisation(float* cost, const float* u0, const float* v0, const float* w0, float* u1, float* v1,
                         float* w1, int hw, int step, float quant, const int* ordered, const int* parents,
                         const float* edgemst, int m, int n, int o) {
    return guarded(nullptr, [&] {
        DEEDS_REQUIRE(cost && u0 && v0 && w0 && u1 && v1 && w1 && ordered && parents && edgemst, "null pointer");
        check_grid(m, n, o, step);
        check_labels(hw, quant);
        deeds_ctx* c = default_ctx();
        DEEDS_HIP(hipSetDevice(c->device));
        int len = 2 * hw + 1;
        size_t N = (size_t)(m / step) * (n / step) * (o / step);
        size_t ncost = N * len * len * len;
        float* dc = c->get<float>("h.cost", ncost);
        float* df = c->get<float>("h.fields", 6 * N);
        h2d(dc, cost, ncost, c->stream);
        h2d(df, u0, N, c->stream);
        h2d(df + N, v0, N, c->stream);
        h2d(df + 2 * N, w0, N, c->stream);
        int rc = deeds_ctx_regularisation(c, dc, df, df + N, df + 2 * N, df + 3 * N, df + 4 * N, df + 5 * N, hw, step,
                                          quant, ordered, parents, edgemst, m, n, o, nullptr);
        if (rc) fail(rc, c->err);
        d2h(u1, df + 3 * N, N, c->stream);
        d2h(v1, df + 4 * N, N, c->stream);
        d2h(w1, df + 5 * N, N, c->stream);
        DEEDS_HIP(hipStreamSynchronize(c->stream));
    });
}

// ---------------------------------------------------------------------------
// field operations: device forms, and the host forms over them
// ---------------------------------------------------------------------------

namespace {
const float kIdentityX[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};

// Logging SSD of dataCostD.h:300-312 (mean squared difference to im1b before
// and after the warp), reduced on the context stream; waits for the result.
void warp_ssd(deeds_ctx* ctx, const float* d_im1b, const float* d_warped, const float* d_before, size_t sz,
              float* ssd01) {
    double* acc = ctx->get<double>("warp.ssd", 2);
    launch_ssd(acc, ctx->get<double>("red.part", kRedPartDoubles), d_im1b, d_warped, d_before, (long long)sz,
               ctx->stream);
    double h[2];
    d2h(h, acc, 2, ctx->stream);
    DEEDS_HIP(hipStreamSynchronize(ctx->stream));
    ssd01[0] = (float)(h[0] / (double)sz);
    ssd01[1] = (float)(h[1] / (double)sz);
}
}  // namespace

int deeds_ctx_upsample_deformations(deeds_ctx* ctx, float* d_u1, float* d_v1, float* d_w1, const float* d_u0,
                                    const float* d_v0, const float* d_w0, int m, int n, int o, int m2, int n2,
                                    int o2) {
    if (!ctx) return DEEDS_ERR_ARG;
    return guarded(ctx, [&] {
        DEEDS_REQUIRE(d_u1 && d_v1 && d_w1 && d_u0 && d_v0 && d_w0, "null pointer");
        check_dims(m, n, o);
        check_dims(m2, n2, o2);
        launch_upsample(d_u1, d_v1, d_w1, d_u0, d_v0, d_w0, m, n, o, m2, n2, o2, ctx->stream);
    });
}

int deeds_ctx_consistent_mapping(deeds_ctx* ctx, float* d_u, float* d_v, float* d_w, float* d_u2, float* d_v2,
                                 float* d_w2, int m, int n, int o, int factor) {
    if (!ctx) return DEEDS_ERR_ARG;
    return guarded(ctx, [&] {
        DEEDS_REQUIRE(d_u && d_v && d_w && d_u2 && d_v2 && d_w2, "null pointer");
        check_dims(m, n, o);
        DEEDS_REQUIRE(factor != 0, "factor must be nonzero");
        const size_t sz = (size_t)m * n * o;
        float* f[6] = {d_u, d_v, d_w, d_u2, d_v2, d_w2};
        launch_consistent(f, ctx->get<float>("cm.tmp", 6 * sz), m, n, o, factor, ctx->stream);
    });
}

int deeds_ctx_jacobian(deeds_ctx* ctx, const float* d_u, const float* d_v, const float* d_w, int m, int n, int o,
                       int factor, float* std_out, float* neg_frac_out) {
    if (!ctx) return DEEDS_ERR_ARG;
    return guarded(ctx, [&] {
        DEEDS_REQUIRE(d_u && d_v && d_w, "null pointer");
        check_dims(m, n, o);
        DEEDS_REQUIRE(factor != 0, "factor must be nonzero");
        const size_t sz = (size_t)m * n * o;
        double* acc = ctx->get<double>("jac.acc", 4);
        launch_jacobian(acc, ctx->get<double>("red.part", kRedPartDoubles), d_u, d_v, d_w, m, n, o, factor,
                        ctx->stream);
        double h[3];
        d2h(h, acc, 3, ctx->stream);
        DEEDS_HIP(hipStreamSynchronize(ctx->stream));
        const double N = (double)sz;
        const double mean = h[0] / N;
        const double var = sz > 1 ? (h[1] - N * mean * mean) / (N - 1) : 0.0;
        if (std_out) *std_out = (float)std::sqrt(std::max(var, 0.0));
        if (neg_frac_out) *neg_frac_out = (float)(h[2] / N);
    });
}

int deeds_ctx_warp_affine(deeds_ctx* ctx, float* d_out, const float* d_in, const float* d_im1b, const float* X,
                          const float* d_u, const float* d_v, const float* d_w, int m, int n, int o, float* ssd01) {
    if (!ctx) return DEEDS_ERR_ARG;
    return guarded(ctx, [&] {
        DEEDS_REQUIRE(d_out && d_in && d_u && d_v && d_w, "null pointer");
        DEEDS_REQUIRE(d_out != d_in, "out must not alias in");
        check_dims(m, n, o);
        const Affine A = affine_from_X(X ? X : kIdentityX);
        launch_warp_affine(d_out, d_in, A, d_u, d_v, d_w, m, n, o, ctx->stream);
        if (ssd01 && d_im1b) warp_ssd(ctx, d_im1b, d_out, d_in, (size_t)m * n * o, ssd01);
    });
}

int deeds_ctx_warp_image(deeds_ctx* ctx, float* d_out, const float* d_im1, const float* d_im1b, const float* d_u,
                         const float* d_v, const float* d_w, int m, int n, int o, float* ssd01) {
    if (!ctx) return DEEDS_ERR_ARG;
    return guarded(ctx, [&] {
        DEEDS_REQUIRE(d_out && d_im1 && d_u && d_v && d_w, "null pointer");
        DEEDS_REQUIRE(d_out != d_im1, "out must not alias im1");
        check_dims(m, n, o);
        launch_interp3(d_out, d_im1, d_u, d_v, d_w, m, n, o, m, n, o, 1, ctx->stream);
        if (ssd01 && d_im1b) warp_ssd(ctx, d_im1b, d_out, d_im1, (size_t)m * n * o, ssd01);
    });
}

int deeds_ctx_warp_affine_s(deeds_ctx* ctx, int16_t* d_out, const int16_t* d_in, const float* X, const float* d_u,
                            const float* d_v, const float* d_w, int m, int n, int o) {
    if (!ctx) return DEEDS_ERR_ARG;
    return guarded(ctx, [&] {
        DEEDS_REQUIRE(d_out && d_in && d_u && d_v && d_w, "null pointer");
        DEEDS_REQUIRE(d_out != d_in, "out must not alias in");
        check_dims(m, n, o);
        const Affine A = affine_from_X(X ? X : kIdentityX);
        launch_warp_affine_s(d_out, d_in, A, d_u, d_v, d_w, m, n, o, ctx->stream);
    });
}

int deeds_upsample_deformations(float* u1, float* v1, float* w1, const float* u0, const float* v0, const float* w0,
                                int m, int n, int o, int m2, int n2, int o2) {
    return guarded(nullptr, [&] {
        DEEDS_REQUIRE(u1 && v1 && w1 && u0 && v0 && w0, "null pointer");
        check_dims(m, n, o);
        check_dims(m2, n2, o2);
        deeds_ctx* c = default_ctx();
        DEEDS_HIP(hipSetDevice(c->device));
        size_t sz = (size_t)m * n * o, sz2 = (size_t)m2 * n2 * o2;
        float* din = c->get<float>("h.up.in", 3 * sz2);
        float* dout = c->get<float>("h.up.out", 3 * sz);
        h2d(din, u0, sz2, c->stream);
        h2d(din + sz2, v0, sz2, c->stream);
        h2d(din + 2 * sz2, w0, sz2, c->stream);
        int rc = deeds_ctx_upsample_deformations(c, dout, dout + sz, dout + 2 * sz, din, din + sz2, din + 2 * sz2, m,
                                                 n, o, m2, n2, o2);
        if (rc) fail(rc, c->err);
        d2h(u1, dout, sz, c->stream);
        d2h(v1, dout + sz, sz, c->stream);
        d2h(w1, dout + 2 * sz, sz, c->stream);
        DEEDS_HIP(hipStreamSynchronize(c->stream));
    });
}

int deeds_consistent_mapping(float* u, float* v, float* w, float* u2, float* v2, float* w2, int m, int n, int o,
                             int factor) {
    return guarded(nullptr, [&] {
        DEEDS_REQUIRE(u && v && w && u2 && v2 && w2, "null pointer");
        check_dims(m, n, o);
        DEEDS_REQUIRE(factor != 0, "factor must be nonzero");
        deeds_ctx* c = default_ctx();
        DEEDS_HIP(hipSetDevice(c->device));
        size_t sz = (size_t)m * n * o;
        float* df = c->get<float>("h.cm", 6 * sz);
        float* hf[6] = {u, v, w, u2, v2, w2};
        float* f[6];
        for (int i = 0; i < 6; i++) {
            f[i] = df + i * sz;
            h2d(f[i], hf[i], sz, c->stream);
        }
        int rc = deeds_ctx_consistent_mapping(c, f[0], f[1], f[2], f[3], f[4], f[5], m, n, o, factor);
        if (rc) fail(rc, c->err);
        for (int i = 0; i < 6; i++) d2h(hf[i], f[i], sz, c->stream);
        DEEDS_HIP(hipStreamSynchronize(c->stream));
    });
}

int deeds_jacobian(const float* u, const float* v, const float* w, int m, int n, int o, int factor, float* std_out,
                   float* neg_frac_out) {
    return guarded(nullptr, [&] {
        DEEDS_REQUIRE(u && v && w, "null pointer");
        check_dims(m, n, o);
        DEEDS_REQUIRE(factor != 0, "factor must be nonzero");
        deeds_ctx* c = default_ctx();
        DEEDS_HIP(hipSetDevice(c->device));
        size_t sz = (size_t)m * n * o;
        float* df = c->get<float>("h.jac", 3 * sz);
        h2d(df, u, sz, c->stream);
        h2d(df + sz, v, sz, c->stream);
        h2d(df + 2 * sz, w, sz, c->stream);
        int rc = deeds_ctx_jacobian(c, df, df + sz, df + 2 * sz, m, n, o, factor, std_out, neg_frac_out);
        if (rc) fail(rc, c->err);
    });
}

int deeds_warp_affine(float* out, const float* in, const float* im1b, const float* X, const float* u, const float* v,
                      const float* w, int m, int n, int o, float* ssd01) {
    return guarded(nullptr, [&] {
        DEEDS_REQUIRE(out && in && u && v && w, "null pointer");
        check_dims(m, n, o);
        deeds_ctx* c = default_ctx();
        DEEDS_HIP(hipSetDevice(c->device));
        size_t sz = (size_t)m * n * o;
        float* d = c->get<float>("h.warp", 6 * sz);
        h2d(d, in, sz, c->stream);
        h2d(d + sz, u, sz, c->stream);
        h2d(d + 2 * sz, v, sz, c->stream);
        h2d(d + 3 * sz, w, sz, c->stream);
        const bool ssd = ssd01 && im1b;
        if (ssd) h2d(d + 5 * sz, im1b, sz, c->stream);
        int rc = deeds_ctx_warp_affine(c, d + 4 * sz, d, ssd ? d + 5 * sz : nullptr, X, d + sz, d + 2 * sz, d + 3 * sz,
                                       m, n, o, ssd ? ssd01 : nullptr);
        if (rc) fail(rc, c->err);
        d2h(out, d + 4 * sz, sz, c->stream);
        DEEDS_HIP(hipStreamSynchronize(c->stream));
    });
}

int deeds_warp_image(float* out, const float* im1, const float* im1b, const float* u, const float* v, const float* w,
                     int m, int n, int o, float* ssd01) {
    return guarded(nullptr, [&] {
        DEEDS_REQUIRE(out && im1 && u && v && w, "null pointer");
        check_dims(m, n, o);
        deeds_ctx* c = default_ctx();
        DEEDS_HIP(hipSetDevice(c->device));
        size_t sz = (size_t)m * n * o;
        float* d = c->get<float>("h.warp", 6 * sz);
        h2d(d, im1, sz, c->stream);
        h2d(d + sz, u, sz, c->stream);
        h2d(d + 2 * sz, v, sz, c->stream);
        h2d(d + 3 * sz, w, sz, c->stream);
        const bool ssd = ssd01 && im1b;
        if (ssd) h2d(d + 5 * sz, im1b, sz, c->stream);
        int rc = deeds_ctx_warp_image(c, d + 4 * sz, d, ssd ? d + 5 * sz : nullptr, d + sz, d + 2 * sz, d + 3 * sz, m,
                                      n, o, ssd ? ssd01 : nullptr);
        if (rc) fail(rc, c->err);
        d2h(out, d + 4 * sz, sz, c->stream);
        DEEDS_HIP(hipStreamSynchronize(c->stream));
    });
}

int deeds_warp_affine_s(int16_t* out, const int16_t* in, const float* X, const float* u, const float* v,
                        const float* w, int m, int n, int o) {
    return guarded(nullptr, [&] {
        DEEDS_REQUIRE(out && in && u && v && w, "null pointer");
        check_dims(m, n, o);
        deeds_ctx* c = default_ctx();
        DEEDS_HIP(hipSetDevice(c->device));
        size_t sz = (size_t)m * n * o;
        float* d = c->get<float>("h.warp", 3 * sz);
        int16_t* s = c->get<int16_t>("h.seg", 2 * sz);
        h2d(s, in, sz, c->stream);
        h2d(d, u, sz, c->stream);
        h2d(d + sz, v, sz, c->stream);
        h2d(d + 2 * sz, w, sz, c->stream);
        int rc = deeds_ctx_warp_affine_s(c, s + sz, s, X, d, d + sz, d + 2 * sz, m, n, o);
        if (rc) fail(rc, c->err);
        d2h(out, s + sz, sz, c->stream);
        DEEDS_HIP(hipStreamSynchronize(c->stream));
    });
}

}  // extern "C"
