// C ABI of the differentiable soft variant (declared in include/deeds_hip.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <stdexcept>

#include "kernels.h"
#include "mst_host.h"
#include "soft.h"

namespace deeds {
namespace {

void check_real(int real) {
    DEEDS_REQUIRE(real == DEEDS_REAL_F32 || real == DEEDS_REAL_F64, "real must be DEEDS_REAL_F32 or DEEDS_REAL_F64");
}

size_t rsize(int real) { return real == DEEDS_REAL_F64 ? sizeof(double) : sizeof(float); }

// Feature arrays are read twelve values at a time with 16-byte vector loads.
void check_aligned(const void* p) {
    DEEDS_REQUIRE(((uintptr_t)p & 15) == 0, "feature arrays must be 16-byte aligned");
}

void check_soft_dims(int m, int n, int o, int qs) {
    DEEDS_REQUIRE(m > 0 && n > 0 && o > 0, "dimensions must be positive");
    DEEDS_REQUIRE((long long)m * n * o <= (1LL << 31) / 12, "volume too large for the soft features");
    DEEDS_REQUIRE(qs >= 1, "qs must be >= 1");
}

void check_soft_grid(int m, int n, int o, int step, int hw, float quant) {
    DEEDS_REQUIRE(m > 0 && n > 0 && o > 0, "dimensions must be positive");
    DEEDS_REQUIRE(step >= 1 && m / step >= 1 && n / step >= 1 && o / step >= 1,
                  "grid step must leave at least one control point per axis");
    // the cost kernels keep one thread per block sample (at most 256: steps up to 11)
    const CostPlan plan = datacost_plan(m, n, o, step, hw, quant, 1.0f, 1);
    DEEDS_REQUIRE(plan.samples <= 256, "grid step too large for the soft cost (at most 256 samples per block)");
    DEEDS_REQUIRE(hw >= 0, "label radius must be >= 0");
    DEEDS_REQUIRE(quant >= 1.0f && (float)(int)quant == quant, "quant must be an integer >= 1");
    DEEDS_REQUIRE((long long)m * n * o <= (1LL << 31) / 12, "volume too large for the soft cost");
}

// The last tree staged on a context: the caller's arrays (the cache key) and
// the re-rooted schedule whose arrays live in the ctx buffer "soft.tree".  The
// forward and backward BP of one step pass the same tree, so the backward
// finds it here.
struct SoftTreeCache {
    std::vector<int> ordered, parents;
    std::vector<float> edgemst;
    SoftTree t;
};

const SoftTree& stage_soft_tree(deeds_ctx* ctx, const int* ordered, const int* parents, const float* edgemst,
                                int N) {
    DEEDS_REQUIRE(ordered && parents && edgemst, "null tree pointer");
    auto* cache = static_cast<SoftTreeCache*>(ctx->soft_cache.get());
    if (cache && (int)cache->ordered.size() == N && std::equal(ordered, ordered + N, cache->ordered.begin()) &&
        std::equal(parents, parents + N, cache->parents.begin()) &&
        std::memcmp(edgemst, cache->edgemst.data(), (size_t)N * sizeof(float)) == 0)
        return cache->t;
    ctx->soft_cache.reset();
    SoftSchedule t;
    try {
        soft_schedule(ordered, parents, edgemst, N, kSoftWindow, t);  // validates the tree
    } catch (const std::runtime_error& e) {
        fail(DEEDS_ERR_ARG, e.what());
    }
    const RoundsHost& up = t.up;
    const RoundsHost& down = t.down;
    const size_t ints = (size_t)N + N + (N + 1) + (N > 0 ? N - 1 : 0) + up.seg_start.size() + up.nodes.size() +
                        down.seg_start.size() + down.nodes.size();
    const size_t bytes = ints * sizeof(int) + (size_t)N * sizeof(float);
    // every soft entry point drains its stream before returning, so the
    // previous copy out of the pinned block has completed
    int* h = static_cast<int*>(ctx->pinned("soft.tree", bytes));
    int* d = static_cast<int*>(ctx->buf("soft.tree", bytes));
    std::shared_ptr<SoftTreeCache> c = std::make_shared<SoftTreeCache>();
    size_t at = 0;
    auto put = [&](const std::vector<int>& v) {
        const int* p = d + at;
        std::copy(v.begin(), v.end(), h + at);
        at += v.size();
        return p;
    };
    SoftTree& st = c->t;
    st.N = N;
    st.maxd = t.maxd;
    st.depth_start = t.depth_start;
    st.d_ordered = put(t.ordered);
    st.d_parents = put(t.parents);
    st.d_child_start = put(t.child_start);
    st.d_child_list = put(t.child_list);
    st.up.d_seg_start = put(up.seg_start);
    st.up.d_nodes = put(up.nodes);
    st.up.round_start = up.round_start;
    st.down.d_seg_start = put(down.seg_start);
    st.down.d_nodes = put(down.nodes);
    st.down.round_start = down.round_start;
    DEEDS_REQUIRE(at == ints, "tree arrays have unexpected sizes");
    DEEDS_REQUIRE((int)up.nodes.size() == N && (int)down.nodes.size() == (N > 0 ? N - 1 : 0),
                  "soft sweep schedule does not cover the tree");
    std::copy(t.edgemst.begin(), t.edgemst.end(), reinterpret_cast<float*>(h + at));
    st.d_edgemst = reinterpret_cast<const float*>(d + at);
    DEEDS_HIP(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, ctx->stream));
    c->ordered.assign(ordered, ordered + N);
    c->parents.assign(parents, parents + N);
    c->edgemst.assign(edgemst, edgemst + N);
    ctx->soft_cache = c;
    return c->t;
}

void check_bp_args(int m, int n, int o, int step, int hw, float quant, float T, int real) {
    check_real(real);
    check_soft_grid(m, n, o, step, hw, quant);
    DEEDS_REQUIRE(T > 0.0f, "temperature must be positive");
    DEEDS_REQUIRE(hw <= kSoftMaxRadius, "soft BP label radius must be <= " + std::to_string(kSoftMaxRadius));
    DEEDS_REQUIRE(soft_bp_smem(2 * hw + 1, rsize(real)) <= (size_t)kMaxLds,
                  "label radius too large for the soft BP (its workgroup keeps six label cubes in LDS)");
}

}  // namespace
}  // namespace deeds

using namespace deeds;

extern "C" {

size_t deeds_soft_saved_count(int N, int hw) {
    const size_t L = 2 * (size_t)(hw < 0 ? 0 : hw) + 1;
    return 3 * (size_t)(N < 0 ? 0 : N) * L * L * L;
}

int deeds_ctx_soft_features(deeds_ctx* ctx, int real, void* d_phi, const void* d_im, int m, int n, int o, int qs) {
    if (!ctx) return DEEDS_ERR_ARG;
    return guarded(ctx, [&] {
        check_real(real);
        check_soft_dims(m, n, o, qs);
        DEEDS_REQUIRE(d_phi && d_im, "null pointer");
        const size_t sz = (size_t)m * n * o;
        if (real == DEEDS_REAL_F32) {
            float* w0 = ctx->get<float>("soft.f.w0", 6 * sz);
            float* w1 = ctx->get<float>("soft.f.w1", 6 * sz);
            soft_features<float>(static_cast<float*>(d_phi), static_cast<const float*>(d_im), m, n, o, qs, w0, w1,
                                 ctx->stream);
        } else {
            double* w0 = ctx->get<double>("soft.d.w0", 6 * sz);
            double* w1 = ctx->get<double>("soft.d.w1", 6 * sz);
            soft_features<double>(static_cast<double*>(d_phi), static_cast<const double*>(d_im), m, n, o, qs, w0, w1,
                                  ctx->stream);
        }
    });
}

int deeds_ctx_soft_features_backward(deeds_ctx* ctx, int real, void* d_gim, const void* d_gphi, const void* d_im,
                                     int m, int n, int o, int qs) {
    if (!ctx) return DEEDS_ERR_ARG;
    return guarded(ctx, [&] {
        check_real(real);
        check_soft_dims(m, n, o, qs);
        DEEDS_REQUIRE(d_gim && d_gphi && d_im, "null pointer");
        const size_t sz = (size_t)m * n * o;
        if (real == DEEDS_REAL_F32) {
            float* w0 = ctx->get<float>("soft.f.w0", 6 * sz);
            float* w1 = ctx->get<float>("soft.f.w1", 6 * sz);
            float* w2 = ctx->get<float>("soft.f.w2", 12 * sz);
            soft_features_backward<float>(static_cast<float*>(d_gim), static_cast<const float*>(d_gphi),
                                          static_cast<const float*>(d_im), m, n, o, qs, w0, w1, w2, ctx->stream);
        } else {
            double* w0 = ctx->get<double>("soft.d.w0", 6 * sz);
            double* w1 = ctx->get<double>("soft.d.w1", 6 * sz);
            double* w2 = ctx->get<double>("soft.d.w2", 12 * sz);
            soft_features_backward<double>(static_cast<double*>(d_gim), static_cast<const double*>(d_gphi),
                                           static_cast<const double*>(d_im), m, n, o, qs, w0, w1, w2, ctx->stream);
        }
    });
}

int deeds_ctx_soft_cost(deeds_ctx* ctx, int real, void* d_cost, const void* d_fixed_phi, const void* d_moving_phi,
                        int m, int n, int o, int step, int hw, float quant, float alpha) {
    if (!ctx) return DEEDS_ERR_ARG;
    return guarded(ctx, [&] {
        check_real(real);
        check_soft_grid(m, n, o, step, hw, quant);
        DEEDS_REQUIRE(d_cost && d_fixed_phi && d_moving_phi, "null pointer");
        check_aligned(d_fixed_phi);
        check_aligned(d_moving_phi);
        if (real == DEEDS_REAL_F32)
            soft_cost<float>(static_cast<float*>(d_cost), static_cast<const float*>(d_fixed_phi),
                             static_cast<const float*>(d_moving_phi), m, n, o, step, hw, quant, alpha, ctx->stream);
        else
            soft_cost<double>(static_cast<double*>(d_cost), static_cast<const double*>(d_fixed_phi),
                              static_cast<const double*>(d_moving_phi), m, n, o, step, hw, quant, alpha, ctx->stream);
    });
}

int deeds_ctx_soft_cost_backward(deeds_ctx* ctx, int real, void* d_gfixed, void* d_gmoving, const void* d_gcost,
                                 const void* d_fixed_phi, const void* d_moving_phi, int m, int n, int o, int step,
                                 int hw, float quant, float alpha) {
    if (!ctx) return DEEDS_ERR_ARG;
    return guarded(ctx, [&] {
        check_real(real);
        check_soft_grid(m, n, o, step, hw, quant);
        DEEDS_REQUIRE(d_gfixed && d_gmoving && d_gcost && d_fixed_phi && d_moving_phi, "null pointer");
        check_aligned(d_fixed_phi);
        check_aligned(d_moving_phi);
        check_aligned(d_gmoving);
        const size_t L = 2 * (size_t)hw + 1;
        DEEDS_REQUIRE(L * L * L * rsize(real) <= (size_t)kMaxLds, "label radius too large for the soft cost backward");
        const size_t part = soft_cost_part_count(m, n, o, hw, quant);
        const size_t gt = soft_cost_grad_t_count(m, n, o, step, hw);
        if (real == DEEDS_REAL_F32)
            soft_cost_backward<float>(static_cast<float*>(d_gfixed), static_cast<float*>(d_gmoving),
                                      static_cast<const float*>(d_gcost), static_cast<const float*>(d_fixed_phi),
                                      static_cast<const float*>(d_moving_phi), m, n, o, step, hw, quant, alpha,
                                      ctx->get<float>("soft.cost.part", part), ctx->get<float>("soft.cost.gt", gt),
                                      ctx->stream);
        else
            soft_cost_backward<double>(static_cast<double*>(d_gfixed), static_cast<double*>(d_gmoving),
                                       static_cast<const double*>(d_gcost), static_cast<const double*>(d_fixed_phi),
                                       static_cast<const double*>(d_moving_phi), m, n, o, step, hw, quant, alpha,
                                       ctx->get<double>("soft.cost.part", part), ctx->get<double>("soft.cost.gt", gt),
                                       ctx->stream);
    });
}

int deeds_ctx_soft_bp(deeds_ctx* ctx, int real, void* d_u1, void* d_saved, const void* d_cost, const void* d_u0,
                      const int* ordered, const int* parents, const float* edgemst, int m, int n, int o, int step,
                      int hw, float quant, float T) {
    if (!ctx) return DEEDS_ERR_ARG;
    return guarded(ctx, [&] {
        check_bp_args(m, n, o, step, hw, quant, T, real);
        DEEDS_REQUIRE(d_u1 && d_saved && d_cost && d_u0, "null pointer");
        const int N = (m / step) * (n / step) * (o / step);
        const SoftTree& st = stage_soft_tree(ctx, ordered, parents, edgemst, N);
        if (real == DEEDS_REAL_F32)
            soft_bp_forward<float>(st, static_cast<const float*>(d_cost), static_cast<const float*>(d_u0),
                                   static_cast<float*>(d_u1), static_cast<float*>(d_saved), hw, quant, T,
                                   ctx->stream);
        else
            soft_bp_forward<double>(st, static_cast<const double*>(d_cost), static_cast<const double*>(d_u0),
                                    static_cast<double*>(d_u1), static_cast<double*>(d_saved), hw, quant, T,
                                    ctx->stream);
        // drained before returning: the next staging may rewrite the pinned tree block
        DEEDS_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int deeds_ctx_soft_bp_backward(deeds_ctx* ctx, int real, void* d_gcost, const void* d_gu1, const void* d_saved,
                               const void* d_cost, const void* d_u0, const int* ordered, const int* parents,
                               const float* edgemst, int m, int n, int o, int step, int hw, float quant, float T) {
    if (!ctx) return DEEDS_ERR_ARG;
    return guarded(ctx, [&] {
        check_bp_args(m, n, o, step, hw, quant, T, real);
        DEEDS_REQUIRE(d_gcost && d_gu1 && d_saved && d_cost && d_u0, "null pointer");
        const int N = (m / step) * (n / step) * (o / step);
        const size_t L = 2 * (size_t)hw + 1;
        const SoftTree& st = stage_soft_tree(ctx, ordered, parents, edgemst, N);
        const size_t scratch = 9 * (size_t)N * L * L * L;
        if (real == DEEDS_REAL_F32)
            soft_bp_backward<float>(st, static_cast<const float*>(d_cost), static_cast<const float*>(d_u0),
                                    static_cast<const float*>(d_saved), static_cast<const float*>(d_gu1),
                                    static_cast<float*>(d_gcost), ctx->get<float>("soft.bp.scratch", scratch), hw,
                                    quant, T, ctx->stream);
        else
            soft_bp_backward<double>(st, static_cast<const double*>(d_cost), static_cast<const double*>(d_u0),
                                     static_cast<const double*>(d_saved), static_cast<const double*>(d_gu1),
                                     static_cast<double*>(d_gcost), ctx->get<double>("soft.bp.scratch", scratch), hw,
                                     quant, T, ctx->stream);
        DEEDS_HIP(hipStreamSynchronize(ctx->stream));
    });
}

}  // extern "C"
