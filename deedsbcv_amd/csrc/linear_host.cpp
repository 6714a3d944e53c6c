// Affine (or rigid) pre-alignment -- the linearBCV program's estimation
// (reference: linearBCV.cpp:144-263, affineLTS1.h, QRsolve.h).
//
// Per level: warp both images by the current estimate and its inverse
// (GPU), MIND codes and the Hamming cost volumes with alpha 1 (the
// registration kernels), smoothing and per-node minima (linear.hip); then on
// the host, the node centres whose smoothed minimum lies below the quantile
// give point pairs to a least-trimmed-squares fit, composed with the
// previous estimate.  The host fit keeps the reference's float / double
// arithmetic and summation order, so the matrix is bit-identical.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "common.h"
#include "kernels.h"

namespace deeds {

namespace {

// QRsolve.h:1-7: float sum of the squares, each exact in double (pow(float, 2)).
float vnorm(const float* v, int len) {
    float s = 0.0f;
    for (int i = 0; i < len; i++) s = (float)((double)s + (double)v[i] * (double)v[i]);
    return std::sqrt(s);
}

// QRsolve.h:9-15
float vdot(const float* a, const float* b, int len) {
    float d = 0.0f;
    for (int i = 0; i < len; i++) d += a[i] * b[i];
    return d;
}

// QRsolve.h:17-70: modified Gram-Schmidt QR of the len x 4 matrix A (column c
// at A + c*len), back substitution for each right-hand side b + j*len.
void qr_solve(float* X, const float* A, const float* b, int len, int len2) {
    std::vector<float> Q(4 * (size_t)len);
    float* q0 = Q.data();
    float* q1 = q0 + len;
    float* q2 = q1 + len;
    float* q3 = q2 + len;
    const float* a1 = A + len;
    const float* a2 = A + 2 * len;
    const float* a3 = A + 3 * len;
    const float r11 = vnorm(A, len);
    for (int i = 0; i < len; i++) q0[i] = A[i] / r11;
    const float r12 = vdot(q0, a1, len), r13 = vdot(q0, a2, len), r14 = vdot(q0, a3, len);
    for (int i = 0; i < len; i++) {
        q1[i] = a1[i] - q0[i] * r12;
        q2[i] = a2[i] - q0[i] * r13;
        q3[i] = a3[i] - q0[i] * r14;
    }
    const float r22 = vnorm(q1, len);
    for (int i = 0; i < len; i++) q1[i] = q1[i] / r22;
    const float r23 = vdot(q1, a2, len), r24 = vdot(q1, a3, len);
    for (int i = 0; i < len; i++) {
        q2[i] = a2[i] - q0[i] * r13 - q1[i] * r23;
        q3[i] = a3[i] - q0[i] * r14 - q1[i] * r24;
    }
    const float r33 = vnorm(q2, len);
    for (int i = 0; i < len; i++) q2[i] = q2[i] / r33;
    const float r34 = vdot(q2, a3, len);
    for (int i = 0; i < len; i++) q3[i] = a3[i] - q0[i] * r14 - q1[i] * r24 - q2[i] * r34;
    const float r44 = vnorm(q3, len);
    for (int i = 0; i < len; i++) q3[i] = q3[i] / r44;
    for (int j = 0; j < len2; j++) {
        const float* bj = b + (size_t)j * len;
        const float d1 = vdot(q0, bj, len), d2 = vdot(q1, bj, len), d3 = vdot(q2, bj, len), d4 = vdot(q3, bj, len);
        const float x4 = d4 / r44;
        const float x3 = (d3 - r34 * x4) / r33;
        const float x2 = (d2 - r23 * x3 - r24 * x4) / r22;
        const float x1 = (d1 - r12 * x2 - r13 * x3 - r14 * x4) / r11;
        X[4 * j + 0] = x1;
        X[4 * j + 1] = x2;
        X[4 * j + 2] = x3;
        X[4 * j + 3] = x4;
    }
}

// QRsolve.h:72-122: one-sided Jacobi on the columns of the 3x3 A (four sweeps
// of the pairs (0,1), (0,2), (1,2)), then unit columns.
void jacobi_svd3(const float* A, float* U, float* V) {
    std::copy(A, A + 9, U);
    std::fill(V, V + 9, 0.0f);
    V[0] = V[4] = V[8] = 1.0f;
    auto rotate = [](float* M, int i, int j, float c, float s) {
        for (int k = 0; k < 3; k++) {
            const float ti = M[k + 3 * i], tj = M[k + 3 * j];
            M[k + 3 * i] = c * ti - s * tj;
            M[k + 3 * j] = s * ti + c * tj;
        }
    };
    for (int sweep = 0; sweep < 4; sweep++)
        for (int j = 1; j < 3; j++)
            for (int i = 0; i < j; i++) {
                const float a = vdot(U + 3 * i, U + 3 * i, 3);
                const float b = vdot(U + 3 * j, U + 3 * j, 3);
                const float g = vdot(U + 3 * i, U + 3 * j, 3);
                const float zeta = (float)((double)(b - a) / (2.0 * (double)g));
                const float t = (float)((zeta > 0 ? 1.0 : -1.0) /
                                        ((double)std::fabs(zeta) + std::sqrt(1.0 + (double)(zeta * zeta))));
                const float c = (float)(1.0 / (double)std::sqrt(1.0f + t * t));
                const float s = c * t;
                rotate(U, i, j, c, s);
                rotate(V, i, j, c, s);
            }
    for (int j = 0; j < 3; j++) {
        const float sv = vnorm(U + 3 * j, 3);
        for (int i = 0; i < 3; i++) U[i + 3 * j] /= sv;
    }
}

// QRsolve.h:124-254: rotation from the SVD of the centred cross-covariance,
// translation from the mean of R^-1 applied to the targets minus the source mean.
void fit_rigid(float* RT, const float* p1t, const float* p2t, int len) {
    std::vector<float> p1(4 * (size_t)len), p2(4 * (size_t)len);
    for (int r = 0; r < 4; r++)
        for (int k = 0; k < len; k++) {
            p1[r + 4 * (size_t)k] = p1t[k + (size_t)r * len];
            p2[r + 4 * (size_t)k] = p2t[k + (size_t)r * len];
        }
    float c1[3] = {0, 0, 0}, c2[3] = {0, 0, 0};
    for (int k = 0; k < len; k++)
        for (int r = 0; r < 3; r++) {
            c1[r] += p1[r + 4 * (size_t)k];
            c2[r] += p2[r + 4 * (size_t)k];
        }
    for (int r = 0; r < 3; r++) {
        c1[r] /= (float)len;
        c2[r] /= (float)len;
    }
    float K[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int k = 0; k < len; k++)
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) K[i + 3 * j] += (p1[j + 4 * (size_t)k] - c1[j]) * (p2[i + 4 * (size_t)k] - c2[i]);
    float U[9], V[9], R[16];
    jacobi_svd3(K, U, V);
    std::fill(R, R + 16, 0.0f);
    std::fill(RT, RT + 16, 0.0f);
    R[15] = 1.0f;
    RT[15] = 1.0f;
    for (int k = 0; k < 3; k++)
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) R[j + 4 * i] += U[j + 3 * k] * V[i + 3 * k];
    std::vector<float> Rp2(4 * (size_t)len);
    qr_solve(Rp2.data(), R, p2.data(), 4, len);
    float t1[3] = {0, 0, 0};
    for (int k = 0; k < len; k++)
        for (int r = 0; r < 3; r++) t1[r] += Rp2[r + 4 * (size_t)k];
    for (int r = 0; r < 3; r++) t1[r] = t1[r] / (float)len - c1[r];
    float t0[3] = {0, 0, 0};
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) t0[i] += R[i + 4 * j] * t1[j];
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++) RT[i + 4 * j] += R[j + 4 * i];
        RT[3 + 4 * i] = t0[i];
    }
}

// the value std::nth_element leaves at position k
float kth_value(const float* v, int len, int k) {
    std::vector<float> c(v, v + len);
    std::nth_element(c.begin(), c.begin() + k, c.end());
    return c[k];
}

// QRsolve.h:258-296: fit to all pairs, then 15 refits to the pairs whose
// squared residual is at most the median (ties beyond half overwrite the last slot).
void robust_fit(float* RT, const float* p1, const float* p2, int len, bool rigid) {
    if (rigid) fit_rigid(RT, p1, p2, len);
    else qr_solve(RT, p1, p2, len, 4);
    const int half = len / 2;
    std::vector<float> err(len);
    std::vector<float> b1(4 * (size_t)std::max(half, 1), 1.0f), b2(4 * (size_t)std::max(half, 1), 1.0f);
    for (int iter = 0; iter < 15; iter++) {
        for (int l = 0; l < len; l++) {
            float y = 0, x = 0, z = 0;
            for (int r = 0; r < 4; r++) {
                const float v = p1[l + (size_t)r * len];
                y += v * RT[r];
                x += v * RT[r + 4];
                z += v * RT[r + 8];
            }
            const double ey = y - p2[l], ex = x - p2[l + len], ez = z - p2[l + 2 * (size_t)len];
            err[l] = (float)(ey * ey + ex * ex + ez * ez);
        }
        const float median = kth_value(err.data(), len, len / 2);
        int count = 0;
        for (int i = 0; i < len; i++) {
            if (!(err[i] <= median)) continue;
            const int l = std::min(count, half - 1);
            for (int r = 0; r < 3; r++) {
                b1[l + (size_t)r * half] = p1[i + (size_t)r * len];
                b2[l + (size_t)r * half] = p2[i + (size_t)r * len];
            }
            count++;
        }
        if (rigid) fit_rigid(RT, b1.data(), b2.data(), half);
        else qr_solve(RT, b1.data(), b2.data(), half, 4);
    }
}

// affineLTS1.h:1-12
void mat4_mult(const float* A, const float* B, float* C) {
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            float s = 0.0f;
            for (int k = 0; k < 4; k++) s += A[i + 4 * k] * B[k + 4 * j];
            C[i + 4 * j] = s;
        }
}

// affineLTS1.h:191-233: the quantile of the masked minima (linearBCV.cpp:36-46),
// the point pairs (fixed-direction: centre -> centre + label displacement;
// moving-direction: the reverse), the robust fit composed with Xprev.
int fit_level(const float* minval, const int* minind, int m1, int n1, int o1, int step, float quant, int hw,
              bool rigid, const float* Xprev, float* X) {
    const int N = m1 * n1 * o1, len = 2 * hw + 1, stephw = (step - 1) / 2;
    const float* mv[2] = {minval, minval + N};
    const int* mi[2] = {minind, minind + N};
    int allcount = 0;
    for (int d = 0; d < 2; d++)
        for (int i = 0; i < N; i++) allcount += (double)mv[d][i] < 1e19;  // affineLTS1.h:183-186
    const float q = (float)(0.5 * allcount / (double)(float)N / 2.0);
    const int qi = (int)((float)N * std::min(std::max(q, 0.01f), 0.99f));
    const float med[2] = {kth_value(mv[0], N, qi), kth_value(mv[1], N, qi)};
    int used = 0;
    for (int i = 0; i < N; i++) used += (mv[0][i] < med[0]) + (mv[1][i] < med[1]);
    std::vector<float> p1(4 * (size_t)std::max(used, 1)), p2(4 * (size_t)std::max(used, 1));
    int at = 0;
    for (int i = 0; i < N; i++) {
        const int z1 = i / (m1 * n1), x1 = (i / m1) % n1, y1 = i % m1;
        const float cy = (float)(y1 * step + stephw), cx = (float)(x1 * step + stephw), cz = (float)(z1 * step + stephw);
        for (int d = 0; d < 2; d++) {
            if (!(mv[d][i] < med[d])) continue;
            const int l = mi[d][i];
            const float dy = (float)(l % len - hw) * quant, dx = (float)((l / len) % len - hw) * quant,
                        dz = (float)(l / (len * len) - hw) * quant;
            float* src = d == 0 ? p1.data() : p2.data();
            float* dst = d == 0 ? p2.data() : p1.data();
            src[at] = cy;
            src[at + used] = cx;
            src[at + 2 * (size_t)used] = cz;
            dst[at] = cy + dy;
            dst[at + used] = cx + dx;
            dst[at + 2 * (size_t)used] = cz + dz;
            p1[at + 3 * (size_t)used] = 1.0f;
            p2[at + 3 * (size_t)used] = 1.0f;
            at++;
        }
    }
    float fit[16];
    robust_fit(fit, p1.data(), p2.data(), used, rigid);
    mat4_mult(fit, Xprev, X);
    return used;
}

}  // namespace

void linear_register(deeds_ctx* ctx, const float* d_fixed, const float* d_moving, int m, int n, int o,
                     const deeds_schedule& S, bool rigid, float* X_out, int* points_used) {
    hipStream_t s = ctx->stream;
    const long long sz = (long long)m * n * o;
    float* im1b = ctx->get<float>("lin.im1b", sz);
    float* im1 = ctx->get<float>("lin.im1", sz);
    float* w1 = ctx->get<float>("lin.w1", sz);
    float* w2 = ctx->get<float>("lin.w2", sz);
    uint64_t* mind_m = ctx->get<uint64_t>("lin.mind.m", sz);
    uint64_t* mind_f = ctx->get<uint64_t>("lin.mind.f", sz);
    uint64_t* mind_w = ctx->get<uint64_t>("lin.mind.w", sz);
    float* t0 = ctx->get<float>("lin.mindtmp.0", 6 * (size_t)sz);
    float* t1 = ctx->get<float>("lin.mindtmp.1", 6 * (size_t)sz);
    size_t max_cost = 0, max_pad = 0, max_n = 0;
    for (int l = 0; l < S.levels; l++) {
        const int st = S.grid[l], len = 2 * S.radius[l] + 1;
        const size_t N = (size_t)(m / st) * (n / st) * (o / st);
        const CostPlan plan = datacost_plan(m, n, o, st, S.radius[l], (float)S.quant[l], 1.0f, 1);
        max_cost = std::max(max_cost, N * len * len * len);
        max_pad = std::max(max_pad, (size_t)plan.mp * plan.np * plan.op);
        max_n = std::max(max_n, N);
    }
    uint64_t* pad = ctx->get<uint64_t>("lin.pad", max_pad);
    float* c1 = ctx->get<float>("lin.cost.0", max_cost);
    float* c2 = ctx->get<float>("lin.cost.1", max_cost);
    float* d_minval = ctx->get<float>("lin.minval", 2 * max_n);
    int* d_minind = ctx->get<int>("lin.minind", 2 * max_n);
    float* h_minval = static_cast<float*>(ctx->pinned("lin.minval", 2 * max_n * sizeof(float)));
    int* h_minind = static_cast<int*>(ctx->pinned("lin.minind", 2 * max_n * sizeof(int)));
    // linearBCV.cpp:146-154: +1024 HU
    launch_lin_offset(im1b, d_fixed, -1024.0f, sz, s);
    launch_lin_offset(im1, d_moving, -1024.0f, sz, s);
    float Xprev[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1}, X[16];
    const float ident[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    std::copy(Xprev, Xprev + 16, X);
    int prev_qs = -1;
    for (int l = 0; l < S.levels; l++) {
        const int step = S.grid[l], hw = S.radius[l], len = 2 * hw + 1, len3 = len * len * len;
        const float quant = (float)S.quant[l];
        const int m1 = m / step, n1 = n / step, o1 = o / step;
        // the inverse of the current estimate by the same QR solve (linearBCV.cpp:204-205)
        float Xinv[16];
        qr_solve(Xinv, Xprev, ident, 4, 4);
        launch_lin_warp(w2, im1b, Xinv, m, n, o, s);
        launch_lin_warp(w1, im1, Xprev, m, n, o, s);
        const int qs = (int)std::floor(0.5f * quant + 1.0f);  // linearBCV.cpp:163-165
        if (l == 0 || qs != prev_qs) {
            launch_descriptor(mind_m, im1, m, n, o, qs, t0, t1, s);
            launch_descriptor(mind_f, im1b, m, n, o, qs, t0, t1, s);
        }
        prev_qs = qs;
        const CostPlan plan = datacost_plan(m, n, o, step, hw, quant, 1.0f, 1);
        // forward: fixed codes against the warped moving image; backward: the reverse
        launch_descriptor(mind_w, w1, m, n, o, qs, t0, t1, s);
        launch_pad_codes(pad, mind_w, m, n, o, plan.pad, s);
        launch_datacost(c1, mind_f, pad, m, n, o, step, hw, (int)quant, plan, s);
        launch_descriptor(mind_w, w2, m, n, o, qs, t0, t1, s);
        launch_pad_codes(pad, mind_w, m, n, o, plan.pad, s);
        launch_datacost(c2, mind_m, pad, m, n, o, step, hw, (int)quant, plan, s);
        launch_lin_smooth(c1, m1, n1, o1, len3, s);
        launch_lin_smooth(c2, m1, n1, o1, len3, s);
        launch_lin_nodes(c1, c2, im1b, im1, m, n, step, m1, n1, o1, len3, d_minval, d_minind, s);
        const size_t N = (size_t)m1 * n1 * o1;
        d2h(h_minval, d_minval, 2 * N, s);
        d2h(h_minind, d_minind, 2 * N, s);
        DEEDS_HIP(hipStreamSynchronize(s));
        const int used = fit_level(h_minval, h_minind, m1, n1, o1, step, quant, hw, rigid, Xprev, X);
        if (points_used) points_used[l] = used;
        std::copy(X, X + 16, Xprev);
    }
    std::copy(X, X + 16, X_out);
}

}  // namespace deeds

using namespace deeds;

namespace {

void check_linear(const deeds_schedule* S, int m, int n, int o) {
    check_schedule(S, m, n, o);
    int max_qs = 0;
    for (int l = 0; l < S->levels; l++) {
        const int st = S->grid[l];
        DEEDS_REQUIRE(m / st >= 1 && n / st >= 1 && o / st >= 1, "grid spacing larger than a dimension");
        max_qs = std::max(max_qs, (int)std::floor(0.5f * (float)S->quant[l] + 1.0f));
    }
    DEEDS_REQUIRE(m >= 2 * max_qs + 1 && n >= 2 * max_qs + 1 && o >= 2 * max_qs + 1,
                  "volume too small for the MIND patch radius of the schedule");
    DEEDS_REQUIRE(m <= mind_max_dim() && n <= mind_max_dim() && o <= mind_max_dim(),
                  "a dimension exceeds the box filter's LDS tile limit");
}

}  // namespace

extern "C" {

int deeds_ctx_linear_register(deeds_ctx* ctx, const float* d_fixed, const float* d_moving, int m, int n, int o,
                              const deeds_schedule* sched, int rigid, float* X, int* points_used) {
    if (!ctx) return DEEDS_ERR_ARG;
    return guarded(ctx, [&] {
        DEEDS_REQUIRE(d_fixed && d_moving && X, "null pointer");
        check_linear(sched, m, n, o);
        linear_register(ctx, d_fixed, d_moving, m, n, o, *sched, rigid != 0, X, points_used);
    });
}

int deeds_linear_register(const float* fixed, const float* moving, int m, int n, int o, const deeds_schedule* sched,
                          int rigid, float* X, int* points_used) {
    return guarded(nullptr, [&] {
        DEEDS_REQUIRE(fixed && moving && X, "null pointer");
        check_linear(sched, m, n, o);
        deeds_ctx* c = default_ctx();
        DEEDS_HIP(hipSetDevice(c->device));
        const size_t sz = (size_t)m * n * o;
        float* d_in = c->get<float>("h.lin.in", 2 * sz);
        h2d(d_in, fixed, sz, c->stream);
        h2d(d_in + sz, moving, sz, c->stream);
        const int rc = deeds_ctx_linear_register(c, d_in, d_in + sz, m, n, o, sched, rigid, X, points_used);
        if (rc) fail(rc, c->err);
    });
}

}  // extern "C"
