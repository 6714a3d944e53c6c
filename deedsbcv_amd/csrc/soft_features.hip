// Soft variant, stage 1: the un-binarised self-similarity features and their
// adjoint (the model is stated in soft.h).
//
// Forward:  S_c = (I(p + t_c) - I(p))^2 for the six shifts t_c, zero where
//           p + t_c leaves the volume (MINDSSCbox.h:79-90, :113-117); D_c = the
//           clamped box sums of radius qs along i, j, k (:2-76, as window
//           sums rather than prefix-sum differences); phi = the twelve SSC
//           values minus their minimum over max(mean, 1e-6) (:184-201).
// Backward: the adjoint of each step in reverse (the 3-D box is symmetric,
//           so it is its own adjoint); the distances D are recomputed.  All
//           kernels are elementwise or short stencils (HBM bound); the three
//           box passes run as one LDS-tiled kernel.
#include "soft_common.h"

namespace deeds {

namespace {

using namespace soft_dev;

// SSD shifts (di, dj, dk) of the six channels in units of qs (MINDSSCbox.h:110-112)
__constant__ int c_sshift[6][3] = {{+1, +1, 0}, {-1, +1, 0}, {0, -1, +1}, {-1, 0, +1}, {0, +1, +1}, {+1, 0, +1}};
// SSC offsets (di, dj, dk) in units of qs and their channel (MINDSSCbox.h:136-140)
__constant__ int c_sssc[12][4] = {{0, -1, 0, 0}, {-1, 0, 0, 0}, {0, -1, 0, 1}, {+1, 0, 0, 1},
                                  {0, 0, -1, 2}, {0, +1, 0, 2}, {0, 0, -1, 3}, {+1, 0, 0, 3},
                                  {0, 0, -1, 4}, {0, -1, 0, 4}, {0, 0, -1, 5}, {-1, 0, 0, 5}};

struct Dims {
    int m, n, o;
    long long sz;
    __device__ bool in(int i, int j, int k) const { return i >= 0 && i < m && j >= 0 && j < n && k >= 0 && k < o; }
    __device__ long long at(int i, int j, int k) const { return i + (long long)j * m + (long long)k * m * n; }
};

// S[c][p] = (im[p + t_c] - im[p])^2, 0 where p + t_c leaves the volume
template <typename R>
__global__ void k_sf_ssd(R* __restrict__ S, const R* __restrict__ im, Dims d, int qs) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= 6 * d.sz) return;
    const int c = (int)(idx / d.sz);
    const long long p = idx - c * d.sz;
    const int i = (int)(p % d.m), j = (int)((p / d.m) % d.n), k = (int)(p / ((long long)d.m * d.n));
    const int ii = i + c_sshift[c][0] * qs, jj = j + c_sshift[c][1] * qs, kk = k + c_sshift[c][2] * qs;
    R v = 0;
    if (d.in(ii, jj, kk)) {
        const R t = im[d.at(ii, jj, kk)] - im[p];
        v = t * t;
    }
    S[idx] = v;
}

// Clamped window sum of radius r along one axis of the six planar channels.
// The operator is symmetric, so it is also its own adjoint.
template <typename R>
__global__ void k_sf_box(R* __restrict__ out, const R* __restrict__ in, Dims d, int r, int axis) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= 6 * d.sz) return;
    const long long p = idx % d.sz;
    const int i = (int)(p % d.m), j = (int)((p / d.m) % d.n), k = (int)(p / ((long long)d.m * d.n));
    const int x = axis == 0 ? i : axis == 1 ? j : k;
    const int L = axis == 0 ? d.m : axis == 1 ? d.n : d.o;
    const long long stride = axis == 0 ? 1 : axis == 1 ? d.m : (long long)d.m * d.n;
    const int lo = x - r < 0 ? 0 : x - r, hi = x + r > L - 1 ? L - 1 : x + r;
    const R* base = in + (idx - (long long)x * stride);
    R acc = 0;
    for (int t = lo; t <= hi; t++) acc += base[t * stride];
    out[idx] = acc;
}

// The three box passes (i, then j, then k) of the six channels in one kernel:
// a workgroup stages a BI x BJ x BK output tile with its RAD halo in LDS (zero
// beyond the volume, which is the clamped window), sums along i into a
// second buffer, along j back into the first, and along k into the output.
// Every window sum adds its terms in ascending order as k_sf_box does, and
// the zeros it adds fall before or after the in-volume terms, so the result
// is k_sf_box's, three passes through HBM fewer.
constexpr int kBI = 32, kBJ = 8, kBK = 8;
template <int RAD>
struct Box3 {
    static constexpr int XI = kBI + 2 * RAD, XJ = kBJ + 2 * RAD, XK = kBK + 2 * RAD;
    static constexpr int NA = XI * XJ * XK, NB = kBI * XJ * XK;
};
template <typename R, int RAD>
__global__ __launch_bounds__(256) void k_sf_box3(R* __restrict__ out, const R* __restrict__ in, Dims d, int ti,
                                                  int tj) {
    using X = Box3<RAD>;
    extern __shared__ __align__(16) unsigned char sf_smem[];
    R* A = reinterpret_cast<R*>(sf_smem);  // [XK][XJ][XI], then [XK][BJ][BI]
    R* B = A + X::NA;                      // [XK][XJ][BI]
    const int tile = blockIdx.x, ch = blockIdx.y;
    const int bi = tile % ti, bj = (tile / ti) % tj, bk = tile / (ti * tj);
    const int i0 = bi * kBI - RAD, j0 = bj * kBJ - RAD, k0 = bk * kBK - RAD;
    const R* src = in + ch * d.sz;
    for (int e = threadIdx.x; e < X::NA; e += blockDim.x) {
        const int ii = e % X::XI, jj = (e / X::XI) % X::XJ, kk = e / (X::XI * X::XJ);
        const int i = i0 + ii, j = j0 + jj, k = k0 + kk;
        A[e] = d.in(i, j, k) ? src[d.at(i, j, k)] : (R)0;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < X::NB; e += blockDim.x) {
        const int ii = e % kBI, r = e / kBI;  // r = jj + XJ * kk
        const R* a = A + r * X::XI + ii;
        R acc = 0;
#pragma unroll
        for (int t = 0; t <= 2 * RAD; t++) acc += a[t];
        B[e] = acc;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < kBI * kBJ * X::XK; e += blockDim.x) {
        const int ii = e % kBI, jo = (e / kBI) % kBJ, kk = e / (kBI * kBJ);
        const R* b = B + (kk * X::XJ + jo) * kBI + ii;
        R acc = 0;
#pragma unroll
        for (int t = 0; t <= 2 * RAD; t++) acc += b[t * kBI];
        A[e] = acc;
    }
    __syncthreads();
    R* dst = out + ch * d.sz;
    for (int e = threadIdx.x; e < kBI * kBJ * kBK; e += blockDim.x) {
        const int ii = e % kBI, jo = (e / kBI) % kBJ, ko = e / (kBI * kBJ);
        const int i = bi * kBI + ii, j = bj * kBJ + jo, k = bk * kBK + ko;
        if (!d.in(i, j, k)) continue;
        const R* c = A + ko * kBJ * kBI + jo * kBI + ii;
        R acc = 0;
#pragma unroll
        for (int t = 0; t <= 2 * RAD; t++) acc += c[t * kBJ * kBI];
        dst[d.at(i, j, k)] = acc;
    }
}

// The 3-D box of radius r over the six channels: fused for radius 1..3, else
// three k_sf_box launches through tmp, in the given axis order (0, 1, 2 or,
// for the adjoint, its reverse).  The 1-D boxes are symmetric and commute, so
// the 3-D box is its own adjoint in either order.
template <typename R>
void box3(R* out, const R* in, R* tmp, const Dims& d, int r, bool forward, hipStream_t s) {
    const unsigned g6 = ceil_div(6 * d.sz, 256);
    if (r >= 1 && r <= 3) {
        const int ti = (int)ceil_div(d.m, kBI), tj = (int)ceil_div(d.n, kBJ), tk = (int)ceil_div(d.o, kBK);
        const dim3 grid((unsigned)(ti * tj * tk), 6);
        auto go = [&](auto kern, int NA, int NB) {
            const size_t smem = (size_t)(NA + NB) * sizeof(R);
            set_smem((const void*)kern, smem);
            kern<<<grid, 256, smem, s>>>(out, in, d, ti, tj);
        };
        if (r == 1) go(k_sf_box3<R, 1>, Box3<1>::NA, Box3<1>::NB);
        else if (r == 2) go(k_sf_box3<R, 2>, Box3<2>::NA, Box3<2>::NB);
        else go(k_sf_box3<R, 3>, Box3<3>::NA, Box3<3>::NB);
        DEEDS_LAUNCH_CHECK();
        return;
    }
    const int a0 = forward ? 0 : 2, a2 = forward ? 2 : 0;
    k_sf_box<R><<<g6, 256, 0, s>>>(out, in, d, r, a0);
    k_sf_box<R><<<g6, 256, 0, s>>>(tmp, out, d, r, 1);
    k_sf_box<R><<<g6, 256, 0, s>>>(out, tmp, d, r, a2);
    DEEDS_LAUNCH_CHECK();
}

// the twelve SSC distances of voxel (i,j,k) with the out-of-volume fallback of MINDSSCbox.h:186-190
template <typename R>
__device__ __forceinline__ void ssc_values(R v[12], const R* __restrict__ D, const Dims& d, int i, int j, int k,
                                           int qs) {
#pragma unroll
    for (int l = 0; l < 12; l++) {
        int ii = i + c_sssc[l][0] * qs, jj = j + c_sssc[l][1] * qs, kk = k + c_sssc[l][2] * qs;
        if (!d.in(ii, jj, kk)) { ii = i; jj = j; kk = k; }
        v[l] = D[c_sssc[l][3] * d.sz + d.at(ii, jj, kk)];
    }
}

template <typename R>
__global__ void k_sf_phi(R* __restrict__ phi, const R* __restrict__ D, Dims d, int qs) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= d.sz) return;
    const int i = (int)(p % d.m), j = (int)((p / d.m) % d.n), k = (int)(p / ((long long)d.m * d.n));
    R v[12];
    ssc_values<R>(v, D, d, i, j, k, qs);
    R mn = v[0];
#pragma unroll
    for (int l = 1; l < 12; l++) mn = v[l] < mn ? v[l] : mn;
    R sum = 0;
#pragma unroll
    for (int l = 0; l < 12; l++) sum += v[l] - mn;
    R noise = sum / (R)12;
    noise = noise > (R)1e-6 ? noise : (R)1e-6;
#pragma unroll
    for (int l = 0; l < 12; l++) phi[12 * p + l] = (v[l] - mn) / noise;
}

// adjoint of the normalisation: gd[p][l] = d(loss)/d(v_l) at voxel p
template <typename R>
__global__ void k_sf_phi_bwd(R* __restrict__ gd, const R* __restrict__ gphi, const R* __restrict__ D, Dims d,
                             int qs) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= d.sz) return;
    const int i = (int)(p % d.m), j = (int)((p / d.m) % d.n), k = (int)(p / ((long long)d.m * d.n));
    R v[12], g[12];
    ssc_values<R>(v, D, d, i, j, k, qs);
    int a = 0;
    R mn = v[0];
#pragma unroll
    for (int l = 1; l < 12; l++)
        if (v[l] < mn) { mn = v[l]; a = l; }  // first minimum, as min_element
    R sum = 0, G = 0, P = 0;
#pragma unroll
    for (int l = 0; l < 12; l++) {
        g[l] = gphi[12 * p + l];
        sum += v[l] - mn;
    }
    const R mu = sum / (R)12;
    const bool clamped = !(mu > (R)1e-6);
    const R noise = clamped ? (R)1e-6 : mu;
#pragma unroll
    for (int l = 0; l < 12; l++) {
        G += g[l];
        P += g[l] * (v[l] - mn) / noise;
    }
#pragma unroll
    for (int l = 0; l < 12; l++) {
        R r = (g[l] - (l == a ? G : (R)0)) / noise;
        if (!clamped) r -= P / noise * ((R)1 / (R)12 - (l == a ? (R)1 : (R)0));
        gd[12 * p + l] = r;
    }
}

// gD[c][q] = sum over the two SSC values l of channel c that read D_c at q
template <typename R>
__global__ void k_sf_gather(R* __restrict__ gD, const R* __restrict__ gd, Dims d, int qs) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= 6 * d.sz) return;
    const int c = (int)(idx / d.sz);
    const long long q = idx - c * d.sz;
    const int i = (int)(q % d.m), j = (int)((q / d.m) % d.n), k = (int)(q / ((long long)d.m * d.n));
    R acc = 0;
#pragma unroll
    for (int e = 0; e < 2; e++) {
        const int l = 2 * c + e;
        const int di = c_sssc[l][0] * qs, dj = c_sssc[l][1] * qs, dk = c_sssc[l][2] * qs;
        if (d.in(i - di, j - dj, k - dk)) acc += gd[12 * d.at(i - di, j - dj, k - dk) + l];
        if (!d.in(i + di, j + dj, k + dk)) acc += gd[12 * q + l];
    }
    gD[idx] = acc;
}

// adjoint of k_sf_ssd: gim[q] from gS (planar six channels)
template <typename R>
__global__ void k_sf_ssd_bwd(R* __restrict__ gim, const R* __restrict__ gS, const R* __restrict__ im, Dims d, int qs) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= d.sz) return;
    const int i = (int)(q % d.m), j = (int)((q / d.m) % d.n), k = (int)(q / ((long long)d.m * d.n));
    const R x = im[q];
    R acc = 0;
#pragma unroll
    for (int c = 0; c < 6; c++) {
        const int di = c_sshift[c][0] * qs, dj = c_sshift[c][1] * qs, dk = c_sshift[c][2] * qs;
        if (d.in(i - di, j - dj, k - dk)) {  // q is the shifted voxel of p = q - t_c
            const long long p = d.at(i - di, j - dj, k - dk);
            acc += (R)2 * (x - im[p]) * gS[c * d.sz + p];
        }
        if (d.in(i + di, j + dj, k + dk))  // q is the centre voxel
            acc -= (R)2 * (im[d.at(i + di, j + dj, k + dk)] - x) * gS[c * d.sz + q];
    }
    gim[q] = acc;
}

Dims dims_of(int m, int n, int o) { return Dims{m, n, o, (long long)m * n * o}; }

}  // namespace

// ---- features ---------------------------------------------------------------

template <typename R>
void soft_features(R* phi, const R* im, int m, int n, int o, int qs, R* w0, R* w1, hipStream_t s) {
    const Dims d = dims_of(m, n, o);
    const unsigned g6 = ceil_div(6 * d.sz, 256), g1 = ceil_div(d.sz, 256);
    k_sf_ssd<R><<<g6, 256, 0, s>>>(w0, im, d, qs);
    DEEDS_LAUNCH_CHECK();
    box3<R>(w1, w0, w0, d, qs, true, s);
    k_sf_phi<R><<<g1, 256, 0, s>>>(phi, w1, d, qs);
    DEEDS_LAUNCH_CHECK();
}

template <typename R>
void soft_features_backward(R* gim, const R* gphi, const R* im, int m, int n, int o, int qs, R* w0, R* w1, R* w2,
                            hipStream_t s) {
    const Dims d = dims_of(m, n, o);
    const unsigned g6 = ceil_div(6 * d.sz, 256), g1 = ceil_div(d.sz, 256);
    // recompute the box-filtered distances D into w1
    k_sf_ssd<R><<<g6, 256, 0, s>>>(w0, im, d, qs);
    DEEDS_LAUNCH_CHECK();
    box3<R>(w1, w0, w0, d, qs, true, s);
    // w2 (12 per voxel) = d/dv, w0 = d/dD, then back through the box (w1) and the SSD
    k_sf_phi_bwd<R><<<g1, 256, 0, s>>>(w2, gphi, w1, d, qs);
    k_sf_gather<R><<<g6, 256, 0, s>>>(w0, w2, d, qs);
    DEEDS_LAUNCH_CHECK();
    box3<R>(w1, w0, w0, d, qs, false, s);
    k_sf_ssd_bwd<R><<<g1, 256, 0, s>>>(gim, w1, im, d, qs);
    DEEDS_LAUNCH_CHECK();
}

#define DEEDS_SOFT_FEATURES(R)                                                                 \
    template void soft_features<R>(R*, const R*, int, int, int, int, R*, R*, hipStream_t); \
    template void soft_features_backward<R>(R*, const R*, const R*, int, int, int, int, R*, R*, R*, hipStream_t);

DEEDS_SOFT_FEATURES(float)
DEEDS_SOFT_FEATURES(double)

}  // namespace deeds
