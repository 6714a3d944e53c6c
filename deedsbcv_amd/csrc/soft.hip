// Differentiable soft variant of the deeds registration step on gfx950 (no
// counterpart in the reference; the model below is this library's own).
//
//   features  phi[p][l], l = 0..11: the un-binarised self-similarity context
//             of MINDSSCbox.h:127-201 -- six box-filtered squared
//             differences D_c (:104-124, clamped window sums of radius qs),
//             read at the twelve SSC offsets (:184-191), minus their minimum,
//             divided by max(mean, 1e-6) (:192-201) -- before thresholding
//   cost      C[n,l] = a1 * sum_{samples s of n} sum_ch (F[s,ch] - M[clamp(s+d(l)),ch])^2
//             with the samples, clamping and a1 of dataCostD.h:141-196
//   energy    E(L) = sum_n C_n(l_n) + sum_{c != root} |l_p - l_c + off_c|^2 / w_c
//             on a tree in BFS order (the MST of primsMST.h), w_c = edgemst[c],
//             off_c = (u0[p] - u0[c]) / quant per axis (regularisation.h:208-210)
//   inference sum-product BP at temperature T: log-sum-exp distance
//             transforms separable over the three label axes, upward then
//             downward messages, marginals P_c = softmax(-g_c / T), output
//             u1_c = u0_c + E_{P_c}[label displacement]
//   backward  hand-written adjoints of all three stages; BP keeps the upward
//             beliefs, upward messages and full beliefs of every node and
//             recomputes each distance transform in the backward sweeps
// Every stage is templated on the real type: float for speed, double for the
// finite-difference gradient checks.
#include <cmath>
#include <vector>

#include "kernels.h"
#include "soft.h"

namespace deeds {

namespace {

constexpr int SB = 256;  // threads per soft BP workgroup (one tree node each)

template <typename R>
__device__ __forceinline__ R r_exp(R x);
template <>
__device__ __forceinline__ float r_exp(float x) { return __expf(x); }
template <>
__device__ __forceinline__ double r_exp(double x) { return exp(x); }
template <typename R>
__device__ __forceinline__ R r_log(R x);
template <>
__device__ __forceinline__ float r_log(float x) { return __logf(x); }
template <>
__device__ __forceinline__ double r_log(double x) { return log(x); }
template <typename R>
__device__ __forceinline__ R r_big();
template <>
__device__ __forceinline__ float r_big() { return 3.0e38f; }
template <>
__device__ __forceinline__ double r_big() { return 1.0e300; }

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// ---------------------------------------------------------------------------
// features
// ---------------------------------------------------------------------------

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

// ---------------------------------------------------------------------------
// cost volume
// ---------------------------------------------------------------------------

struct CostGeom {
    int m, n, o, m1, n1, o1, step, len, q, hw, si, sj, sk;
    double scale;
};

CostGeom cost_geom(int m, int n, int o, int step, int hw, float quant, float alpha) {
    CostPlan plan = datacost_plan(m, n, o, step, hw, quant, alpha, 1);
    CostGeom g{};
    g.m = m; g.n = n; g.o = o;
    g.m1 = m / step; g.n1 = n / step; g.o1 = o / step;
    g.step = step; g.len = 2 * hw + 1; g.q = (int)quant; g.hw = hw;
    g.si = plan.skip_i; g.sj = plan.skip_j; g.sk = plan.skip_k;
    g.scale = (double)plan.scale;
    return g;
}

// one thread per (node, label); features are [voxel][12]
template <typename R>
__global__ void k_sc_cost(R* __restrict__ cost, const R* __restrict__ F, const R* __restrict__ M, CostGeom g) {
    const int L3 = g.len * g.len * g.len;
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long nodes = (long long)g.m1 * g.n1 * g.o1;
    if (idx >= nodes * L3) return;
    const int node = (int)(idx / L3), l = (int)(idx - (long long)node * L3);
    const int y = node % g.m1, x = (node / g.m1) % g.n1, z = node / (g.m1 * g.n1);
    const int ly = l % g.len, lx = (l / g.len) % g.len, lz = l / (g.len * g.len);
    const int dy = (ly - g.hw) * g.q, dx = (lx - g.hw) * g.q, dz = (lz - g.hw) * g.q;
    const long long mn = (long long)g.m * g.n;
    R acc = 0;
    for (int k = 0; k < g.step; k += g.sk) {
        const int fk = z * g.step + k, mk = clampi(fk + dz, 0, g.o - 1);
        for (int j = 0; j < g.step; j += g.sj) {
            const int fj = x * g.step + j, mj = clampi(fj + dx, 0, g.n - 1);
            for (int i = 0; i < g.step; i += g.si) {
                const int fi = y * g.step + i, mi = clampi(fi + dy, 0, g.m - 1);
                const R* fp = F + 12 * (fi + (long long)fj * g.m + fk * mn);
                const R* mp = M + 12 * (mi + (long long)mj * g.m + mk * mn);
#pragma unroll
                for (int ch = 0; ch < 12; ch++) {
                    const R t = fp[ch] - mp[ch];
                    acc += t * t;
                }
            }
        }
    }
    cost[idx] = acc * (R)g.scale;
}

// dF at the samples of one node (one workgroup per node, one thread per (sample, channel))
template <typename R>
__global__ void k_sc_dF(R* __restrict__ dF, const R* __restrict__ gc, const R* __restrict__ F,
                        const R* __restrict__ M, CostGeom g) {
    extern __shared__ unsigned char sc_smem[];
    R* gl = reinterpret_cast<R*>(sc_smem);
    const int node = blockIdx.x;
    const int L = g.len, L3 = L * L * L;
    for (int l = threadIdx.x; l < L3; l += blockDim.x) gl[l] = gc[(long long)node * L3 + l];
    __syncthreads();
    const int y = node % g.m1, x = (node / g.m1) % g.n1, z = node / (g.m1 * g.n1);
    const int ni = (g.step + g.si - 1) / g.si, nj = (g.step + g.sj - 1) / g.sj, nk = (g.step + g.sk - 1) / g.sk;
    const long long mn = (long long)g.m * g.n;
    for (int w = threadIdx.x; w < ni * nj * nk * 12; w += blockDim.x) {
        const int ch = w % 12, s = w / 12;
        const int fi = y * g.step + (s % ni) * g.si, fj = x * g.step + ((s / ni) % nj) * g.sj,
                  fk = z * g.step + (s / (ni * nj)) * g.sk;
        const long long fidx = 12 * (fi + (long long)fj * g.m + fk * mn) + ch;
        const R f = F[fidx];
        R acc = 0;
        int l = 0;
        for (int lz = 0; lz < L; lz++) {
            const long long zo = clampi(fk + (lz - g.hw) * g.q, 0, g.o - 1) * mn;
            for (int lx = 0; lx < L; lx++) {
                const long long xo = zo + (long long)clampi(fj + (lx - g.hw) * g.q, 0, g.n - 1) * g.m;
                for (int ly = 0; ly < L; ly++, l++) {
                    const int mi = clampi(fi + (ly - g.hw) * g.q, 0, g.m - 1);
                    acc += gl[l] * (f - M[12 * (xo + mi) + ch]);
                }
            }
        }
        dF[fidx] = (R)(2.0 * g.scale) * acc;
    }
}

// Sample coordinates s along one axis with clamp(s + dd, 0, len-1) == t, restricted
// to the control-grid region [0, extent).
__device__ __forceinline__ void axis_range(int t, int dd, int len, int extent, int& lo, int& hi) {
    lo = t - dd;
    hi = t - dd;
    if (t == 0) lo = -(1 << 20);
    if (t == len - 1) hi = 1 << 20;
    lo = lo < 0 ? 0 : lo;
    hi = hi > extent - 1 ? extent - 1 : hi;
}

// dM at one moving voxel t (one thread per voxel, all twelve channels): every
// (sample s, label l) whose clamped displaced position is t contributes
// 2 a1 g[n(s), l] (M_t - F_s)
template <typename R>
__global__ void k_sc_dM(R* __restrict__ dM, const R* __restrict__ gc, const R* __restrict__ F,
                        const R* __restrict__ M, CostGeom g) {
    const long long sz = (long long)g.m * g.n * g.o;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= sz) return;
    const int ti = (int)(t % g.m), tj = (int)((t / g.m) % g.n), tk = (int)(t / ((long long)g.m * g.n));
    const int L = g.len, L3 = L * L * L;
    const long long mn = (long long)g.m * g.n;
    R accg = 0, accf[12];
#pragma unroll
    for (int ch = 0; ch < 12; ch++) accf[ch] = 0;
    for (int lz = 0; lz < L; lz++) {
        int zlo, zhi;
        axis_range(tk, (lz - g.hw) * g.q, g.o, g.o1 * g.step, zlo, zhi);
        for (int sk = zlo; sk <= zhi; sk++) {
            if ((sk % g.step) % g.sk) continue;
            const int zn = (sk / g.step) * g.m1 * g.n1;
            for (int lx = 0; lx < L; lx++) {
                int xlo, xhi;
                axis_range(tj, (lx - g.hw) * g.q, g.n, g.n1 * g.step, xlo, xhi);
                for (int sj = xlo; sj <= xhi; sj++) {
                    if ((sj % g.step) % g.sj) continue;
                    const int xn = zn + (sj / g.step) * g.m1;
                    for (int ly = 0; ly < L; ly++) {
                        int ylo, yhi;
                        axis_range(ti, (ly - g.hw) * g.q, g.m, g.m1 * g.step, ylo, yhi);
                        for (int si = ylo; si <= yhi; si++) {
                            if ((si % g.step) % g.si) continue;
                            const int node = xn + si / g.step;
                            const R w = gc[(long long)node * L3 + ly + lx * L + lz * L * L];
                            const R* fp = F + 12 * (si + (long long)sj * g.m + sk * mn);
                            accg += w;
#pragma unroll
                            for (int ch = 0; ch < 12; ch++) accf[ch] += w * fp[ch];
                        }
                    }
                }
            }
        }
    }
#pragma unroll
    for (int ch = 0; ch < 12; ch++) dM[12 * t + ch] = (R)(2.0 * g.scale) * (M[12 * t + ch] * accg - accf[ch]);
}

template <typename R>
__global__ void k_fill(R* p, R v, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ---------------------------------------------------------------------------
// sum-product BP
// ---------------------------------------------------------------------------

template <typename R>
__device__ R block_reduce(R v, R* red, bool is_min) {
    for (int off = 32; off > 0; off >>= 1) {
        const R o = __shfl_xor(v, off, 64);
        v = is_min ? (o < v ? o : v) : v + o;
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[wid] = v;
    __syncthreads();
    R r = red[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) r = is_min ? (red[w] < r ? red[w] : r) : r + red[w];
    __syncthreads();
    return r;
}

// out(a) = -T log sum_b exp(-(in(b) + z[a - b + L]) / T) along axis ax of the L^3 cube
template <typename R>
__device__ void lse_pass(const R* in, R* out, const R* z, int L, int ax, R T) {
    const int L3 = L * L * L;
    const int stride = ax == 0 ? 1 : ax == 1 ? L : L * L;
    const R iT = (R)1 / T;
    for (int o = threadIdx.x; o < L3; o += blockDim.x) {
        const int a = (o / stride) % L;
        const int base = o - a * stride;
        const R* zz = z + a + L;  // zz[-b] = z[a - b + L]
        R mn = r_big<R>();
        for (int b = 0; b < L; b++) {
            const R v = in[base + b * stride] + zz[-b];
            mn = v < mn ? v : mn;
        }
        R s = 0;
        for (int b = 0; b < L; b++) s += r_exp<R>((mn - (in[base + b * stride] + zz[-b])) * iT);
        out[o] = mn - T * r_log<R>(s);
    }
}

// adjoint: gin(b) = sum_a gout(a) exp(-(in(b) + z[a - b + L] - out(a)) / T)
template <typename R>
__device__ void lse_pass_adj(const R* in, const R* out, const R* gout, R* gin, const R* z, int L, int ax, R T) {
    const int L3 = L * L * L;
    const int stride = ax == 0 ? 1 : ax == 1 ? L : L * L;
    const R iT = (R)1 / T;
    for (int idx = threadIdx.x; idx < L3; idx += blockDim.x) {
        const int b = (idx / stride) % L;
        const int base = idx - b * stride;
        const R x = in[idx];
        const R* zz = z + L - b;  // zz[a] = z[a - b + L]
        R acc = 0;
        for (int a = 0; a < L; a++) {
            const int oa = base + a * stride;
            acc += gout[oa] * r_exp<R>((out[oa] - x - zz[a]) * iT);
        }
        gin[idx] = acc;
    }
}

struct SoftTreeDev {
    const int* ordered;
    const int* parents;
    const int* child_start;  // N + 1
    const int* child_list;   // N - 1
    const float* edgemst;
};

template <typename R>
struct BPState {
    const R* cost;  // N x L3
    R* H;           // upward beliefs h_c = C_c + sum of the children's messages
    R* Msg;         // upward messages, shifted to minimum 0
    R* G;           // full beliefs g_c = h_c + downward message
    const R* u0;    // 3 x N (u, v, w)
    R* u1;          // 3 x N
    SoftTreeDev t;
    int N, L, hw;
    R quant, T;
};

template <typename R>
struct Lds {
    R *X0, *X1, *X2, *X3, *Ga, *Gb, *zy, *zx, *zz, *red;
    __device__ Lds(R* base, int L) {
        const int L3 = L * L * L;
        X0 = base; X1 = X0 + L3; X2 = X1 + L3; X3 = X2 + L3; Ga = X3 + L3; Gb = Ga + L3;
        zy = Gb + L3; zx = zy + (2 * L + 1); zz = zx + (2 * L + 1); red = zz + (2 * L + 1);
    }
};

// quadratic tables of the edge c -> p per axis; sign +1 for the upward
// message (a = parent label, b = child label), -1 for the downward one
template <typename R>
__device__ void edge_tables(const BPState<R>& a, int c, int p, R sign, Lds<R>& s) {
    const int L = a.L;
    const float wf = a.t.edgemst[c];
    const R w = (R)(wf > kSoftMinEdgeWeight ? wf : kSoftMinEdgeWeight);
    const R ou = sign * (a.u0[p] - a.u0[c]) / a.quant;
    const R ov = sign * (a.u0[a.N + p] - a.u0[a.N + c]) / a.quant;
    const R ow = sign * (a.u0[2 * a.N + p] - a.u0[2 * a.N + c]) / a.quant;
    for (int q = threadIdx.x; q < 2 * L + 1; q += blockDim.x) {
        const R t = (R)(q - L);
        s.zy[q] = (t + ov) * (t + ov) / w;  // label axis 0 is i (v)
        s.zx[q] = (t + ou) * (t + ou) / w;  // axis 1 is j (u)
        s.zz[q] = (t + ow) * (t + ow) / w;  // axis 2 is k (w)
    }
}

template <typename R>
__device__ void dt3(Lds<R>& s, int L, R T) {
    __syncthreads();
    lse_pass<R>(s.X0, s.X1, s.zy, L, 0, T);
    __syncthreads();
    lse_pass<R>(s.X1, s.X2, s.zx, L, 1, T);
    __syncthreads();
    lse_pass<R>(s.X2, s.X3, s.zz, L, 2, T);
    __syncthreads();
}

// adjoint of dt3 for d/dX3 in Ga; leaves d/dX0 in Gb
template <typename R>
__device__ void dt3_adj(Lds<R>& s, int L, R T) {
    __syncthreads();
    lse_pass_adj<R>(s.X2, s.X3, s.Ga, s.Gb, s.zz, L, 2, T);
    __syncthreads();
    lse_pass_adj<R>(s.X1, s.X2, s.Gb, s.Ga, s.zx, L, 1, T);
    __syncthreads();
    lse_pass_adj<R>(s.X0, s.X1, s.Ga, s.Gb, s.zy, L, 0, T);
    __syncthreads();
}

template <typename R>
__device__ R block_min_of(const R* v, int L3, R* red) {
    R mn = r_big<R>();
    for (int l = threadIdx.x; l < L3; l += blockDim.x) mn = v[l] < mn ? v[l] : mn;
    return block_reduce<R>(mn, red, true);
}

// upward sweep at one depth (deepest first), one workgroup per node
template <typename R>
__global__ __launch_bounds__(SB) void k_sp_up(BPState<R> a, int pos0) {
    extern __shared__ unsigned char sp_smem[];
    Lds<R> s(reinterpret_cast<R*>(sp_smem), a.L);
    const int L3 = a.L * a.L * a.L;
    const int c = a.t.ordered[pos0 + blockIdx.x];
    const long long cb = (long long)c * L3;
    const int c0 = a.t.child_start[c], c1 = a.t.child_start[c + 1];
    for (int l = threadIdx.x; l < L3; l += blockDim.x) {
        R h = a.cost[cb + l];
        for (int ci = c0; ci < c1; ci++) h += a.Msg[(long long)a.t.child_list[ci] * L3 + l];
        a.H[cb + l] = h;
        s.X0[l] = h;
    }
    const int p = a.t.parents[c];
    if (p < 0) {  // the root's full belief is its upward belief
        for (int l = threadIdx.x; l < L3; l += blockDim.x) a.G[cb + l] = s.X0[l];
        return;
    }
    edge_tables<R>(a, c, p, (R)1, s);
    dt3<R>(s, a.L, a.T);
    const R mn = block_min_of<R>(s.X3, L3, s.red);
    for (int l = threadIdx.x; l < L3; l += blockDim.x) a.Msg[cb + l] = s.X3[l] - mn;
}

// downward sweep at one depth >= 1 (root side first)
template <typename R>
__global__ __launch_bounds__(SB) void k_sp_down(BPState<R> a, int pos0) {
    extern __shared__ unsigned char sp_smem[];
    Lds<R> s(reinterpret_cast<R*>(sp_smem), a.L);
    const int L3 = a.L * a.L * a.L;
    const int c = a.t.ordered[pos0 + blockIdx.x];
    const int p = a.t.parents[c];
    const long long cb = (long long)c * L3, pb = (long long)p * L3;
    for (int l = threadIdx.x; l < L3; l += blockDim.x) s.X0[l] = a.G[pb + l] - a.Msg[cb + l];
    edge_tables<R>(a, c, p, (R)-1, s);
    dt3<R>(s, a.L, a.T);
    const R mn = block_min_of<R>(s.X3, L3, s.red);
    for (int l = threadIdx.x; l < L3; l += blockDim.x) a.G[cb + l] = a.H[cb + l] + (s.X3[l] - mn);
}

// Marginal of node c: P = softmax(-g_c / T) into P, expected label displacement into e[3] (u, v, w).
template <typename R>
__device__ void marginal(const BPState<R>& a, int c, R* P, R* red, R e[3]) {
    const int L = a.L, L3 = L * L * L;
    const long long cb = (long long)c * L3;
    const R mn = block_min_of<R>(a.G + cb, L3, red);
    const R iT = (R)1 / a.T;
    R z = 0, eu = 0, ev = 0, ew = 0;
    for (int l = threadIdx.x; l < L3; l += blockDim.x) {
        const R pw = r_exp<R>((mn - a.G[cb + l]) * iT);
        P[l] = pw;
        const int li = l % L, lj = (l / L) % L, lk = l / (L * L);
        z += pw;
        eu += pw * (R)(lj - a.hw);
        ev += pw * (R)(li - a.hw);
        ew += pw * (R)(lk - a.hw);
    }
    z = block_reduce<R>(z, red, false);
    eu = block_reduce<R>(eu, red, false);
    ev = block_reduce<R>(ev, red, false);
    ew = block_reduce<R>(ew, red, false);
    const R iz = (R)1 / z;
    for (int l = threadIdx.x; l < L3; l += blockDim.x) P[l] *= iz;
    __syncthreads();
    e[0] = eu * iz * a.quant;
    e[1] = ev * iz * a.quant;
    e[2] = ew * iz * a.quant;
}

template <typename R>
__global__ __launch_bounds__(SB) void k_sp_marginals(BPState<R> a) {
    extern __shared__ unsigned char sp_smem[];
    Lds<R> s(reinterpret_cast<R*>(sp_smem), a.L);
    const int c = blockIdx.x;
    R e[3];
    marginal<R>(a, c, s.X0, s.red, e);
    if (threadIdx.x == 0) {
        a.u1[c] = a.u0[c] + e[0];
        a.u1[a.N + c] = a.u0[a.N + c] + e[1];
        a.u1[2 * a.N + c] = a.u0[2 * a.N + c] + e[2];
    }
}

template <typename R>
struct BPGrad {
    const R* gu1;  // 3 x N
    R* A;          // total d/dg_c
    R* Dt;         // d/d(input of the downward transform of c)
    R* dH;         // d/dh_c == d/dcost_c
};

// backward of the marginals and of the downward sweep (deepest first):
// A_c = marginal term + sum over children k of Dt_k; Dt_c = DTdown_c^T(A_c)
template <typename R>
__global__ __launch_bounds__(SB) void k_sp_bwd_down(BPState<R> a, BPGrad<R> b, int pos0) {
    extern __shared__ unsigned char sp_smem[];
    Lds<R> s(reinterpret_cast<R*>(sp_smem), a.L);
    const int L = a.L, L3 = L * L * L;
    const int c = a.t.ordered[pos0 + blockIdx.x];
    const long long cb = (long long)c * L3;
    R e[3];
    marginal<R>(a, c, s.X0, s.red, e);
    const R gu = b.gu1[c], gv = b.gu1[a.N + c], gw = b.gu1[2 * a.N + c];
    const R iT = (R)1 / a.T;
    const int c0 = a.t.child_start[c], c1 = a.t.child_start[c + 1];
    for (int l = threadIdx.x; l < L3; l += blockDim.x) {
        const int li = l % L, lj = (l / L) % L, lk = l / (L * L);
        const R du = (R)(lj - a.hw) * a.quant - e[0];
        const R dv = (R)(li - a.hw) * a.quant - e[1];
        const R dw = (R)(lk - a.hw) * a.quant - e[2];
        R g = -(s.X0[l] * iT) * (gu * du + gv * dv + gw * dw);
        for (int ci = c0; ci < c1; ci++) g += b.Dt[(long long)a.t.child_list[ci] * L3 + l];
        b.A[cb + l] = g;
        s.Ga[l] = g;
    }
    const int p = a.t.parents[c];
    if (p < 0) return;
    __syncthreads();
    const long long pb = (long long)p * L3;
    for (int l = threadIdx.x; l < L3; l += blockDim.x) s.X0[l] = a.G[pb + l] - a.Msg[cb + l];
    edge_tables<R>(a, c, p, (R)-1, s);
    dt3<R>(s, L, a.T);
    dt3_adj<R>(s, L, a.T);
    for (int l = threadIdx.x; l < L3; l += blockDim.x) b.Dt[cb + l] = s.Gb[l];
}

// backward of the upward sweep (root side first):
// dH_c = A_c + DTup_c^T(dH_p - Dt_c); the root's dH is its A
template <typename R>
__global__ __launch_bounds__(SB) void k_sp_bwd_up(BPState<R> a, BPGrad<R> b, int pos0) {
    extern __shared__ unsigned char sp_smem[];
    Lds<R> s(reinterpret_cast<R*>(sp_smem), a.L);
    const int L = a.L, L3 = L * L * L;
    const int c = a.t.ordered[pos0 + blockIdx.x];
    const long long cb = (long long)c * L3;
    const int p = a.t.parents[c];
    if (p < 0) {
        for (int l = threadIdx.x; l < L3; l += blockDim.x) b.dH[cb + l] = b.A[cb + l];
        return;
    }
    const long long pb = (long long)p * L3;
    for (int l = threadIdx.x; l < L3; l += blockDim.x) {
        s.X0[l] = a.H[cb + l];
        s.Ga[l] = b.dH[pb + l] - b.Dt[cb + l];
    }
    edge_tables<R>(a, c, p, (R)1, s);
    dt3<R>(s, L, a.T);
    dt3_adj<R>(s, L, a.T);
    for (int l = threadIdx.x; l < L3; l += blockDim.x) b.dH[cb + l] = b.A[cb + l] + s.Gb[l];
}

template <typename R>
void set_smem(const void* fn, size_t bytes) {
    DEEDS_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
}

Dims dims_of(int m, int n, int o) { return Dims{m, n, o, (long long)m * n * o}; }

template <typename R>
BPState<R> bp_state(const SoftTree& t, const R* cost, const R* u0, R* u1, R* saved, int hw, float quant, float T) {
    const int L = 2 * hw + 1;
    const size_t L3 = (size_t)L * L * L;
    BPState<R> a{};
    a.cost = cost;
    a.H = saved;
    a.Msg = saved + (size_t)t.N * L3;
    a.G = saved + 2 * (size_t)t.N * L3;
    a.u0 = u0;
    a.u1 = u1;
    a.t = SoftTreeDev{t.d_ordered, t.d_parents, t.d_child_start, t.d_child_list, t.d_edgemst};
    a.N = t.N;
    a.L = L;
    a.hw = hw;
    a.quant = (R)quant;
    a.T = (R)T;
    return a;
}

}  // namespace

size_t soft_bp_smem(int L, size_t rsize) {
    const size_t L3 = (size_t)L * L * L;
    return (6 * L3 + 3 * (2 * (size_t)L + 1) + 16) * rsize;
}

// ---- features ---------------------------------------------------------------

template <typename R>
void soft_features(R* phi, const R* im, int m, int n, int o, int qs, R* w0, R* w1, hipStream_t s) {
    const Dims d = dims_of(m, n, o);
    const unsigned g6 = ceil_div(6 * d.sz, 256), g1 = ceil_div(d.sz, 256);
    k_sf_ssd<R><<<g6, 256, 0, s>>>(w0, im, d, qs);
    DEEDS_LAUNCH_CHECK();
    k_sf_box<R><<<g6, 256, 0, s>>>(w1, w0, d, qs, 0);
    k_sf_box<R><<<g6, 256, 0, s>>>(w0, w1, d, qs, 1);
    k_sf_box<R><<<g6, 256, 0, s>>>(w1, w0, d, qs, 2);
    DEEDS_LAUNCH_CHECK();
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
    k_sf_box<R><<<g6, 256, 0, s>>>(w1, w0, d, qs, 0);
    k_sf_box<R><<<g6, 256, 0, s>>>(w0, w1, d, qs, 1);
    k_sf_box<R><<<g6, 256, 0, s>>>(w1, w0, d, qs, 2);
    DEEDS_LAUNCH_CHECK();
    // w2 (12 per voxel) = d/dv, w0 = d/dD, then back through the box passes (w1, w0) and the SSD
    k_sf_phi_bwd<R><<<g1, 256, 0, s>>>(w2, gphi, w1, d, qs);
    k_sf_gather<R><<<g6, 256, 0, s>>>(w0, w2, d, qs);
    k_sf_box<R><<<g6, 256, 0, s>>>(w1, w0, d, qs, 2);
    k_sf_box<R><<<g6, 256, 0, s>>>(w0, w1, d, qs, 1);
    k_sf_box<R><<<g6, 256, 0, s>>>(w1, w0, d, qs, 0);
    k_sf_ssd_bwd<R><<<g1, 256, 0, s>>>(gim, w1, im, d, qs);
    DEEDS_LAUNCH_CHECK();
}

// ---- cost -------------------------------------------------------------------

template <typename R>
void soft_cost(R* cost, const R* F, const R* M, int m, int n, int o, int step, int hw, float quant, float alpha,
               hipStream_t s) {
    const CostGeom g = cost_geom(m, n, o, step, hw, quant, alpha);
    const long long total = (long long)g.m1 * g.n1 * g.o1 * g.len * g.len * g.len;
    if (total == 0) return;
    k_sc_cost<R><<<ceil_div(total, 256), 256, 0, s>>>(cost, F, M, g);
    DEEDS_LAUNCH_CHECK();
}

template <typename R>
void soft_cost_backward(R* dF, R* dM, const R* gcost, const R* F, const R* M, int m, int n, int o, int step, int hw,
                        float quant, float alpha, hipStream_t s) {
    const CostGeom g = cost_geom(m, n, o, step, hw, quant, alpha);
    const long long sz = (long long)m * n * o;
    // voxels outside every sampled block get no gradient from the cost
    k_fill<R><<<ceil_div(12 * sz, 256), 256, 0, s>>>(dF, (R)0, 12 * sz);
    DEEDS_LAUNCH_CHECK();
    const int nodes = g.m1 * g.n1 * g.o1;
    if (nodes > 0) {
        const size_t smem = (size_t)g.len * g.len * g.len * sizeof(R);
        set_smem<R>((const void*)k_sc_dF<R>, smem);
        k_sc_dF<R><<<nodes, 256, smem, s>>>(dF, gcost, F, M, g);
        DEEDS_LAUNCH_CHECK();
    }
    k_sc_dM<R><<<ceil_div(sz, 256), 256, 0, s>>>(dM, gcost, F, M, g);
    DEEDS_LAUNCH_CHECK();
}

// ---- BP ---------------------------------------------------------------------

template <typename R>
void soft_bp_forward(const SoftTree& t, const R* cost, const R* u0, R* u1, R* saved, int hw, float quant, float T,
                     hipStream_t s) {
    BPState<R> a = bp_state<R>(t, cost, u0, u1, saved, hw, quant, T);
    const size_t smem = soft_bp_smem(a.L, sizeof(R));
    set_smem<R>((const void*)k_sp_up<R>, smem);
    set_smem<R>((const void*)k_sp_down<R>, smem);
    set_smem<R>((const void*)k_sp_marginals<R>, smem);
    for (int d = t.maxd; d >= 0; d--) {
        const int b = t.depth_start[d], n = t.depth_start[d + 1] - b;
        k_sp_up<R><<<n, SB, smem, s>>>(a, b);
        DEEDS_LAUNCH_CHECK();
    }
    for (int d = 1; d <= t.maxd; d++) {
        const int b = t.depth_start[d], n = t.depth_start[d + 1] - b;
        k_sp_down<R><<<n, SB, smem, s>>>(a, b);
        DEEDS_LAUNCH_CHECK();
    }
    k_sp_marginals<R><<<t.N, SB, smem, s>>>(a);
    DEEDS_LAUNCH_CHECK();
}

template <typename R>
void soft_bp_backward(const SoftTree& t, const R* cost, const R* u0, const R* saved, const R* gu1, R* gcost,
                      R* scratch, int hw, float quant, float T, hipStream_t s) {
    BPState<R> a = bp_state<R>(t, cost, u0, nullptr, const_cast<R*>(saved), hw, quant, T);
    const size_t L3 = (size_t)a.L * a.L * a.L;
    BPGrad<R> b{gu1, scratch, scratch + (size_t)t.N * L3, gcost};
    const size_t smem = soft_bp_smem(a.L, sizeof(R));
    set_smem<R>((const void*)k_sp_bwd_down<R>, smem);
    set_smem<R>((const void*)k_sp_bwd_up<R>, smem);
    for (int d = t.maxd; d >= 0; d--) {
        const int bb = t.depth_start[d], n = t.depth_start[d + 1] - bb;
        k_sp_bwd_down<R><<<n, SB, smem, s>>>(a, b, bb);
        DEEDS_LAUNCH_CHECK();
    }
    for (int d = 0; d <= t.maxd; d++) {
        const int bb = t.depth_start[d], n = t.depth_start[d + 1] - bb;
        k_sp_bwd_up<R><<<n, SB, smem, s>>>(a, b, bb);
        DEEDS_LAUNCH_CHECK();
    }
}

#define DEEDS_SOFT_INSTANTIATE(R)                                                                                   \
    template void soft_features<R>(R*, const R*, int, int, int, int, R*, R*, hipStream_t);                          \
    template void soft_features_backward<R>(R*, const R*, const R*, int, int, int, int, R*, R*, R*, hipStream_t);   \
    template void soft_cost<R>(R*, const R*, const R*, int, int, int, int, int, float, float, hipStream_t);         \
    template void soft_cost_backward<R>(R*, R*, const R*, const R*, const R*, int, int, int, int, int, float, float, \
                                        hipStream_t);                                                               \
    template void soft_bp_forward<R>(const SoftTree&, const R*, const R*, R*, R*, int, float, float, hipStream_t);  \
    template void soft_bp_backward<R>(const SoftTree&, const R*, const R*, const R*, const R*, R*, R*, int, float,   \
                                      float, hipStream_t);

DEEDS_SOFT_INSTANTIATE(float)
DEEDS_SOFT_INSTANTIATE(double)

}  // namespace deeds
