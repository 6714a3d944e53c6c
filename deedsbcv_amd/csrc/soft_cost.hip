// Soft variant, stage 2: the SSD cost volume and its adjoint (the model is
// stated in soft.h).
//
// C[n,l] = a1 * sum over the block samples s of n of sum_ch (F[s,ch] - M[clamp(s + d_l),ch])^2,
// with the block samples, clamping and a1 of the Hamming cost (dataCostD.h:141-196).
// Features are [voxel][12]; the cost volume is node-major [node][label].
// Backward: dF at the samples of each node (a correlation of the node's cost
// gradient with the moving features), dM by gathering, for each moving
// voxel, every (sample, label) pair whose clamped displaced position it is.
#include "soft_common.h"

namespace deeds {

namespace {

using namespace soft_dev;

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

}  // namespace

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
        set_smem((const void*)k_sc_dF<R>, smem);
        k_sc_dF<R><<<nodes, 256, smem, s>>>(dF, gcost, F, M, g);
        DEEDS_LAUNCH_CHECK();
    }
    k_sc_dM<R><<<ceil_div(sz, 256), 256, 0, s>>>(dM, gcost, F, M, g);
    DEEDS_LAUNCH_CHECK();
}

#define DEEDS_SOFT_COST(R)                                                                                     \
    template void soft_cost<R>(R*, const R*, const R*, int, int, int, int, int, float, float, hipStream_t);         \
    template void soft_cost_backward<R>(R*, R*, const R*, const R*, const R*, int, int, int, int, int, float, float, \
                                        hipStream_t);

DEEDS_SOFT_COST(float)
DEEDS_SOFT_COST(double)

}  // namespace deeds
