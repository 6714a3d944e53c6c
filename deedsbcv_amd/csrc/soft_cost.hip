// Soft variant, stage 2: the SSD cost volume and its adjoint (the model is
// stated in soft.h).
//
// C[n,l] = a1 * sum over the block samples s of n of sum_ch (F[s,ch] - M[clamp(s + d_l),ch])^2,
// with the block samples, clamping and a1 of the Hamming cost (dataCostD.h:141-196).
// Features are [voxel][12]; the cost volume is node-major [node][label].
// Backward: dF at the samples of each node (a correlation of the node's cost
// gradient with the moving features); dM by gathering, for every position of
// the zero-padded box, the (sample, label) pairs displaced onto it, then
// folding the pad region onto the face voxels it clamps to.
#include "soft_common.h"

// The soft cost reproduces no reference rounding: let a*b+c contract to FMA here.
#pragma clang fp contract(fast)

namespace deeds {

namespace {

using namespace soft_dev;

constexpr int CB = 256;  // threads per cost workgroup

struct CostGeom {
    int m, n, o, m1, n1, o1, step, len, q, hw, si, sj, sk;
    int ni, nj, nk, ns;       // samples per axis and per node
    unsigned lati, latj, latk;  // bit r set when block offset r is sampled along the axis
    double scale;
};

CostGeom cost_geom(int m, int n, int o, int step, int hw, float quant, float alpha) {
    CostPlan plan = datacost_plan(m, n, o, step, hw, quant, alpha, 1);
    CostGeom g{};
    g.m = m; g.n = n; g.o = o;
    g.m1 = m / step; g.n1 = n / step; g.o1 = o / step;
    g.step = step; g.len = 2 * hw + 1; g.q = (int)quant; g.hw = hw;
    g.si = plan.skip_i; g.sj = plan.skip_j; g.sk = plan.skip_k;
    g.ni = (step + g.si - 1) / g.si;
    g.nj = (step + g.sj - 1) / g.sj;
    g.nk = (step + g.sk - 1) / g.sk;
    g.ns = g.ni * g.nj * g.nk;
    for (int r = 0; r < step && r < 32; r++) {
        if (r % g.si == 0) g.lati |= 1u << r;
        if (r % g.sj == 0) g.latj |= 1u << r;
        if (r % g.sk == 0) g.latk |= 1u << r;
    }
    g.scale = (double)plan.scale;
    return g;
}

__device__ __forceinline__ void load12(float (&d)[12], const float* __restrict__ p) {
    const float4* q = reinterpret_cast<const float4*>(p);
    const float4 a = q[0], b = q[1], c = q[2];
    d[0] = a.x; d[1] = a.y; d[2] = a.z; d[3] = a.w;
    d[4] = b.x; d[5] = b.y; d[6] = b.z; d[7] = b.w;
    d[8] = c.x; d[9] = c.y; d[10] = c.z; d[11] = c.w;
}
__device__ __forceinline__ void load12(double (&d)[12], const double* __restrict__ p) {
    const double2* q = reinterpret_cast<const double2*>(p);
#pragma unroll
    for (int e = 0; e < 6; e++) {
        const double2 a = q[e];
        d[2 * e] = a.x;
        d[2 * e + 1] = a.y;
    }
}

// Forward: one workgroup per (node, chunk of CB labels), one thread per label.
// The node's fixed-image samples are staged in LDS in sample order (k, j, i).
template <typename R>
__global__ __launch_bounds__(CB) void k_sc_cost(R* __restrict__ cost, const R* __restrict__ F,
                                                const R* __restrict__ M, CostGeom g) {
    extern __shared__ __align__(16) unsigned char sc_smem[];
    R* Fs = reinterpret_cast<R*>(sc_smem);
    const int L3 = g.len * g.len * g.len;
    const int chunks = (L3 + CB - 1) / CB;
    const int node = blockIdx.x / chunks;
    const int l = (blockIdx.x - node * chunks) * CB + threadIdx.x;
    const int y = node % g.m1, x = (node / g.m1) % g.n1, z = node / (g.m1 * g.n1);
    const long long mn = (long long)g.m * g.n;
    for (int w = threadIdx.x; w < g.ns * 12; w += CB) {
        const int ch = w % 12, s = w / 12;
        const int i = (s % g.ni) * g.si, j = ((s / g.ni) % g.nj) * g.sj, k = (s / (g.ni * g.nj)) * g.sk;
        Fs[w] = F[12 * (y * g.step + i + (long long)(x * g.step + j) * g.m + (z * g.step + k) * mn) + ch];
    }
    __syncthreads();
    if (l >= L3) return;
    const int ly = l % g.len, lx = (l / g.len) % g.len, lz = l / (g.len * g.len);
    const int dy = (ly - g.hw) * g.q, dx = (lx - g.hw) * g.q, dz = (lz - g.hw) * g.q;
    R acc = 0;
    const R* fs = Fs;
    for (int k = 0; k < g.step; k += g.sk) {
        const long long zo = clampi(z * g.step + k + dz, 0, g.o - 1) * mn;
        for (int j = 0; j < g.step; j += g.sj) {
            const long long xo = zo + (long long)clampi(x * g.step + j + dx, 0, g.n - 1) * g.m;
            for (int i = 0; i < g.step; i += g.si, fs += 12) {
                R mv[12];
                load12(mv, M + 12 * (xo + clampi(y * g.step + i + dy, 0, g.m - 1)));
#pragma unroll
                for (int ch = 0; ch < 12; ch++) {
                    const R d = fs[ch] - mv[ch];
                    acc += d * d;
                }
            }
        }
    }
    cost[(long long)node * L3 + l] = acc * (R)g.scale;
}

// dF at the samples of one node: one workgroup per node; thread (s, c) sums
// g[l] * M[clamp(s + d_l)] over label chunk c for all twelve channels, then
// the chunks are reduced in LDS: dF = 2 a1 (F * sum_l g - sum_l g M).
template <typename R>
__global__ __launch_bounds__(CB) void k_sc_dF(R* __restrict__ dF, const R* __restrict__ gc, const R* __restrict__ F,
                                              const R* __restrict__ M, CostGeom g) {
    extern __shared__ __align__(16) unsigned char sc_smem[];
    const int L = g.len, L3 = L * L * L;
    R* gl = reinterpret_cast<R*>(sc_smem);
    R* part = gl + ((L3 + 1) & ~1);  // [chunk][sample][13]
    const int node = blockIdx.x;
    for (int l = threadIdx.x; l < L3; l += CB) gl[l] = gc[(long long)node * L3 + l];
    __syncthreads();
    const int y = node % g.m1, x = (node / g.m1) % g.n1, z = node / (g.m1 * g.n1);
    const long long mn = (long long)g.m * g.n;
    const int CH = CB / g.ns;  // label chunks (ns <= CB)
    const int t = threadIdx.x;
    if (t < g.ns * CH) {
        const int s = t % g.ns, c = t / g.ns;
        const int fi = y * g.step + (s % g.ni) * g.si, fj = x * g.step + ((s / g.ni) % g.nj) * g.sj,
                  fk = z * g.step + (s / (g.ni * g.nj)) * g.sk;
        const int l0 = (int)((long long)c * L3 / CH), l1 = (int)((long long)(c + 1) * L3 / CH);
        R acc[12], G = 0;
#pragma unroll
        for (int ch = 0; ch < 12; ch++) acc[ch] = 0;
        int ly = l0 % L, lx = (l0 / L) % L, lz = l0 / (L * L);
        long long xo = clampi(fk + (lz - g.hw) * g.q, 0, g.o - 1) * mn +
                       (long long)clampi(fj + (lx - g.hw) * g.q, 0, g.n - 1) * g.m;
        for (int l = l0; l < l1; l++) {
            const R w = gl[l];
            R mv[12];
            load12(mv, M + 12 * (xo + clampi(fi + (ly - g.hw) * g.q, 0, g.m - 1)));
            G += w;
#pragma unroll
            for (int ch = 0; ch < 12; ch++) acc[ch] += w * mv[ch];
            if (++ly == L) {
                ly = 0;
                if (++lx == L) {
                    lx = 0;
                    lz++;
                }
                xo = clampi(fk + (lz - g.hw) * g.q, 0, g.o - 1) * mn +
                     (long long)clampi(fj + (lx - g.hw) * g.q, 0, g.n - 1) * g.m;
            }
        }
        R* pp = part + 13 * (c * g.ns + s);
#pragma unroll
        for (int ch = 0; ch < 12; ch++) pp[ch] = acc[ch];
        pp[12] = G;
    }
    __syncthreads();
    for (int w = t; w < g.ns * 12; w += CB) {
        const int s = w / 12, ch = w % 12;
        R a = 0, G = 0;
        for (int c = 0; c < CH; c++) {
            a += part[13 * (c * g.ns + s) + ch];
            G += part[13 * (c * g.ns + s) + 12];
        }
        const int fi = y * g.step + (s % g.ni) * g.si, fj = x * g.step + ((s / g.ni) % g.nj) * g.sj,
                  fk = z * g.step + (s / (g.ni * g.nj)) * g.sk;
        const long long fidx = 12 * (fi + (long long)fj * g.m + fk * mn) + ch;
        dF[fidx] = (R)(2.0 * g.scale) * (F[fidx] * G - a);
    }
}

// Position s = div * step + mod of a sample coordinate that decreases by q per label.
struct Walk {
    int s, div, mod;
    __device__ Walk(int s0, int step) : s(s0), div(s0 / step), mod(s0 - (s0 / step) * step) {}
    __device__ __forceinline__ void next(int q, int step) {
        s -= q;
        mod -= q;
        while (mod < 0) {
            mod += step;
            div--;
        }
    }
    __device__ __forceinline__ bool ok(int cells, unsigned lat) const { return div < cells && ((lat >> mod) & 1u); }
};

// dM at a moving voxel t (one thread per voxel, all twelve channels): every
// (sample s, label l) whose clamped displaced position is t contributes
// 2 a1 g[n(s), l] (M_t - F_s).
template <typename R>
struct DMAcc {
    R g = 0, f[12];
    __device__ DMAcc() {
#pragma unroll
        for (int ch = 0; ch < 12; ch++) f[ch] = 0;
    }
    __device__ __forceinline__ void add(const R* __restrict__ gc, const R* __restrict__ F, int L3, int node, int l,
                                        long long fvox) {
        const R w = gc[(long long)node * L3 + l];
        R fv[12];
        load12(fv, F + 12 * fvox);
        g += w;
#pragma unroll
        for (int ch = 0; ch < 12; ch++) f[ch] += w * fv[ch];
    }
    __device__ __forceinline__ void store(R* __restrict__ dM, const R* __restrict__ M, long long t, double scale) {
        R mv[12];
        load12(mv, M + 12 * t);
#pragma unroll
        for (int ch = 0; ch < 12; ch++) dM[12 * t + ch] = (R)(2.0 * scale) * (mv[ch] * g - f[ch]);
    }
};

// Pull over the padded box: every position P = s + d_l of a (sample, label)
// pair lies in [-pad, len-1+pad] per axis (pad = q*hw), and for a padded
// position P each label has at most one sample per axis, s = P - d_l, walked
// without divisions.  Voxels strictly inside the volume receive exactly the
// pairs at their own position and write dM; positions on or beyond a face
// write their sums (g, g*F) to `part` for k_sc_dM_fold, since the clamping
// sends all of them to the face voxel they clamp to.
template <typename R>
__global__ __launch_bounds__(CB) void k_sc_dM_pull(R* __restrict__ dM, R* __restrict__ part, const R* __restrict__ gc,
                                                   const R* __restrict__ F, const R* __restrict__ M, CostGeom g) {
    const int P = g.q * g.hw;
    const int mp = g.m + 2 * P, np = g.n + 2 * P, op = g.o + 2 * P;
    const long long e = (long long)blockIdx.x * CB + threadIdx.x;
    if (e >= (long long)mp * np * op) return;
    const int ip = (int)(e % mp), jp = (int)((e / mp) % np), kp = (int)(e / ((long long)mp * np));
    const int L = g.len, L3 = L * L * L;
    const long long mn = (long long)g.m * g.n;
    DMAcc<R> acc;
    const Walk wi0(ip, g.step);  // s = (ip - P) - d_l = ip - q*l
    Walk wk(kp, g.step);
    for (int lz = 0; lz < L && wk.s >= 0; lz++, wk.next(g.q, g.step)) {
        if (!wk.ok(g.o1, g.latk)) continue;
        Walk wj(jp, g.step);
        for (int lx = 0; lx < L && wj.s >= 0; lx++, wj.next(g.q, g.step)) {
            if (!wj.ok(g.n1, g.latj)) continue;
            const int nzx = (wk.div * g.n1 + wj.div) * g.m1;
            const long long fzx = (long long)wk.s * mn + (long long)wj.s * g.m;
            const int lzx = (lz * L + lx) * L;
            Walk wi = wi0;
            for (int ly = 0; ly < L && wi.s >= 0; ly++, wi.next(g.q, g.step)) {
                if (wi.ok(g.m1, g.lati)) acc.add(gc, F, L3, nzx + wi.div, lzx + ly, fzx + wi.s);
            }
        }
    }
    const int ti = ip - P, tj = jp - P, tk = kp - P;
    if (ti >= 1 && ti <= g.m - 2 && tj >= 1 && tj <= g.n - 2 && tk >= 1 && tk <= g.o - 2) {
        acc.store(dM, M, ti + (long long)tj * g.m + tk * mn, g.scale);
    } else {
        R* pp = part + 13 * e;
#pragma unroll
        for (int ch = 0; ch < 12; ch++) pp[ch] = acc.f[ch];
        pp[12] = acc.g;
    }
}

// dM of the face voxels: the sums of every padded position that clamps to
// the voxel (a line, square or cube of the pad region beyond the faces).
// Faces are enumerated without overlap: both i-faces whole, then the
// j-faces and k-faces without the voxels already covered.
template <typename R>
__global__ __launch_bounds__(CB) void k_sc_dM_fold(R* __restrict__ dM, const R* __restrict__ part,
                                                   const R* __restrict__ M, CostGeom g) {
    const long long e0 = (long long)blockIdx.x * CB + threadIdx.x;
    const long long mi = g.m > 2 ? g.m - 2 : 0, ni = g.n > 2 ? g.n - 2 : 0;
    const long long ci = 2LL * g.n * g.o, cj = 2 * mi * g.o, ck = 2 * mi * ni;
    long long e = e0;
    int ti, tj, tk;
    if (e < ci) {
        const long long half = (long long)g.n * g.o;
        const int side = (int)(e / half);
        e -= side * half;
        ti = side ? g.m - 1 : 0;
        tj = (int)(e % g.n);
        tk = (int)(e / g.n);
    } else if ((e -= ci) < cj) {
        const long long half = mi * g.o;
        const int side = (int)(e / half);
        e -= side * half;
        tj = side ? g.n - 1 : 0;
        ti = 1 + (int)(e % mi);
        tk = (int)(e / mi);
    } else if ((e -= cj) < ck) {
        const long long half = mi * ni;
        const int side = (int)(e / half);
        e -= side * half;
        tk = side ? g.o - 1 : 0;
        ti = 1 + (int)(e % mi);
        tj = 1 + (int)(e / mi);
    } else {
        return;
    }
    const int P = g.q * g.hw;
    const int mp = g.m + 2 * P, np = g.n + 2 * P;
    // padded index ranges that clamp to (ti, tj, tk)
    const int i0 = ti == 0 ? 0 : ti + P, i1 = ti == g.m - 1 ? g.m - 1 + 2 * P : ti + P;
    const int j0 = tj == 0 ? 0 : tj + P, j1 = tj == g.n - 1 ? g.n - 1 + 2 * P : tj + P;
    const int k0 = tk == 0 ? 0 : tk + P, k1 = tk == g.o - 1 ? g.o - 1 + 2 * P : tk + P;
    DMAcc<R> acc;
    for (int kp = k0; kp <= k1; kp++)
        for (int jp = j0; jp <= j1; jp++)
            for (int ip = i0; ip <= i1; ip++) {
                const R* pp = part + 13 * (ip + (long long)jp * mp + (long long)kp * mp * np);
#pragma unroll
                for (int ch = 0; ch < 12; ch++) acc.f[ch] += pp[ch];
                acc.g += pp[12];
            }
    acc.store(dM, M, ti + (long long)tj * g.m + (long long)tk * g.m * g.n, g.scale);
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
    const long long nodes = (long long)g.m1 * g.n1 * g.o1;
    const int L3 = g.len * g.len * g.len;
    if (nodes == 0) return;
    const long long blocks = nodes * ((L3 + CB - 1) / CB);
    if (blocks > 0x7fffffffLL) fail(DEEDS_ERR_ARG, "soft cost volume too large");
    const size_t smem = (size_t)g.ns * 12 * sizeof(R);
    set_smem((const void*)k_sc_cost<R>, smem);
    k_sc_cost<R><<<(unsigned)blocks, CB, smem, s>>>(cost, F, M, g);
    DEEDS_LAUNCH_CHECK();
}

template <typename R>
void soft_cost_backward(R* dF, R* dM, const R* gcost, const R* F, const R* M, int m, int n, int o, int step, int hw,
                        float quant, float alpha, R* part, hipStream_t s) {
    const CostGeom g = cost_geom(m, n, o, step, hw, quant, alpha);
    const long long sz = (long long)m * n * o;
    // voxels outside every sampled block get no gradient from the cost
    k_fill<R><<<ceil_div(12 * sz, 256), 256, 0, s>>>(dF, (R)0, 12 * sz);
    DEEDS_LAUNCH_CHECK();
    const int nodes = g.m1 * g.n1 * g.o1;
    if (nodes > 0) {
        const size_t L3 = (size_t)g.len * g.len * g.len;
        const size_t smem = (((L3 + 1) & ~(size_t)1) + (size_t)(CB / g.ns) * g.ns * 13) * sizeof(R);
        set_smem((const void*)k_sc_dF<R>, smem);
        k_sc_dF<R><<<nodes, CB, smem, s>>>(dF, gcost, F, M, g);
        DEEDS_LAUNCH_CHECK();
    }
    // pull over the padded box (interior voxels finish there), then fold the pad region onto the faces
    const long long P = (long long)g.q * g.hw;
    const long long padded = (m + 2 * P) * (n + 2 * P) * (o + 2 * P);
    k_sc_dM_pull<R><<<ceil_div(padded, CB), CB, 0, s>>>(dM, part, gcost, F, M, g);
    DEEDS_LAUNCH_CHECK();
    const long long mi = m > 2 ? m - 2 : 0, ni = n > 2 ? n - 2 : 0;
    const long long faces = 2LL * n * o + 2 * mi * o + 2 * mi * ni;
    k_sc_dM_fold<R><<<ceil_div(faces, CB), CB, 0, s>>>(dM, part, M, g);
    DEEDS_LAUNCH_CHECK();
}

#define DEEDS_SOFT_COST(R)                                                                                     \
    template void soft_cost<R>(R*, const R*, const R*, int, int, int, int, int, float, float, hipStream_t);         \
    template void soft_cost_backward<R>(R*, R*, const R*, const R*, const R*, int, int, int, int, int, float, float, \
                                        R*, hipStream_t);

DEEDS_SOFT_COST(float)
DEEDS_SOFT_COST(double)

}  // namespace deeds
