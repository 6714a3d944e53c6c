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
    // dM[t] = k (M_t g - f)
    __device__ __forceinline__ void store(R* __restrict__ dM, const R* __restrict__ M, long long t, R k) {
        R mv[12];
        load12(mv, M + 12 * t);
#pragma unroll
        for (int ch = 0; ch < 12; ch++) dM[12 * t + ch] = k * (mv[ch] * g - f[ch]);
    }
};

// Pull over the padded box: every position P = s + d_l of a (sample, label)
// pair lies in [-pad, len-1+pad] per axis (pad = q*hw), and for a padded
// position P each label has at most one sample per axis, s = P - d_l, walked
// without divisions.  Voxels strictly inside the volume receive exactly the
// pairs at their own position and write dM; positions on or beyond a face
// write their sums (g, g*F), times 2 a1, to `part` for the folds, since
// the clamping sends all of them to the face voxel they clamp to.
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
        acc.store(dM, M, ti + (long long)tj * g.m + tk * mn, (R)(2.0 * g.scale));
    } else {
        R* pp = part + 13 * e;
        const R k = (R)(2.0 * g.scale);
#pragma unroll
        for (int ch = 0; ch < 12; ch++) pp[ch] = k * acc.f[ch];
        pp[12] = k * acc.g;
    }
}

// dM of the face voxels: the sums (already times 2 a1) of every padded
// position that clamps to the voxel (a line, square or cube of the pad
// region beyond the faces).  The clamp acts on each axis alone, so the pad
// is folded one axis at a time, in place: along i for every row, then along
// j for the positions inside [0, m) along i, then along k inside [0, m) x
// [0, n); each fold adds a pad's sums onto the face slot next to it (one
// thread per channel and line: both sides in order, so a one-voxel axis
// folds both pads into its single slot).  Face voxels then read their slot.
template <typename R, int AX>
__global__ __launch_bounds__(CB) void k_sc_dM_fold(R* __restrict__ part, CostGeom g) {
    const int P = g.q * g.hw;
    const int mp = g.m + 2 * P, np = g.n + 2 * P, op = g.o + 2 * P;
    // lines: AX 0 over (jp, kp) of the padded box, AX 1 over (i, kp), AX 2 over (i, j) inside the volume
    const int w0 = AX == 0 ? np : g.m;
    const int w1 = AX == 0 ? op : (AX == 1 ? op : g.n);
    const long long e = (long long)blockIdx.x * CB + threadIdx.x;
    if (e >= 13LL * w0 * w1) return;
    const int ch = (int)(e % 13);
    const long long r = e / 13;
    const int a = (int)(r % w0), b = (int)(r / w0);
    long long base, stride;
    int len;
    if (AX == 0) {
        base = (long long)a * mp + (long long)b * mp * np;
        stride = 1;
        len = g.m;
    } else if (AX == 1) {
        base = (a + P) + (long long)b * mp * np;
        stride = mp;
        len = g.n;
    } else {
        base = (a + P) + (long long)(b + P) * mp;
        stride = (long long)mp * np;
        len = g.o;
    }
    R* q = part + 13 * base + ch;
    R lo = 0, hi = 0;
    for (int t = 0; t < P; t++) {
        lo += q[13 * stride * t];
        hi += q[13 * stride * (len + P + t)];
    }
    q[13 * stride * P] += lo;
    q[13 * stride * (len - 1 + P)] += hi;
}

// Face voxels enumerated without overlap (both i-faces whole, then the
// j-faces and k-faces without the voxels already covered) write dM from
// their folded slot.
template <typename R>
__global__ __launch_bounds__(CB) void k_sc_dM_faces(R* __restrict__ dM, const R* __restrict__ part,
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
    const R* pp = part + 13 * ((ti + P) + (long long)(tj + P) * mp + (long long)(tk + P) * mp * np);
    DMAcc<R> acc;
#pragma unroll
    for (int ch = 0; ch < 12; ch++) acc.f[ch] = pp[ch];
    acc.g = pp[12];
    acc.store(dM, M, ti + (long long)tj * g.m + (long long)tk * g.m * g.n, (R)1);
}

// ---- row forms of the adjoint ----------------------------------------------------
// With every voxel of a block sampled along i (skip_i 1), both adjoints are
// correlations along i of a source row with the nodes' cost-gradient slices:
//   dM: output t (padded box) gathers g[n(s), l] F[s] over s = t - d_l;
//   dF: output s (a sample) gathers g[n(s), l] M[clamp(s + d_l)].
// A workgroup is one wavefront that owns 64 consecutive positions along i,
// one per lane, and J output rows along j that lie Q apart, so that one
// source row pairs with every output row of the group, each through its own
// label lx.  With output row jj and source row sr (both Q apart, the source
// rows starting HW*Q before the first output row) the label is lx = jj - sr
// + 2 HW for dM and lx = sr - jj for dF, valid for jj in [sr - 2 HW, sr].
// Per step (lz, source row) the wavefront holds in LDS the source row (zero
// beyond the volume for dM, clamped for dF) and the gradient slab
// g[jj][ly][node slot] of the nodes under its window (zero where a row has
// no label for the source row or a slot no node); each lane then adds, per
// label ly, one source voxel (lanes 48 bytes apart: no bank conflicts) into
// its J accumulators.  The next step's row and slab are loaded into
// registers while the current step computes (the gradient is read from its
// label-major copy gT[l][node]: the slots of one (jj, ly) are consecutive).

constexpr int floor_div(int a, int b) { return a >= 0 ? a / b : -((-a + b - 1) / b); }

constexpr int kRowJ = 2;    // output rows per workgroup of the row forms
constexpr int kRowNT = 64;  // lanes per workgroup of the row forms

// gT: the cost gradient gc [N][L3] as [L3][N]: row l, column node.
// 64 x 64 tiles through LDS.
template <typename R>
__global__ __launch_bounds__(256) void k_sc_grad_t(R* __restrict__ gT, const R* __restrict__ gc, int N, int L3) {
    __shared__ R tile[64][65];
    const int l0 = blockIdx.x * 64, n0 = blockIdx.y * 64;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 16; k++) {
        const int r = ty + 4 * k, node = n0 + r, l = l0 + tx;
        tile[r][tx] = node < N && l < L3 ? gc[(long long)node * L3 + l] : (R)0;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; k++) {
        const int r = ty + 4 * k, l = l0 + r, node = n0 + tx;
        if (l < L3 && node < N) gT[(long long)l * N + node] = tile[tx][r];
    }
}

// The adjoint sums of one lane: J rows of twelve channels and the gradient sum.
template <typename R, int J>
struct RowSums {
    R acc[J][12], G[J];
    __device__ RowSums() {
#pragma unroll
        for (int jj = 0; jj < J; jj++) {
            G[jj] = 0;
#pragma unroll
            for (int ch = 0; ch < 12; ch++) acc[jj][ch] = 0;
        }
    }
    __device__ __forceinline__ void add(int jj, R w, const R (&fv)[12]) {
        G[jj] += w;
#pragma unroll
        for (int ch = 0; ch < 12; ch++) acc[jj][ch] += w * fv[ch];
    }
};

// One step's data of a row-form wavefront: the source row ((kRowNT + 2P)
// voxels of twelve channels) and the gradient slab [J][LEN][NS], loaded
// into registers (load) and then written to LDS (store).
template <typename R, int P, int J, int LEN, int NS>
struct RowStage {
    static constexpr int V = 12 * (int)sizeof(R) / 16;  // 16-byte chunks per voxel
    static constexpr int NF = (kRowNT + 2 * P) * V, PF = (NF + kRowNT - 1) / kRowNT;
    static constexpr int NG = J * LEN * NS, PG = (NG + kRowNT - 1) / kRowNT;
    static constexpr size_t ROW_BYTES = (size_t)NF * 16;
    uint4 f[PF];
    R w[PG];
    // row: source row index of the [o][n][m][12] volume; p0: its first position;
    // zero positions outside [0, m) (ZERO) or clamp them; slab(jj, ly, slot)
    // returns the gradient (a read of gT or zero)
    template <bool ZERO, typename Slab>
    __device__ __forceinline__ void load(const R* __restrict__ src, long long row, int m, int p0, Slab slab) {
        const int lane = threadIdx.x;
#pragma unroll
        for (int k = 0; k < PF; k++) {
            const int e = lane + k * kRowNT, v = e / V, c = e - v * V;
            int p = p0 + v;
            f[k] = make_uint4(0u, 0u, 0u, 0u);
            if (e < NF && (!ZERO || (p >= 0 && p < m))) {
                p = clampi(p, 0, m - 1);
                f[k] = reinterpret_cast<const uint4*>(src + 12 * (row * m + p))[c];
            }
        }
#pragma unroll
        for (int k = 0; k < PG; k++) {
            const int e = lane + k * kRowNT, slot = e % NS, r = e / NS, ly = r % LEN, jj = r / LEN;
            w[k] = e < NG ? slab(jj, ly, slot) : (R)0;
        }
    }
    __device__ __forceinline__ void store(unsigned char* lds) const {
        const int lane = threadIdx.x;
#pragma unroll
        for (int k = 0; k < PF; k++) {
            const int e = lane + k * kRowNT;
            if (e < NF) reinterpret_cast<uint4*>(lds)[e] = f[k];
        }
        R* gs = reinterpret_cast<R*>(lds + ROW_BYTES);
#pragma unroll
        for (int k = 0; k < PG; k++) {
            const int e = lane + k * kRowNT;
            if (e < NG) gs[e] = w[k];
        }
    }
};

// dM: workgroup (i tile, row group, kp); output rows tj = tjb + Q jj of the
// padded box [-P, n + P), tk = kp - P.  Interior voxels write dM, positions
// on or beyond a face their sums, times 2 a1, to `part` for the folds
// (every padded position is written exactly once).
template <typename R, int STEP, int LEN, int Q, int J>
__global__ __launch_bounds__(kRowNT) void k_sc_dM_rows(R* __restrict__ dM, R* __restrict__ part,
                                                       const R* __restrict__ gT, const R* __restrict__ F,
                                                       const R* __restrict__ M, CostGeom g, int itiles, int groups) {
    constexpr int HW = (LEN - 1) / 2, P = HW * Q, L2 = LEN * LEN, SR = J + 2 * HW, STEPS = LEN * SR;
    constexpr int NS = (kRowNT + 2 * P) / STEP + 2;  // node slots under the source window
    using Stage = RowStage<R, P, J, LEN, NS>;
    __shared__ __align__(16) unsigned char lds[Stage::ROW_BYTES + Stage::NG * sizeof(R)];
    const R* Fs = reinterpret_cast<const R*>(lds);
    const R* gs = reinterpret_cast<const R*>(lds + Stage::ROW_BYTES);
    int b = blockIdx.x;
    const int it = b % itiles;
    b /= itiles;
    const int grp = b % groups, kp = b / groups;
    const int tk = kp - P, tjb = -P + grp % Q + Q * J * (grp / Q);
    const int t_lo = -P + it * kRowNT, lane = threadIdx.x, ti = t_lo + lane;
    const int w0 = t_lo - P, nlo = floor_div(w0, STEP), off0 = w0 - nlo * STEP;  // window [w0, w0 + 64 + 2P)
    const int N = g.m1 * g.n1 * g.o1;
    // steps (lz, source row) with a sample plane and a sample row
    auto valid = [&](int pr) {
        const int lz = pr / SR, sr = pr - lz * SR;
        const int s_k = tk - Q * (lz - HW), s_j = tjb - P + Q * sr;
        return s_k >= 0 && s_k < g.o1 * STEP && (s_k % STEP) % g.sk == 0 && s_j >= 0 && s_j < g.n1 * STEP &&
               (s_j % STEP) % g.sj == 0;
    };
    auto next = [&](int pr) {
        while (pr < STEPS && !valid(pr)) pr++;
        return pr;
    };
    auto fetch = [&](Stage& st, int pr) {
        const int lz = pr / SR, sr = pr - lz * SR;
        const int s_k = tk - Q * (lz - HW), s_j = tjb - P + Q * sr;
        const int nrow = ((s_k / STEP) * g.n1 + s_j / STEP) * g.m1;
        st.template load<true>(F, (long long)s_k * g.n + s_j, g.m, w0, [&](int jj, int ly, int slot) {
            const int lx = jj - sr + 2 * HW, node_i = nlo + slot;
            return lx >= 0 && lx < LEN && node_i >= 0 && node_i < g.m1
                       ? gT[(long long)(lz * L2 + lx * LEN + ly) * N + nrow + node_i] : (R)0;
        });
    };
    RowSums<R, J> S;
    Stage st;
    int pr = next(0);
    if (pr < STEPS) fetch(st, pr);
    while (pr < STEPS) {
        __syncthreads();
        st.store(lds);
        __syncthreads();
        const int nx = next(pr + 1);
        if (nx < STEPS) fetch(st, nx);  // in flight while this step computes
#pragma unroll
        for (int ly = 0; ly < LEN; ly++) {
            // source position ti + P - Q ly: window entry lane + 2P - Q ly, node slot below
            const int e = lane + 2 * P - Q * ly;
            R fv[12];
            load12(fv, Fs + 12 * e);
            const R* gl = gs + ly * NS + (e + off0) / STEP;
#pragma unroll
            for (int jj = 0; jj < J; jj++) S.add(jj, gl[jj * LEN * NS], fv);
            __builtin_amdgcn_sched_barrier(0);  // one label's reads at a time (registers)
        }
        pr = nx;
    }
    if (ti >= g.m + P) return;
    const int mp = g.m + 2 * P, np = g.n + 2 * P;
    const long long mn = (long long)g.m * g.n;
    const bool inner_ik = ti >= 1 && ti <= g.m - 2 && tk >= 1 && tk <= g.o - 2;
    const R k2 = (R)(2.0 * g.scale);
#pragma unroll
    for (int jj = 0; jj < J; jj++) {
        const int tj = tjb + Q * jj;
        if (tj >= g.n + P) continue;
        if (inner_ik && tj >= 1 && tj <= g.n - 2) {
            R mv[12];
            const long long v = ti + (long long)tj * g.m + tk * mn;
            load12(mv, M + 12 * v);
#pragma unroll
            for (int ch = 0; ch < 12; ch++) dM[12 * v + ch] = k2 * (mv[ch] * S.G[jj] - S.acc[jj][ch]);
        } else {
            R* pp = part + 13 * ((ti + P) + (long long)(tj + P) * mp + (long long)kp * mp * np);
#pragma unroll
            for (int ch = 0; ch < 12; ch++) pp[ch] = k2 * S.acc[jj][ch];
            pp[12] = k2 * S.G[jj];
        }
    }
}

// dF: workgroup (i tile, row group, s_k) over the sampled positions
// [0, m1 STEP) x [0, n1 STEP) x [0, o1 STEP); output rows s_j = sjb + Q jj
// that are sample rows write dF = 2 a1 (F G - sum g M) (the rest of dF was
// cleared before).  Lanes past the last sample write nothing.
template <typename R, int STEP, int LEN, int Q, int J>
__global__ __launch_bounds__(kRowNT) void k_sc_dF_rows(R* __restrict__ dF, const R* __restrict__ gT,
                                                       const R* __restrict__ F, const R* __restrict__ M, CostGeom g,
                                                       int itiles, int groups) {
    constexpr int HW = (LEN - 1) / 2, P = HW * Q, L2 = LEN * LEN, SR = J + 2 * HW, STEPS = LEN * SR;
    constexpr int NS = kRowNT / STEP + 2;  // node slots under the lanes
    using Stage = RowStage<R, P, J, LEN, NS>;
    __shared__ __align__(16) unsigned char lds[Stage::ROW_BYTES + Stage::NG * sizeof(R)];
    const R* Ms = reinterpret_cast<const R*>(lds);
    const R* gs = reinterpret_cast<const R*>(lds + Stage::ROW_BYTES);
    int b = blockIdx.x;
    const int it = b % itiles;
    b /= itiles;
    const int grp = b % groups, s_k = b / groups;
    if ((s_k % STEP) % g.sk) return;  // not a sample plane
    const int sjb = grp % Q + Q * J * (grp / Q);
    unsigned rows = 0;  // bit jj: output row jj is a sample row
#pragma unroll
    for (int jj = 0; jj < J; jj++) {
        const int s_j = sjb + Q * jj;
        if (s_j < g.n1 * STEP && (s_j % STEP) % g.sj == 0) rows |= 1u << jj;
    }
    if (!rows) return;
    const int s_lo = it * kRowNT, lane = threadIdx.x, s_i = s_lo + lane;
    const int nlo = s_lo / STEP, slot = (lane + s_lo - nlo * STEP) / STEP;
    const int N = g.m1 * g.n1 * g.o1;
    const int zrow = (s_k / STEP) * g.n1;
    // steps (lz, source row) that pair with a sample row
    auto valid = [&](int pr) {
        const int sr = pr % SR, jlo = max(0, sr - 2 * HW), jhi = min(J - 1, sr);
        return ((rows >> jlo) & ((2u << (jhi - jlo)) - 1u)) != 0;
    };
    auto next = [&](int pr) {
        while (pr < STEPS && !valid(pr)) pr++;
        return pr;
    };
    auto fetch = [&](Stage& st, int pr) {
        const int lz = pr / SR, sr = pr - lz * SR;
        const int m_k = clampi(s_k + Q * (lz - HW), 0, g.o - 1), m_j = clampi(sjb - P + Q * sr, 0, g.n - 1);
        st.template load<false>(M, (long long)m_k * g.n + m_j, g.m, s_lo - P, [&](int jj, int ly, int sl) {
            const int lx = sr - jj, node_i = nlo + sl;
            const int nrow = (zrow + (sjb + Q * jj) / STEP) * g.m1;
            return lx >= 0 && lx < LEN && ((rows >> jj) & 1u) && node_i < g.m1
                       ? gT[(long long)(lz * L2 + lx * LEN + ly) * N + nrow + node_i] : (R)0;
        });
    };
    RowSums<R, J> S;
    Stage st;
    int pr = next(0);
    if (pr < STEPS) fetch(st, pr);
    while (pr < STEPS) {
        __syncthreads();
        st.store(lds);
        __syncthreads();
        const int nx = next(pr + 1);
        if (nx < STEPS) fetch(st, nx);  // in flight while this step computes
#pragma unroll
        for (int ly = 0; ly < LEN; ly++) {
            // moving position s_i + Q (ly - HW): window entry lane + Q ly
            R mv[12];
            load12(mv, Ms + 12 * (lane + Q * ly));
            const R* gl = gs + ly * NS + slot;
#pragma unroll
            for (int jj = 0; jj < J; jj++) S.add(jj, gl[jj * LEN * NS], mv);
            __builtin_amdgcn_sched_barrier(0);  // one label's reads at a time (registers)
        }
        pr = nx;
    }
    if (s_i >= g.m1 * STEP) return;
    const long long mn = (long long)g.m * g.n;
    const R k2 = (R)(2.0 * g.scale);
#pragma unroll
    for (int jj = 0; jj < J; jj++) {
        if (!((rows >> jj) & 1u)) continue;
        const long long fidx = 12 * ((long long)s_k * mn + (long long)(sjb + Q * jj) * g.m + s_i);
        R fv[12];
        load12(fv, F + fidx);
#pragma unroll
        for (int ch = 0; ch < 12; ch++) dF[fidx + ch] = k2 * (fv[ch] * S.G[jj] - S.acc[jj][ch]);
    }
}

// ---- per-node forms ---------------------------------------------------------------
// A node's moving box (the voxels its samples and labels reach, on the lattice
// the sample skips and the quantisation step span) is staged in LDS one
// k-plane at a time with coalesced 16-byte loads; a workgroup covers NY
// consecutive nodes along i, whose boxes overlap in all but NY*STEP voxels of
// each row.  Workgroups run in XCD-contiguous ranges, so neighbouring boxes
// meet in one L2.

constexpr int kNodeMaxNk = 4;  // sample planes per node (k) the per-node kernels hold

struct BoxGeom {
    int dk, dj;       // plane / row step of the used lattice, gcd(skip, quantisation)
    int nw, nv, np;   // planes, rows per plane, voxels per row of the staged box
    int ny, nyb;      // nodes per workgroup along i, workgroups along i
    int per;          // workgroups per XCD
};

__device__ __forceinline__ int xcd_tile(int b, int per) { return (b & 7) * per + (b >> 3); }

// Stages plane koff of the box of nodes y0.. into Ms ([nv][np][12]).
template <typename R>
__device__ __forceinline__ void stage_plane(R* __restrict__ Ms, const R* __restrict__ M, const CostGeom& g,
                                            const BoxGeom& b, int kb, int jb, int ib, int koff) {
    constexpr int V = 12 * (int)sizeof(R) / 16;  // 16-byte chunks per voxel
    const long long mn = (long long)g.m * g.n;
    const long long ko = (long long)clampi(kb + koff, 0, g.o - 1) * mn;
    const int n = b.nv * b.np * V;
    for (int e = threadIdx.x; e < n; e += blockDim.x) {
        const int v = e / (b.np * V), r = e - v * (b.np * V), p = r / V, c = r - p * V;
        const long long vox = ko + (long long)clampi(jb + v * b.dj, 0, g.n - 1) * g.m + clampi(ib + p, 0, g.m - 1);
        reinterpret_cast<uint4*>(Ms)[(v * b.np + p) * V + c] = reinterpret_cast<const uint4*>(M + 12 * vox)[c];
    }
}

// The label lz with sk * skip_k + lz * Q == koff for each sample plane sk, or -1.
__device__ __forceinline__ bool plane_labels(const CostGeom& g, int koff, int (&lz)[kNodeMaxNk]) {
    bool any = false;
#pragma unroll
    for (int sk = 0; sk < kNodeMaxNk; sk++) {
        const int r = koff - sk * g.sk;
        lz[sk] = sk < g.nk && r >= 0 && r % g.q == 0 && r / g.q < g.len ? r / g.q : -1;
        any |= lz[sk] >= 0;
    }
    return any;
}

// Cost: the waves of node slot ny hold its labels (lx, ly); each plane adds
// the partial sums of its (sample plane, lz) pairs to the node's label sums
// in LDS (a slot per (lz, thread): no two threads share one).  The node's
// fixed samples are the same for the whole wave (scalar loads).
template <typename R>
__global__ __launch_bounds__(1024) void k_sc_cost_node(R* __restrict__ cost, const R* __restrict__ F,
                                                       const R* __restrict__ M, CostGeom g, BoxGeom b, int tpn) {
    extern __shared__ __align__(16) unsigned char sc_smem[];
    const int tile = xcd_tile((int)blockIdx.x, b.per);
    if (tile >= b.nyb * g.n1 * g.o1) return;
    const int L = g.len, L2 = L * L, L3 = L2 * L;
    R* Ms = reinterpret_cast<R*>(sc_smem);
    R* acc = Ms + (size_t)b.nv * b.np * 12;  // [ny][L3]
    const int yb = tile % b.nyb, x = (tile / b.nyb) % g.n1, z = tile / (b.nyb * g.n1);
    // tpn is a multiple of the wavefront size: the slot is the same over a wave
    const int slot = __builtin_amdgcn_readfirstlane(threadIdx.x / tpn), t = threadIdx.x - slot * tpn;
    const int y0 = yb * b.ny, y = y0 + slot;
    const bool active = t < L2 && y < g.m1;
    const int lx = t / L, ly = t - lx * L;
    for (int l = threadIdx.x; l < b.ny * L3; l += blockDim.x) acc[l] = 0;
    const long long mn = (long long)g.m * g.n;
    const int P = g.hw * g.q;
    const int kb = z * g.step - P, jb = x * g.step - P, ib = y0 * g.step - P;
    // the node's samples, uniform over the wave: read through the constant
    // address space, so the loads are scalar (F is not written in the kernel)
    using CR = const R __attribute__((address_space(4)));
    CR* Fn = (CR*)(F + 12 * ((long long)z * g.step * mn + (long long)x * g.step * g.m +
                              (long long)min(y, g.m1 - 1) * g.step));
    R* an = acc + (size_t)slot * L3;
    for (int w = 0; w < b.nw; w++) {
        int lz[kNodeMaxNk];
        if (!plane_labels(g, w * b.dk, lz)) continue;
        __syncthreads();
        stage_plane(Ms, M, g, b, kb, jb, ib, w * b.dk);
        __syncthreads();
        if (!active) continue;
        R part[kNodeMaxNk];
#pragma unroll
        for (int sk = 0; sk < kNodeMaxNk; sk++) part[sk] = 0;
        for (int sj = 0; sj < g.nj; sj++) {
            const R* mrow = Ms + (size_t)(((sj * g.sj + lx * g.q) / b.dj) * b.np + slot * g.step + ly * g.q) * 12;
            CR* frow = Fn + 12LL * sj * g.sj * g.m;
            for (int si = 0; si < g.step; si++) {
                R mv[12];
                load12(mv, mrow + 12 * si);
#pragma unroll
                for (int sk = 0; sk < kNodeMaxNk; sk++) {
                    if (lz[sk] < 0) continue;
                    CR* fs = frow + 12 * ((long long)sk * g.sk * mn + si);
#pragma unroll
                    for (int ch = 0; ch < 12; ch++) {
                        const R d = fs[ch] - mv[ch];
                        part[sk] += d * d;
                    }
                }
            }
        }
#pragma unroll
        for (int sk = 0; sk < kNodeMaxNk; sk++)
            if (lz[sk] >= 0) an[lz[sk] * L2 + t] += part[sk];
    }
    __syncthreads();
    for (int e = threadIdx.x; e < b.ny * L3; e += blockDim.x) {
        const int s2 = e / L3, l = e - s2 * L3;
        if (y0 + s2 < g.m1) {
            const long long node = y0 + s2 + (long long)g.m1 * (x + (long long)g.n1 * z);
            cost[node * L3 + l] = acc[e] * (R)g.scale;
        }
    }
}

// dF: the waves of node slot ny hold its samples (si, sj, sk); each thread
// sums g[l] M[s + d_l] and g[l] over the labels, plane by plane, with the
// node's cost-gradient slices of the plane's lz values staged beside it.
template <typename R>
__global__ __launch_bounds__(1024) void k_sc_dF_node(R* __restrict__ dF, const R* __restrict__ gc,
                                                     const R* __restrict__ F, const R* __restrict__ M, CostGeom g,
                                                     BoxGeom b, int tpn) {
    extern __shared__ __align__(16) unsigned char sc_smem[];
    const int tile = xcd_tile((int)blockIdx.x, b.per);
    if (tile >= b.nyb * g.n1 * g.o1) return;
    const int L = g.len, L2 = L * L, L3 = L2 * L;
    R* Ms = reinterpret_cast<R*>(sc_smem);
    R* gs = Ms + (size_t)b.nv * b.np * 12;  // [ny][kNodeMaxNk][L2]
    const int yb = tile % b.nyb, x = (tile / b.nyb) % g.n1, z = tile / (b.nyb * g.n1);
    const int slot = threadIdx.x / tpn, t = threadIdx.x - slot * tpn;
    const int y0 = yb * b.ny, y = y0 + slot;
    const bool active = t < g.ns && y < g.m1;
    const int si = t % g.step, sj = (t / g.step) % g.nj, sk = t / (g.step * g.nj);
    const long long mn = (long long)g.m * g.n;
    const int P = g.hw * g.q;
    const int kb = z * g.step - P, jb = x * g.step - P, ib = y0 * g.step - P;
    const long long nrow = (long long)g.m1 * (x + (long long)g.n1 * z);
    R A[12], G = 0;
#pragma unroll
    for (int ch = 0; ch < 12; ch++) A[ch] = 0;
    for (int w = 0; w < b.nw; w++) {
        int lz[kNodeMaxNk];
        if (!plane_labels(g, w * b.dk, lz)) continue;
        __syncthreads();
        stage_plane(Ms, M, g, b, kb, jb, ib, w * b.dk);
        for (int e = threadIdx.x; e < b.ny * kNodeMaxNk * L2; e += blockDim.x) {
            const int s2 = e / (kNodeMaxNk * L2), r = e - s2 * (kNodeMaxNk * L2), k2 = r / L2, l2 = r - k2 * L2;
            int lzk = -1;
#pragma unroll
            for (int q = 0; q < kNodeMaxNk; q++)
                if (q == k2) lzk = lz[q];
            if (lzk >= 0 && y0 + s2 < g.m1) gs[e] = gc[(nrow + y0 + s2) * L3 + lzk * L2 + l2];
        }
        __syncthreads();
        int mylz = -1;
#pragma unroll
        for (int q = 0; q < kNodeMaxNk; q++)
            if (q == sk) mylz = lz[q];
        if (!active || mylz < 0) continue;
        const R* gsl = gs + (size_t)(slot * kNodeMaxNk + sk) * L2;
        for (int lx = 0; lx < L; lx++) {
            const R* mrow = Ms + (size_t)(((sj * g.sj + lx * g.q) / b.dj) * b.np + slot * g.step + si) * 12;
            for (int ly = 0; ly < L; ly++) {
                const R gv = gsl[lx * L + ly];
                R mv[12];
                load12(mv, mrow + 12 * ly * g.q);
                G += gv;
#pragma unroll
                for (int ch = 0; ch < 12; ch++) A[ch] += gv * mv[ch];
            }
        }
    }
    if (!active) return;
    const long long fidx = 12 * ((long long)(z * g.step + sk * g.sk) * mn + (long long)(x * g.step + sj * g.sj) * g.m +
                                 (long long)y * g.step + si);
    R fv[12];
    load12(fv, F + fidx);
#pragma unroll
    for (int ch = 0; ch < 12; ch++) dF[fidx + ch] = (R)(2.0 * g.scale) * (fv[ch] * G - A[ch]);
}

template <typename R>
__global__ void k_fill(R* p, R v, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ---- per-node dispatch ----------------------------------------------------------

inline int gcd_int(int a, int b) { return b == 0 ? a : gcd_int(b, a % b); }

constexpr int kNodeY = 2;  // nodes per workgroup along i

inline BoxGeom box_geom(const CostGeom& g) {
    BoxGeom b{};
    b.dk = gcd_int(g.sk, g.q);
    b.dj = gcd_int(g.sj, g.q);
    b.nw = ((g.nk - 1) * g.sk + (g.len - 1) * g.q) / b.dk + 1;
    b.nv = ((g.nj - 1) * g.sj + (g.len - 1) * g.q) / b.dj + 1;
    b.ny = g.m1 < kNodeY ? (g.m1 > 0 ? g.m1 : 1) : kNodeY;
    b.np = b.ny * g.step + (g.len - 1) * g.q;
    b.nyb = (g.m1 + b.ny - 1) / b.ny;
    b.per = (b.nyb * g.n1 * g.o1 + 7) / 8;
    return b;
}

// LDS bytes of the per-node forms, or 0 when the shape has none
// (samples skipped along i, more than kNodeMaxNk sample planes, more threads
// than a workgroup holds, or more LDS than 64 KB, which would leave a CU
// fewer than two workgroups).
template <typename R>
size_t node_smem_cost(const CostGeom& g, const BoxGeom& b) {
    if (g.si != 1 || g.nk > kNodeMaxNk) return 0;
    const int tpn = (g.len * g.len + 63) / 64 * 64;
    if (tpn * b.ny > 1024) return 0;
    const size_t bytes = ((size_t)b.nv * b.np * 12 + (size_t)b.ny * g.len * g.len * g.len) * sizeof(R);
    return bytes <= 64 * 1024 ? bytes : 0;
}
template <typename R>
size_t node_smem_dF(const CostGeom& g, const BoxGeom& b) {
    if (g.si != 1 || g.nk > kNodeMaxNk) return 0;
    const int tpn = (g.ns + 63) / 64 * 64;
    if (tpn * b.ny > 1024) return 0;
    const size_t bytes = ((size_t)b.nv * b.np * 12 + (size_t)b.ny * kNodeMaxNk * g.len * g.len) * sizeof(R);
    return bytes <= 64 * 1024 ? bytes : 0;
}

// ---- row-form dispatch ---------------------------------------------------------

// (STEP, LEN, Q) shapes with compiled row kernels; other shapes run the
// per-label kernels above.
#define DEEDS_SOFT_ROW_SHAPES(X) X(6, 13, 2) X(6, 5, 2) X(4, 5, 2) X(5, 3, 2)

inline unsigned grid_of(long long blocks) {
    if (blocks > 0x7fffffffLL) fail(DEEDS_ERR_ARG, "soft cost volume too large");
    return (unsigned)blocks;
}

template <typename R, int STEP, int LEN, int Q>
struct RowForms {
    static constexpr int P = (LEN - 1) / 2 * Q, J = kRowJ;
    // row groups: per residue of the row index modulo Q, J rows Q apart
    static int groups(int rows) { return Q * (((rows + Q - 1) / Q + J - 1) / J); }
    static int tiles(int W) { return (W + kRowNT - 1) / kRowNT; }
    static void dM(R* dM, R* part, const R* gT, const R* F, const R* M, const CostGeom& g, hipStream_t s) {
        const int it = tiles(g.m + 2 * P), gr = groups(g.n + 2 * P);
        k_sc_dM_rows<R, STEP, LEN, Q, J><<<grid_of((long long)it * gr * (g.o + 2 * P)), kRowNT, 0, s>>>(
            dM, part, gT, F, M, g, it, gr);
        DEEDS_LAUNCH_CHECK();
    }
    static void dF(R* dF, const R* gT, const R* F, const R* M, const CostGeom& g, hipStream_t s) {
        const int it = tiles(g.m1 * STEP), gr = groups(g.n1 * STEP);
        k_sc_dF_rows<R, STEP, LEN, Q, J><<<grid_of((long long)it * gr * (g.o1 * STEP)), kRowNT, 0, s>>>(
            dF, gT, F, M, g, it, gr);
        DEEDS_LAUNCH_CHECK();
    }
};

// Calls Op<R, STEP, LEN, Q>::f(args...) for a shape with row kernels; false otherwise.
template <typename R, typename Fn>
bool with_rows(const CostGeom& g, Fn&& fn) {
    if (g.si != 1) return false;
#define DEEDS_ROW_CASE(S, L, QQ)                          \
    if (g.step == S && g.len == L && g.q == QQ) {          \
        fn(RowForms<R, S, L, QQ>{});                       \
        return true;                                       \
    }
    DEEDS_SOFT_ROW_SHAPES(DEEDS_ROW_CASE)
#undef DEEDS_ROW_CASE
    return false;
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
    const BoxGeom b = box_geom(g);
    if (const size_t smem = node_smem_cost<R>(g, b)) {
        const int tpn = (g.len * g.len + 63) / 64 * 64;
        set_smem((const void*)k_sc_cost_node<R>, smem);
        k_sc_cost_node<R><<<8 * b.per, tpn * b.ny, smem, s>>>(cost, F, M, g, b, tpn);
        DEEDS_LAUNCH_CHECK();
        return;
    }
    const long long blocks = nodes * ((L3 + CB - 1) / CB);
    if (blocks > 0x7fffffffLL) fail(DEEDS_ERR_ARG, "soft cost volume too large");
    const size_t smem = (size_t)g.ns * 12 * sizeof(R);
    set_smem((const void*)k_sc_cost<R>, smem);
    k_sc_cost<R><<<(unsigned)blocks, CB, smem, s>>>(cost, F, M, g);
    DEEDS_LAUNCH_CHECK();
}

template <typename R>
void soft_cost_backward(R* dF, R* dM, const R* gcost, const R* F, const R* M, int m, int n, int o, int step, int hw,
                        float quant, float alpha, R* part, R* gT, hipStream_t s) {
    const CostGeom g = cost_geom(m, n, o, step, hw, quant, alpha);
    const long long sz = (long long)m * n * o;
    // voxels outside every sampled block get no gradient from the cost
    k_fill<R><<<ceil_div(12 * sz, 256), 256, 0, s>>>(dF, (R)0, 12 * sz);
    DEEDS_LAUNCH_CHECK();
    const int nodes = g.m1 * g.n1 * g.o1;
    // shapes with row kernels: both adjoints as row correlations over the
    // label-major gradient
    const bool rows = nodes > 0 && with_rows<R>(g, [&](auto rf) {
        const int L3 = g.len * g.len * g.len;
        k_sc_grad_t<R><<<dim3((L3 + 63) / 64, (nodes + 63) / 64), 256, 0, s>>>(gT, gcost, nodes, L3);
        DEEDS_LAUNCH_CHECK();
        rf.dF(dF, gT, F, M, g, s);
        rf.dM(dM, part, gT, F, M, g, s);
    });
    if (!rows) {
        const BoxGeom b = box_geom(g);
        if (const size_t smem = nodes > 0 ? node_smem_dF<R>(g, b) : 0) {
            const int tpn = (g.ns + 63) / 64 * 64;
            set_smem((const void*)k_sc_dF_node<R>, smem);
            k_sc_dF_node<R><<<8 * b.per, tpn * b.ny, smem, s>>>(dF, gcost, F, M, g, b, tpn);
            DEEDS_LAUNCH_CHECK();
        } else if (nodes > 0) {
            const size_t L3 = (size_t)g.len * g.len * g.len;
            const size_t smem = (((L3 + 1) & ~(size_t)1) + (size_t)(CB / g.ns) * g.ns * 13) * sizeof(R);
            set_smem((const void*)k_sc_dF<R>, smem);
            k_sc_dF<R><<<nodes, CB, smem, s>>>(dF, gcost, F, M, g);
            DEEDS_LAUNCH_CHECK();
        }
        // pull over the padded box (interior voxels finish there)
        const long long P = (long long)g.q * g.hw;
        const long long padded = (m + 2 * P) * (n + 2 * P) * (o + 2 * P);
        k_sc_dM_pull<R><<<ceil_div(padded, CB), CB, 0, s>>>(dM, part, gcost, F, M, g);
        DEEDS_LAUNCH_CHECK();
    }
    // fold the pad region onto the faces, one axis at a time, then the face voxels
    {
        const long long P = (long long)g.q * g.hw;
        k_sc_dM_fold<R, 0><<<ceil_div(13 * (n + 2 * P) * (o + 2 * P), CB), CB, 0, s>>>(part, g);
        DEEDS_LAUNCH_CHECK();
        k_sc_dM_fold<R, 1><<<ceil_div(13LL * m * (o + 2 * P), CB), CB, 0, s>>>(part, g);
        DEEDS_LAUNCH_CHECK();
        k_sc_dM_fold<R, 2><<<ceil_div(13LL * m * n, CB), CB, 0, s>>>(part, g);
        DEEDS_LAUNCH_CHECK();
    }
    const long long mi = m > 2 ? m - 2 : 0, ni = n > 2 ? n - 2 : 0;
    const long long faces = 2LL * n * o + 2 * mi * o + 2 * mi * ni;
    k_sc_dM_faces<R><<<ceil_div(faces, CB), CB, 0, s>>>(dM, part, M, g);
    DEEDS_LAUNCH_CHECK();
}

#define DEEDS_SOFT_COST(R)                                                                                     \
    template void soft_cost<R>(R*, const R*, const R*, int, int, int, int, int, float, float, hipStream_t);         \
    template void soft_cost_backward<R>(R*, R*, const R*, const R*, const R*, int, int, int, int, int, float, float, \
                                        R*, R*, hipStream_t);

DEEDS_SOFT_COST(float)
DEEDS_SOFT_COST(double)

}  // namespace deeds
