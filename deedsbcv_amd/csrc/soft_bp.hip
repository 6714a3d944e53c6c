// Soft variant, stage 3: sum-product belief propagation on the tree and its
// adjoint (the model is stated in soft.h).
//
// Each sweep runs in windowed segment rounds (SoftRounds): one launch per
// round, one workgroup per segment, the segment's nodes in order with the
// message of the node before kept on chip for the next one.
// Forward sweeps:
//   upward   (children first)       h_c = C_c + sum of the children's messages,
//            m_c = DTup_c(h_c) - min
//   downward (parents first)        g_c = h_c + DTdown_c(g_p - m_c) - min
//   marginals P_c = softmax(-g_c / T), u1_c = u0_c + E_{P_c}[displacement]
// DT is three log-sum-exp passes, one per label axis, against the quadratic
// pairwise table of the edge (the soft counterpart of messageDT,
// regularisation.h:6-113).  The label count per axis is a template
// parameter: the S threads of a line keep the line's values and a window of
// the table in registers and produce OUT outputs each, so a pass costs no
// index arithmetic or LDS traffic beyond one read of the line.
//
// Backward sweeps: the marginals' own terms of A_c = d/dg_c for all nodes in
// one launch, and in one more launch every node's upward and downward
// transforms recomputed from the saved beliefs and messages (all four pass
// buffers kept); then children first, A_c += the adjoints Dt_k of the
// children's downward transforms and Dt_c = DTdown_c^T(A_c); then parents
// first, dH_c = A_c + DTup_c^T(dH_p - Dt_c), which is d/dC_c.  A step of the
// sweeps reads its node's recomputed passes and differentiates them pass by
// pass in reverse.
// The minimum subtracted from each message only shifts beliefs by constants
// the marginals ignore, so it has no adjoint term.
#include "soft_common.h"

// The soft BP reproduces no reference rounding: let a*b+c contract to FMA here.
#pragma clang fp contract(fast)

namespace deeds {

namespace {

using namespace soft_dev;

// The tree arrays and schedules are written before a sweep and only read in
// it; read through the constant address space, their uniform loads are scalar.
using cint = const int __attribute__((address_space(4)));
__device__ __forceinline__ cint* ro(const int* p) { return (cint*)p; }
template <typename R>
__device__ __forceinline__ const R __attribute__((address_space(4)))* ro(const R* p) {
    return (const R __attribute__((address_space(4)))*)p;
}

template <typename R, int LEN>
struct SS {
    static constexpr int HW = (LEN - 1) / 2;
    static constexpr int L2 = LEN * LEN, L3 = L2 * LEN;
    static constexpr int OUT = soft_bp_out(LEN), S = soft_bp_s(LEN);
    static constexpr int BLOCK = ((L2 * S + 63) / 64) * 64;
    static constexpr int PAD = S * OUT;  // table entries beyond both ends read by the last windows
    static constexpr int ZT = 2 * LEN + 1 + 2 * PAD;
    static constexpr size_t SMEM = soft_bp_smem(LEN, sizeof(R));
    static constexpr int KL = (L3 + BLOCK - 1) / BLOCK;  // label-cube entries per thread
};

// Label cubes, the quadratic tables z of an edge per axis, their exponential
// forms k = exp(-(z - zmin) / T) and a reduction area whose slots 16..18
// hold zmin per axis.
template <typename R, int LEN>
struct Cubes {
    R *X0, *X1, *X2, *X3, *Ga, *Gb, *zy, *zx, *zz, *ky, *kx, *kz, *red;
    __device__ explicit Cubes(R* b) {
        using H = SS<R, LEN>;
        X0 = b;
        X1 = X0 + H::L3;
        X2 = X1 + H::L3;
        X3 = X2 + H::L3;
        Ga = X3 + H::L3;
        Gb = Ga + H::L3;
        zy = Gb + H::L3 + H::PAD;  // zy[q] for q = -PAD .. 2*LEN+PAD: offset q - LEN
        zx = zy + H::ZT;
        zz = zx + H::ZT;
        ky = zz + H::ZT;
        kx = ky + H::ZT;
        kz = kx + H::ZT;
        red = kz - H::PAD + H::ZT;
    }
};

// A log-sum-exp pass runs as a convolution in the exponential domain,
// sum_b exp(-(in(b) - min in) / T) exp(-(z(a - b) - zmin) / T), when that sum
// stays above 2^-ConvLimit (every term of it is then a normal number and the
// largest one is at least the sum / LEN); below it, the output is computed
// term by term, shifted by its own minimum, as the exact form.
template <typename R>
struct Conv;
template <>
struct Conv<float> {
    static constexpr float limit = 100.0f;       // log2 units
    static constexpr float floor = 0x1p-100f;
};
template <>
struct Conv<double> {
    static constexpr double limit = 900.0;
    static constexpr double floor = 0x1p-900;
};

// Lines of the cube along axis AX (0: i / stride 1, 1: j / stride LEN, 2: k / stride LEN^2).
template <int LEN, int AX>
struct Line {
    static constexpr int STRIDE = AX == 0 ? 1 : AX == 1 ? LEN : LEN * LEN;
    static __device__ __forceinline__ int base(int line) {
        return AX == 0 ? line * LEN : AX == 1 ? (line % LEN) + (line / LEN) * LEN * LEN : line;
    }
};

// out(a) = -T ln sum_b exp(-(in(b) + z[a - b]) / T) along axis AX (z the
// quadratic table, k its exponential form, zmin its minimum over the label
// differences); returns the smallest output this thread wrote.
template <typename R, int LEN, int AX>
__device__ __forceinline__ R lse_pass(const R* __restrict__ in, R* __restrict__ out, const R* __restrict__ zt,
                                      const R* __restrict__ kt, R zmin, R iT2, R Tln2) {
    using H = SS<R, LEN>;
    using G = Line<LEN, AX>;
    const int line = threadIdx.x / H::S, sub = threadIdx.x - line * H::S;
    R lo = big<R>();
    if (line >= H::L2) return lo;
    const int base = G::base(line);
    R v[LEN], mi = big<R>();
#pragma unroll
    for (int j = 0; j < LEN; j++) {
        v[j] = in[base + j * G::STRIDE];
        mi = rmin(v[j], mi);
    }
    R e[LEN];
#pragma unroll
    for (int j = 0; j < LEN; j++) e[j] = FastMath<R>::exp2((mi - v[j]) * iT2);
    const int i0 = sub * H::OUT;
    R kw[H::OUT + LEN];  // kw[r - j + LEN] = kt[(i0 + r) - j + LEN]
#pragma unroll
    for (int q = 0; q < H::OUT + LEN; q++) kw[q] = kt[i0 + q];
    const R shift = mi + zmin;
#pragma unroll
    for (int r = 0; r < H::OUT; r++) {
        if (i0 + r < LEN) {
            R sum = 0;
#pragma unroll
            for (int j = 0; j < LEN; j++) sum += e[j] * kw[r - j + LEN];
            R o;
            if (sum >= Conv<R>::floor) {
                o = shift - Tln2 * FastMath<R>::log2(sum);
            } else {
                R c[LEN], mn = big<R>();
#pragma unroll
                for (int j = 0; j < LEN; j++) {
                    c[j] = v[j] + zt[i0 + r - j + LEN];
                    mn = rmin(c[j], mn);
                }
                R s2 = 0;
#pragma unroll
                for (int j = 0; j < LEN; j++) s2 += FastMath<R>::exp2((mn - c[j]) * iT2);
                o = mn - Tln2 * FastMath<R>::log2(s2);
            }
            out[base + (i0 + r) * G::STRIDE] = o;
            lo = rmin(o, lo);
        }
    }
    return lo;
}

// Adjoint of lse_pass: gin(b) = sum_a gout(a) exp(-(in(b) + z[a - b] - out(a)) / T)
//   = exp(-(in(b) - min in) / T) sum_a w(a) k[a - b],  w(a) = gout(a) exp((out(a) - min in - zmin) / T),
// with the exact term for every a whose weight factor exceeds 2^ConvLimit and
// for every b whose factor is below 2^-ConvLimit.
template <typename R, int LEN, int AX>
__device__ __forceinline__ void lse_adj(const R* __restrict__ in, const R* __restrict__ out,
                                        const R* __restrict__ gout, R* __restrict__ gin, const R* __restrict__ zt,
                                        const R* __restrict__ kt, R zmin, R iT2) {
    using H = SS<R, LEN>;
    using G = Line<LEN, AX>;
    const int line = threadIdx.x / H::S, sub = threadIdx.x - line * H::S;
    if (line >= H::L2) return;
    const int base = G::base(line);
    R go[LEN], ou[LEN], mi = big<R>();
#pragma unroll
    for (int a = 0; a < LEN; a++) {
        go[a] = gout[base + a * G::STRIDE];
        ou[a] = out[base + a * G::STRIDE];
        const R x = in[base + a * G::STRIDE];
        mi = rmin(x, mi);
    }
    R w[LEN];
    bool any_big = false;
#pragma unroll
    for (int a = 0; a < LEN; a++) {
        const R ex = (ou[a] - mi - zmin) * iT2;
        const bool bg = ex > Conv<R>::limit;
        any_big |= bg;
        w[a] = bg ? (R)0 : go[a] * FastMath<R>::exp2(ex);
    }
    const int i0 = sub * H::OUT;
    R kw[H::OUT + LEN - 1];  // kw[a - r + OUT - 1] = kt[a - (i0 + r) + LEN]
#pragma unroll
    for (int q = 0; q < H::OUT + LEN - 1; q++) kw[q] = kt[LEN - i0 - (H::OUT - 1) + q];
#pragma unroll
    for (int r = 0; r < H::OUT; r++) {
        const int b = i0 + r;
        if (b < LEN) {
            const R x = in[base + b * G::STRIDE];
            const R eb = (mi - x) * iT2;
            R acc = 0;
            if (eb >= -Conv<R>::limit) {
#pragma unroll
                for (int a = 0; a < LEN; a++) acc += w[a] * kw[a - r + H::OUT - 1];
                acc *= FastMath<R>::exp2(eb);
                if (any_big) {
#pragma unroll
                    for (int a = 0; a < LEN; a++)
                        if ((ou[a] - mi - zmin) * iT2 > Conv<R>::limit)
                            acc += go[a] * FastMath<R>::exp2((ou[a] - x - zt[a - b + LEN]) * iT2);
                }
            } else {
#pragma unroll
                for (int a = 0; a < LEN; a++) acc += go[a] * FastMath<R>::exp2((ou[a] - x - zt[a - b + LEN]) * iT2);
            }
            gin[base + b * G::STRIDE] = acc;
        }
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
    R* H;           // upward beliefs h_c
    R* Msg;         // upward messages, shifted to minimum 0
    R* G;           // full beliefs g_c
    const R* u0;    // 3 x N (u, v, w)
    R* u1;          // 3 x N
    SoftTreeDev t;
    int N;
    R quant, T;
};

template <typename R>
struct BPGrad {
    const R* gu1;  // 3 x N
    R* A;          // total d/dg_c
    R* Dt;         // d/d(input of the downward transform of c)
    R* dH;         // d/dh_c == d/dC_c
    R* RX;         // N x 7 x L3: recomputed passes X1..X3 of the upward and X0..X3 of the downward transform
};

// Offsets in RX of a node's recomputed passes.
template <int L3>
struct RXAt {
    static constexpr int UP = 0, DOWN = 3 * L3, STRIDE = 7 * L3;
};

// The scalars behind the quadratic tables of the edge c -> p: its weight
// and the offsets (u0_p - u0_c) / quant per axis, times sign (+1 for the
// upward transform, a = parent label, b = child label; -1 for the downward
// one).  Loaded on their own so that a sweep can fetch them a step ahead.
template <typename R>
struct EdgeIn {
    R w, ou, ov, ow;
};
template <typename R>
__device__ __forceinline__ EdgeIn<R> edge_in(const BPState<R>& a, int c, int p, R sign) {
    const float wf = ro(a.t.edgemst)[c];
    const auto u0 = ro(a.u0);
    EdgeIn<R> e;
    e.w = (R)(wf > kSoftMinEdgeWeight ? wf : kSoftMinEdgeWeight);
    e.ou = sign * (u0[p] - u0[c]) / a.quant;
    e.ov = sign * (u0[a.N + p] - u0[a.N + c]) / a.quant;
    e.ow = sign * (u0[2 * a.N + p] - u0[2 * a.N + c]) / a.quant;
    return e;
}

// Quadratic tables of the edge: z[a - b] = (a - b + off)^2 / w per axis
// (label axis 0 is i / v, 1 is j / u, 2 is k / w).
template <typename R, int LEN>
__device__ void edge_tables(const BPState<R>& a, const EdgeIn<R>& e, const Cubes<R, LEN>& s) {
    using H = SS<R, LEN>;
    const R w = e.w, ou = e.ou, ov = e.ov, ow = e.ow;
    // zmin: the table's minimum over the label differences -(LEN-1) .. LEN-1,
    // at the integer next to -off on either side
    R zm[3];
    const R offs[3] = {ov, ou, ow};
#pragma unroll
    for (int ax = 0; ax < 3; ax++) {
        const R o = offs[ax];
        R t0 = floor(-o), t1 = t0 + 1;
        t0 = t0 < (R)(1 - LEN) ? (R)(1 - LEN) : (t0 > (R)(LEN - 1) ? (R)(LEN - 1) : t0);
        t1 = t1 < (R)(1 - LEN) ? (R)(1 - LEN) : (t1 > (R)(LEN - 1) ? (R)(LEN - 1) : t1);
        const R z0 = (t0 + o) * (t0 + o) / w, z1 = (t1 + o) * (t1 + o) / w;
        zm[ax] = z0 < z1 ? z0 : z1;
    }
    const R iT2 = (R)kLog2e / a.T;
    for (int q = threadIdx.x; q < H::ZT; q += blockDim.x) {
        const R t = (R)(q - H::PAD - LEN);
        const R y = (t + ov) * (t + ov) / w, x = (t + ou) * (t + ou) / w, z = (t + ow) * (t + ow) / w;
        s.zy[q - H::PAD] = y;
        s.zx[q - H::PAD] = x;
        s.zz[q - H::PAD] = z;
        s.ky[q - H::PAD] = FastMath<R>::exp2(-(y - zm[0]) * iT2);
        s.kx[q - H::PAD] = FastMath<R>::exp2(-(x - zm[1]) * iT2);
        s.kz[q - H::PAD] = FastMath<R>::exp2(-(z - zm[2]) * iT2);
    }
    if (threadIdx.x == 0) {
        s.red[16] = zm[0];
        s.red[17] = zm[1];
        s.red[18] = zm[2];
    }
}

// X0 -> X1 -> X2 -> X3 with the current tables; returns the minimum of X3.
template <typename R, int LEN>
__device__ R transform(const Cubes<R, LEN>& s, R T) {
    const R iT2 = (R)kLog2e / T, Tln2 = T * (R)kLn2;
    __syncthreads();
    lse_pass<R, LEN, 0>(s.X0, s.X1, s.zy, s.ky, s.red[16], iT2, Tln2);
    __syncthreads();
    lse_pass<R, LEN, 1>(s.X1, s.X2, s.zx, s.kx, s.red[17], iT2, Tln2);
    __syncthreads();
    const R lo = lse_pass<R, LEN, 2>(s.X2, s.X3, s.zz, s.kz, s.red[18], iT2, Tln2);
    return block_reduce<R>(lo, s.red, true);
}

// X0 -> X1 -> X2 -> X3 with the current tables, without the minimum.
template <typename R, int LEN>
__device__ void passes(const Cubes<R, LEN>& s, R T) {
    const R iT2 = (R)kLog2e / T, Tln2 = T * (R)kLn2;
    __syncthreads();
    lse_pass<R, LEN, 0>(s.X0, s.X1, s.zy, s.ky, s.red[16], iT2, Tln2);
    __syncthreads();
    lse_pass<R, LEN, 1>(s.X1, s.X2, s.zx, s.kx, s.red[17], iT2, Tln2);
    __syncthreads();
    lse_pass<R, LEN, 2>(s.X2, s.X3, s.zz, s.kz, s.red[18], iT2, Tln2);
    __syncthreads();
}

// d/dX3 in Ga -> d/dX0 in Gb (uses Ga as scratch).
template <typename R, int LEN>
__device__ void transform_adj(const Cubes<R, LEN>& s, R T) {
    const R iT2 = (R)kLog2e / T;
    __syncthreads();
    lse_adj<R, LEN, 2>(s.X2, s.X3, s.Ga, s.Gb, s.zz, s.kz, s.red[18], iT2);
    __syncthreads();
    lse_adj<R, LEN, 1>(s.X1, s.X2, s.Gb, s.Ga, s.zx, s.kx, s.red[17], iT2);
    __syncthreads();
    lse_adj<R, LEN, 0>(s.X0, s.X1, s.Ga, s.Gb, s.zy, s.ky, s.red[16], iT2);
    __syncthreads();
}

// Schedule of one launch: segment blockIdx.x of the round that starts at seg0.
struct SegArgs {
    const int* seg_start;
    const int* nodes;
    int seg0;
};

// Every sweep step loads its node's inputs one step ahead, while the step
// before computes (the loads stay in flight across the workgroup barriers,
// which wait for LDS only): rows of the node itself, the stored rows of up
// to two children other than the one before it in the segment (further
// ones, rare in an MST, are read when the step starts), and its edge's
// table scalars.  The chained neighbour's row comes from LDS or registers.

// Whether a sweep loads a step ahead: where the inputs of a step fit in
// registers beside the transform's (f32 with at most three label entries
// per thread); elsewhere a step loads its inputs when it starts.
template <typename R, int LEN>
constexpr bool sweep_prefetch() {
    return sizeof(R) == 4 && SS<R, LEN>::KL <= 3;
}

// The first two children of c other than `chained` (or -1), and whether there are more.
struct OtherKids {
    int k0 = -1, k1 = -1;
    bool more = false;
};
__device__ __forceinline__ OtherKids other_kids(cint* cs, cint* cl, int c, int chained) {
    OtherKids o;
    for (int ci = cs[c]; ci < cs[c + 1]; ci++) {
        const int k = cl[ci];
        if (k == chained) continue;
        if (o.k0 < 0) o.k0 = k;
        else if (o.k1 < 0) o.k1 = k;
        else o.more = true;
    }
    return o;
}

// Upward: h_c = C_c + the children's messages (the other children's stored
// messages, then the message of the child just before in the segment from
// X3 - its minimum), m_c = DTup(h_c) - min.
template <typename R, int LEN>
__global__ __launch_bounds__((SS<R, LEN>::BLOCK)) void k_sp_up(BPState<R> a, SegArgs g) {
    using H = SS<R, LEN>;
    extern __shared__ __align__(16) unsigned char sp_smem[];
    const Cubes<R, LEN> s(reinterpret_cast<R*>(sp_smem));
    cint* nd = ro(g.nodes);
    cint* cs = ro(a.t.child_start);
    cint* cl = ro(a.t.child_list);
    const int i0 = ro(g.seg_start)[g.seg0 + blockIdx.x], i1 = ro(g.seg_start)[g.seg0 + blockIdx.x + 1];
    struct In {
        int c, p;
        OtherKids kid;
        R cost[H::KL], m0[H::KL], m1[H::KL];
        EdgeIn<R> e;
    };
    auto fetch = [&](In& in, int c, int chained) {
        in.c = c;
        in.p = ro(a.t.parents)[c];
        in.kid = other_kids(cs, cl, c, chained);
        const long long cb = (long long)c * H::L3;
#pragma unroll
        for (int k = 0; k < H::KL; k++) {
            const int l = threadIdx.x + k * H::BLOCK;
            if (l < H::L3) {
                in.cost[k] = a.cost[cb + l];
                in.m0[k] = in.kid.k0 >= 0 ? a.Msg[(long long)in.kid.k0 * H::L3 + l] : (R)0;
                in.m1[k] = in.kid.k1 >= 0 ? a.Msg[(long long)in.kid.k1 * H::L3 + l] : (R)0;
            }
        }
        if (in.p >= 0) in.e = edge_in(a, c, in.p, (R)1);
    };
    constexpr bool PF = sweep_prefetch<R, LEN>();
    In cur, nxt;
    if (PF) fetch(cur, nd[i0], -1);
    int prev = -1;
    R pmn = 0;
    for (int i = i0; i < i1; i++) {
        if (!PF) fetch(cur, nd[i], prev);
        const int c = cur.c, p = cur.p;
        const long long cb = (long long)c * H::L3;
        const bool chained = prev >= 0;  // the node before in the segment is a child of c
#pragma unroll
        for (int k = 0; k < H::KL; k++) {
            const int l = threadIdx.x + k * H::BLOCK;
            if (l < H::L3) {
                R h = cur.cost[k] + cur.m0[k] + cur.m1[k];
                if (cur.kid.more) {
                    int seen = 0;
                    for (int ci = cs[c]; ci < cs[c + 1]; ci++) {
                        const int kk = cl[ci];
                        if (kk != prev && seen++ >= 2) h += a.Msg[(long long)kk * H::L3 + l];
                    }
                }
                if (chained) h += s.X3[l] - pmn;
                a.H[cb + l] = h;
                s.X0[l] = h;
            }
        }
        if (p < 0) {  // the root's full belief is its upward belief
            for (int l = threadIdx.x; l < H::L3; l += H::BLOCK) a.G[cb + l] = s.X0[l];
            return;
        }
        if (PF && i + 1 < i1) fetch(nxt, nd[i + 1], c);
        edge_tables<R, LEN>(a, cur.e, s);
        pmn = transform<R, LEN>(s, a.T);
        for (int l = threadIdx.x; l < H::L3; l += H::BLOCK) a.Msg[cb + l] = s.X3[l] - pmn;
        prev = c;
        if (PF) cur = nxt;
    }
}

// Downward: g_c = h_c + DTdown(g_p - m_c) - min, g_p of the node just before
// in the segment from registers.
template <typename R, int LEN>
__global__ __launch_bounds__((SS<R, LEN>::BLOCK)) void k_sp_down(BPState<R> a, SegArgs g) {
    using H = SS<R, LEN>;
    extern __shared__ __align__(16) unsigned char sp_smem[];
    const Cubes<R, LEN> s(reinterpret_cast<R*>(sp_smem));
    cint* nd = ro(g.nodes);
    const int i0 = ro(g.seg_start)[g.seg0 + blockIdx.x], i1 = ro(g.seg_start)[g.seg0 + blockIdx.x + 1];
    struct In {
        int c, p;
        R msg[H::KL], h[H::KL], gp[H::KL];
        EdgeIn<R> e;
    };
    auto fetch = [&](In& in, int c, int chained) {
        in.c = c;
        in.p = ro(a.t.parents)[c];
        const long long cb = (long long)c * H::L3, pb = (long long)in.p * H::L3;
#pragma unroll
        for (int k = 0; k < H::KL; k++) {
            const int l = threadIdx.x + k * H::BLOCK;
            if (l < H::L3) {
                in.msg[k] = a.Msg[cb + l];
                in.h[k] = a.H[cb + l];
                in.gp[k] = in.p != chained ? a.G[pb + l] : (R)0;
            }
        }
        in.e = edge_in(a, c, in.p, (R)-1);
    };
    constexpr bool PF = sweep_prefetch<R, LEN>();
    In cur, nxt;
    if (PF) fetch(cur, nd[i0], -1);
    int prev = -1;
    R gk[H::KL];
    for (int i = i0; i < i1; i++) {
        if (!PF) fetch(cur, nd[i], prev);
        const int c = cur.c;
        const bool chained = cur.p == prev;
        const long long cb = (long long)c * H::L3;
#pragma unroll
        for (int k = 0; k < H::KL; k++) {
            const int l = threadIdx.x + k * H::BLOCK;
            if (l < H::L3) s.X0[l] = (chained ? gk[k] : cur.gp[k]) - cur.msg[k];
        }
        if (PF && i + 1 < i1) fetch(nxt, nd[i + 1], c);
        edge_tables<R, LEN>(a, cur.e, s);
        const R mn = transform<R, LEN>(s, a.T);
#pragma unroll
        for (int k = 0; k < H::KL; k++) {
            const int l = threadIdx.x + k * H::BLOCK;
            if (l < H::L3) {
                gk[k] = cur.h[k] + (s.X3[l] - mn);
                a.G[cb + l] = gk[k];
            }
        }
        prev = c;
        if (PF) cur = nxt;
    }
}

// Marginal of node c into P (softmax of -g_c / T); e = expected displacement (u, v, w).
template <typename R, int LEN>
__device__ void marginal(const BPState<R>& a, int c, R* P, R* red, R e[3]) {
    using H = SS<R, LEN>;
    const long long cb = (long long)c * H::L3;
    R lo = big<R>();
    for (int l = threadIdx.x; l < H::L3; l += H::BLOCK) {
        const R g = a.G[cb + l];
        P[l] = g;
        lo = rmin(g, lo);
    }
    const R mn = block_reduce<R>(lo, red, true);
    const R iT2 = (R)kLog2e / a.T;
    R z = 0, eu = 0, ev = 0, ew = 0;
    for (int l = threadIdx.x; l < H::L3; l += H::BLOCK) {
        const R pw = FastMath<R>::exp2((mn - P[l]) * iT2);
        P[l] = pw;
        const int li = l % LEN, lj = (l / LEN) % LEN, lk = l / H::L2;
        z += pw;
        eu += pw * (R)(lj - H::HW);
        ev += pw * (R)(li - H::HW);
        ew += pw * (R)(lk - H::HW);
    }
    z = block_reduce<R>(z, red, false);
    eu = block_reduce<R>(eu, red, false);
    ev = block_reduce<R>(ev, red, false);
    ew = block_reduce<R>(ew, red, false);
    const R iz = (R)1 / z;
    for (int l = threadIdx.x; l < H::L3; l += H::BLOCK) P[l] *= iz;
    __syncthreads();
    e[0] = eu * iz * a.quant;
    e[1] = ev * iz * a.quant;
    e[2] = ew * iz * a.quant;
}

template <typename R, int LEN>
__global__ __launch_bounds__((SS<R, LEN>::BLOCK)) void k_sp_marginals(BPState<R> a) {
    extern __shared__ __align__(16) unsigned char sp_smem[];
    const Cubes<R, LEN> s(reinterpret_cast<R*>(sp_smem));
    const int c = blockIdx.x;
    R e[3];
    marginal<R, LEN>(a, c, s.X0, s.red, e);
    if (threadIdx.x == 0) {
        a.u1[c] = a.u0[c] + e[0];
        a.u1[a.N + c] = a.u0[a.N + c] + e[1];
        a.u1[2 * a.N + c] = a.u0[2 * a.N + c] + e[2];
    }
}

// The marginal's own term of A_c = d/dg_c for every node (one workgroup per
// node, all nodes at once): the gradient of u1_c through P_c.
template <typename R, int LEN>
__global__ __launch_bounds__((SS<R, LEN>::BLOCK)) void k_sp_bwd_marg(BPState<R> a, BPGrad<R> b) {
    using H = SS<R, LEN>;
    extern __shared__ __align__(16) unsigned char sp_smem[];
    const Cubes<R, LEN> s(reinterpret_cast<R*>(sp_smem));
    const int c = blockIdx.x;
    const long long cb = (long long)c * H::L3;
    R e[3];
    marginal<R, LEN>(a, c, s.X3, s.red, e);
    const R gu = b.gu1[c], gv = b.gu1[a.N + c], gw = b.gu1[2 * a.N + c];
    const R iT = (R)1 / a.T;
    for (int l = threadIdx.x; l < H::L3; l += H::BLOCK) {
        const int li = l % LEN, lj = (l / LEN) % LEN, lk = l / H::L2;
        const R du = (R)(lj - H::HW) * a.quant - e[0];
        const R dv = (R)(li - H::HW) * a.quant - e[1];
        const R dw = (R)(lk - H::HW) * a.quant - e[2];
        b.A[cb + l] = -(s.X3[l] * iT) * (gu * du + gv * dv + gw * dw);
    }
}

// The forward transforms the backward steps differentiate, recomputed for
// every node at once (their inputs are the saved beliefs and messages, so
// they do not wait on the sequential sweeps): workgroup 2c + 1 recomputes
// node c's downward transform (input g_p - m_c, stored with its passes),
// workgroup 2c its upward one (input h_c, passes stored).
template <typename R, int LEN>
__global__ __launch_bounds__((SS<R, LEN>::BLOCK)) void k_sp_bwd_passes(BPState<R> a, BPGrad<R> b) {
    using H = SS<R, LEN>;
    using X = RXAt<H::L3>;
    extern __shared__ __align__(16) unsigned char sp_smem[];
    const Cubes<R, LEN> s(reinterpret_cast<R*>(sp_smem));
    const int c = blockIdx.x >> 1, down = blockIdx.x & 1;
    const int p = ro(a.t.parents)[c];
    if (p < 0) return;
    const long long cb = (long long)c * H::L3, pb = (long long)p * H::L3;
    R* rx = b.RX + (long long)c * X::STRIDE + (down ? X::DOWN : X::UP);
    for (int l = threadIdx.x; l < H::L3; l += H::BLOCK) {
        if (down) {
            const R x = a.G[pb + l] - a.Msg[cb + l];
            s.X0[l] = x;
            rx[l] = x;
        } else {
            s.X0[l] = a.H[cb + l];
        }
    }
    edge_tables<R, LEN>(a, edge_in(a, c, p, down ? (R)-1 : (R)1), s);
    passes<R, LEN>(s, a.T);
    R* o = rx + (down ? H::L3 : 0);
    for (int l = threadIdx.x; l < H::L3; l += H::BLOCK) {
        o[l] = s.X1[l];
        o[H::L3 + l] = s.X2[l];
        o[2 * H::L3 + l] = s.X3[l];
    }
}

// Children first: A_c = own term + the children's Dt (the other children's
// stored ones, then the one of the child just before in the segment from
// Gb), then Dt_c = DTdown_c^T(A_c) over the recomputed downward transform.
// At the root, dH = A.
template <typename R, int LEN>
__global__ __launch_bounds__((SS<R, LEN>::BLOCK)) void k_sp_bwd_down(BPState<R> a, BPGrad<R> b, SegArgs g) {
    using H = SS<R, LEN>;
    using X = RXAt<H::L3>;
    extern __shared__ __align__(16) unsigned char sp_smem[];
    const Cubes<R, LEN> s(reinterpret_cast<R*>(sp_smem));
    cint* nd = ro(g.nodes);
    cint* cs = ro(a.t.child_start);
    cint* cl = ro(a.t.child_list);
    const int i0 = ro(g.seg_start)[g.seg0 + blockIdx.x], i1 = ro(g.seg_start)[g.seg0 + blockIdx.x + 1];
    struct In {
        int c, p;
        OtherKids kid;
        R A[H::KL], d0[H::KL], d1[H::KL], x[4][H::KL];
        EdgeIn<R> e;
    };
    auto fetch = [&](In& in, int c, int chained) {
        in.c = c;
        in.p = ro(a.t.parents)[c];
        in.kid = other_kids(cs, cl, c, chained);
        const long long cb = (long long)c * H::L3;
        const R* rx = b.RX + (long long)c * X::STRIDE + X::DOWN;
#pragma unroll
        for (int k = 0; k < H::KL; k++) {
            const int l = threadIdx.x + k * H::BLOCK;
            if (l < H::L3) {
                in.A[k] = b.A[cb + l];
                in.d0[k] = in.kid.k0 >= 0 ? b.Dt[(long long)in.kid.k0 * H::L3 + l] : (R)0;
                in.d1[k] = in.kid.k1 >= 0 ? b.Dt[(long long)in.kid.k1 * H::L3 + l] : (R)0;
                if (in.p >= 0) {  // the recomputed downward transform of c
#pragma unroll
                    for (int q = 0; q < 4; q++) in.x[q][k] = rx[q * H::L3 + l];
                }
            }
        }
        if (in.p >= 0) in.e = edge_in(a, c, in.p, (R)-1);
    };
    constexpr bool PF = sweep_prefetch<R, LEN>();
    In cur, nxt;
    if (PF) fetch(cur, nd[i0], -1);
    int prev = -1;
    for (int i = i0; i < i1; i++) {
        if (!PF) fetch(cur, nd[i], prev);
        const int c = cur.c, p = cur.p;
        const long long cb = (long long)c * H::L3;
        const bool chained = prev >= 0;
#pragma unroll
        for (int k = 0; k < H::KL; k++) {
            const int l = threadIdx.x + k * H::BLOCK;
            if (l < H::L3) {
                R v = cur.A[k] + cur.d0[k] + cur.d1[k];
                if (cur.kid.more) {
                    int seen = 0;
                    for (int ci = cs[c]; ci < cs[c + 1]; ci++) {
                        const int kk = cl[ci];
                        if (kk != prev && seen++ >= 2) v += b.Dt[(long long)kk * H::L3 + l];
                    }
                }
                if (chained) v += s.Gb[l];
                s.X0[l] = cur.x[0][k];
                s.X1[l] = cur.x[1][k];
                s.X2[l] = cur.x[2][k];
                s.X3[l] = cur.x[3][k];
                b.A[cb + l] = v;
                s.Ga[l] = v;
            }
        }
        if (p < 0) {
            for (int l = threadIdx.x; l < H::L3; l += H::BLOCK) b.dH[cb + l] = s.Ga[l];
            return;
        }
        if (PF && i + 1 < i1) fetch(nxt, nd[i + 1], c);
        edge_tables<R, LEN>(a, cur.e, s);
        transform_adj<R, LEN>(s, a.T);
        for (int l = threadIdx.x; l < H::L3; l += H::BLOCK) b.Dt[cb + l] = s.Gb[l];
        prev = c;
        if (PF) cur = nxt;
    }
}

// Parents first: dH_c = A_c + DTup_c^T(dH_p - Dt_c) over the recomputed
// upward transform, dH_p of the node just before in the segment from
// registers.
template <typename R, int LEN>
__global__ __launch_bounds__((SS<R, LEN>::BLOCK)) void k_sp_bwd_up(BPState<R> a, BPGrad<R> b, SegArgs g) {
    using H = SS<R, LEN>;
    using X = RXAt<H::L3>;
    extern __shared__ __align__(16) unsigned char sp_smem[];
    const Cubes<R, LEN> s(reinterpret_cast<R*>(sp_smem));
    cint* nd = ro(g.nodes);
    const int i0 = ro(g.seg_start)[g.seg0 + blockIdx.x], i1 = ro(g.seg_start)[g.seg0 + blockIdx.x + 1];
    struct In {
        int c, p;
        R h[H::KL], x[3][H::KL], dt[H::KL], A[H::KL], dhp[H::KL];
        EdgeIn<R> e;
    };
    auto fetch = [&](In& in, int c, int chained) {
        in.c = c;
        in.p = ro(a.t.parents)[c];
        const long long cb = (long long)c * H::L3, pb = (long long)in.p * H::L3;
        const R* rx = b.RX + (long long)c * X::STRIDE + X::UP;
#pragma unroll
        for (int k = 0; k < H::KL; k++) {
            const int l = threadIdx.x + k * H::BLOCK;
            if (l < H::L3) {
                in.h[k] = a.H[cb + l];
#pragma unroll
                for (int q = 0; q < 3; q++) in.x[q][k] = rx[q * H::L3 + l];
                in.dt[k] = b.Dt[cb + l];
                in.A[k] = b.A[cb + l];
                in.dhp[k] = in.p != chained ? b.dH[pb + l] : (R)0;
            }
        }
        in.e = edge_in(a, c, in.p, (R)1);
    };
    constexpr bool PF = sweep_prefetch<R, LEN>();
    In cur, nxt;
    if (PF) fetch(cur, nd[i0], -1);
    int prev = -1;
    R hk[H::KL];
    for (int i = i0; i < i1; i++) {
        if (!PF) fetch(cur, nd[i], prev);
        const int c = cur.c;
        const bool chained = cur.p == prev;
        const long long cb = (long long)c * H::L3;
#pragma unroll
        for (int k = 0; k < H::KL; k++) {
            const int l = threadIdx.x + k * H::BLOCK;
            if (l < H::L3) {
                s.X0[l] = cur.h[k];
                s.X1[l] = cur.x[0][k];
                s.X2[l] = cur.x[1][k];
                s.X3[l] = cur.x[2][k];
                s.Ga[l] = (chained ? hk[k] : cur.dhp[k]) - cur.dt[k];
            }
        }
        if (PF && i + 1 < i1) fetch(nxt, nd[i + 1], c);
        edge_tables<R, LEN>(a, cur.e, s);
        transform_adj<R, LEN>(s, a.T);
#pragma unroll
        for (int k = 0; k < H::KL; k++) {
            const int l = threadIdx.x + k * H::BLOCK;
            if (l < H::L3) {
                hk[k] = cur.A[k] + s.Gb[l];
                b.dH[cb + l] = hk[k];
            }
        }
        prev = c;
        if (PF) cur = nxt;
    }
}

template <typename R>
BPState<R> bp_state(const SoftTree& t, const R* cost, const R* u0, R* u1, R* saved, int hw, float quant, float T) {
    const size_t L = 2 * (size_t)hw + 1, L3 = L * L * L;
    BPState<R> a{};
    a.cost = cost;
    a.H = saved;
    a.Msg = saved + (size_t)t.N * L3;
    a.G = saved + 2 * (size_t)t.N * L3;
    a.u0 = u0;
    a.u1 = u1;
    a.t = SoftTreeDev{t.d_ordered, t.d_parents, t.d_child_start, t.d_child_list, t.d_edgemst};
    a.N = t.N;
    a.quant = (R)quant;
    a.T = (R)T;
    return a;
}

template <typename R, int LEN>
struct Sweeps {
    using H = SS<R, LEN>;
    // one launch per round of r, the kernel's arguments followed by the round's SegArgs
    template <typename... Args>
    static void rounds(const SoftRounds& r, void (*kernel)(Args..., SegArgs), hipStream_t s, const Args&... args) {
        for (int k = 0; k < r.rounds(); k++) {
            const SegArgs g{r.d_seg_start, r.d_nodes, r.round_start[k]};
            const unsigned n = (unsigned)(r.round_start[k + 1] - r.round_start[k]);
            kernel<<<n, H::BLOCK, H::SMEM, s>>>(args..., g);
            DEEDS_LAUNCH_CHECK();
        }
    }
    static void forward(const SoftTree& t, const BPState<R>& a, hipStream_t s) {
        set_smem((const void*)k_sp_up<R, LEN>, H::SMEM);
        set_smem((const void*)k_sp_down<R, LEN>, H::SMEM);
        set_smem((const void*)k_sp_marginals<R, LEN>, H::SMEM);
        rounds<BPState<R>>(t.up, k_sp_up<R, LEN>, s, a);
        rounds<BPState<R>>(t.down, k_sp_down<R, LEN>, s, a);
        k_sp_marginals<R, LEN><<<t.N, H::BLOCK, H::SMEM, s>>>(a);
        DEEDS_LAUNCH_CHECK();
    }
    static void backward(const SoftTree& t, const BPState<R>& a, const BPGrad<R>& g, hipStream_t s) {
        set_smem((const void*)k_sp_bwd_marg<R, LEN>, H::SMEM);
        set_smem((const void*)k_sp_bwd_down<R, LEN>, H::SMEM);
        set_smem((const void*)k_sp_bwd_up<R, LEN>, H::SMEM);
        set_smem((const void*)k_sp_bwd_passes<R, LEN>, H::SMEM);
        k_sp_bwd_marg<R, LEN><<<t.N, H::BLOCK, H::SMEM, s>>>(a, g);
        DEEDS_LAUNCH_CHECK();
        k_sp_bwd_passes<R, LEN><<<2 * t.N, H::BLOCK, H::SMEM, s>>>(a, g);
        DEEDS_LAUNCH_CHECK();
        rounds<BPState<R>, BPGrad<R>>(t.up, k_sp_bwd_down<R, LEN>, s, a, g);
        rounds<BPState<R>, BPGrad<R>>(t.down, k_sp_bwd_up<R, LEN>, s, a, g);
    }
};

template <typename R, template <typename, int> class Op, typename... Args>
void dispatch_len(int L, Args&&... args) {
    switch (L) {
        case 1: Op<R, 1>::run(args...); break;
        case 3: Op<R, 3>::run(args...); break;
        case 5: Op<R, 5>::run(args...); break;
        case 7: Op<R, 7>::run(args...); break;
        case 9: Op<R, 9>::run(args...); break;
        case 11: Op<R, 11>::run(args...); break;
        case 13: Op<R, 13>::run(args...); break;
        case 15: Op<R, 15>::run(args...); break;
        case 17: Op<R, 17>::run(args...); break;
        default: fail(DEEDS_ERR_ARG, "soft BP label radius must be 0.." + std::to_string(kSoftMaxRadius));
    }
}

template <typename R, int LEN>
struct RunForward {
    static void run(const SoftTree& t, const BPState<R>& a, hipStream_t s) { Sweeps<R, LEN>::forward(t, a, s); }
};
template <typename R, int LEN>
struct RunBackward {
    static void run(const SoftTree& t, const BPState<R>& a, const BPGrad<R>& g, hipStream_t s) {
        Sweeps<R, LEN>::backward(t, a, g, s);
    }
};

}  // namespace

template <typename R>
void soft_bp_forward(const SoftTree& t, const R* cost, const R* u0, R* u1, R* saved, int hw, float quant, float T,
                     hipStream_t s) {
    const BPState<R> a = bp_state<R>(t, cost, u0, u1, saved, hw, quant, T);
    dispatch_len<R, RunForward>(2 * hw + 1, t, a, s);
}

template <typename R>
void soft_bp_backward(const SoftTree& t, const R* cost, const R* u0, const R* saved, const R* gu1, R* gcost,
                      R* scratch, int hw, float quant, float T, hipStream_t s) {
    const BPState<R> a = bp_state<R>(t, cost, u0, nullptr, const_cast<R*>(saved), hw, quant, T);
    const size_t L = 2 * (size_t)hw + 1;
    const size_t NL3 = (size_t)t.N * L * L * L;
    const BPGrad<R> g{gu1, scratch, scratch + NL3, gcost, scratch + 2 * NL3};
    dispatch_len<R, RunBackward>(2 * hw + 1, t, a, g, s);
}

template void soft_bp_forward<float>(const SoftTree&, const float*, const float*, float*, float*, int, float, float,
                                     hipStream_t);
template void soft_bp_forward<double>(const SoftTree&, const double*, const double*, double*, double*, int, float,
                                      float, hipStream_t);
template void soft_bp_backward<float>(const SoftTree&, const float*, const float*, const float*, const float*, float*,
                                      float*, int, float, float, hipStream_t);
template void soft_bp_backward<double>(const SoftTree&, const double*, const double*, const double*, const double*,
                                       double*, double*, int, float, float, hipStream_t);

}  // namespace deeds
