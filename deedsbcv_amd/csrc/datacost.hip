// Hamming data-cost volume on gfx950 (reference: dataCostD.h:108-210).
//
// cost[node*len3 + l] = scale * sum_{block samples s} popcount(F[s] ^ M[s + d(l)])
// with the moving codes clamp-padded by quant*hw (:126-138) so the inner loop
// needs no bounds logic.  One workgroup per control point; its block of fixed
// codes is wave-uniform (scalar loads), each lane owns labels l, l+256, ...
// and accumulates an exact integer sum with v_xor + v_bcnt (two 32-bit halves).
#include "kernels.h"

#include <cmath>

namespace deeds {

namespace {

__global__ void k_pad_codes(uint64_t* __restrict__ Mp, const uint64_t* __restrict__ M, int m, int n, int o,
                            int pad) {
    int mp = m + 2 * pad, np = n + 2 * pad, op = o + 2 * pad;
    long long total = (long long)mp * np * op;
    long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= total) return;
    int i = (int)(p % mp);
    int j = (int)((p / mp) % np);
    int k = (int)(p / ((long long)mp * np));
    i = min(max(i - pad, 0), m - 1);
    j = min(max(j - pad, 0), n - 1);
    k = min(max(k - pad, 0), o - 1);
    Mp[p] = M[i + (long long)j * m + (long long)k * m * n];
}

// Blocks are dispatched round-robin over the 8 XCDs; hand each XCD a
// contiguous run of control points so neighbouring points, whose moving-code
// neighbourhoods overlap almost entirely, share that XCD's L2.
__device__ __forceinline__ int xcd_node(int b, int nodes) {
    const int per = (nodes + 7) / 8;
    return (b & 7) * per + (b >> 3);
}

// Sample loops unrolled at compile time (STEP and the skips are template
// arguments), so every wave keeps the moving-code loads of a whole block in
// flight instead of one at a time.
template <int STEP, int SI, int SJ, int SK>
__global__ __launch_bounds__(256) void k_datacost_t(float* __restrict__ cost, const uint64_t* __restrict__ F,
                                                    const uint64_t* __restrict__ Mp, int m, int n, int m1, int n1,
                                                    int nodes, int len, int q, float scale, int mp, int np) {
    const int node = xcd_node(blockIdx.x, nodes);
    if (node >= nodes) return;
    const int y = node % m1;
    const int x = (node / m1) % n1;
    const int z = node / (m1 * n1);
    const int len3 = len * len * len;
    const long long mn = (long long)m * n, mpnp = (long long)mp * np;
    const uint64_t* fblk = F + (long long)y * STEP + (long long)x * STEP * m + (long long)z * STEP * mn;
    const uint64_t* mblk = Mp + (long long)y * STEP + (long long)x * STEP * mp + (long long)z * STEP * mpnp;
    for (int l = threadIdx.x; l < len3; l += blockDim.x) {
        const int ls = l % len, lx = (l / len) % len, lz = l / (len * len);
        const uint64_t* mb = mblk + ls * q + (long long)(lx * q) * mp + (long long)(lz * q) * mpnp;
        unsigned sum = 0;
#pragma unroll
        for (int k = 0; k < STEP; k += SK) {
#pragma unroll
            for (int j = 0; j < STEP; j += SJ) {
                const uint64_t* fr = fblk + k * mn + (long long)j * m;
                const uint64_t* mr = mb + k * mpnp + (long long)j * mp;
#pragma unroll
                for (int i = 0; i < STEP; i += SI) sum += __builtin_popcountll(fr[i] ^ mr[i]);
            }
        }
        cost[(long long)node * len3 + l] = (float)sum * scale;
    }
}

__global__ __launch_bounds__(256) void k_datacost(float* __restrict__ cost, const uint64_t* __restrict__ F,
                                                  const uint64_t* __restrict__ Mp, int m, int n, int m1, int n1,
                                                  int step, int len, int q, int si, int sj, int sk, float scale,
                                                  int mp, int np) {
    const int node = blockIdx.x;
    const int y = node % m1;
    const int x = (node / m1) % n1;
    const int z = node / (m1 * n1);
    const int len3 = len * len * len;
    const long long mn = (long long)m * n, mpnp = (long long)mp * np;
    const uint64_t* fblk = F + (long long)y * step + (long long)x * step * m + (long long)z * step * mn;
    // padded coordinate of (block origin + label displacement) = y*step + ls*q
    const uint64_t* mblk = Mp + (long long)y * step + (long long)x * step * mp + (long long)z * step * mpnp;
    for (int l = threadIdx.x; l < len3; l += blockDim.x) {
        int ls = l % len, lx = (l / len) % len, lz = l / (len * len);
        const uint64_t* mb = mblk + ls * q + (long long)(lx * q) * mp + (long long)(lz * q) * mpnp;
        unsigned sum = 0;
        for (int k = 0; k < step; k += sk) {
            for (int j = 0; j < step; j += sj) {
                const uint64_t* fr = fblk + k * mn + (long long)j * m;
                const uint64_t* mr = mb + k * mpnp + (long long)j * mp;
                for (int i = 0; i < step; i += si) {
                    uint64_t t = fr[i] ^ mr[i];
                    sum += __builtin_popcountll(t);
                }
            }
        }
        cost[(long long)node * len3 + l] = (float)sum * scale;
    }
}

}  // namespace

CostPlan datacost_plan(int m, int n, int o, int step, int hw, float quant, float alpha, int randnum) {
    // dataCostD.h:141-163 with RAND_SAMPLES = randnum
    CostPlan p{};
    int sz = 1, sx = 1, sy = 1;
    if (step > 4) {
        if (randnum > 0) { sz = 2; sx = 2; }
        if (randnum > 1) sy = 2;
    }
    if (randnum > 1 && step > 7) { sz = 3; sx = 3; sy = 3; }
    if (step == 4 && randnum > 1) sz = 2;
    p.skip_i = sy; p.skip_j = sx; p.skip_k = sz;
    float maxsamp = std::ceil((float)step / (float)sx) * std::ceil((float)step / (float)sz) *
                    std::ceil((float)step / (float)sy);
    float alphai = (float)step / (alpha * quant);
    p.scale = (float)(0.5 * (double)alphai / (double)maxsamp);
    int q = (int)quant;
    p.pad = q * hw;
    p.mp = m + 2 * p.pad; p.np = n + 2 * p.pad; p.op = o + 2 * p.pad;
    p.samples = (long long)((step + sy - 1) / sy) * ((step + sx - 1) / sx) * ((step + sz - 1) / sz);
    return p;
}

void launch_pad_codes(uint64_t* Mp, const uint64_t* M, int m, int n, int o, int pad, hipStream_t s) {
    long long total = (long long)(m + 2 * pad) * (n + 2 * pad) * (o + 2 * pad);
    k_pad_codes<<<ceil_div(total, 256), 256, 0, s>>>(Mp, M, m, n, o, pad);
    DEEDS_LAUNCH_CHECK();
}

void launch_datacost(float* cost, const uint64_t* F, const uint64_t* Mp, int m, int n, int o, int step, int hw,
                     int q, const CostPlan& plan, hipStream_t s) {
    int m1 = m / step, n1 = n / step, o1 = o / step;
    long long nodes = (long long)m1 * n1 * o1;
    if (nodes == 0) return;
    int len = 2 * hw + 1;
    const unsigned grid = (unsigned)(8 * ((nodes + 7) / 8));
    const int N = (int)nodes;
#define DEEDS_DC(ST, SI, SJ, SK)                                                                              \
    if (step == ST && plan.skip_i == SI && plan.skip_j == SJ && plan.skip_k == SK) {                        \
        k_datacost_t<ST, SI, SJ, SK><<<grid, 256, 0, s>>>(cost, F, Mp, m, n, m1, n1, N, len, q, plan.scale, \
                                                          plan.mp, plan.np);                                \
        DEEDS_LAUNCH_CHECK();                                                                               \
        return;                                                                                             \
    }
    // the schedules of deedsBCV (RAND_SAMPLES = 1: skip 2 in j and k above step 4)
    DEEDS_DC(8, 1, 2, 2) DEEDS_DC(7, 1, 2, 2) DEEDS_DC(6, 1, 2, 2) DEEDS_DC(5, 1, 2, 2)
    DEEDS_DC(4, 1, 1, 1) DEEDS_DC(3, 1, 1, 1) DEEDS_DC(2, 1, 1, 1)
#undef DEEDS_DC
    k_datacost<<<(unsigned)nodes, 256, 0, s>>>(cost, F, Mp, m, n, m1, n1, step, len, q, plan.skip_i, plan.skip_j,
                                               plan.skip_k, plan.scale, plan.mp, plan.np);
    DEEDS_LAUNCH_CHECK();
}

}  // namespace deeds
