// GPU half of the affine pre-alignment (reference: linearBCV.cpp:144-263,
// affineLTS1.h:39-236).  The MIND codes and Hamming cost volumes come from
// the registration kernels (alpha 1); these kernels add
//   k_lin_warp     the affine-only trilinear warp, zero where the 2x2x2
//                  neighbourhood leaves the volume (affineLTS1.h:39-75)
//   k_lin_smooth   the in-place cost smoothing c[p] += 0.5 c[p+1] + 0.5 c[p-1]
//                  along one node axis (affineLTS1.h:122-148): nodes are
//                  visited in index order, so p-1 is already smoothed and p+1
//                  is not -- one thread walks one (line, label) in order
//   k_lin_nodes    per node the first minimum of both smoothed volumes and
//                  the background / border masks (affineLTS1.h:150-190)
// The robust fit itself (quantile, point lists, trimmed least squares) runs
// on the host (linear_host.cpp).
#include "kernels.h"

namespace deeds {

namespace {

// affineLTS1.h:47-71 for one voxel: float coordinates, double weights except
// the float products dx*dy (term 5) and dx*dy*dz*in (term 8), summed in order.
__global__ void k_lin_warp(float* __restrict__ out, const float* __restrict__ in, Affine A, int m, int n, int o) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long sz = (long long)m * n * o;
    if (p >= sz) return;
    const int i = (int)(p % m), j = (int)((p / m) % n), k = (int)(p / ((long long)m * n));
    const float* X = A.r;
    const float y1 = (float)i * X[0] + (float)j * X[1] + (float)k * X[2] + X[3];
    const float x1 = (float)i * X[4] + (float)j * X[5] + (float)k * X[6] + X[7];
    const float z1 = (float)i * X[8] + (float)j * X[9] + (float)k * X[10] + X[11];
    const int x = (int)floorf(x1), y = (int)floorf(y1), z = (int)floorf(z1);
    if (y < 0 || y >= m - 1 || x < 0 || x >= n - 1 || z < 0 || z >= o - 1) {
        out[p] = 0.0f;
        return;
    }
    const float dx = x1 - (float)x, dy = y1 - (float)y, dz = z1 - (float)z;
    const double ddx = dx, ddy = dy, ddz = dz;
    const long long s1 = m, s2 = (long long)m * n;
    const float* b = in + y + x * s1 + z * s2;
    double s = (1.0 - ddx) * (1.0 - ddy) * (1.0 - ddz) * (double)b[0];
    s += (1.0 - ddx) * ddy * (1.0 - ddz) * (double)b[1];
    s += ddx * (1.0 - ddy) * (1.0 - ddz) * (double)b[s1];
    s += (1.0 - ddx) * (1.0 - ddy) * ddz * (double)b[s2];
    s += (double)(dx * dy) * (1.0 - ddz) * (double)b[1 + s1];
    s += (1.0 - ddx) * ddy * ddz * (double)b[1 + s2];
    s += ddx * (1.0 - ddy) * ddz * (double)b[s1 + s2];
    s += (double)(dx * dy * dz * b[1 + s1 + s2]);
    out[p] = (float)s;
}

// Axis AX (0: y / stride 1, 1: x / stride m1, 2: z / stride m1*n1).  Thread =
// (line, label), labels fastest so a wave reads consecutive floats of a node.
template <int AX>
__global__ void k_lin_smooth(float* __restrict__ c, int m1, int n1, int o1, int len3) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int lines = AX == 0 ? n1 * o1 : AX == 1 ? m1 * o1 : m1 * n1;
    if (t >= (long long)lines * len3) return;
    const int line = (int)(t / len3), l = (int)(t % len3);
    int base, L;
    long long stride;
    if (AX == 0) {
        base = line * m1;  // (x1, z1) -> x1*m1 + z1*m1*n1 == line*m1
        L = m1;
        stride = 1;
    } else if (AX == 1) {
        const int y1 = line % m1, z1 = line / m1;
        base = y1 + z1 * m1 * n1;
        L = n1;
        stride = m1;
    } else {
        base = line;  // y1 + x1*m1
        L = o1;
        stride = (long long)m1 * n1;
    }
    float* p = c + (long long)base * len3 + l;
    const long long st = stride * len3;
    float cur = p[0];
    float prev = cur;  // at the first node the "previous" neighbour is the node itself, unsmoothed
    for (int q = 0; q < L; q++) {
        const float next = q + 1 < L ? p[(q + 1) * st] : cur;  // the last node's next is itself
        const double add = 0.5 * (double)next + 0.5 * (double)prev;
        const float nv = (float)((double)cur + add);
        p[q * st] = nv;
        prev = nv;
        cur = next;
    }
}

// One workgroup per node: first minimum (value, index) of both volumes,
// block means of both images (sequential float sums in k, j, i order) and the
// masks: background (mean < 5) and the two-node border get 1e20.
__global__ __launch_bounds__(256) void k_lin_nodes(const float* __restrict__ c1, const float* __restrict__ c2,
                                                   const float* __restrict__ im1, const float* __restrict__ im2,
                                                   int m, int n, int step, int m1, int n1, int o1, int len3,
                                                   float* __restrict__ minval, int* __restrict__ minind) {
    __shared__ float rv[2][4];
    __shared__ int ri[2][4];
    const int node = blockIdx.x;
    const float* a = c1 + (long long)node * len3;
    const float* b = c2 + (long long)node * len3;
    float va = 3.402823466e38f, vb = 3.402823466e38f;
    int ia = 0x7fffffff, ib = 0x7fffffff;
    for (int l = threadIdx.x; l < len3; l += blockDim.x) {
        const float x = a[l], y = b[l];
        if (x < va) { va = x; ia = l; }  // l increases: the first minimum of this thread
        if (y < vb) { vb = y; ib = l; }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const float ova = __shfl_xor(va, off, 64), ovb = __shfl_xor(vb, off, 64);
        const int oia = __shfl_xor(ia, off, 64), oib = __shfl_xor(ib, off, 64);
        if (ova < va || (ova == va && oia < ia)) { va = ova; ia = oia; }
        if (ovb < vb || (ovb == vb && oib < ib)) { vb = ovb; ib = oib; }
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) {
        rv[0][wid] = va; ri[0][wid] = ia;
        rv[1][wid] = vb; ri[1][wid] = ib;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) {
        if (rv[0][w] < va || (rv[0][w] == va && ri[0][w] < ia)) { va = rv[0][w]; ia = ri[0][w]; }
        if (rv[1][w] < vb || (rv[1][w] == vb && ri[1][w] < ib)) { vb = rv[1][w]; ib = ri[1][w]; }
    }
    const int z1 = node / (m1 * n1), x1 = (node - z1 * m1 * n1) / m1, y1 = node - z1 * m1 * n1 - x1 * m1;
    float mean1 = 0.0f, mean2 = 0.0f;
    const long long mn = (long long)m * n;
    for (int k1 = 0; k1 < step; k1++)
        for (int j1 = 0; j1 < step; j1++)
            for (int i1 = 0; i1 < step; i1++) {
                const long long q = (y1 * step + i1) + (long long)(x1 * step + j1) * m + (long long)(z1 * step + k1) * mn;
                mean1 += im1[q];
                mean2 += im2[q];
            }
    const float vol = (float)(step * step * step);
    mean1 /= vol;
    mean2 /= vol;
    if (mean1 < 5.0f) va = 1e20f;
    if (mean2 < 5.0f) vb = 1e20f;
    if (z1 < 2 || z1 > o1 - 2 || x1 < 2 || x1 > n1 - 2 || y1 < 2 || y1 > m1 - 2) {
        va = 1e20f;
        vb = 1e20f;
    }
    const int N = m1 * n1 * o1;
    minval[node] = va;
    minind[node] = ia;
    minval[N + node] = vb;
    minind[N + node] = ib;
}

__global__ void k_lin_offset(float* __restrict__ out, const float* __restrict__ in, float sub, long long n) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) out[p] = in[p] - sub;
}

}  // namespace

void launch_lin_offset(float* out, const float* in, float sub, long long n, hipStream_t s) {
    if (n == 0) return;
    k_lin_offset<<<ceil_div(n, 256), 256, 0, s>>>(out, in, sub, n);
    DEEDS_LAUNCH_CHECK();
}

void launch_lin_warp(float* out, const float* in, const float* X, int m, int n, int o, hipStream_t s) {
    const long long sz = (long long)m * n * o;
    if (sz == 0) return;
    Affine A;
    for (int t = 0; t < 12; t++) A.r[t] = X[t];
    k_lin_warp<<<ceil_div(sz, 256), 256, 0, s>>>(out, in, A, m, n, o);
    DEEDS_LAUNCH_CHECK();
}

void launch_lin_smooth(float* cost, int m1, int n1, int o1, int len3, hipStream_t s) {
    const long long total[3] = {(long long)n1 * o1 * len3, (long long)m1 * o1 * len3, (long long)m1 * n1 * len3};
    k_lin_smooth<0><<<ceil_div(total[0], 256), 256, 0, s>>>(cost, m1, n1, o1, len3);
    DEEDS_LAUNCH_CHECK();
    k_lin_smooth<1><<<ceil_div(total[1], 256), 256, 0, s>>>(cost, m1, n1, o1, len3);
    DEEDS_LAUNCH_CHECK();
    k_lin_smooth<2><<<ceil_div(total[2], 256), 256, 0, s>>>(cost, m1, n1, o1, len3);
    DEEDS_LAUNCH_CHECK();
}

void launch_lin_nodes(const float* c1, const float* c2, const float* im1, const float* im2, int m, int n, int step,
                      int m1, int n1, int o1, int len3, float* minval, int* minind, hipStream_t s) {
    const int N = m1 * n1 * o1;
    if (N == 0) return;
    k_lin_nodes<<<N, 256, 0, s>>>(c1, c2, im1, im2, m, n, step, m1, n1, o1, len3, minval, minind);
    DEEDS_LAUNCH_CHECK();
}

}  // namespace deeds
