// MST edge costs on gfx950 (reference: primsMST.h:75-100).
//
// For each control point and each of its six grid neighbours, the float sum of
// |I(a) - I(b)| over corresponding voxels of the two step^3 blocks, summed in
// the reference's (k1, j1, i1) order so the float result is bit-identical.
// The tree itself (edge weights, Prim's, BFS order) is built on the host in
// mst_host.cpp from these costs.
#include "kernels.h"

namespace deeds {

namespace {

__constant__ int c_nb[6][3] = {{0, -1, 0}, {0, 1, 0}, {-1, 0, 0}, {1, 0, 0}, {0, 0, -1}, {0, 0, 1}};  // (di,dj,dk)

__global__ void k_edge_sad(float* __restrict__ ec, const float* __restrict__ im, int step, int m, int n, int m1,
                           int n1, int o1) {
    long long N = (long long)m1 * n1 * o1;
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 6 * N) return;
    int nb = (int)(t / N);
    long long node = t - nb * N;
    int i = (int)(node % m1);
    int j = (int)((node / m1) % n1);
    int k = (int)(node / ((long long)m1 * n1));
    int ii = i + c_nb[nb][0], jj = j + c_nb[nb][1], kk = k + c_nb[nb][2];
    float acc = 0.0f;
    if (ii >= 0 && ii < m1 && jj >= 0 && jj < n1 && kk >= 0 && kk < o1) {
        long long mn = (long long)m * n;
        const float* a = im + (long long)i * step + (long long)j * step * m + (long long)k * step * mn;
        const float* b = im + (long long)ii * step + (long long)jj * step * m + (long long)kk * step * mn;
        for (int k1 = 0; k1 < step; k1++)
            for (int j1 = 0; j1 < step; j1++) {
                long long r = k1 * mn + (long long)j1 * m;
                for (int i1 = 0; i1 < step; i1++) acc = acc + fabsf(a[r + i1] - b[r + i1]);
            }
    }
    ec[t] = acc;
}

}  // namespace

void launch_edge_sad(float* ec, const float* im, int step, int m, int n, int o, hipStream_t s) {
    int m1 = m / step, n1 = n / step, o1 = o / step;
    long long total = 6LL * m1 * n1 * o1;
    if (total == 0) return;
    k_edge_sad<<<ceil_div(total, 128), 128, 0, s>>>(ec, im, step, m, n, m1, n1, o1);
    DEEDS_LAUNCH_CHECK();
}

}  // namespace deeds
