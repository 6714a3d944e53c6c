// MST edge costs on gfx950 (reference: primsMST.h:75-100).
//
// For each control point and each of its six grid neighbours, the float sum of
// |I(a) - I(b)| over corresponding voxels of the two step^3 blocks, summed in
// the reference's (k1, j1, i1) order so the float result is bit-identical
// (each edge is summed once and written to both of its ends).
// The tree itself (edge weights, Prim's, BFS order) is built on the host in
// mst_host.cpp from these costs.
#include "kernels.h"

namespace deeds {

namespace {

// Neighbour slots (the reference's order): 0 -j, 1 +j, 2 -i, 3 +i, 4 -k, 5 +k.

// One thread per control point: the sums towards its +j, +i and +k
// neighbours (slots 1, 3, 5), with the block's voxels read once for all
// three.  Slot 0, 2, 4 of the neighbour on the other side of each edge is the
// same sum -- the same voxel pairs in the same order, and |a - b| == |b - a|
// in float -- so the thread writes it there too; slots whose neighbour lies
// outside the grid are 0, as the reference leaves them.
__global__ void k_edge_sad(float* __restrict__ ec, const float* __restrict__ im, int step, int m, int n, int m1,
                           int n1, int o1) {
    const long long N = (long long)m1 * n1 * o1;
    const long long node = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= N) return;
    const int i = (int)(node % m1);
    const int j = (int)((node / m1) % n1);
    const int k = (int)(node / ((long long)m1 * n1));
    const bool hj = j + 1 < n1, hi = i + 1 < m1, hk = k + 1 < o1;
    const long long mn = (long long)m * n;
    const float* a = im + (long long)i * step + (long long)j * step * m + (long long)k * step * mn;
    // the neighbour blocks (clamped to the node's own block where absent: their sums are not used)
    const float* bj = hj ? a + (long long)step * m : a;
    const float* bi = hi ? a + step : a;
    const float* bk = hk ? a + (long long)step * mn : a;
    float sj = 0.0f, si = 0.0f, sk = 0.0f;
    for (int k1 = 0; k1 < step; k1++)
        for (int j1 = 0; j1 < step; j1++) {
            const long long r = k1 * mn + (long long)j1 * m;
            for (int i1 = 0; i1 < step; i1++) {
                const float x = a[r + i1];
                sj = sj + fabsf(x - bj[r + i1]);
                si = si + fabsf(x - bi[r + i1]);
                sk = sk + fabsf(x - bk[r + i1]);
            }
        }
    sj = hj ? sj : 0.0f;
    si = hi ? si : 0.0f;
    sk = hk ? sk : 0.0f;
    ec[1 * N + node] = sj;
    ec[3 * N + node] = si;
    ec[5 * N + node] = sk;
    if (hj) ec[0 * N + node + m1] = sj;
    if (hi) ec[2 * N + node + 1] = si;
    if (hk) ec[4 * N + node + (long long)m1 * n1] = sk;
    if (j == 0) ec[0 * N + node] = 0.0f;
    if (i == 0) ec[2 * N + node] = 0.0f;
    if (k == 0) ec[4 * N + node] = 0.0f;
}

}  // namespace

void launch_edge_sad(float* ec, const float* im, int step, int m, int n, int o, hipStream_t s) {
    int m1 = m / step, n1 = n / step, o1 = o / step;
    long long total = (long long)m1 * n1 * o1;
    if (total == 0) return;
    k_edge_sad<<<ceil_div(total, 128), 128, 0, s>>>(ec, im, step, m, n, m1, n1, o1);
    DEEDS_LAUNCH_CHECK();
}

}  // namespace deeds
