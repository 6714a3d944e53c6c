// MIND-SSC descriptor on gfx950 (reference: MINDSSCbox.h:2-232).
//
// Pipeline per call, all in HBM-resident float planes (6 channels):
//   k_ssd6      squared difference to the six shifted copies      (:79-90, :113-117)
//   k_box_line  per-line sequential prefix sum + clamped window    (:2-76), once per axis
//   k_mind_code 12 SSC values -> min/noise normalise -> 5-level    (:172-222)
//               thermometer fields packed into a 64-bit code
// The box filter must reproduce the reference's float rounding: each line is
// one sequential prefix sum (values exceed 2^24, so a tree scan would change
// bits), hence one thread per line.  Lines along j and k are coalesced across
// the wavefront (consecutive lanes = consecutive i); lines along i are
// per-lane sequential and served from L1/L2.
#include "kernels.h"

#include <cmath>

namespace deeds {

namespace {

// shift of the six SSD channels in units of qs, as (di, dj, dk)
__constant__ int c_ssd_shift[6][3] = {{+1, +1, 0}, {-1, +1, 0}, {0, -1, +1}, {-1, 0, +1}, {0, +1, +1}, {+1, 0, +1}};
// twelve SSC pairs: neighbour offset (di, dj, dk) in units of qs and channel
__constant__ int c_ssc[12][4] = {{0, -1, 0, 0}, {-1, 0, 0, 0}, {0, -1, 0, 1}, {+1, 0, 0, 1},
                                 {0, 0, -1, 2}, {0, +1, 0, 2}, {0, 0, -1, 3}, {+1, 0, 0, 3},
                                 {0, 0, -1, 4}, {0, -1, 0, 4}, {0, 0, -1, 5}, {-1, 0, 0, 5}};

__global__ void k_ssd6(float* __restrict__ out, const float* __restrict__ im, int m, int n, int o, int qs) {
    long long sz = (long long)m * n * o;
    long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= sz) return;
    int i = (int)(p % m);
    int j = (int)((p / m) % n);
    int k = (int)(p / ((long long)m * n));
    float c = im[p];
#pragma unroll
    for (int l = 0; l < 6; l++) {
        int ii = i + c_ssd_shift[l][0] * qs, jj = j + c_ssd_shift[l][1] * qs, kk = k + c_ssd_shift[l][2] * qs;
        bool in = ii >= 0 && ii < m && jj >= 0 && jj < n && kk >= 0 && kk < o;
        float s = in ? im[ii + (long long)jj * m + (long long)kk * m * n] : c;
        float d = s - c;
        out[l * sz + p] = d * d;
    }
}

// One thread per (line, channel).  axis 0: lines along i (stride 1),
// axis 1: along j (stride m), axis 2: along k (stride m*n).  The prefix sums
// go to `pre`, the windowed differences to `dst`.
__global__ void k_box_line(float* __restrict__ dst, float* __restrict__ pre, const float* __restrict__ src,
                           int m, int n, int o, int axis, int hw) {
    long long sz = (long long)m * n * o;
    long long lines = axis == 0 ? (long long)n * o : axis == 1 ? (long long)m * o : (long long)m * n;
    long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= lines * 6) return;
    int ch = (int)(tid / lines);
    long long line = tid - ch * lines;
    long long base, stride;
    int L;
    if (axis == 0) {
        base = line * m; stride = 1; L = m;
    } else if (axis == 1) {
        long long i = line % m, k = line / m;
        base = i + k * (long long)m * n; stride = m; L = n;
    } else {
        base = line; stride = (long long)m * n; L = o;
    }
    base += ch * sz;
    const float* x = src + base;
    float* P = pre + base;
    float* y = dst + base;
    float acc = x[0];
    P[0] = acc;
    for (int t = 1; t < L; t++) {
        acc = acc + x[t * stride];
        P[t * stride] = acc;
    }
    for (int t = 0; t < L; t++) {
        int hi = t + hw < L - 1 ? t + hw : L - 1;
        float v = P[hi * stride];
        if (t > hw) v = v - P[(t - hw - 1) * stride];
        y[t * stride] = v;
    }
}

struct Thr {
    float t[5];
};

__global__ void k_mind_code(uint64_t* __restrict__ out, const float* __restrict__ d, int m, int n, int o, int qs,
                            Thr thr) {
    long long sz = (long long)m * n * o;
    long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= sz) return;
    int i = (int)(p % m);
    int j = (int)((p / m) % n);
    int k = (int)(p / ((long long)m * n));
    float v[12];
#pragma unroll
    for (int l = 0; l < 12; l++) {
        int ii = i + c_ssc[l][0] * qs, jj = j + c_ssc[l][1] * qs, kk = k + c_ssc[l][2] * qs;
        bool in = ii >= 0 && ii < m && jj >= 0 && jj < n && kk >= 0 && kk < o;
        long long q = in ? ii + (long long)jj * m + (long long)kk * m * n : p;
        v[l] = d[c_ssc[l][3] * sz + q];
    }
    float mn = v[0];
#pragma unroll
    for (int l = 1; l < 12; l++) mn = v[l] < mn ? v[l] : mn;
    float sum = 0.0f;
#pragma unroll
    for (int l = 0; l < 12; l++) {
        v[l] = v[l] - mn;
        sum = sum + v[l];
    }
    float noise = sum / 12.0f;
    noise = noise < 1e-6f ? 1e-6f : noise;
    uint64_t code = 0;
#pragma unroll
    for (int l = 0; l < 12; l++) {
        float qv = v[l] / noise;
        int lev = 0;
#pragma unroll
        for (int c = 0; c < 5; c++) lev += thr.t[c] > qv ? 1 : 0;
        code |= (uint64_t)((1u << lev) - 1u) << (5 * l);
    }
    out[p] = code;
}

}  // namespace

void mind_thresholds(float thr[5]) {
    // -log((c+1.5)/6) rounded once to float (the reference evaluates the float
    // overload of log on a float argument; both agree for these five values,
    // which a test pins).
    for (int c = 0; c < 5; c++) thr[c] = (float)(-std::log((double)((c + 1.5f) / 6)));
}

void launch_descriptor(uint64_t* out, const float* im, int m, int n, int o, int qs, float* w0, float* w1,
                       float* w2, hipStream_t s) {
    long long sz = (long long)m * n * o;
    const int B = 256;
    k_ssd6<<<ceil_div(sz, B), B, 0, s>>>(w0, im, m, n, o, qs);
    DEEDS_LAUNCH_CHECK();
    // i pass: w0 -> w1, j pass: w1 -> w0, k pass: w0 -> w1 (w2 holds prefix sums)
    long long li = (long long)n * o * 6, lj = (long long)m * o * 6, lk = (long long)m * n * 6;
    k_box_line<<<ceil_div(li, 64), 64, 0, s>>>(w1, w2, w0, m, n, o, 0, qs);
    DEEDS_LAUNCH_CHECK();
    k_box_line<<<ceil_div(lj, B), B, 0, s>>>(w0, w2, w1, m, n, o, 1, qs);
    DEEDS_LAUNCH_CHECK();
    k_box_line<<<ceil_div(lk, B), B, 0, s>>>(w1, w2, w0, m, n, o, 2, qs);
    DEEDS_LAUNCH_CHECK();
    Thr thr;
    mind_thresholds(thr.t);
    k_mind_code<<<ceil_div(sz, B), B, 0, s>>>(out, w1, m, n, o, qs, thr);
    DEEDS_LAUNCH_CHECK();
}

}  // namespace deeds
