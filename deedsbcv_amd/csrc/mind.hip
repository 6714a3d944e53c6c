// MIND-SSC descriptor on gfx950 (reference: MINDSSCbox.h:2-232).
//
// Pipeline per call (6 float channels in HBM):
//   k_ssd_box_ij  squared difference to the shifted image (:79-90, :113-117)
//                 fused with the box passes along i and j (:2-50), one
//                 workgroup per slice and channel (box radius 1..3, m <= 1024)
//   k_box_lines   box pass along k (:52-73)
//   k_mind_code   12 SSC values -> min/noise normalisation -> 6-level
//                 thermometer, packed into a 64-bit code (:172-222)
// Other radii and sizes take k_ssd_box_i (the SSD and the i pass) and
// k_box_cols for j and k.
// The box filter must reproduce the reference's float rounding: every line is
// one sequential float prefix sum (the sums exceed 2^24, so a tree scan would
// change bits) followed by the clamped window difference.  Along i (lines
// contiguous in memory) a workgroup stages TILE lines in LDS with coalesced
// loads, one thread per line runs the sequential prefix in LDS, and all
// threads write the window differences back coalesced.  Along j and k the
// lanes of a wave are consecutive i, so one thread per line streams its line
// straight from HBM with coalesced loads and keeps the prefix in registers.
#include "kernels.h"

#include <cmath>

namespace deeds {

namespace {

constexpr int TILE = 32;      // lines per workgroup (j, k passes of k_box_cols)
constexpr int TILE_I = 16;    // lines per workgroup of the i pass
constexpr int BOX_THREADS = 256;

// shift of the six SSD channels in units of qs, as (di, dj, dk)
__constant__ int c_ssd_shift[6][3] = {{+1, +1, 0}, {-1, +1, 0}, {0, -1, +1}, {-1, 0, +1}, {0, +1, +1}, {+1, 0, +1}};
// twelve SSC pairs: neighbour offset (di, dj, dk) in units of qs and channel
__constant__ int c_ssc[12][4] = {{0, -1, 0, 0}, {-1, 0, 0, 0}, {0, -1, 0, 1}, {+1, 0, 0, 1},
                                 {0, 0, -1, 2}, {0, +1, 0, 2}, {0, 0, -1, 3}, {+1, 0, 0, 3},
                                 {0, 0, -1, 4}, {0, -1, 0, 4}, {0, 0, -1, 5}, {-1, 0, 0, 5}};

// Sequential prefix sum of `L` values at p[0], p[stride], ... in place.
__device__ __forceinline__ void prefix_in_place(float* p, int L, int stride) {
    float acc = p[0];
    int t = 1;
    for (; t + 4 <= L; t += 4) {
        float x0 = p[t * stride], x1 = p[(t + 1) * stride], x2 = p[(t + 2) * stride], x3 = p[(t + 3) * stride];
        acc = acc + x0;
        p[t * stride] = acc;
        acc = acc + x1;
        p[(t + 1) * stride] = acc;
        acc = acc + x2;
        p[(t + 2) * stride] = acc;
        acc = acc + x3;
        p[(t + 3) * stride] = acc;
    }
    for (; t < L; t++) {
        acc = acc + p[t * stride];
        p[t * stride] = acc;
    }
}

// Clamped window of the prefix sums P (MINDSSCbox.h:17-27): P[min(t+hw,L-1)] - P[t-hw-1], second term absent for t <= hw.
__device__ __forceinline__ float window(const float* P, int t, int L, int hw, int stride) {
    int hi = t + hw < L - 1 ? t + hw : L - 1;
    float v = P[hi * stride];
    if (t > hw) v = v - P[(t - hw - 1) * stride];
    return v;
}

// i pass: tile = TILE rows (j0 .. j0+TILE-1 at fixed k) of channel ch; LDS row stride m+1.
__global__ __launch_bounds__(BOX_THREADS) void k_ssd_box_i(float* __restrict__ out, const float* __restrict__ im,
                                                            int m, int n, int o, int qs) {
    extern __shared__ float tile[];
    const int j0 = blockIdx.x * TILE_I, k = blockIdx.y, ch = blockIdx.z;
    const int rows = min(TILE_I, n - j0);
    const int ld = m + 1;
    const long long mn = (long long)m * n;
    const int di = c_ssd_shift[ch][0] * qs, dj = c_ssd_shift[ch][1] * qs, dk = c_ssd_shift[ch][2] * qs;
    const int kk = k + dk;
    const bool k_in = kk >= 0 && kk < o;
    // thread = column i; the TILE_I rows' loads are unrolled so they are all in flight together
    for (int ib = 0; ib < m; ib += BOX_THREADS) {
        const int i = ib + (int)threadIdx.x, ii = i + di;
        const bool i_ok = i < m, ii_in = ii >= 0 && ii < m;
        float cv[TILE_I], sv[TILE_I];
#pragma unroll
        for (int r = 0; r < TILE_I; r++) {
            const int j = j0 + r, jj = j + dj;
            const bool ok = i_ok && r < rows;
            const bool in = ok && ii_in && jj >= 0 && jj < n && k_in;
            cv[r] = ok ? im[i + (long long)j * m + k * mn] : 0.0f;
            sv[r] = in ? im[ii + (long long)jj * m + kk * mn] : cv[r];
        }
#pragma unroll
        for (int r = 0; r < TILE_I; r++) {
            if (i_ok && r < rows) {
                const float d = sv[r] - cv[r];
                tile[r * ld + i] = d * d;
            }
        }
    }
    __syncthreads();
    if ((int)threadIdx.x < rows) prefix_in_place(tile + threadIdx.x * ld, m, 1);
    __syncthreads();
    float* dst = out + (long long)ch * mn * o;
    for (int ib = 0; ib < m; ib += BOX_THREADS) {
        const int i = ib + (int)threadIdx.x;
        if (i >= m) break;
#pragma unroll
        for (int r = 0; r < TILE_I; r++)
            if (r < rows) dst[i + (long long)(j0 + r) * m + k * mn] = window(tile + r * ld, i, m, qs, 1);
    }
}

// j pass (axis 1) or k pass (axis 2): tile = TILE consecutive i at fixed
// (k) resp. (j), lines along the axis; LDS layout [t][r] with stride TILE+1.
__global__ __launch_bounds__(BOX_THREADS) void k_box_cols(float* __restrict__ dst, const float* __restrict__ src,
                                                           int m, int n, int o, int axis, int hw) {
    extern __shared__ float tile[];
    const int i0 = blockIdx.x * TILE, other = blockIdx.y, ch = blockIdx.z;
    const int cols = min(TILE, m - i0);
    const long long mn = (long long)m * n;
    const int L = axis == 1 ? n : o;
    const long long stride = axis == 1 ? (long long)m : mn;
    const long long base = (long long)ch * mn * o + i0 + (axis == 1 ? other * mn : (long long)other * m);
    constexpr int ld = TILE + 1;
    for (int idx = threadIdx.x; idx < L * TILE; idx += blockDim.x) {
        const int t = idx / TILE, r = idx - t * TILE;
        if (r < cols) tile[t * ld + r] = src[base + t * stride + r];
    }
    __syncthreads();
    if ((int)threadIdx.x < cols) prefix_in_place(tile + threadIdx.x, L, ld);
    __syncthreads();
    for (int idx = threadIdx.x; idx < L * TILE; idx += blockDim.x) {
        const int t = idx / TILE, r = idx - t * TILE;
        if (r < cols) dst[base + t * stride + r] = window(tile + r, t, L, hw, ld);
    }
}

// j pass (axis 1) or k pass (axis 2) with one thread per line, streaming:
// lanes are consecutive i, so every load and store is coalesced, and the
// line's prefix sums never leave registers -- a running float sum (the
// reference's sequential order) and the last 2*HW + 2 prefix values, enough
// for the clamped window difference of the output HW steps behind.
template <int HW>
__global__ __launch_bounds__(256) void k_box_lines(float* __restrict__ dst, const float* __restrict__ src, int m,
                                                   int n, int o, int axis) {
    const long long mn = (long long)m * n, sz = mn * o;
    const int other_count = axis == 1 ? o : n;
    const long long lines = (long long)m * other_count * 6;
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= lines) return;
    const int i = (int)(id % m);
    const long long rest = id / m;
    const int other = (int)(rest % other_count), ch = (int)(rest / other_count);
    const int L = axis == 1 ? n : o;
    const long long stride = axis == 1 ? (long long)m : mn;
    const long long base = ch * sz + i + (axis == 1 ? other * mn : (long long)other * m);
    const float* sp = src + base;
    float* dp = dst + base;
    constexpr int H = 2 * HW + 2;
    float hist[H];  // after step t: hist[q] = P[t - 2*HW - 1 + q]
#pragma unroll
    for (int q = 0; q < H; q++) hist[q] = 0.0f;
    float acc = 0.0f;
    constexpr int U = 8;
    for (int tb = 0; tb < L; tb += U) {
        float x[U];
#pragma unroll
        for (int u = 0; u < U; u++) x[u] = tb + u < L ? sp[(tb + u) * stride] : 0.0f;
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int t = tb + u;
            if (t < L) {
                acc = t == 0 ? x[u] : acc + x[u];
#pragma unroll
                for (int q = 0; q + 1 < H; q++) hist[q] = hist[q + 1];
                hist[H - 1] = acc;
                // output t - HW: P[min(t, L-1)] - P[t - 2HW - 1] (MINDSSCbox.h:17-27)
                const int so = t - HW;
                if (so >= 0) {
                    float v = acc;
                    if (so > HW) v = v - hist[0];
                    dp[so * stride] = v;
                }
            }
        }
    }
    // the last HW outputs clamp the window's upper end at L - 1
#pragma unroll
    for (int e = 0; e < HW; e++) {
        const int so = L - HW + e;
        float v = hist[H - 1];
        if (so > HW) v = v - hist[e + 1];
        dp[so * stride] = v;
    }
}

// Logical tile of workgroup b in a 1-D launch of 8 * per workgroups: the
// GPU's eight XCDs take workgroups round robin (b mod 8), so XCD x gets the
// contiguous tiles [x * per, (x + 1) * per) and the neighbouring lines a tile
// rereads sit in that XCD's L2.  Tiles >= the real count are skipped by the caller.
__device__ __forceinline__ int xcd_tile(int b, int per) { return (b & 7) * per + (b >> 3); }

// i and j passes fused, with the SSD to the shifted image (MINDSSCbox.h:
// 79-90, 113-117, then the box passes :2-50 along i and j).  One workgroup
// per (slice k, channel), one thread per column i (blockDim >= m).  The
// slice goes through LDS in chunks of R rows: the threads fill the chunk
// with squared differences (coalesced), one thread per row runs the row's
// sequential i prefix in place, then every thread walks its column down the
// chunk, taking the i window from LDS and carrying the j prefix (a running
// float sum and the last 2*HW + 2 prefix values) in registers from chunk to
// chunk, and writes the j window HW rows behind.  The arithmetic of every
// value is that of the separate passes; only the i-filtered planes no longer
// go through HBM.  Tiles: (k, ch) in xcd_tile order, k-major with the six
// channels of a slice adjacent, so the slice and its shifted copy are read
// from HBM about once.
template <int HW>
__global__ __launch_bounds__(1024) void k_ssd_box_ij(float* __restrict__ out, const float* __restrict__ im, int m,
                                                      int n, int o, int R, int per) {
    extern __shared__ float tile[];
    const int tid = xcd_tile((int)blockIdx.x, per);
    if (tid >= 6 * o) return;
    const int k = tid / 6, ch = tid - 6 * (tid / 6);
    const int ld = m + 1;
    const long long mn = (long long)m * n;
    const int di = c_ssd_shift[ch][0] * HW, dj = c_ssd_shift[ch][1] * HW, dk = c_ssd_shift[ch][2] * HW;
    const int kk = k + dk;
    const bool k_in = kk >= 0 && kk < o;
    const int i = threadIdx.x;
    const bool col = i < m;
    const int ic = col ? i : m - 1;  // clamped column: loads need no branch
    const int ii = ic + di;
    const bool ii_in = ii >= 0 && ii < m;
    const float* cur = im + (long long)k * mn;
    const float* sh = im + (long long)(k_in ? kk : k) * mn;
    float* dst = out + (long long)ch * mn * o + (long long)k * mn;
    constexpr int H = 2 * HW + 2;
    float hist[H];  // after row t: hist[q] = P[t - 2*HW - 1 + q] of the column's j prefix
#pragma unroll
    for (int q = 0; q < H; q++) hist[q] = 0.0f;
    float acc = 0.0f;
    constexpr int U = 8;
    for (int j0 = 0; j0 < n; j0 += R) {
        const int rows = min(R, n - j0);
        // fill the chunk: every load of a group of U rows in flight before its stores
        for (int rb = 0; rb < rows; rb += U) {
            float cv[U], sv[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int j = min(j0 + rb + u, n - 1), jj = j + dj;
                const bool in = ii_in && k_in && jj >= 0 && jj < n;
                cv[u] = cur[ic + (long long)j * m];
                sv[u] = sh[(in ? ii : ic) + (long long)(in ? jj : j) * m];
                if (!in) sv[u] = cv[u];
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
                if (col && rb + u < rows) {
                    const float d = sv[u] - cv[u];
                    tile[(rb + u) * ld + i] = d * d;
                }
            }
        }
        __syncthreads();
        if (i < rows) prefix_in_place(tile + i * ld, m, 1);
        __syncthreads();
        if (col) {
            for (int r = 0; r < rows; r++) {
                const float x = window(tile + r * ld, i, m, HW, 1);
                const int t = j0 + r;
                acc = t == 0 ? x : acc + x;
#pragma unroll
                for (int q = 0; q + 1 < H; q++) hist[q] = hist[q + 1];
                hist[H - 1] = acc;
                // output t - HW: P[min(t, n-1)] - P[t - 2HW - 1] (MINDSSCbox.h:17-27)
                const int so = t - HW;
                if (so >= 0) {
                    float v = acc;
                    if (so > HW) v = v - hist[0];
                    dst[i + (long long)so * m] = v;
                }
            }
        }
        __syncthreads();  // the next chunk overwrites the tile
    }
    if (!col) return;
    // the last HW rows clamp the window's upper end at n - 1
#pragma unroll
    for (int e = 0; e < HW; e++) {
        const int so = n - HW + e;
        if (so < 0) continue;
        float v = hist[H - 1];
        if (so > HW) v = v - hist[e + 1];
        dst[i + (long long)so * m] = v;
    }
}

struct Thr {
    float t[5];
};

// One thread per voxel; a workgroup takes `nbi` consecutive chunks of one
// row (j, k).  The rows are ordered in blocks of kCodeRows j-rows, every k
// of a block before the next block, and spread over the XCDs in contiguous
// ranges (xcd_tile): the neighbours a voxel reads at j +- qs and k - qs are
// then rows its XCD read shortly before, found in its L2.
constexpr int kCodeRows = 32;

__global__ __launch_bounds__(256) void k_mind_code(uint64_t* __restrict__ out, const float* __restrict__ d, int m,
                                                   int n, int o, int qs, Thr thr, int nbi, int per) {
    const int tile = xcd_tile((int)blockIdx.x, per);
    if (tile >= nbi * n * o) return;
    const int r = tile / nbi;
    const int i = (tile - r * nbi) * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const int jb = r / (kCodeRows * o), rem = r - jb * kCodeRows * o;
    const int rows_b = min(kCodeRows, n - jb * kCodeRows);
    const int k = rem / rows_b, j = jb * kCodeRows + (rem - k * rows_b);
    const long long sz = (long long)m * n * o;
    const long long p = i + (long long)j * m + (long long)k * m * n;
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

size_t rows_smem(int m) { return (size_t)TILE_I * (m + 1) * sizeof(float); }

// the fused i/j pass: one thread per column (at most 1024), and rows per LDS
// chunk so that the chunk stays near 40 KiB (several workgroups per CU)
constexpr int kIJMaxM = 1024;
int ij_rows(int m) { return std::max(8, std::min(32, 10240 / (m + 1))); }
size_t cols_smem(int L) { return (size_t)L * (TILE + 1) * sizeof(float); }

}  // namespace

int mind_max_dim() { return (int)(kMaxLds / ((TILE + 1) * sizeof(float))) - 1; }

void mind_thresholds(float thr[5]) {
    // -log((c+1.5)/6) rounded once to float (the reference evaluates the float
    // overload of log on a float argument; both agree for these five values,
    // which a test pins).
    for (int c = 0; c < 5; c++) thr[c] = (float)(-std::log((double)((c + 1.5f) / 6)));
}

void launch_descriptor(uint64_t* out, const float* im, int m, int n, int o, int qs, float* w0, float* w1,
                       hipStream_t s) {
    static bool attr = false;
    if (!attr) {
        DEEDS_HIP(hipFuncSetAttribute((const void*)k_ssd_box_i, hipFuncAttributeMaxDynamicSharedMemorySize, kMaxLds));
        DEEDS_HIP(hipFuncSetAttribute((const void*)k_box_cols, hipFuncAttributeMaxDynamicSharedMemorySize, kMaxLds));
        attr = true;
    }
    if (qs >= 1 && qs <= 3 && m <= kIJMaxM) {
        // i and j passes fused: im -> w1; k pass: w1 -> w0
        const int R = ij_rows(m);
        const int per = (int)ceil_div(6LL * o, 8);
        const int threads = 64 * (int)ceil_div(m, 64);
        const size_t smem = (size_t)R * (m + 1) * sizeof(float);
        const unsigned gk = (unsigned)ceil_div((long long)m * n * 6, 256);
        if (qs == 1) {
            k_ssd_box_ij<1><<<8 * per, threads, smem, s>>>(w1, im, m, n, o, R, per);
            k_box_lines<1><<<gk, 256, 0, s>>>(w0, w1, m, n, o, 2);
        } else if (qs == 2) {
            k_ssd_box_ij<2><<<8 * per, threads, smem, s>>>(w1, im, m, n, o, R, per);
            k_box_lines<2><<<gk, 256, 0, s>>>(w0, w1, m, n, o, 2);
        } else {
            k_ssd_box_ij<3><<<8 * per, threads, smem, s>>>(w1, im, m, n, o, R, per);
            k_box_lines<3><<<gk, 256, 0, s>>>(w0, w1, m, n, o, 2);
        }
        DEEDS_LAUNCH_CHECK();
    } else {
        // i pass (with the SSD): im -> w0; j pass: w0 -> w1; k pass: w1 -> w0
        k_ssd_box_i<<<dim3(ceil_div(n, TILE_I), o, 6), BOX_THREADS, rows_smem(m), s>>>(w0, im, m, n, o, qs);
        DEEDS_LAUNCH_CHECK();
        k_box_cols<<<dim3(ceil_div(m, TILE), o, 6), BOX_THREADS, cols_smem(n), s>>>(w1, w0, m, n, o, 1, qs);
        DEEDS_LAUNCH_CHECK();
        k_box_cols<<<dim3(ceil_div(m, TILE), n, 6), BOX_THREADS, cols_smem(o), s>>>(w0, w1, m, n, o, 2, qs);
        DEEDS_LAUNCH_CHECK();
    }
    Thr thr;
    mind_thresholds(thr.t);
    const int ct = std::min(256, 64 * (int)ceil_div(m, 64));
    const int nbi = (int)ceil_div(m, ct);
    const int per = (int)ceil_div((long long)nbi * n * o, 8);
    k_mind_code<<<8 * per, ct, 0, s>>>(out, w0, m, n, o, qs, thr, nbi, per);
    DEEDS_LAUNCH_CHECK();
}

}  // namespace deeds
