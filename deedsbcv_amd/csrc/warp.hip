// Field transforms and image warps on gfx950 (reference: transformations.h,
// dataCostD.h:213-315).
//
// All interpolation reproduces the reference's arithmetic bit for bit: float
// coordinates, floor, clamped neighbour indices and the mixed float/double
// weight expression of transformations.h:13-20 (the TU is compiled with
// -ffp-contract=off so no product/sum pair is fused).  The "_grid" warps fuse
// upsampleDeformationsCL (transformations.h:245-279) into the warp: the full
// resolution displacement of each voxel is computed in registers with the same
// operations and never written to HBM.
#include "kernels.h"

namespace deeds {

namespace {

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// A 1-D launch over voxels in blocks of 256, in 8 * per workgroups: the GPU's
// eight XCDs take workgroups round robin, so XCD x gets the contiguous voxel
// blocks [x * per, (x + 1) * per) -- a slab of rows whose neighbouring rows,
// which the trilinear gathers also read, sit in the same L2.
struct Span {
    int per, blocks;
    __device__ __forceinline__ long long voxel() const {
        const int b = ((int)blockIdx.x & 7) * per + ((int)blockIdx.x >> 3);
        return b < blocks ? (long long)b * blockDim.x + threadIdx.x : -1;
    }
    unsigned grid() const { return 8u * (unsigned)per; }
};

// transformations.h:5-26 / dataCostD.h:280-295 for one sample point.
__device__ __forceinline__ float trilinear(const float* __restrict__ in, float x1, float y1, float z1, int addj,
                                           int addi, int addk, int m2, int n2, int o2) {
    int x = (int)floorf(x1), y = (int)floorf(y1), z = (int)floorf(z1);
    float dx = x1 - (float)x, dy = y1 - (float)y, dz = z1 - (float)z;
    x += addj; y += addi; z += addk;
    // 32-bit offsets: every volume holds fewer than 2^31 voxels (checked by the callers)
    const int s1 = m2, s2 = m2 * n2;
    const int y0 = clampi(y, 0, m2 - 1), y1c = clampi(y + 1, 0, m2 - 1);
    const int x0 = clampi(x, 0, n2 - 1) * s1, x1c = clampi(x + 1, 0, n2 - 1) * s1;
    const int z0 = clampi(z, 0, o2 - 1) * s2, z1c = clampi(z + 1, 0, o2 - 1) * s2;
    float i000 = in[y0 + x0 + z0], i100 = in[y1c + x0 + z0], i010 = in[y0 + x1c + z0], i001 = in[y0 + x0 + z1c];
    float i110 = in[y1c + x1c + z0], i101 = in[y1c + x0 + z1c], i011 = in[y0 + x1c + z1c];
    float i111 = in[y1c + x1c + z1c];
    double ax = 1.0 - (double)dx, ay = 1.0 - (double)dy, az = 1.0 - (double)dz;
    double s = ax * ay * az * (double)i000;
    s = s + ax * (double)dy * az * (double)i100;
    s = s + (double)dx * ay * az * (double)i010;
    s = s + ax * ay * (double)dz * (double)i001;
    s = s + (double)(dx * dy) * az * (double)i110;
    s = s + ax * (double)dy * (double)dz * (double)i101;
    s = s + (double)dx * ay * (double)dz * (double)i011;
    s = s + (double)(dx * dy * dz * i111);
    return (float)s;
}

// (i, j, k) of voxel p in 32-bit unsigned arithmetic (fewer than 2^31 voxels)
__device__ __forceinline__ void voxel_ijk(long long p, int m, int n, int& i, int& j, int& k) {
    const unsigned q = (unsigned)p, um = (unsigned)m, un = (unsigned)n;
    const unsigned r = q / um;
    i = (int)(q - r * um);
    const unsigned kk = r / un;
    j = (int)(r - kk * un);
    k = (int)kk;
}

// float affine coordinate, summed left to right (dataCostD.h:285-287)
__device__ __forceinline__ float affine_coord(const float* r, int i, int j, int k, float disp) {
    return (float)i * r[0] + (float)j * r[1] + (float)k * r[2] + r[3] + disp;
}

struct UpScale {
    float sm, sn, so;
};

// the full-resolution displacement of voxel (i,j,k) upsampled from a coarse grid
__device__ __forceinline__ void upsampled(const float* u0, const float* v0, const float* w0, int i, int j, int k,
                                          UpScale sc, int m2, int n2, int o2, float& u, float& v, float& w) {
    float x = (float)j / sc.sn, y = (float)i / sc.sm, z = (float)k / sc.so;
    u = trilinear(u0, x, y, z, 0, 0, 0, m2, n2, o2);
    v = trilinear(v0, x, y, z, 0, 0, 0, m2, n2, o2);
    w = trilinear(w0, x, y, z, 0, 0, 0, m2, n2, o2);
}

__global__ void k_upsample(float* __restrict__ u1, float* __restrict__ v1, float* __restrict__ w1,
                           const float* __restrict__ u0, const float* __restrict__ v0, const float* __restrict__ w0,
                           int m, int n, int o, int m2, int n2, int o2, UpScale sc, Span sp) {
    long long sz = (long long)m * n * o;
    const long long p = sp.voxel();
    if (p < 0 || p >= sz) return;
    int i, j, k;
    voxel_ijk(p, m, n, i, j, k);
    float u, v, w;
    upsampled(u0, v0, w0, i, j, k, sc, m2, n2, o2, u, v, w);
    u1[p] = u;
    v1[p] = v;
    w1[p] = w;
}

__global__ void k_interp3(float* __restrict__ out, const float* __restrict__ in, const float* __restrict__ x1,
                          const float* __restrict__ y1, const float* __restrict__ z1, int m, int n, int o, int m2,
                          int n2, int o2, int add_identity, Span sp) {
    long long sz = (long long)m * n * o;
    const long long p = sp.voxel();
    if (p < 0 || p >= sz) return;
    int i, j, k;
    voxel_ijk(p, m, n, i, j, k);
    out[p] = add_identity ? trilinear(in, x1[p], y1[p], z1[p], j, i, k, m2, n2, o2)
                          : trilinear(in, x1[p], y1[p], z1[p], 0, 0, 0, m2, n2, o2);
}

__global__ void k_warp_affine(float* __restrict__ out, const float* __restrict__ in, Affine A,
                              const float* __restrict__ u, const float* __restrict__ v, const float* __restrict__ w,
                              int m, int n, int o, Span sp) {
    long long sz = (long long)m * n * o;
    const long long p = sp.voxel();
    if (p < 0 || p >= sz) return;
    int i, j, k;
    voxel_ijk(p, m, n, i, j, k);
    float y1 = affine_coord(A.r, i, j, k, v[p]);
    float x1 = affine_coord(A.r + 4, i, j, k, u[p]);
    float z1 = affine_coord(A.r + 8, i, j, k, w[p]);
    out[p] = trilinear(in, x1, y1, z1, 0, 0, 0, m, n, o);
}

__global__ void k_warp_affine_grid(float* __restrict__ out, const float* __restrict__ in, Affine A,
                                   const float* __restrict__ u0, const float* __restrict__ v0,
                                   const float* __restrict__ w0, int m, int n, int o, int m2, int n2, int o2,
                                   UpScale sc, Span sp) {
    long long sz = (long long)m * n * o;
    const long long p = sp.voxel();
    if (p < 0 || p >= sz) return;
    int i, j, k;
    voxel_ijk(p, m, n, i, j, k);
    float u, v, w;
    upsampled(u0, v0, w0, i, j, k, sc, m2, n2, o2, u, v, w);
    float y1 = affine_coord(A.r, i, j, k, v);
    float x1 = affine_coord(A.r + 4, i, j, k, u);
    float z1 = affine_coord(A.r + 8, i, j, k, w);
    out[p] = trilinear(in, x1, y1, z1, 0, 0, 0, m, n, o);
}

__global__ void k_warp_image_grid(float* __restrict__ out, const float* __restrict__ in,
                                  const float* __restrict__ u0, const float* __restrict__ v0,
                                  const float* __restrict__ w0, int m, int n, int o, int m2, int n2, int o2,
                                  UpScale sc, Span sp) {
    long long sz = (long long)m * n * o;
    const long long p = sp.voxel();
    if (p < 0 || p >= sz) return;
    int i, j, k;
    voxel_ijk(p, m, n, i, j, k);
    float u, v, w;
    upsampled(u0, v0, w0, i, j, k, sc, m2, n2, o2, u, v, w);
    out[p] = trilinear(in, u, v, w, j, i, k, m, n, o);
}

__global__ void k_warp_affine_s(int16_t* __restrict__ out, const int16_t* __restrict__ in, Affine A,
                                const float* __restrict__ u, const float* __restrict__ v,
                                const float* __restrict__ w, int m, int n, int o, Span sp) {
    long long sz = (long long)m * n * o;
    const long long p = sp.voxel();
    if (p < 0 || p >= sz) return;
    int i, j, k;
    voxel_ijk(p, m, n, i, j, k);
    int y = (int)roundf(affine_coord(A.r, i, j, k, v[p]));
    int x = (int)roundf(affine_coord(A.r + 4, i, j, k, u[p]));
    int z = (int)roundf(affine_coord(A.r + 8, i, j, k, w[p]));
    out[p] = in[clampi(y, 0, m - 1) + (long long)clampi(x, 0, n - 1) * m + (long long)clampi(z, 0, o - 1) * m * n];
}

// consistentMappingCL helpers (transformations.h:189-242)
__global__ void k_scale6(float* __restrict__ dst, const float* __restrict__ a, const float* __restrict__ b,
                         const float* __restrict__ c, const float* __restrict__ d, const float* __restrict__ e,
                         const float* __restrict__ f, long long sz, float s) {
    long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= sz) return;
    dst[p] = a[p] * s;
    dst[sz + p] = b[p] * s;
    dst[2 * sz + p] = c[p] * s;
    dst[3 * sz + p] = d[p] * s;
    dst[4 * sz + p] = e[p] * s;
    dst[5 * sz + p] = f[p] * s;
}

struct Six {
    float* f[6];
};

// One Jacobi iteration from the previous iterate `s` into `d`.
__global__ void k_consistent_iter(Six d, Six s, int m, int n, int o, Span sp) {
    long long sz = (long long)m * n * o;
    const long long p = sp.voxel();
    if (p < 0 || p >= sz) return;
    int i, j, k;
    voxel_ijk(p, m, n, i, j, k);
    const float *us = s.f[0], *vs = s.f[1], *ws = s.f[2], *us2 = s.f[3], *vs2 = s.f[4], *ws2 = s.f[5];
    float a = us[p], b = vs[p], c = ws[p];
    float iu = trilinear(us2, a, b, c, j, i, k, m, n, o);
    float iv = trilinear(vs2, a, b, c, j, i, k, m, n, o);
    float iw = trilinear(ws2, a, b, c, j, i, k, m, n, o);
    d.f[0][p] = (float)(0.5 * (double)a - 0.5 * (double)iu);
    d.f[1][p] = (float)(0.5 * (double)b - 0.5 * (double)iv);
    d.f[2][p] = (float)(0.5 * (double)c - 0.5 * (double)iw);
    float a2 = us2[p], b2 = vs2[p], c2 = ws2[p];
    float ju = trilinear(us, a2, b2, c2, j, i, k, m, n, o);
    float jv = trilinear(vs, a2, b2, c2, j, i, k, m, n, o);
    float jw = trilinear(ws, a2, b2, c2, j, i, k, m, n, o);
    d.f[3][p] = (float)(0.5 * (double)a2 - 0.5 * (double)ju);
    d.f[4][p] = (float)(0.5 * (double)b2 - 0.5 * (double)jv);
    d.f[5][p] = (float)(0.5 * (double)c2 - 0.5 * (double)jw);
}

__global__ void k_unscale6(Six d, const float* __restrict__ src, long long sz, float s) {
    long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= sz) return;
#pragma unroll
    for (int c = 0; c < 6; c++) d.f[c][p] = src[c * sz + p] * s;
}

__device__ __forceinline__ double block_sum(double v, double* red) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) red[wid] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) r += red[w];
    __syncthreads();
    return r;
}

__global__ void k_ssd(double* __restrict__ part, const float* __restrict__ ref, const float* __restrict__ warped,
                      const float* __restrict__ before, long long sz) {
    __shared__ double red[4];
    double s0 = 0.0, s1 = 0.0;
    for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < sz; p += (long long)gridDim.x * blockDim.x) {
        double a = (double)(ref[p] - before[p]);
        double b = (double)(ref[p] - warped[p]);
        s0 += a * a;
        s1 += b * b;
    }
    s0 = block_sum(s0, red);
    s1 = block_sum(s1, red);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = s0;
        part[2 * blockIdx.x + 1] = s1;
    }
}

// acc[c] = sum over the `blocks` partials part[b * stride + c], c < stride:
// each thread adds its blocks in order, then a fixed tree over the threads.
__global__ __launch_bounds__(256) void k_sum_parts(double* __restrict__ acc, const double* __restrict__ part,
                                                   int blocks, int stride) {
    __shared__ double red[256];
    for (int c = 0; c < stride; c++) {
        double v = 0.0;
        for (int b = threadIdx.x; b < blocks; b += blockDim.x) v += part[(long long)b * stride + c];
        red[threadIdx.x] = v;
        __syncthreads();
        for (int h = 128; h > 0; h >>= 1) {
            if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
            __syncthreads();
        }
        if (threadIdx.x == 0) acc[c] = red[0];
        __syncthreads();
    }
}

// det J of x + u/factor with central differences (transformations.h:84-185),
// accumulated in double: acc[0] += sum det, acc[1] += sum det^2, acc[2] += #neg.
__device__ __forceinline__ float cdiff(const float* f, int i, int j, int k, int m, int n, int o, int dim) {
    const float g0 = -0.5f, g2 = 0.5f;
    long long mn = (long long)m * n;
    float a, b, c;
    if (dim == 1) {
        a = f[clampi(i - 1, 0, m - 1) + (long long)j * m + k * mn];
        b = f[i + (long long)j * m + k * mn];
        c = f[clampi(i + 1, 0, m - 1) + (long long)j * m + k * mn];
    } else if (dim == 2) {
        a = f[i + (long long)clampi(j - 1, 0, n - 1) * m + k * mn];
        b = f[i + (long long)j * m + k * mn];
        c = f[i + (long long)clampi(j + 1, 0, n - 1) * m + k * mn];
    } else {
        a = f[i + (long long)j * m + clampi(k - 1, 0, o - 1) * mn];
        b = f[i + (long long)j * m + k * mn];
        c = f[i + (long long)j * m + clampi(k + 1, 0, o - 1) * mn];
    }
    float acc = 0.0f;
    acc = acc + g0 * a;
    acc = acc + 0.0f * b;
    acc = acc + g2 * c;
    return acc;
}

__global__ void k_jacobian(double* __restrict__ part, const float* __restrict__ u, const float* __restrict__ v,
                           const float* __restrict__ w, int m, int n, int o, float f1) {
    __shared__ double red[4];
    long long sz = (long long)m * n * o;
    double s = 0.0, s2 = 0.0, neg = 0.0;
    for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < sz; p += (long long)gridDim.x * blockDim.x) {
        int i, j, k;
        voxel_ijk(p, m, n, i, j, k);
        float J11 = cdiff(u, i, j, k, m, n, o, 2) * f1, J12 = cdiff(u, i, j, k, m, n, o, 1) * f1;
        float J13 = cdiff(u, i, j, k, m, n, o, 3) * f1;
        float J21 = cdiff(v, i, j, k, m, n, o, 2) * f1, J22 = cdiff(v, i, j, k, m, n, o, 1) * f1;
        float J23 = cdiff(v, i, j, k, m, n, o, 3) * f1;
        float J31 = cdiff(w, i, j, k, m, n, o, 2) * f1, J32 = cdiff(w, i, j, k, m, n, o, 1) * f1;
        float J33 = cdiff(w, i, j, k, m, n, o, 3) * f1;
        J11 = (float)((double)J11 + 1.0);
        J22 = (float)((double)J22 + 1.0);
        J33 = (float)((double)J33 + 1.0);
        float d = J11 * (J22 * J33 - J23 * J32) - J21 * (J12 * J33 - J13 * J32) + J31 * (J12 * J23 - J13 * J22);
        s += d;
        s2 += (double)d * d;
        neg += d < 0 ? 1.0 : 0.0;
    }
    s = block_sum(s, red);
    s2 = block_sum(s2, red);
    neg = block_sum(neg, red);
    if (threadIdx.x == 0) {
        part[3 * blockIdx.x] = s;
        part[3 * blockIdx.x + 1] = s2;
        part[3 * blockIdx.x + 2] = neg;
    }
}

__global__ void k_zero_d(double* a, int n) {
    if ((int)threadIdx.x < n) a[threadIdx.x] = 0.0;
}

UpScale make_scale(int m, int n, int o, int m2, int n2, int o2) {
    return UpScale{(float)m / (float)m2, (float)n / (float)n2, (float)o / (float)o2};
}

constexpr int TB = 256;

Span span_of(long long sz) {
    const int blocks = (int)ceil_div(sz, TB);
    return Span{(int)ceil_div(blocks, 8), blocks};
}

}  // namespace

Affine affine_from_X(const float* X) {
    // the warps index X flat: y uses X[0..3], x X[4..7], z X[8..11]
    // (dataCostD.h:285-287); X holds the text file's rows column-wise
    Affine a{};
    for (int c = 0; c < 12; c++) a.r[c] = X[c];
    return a;
}

void launch_upsample(float* u1, float* v1, float* w1, const float* u0, const float* v0, const float* w0, int m, int n,
                     int o, int m2, int n2, int o2, hipStream_t s) {
    long long sz = (long long)m * n * o;
    if (sz == 0) return;
    k_upsample<<<span_of(sz).grid(), TB, 0, s>>>(u1, v1, w1, u0, v0, w0, m, n, o, m2, n2, o2,
                                               make_scale(m, n, o, m2, n2, o2), span_of(sz));
    DEEDS_LAUNCH_CHECK();
}

void launch_interp3(float* out, const float* in, const float* x1, const float* y1, const float* z1, int m, int n,
                    int o, int m2, int n2, int o2, int add_identity, hipStream_t s) {
    long long sz = (long long)m * n * o;
    if (sz == 0) return;
    k_interp3<<<span_of(sz).grid(), TB, 0, s>>>(out, in, x1, y1, z1, m, n, o, m2, n2, o2, add_identity, span_of(sz));
    DEEDS_LAUNCH_CHECK();
}

void launch_warp_affine(float* out, const float* in, const Affine& A, const float* u, const float* v, const float* w,
                        int m, int n, int o, hipStream_t s) {
    long long sz = (long long)m * n * o;
    if (sz == 0) return;
    k_warp_affine<<<span_of(sz).grid(), TB, 0, s>>>(out, in, A, u, v, w, m, n, o, span_of(sz));
    DEEDS_LAUNCH_CHECK();
}

void launch_warp_affine_grid(float* out, const float* in, const Affine& A, const float* u0, const float* v0,
                             const float* w0, int m, int n, int o, int m2, int n2, int o2, hipStream_t s) {
    long long sz = (long long)m * n * o;
    if (sz == 0) return;
    k_warp_affine_grid<<<span_of(sz).grid(), TB, 0, s>>>(out, in, A, u0, v0, w0, m, n, o, m2, n2, o2,
                                                       make_scale(m, n, o, m2, n2, o2), span_of(sz));
    DEEDS_LAUNCH_CHECK();
}

void launch_warp_image_grid(float* out, const float* in, const float* u0, const float* v0, const float* w0, int m,
                            int n, int o, int m2, int n2, int o2, hipStream_t s) {
    long long sz = (long long)m * n * o;
    if (sz == 0) return;
    k_warp_image_grid<<<span_of(sz).grid(), TB, 0, s>>>(out, in, u0, v0, w0, m, n, o, m2, n2, o2,
                                                      make_scale(m, n, o, m2, n2, o2), span_of(sz));
    DEEDS_LAUNCH_CHECK();
}

void launch_warp_affine_s(int16_t* out, const int16_t* in, const Affine& A, const float* u, const float* v,
                          const float* w, int m, int n, int o, hipStream_t s) {
    long long sz = (long long)m * n * o;
    if (sz == 0) return;
    k_warp_affine_s<<<span_of(sz).grid(), TB, 0, s>>>(out, in, A, u, v, w, m, n, o, span_of(sz));
    DEEDS_LAUNCH_CHECK();
}

void launch_consistent(float* f[6], float* tmp, int m, int n, int o, int factor, hipStream_t s) {
    long long sz = (long long)m * n * o;
    if (sz == 0) return;
    float f1 = (float)(1.0 / (float)factor);
    unsigned g = ceil_div(sz, TB);
    const Span sp = span_of(sz);
    k_scale6<<<g, TB, 0, s>>>(tmp, f[0], f[1], f[2], f[3], f[4], f[5], sz, f1);
    DEEDS_LAUNCH_CHECK();
    Six a, b;
    for (int c = 0; c < 6; c++) {
        a.f[c] = tmp + c * sz;
        b.f[c] = f[c];
    }
    // ten iterations ping-pong tmp -> f -> tmp ...; the tenth lands in tmp
    for (int it = 0; it < 10; it++) {
        if (it % 2 == 0) k_consistent_iter<<<sp.grid(), TB, 0, s>>>(b, a, m, n, o, sp);
        else k_consistent_iter<<<sp.grid(), TB, 0, s>>>(a, b, m, n, o, sp);
        DEEDS_LAUNCH_CHECK();
    }
    k_unscale6<<<g, TB, 0, s>>>(b, tmp, sz, (float)factor);
    DEEDS_LAUNCH_CHECK();
}

void launch_ssd(double* acc, double* part, const float* ref, const float* warped, const float* before, long long sz,
                hipStream_t s) {
    const unsigned g = (unsigned)std::min<long long>(ceil_div(sz, TB), kRedBlocks);
    if (g == 0) {
        k_zero_d<<<1, 64, 0, s>>>(acc, 2);
        DEEDS_LAUNCH_CHECK();
        return;
    }
    k_ssd<<<g, TB, 0, s>>>(part, ref, warped, before, sz);
    DEEDS_LAUNCH_CHECK();
    k_sum_parts<<<1, 256, 0, s>>>(acc, part, (int)g, 2);
    DEEDS_LAUNCH_CHECK();
}

void launch_jacobian(double* acc, double* part, const float* u, const float* v, const float* w, int m, int n, int o,
                     int factor, hipStream_t s) {
    long long sz = (long long)m * n * o;
    if (sz == 0) {
        k_zero_d<<<1, 64, 0, s>>>(acc, 3);
        DEEDS_LAUNCH_CHECK();
        return;
    }
    float f1 = (float)(1.0 / (float)factor);
    const unsigned g = (unsigned)std::min<long long>(ceil_div(sz, TB), kRedBlocks);
    k_jacobian<<<g, TB, 0, s>>>(part, u, v, w, m, n, o, f1);
    DEEDS_LAUNCH_CHECK();
    k_sum_parts<<<1, 256, 0, s>>>(acc, part, (int)g, 3);
    DEEDS_LAUNCH_CHECK();
}

}  // namespace deeds
