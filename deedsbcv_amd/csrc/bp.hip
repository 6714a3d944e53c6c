// Tree belief propagation on gfx950 (reference: regularisation.h:6-303).
//
// Bottom-up, one launch per BFS depth d (deepest first): one workgroup per
// parent p at depth d-1 that has children at depth d.  The workgroup walks its
// children in `ordered` order (the reference's serial fold order,
// regularisation.h:221-232), and for each child c:
//   - stages cost[c] * edgemst[c] in LDS                              (:197-202)
//   - runs the three separable min-convolution passes y, x, z with the
//     quadratic (t - len + offset)^2 in label units and the reference's
//     strict-< lowest-index tie rule, tracking the composite argmin   (:6-113)
//   - folds  acc[p] += msg - min(msg)  into an LDS accumulator        (:221-232)
// and finally writes the accumulator back to cost[p].  Each 1-D pass is the
// exact brute-force loop of the reference (same float operations in the same
// order), so messages and argmin tables are bit-identical; with len <= 17 the
// O(len) inner loop is cheap next to the per-depth latency.
// Top-down: a single workgroup selects labels depth by depth (:259-293).
#include "kernels.h"

namespace deeds {

namespace {

constexpr int BP_THREADS = 256;

__device__ __forceinline__ float block_min(float v, float* red) {
    // exact: min is order independent
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        float o = __shfl_xor(v, off, 64);
        v = o < v ? o : v;
    }
    int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) red[wid] = v;
    __syncthreads();
    float r = red[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) r = red[w] < r ? red[w] : r;
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(BP_THREADS) void k_bp_level(float* __restrict__ cost, int16_t* __restrict__ allinds,
                                                         const int* __restrict__ grp_parent,
                                                         const int* __restrict__ grp_child,
                                                         const int* __restrict__ child_list,
                                                         const float* __restrict__ edgemst,
                                                         const float* __restrict__ u0, const float* __restrict__ v0,
                                                         const float* __restrict__ w0, int grp_begin, int len,
                                                         float quant) {
    extern __shared__ float smem[];
    __shared__ float red[BP_THREADS / 64];
    const int len2 = len * len, len3 = len2 * len;
    float* A = smem;
    float* B = A + len3;
    float* C = B + len3;
    const int len3e = (len3 + 1) & ~1;  // int16 regions padded to keep the float tables 4-byte aligned
    int16_t* I1 = (int16_t*)(C + len3);
    int16_t* I2 = I1 + len3e;
    float* zt = (float*)(I2 + len3e);  // three tables of 2*len+1
    float* zy = zt;
    float* zx = zt + (2 * len + 1);
    float* zz = zx + (2 * len + 1);

    const int g = grp_begin + blockIdx.x;
    const int p = grp_parent[g];
    const long long pb = (long long)p * len3;
    for (int l = threadIdx.x; l < len3; l += blockDim.x) C[l] = cost[pb + l];

    for (int ci = grp_child[g]; ci < grp_child[g + 1]; ci++) {
        const int c = child_list[ci];
        const long long cb = (long long)c * len3;
        const float wgt = edgemst[c];
        for (int l = threadIdx.x; l < len3; l += blockDim.x) A[l] = cost[cb + l] * wgt;
        // offsets in label units (regularisation.h:208-210)
        const float ox = (u0[p] - u0[c]) / quant;
        const float oy = (v0[p] - v0[c]) / quant;
        const float oz = (w0[p] - w0[c]) / quant;
        for (int t = threadIdx.x; t < 2 * len + 1; t += blockDim.x) {
            float dy = (float)(t - len) + oy;
            float dx = (float)(t - len) + ox;
            float dz = (float)(t - len) + oz;
            zy[t] = dy * dy;
            zx[t] = dx * dx;
            zz[t] = dz * dz;
        }
        __syncthreads();
        // pass y (contiguous axis): A -> B, I1 = composite index of the source
        for (int out = threadIdx.x; out < len3; out += blockDim.x) {
            const int line = out / len, i = out - line * len;
            const float* src = A + line * len;
            float best = src[0] + zy[i + len];
            int bj = 0;
            for (int j = 0; j < len; j++) {
                float cnd = src[j] + zy[i - j + len];
                if (cnd < best) { best = cnd; bj = j; }
            }
            B[out] = best;
            I1[out] = (int16_t)(bj + line * len);
        }
        __syncthreads();
        // pass x (stride len): B -> A, I2
        for (int out = threadIdx.x; out < len3; out += blockDim.x) {
            const int k1 = out / len2, rem = out - k1 * len2, i = rem / len, i1 = rem - i * len;
            const int base = i1 + k1 * len2;
            float best = B[base] + zx[i + len];
            int bj = 0;
            for (int j = 0; j < len; j++) {
                float cnd = B[base + j * len] + zx[i - j + len];
                if (cnd < best) { best = cnd; bj = j; }
            }
            A[out] = best;
            I2[out] = I1[base + bj * len];
        }
        __syncthreads();
        // pass z (stride len^2): A -> B, argmin table to global
        for (int out = threadIdx.x; out < len3; out += blockDim.x) {
            const int i = out / len2, c2 = out - i * len2;
            float best = A[c2] + zz[i + len];
            int bj = 0;
            for (int j = 0; j < len; j++) {
                float cnd = A[c2 + j * len2] + zz[i - j + len];
                if (cnd < best) { best = cnd; bj = j; }
            }
            B[out] = best;
            allinds[cb + out] = I2[c2 + bj * len2];
        }
        __syncthreads();
        float mn = 3.402823466e38f;
        for (int l = threadIdx.x; l < len3; l += blockDim.x) mn = B[l] < mn ? B[l] : mn;
        mn = block_min(mn, red);
        for (int l = threadIdx.x; l < len3; l += blockDim.x) C[l] += (B[l] - mn);
        __syncthreads();
    }
    for (int l = threadIdx.x; l < len3; l += blockDim.x) cost[pb + l] = C[l];
}

// Root argmin (first strict minimum), then label selection depth by depth
// and the displacement update u1 = label displacement + u0.
__global__ __launch_bounds__(1024) void k_bp_select(const float* __restrict__ cost,
                                                    const int16_t* __restrict__ allinds,
                                                    const int* __restrict__ ordered, const int* __restrict__ parents,
                                                    const int* __restrict__ depth_start, int N, int maxd,
                                                    const float* __restrict__ u0, const float* __restrict__ v0,
                                                    const float* __restrict__ w0, float* __restrict__ u1,
                                                    float* __restrict__ v1, float* __restrict__ w1,
                                                    int* __restrict__ sel, int len, int hw, float quant) {
    __shared__ float rv[16];
    __shared__ int ri[16];
    const int len3 = len * len * len;
    const int root = ordered[0];
    const float* rc = cost + (long long)root * len3;
    float bv = 3.402823466e38f;
    int bi = 0x7fffffff;
    for (int l = threadIdx.x; l < len3; l += blockDim.x) {
        float v = rc[l];
        if (v < bv || (v == bv && l < bi)) { bv = v; bi = l; }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        float ov = __shfl_xor(bv, off, 64);
        int oi = __shfl_xor(bi, off, 64);
        if (ov < bv || (ov == bv && oi < bi)) { bv = ov; bi = oi; }
    }
    int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) { rv[wid] = bv; ri[wid] = bi; }
    __syncthreads();
    if (threadIdx.x == 0) {
        float v = rv[0];
        int id = ri[0];
        for (int w = 1; w < (int)(blockDim.x >> 6); w++)
            if (rv[w] < v || (rv[w] == v && ri[w] < id)) { v = rv[w]; id = ri[w]; }
        sel[root] = id;
    }
    __syncthreads();
    for (int d = 1; d <= maxd; d++) {
        for (int t = depth_start[d] + threadIdx.x; t < depth_start[d + 1]; t += blockDim.x) {
            int c = ordered[t];
            sel[c] = allinds[(long long)c * len3 + sel[parents[c]]];
        }
        __syncthreads();
    }
    for (int t = threadIdx.x; t < N; t += blockDim.x) {
        int lab = sel[t];
        int li = lab % len, lj = (lab / len) % len, lk = lab / (len * len);
        u1[t] = (float)(lj - hw) * quant + u0[t];
        v1[t] = (float)(li - hw) * quant + v0[t];
        w1[t] = (float)(lk - hw) * quant + w0[t];
    }
}

}  // namespace

size_t bp_level_smem(int len) {
    size_t len3 = (size_t)len * len * len;
    return 3 * len3 * sizeof(float) + 2 * ((len3 + 1) & ~(size_t)1) * sizeof(int16_t) +
           3 * (2 * len + 1) * sizeof(float);
}

void bp_set_smem_limit() {
    int len = 2 * kMaxBPRadius + 1;
    DEEDS_HIP(hipFuncSetAttribute((const void*)k_bp_level, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)bp_level_smem(len)));
}

void launch_bp_level(float* cost, int16_t* allinds, const BPDevTree& t, int grp_begin, int grp_count,
                     const float* u0, const float* v0, const float* w0, int len, float quant, hipStream_t s) {
    if (grp_count <= 0) return;
    size_t smem = bp_level_smem(len);
    k_bp_level<<<grp_count, BP_THREADS, smem, s>>>(cost, allinds, t.grp_parent, t.grp_child, t.child_list,
                                                   t.edgemst, u0, v0, w0, grp_begin, len, quant);
    DEEDS_LAUNCH_CHECK();
}

void launch_bp_select(const float* cost, const int16_t* allinds, const BPDevTree& t, int N, int maxd,
                      const float* u0, const float* v0, const float* w0, float* u1, float* v1, float* w1, int* sel,
                      int hw, float quant, hipStream_t s) {
    int len = 2 * hw + 1;
    k_bp_select<<<1, 1024, 0, s>>>(cost, allinds, t.ordered, t.parents, t.depth_start, N, maxd, u0, v0, w0, u1, v1,
                                   w1, sel, len, hw, quant);
    DEEDS_LAUNCH_CHECK();
}

}  // namespace deeds
