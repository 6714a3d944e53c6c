// Launchers of the differentiable soft variant (soft.hip).  All pointers are
// device pointers; R is float or double.  Features are laid out [voxel][12],
// cost volumes node-major [node][label], displacement fields [3][N] (u, v, w).
#pragma once

#include <cstddef>
#include <vector>

#include "common.h"

namespace deeds {

// Edge weights below this are raised to it before dividing the pairwise term.
constexpr float kSoftMinEdgeWeight = 1e-6f;

// A tree in BFS order with its depth schedule on the host and its arrays on the device.
struct SoftTree {
    int N = 0, maxd = 0;
    std::vector<int> depth_start;  // host: maxd + 2 offsets into ordered
    const int* d_ordered = nullptr;
    const int* d_parents = nullptr;
    const int* d_child_start = nullptr;  // N + 1
    const int* d_child_list = nullptr;   // N - 1
    const float* d_edgemst = nullptr;
};

// LDS bytes of one soft BP workgroup for L labels per axis.
size_t soft_bp_smem(int L, size_t rsize);

// phi from image im; w0, w1: 6*m*n*o reals of workspace each.
template <typename R>
void soft_features(R* phi, const R* im, int m, int n, int o, int qs, R* w0, R* w1, hipStream_t s);
// gim = d/dim given gphi; w0, w1: 6*m*n*o reals, w2: 12*m*n*o reals.
template <typename R>
void soft_features_backward(R* gim, const R* gphi, const R* im, int m, int n, int o, int qs, R* w0, R* w1, R* w2,
                            hipStream_t s);

template <typename R>
void soft_cost(R* cost, const R* F, const R* M, int m, int n, int o, int step, int hw, float quant, float alpha,
               hipStream_t s);
template <typename R>
void soft_cost_backward(R* dF, R* dM, const R* gcost, const R* F, const R* M, int m, int n, int o, int step, int hw,
                        float quant, float alpha, hipStream_t s);

// saved: 3*N*L^3 reals (upward beliefs, upward messages, full beliefs).
template <typename R>
void soft_bp_forward(const SoftTree& t, const R* cost, const R* u0, R* u1, R* saved, int hw, float quant, float T,
                     hipStream_t s);
// scratch: 2*N*L^3 reals.
template <typename R>
void soft_bp_backward(const SoftTree& t, const R* cost, const R* u0, const R* saved, const R* gu1, R* gcost,
                      R* scratch, int hw, float quant, float T, hipStream_t s);

}  // namespace deeds
