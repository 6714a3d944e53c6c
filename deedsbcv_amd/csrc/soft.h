// Differentiable soft variant of one deeds registration level (no counterpart
// in the reference; the model below is this library's own).
//
//   features  phi[p][l], l = 0..11: the un-binarised self-similarity context
//             of MINDSSCbox.h:127-201 -- six box-filtered squared
//             differences D_c (:104-124, clamped window sums of radius qs),
//             read at the twelve SSC offsets (:184-191), minus their minimum,
//             divided by max(mean, 1e-6) (:192-201) -- before thresholding
//   cost      C[n,l] = a1 * sum_{samples s of n} sum_ch (F[s,ch] - M[clamp(s+d(l)),ch])^2
//             with the samples, clamping and a1 of dataCostD.h:141-196
//   energy    E(L) = sum_n C_n(l_n) + sum_{c != root} |l_p - l_c + off_c|^2 / w_c
//             on a tree in BFS order (the MST of primsMST.h), w_c = edgemst[c],
//             off_c = (u0[p] - u0[c]) / quant per axis (regularisation.h:208-210)
//   inference sum-product BP at temperature T: log-sum-exp distance
//             transforms separable over the three label axes, upward then
//             downward messages, marginals P_c = softmax(-g_c / T), output
//             u1_c = u0_c + E_{P_c}[label displacement]; the sweeps run on
//             the tree re-rooted at its centre (the marginals do not depend
//             on the root; the centre halves the sequential depth)
//   backward  hand-written adjoints of all three stages; BP keeps the upward
//             beliefs, upward messages and full beliefs of every node and
//             recomputes each distance transform in the backward sweeps
//
// Launchers of the three stages (soft_features.hip, soft_cost.hip,
// soft_bp.hip).  All pointers are device pointers; R is float or double.
// Features are laid out [voxel][12], cost volumes node-major [node][label],
// displacement fields [3][N] (u, v, w).
#pragma once

#include <cstddef>
#include <vector>

#include "common.h"

namespace deeds {

// Edge weights below this are raised to it before dividing the pairwise term.
constexpr float kSoftMinEdgeWeight = 1e-6f;

// Rounds of one sweep direction: round r runs segments round_start[r] ..
// round_start[r+1]-1, one workgroup each, in one launch; segment s is the
// node list d_nodes[d_seg_start[s] .. d_seg_start[s+1]), each node the parent
// (upward) or a child (downward) of the one before it, so a workgroup keeps
// the message it just computed on chip for the next node.
struct SoftRounds {
    std::vector<int> round_start;
    const int* d_seg_start = nullptr;
    const int* d_nodes = nullptr;
    int rounds() const { return round_start.empty() ? 0 : (int)round_start.size() - 1; }
};

// Steps per round of the soft sweeps (one launch each).
constexpr int kSoftWindow = 16;

// A tree in BFS order with its sweep schedules on the host and its arrays on the device.
struct SoftTree {
    int N = 0, maxd = 0;
    std::vector<int> depth_start;  // host: maxd + 2 offsets into ordered
    const int* d_ordered = nullptr;
    const int* d_parents = nullptr;
    const int* d_child_start = nullptr;  // N + 1
    const int* d_child_list = nullptr;   // N - 1
    const float* d_edgemst = nullptr;
    SoftRounds up;    // children before parents, every node (the root last)
    SoftRounds down;  // parents before children, every node but the root
};

// Workgroup geometry of the soft BP for L labels per axis: each line of the
// label cube is shared by S threads that produce OUT outputs each, OUT the
// smallest count that keeps L^2 * S threads within 1024.
constexpr int soft_bp_out(int L) {
    int o = 1;
    while (L * L * ((L + o - 1) / o) > 1024) o++;
    return o;
}
constexpr int soft_bp_s(int L) { return (L + soft_bp_out(L) - 1) / soft_bp_out(L); }
// LDS bytes of one soft BP workgroup: six label cubes, three padded quadratic
// tables and their exponential forms, and a reduction area.
constexpr size_t soft_bp_smem(int L, size_t rsize) {
    return (6 * (size_t)L * L * L + 6 * (2 * (size_t)L + 1 + 2 * (size_t)soft_bp_s(L) * soft_bp_out(L)) + 32) *
           rsize;
}
// Largest label radius with a compiled soft BP kernel.
constexpr int kSoftMaxRadius = 8;

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
// part: soft_cost_part_count(...) reals of workspace (sums of the padded positions);
// gT: soft_cost_grad_t_count(...) reals (the cost gradient, label-major).
template <typename R>
void soft_cost_backward(R* dF, R* dM, const R* gcost, const R* F, const R* M, int m, int n, int o, int step, int hw,
                        float quant, float alpha, R* part, R* gT, hipStream_t s);
inline size_t soft_cost_part_count(int m, int n, int o, int hw, float quant) {
    const size_t P = (size_t)quant * hw;
    return 13 * (m + 2 * P) * (n + 2 * P) * (o + 2 * P);
}
inline size_t soft_cost_grad_t_count(int m, int n, int o, int step, int hw) {
    const size_t L = 2 * (size_t)hw + 1, N = (size_t)(m / step) * (n / step) * (o / step);
    return L * L * L * N;
}

// saved: 3*N*L^3 reals (upward beliefs, upward messages, full beliefs).
template <typename R>
void soft_bp_forward(const SoftTree& t, const R* cost, const R* u0, R* u1, R* saved, int hw, float quant, float T,
                     hipStream_t s);
// scratch: 9*N*L^3 reals (A, Dt and the transforms recomputed for the backward steps).
template <typename R>
void soft_bp_backward(const SoftTree& t, const R* cost, const R* u0, const R* saved, const R* gu1, R* gcost,
                      R* scratch, int hw, float quant, float T, hipStream_t s);

}  // namespace deeds
