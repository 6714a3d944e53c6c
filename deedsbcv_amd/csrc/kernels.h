// Host-side launchers of the deeds_hip kernels (all take device pointers and
// enqueue on the given stream; shapes are validated by the callers).
#pragma once

#include "common.h"

namespace deeds {

// LDS available to one workgroup on gfx950 (160 KiB per CU).
constexpr int kMaxLds = 160 * 1024;

// ---- MIND-SSC (mind.hip) --------------------------------------------------
// The five quantisation thresholds -log((c+1.5)/6) in float.
void mind_thresholds(float thr[5]);
// Workspace: two buffers of 6*m*n*o floats.
void launch_descriptor(uint64_t* out, const float* im, int m, int n, int o, int qs,
                       float* w0, float* w1, hipStream_t s);
// Largest dimension the LDS-tiled box filter accepts.
int mind_max_dim();

// ---- Hamming data cost (datacost.hip) -------------------------------------
struct CostPlan {
    int skip_i, skip_j, skip_k;  // sample strides along i, j, k
    float scale;                 // float factor applied to the integer sum
    int pad;                     // quant*hw padding of the moving codes
    int mp, np, op;              // padded dims
    long long samples;           // block samples per label
};
CostPlan datacost_plan(int m, int n, int o, int step, int hw, float quant, float alpha, int randnum);
void launch_pad_codes(uint64_t* Mp, const uint64_t* M, int m, int n, int o, int pad, hipStream_t s);
void launch_datacost(float* cost, const uint64_t* F, const uint64_t* Mp, int m, int n, int o,
                     int step, int hw, int q, const CostPlan& plan, hipStream_t s);
// The row-kernel tile shapes the autotuner chose so far, as a JSON array.
std::string datacost_tuning_report();
// Pins the row kernel's label planes per workgroup (0: autotune) and pipelining.
void pin_datacost_tiles(int planes, int pipe);

// ---- MST edge costs (mst.hip) ----------------------------------------------
// ec[node + nb*N] = float SAD of the two step^3 blocks (0 off-grid).
void launch_edge_sad(float* ec, const float* im, int step, int m, int n, int o, hipStream_t s);

// ---- belief propagation (bp.hip) ------------------------------------------
struct BPDevTree {
    const int* ordered;      // N
    const int* parents;      // N
    const int* depth_start;  // maxd+2
    const int* prog_start;   // segments+1 (segments in round order)
    int* prog;               // per segment a program block (mst_host.h); the offsets are written per pass
    int n_segments;
    const int* leaves;       // n_leaves leaf nodes (round 0)
    const int* leaf_w;       // their edgemst bits
    int n_leaves;
    const int* sel_info;     // 2*N: per position of ordered, (node, parent's index within its depth)
    const int* sel_jump;     // N: per position, its kSelJump-th ancestor's index within its depth (anchors) or -(depth mod kSelJump)
    int max_width;           // largest number of nodes at one depth
    float* off3;             // 3*N floats of scratch (parent offsets)
    int* sel_labs;           // 2*N ints of scratch: the selection's label rows when a depth is too wide for LDS
    int root;
};
// Largest label radius the LDS-resident BP kernel supports (3 len^3 floats + 2 len^3 int16 in LDS).
constexpr int kMaxBPRadius = 10;
// Bottom-up pass: one launch per round below tail_round (round r covers
// segments round_start[r] .. round_start[r+1], host array), then the rounds
// from tail_round on (one segment each) in one single-workgroup launch.  off3: 3*N
// floats of scratch for the parent offsets.
void launch_bp_rounds(float* cost, int16_t* allinds, float* off3, const BPDevTree& t, int N,
                      const int* round_start, int rounds, int tail_round, const float* u0, const float* v0,
                      const float* w0, int hw, float quant, hipStream_t s);
// Top-down selection.  With `comp` (N*len^3 int16 of scratch) and a tree
// whose depths fit LDS it walks only the anchor depths (multiples of
// kSelJump) on composed tables (first_anchor: the first position at depth
// kSelJump, N when there is none) and fills the depths between in one
// parallel pass, else it walks every depth.
void launch_bp_select(const float* cost, const int16_t* allinds, int16_t* comp, int first_anchor, const BPDevTree& t,
                      int N, int maxd, const float* u0, const float* v0, const float* w0, float* u1, float* v1,
                      float* w1, int hw, float quant, hipStream_t s);
// regularisation.h:273-274: the root's argmin row is the identity (allinds[root][l] = l).
void launch_root_identity(int16_t* allinds, int root, int len3, hipStream_t s);

// One tree of a batched BP (device pointers; the schedule as in BPDevTree).
struct BPBatchTree {
    float* cost;
    int16_t* allinds;
    const int* prog_start;
    int* prog;
    const int* leaves;
    const int* leaf_w;
    float* off3;
    const int* parents;
    const int* ordered;
    const int* depth_start;
    const int* sel_info;
    const float *u0, *v0, *w0;
    float *u1, *v1, *w1;
    int* sel_labs;  // 2*N ints of scratch (BPDevTree::sel_labs)
    int root, N, maxd, n_segments;
};
// BP of several trees with the same label radius in one launch sequence:
// the parent offsets of all trees, then one launch per round holding every
// tree's items of that round -- work[round_off[0] .. round_off[1]) are the
// leaves (tree, leaf position), each later range a round's (tree, offset of
// the segment's program in prog), the last range the trees' spines as item
// pairs ((tree, first segment), (end segment, 0)) run by one workgroup each -- then one
// selection workgroup per tree.
// d_trees: device array of `trees` descriptors; max_n: largest N; width:
// largest depth width.
void launch_bp_batch(const BPBatchTree* d_trees, int trees, int max_n, const int2* work,
                     const std::vector<int>& round_off, int width, int hw, float quant, hipStream_t s);

// ---- transforms (warp.hip) -------------------------------------------------
struct Affine {
    float r[12];  // coefficients of y = r[0..3], x = r[4..7], z = r[8..11] (dataCostD.h:285-287)
};
Affine affine_from_X(const float* X);  // X: the 16 floats as stored by deedsBCV0.cpp:161-165
void launch_upsample(float* u1, float* v1, float* w1, const float* u0, const float* v0, const float* w0,
                     int m, int n, int o, int m2, int n2, int o2, hipStream_t s);
void launch_interp3(float* out, const float* in, const float* x1, const float* y1, const float* z1,
                    int m, int n, int o, int m2, int n2, int o2, int add_identity, hipStream_t s);
void launch_warp_affine(float* out, const float* in, const Affine& A, const float* u, const float* v,
                        const float* w, int m, int n, int o, hipStream_t s);
// Fused: trilinear upsample of the (m2,n2,o2) field to full resolution on the
// fly, then the affine warp (identical per-voxel arithmetic to the two-step path).
void launch_warp_affine_grid(float* out, const float* in, const Affine& A, const float* u0, const float* v0,
                             const float* w0, int m, int n, int o, int m2, int n2, int o2, hipStream_t s);
void launch_warp_image_grid(float* out, const float* in, const float* u0, const float* v0, const float* w0,
                            int m, int n, int o, int m2, int n2, int o2, hipStream_t s);
void launch_warp_affine_s(int16_t* out, const int16_t* in, const Affine& A, const float* u, const float* v,
                          const float* w, int m, int n, int o, hipStream_t s);
// In place on the six fields; tmp holds 6*m*n*o floats.
void launch_consistent(float* f[6], float* tmp, int m, int n, int o, int factor, hipStream_t s);
// Partial sums of the logging reductions below: per workgroup, then added
// in a fixed order, so the figures do not vary from run to run.
constexpr int kRedBlocks = 1024;
constexpr int kRedPartDoubles = 3 * kRedBlocks;
// acc[0] = sum (ref-before)^2, acc[1] = sum (ref-warped)^2 in double (logging
// only); part: kRedPartDoubles of scratch.
void launch_ssd(double* acc, double* part, const float* ref, const float* warped, const float* before, long long sz,
                hipStream_t s);
// acc[0] = sum det J, acc[1] = sum det J^2, acc[2] = #(det J < 0) in double
// (logging only); part: kRedPartDoubles of scratch.
void launch_jacobian(double* acc, double* part, const float* u, const float* v, const float* w, int m, int n, int o,
                     int factor, hipStream_t s);

// ---- affine pre-alignment (linear.hip) ------------------------------------
// out = in - sub (the +1024 HU offset of linearBCV.cpp:146-154 with sub = -1024).
void launch_lin_offset(float* out, const float* in, float sub, long long n, hipStream_t s);
// affineLTS1.h:39-75: trilinear warp of `in` by the affine X (16 floats,
// column-wise), zero where the neighbourhood leaves the volume.
void launch_lin_warp(float* out, const float* in, const float* X, int m, int n, int o, hipStream_t s);
// affineLTS1.h:122-148: the three in-place smoothing passes of one cost volume.
void launch_lin_smooth(float* cost, int m1, int n1, int o1, int len3, hipStream_t s);
// affineLTS1.h:150-190: per node the first minimum of c1 / c2 (masked to 1e20
// for background and border nodes) into minval[0..N) / [N..2N), the label
// index into minind.
void launch_lin_nodes(const float* c1, const float* c2, const float* im1, const float* im2, int m, int n, int step,
                      int m1, int n1, int o1, int len3, float* minval, int* minind, hipStream_t s);

}  // namespace deeds
