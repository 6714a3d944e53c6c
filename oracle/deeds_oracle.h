/* CPU restatement of deedsBCV's registration hot path -- TEST INFRASTRUCTURE.
 *
 * This library is the oracle the HIP product is checked against.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it, and only as
 * the checker; the product library never links or calls it.
 *
 * Volumes use the reference's voxel order  idx = i + j*m + k*m*n  (i along the
 * first NIfTI dimension).  Every function names the reference file:line it
 * restates.  Arithmetic follows the reference's C++ evaluation exactly
 * (float vs double promotion of each sub-expression, left-to-right sums, no FMA).
 */
#ifndef DEEDS_ORACLE_H
#define DEEDS_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* MIND-SSC (MINDSSCbox.h) */
void orc_boxfilter(float* vol, int hw, int m, int n, int o);
void orc_ssd_channel(const float* im, float* out, int m, int n, int o, int qs, int l);
void orc_mind_thresholds(float thr[5]);
void orc_descriptor(uint64_t* codes, const float* im, int m, int n, int o, int qs);

/* Hamming data cost (dataCostD.h) */
float orc_datacost_scale(int step, float quant, float alpha, int randnum, int skip[3]);
void orc_datacost(const uint64_t* F, const uint64_t* M, float* cost, int m, int n, int o,
                  int step, int hw, float quant, float alpha, int randnum);

/* MST (primsMST.h) */
void orc_edge_sad(const float* im, float* edgecost, int* nbr, int step, int m, int n, int o);
void orc_image_stats(const float* im, long long count, float* mean, float* stddev);
void orc_prims(const float* im, int* ordered, int* parents, float* edgemst, int step, int m, int n, int o);

/* belief propagation (regularisation.h) */
void orc_message_dt(float* cube, short* inds, int len, float offx, float offy, float offz);
void orc_regularisation(float* cost, const float* u0, const float* v0, const float* w0,
                        float* u1, float* v1, float* w1, int hw, int step, float quant,
                        const int* ordered, const int* parents, const float* edgemst,
                        int m, int n, int o, short* allinds_out, int* selected_out);

/* field transforms and warps (transformations.h, dataCostD.h) */
void orc_interp3(float* out, const float* in, const float* x1, const float* y1, const float* z1,
                 int m, int n, int o, int m2, int n2, int o2, int add_identity);
void orc_upsample(float* u1, float* v1, float* w1, const float* u0, const float* v0, const float* w0,
                  int m, int n, int o, int m2, int n2, int o2);
void orc_consistent(float* u, float* v, float* w, float* u2, float* v2, float* w2,
                    int m, int n, int o, int factor);
float orc_jacobian(const float* u, const float* v, const float* w, int m, int n, int o, int factor,
                   float* neg_frac);
void orc_warp_affine(float* out, const float* in, const float* im1b, const float* X,
                     const float* u, const float* v, const float* w, int m, int n, int o, float* ssd01);
void orc_warp_image(float* out, const float* im1, const float* im1b,
                    const float* u, const float* v, const float* w, int m, int n, int o, float* ssd01);
void orc_warp_affine_s(short* out, const short* in, const float* X,
                       const float* u, const float* v, const float* w, int m, int n, int o);

/* diceMulti.cpp:38-55: Dice of every label of seg1 but its smallest, in
 * increasing label order; returns the number of labels (entries written up to
 * capacity). */
int orc_dice(const short* seg1, const short* seg2, long long sz, int capacity, short* labels, float* dice);

/* affine pre-alignment (linearBCV.cpp, affineLTS1.h, QRsolve.h) */
void orc_lin_warp(float* out, const float* in, const float* X, int m, int n, int o);
void orc_qrsolve(float* X, const float* A, const float* b, int len, int len2);
void orc_find_rigid(float* RT, const float* pts1t, const float* pts2t, int len);
void orc_affine_robust(float* RT, const float* pts1, const float* pts2, int len, int rigid);
void orc_estimate_affine(float* X, const float* Xprev, const float* im1, const float* im2, float* cost, float* cost2,
                         int m, int n, int o, int step, float quant, int hw, int rigid, int* used);
void orc_linear_register(const float* fixed, const float* moving, int m, int n, int o, int levels, const int* grid,
                         const int* radius, const int* quant, int rigid, float* X_out);

/* the symmetric multi-level registration loop (deedsBCV0.cpp:183-349).
 * fixed/moving are HU volumes (the +1024 offset of deedsBCV0.cpp:139-144 is
 * applied inside).  flow_out receives the final control-grid field [u;v;w]
 * (3*sz1 floats, the _displacements.dat payload); warped_out, if not NULL,
 * the deformed moving image in HU (m*n*o floats).  Returns sz1 or -1. */
int orc_register(const float* fixed, const float* moving, const float* X, int m, int n, int o,
                 int levels, const int* grid, const int* radius, const int* quant, float alpha,
                 float* flow_out, float* warped_out, int verbose);

#ifdef __cplusplus
}
#endif
#endif
