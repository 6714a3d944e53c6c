/* deeds_hip -- MI355X (gfx950) implementation of deedsBCV's dense-displacement
 * registration core, exported as a plain C ABI.
 *
 * Conventions (all entry points):
 *   - Volumes use the reference voxel order idx = i + j*m + k*m*n, i along the
 *     first NIfTI dimension; m, n, o are the three dimensions.
 *   - Displacements: u along j (dim 2), v along i (dim 1), w along k (dim 3)
 *     (deedsBCV0.cpp:193-194).  Labels l = ys + xs*len + zs*len^2, len = 2*hw+1.
 *   - Every function returns 0 on success and a nonzero deeds_status on error;
 *     deeds_last_error() (thread-local) / deeds_ctx_last_error() give the text.
 *     Nothing aborts the process.
 *   - Host-pointer forms (no ctx argument) take and return host memory, run
 *     synchronously on a per-thread default context (device 0 unless
 *     deeds_set_default_device was called) and mirror one reference function
 *     each, with the reference's globals (image_m/n/o) made explicit arguments.
 *   - deeds_ctx_* forms take DEVICE pointers, enqueue on the context's stream
 *     and return before completion (deeds_ctx_synchronize waits), except the
 *     registration drivers, which return when the result is complete.
 *   - The caller owns every buffer passed in; a context owns its workspace.
 *     A context may be used by one host thread at a time; distinct contexts
 *     are independent.
 */
#ifndef DEEDS_HIP_H
#define DEEDS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum deeds_status {
    DEEDS_OK = 0,
    DEEDS_ERR_ARG = 1,      /* invalid argument (sizes, radii, null pointers) */
    DEEDS_ERR_HIP = 2,      /* a HIP runtime call failed */
    DEEDS_ERR_ALLOC = 3,    /* device or host allocation failed */
    DEEDS_ERR_INTERNAL = 4  /* internal consistency check failed */
};

typedef struct deeds_ctx deeds_ctx;

/* Registration schedule: deedsBCV's -a -G -L -Q flags (parseArguments.h:12-174,
 * defaults deedsBCV0.cpp:76-79). */
#define DEEDS_MAX_LEVELS 10
typedef struct deeds_schedule {
    int levels;                       /* number of levels (-l / count of -G values) */
    int grid[DEEDS_MAX_LEVELS];       /* control-point spacing per level (-G) */
    int radius[DEEDS_MAX_LEVELS];     /* label search radius hw per level (-L) */
    int quant[DEEDS_MAX_LEVELS];      /* label quantisation per level (-Q) */
    float alpha;                      /* regularisation weight (-a) */
} deeds_schedule;

/* Per-level timings in seconds plus work counters, the figures the reference
 * prints per level (deedsBCV0.cpp:308) and in total (deedsBCV0.cpp:381). */
typedef struct deeds_timing {
    int levels;
    double mind[DEEDS_MAX_LEVELS];
    double data[DEEDS_MAX_LEVELS];
    double mst_reg[DEEDS_MAX_LEVELS];
    double transf[DEEDS_MAX_LEVELS];
    double total;                     /* the reference's "Total time" span (:217 -> :336) */
    double hamming_evals;             /* algorithmic 64-bit Hamming evaluations, both directions */
    float ssd_before, ssd_after;      /* logging-only SSD of the last warp (dataCostD.h:300-312) */
    float jac_std[DEEDS_MAX_LEVELS];  /* std(J) logged per level (transformations.h:166) */
    float jac_neg[DEEDS_MAX_LEVELS];  /* fraction of J<0 per level (transformations.h:168) */
    /* Hamming data-cost kernel: summed launch durations (HIP events on the
     * launching stream), number of launches and evaluations per level. */
    double datacost_kernel_s;
    int datacost_launches;
    double datacost_level_s[DEEDS_MAX_LEVELS];
    double datacost_level_evals[DEEDS_MAX_LEVELS];
    /* host MST work: summed build time of all trees, and the time the
     * launching thread spent waiting for trees that were not ready yet */
    double tree_build_s;
    double tree_wait_s;
    double tree_build_level_s[DEEDS_MAX_LEVELS];  /* per level: max of the two directions' build times */
    double tree_wait_level_s[DEEDS_MAX_LEVELS];   /* per level: host wait for that level's trees */
    double stats_s;                               /* host image mean/std (both images, concurrent) */
    double bp_idle_level_s[DEEDS_MAX_LEVELS];     /* per level: GPU idle between each direction's data cost
                                                     and its BP (waiting for the host tree), summed */
    double bp_level_s[DEEDS_MAX_LEVELS];          /* per level: forward-direction BP time on its stream */
    double tree_ready_level_s[DEEDS_MAX_LEVELS];  /* per level: host time from the registration's start
                                                     until both trees were built */
    double tree_shape_level_s[DEEDS_MAX_LEVELS];  /* per level: the part of the build spent in Prim's
                                                     (edge ranks, the heap loop, the BFS order), max of both */
} deeds_timing;

/* ---- library / context ------------------------------------------------- */
const char* deeds_version(void);
const char* deeds_last_error(void);
int deeds_set_default_device(int device);
int deeds_ctx_create(int device, deeds_ctx** out);
int deeds_ctx_destroy(deeds_ctx* ctx);
const char* deeds_ctx_last_error(const deeds_ctx* ctx);
/* Run subsequent work of ctx on an existing HIP stream (e.g. torch's); NULL restores the own stream. */
int deeds_ctx_set_stream(deeds_ctx* ctx, void* hip_stream);
int deeds_ctx_synchronize(deeds_ctx* ctx);
/* Release the context's cached workspace (it is re-grown on demand). */
int deeds_ctx_trim(deeds_ctx* ctx);

/* The tile shapes the library's autotuner has chosen in this process so far
 * (per data-cost shape: label planes per workgroup, pipelined inner loop, the
 * measured kernel time) and the number of batch level graphs captured and
 * replayed, as a JSON object; writes at most cap-1 bytes plus a
 * terminating zero into buf (may be NULL) and returns the full length.  No
 * reference counterpart (the reference has fixed loops). */
size_t deeds_tuning_report(char* buf, size_t cap);
/* Pins the data-cost tile shape for later calls in this process instead of
 * the autotuner's choice: label planes per workgroup (0 restores tuning) and
 * the form, bit 0 the pipelined inner loop, bit 1 sample rows kept from plane
 * to plane (each used where the shape allows it).  Every shape gives the same
 * cost volume. */
int deeds_set_datacost_tiles(int planes_per_workgroup, int pipelined);

/* Fault injection for testing callers' error paths (the reference's failure
 * points are its I/O exits, imageIOgzType.h:15-17,94-96, deedsBCV0.cpp:92-99):
 * after `after` further events of `kind` succeed, the next one fails with a
 * status -- DEEDS_ERR_ALLOC for an allocation (device or pinned workspace),
 * DEEDS_ERR_HIP for a kernel launch (checked after the launch, whose kernel
 * runs normally) -- and the kind disarms; after < 0 disarms it.  The same is
 * armed at load time by DEEDS_FAULT_INJECT=alloc:N or launch:N.  A failed
 * call leaves no work of its own running (its context streams are drained). */
enum deeds_fault_kind { DEEDS_FAULT_ALLOC = 1, DEEDS_FAULT_LAUNCH = 2 };
int deeds_inject_fault(int kind, long long after);

/* ---- host-pointer forms: one per reference function ---------------------- */

/* MINDSSCbox.h:127 descriptor(): 12-channel MIND-SSC, 60-bit codes. */
int deeds_descriptor(uint64_t* mindq, const float* im, int m, int n, int o, int qs);

/* dataCostD.h:108 dataCostCL(): cost[node*len3 + l], len3 = (2hw+1)^3, node over
 * the (m/step)x(n/step)x(o/step) control grid. */
int deeds_datacost(const uint64_t* fixed_mind, const uint64_t* moving_mind, float* cost,
                   int m, int n, int o, int step, int hw, float quant, float alpha, int randnum);

/* primsMST.h:52 primsGraph(): MST of the control grid, BFS-ordered. */
int deeds_prims_graph(const float* im, int* ordered, int* parents, float* edgemst,
                      int step, int m, int n, int o);

/* primsMST.h:101-110 (inside primsGraph): the image's float mean and standard
 * deviation, bit-identical to the reference's sequential accumulation.  Host
 * only (no GPU work). */
int deeds_image_stats(const float* im, long long count, float* mean, float* sd);

/* regularisation.h:115 regularisationCL(): tree BP; cost is consumed (overwritten). */
int deeds_regularisation(float* cost, const float* u0, const float* v0, const float* w0,
                         float* u1, float* v1, float* w1, int hw, int step, float quant,
                         const int* ordered, const int* parents, const float* edgemst,
                         int m, int n, int o);

/* transformations.h:245 upsampleDeformationsCL(): (m2,n2,o2) field -> (m,n,o). */
int deeds_upsample_deformations(float* u1, float* v1, float* w1,
                                const float* u0, const float* v0, const float* w0,
                                int m, int n, int o, int m2, int n2, int o2);

/* transformations.h:189 consistentMappingCL(): in place on both fields. */
int deeds_consistent_mapping(float* u, float* v, float* w, float* u2, float* v2, float* w2,
                             int m, int n, int o, int factor);

/* transformations.h:84 jacobian(): std of det(J) of x + u/factor, and the J<0 fraction. */
int deeds_jacobian(const float* u, const float* v, const float* w, int m, int n, int o, int factor,
                   float* std_out, float* neg_frac_out);

/* dataCostD.h:267 warpAffine(): trilinear warp of `in` by X (16 floats,
 * column-wise as read at deedsBCV0.cpp:161-165) plus the field; ssd01
 * (optional, 2 floats) receives the logging SSD before/after vs im1b. */
int deeds_warp_affine(float* out, const float* in, const float* im1b, const float* X,
                      const float* u, const float* v, const float* w, int m, int n, int o, float* ssd01);

/* dataCostD.h:213 warpImageCL(): trilinear warp of im1 by x + field. */
int deeds_warp_image(float* out, const float* im1, const float* im1b,
                     const float* u, const float* v, const float* w, int m, int n, int o, float* ssd01);

/* dataCostD.h:241 warpAffineS(): nearest-neighbour label warp. */
int deeds_warp_affine_s(int16_t* out, const int16_t* in, const float* X,
                        const float* u, const float* v, const float* w, int m, int n, int o);

/* deedsBCV0.cpp:139-356: the full symmetric registration of fixed/moving
 * (HU values, m*n*o floats each).  flow_out receives the final control-grid
 * field [u;v;w] (3 * (m/g)(n/g)(o/g) floats for the last grid g) -- the
 * _displacements.dat payload; warped_out (optional) the deformed moving image
 * in HU.  X may be NULL for identity.  timing is optional. */
int deeds_register(const float* fixed, const float* moving, const float* X, int m, int n, int o,
                   const deeds_schedule* sched, float* flow_out, float* warped_out, deeds_timing* timing);

/* Multi-atlas batch: register `count` independent pairs (each with its own
 * dimensions and affine) with the same schedule, in groups of `concurrency`
 * pairs that run level-synchronously: every pair's MIND / data-cost kernels
 * of a level, then the belief propagation of all the group's trees in one
 * launch sequence (one launch per round for all trees), then every pair's
 * field updates.  Each result is bit-identical to a deeds_register call on
 * that pair alone (no reference counterpart: the reference registers one
 * pair per process, deedsBCV0.cpp:55).  Device memory grows with the group
 * (about 1.5 GB per 200x170x140 pair).  Pointers are host memory for
 * deeds_register_batch and device memory of `device` for
 * deeds_register_batch_device; timings (optional) has `count` entries. */
typedef struct deeds_pair {
    const float* fixed;   /* m*n*o HU values */
    const float* moving;  /* m*n*o HU values */
    const float* X;       /* 16 floats or NULL (identity); always host memory */
    int m, n, o;
    float* flow;          /* 3*sz1 floats (final control grid) */
    float* warped;        /* m*n*o floats or NULL */
} deeds_pair;
int deeds_register_batch(const deeds_pair* pairs, int count, const deeds_schedule* sched, int concurrency,
                         deeds_timing* timings);
int deeds_register_batch_device(int device, const deeds_pair* pairs, int count, const deeds_schedule* sched,
                                int concurrency, deeds_timing* timings);
/* Batch level graphs (no reference counterpart).  A group's pre- and post-BP
 * work of every level is a fixed launch sequence for given pairs and
 * schedule; the second batch call with the same pairs (same buffers, sizes,
 * affines) captures it into one HIP graph per level and phase, and later calls
 * replay the graphs (the BP between them, which depends on the host-built
 * trees, stays on the streams).  Results are identical either way.  On by
 * default (DEEDS_GRAPHS=0 in the environment turns it off at load). */
int deeds_set_batch_graphs(int enabled);

/* applyBCV.cpp:161-198 / applyBCVfloat.cpp:186-198: apply a _displacements.dat
 * field (3*sz1 floats on grid `grid_step`) to a label volume (nearest) or an
 * image (trilinear). */
int deeds_apply_seg(int16_t* out, const int16_t* in, const float* X, const float* flow, int grid_step,
                    int m, int n, int o);
int deeds_apply_float(float* out, const float* in, const float* X, const float* flow, int grid_step,
                      int m, int n, int o);

/* diceMulti.cpp:17-79: Dice overlap 2|A_l & B_l| / (|A_l| + |B_l|) of every
 * label l of seg1 except its smallest value (the background), in increasing
 * label order, computed in float as the reference does.  labels / dice
 * (either may be NULL) receive up to `capacity` entries; *count is the number
 * of labels, which may exceed capacity. */
int deeds_dice(const int16_t* seg1, const int16_t* seg2, long long voxels, int capacity, int16_t* labels,
               float* dice, int* count);

/* linearBCV.cpp:144-263 (linearBCV's estimation): multi-level affine (or,
 * with rigid != 0, rigid) alignment of moving to fixed (HU volumes; the +1024
 * offset is applied inside) using sched's grid / radius / quant per level
 * (alpha is not used: the data cost runs with alpha 1).  X receives the 16
 * floats column-wise, as linearBCV writes them to _matrix.txt and deedsBCV
 * reads them with -A.  points_used (optional, sched->levels entries) receives
 * the number of point pairs fitted per level (the "# points used" figure). */
int deeds_linear_register(const float* fixed, const float* moving, int m, int n, int o,
                          const deeds_schedule* sched, int rigid, float* X, int* points_used);

/* ---- device-pointer forms (asynchronous on the context stream) ---------- */
int deeds_ctx_descriptor(deeds_ctx* ctx, uint64_t* d_mindq, const float* d_im, int m, int n, int o, int qs);
/* linearBCV.cpp:144-263 on device volumes (synchronous): as deeds_linear_register. */
int deeds_ctx_linear_register(deeds_ctx* ctx, const float* d_fixed, const float* d_moving, int m, int n, int o,
                              const deeds_schedule* sched, int rigid, float* X, int* points_used);
/* diceMulti.cpp:17-79 on device volumes (synchronous): as deeds_dice. */
int deeds_ctx_dice(deeds_ctx* ctx, const int16_t* d_seg1, const int16_t* d_seg2, long long voxels, int capacity,
                   int16_t* labels, float* dice, int* count);
/* Allocation-free Dice.  No counterpart in the reference: a variant of
 * deeds_ctx_dice for callers that cannot allocate device or pinned memory
 * (code under HIP graph capture, tight serving loops).  Every intermediate
 * (presence bitmap, value->label table, per-label counters) lives in a fixed
 * file-scope device array of about 1 MiB per GPU, held until the process ends;
 * the call makes no hipMalloc/hipFree/hipHostMalloc and no host round trip:
 * it only enqueues kernels on ctx's stream and returns.  Outputs are DEVICE
 * memory: d_labels / d_dice (capacity entries; may be NULL when capacity is 0)
 * receive the first `capacity` labels after the smallest and their Dice, as
 * deeds_ctx_dice returns them, and *d_count (device int) their number.
 * Limits: voxels < 2^32 (32-bit counters), capacity <= 65536; larger values
 * are rejected with DEEDS_ERR_ARG.  All calls on one GPU share the array, so
 * calls that overlap in time on one GPU (two streams, two threads) race and
 * must be serialised by the caller; each call clears what it needs first.
 * The first call on a device resolves the array's address (before any graph
 * capture). */
int deeds_ctx_dice_noalloc(deeds_ctx* ctx, const int16_t* d_seg1, const int16_t* d_seg2, long long voxels,
                           int capacity, int16_t* d_labels, float* d_dice, int* d_count);
/* primsMST.h:52 primsGraph() on a device image: the same tree as
 * deeds_prims_graph (edge SADs and the exact image statistics on the GPU,
 * Prim's algorithm on the host); ordered/parents/edgemst are host arrays of
 * (m/step)(n/step)(o/step) entries.  Synchronous. */
int deeds_ctx_prims_graph(deeds_ctx* ctx, const float* d_im, int* ordered, int* parents, float* edgemst, int step,
                          int m, int n, int o);
/* primsMST.h:101-110 on a device image: the same bit-identical mean and
 * standard deviation as deeds_image_stats, the per-chunk work on the GPU.
 * Returns when the result is complete. */
int deeds_ctx_image_stats(deeds_ctx* ctx, const float* d_im, long long count, float* mean, float* sd);
int deeds_ctx_datacost(deeds_ctx* ctx, const uint64_t* d_fixed_mind, const uint64_t* d_moving_mind, float* d_cost,
                       int m, int n, int o, int step, int hw, float quant, float alpha, int randnum);
/* BP over a tree given on the host (ordered/parents/edgemst as returned by
 * deeds_prims_graph); d_cost is consumed.  d_allinds (optional, sz1*len3
 * int16) receives the per-node argmin tables of messageDT
 * (regularisation.h:6-113), the root's row being the identity
 * (regularisation.h:273-274).  Returns when the work is complete. */
int deeds_ctx_regularisation(deeds_ctx* ctx, float* d_cost, const float* d_u0, const float* d_v0, const float* d_w0,
                             float* d_u1, float* d_v1, float* d_w1, int hw, int step, float quant,
                             const int* ordered, const int* parents, const float* edgemst,
                             int m, int n, int o, int16_t* d_allinds);
/* Field operations on device arrays (asynchronous on the context stream,
 * except where a host result is returned).  Arguments as in the host forms
 * above; X (16 floats or NULL for identity) and ssd01 are host memory; out
 * must not alias the input volume. */
/* transformations.h:245 upsampleDeformationsCL() */
int deeds_ctx_upsample_deformations(deeds_ctx* ctx, float* d_u1, float* d_v1, float* d_w1, const float* d_u0,
                                    const float* d_v0, const float* d_w0, int m, int n, int o, int m2, int n2,
                                    int o2);
/* transformations.h:189 consistentMappingCL(): in place on the six fields */
int deeds_ctx_consistent_mapping(deeds_ctx* ctx, float* d_u, float* d_v, float* d_w, float* d_u2, float* d_v2,
                                 float* d_w2, int m, int n, int o, int factor);
/* transformations.h:84 jacobian() (synchronous: returns the two figures) */
int deeds_ctx_jacobian(deeds_ctx* ctx, const float* d_u, const float* d_v, const float* d_w, int m, int n, int o,
                       int factor, float* std_out, float* neg_frac_out);
/* dataCostD.h:267 warpAffine(); with ssd01 and d_im1b given it waits for the logging SSD */
int deeds_ctx_warp_affine(deeds_ctx* ctx, float* d_out, const float* d_in, const float* d_im1b, const float* X,
                          const float* d_u, const float* d_v, const float* d_w, int m, int n, int o, float* ssd01);
/* dataCostD.h:213 warpImageCL() */
int deeds_ctx_warp_image(deeds_ctx* ctx, float* d_out, const float* d_im1, const float* d_im1b, const float* d_u,
                         const float* d_v, const float* d_w, int m, int n, int o, float* ssd01);
/* dataCostD.h:241 warpAffineS() */
int deeds_ctx_warp_affine_s(deeds_ctx* ctx, int16_t* d_out, const int16_t* d_in, const float* X, const float* d_u,
                            const float* d_v, const float* d_w, int m, int n, int o);
int deeds_ctx_register(deeds_ctx* ctx, const float* d_fixed, const float* d_moving, const float* X,
                       int m, int n, int o, const deeds_schedule* sched,
                       float* d_flow, float* d_warped, deeds_timing* timing);

/* applyBCV.cpp:161-198 / applyBCVfloat.cpp:186-198 on device volumes: d_flow
 * is the 3*sz1-float _displacements.dat payload on grid `grid_step`; X host
 * memory (16 floats or NULL).  Asynchronous; d_out must not alias d_in. */
int deeds_ctx_apply_seg(deeds_ctx* ctx, int16_t* d_out, const int16_t* d_in, const float* X, const float* d_flow,
                        int grid_step, int m, int n, int o);
int deeds_ctx_apply_float(deeds_ctx* ctx, float* d_out, const float* d_in, const float* X, const float* d_flow,
                          int grid_step, int m, int n, int o);

/* ---- differentiable soft variant (no reference counterpart) --------------
 * A smooth relaxation of one registration level beside the plain path, for
 * callers that need gradients (learning pipelines, torch.autograd):
 *   features  phi[p][l] (l = 0..11): the un-binarised self-similarity values
 *             that descriptor() thresholds (MINDSSCbox.h:184-201), i.e. the
 *             twelve box-filtered SSDs minus their minimum over max(mean, 1e-6)
 *   cost      C[n][l] = a1 * sum over the block samples of sum_ch
 *             (phi_F[s][ch] - phi_M[clamp(s + d_l)][ch])^2 with the samples,
 *             clamping and a1 of dataCostCL (dataCostD.h:141-196)
 *   BP        sum-product belief propagation at temperature T on a tree in
 *             BFS order (ordered/parents/edgemst as from deeds_prims_graph) for
 *             E = sum_n C_n + sum_c |l_p - l_c + (u0_p - u0_c)/quant|^2 / edgemst[c];
 *             d_u1 = d_u0 + the expected label displacement of each marginal
 * Each forward has an exact backward (vector-Jacobian product).  `real`
 * selects float (DEEDS_REAL_F32) or double (DEEDS_REAL_F64) for every void*
 * array of the call.  Layouts: images m*n*o, features [voxel][12], cost
 * [node][label], fields [3][N] (u, v, w), d_saved deeds_soft_saved_count(N, hw)
 * reals written by deeds_ctx_soft_bp and read by its backward.  The BP
 * calls return when their work is complete; the others are asynchronous. */
enum deeds_real { DEEDS_REAL_F32 = 0, DEEDS_REAL_F64 = 1 };
size_t deeds_soft_saved_count(int N, int hw);
int deeds_ctx_soft_features(deeds_ctx* ctx, int real, void* d_phi, const void* d_im, int m, int n, int o, int qs);
int deeds_ctx_soft_features_backward(deeds_ctx* ctx, int real, void* d_gim, const void* d_gphi, const void* d_im,
                                     int m, int n, int o, int qs);
int deeds_ctx_soft_cost(deeds_ctx* ctx, int real, void* d_cost, const void* d_fixed_phi, const void* d_moving_phi,
                        int m, int n, int o, int step, int hw, float quant, float alpha);
int deeds_ctx_soft_cost_backward(deeds_ctx* ctx, int real, void* d_gfixed, void* d_gmoving, const void* d_gcost,
                                 const void* d_fixed_phi, const void* d_moving_phi, int m, int n, int o, int step,
                                 int hw, float quant, float alpha);
int deeds_ctx_soft_bp(deeds_ctx* ctx, int real, void* d_u1, void* d_saved, const void* d_cost, const void* d_u0,
                      const int* ordered, const int* parents, const float* edgemst, int m, int n, int o, int step,
                      int hw, float quant, float T);
int deeds_ctx_soft_bp_backward(deeds_ctx* ctx, int real, void* d_gcost, const void* d_gu1, const void* d_saved,
                               const void* d_cost, const void* d_u0, const int* ordered, const int* parents,
                               const float* edgemst, int m, int n, int o, int step, int hw, float quant, float T);

#ifdef __cplusplus
}
#endif
#endif
