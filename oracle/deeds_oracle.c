/* CPU restatement of deedsBCV's registration hot path -- TEST INFRASTRUCTURE.
 *
 * The checker for the HIP product (never linked into it).  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 *
 * Each function restates the reference function named in its comment.  The
 * arithmetic reproduces the reference's C++ evaluation exactly: which
 * sub-expressions are float and which are promoted to double by a double
 * literal (1.0, 0.5), left-to-right summation order, sequential prefix sums,
 * strict-< argmin tie rules, and no FMA contraction (Makefile builds with
 * -ffp-contract=off, the reference build has no -mfma).  Function-call
 * overloads of the reference's C++ (log/exp/sqrt/floor/round/ceil on a float
 * argument pick the float versions; pow(float,int) promotes to double) are
 * spelled out explicitly as the C functions they resolve to.
 *
 * Parity pins (tests/): every function here is compared bit-for-bit with the
 * unmodified reference compiled from its own sources (oracle/_ref) and with
 * the golden vectors in tests/golden/, plus closed-form / brute-force checks.
 */
#include "deeds_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define IDX(i, j, k, m, n) ((long long)(i) + (long long)(j) * (m) + (long long)(k) * (m) * (n))

static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
static void* xmalloc(size_t bytes) {
    void* p = malloc(bytes ? bytes : 1);
    if (!p) { fprintf(stderr, "oracle: out of memory (%zu bytes)\n", bytes); abort(); }
    return p;
}

/* ------------------------------------------------------------------------- */
/* MIND-SSC                                                                   */
/* ------------------------------------------------------------------------- */

/* One strided line: sequential float prefix sum, then the clamped window
 * difference  out[t] = P[min(t+hw, L-1)] - P[t-hw-1]  (second term absent for
 * t <= hw).  MINDSSCbox.h:9-29 (and the j / k copies at :31-73). */
static void box_line(const float* src, float* dst, long long stride, int L, int hw, float* P) {
    float acc = src[0];
    P[0] = acc;
    for (int t = 1; t < L; t++) { acc = acc + src[(long long)t * stride]; P[t] = acc; }
    for (int t = 0; t < L; t++) {
        int hi = t + hw < L - 1 ? t + hw : L - 1;
        float v = (t <= hw) ? P[hi] : P[hi] - P[t - hw - 1];
        dst[(long long)t * stride] = v;
    }
}

/* MINDSSCbox.h:2-76: separable clamped box SUM of radius hw along i, j, k. */
void orc_boxfilter(float* vol, int hw, int m, int n, int o) {
    long long sz = (long long)m * n * o;
    float* tmp = (float*)xmalloc(sizeof(float) * sz);
    int maxl = m > n ? m : n; maxl = maxl > o ? maxl : o;
    float* P = (float*)xmalloc(sizeof(float) * maxl);
    for (int k = 0; k < o; k++)
        for (int j = 0; j < n; j++) box_line(vol + IDX(0, j, k, m, n), tmp + IDX(0, j, k, m, n), 1, m, hw, P);
    for (int k = 0; k < o; k++)
        for (int i = 0; i < m; i++) box_line(tmp + IDX(i, 0, k, m, n), vol + IDX(i, 0, k, m, n), m, n, hw, P);
    for (int j = 0; j < n; j++)
        for (int i = 0; i < m; i++) {
            /* k pass reads vol and writes tmp, then copy back (reference ping-pongs) */
            box_line(vol + IDX(i, j, 0, m, n), tmp + IDX(i, j, 0, m, n), (long long)m * n, o, hw, P);
        }
    memcpy(vol, tmp, sizeof(float) * sz);
    free(tmp); free(P);
}

/* Six self-similarity shifts, in (j, i, k) = (dx, dy, dz) units of qs
 * (MINDSSCbox.h:108-110 and :131-133). */
static const int SSD_DJ[6] = {+1, +1, -1, 0, +1, 0};
static const int SSD_DI[6] = {+1, -1, 0, -1, 0, +1};
static const int SSD_DK[6] = {0, 0, +1, +1, +1, +1};
/* Twelve SSC pairs: neighbour offset (j, i, k) in units of qs and the SSD
 * channel it reads (MINDSSCbox.h:135-139). */
static const int SSC_DJ[12] = {-1, 0, -1, 0, 0, +1, 0, 0, 0, -1, 0, 0};
static const int SSC_DI[12] = {0, -1, 0, +1, 0, 0, 0, +1, 0, 0, 0, -1};
static const int SSC_DK[12] = {0, 0, 0, 0, -1, 0, -1, 0, -1, 0, -1, 0};
static const int SSC_CH[12] = {0, 0, 1, 1, 2, 2, 3, 3, 4, 4, 5, 5};

/* MINDSSCbox.h:79-90 (imshift) + :104-124 (distances): squared difference to
 * the shifted image (0 where the shift leaves the volume), then box-summed. */
void orc_ssd_channel(const float* im, float* out, int m, int n, int o, int qs, int l) {
    int di = SSD_DI[l] * qs, dj = SSD_DJ[l] * qs, dk = SSD_DK[l] * qs;
    for (int k = 0; k < o; k++)
        for (int j = 0; j < n; j++)
            for (int i = 0; i < m; i++) {
                long long p = IDX(i, j, k, m, n);
                int inside = i + di >= 0 && i + di < m && j + dj >= 0 && j + dj < n && k + dk >= 0 && k + dk < o;
                float s = inside ? im[IDX(i + di, j + dj, k + dk, m, n)] : im[p];
                float d = s - im[p];
                out[p] = d * d;
            }
    orc_boxfilter(out, qs, m, n, o);
}

/* MINDSSCbox.h:176-180: thresholds -log((c+1.5)/6) evaluated in float. */
void orc_mind_thresholds(float thr[5]) {
    for (int c = 0; c < 5; c++) thr[c] = -logf((c + 1.5f) / 6);
}

/* MINDSSCbox.h:127-232: 12 SSC values, min-subtract, noise-normalise,
 * 6-level thermometer quantisation into 5-bit fields of a 64-bit code. */
void orc_descriptor(uint64_t* codes, const float* im, int m, int n, int o, int qs) {
    long long sz = (long long)m * n * o;
    float* d = (float*)xmalloc(sizeof(float) * sz * 6);
    for (int l = 0; l < 6; l++) orc_ssd_channel(im, d + l * sz, m, n, o, qs, l);
    float thr[5];
    orc_mind_thresholds(thr);
    for (int k = 0; k < o; k++)
        for (int j = 0; j < n; j++)
            for (int i = 0; i < m; i++) {
                float v[12];
                for (int l = 0; l < 12; l++) {
                    int ii = i + SSC_DI[l] * qs, jj = j + SSC_DJ[l] * qs, kk = k + SSC_DK[l] * qs;
                    int inside = ii >= 0 && ii < m && jj >= 0 && jj < n && kk >= 0 && kk < o;
                    long long p = inside ? IDX(ii, jj, kk, m, n) : IDX(i, j, k, m, n);
                    v[l] = d[p + SSC_CH[l] * sz];
                }
                float mn = v[0];
                for (int l = 1; l < 12; l++) if (v[l] < mn) mn = v[l];
                float sum = 0.0f;
                for (int l = 0; l < 12; l++) { v[l] = v[l] - mn; sum = sum + v[l]; }
                float noise = sum / (float)12;
                if (noise < 1e-6f) noise = 1e-6f;
                uint64_t code = 0;
                for (int l = 0; l < 12; l++) {
                    float q = v[l] / noise;
                    int level = 0;
                    for (int c = 0; c < 5; c++) level += thr[c] > q ? 1 : 0;
                    code |= (uint64_t)((1u << level) - 1u) << (5 * l);
                }
                codes[IDX(i, j, k, m, n)] = code;
            }
    free(d);
}

/* ------------------------------------------------------------------------- */
/* Hamming data cost                                                          */
/* ------------------------------------------------------------------------- */

/* dataCostD.h:141-163: sample skips and the float scale applied to the
 * integer Hamming sum.  skip[] = {i (y), j (x), k (z)}. */
float orc_datacost_scale(int step, float quant, float alpha, int randnum, int skip[3]) {
    int sz_ = 1, sx = 1, sy = 1;
    if (step > 4) {
        if (randnum > 0) { sz_ = 2; sx = 2; }
        if (randnum > 1) sy = 2;
    }
    if (randnum > 1 && step > 7) { sz_ = 3; sx = 3; sy = 3; }
    if (step == 4 && randnum > 1) sz_ = 2;
    skip[0] = sy; skip[1] = sx; skip[2] = sz_;
    float maxsamp = ceilf((float)step / (float)sx) * ceilf((float)step / (float)sz_) * ceilf((float)step / (float)sy);
    float alphai = (float)step / (alpha * quant);
    return (float)(0.5 * (double)alphai / (double)maxsamp);
}

/* dataCostD.h:108-210.  cost[node*len3 + l] = scale * sum over block samples
 * of popcount(F ^ M(clamped displaced position)). */
void orc_datacost(const uint64_t* F, const uint64_t* M, float* cost, int m, int n, int o,
                  int step, int hw, float quant, float alpha, int randnum) {
    int len = 2 * hw + 1, len3 = len * len * len;
    int m1 = m / step, n1 = n / step, o1 = o / step;
    int q = (int)quant;
    int pad = q * hw;
    int skip[3];
    float scale = orc_datacost_scale(step, quant, alpha, randnum, skip);
    for (int z = 0; z < o1; z++)
        for (int x = 0; x < n1; x++)
            for (int y = 0; y < m1; y++) {
                long long node = IDX(y, x, z, m1, n1);
                for (int l = 0; l < len3; l++) {
                    int ls = l % len, lx = (l / len) % len, lz = l / (len * len);
                    int oy = (int)(ls * quant) - pad, ox = (int)(lx * quant) - pad, oz = (int)(lz * quant) - pad;
                    int sum = 0;
                    for (int k = 0; k < step; k += skip[2])
                        for (int j = 0; j < step; j += skip[1])
                            for (int i = 0; i < step; i += skip[0]) {
                                int fi = y * step + i, fj = x * step + j, fk = z * step + k;
                                uint64_t a = F[IDX(fi, fj, fk, m, n)];
                                uint64_t b = M[IDX(clampi(fi + oy, 0, m - 1), clampi(fj + ox, 0, n - 1),
                                                   clampi(fk + oz, 0, o - 1), m, n)];
                                sum += __builtin_popcountll(a ^ b);
                            }
                    cost[node * len3 + l] = (float)sum * scale;
                }
            }
}

/* ------------------------------------------------------------------------- */
/* Minimum spanning tree                                                      */
/* ------------------------------------------------------------------------- */

/* Six grid neighbours in (i, j, k) order of primsMST.h:72-74. */
static const int NB_DI[6] = {0, 0, -1, 1, 0, 0};
static const int NB_DJ[6] = {-1, 1, 0, 0, 0, 0};
static const int NB_DK[6] = {0, 0, 0, 0, -1, 1};

/* primsMST.h:75-100: per control point and neighbour, the SAD between the two
 * step^3 voxel blocks accumulated in float in (k1, j1, i1) order; nbr = -1 and
 * cost 0 where the neighbour leaves the grid. */
void orc_edge_sad(const float* im, float* edgecost, int* nbr, int step, int m2, int n2, int o2) {
    int m = m2 / step, n = n2 / step, o = o2 / step;
    long long N = (long long)m * n * o;
    for (long long e = 0; e < 6 * N; e++) { edgecost[e] = 0.0f; nbr[e] = -1; }
    for (int k = 0; k < o; k++)
        for (int j = 0; j < n; j++)
            for (int i = 0; i < m; i++)
                for (int nb = 0; nb < 6; nb++) {
                    int ii = i + NB_DI[nb], jj = j + NB_DJ[nb], kk = k + NB_DK[nb];
                    if (!(ii >= 0 && ii < m && jj >= 0 && jj < n && kk >= 0 && kk < o)) continue;
                    long long e = IDX(i, j, k, m, n) + nb * N;
                    nbr[e] = (int)IDX(ii, jj, kk, m, n);
                    float acc = 0.0f;
                    for (int k1 = 0; k1 < step; k1++)
                        for (int j1 = 0; j1 < step; j1++)
                            for (int i1 = 0; i1 < step; i1++) {
                                float a = im[IDX(i * step + i1, j * step + j1, k * step + k1, m2, n2)];
                                float b = im[IDX(ii * step + i1, jj * step + j1, kk * step + k1, m2, n2)];
                                acc = acc + fabsf(a - b);
                            }
                    edgecost[e] = acc;
                }
}

/* primsMST.h:101-110: sequential float mean; variance accumulated as
 * float += double(pow(float,2)); float sqrt. */
void orc_image_stats(const float* im, long long count, float* mean, float* stddev) {
    float mu = 0.0f;
    for (long long t = 0; t < count; t++) mu = mu + im[t];
    mu = mu / (float)count;
    float var = 0.0f;
    for (long long t = 0; t < count; t++) {
        double d = (double)(im[t] - mu);
        var = (float)((double)var + d * d);
    }
    *mean = mu;
    *stddev = sqrtf(var / (float)count);
}

/* Binary heap with the exact sift rules of the C++ standard library the
 * reference links (libstdc++ std::push_heap / std::pop_heap), ordered by the
 * reference's Edge::operator< (primsMST.h:24-26: a<b iff a.weight > b.weight,
 * so the top is the lightest edge).  Ties resolve as the reference's do only
 * because the sift sequence is the same. */
typedef struct { double w; int a, b; } edge_t;
static inline int edge_less(const edge_t* x, const edge_t* y) { return x->w > y->w; }

static void heap_sift_up(edge_t* h, long long hole, long long top, edge_t v) {
    long long parent = (hole - 1) / 2;
    while (hole > top && edge_less(&h[parent], &v)) {
        h[hole] = h[parent];
        hole = parent;
        parent = (hole - 1) / 2;
    }
    h[hole] = v;
}
static void heap_push_back(edge_t* h, long long len) { heap_sift_up(h, len - 1, 0, h[len - 1]); }
static void heap_pop_to_back(edge_t* h, long long len) {
    if (len < 2) return;
    edge_t v = h[len - 1];
    h[len - 1] = h[0];
    long long n = len - 1, hole = 0, child = 0;
    while (child < (n - 1) / 2) {
        child = 2 * (child + 1);
        if (edge_less(&h[child], &h[child - 1])) child--;
        h[hole] = h[child];
        hole = child;
    }
    if ((n & 1) == 0 && child == (n - 2) / 2) {
        child = 2 * (child + 1);
        h[hole] = h[child - 1];
        hole = child - 1;
    }
    heap_sift_up(h, hole, 0, v);
}

/* primsMST.h:52-223: edge weights -exp(-c/(2 sigma)), Prim's from the grid
 * centre with a lazy heap, BFS-depth ordering (stable in vertex id), and the
 * weight of each node's edge to its parent. */
void orc_prims(const float* im, int* ordered, int* parents, float* edgemst, int step, int m2, int n2, int o2) {
    int m = m2 / step, n = n2 / step, o = o2 / step;
    int N = m * n * o;
    float* ec = (float*)xmalloc(sizeof(float) * 6 * (size_t)N);
    int* nbr = (int*)xmalloc(sizeof(int) * 6 * (size_t)N);
    orc_edge_sad(im, ec, nbr, step, m2, n2, o2);
    float mean, sd;
    orc_image_stats(im, (long long)m2 * n2 * o2, &mean, &sd);
    float vol = (float)((double)step * step * step);
    for (long long e = 0; e < 6LL * N; e++) ec[e] = ec[e] / vol;
    float two_sd = 2.0f * sd;
    for (long long e = 0; e < 6LL * N; e++) ec[e] = -expf(-ec[e] / two_sd);

    int root = m / 2 + (n / 2) * m + (o / 2) * m * n;
    unsigned char* in_tree = (unsigned char*)calloc(N, 1);
    int* depth = (int*)xmalloc(sizeof(int) * N);
    edge_t* heap = (edge_t*)xmalloc(sizeof(edge_t) * (12 * (size_t)N + 16));
    long long hlen = 0;
    for (int v = 0; v < N; v++) parents[v] = -1;
    depth[root] = 0;
    in_tree[root] = 1;
    int last = root;
    for (int it = 0; it < N - 1; it++) {
        for (int nb = 0; nb < 6; nb++) {
            int t = nbr[last + nb * N];
            if (t >= 0) {
                heap[hlen].w = (double)ec[last + nb * N];
                heap[hlen].a = last; heap[hlen].b = t;
                hlen++;
                heap_push_back(heap, hlen);
            }
        }
        int from = -1, to = -1;
        while (to < 0) {
            edge_t e = heap[0];
            heap_pop_to_back(heap, hlen);
            hlen--;
            int ia = in_tree[e.a], ib = in_tree[e.b];
            if (ia ^ ib) { if (ia) { from = e.a; to = e.b; } else { from = e.b; to = e.a; } }
        }
        in_tree[to] = 1;
        depth[to] = depth[from] + 1;
        parents[to] = from;
        last = to;
    }
    /* counting sort by depth, stable in vertex id (primsMST.h:166-190) */
    int maxd = 0;
    for (int v = 0; v < N; v++) if (depth[v] > maxd) maxd = depth[v];
    int* start = (int*)calloc(maxd + 2, sizeof(int));
    for (int v = 0; v < N; v++) start[depth[v] + 1]++;
    for (int d = 1; d <= maxd + 1; d++) start[d] += start[d - 1];
    for (int v = 0; v < N; v++) ordered[start[depth[v]]++] = v;
    /* edgemst[child] = exp(-c/(2 sigma)) of the edge to the parent; the scan
     * over the six neighbour slots keeps the last match (primsMST.h:198-211) */
    for (int v = 0; v < N; v++) edgemst[v] = 0.0f;
    for (int t = 1; t < N; t++) {
        int c = ordered[t], p = parents[c];
        int z = c / (m * n), x = (c - z * m * n) / m, y = c - z * m * n - x * m;
        for (int nb = 0; nb < 6; nb++) {
            int idx = y + NB_DI[nb] + (x + NB_DJ[nb]) * m + (z + NB_DK[nb]) * m * n;
            if (idx == p) edgemst[c] = -ec[c + nb * N];
        }
    }
    free(ec); free(nbr); free(in_tree); free(depth); free(heap); free(start);
}

/* ------------------------------------------------------------------------- */
/* Belief propagation                                                         */
/* ------------------------------------------------------------------------- */

/* One 1-D min-convolution with (t - len + off)^2 along a strided line: the
 * brute-force loop of regularisation.h:42-53 with its lowest-j tie rule. */
static void dt_line(const float* src, long long stride, float* dst, int* arg, const float* zt, int len) {
    for (int i = 0; i < len; i++) {
        float best = src[0] + zt[i + len];
        int bj = 0;
        for (int j = 0; j < len; j++) {
            float c = src[(long long)j * stride] + zt[i - j + len];
            if (c < best) { best = c; bj = j; }
        }
        dst[(long long)i * stride] = best;
        arg[i] = bj;
    }
}

/* regularisation.h:6-113: three separable passes (y = contiguous axis, then
 * x, then z) in label units with fractional offsets; writes the message in
 * place and the composite argmin label of each output label into inds. */
void orc_message_dt(float* cube, short* inds, int len, float offx, float offy, float offz) {
    int len2 = len * len, len3 = len2 * len;
    float* b1 = (float*)xmalloc(sizeof(float) * len3);
    float* b2 = (float*)xmalloc(sizeof(float) * len3);
    int* i1 = (int*)xmalloc(sizeof(int) * len3);
    int* i2 = (int*)xmalloc(sizeof(int) * len3);
    float zt[64 + 1];
    int arg[64];
    for (int t = 0; t < 2 * len + 1; t++) { float d = (float)(t - len) + offy; zt[t] = d * d; }
    for (int c = 0; c < len2; c++) {
        dt_line(cube + c * len, 1, b1 + c * len, arg, zt, len);
        for (int i = 0; i < len; i++) i1[c * len + i] = arg[i] + c * len;
    }
    for (int t = 0; t < 2 * len; t++) { float d = (float)(t - len) + offx; zt[t] = d * d; }
    for (int k1 = 0; k1 < len; k1++)
        for (int a = 0; a < len; a++) {
            long long base = a + (long long)k1 * len2;
            dt_line(b1 + base, len, b2 + base, arg, zt, len);
            for (int i = 0; i < len; i++) i2[base + i * len] = i1[base + arg[i] * len];
        }
    for (int t = 0; t < 2 * len; t++) { float d = (float)(t - len) + offz; zt[t] = d * d; }
    for (int c = 0; c < len2; c++) {
        dt_line(b2 + c, len2, cube + c, arg, zt, len);
        for (int i = 0; i < len; i++) inds[c + i * len2] = (short)i2[c + arg[i] * len2];
    }
    free(b1); free(b2); free(i1); free(i2);
}

/* regularisation.h:115-303: bottom-up messages per BFS depth (cost scaled by
 * the edge weight, min-convolved, then added to the parent minus its minimum,
 * children folded in `ordered` order), root argmin (first strict minimum),
 * top-down label selection; u1 = displacement of the label + u0. */
void orc_regularisation(float* cost, const float* u0, const float* v0, const float* w0,
                        float* u1, float* v1, float* w1, int hw, int step, float quant,
                        const int* ordered, const int* parents, const float* edgemst,
                        int m2, int n2, int o2, short* allinds_out, int* selected_out) {
    int m = m2 / step, n = n2 / step, o = o2 / step;
    int N = m * n * o;
    int len = 2 * hw + 1, len3 = len * len * len;
    short* allinds = allinds_out ? allinds_out : (short*)xmalloc(sizeof(short) * (size_t)N * len3);
    int* depth = (int*)calloc(N, sizeof(int));
    for (int t = 1; t < N; t++) depth[ordered[t]] = depth[parents[ordered[t]]] + 1;
    int maxd = 0;
    for (int v = 0; v < N; v++) if (depth[v] > maxd) maxd = depth[v];
    int* first = (int*)calloc(maxd + 2, sizeof(int)); /* first[d] = position of depth d in ordered */
    for (int v = 0; v < N; v++) first[depth[v] + 1]++;
    for (int d = 1; d <= maxd + 1; d++) first[d] += first[d - 1];

    for (int d = maxd; d > 0; d--) {
        int s = first[d], e = first[d + 1];
        for (int t = s; t < e; t++) {
            int c = ordered[t];
            for (int l = 0; l < len3; l++) cost[(long long)c * len3 + l] *= edgemst[c];
        }
        for (int t = s; t < e; t++) {
            int c = ordered[t], p = parents[c];
            float ox = (u0[p] - u0[c]) / quant;
            float oy = (v0[p] - v0[c]) / quant;
            float oz = (w0[p] - w0[c]) / quant;
            orc_message_dt(cost + (long long)c * len3, allinds + (long long)c * len3, len, ox, oy, oz);
        }
        for (int t = s; t < e; t++) {
            int c = ordered[t], p = parents[c];
            const float* msg = cost + (long long)c * len3;
            float mn = msg[0];
            for (int l = 1; l < len3; l++) if (msg[l] < mn) mn = msg[l];
            float* acc = cost + (long long)p * len3;
            for (int l = 0; l < len3; l++) acc[l] += (msg[l] - mn);
        }
    }
    int* sel = selected_out ? selected_out : (int*)xmalloc(sizeof(int) * N);
    int root = ordered[0];
    const float* rc = cost + (long long)root * len3;
    float best = rc[0];
    int bl = 0;
    for (int l = 0; l < len3; l++) {
        if (rc[l] < best) { best = rc[l]; bl = l; }
        allinds[(long long)root * len3 + l] = (short)l;
    }
    for (int t = 0; t < N; t++) {
        int c = ordered[t];
        int lab = (t == 0) ? bl : allinds[(long long)c * len3 + sel[parents[c]]];
        sel[c] = lab;
        int li = lab % len, lj = (lab / len) % len, lk = lab / (len * len);
        float dx = (float)(lj - hw) * quant, dy = (float)(li - hw) * quant, dz = (float)(lk - hw) * quant;
        u1[c] = dx + u0[c];
        v1[c] = dy + v0[c];
        w1[c] = dz + w0[c];
    }
    if (!allinds_out) free(allinds);
    if (!selected_out) free(sel);
    free(depth); free(first);
}

/* ------------------------------------------------------------------------- */
/* Transforms                                                                 */
/* ------------------------------------------------------------------------- */

/* Trilinear sample of `in` (m2 x n2 x o2, clamped) at float coordinates
 * (x = along j, y = along i, z = along k).  The weight products are the
 * reference's mixed float/double expression tree (transformations.h:13-20,
 * dataCostD.h:288-295): terms 1-4,6,7 promote to double at the first (1.0-d)
 * factor, term 5 multiplies dx*dy in float first, term 8 is float throughout;
 * the eight terms are summed left to right in double, then stored as float. */
static inline float trilinear(const float* in, float x1, float y1, float z1, int addj, int addi, int addk,
                              int m2, int n2, int o2) {
    int x = (int)floorf(x1), y = (int)floorf(y1), z = (int)floorf(z1);
    float dx = x1 - (float)x, dy = y1 - (float)y, dz = z1 - (float)z;
    x += addj; y += addi; z += addk;
    int y0 = clampi(y, 0, m2 - 1), y1c = clampi(y + 1, 0, m2 - 1);
    int x0 = clampi(x, 0, n2 - 1), x1c = clampi(x + 1, 0, n2 - 1);
    int z0 = clampi(z, 0, o2 - 1), z1c = clampi(z + 1, 0, o2 - 1);
    double ax = 1.0 - dx, ay = 1.0 - dy, az = 1.0 - dz;
    double s = ax * ay * az * (double)in[IDX(y0, x0, z0, m2, n2)];
    s = s + ax * (double)dy * az * (double)in[IDX(y1c, x0, z0, m2, n2)];
    s = s + (double)dx * ay * az * (double)in[IDX(y0, x1c, z0, m2, n2)];
    s = s + ax * ay * (double)dz * (double)in[IDX(y0, x0, z1c, m2, n2)];
    s = s + (double)(dx * dy) * az * (double)in[IDX(y1c, x1c, z0, m2, n2)];
    s = s + ax * (double)dy * (double)dz * (double)in[IDX(y1c, x0, z1c, m2, n2)];
    s = s + (double)dx * ay * (double)dz * (double)in[IDX(y0, x1c, z1c, m2, n2)];
    s = s + (double)(dx * dy * dz * in[IDX(y1c, x1c, z1c, m2, n2)]);
    return (float)s;
}

/* transformations.h:5-26 */
void orc_interp3(float* out, const float* in, const float* x1, const float* y1, const float* z1,
                 int m, int n, int o, int m2, int n2, int o2, int add_identity) {
    for (int k = 0; k < o; k++)
        for (int j = 0; j < n; j++)
            for (int i = 0; i < m; i++) {
                long long p = IDX(i, j, k, m, n);
                out[p] = add_identity ? trilinear(in, x1[p], y1[p], z1[p], j, i, k, m2, n2, o2)
                                      : trilinear(in, x1[p], y1[p], z1[p], 0, 0, 0, m2, n2, o2);
            }
}

/* transformations.h:245-279: sample the coarse field at (j/sn, i/sm, k/so). */
void orc_upsample(float* u1, float* v1, float* w1, const float* u0, const float* v0, const float* w0,
                  int m, int n, int o, int m2, int n2, int o2) {
    float sm = (float)m / (float)m2, sn = (float)n / (float)n2, so = (float)o / (float)o2;
    for (int k = 0; k < o; k++)
        for (int j = 0; j < n; j++)
            for (int i = 0; i < m; i++) {
                long long p = IDX(i, j, k, m, n);
                float x = (float)j / sn, y = (float)i / sm, z = (float)k / so;
                u1[p] = trilinear(u0, x, y, z, 0, 0, 0, m2, n2, o2);
                v1[p] = trilinear(v0, x, y, z, 0, 0, 0, m2, n2, o2);
                w1[p] = trilinear(w0, x, y, z, 0, 0, 0, m2, n2, o2);
            }
}

/* transformations.h:189-242: fields scaled by 1/factor, ten Jacobi
 * iterations u <- 0.5 u - 0.5 u2(x+u) and u2 <- 0.5 u2 - 0.5 u(x+u2) from the
 * previous iterate (double arithmetic, stored as float), scaled back. */
void orc_consistent(float* u, float* v, float* w, float* u2, float* v2, float* w2,
                    int m, int n, int o, int factor) {
    long long sz = (long long)m * n * o;
    float f1 = (float)(1.0 / (float)factor);
    float* s = (float*)xmalloc(sizeof(float) * sz * 6);
    float *us = s, *vs = s + sz, *ws = s + 2 * sz, *us2 = s + 3 * sz, *vs2 = s + 4 * sz, *ws2 = s + 5 * sz;
    for (long long p = 0; p < sz; p++) {
        us[p] = u[p] * f1; vs[p] = v[p] * f1; ws[p] = w[p] * f1;
        us2[p] = u2[p] * f1; vs2[p] = v2[p] * f1; ws2[p] = w2[p] * f1;
    }
    for (int it = 0; it < 10; it++) {
        orc_interp3(u, us2, us, vs, ws, m, n, o, m, n, o, 1);
        orc_interp3(v, vs2, us, vs, ws, m, n, o, m, n, o, 1);
        orc_interp3(w, ws2, us, vs, ws, m, n, o, m, n, o, 1);
        for (long long p = 0; p < sz; p++) {
            u[p] = (float)(0.5 * us[p] - 0.5 * u[p]);
            v[p] = (float)(0.5 * vs[p] - 0.5 * v[p]);
            w[p] = (float)(0.5 * ws[p] - 0.5 * w[p]);
        }
        orc_interp3(u2, us, us2, vs2, ws2, m, n, o, m, n, o, 1);
        orc_interp3(v2, vs, us2, vs2, ws2, m, n, o, m, n, o, 1);
        orc_interp3(w2, ws, us2, vs2, ws2, m, n, o, m, n, o, 1);
        for (long long p = 0; p < sz; p++) {
            u2[p] = (float)(0.5 * us2[p] - 0.5 * u2[p]);
            v2[p] = (float)(0.5 * vs2[p] - 0.5 * v2[p]);
            w2[p] = (float)(0.5 * ws2[p] - 0.5 * w2[p]);
        }
        memcpy(us, u, sizeof(float) * sz); memcpy(vs, v, sizeof(float) * sz); memcpy(ws, w, sizeof(float) * sz);
        memcpy(us2, u2, sizeof(float) * sz); memcpy(vs2, v2, sizeof(float) * sz); memcpy(ws2, w2, sizeof(float) * sz);
    }
    float ff = (float)factor;
    for (long long p = 0; p < sz; p++) {
        u[p] *= ff; v[p] *= ff; w[p] *= ff;
        u2[p] *= ff; v2[p] *= ff; w2[p] *= ff;
    }
    free(s);
}

/* transformations.h:30-54 with the central-difference kernel {-0.5,0,0.5}. */
static void grad_filter(const float* in, float* out, int m, int n, int o, int dim) {
    static const float g[3] = {-0.5f, 0.0f, 0.5f};
    for (int k = 0; k < o; k++)
        for (int j = 0; j < n; j++)
            for (int i = 0; i < m; i++) {
                float acc = 0.0f;
                for (int f = 0; f < 3; f++) {
                    long long p;
                    if (dim == 1) p = IDX(clampi(i + f - 1, 0, m - 1), j, k, m, n);
                    else if (dim == 2) p = IDX(i, clampi(j + f - 1, 0, n - 1), k, m, n);
                    else p = IDX(i, j, clampi(k + f - 1, 0, o - 1), m, n);
                    acc = acc + g[f] * in[p];
                }
                out[IDX(i, j, k, m, n)] = acc;
            }
}

/* transformations.h:84-185 (diagnostic): std of the Jacobian determinant of
 * x + u/factor, and the fraction of negative determinants. */
float orc_jacobian(const float* u, const float* v, const float* w, int m, int n, int o, int factor,
                   float* neg_frac) {
    long long sz = (long long)m * n * o;
    float f1 = (float)(1.0 / (float)factor);
    float* J = (float*)xmalloc(sizeof(float) * sz * 10);
    float* Jd = J + 9 * sz;
    const float* fld[3] = {u, v, w};
    static const int dims[3] = {2, 1, 3};
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) grad_filter(fld[a], J + (a * 3 + b) * sz, m, n, o, dims[b]);
    for (long long p = 0; p < 9 * sz; p++) J[p] *= f1;
    for (long long p = 0; p < sz; p++) {
        J[0 * sz + p] = (float)((double)J[0 * sz + p] + 1.0);
        J[4 * sz + p] = (float)((double)J[4 * sz + p] + 1.0);
        J[8 * sz + p] = (float)((double)J[8 * sz + p] + 1.0);
    }
    float mean = 0.0f, neg = 0.0f, count = 0.0f;
    for (long long p = 0; p < sz; p++) {
        float J11 = J[p], J12 = J[sz + p], J13 = J[2 * sz + p];
        float J21 = J[3 * sz + p], J22 = J[4 * sz + p], J23 = J[5 * sz + p];
        float J31 = J[6 * sz + p], J32 = J[7 * sz + p], J33 = J[8 * sz + p];
        float d = J11 * (J22 * J33 - J23 * J32) - J21 * (J12 * J33 - J13 * J32) + J31 * (J12 * J23 - J13 * J22);
        mean += d;
        if (d < 0) neg++;
        count++;
        Jd[p] = d;
    }
    mean /= (float)sz;
    float sd = 0.0f;
    for (long long p = 0; p < sz; p++) {
        double e = (double)(Jd[p] - mean);
        sd = (float)((double)sd + e * e);
    }
    sd /= (float)(sz - 1);
    sd = sqrtf(sd);
    if (neg_frac) *neg_frac = neg / count;
    free(J);
    return sd;
}

/* dataCostD.h:300-312 / :225-237: logging-only mean squared differences,
 * accumulated as float += double in (i, j, k) loop order. */
static void ssd_log(const float* ref, const float* warped, const float* before, int m, int n, int o, float* ssd01) {
    if (!ssd01) return;
    float s = 0.0f, s0 = 0.0f;
    for (int i = 0; i < m; i++)
        for (int j = 0; j < n; j++)
            for (int k = 0; k < o; k++) {
                long long p = IDX(i, j, k, m, n);
                double a = (double)(ref[p] - warped[p]);
                double b = (double)(ref[p] - before[p]);
                s = (float)((double)s + a * a);
                s0 = (float)((double)s0 + b * b);
            }
    ssd01[0] = s0 / (float)(m * n * o);
    ssd01[1] = s / (float)(m * n * o);
}

/* float affine coordinate of dataCostD.h:285-287, summed left to right */
static inline float affine_coord(const float* r, int i, int j, int k, float disp) {
    return (float)i * r[0] + (float)j * r[1] + (float)k * r[2] + r[3] + disp;
}

/* dataCostD.h:267-315 */
void orc_warp_affine(float* out, const float* in, const float* im1b, const float* X,
                     const float* u, const float* v, const float* w, int m, int n, int o, float* ssd01) {
    for (int k = 0; k < o; k++)
        for (int j = 0; j < n; j++)
            for (int i = 0; i < m; i++) {
                long long p = IDX(i, j, k, m, n);
                float y1 = affine_coord(X, i, j, k, v[p]);
                float x1 = affine_coord(X + 4, i, j, k, u[p]);
                float z1 = affine_coord(X + 8, i, j, k, w[p]);
                out[p] = trilinear(in, x1, y1, z1, 0, 0, 0, m, n, o);
            }
    ssd_log(im1b, out, in, m, n, o, ssd01);
}

/* dataCostD.h:213-239 */
void orc_warp_image(float* out, const float* im1, const float* im1b,
                    const float* u, const float* v, const float* w, int m, int n, int o, float* ssd01) {
    orc_interp3(out, im1, u, v, w, m, n, o, m, n, o, 1);
    ssd_log(im1b, out, im1, m, n, o, ssd01);
}

/* dataCostD.h:241-266: nearest neighbour (round half away from zero), clamped. */
void orc_warp_affine_s(short* out, const short* in, const float* X,
                       const float* u, const float* v, const float* w, int m, int n, int o) {
    for (int k = 0; k < o; k++)
        for (int j = 0; j < n; j++)
            for (int i = 0; i < m; i++) {
                long long p = IDX(i, j, k, m, n);
                int y = (int)roundf(affine_coord(X, i, j, k, v[p]));
                int x = (int)roundf(affine_coord(X + 4, i, j, k, u[p]));
                int z = (int)roundf(affine_coord(X + 8, i, j, k, w[p]));
                out[p] = in[IDX(clampi(y, 0, m - 1), clampi(x, 0, n - 1), clampi(z, 0, o - 1), m, n)];
            }
}

/* ------------------------------------------------------------------------- */
/* The symmetric level loop                                                   */
/* ------------------------------------------------------------------------- */

/* deedsBCV0.cpp:139-349 with RAND_SAMPLES = 1. */
int orc_register(const float* fixed, const float* moving, const float* X, int m, int n, int o,
                 int levels, const int* grid, const int* radius, const int* quant, float alpha,
                 float* flow_out, float* warped_out, int verbose) {
    long long sz = (long long)m * n * o;
    float* im1b = (float*)xmalloc(sizeof(float) * sz);
    float* im1 = (float*)xmalloc(sizeof(float) * sz);
    for (long long p = 0; p < sz; p++) { im1b[p] = fixed[p] - (-1024.0f); im1[p] = moving[p] - (-1024.0f); }
    int mind_step[16];
    for (int l = 0; l < levels; l++) mind_step[l] = (int)floorf(0.5f * (float)quant[l] + 1.0f);
    float* ux = (float*)calloc(sz, sizeof(float));
    float* vx = (float*)calloc(sz, sizeof(float));
    float* wx = (float*)calloc(sz, sizeof(float));
    float* warped0 = (float*)xmalloc(sizeof(float) * sz);
    float* warped1 = (float*)xmalloc(sizeof(float) * sz);
    uint64_t* mind_mov = (uint64_t*)xmalloc(sizeof(uint64_t) * sz);
    uint64_t* mind_fix = (uint64_t*)xmalloc(sizeof(uint64_t) * sz);
    uint64_t* mind_w = (uint64_t*)xmalloc(sizeof(uint64_t) * sz);
    int m2 = m / grid[0], n2 = n / grid[0], o2 = o / grid[0];
    int sz2 = m2 * n2 * o2;
    float* f = (float*)calloc(6 * (size_t)sz2, sizeof(float)); /* u1,v1,w1,u1i,v1i,w1i */
    float ssd[2];
    orc_warp_affine(warped0, im1, im1b, X, ux, vx, wx, m, n, o, ssd);
    int m1 = m2, n1 = n2, o1 = o2, sz1 = sz2;
    for (int level = 0; level < levels; level++) {
        float q1 = (float)quant[level];
        int prev = mind_step[level > 0 ? level - 1 : 0], curr = mind_step[level];
        if (level == 0 || prev != curr) {
            orc_descriptor(mind_mov, warped0, m, n, o, curr);
            orc_descriptor(mind_fix, im1b, m, n, o, curr);
        }
        int step = grid[level], hw = radius[level];
        int len = 2 * hw + 1, len3 = len * len * len;
        m1 = m / step; n1 = n / step; o1 = o / step; sz1 = m1 * n1 * o1;
        float* cost = (float*)xmalloc(sizeof(float) * (size_t)sz1 * len3);
        float* g0 = (float*)xmalloc(sizeof(float) * 3 * (size_t)sz1);
        float* nf = (float*)xmalloc(sizeof(float) * 6 * (size_t)sz1);
        int* ordered = (int*)xmalloc(sizeof(int) * sz1);
        int* parents = (int*)xmalloc(sizeof(int) * sz1);
        float* edgemst = (float*)xmalloc(sizeof(float) * sz1);
        if (verbose) printf("oracle level %d grid=%d %dx%dx%d hw=%d quant=%d\n", level, step, m1, n1, o1, hw, quant[level]);
        for (int dir = 0; dir < 2; dir++) {
            const float* prevf = f + (dir ? 3 * (size_t)sz2 : 0);
            orc_upsample(g0, g0 + sz1, g0 + 2 * (size_t)sz1, prevf, prevf + sz2, prevf + 2 * (size_t)sz2,
                         m1, n1, o1, m2, n2, o2);
            orc_upsample(ux, vx, wx, g0, g0 + sz1, g0 + 2 * (size_t)sz1, m, n, o, m1, n1, o1);
            if (dir == 0) orc_warp_affine(warped1, im1, im1b, X, ux, vx, wx, m, n, o, ssd);
            else orc_warp_image(warped1, im1b, warped0, ux, vx, wx, m, n, o, ssd);
            orc_descriptor(mind_w, warped1, m, n, o, curr);
            orc_datacost(dir ? mind_mov : mind_fix, mind_w, cost, m, n, o, step, hw, q1, alpha, 1);
            orc_prims(dir ? warped0 : im1b, ordered, parents, edgemst, step, m, n, o);
            float* out = nf + (dir ? 3 * (size_t)sz1 : 0);
            orc_regularisation(cost, g0, g0 + sz1, g0 + 2 * (size_t)sz1, out, out + sz1, out + 2 * (size_t)sz1,
                               hw, step, q1, ordered, parents, edgemst, m, n, o, NULL, NULL);
        }
        orc_consistent(nf, nf + sz1, nf + 2 * (size_t)sz1, nf + 3 * (size_t)sz1, nf + 4 * (size_t)sz1,
                       nf + 5 * (size_t)sz1, m1, n1, o1, step);
        free(f);
        f = nf;
        m2 = m1; n2 = n1; o2 = o1; sz2 = sz1;
        free(cost); free(g0); free(ordered); free(parents); free(edgemst);
    }
    orc_upsample(ux, vx, wx, f, f + sz1, f + 2 * (size_t)sz1, m, n, o, m1, n1, o1);
    if (flow_out) memcpy(flow_out, f, sizeof(float) * 3 * (size_t)sz1);
    if (warped_out) {
        orc_warp_affine(warped1, im1, im1b, X, ux, vx, wx, m, n, o, ssd);
        for (long long p = 0; p < sz; p++) warped_out[p] = warped1[p] + (-1024.0f);
    }
    free(im1b); free(im1); free(ux); free(vx); free(wx); free(warped0); free(warped1);
    free(mind_mov); free(mind_fix); free(mind_w); free(f);
    return sz1;
}

/* ------------------------------------------------------------------------- */
/* Evaluation                                                                 */
/* ------------------------------------------------------------------------- */

/* diceMulti.cpp:38-55: the distinct values of seg1 in increasing order
 * (sort + unique, :38-40), all but the smallest (:44), each with
 * 2*|A & B| / (|A| + |B|) from counts converted to float (:46-55). */
int orc_dice(const short* seg1, const short* seg2, long long sz, int capacity, short* labels, float* dice) {
    long long* c1 = (long long*)calloc(65536, sizeof(long long));
    long long* c2 = (long long*)calloc(65536, sizeof(long long));
    long long* both = (long long*)calloc(65536, sizeof(long long));
    if (!c1 || !c2 || !both) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
    for (long long p = 0; p < sz; p++) {
        int a = seg1[p] + 32768, b = seg2[p] + 32768;
        c1[a]++;
        c2[b]++;
        if (a == b) both[a]++;
    }
    int total = 0, seen = 0;
    for (int u = 0; u < 65536; u++) {
        if (!c1[u]) continue;
        if (seen++ == 0) continue; /* the smallest label (background) */
        if (total < capacity) {
            float unionval = (float)both[u], card1 = (float)c1[u], card2 = (float)c2[u];
            labels[total] = (short)(u - 32768);
            dice[total] = 2.0f * unionval / (card1 + card2);
        }
        total++;
    }
    free(c1); free(c2); free(both);
    return total;
}

/* ------------------------------------------------------------------------- */
/* Affine pre-alignment (linearBCV)                                           */
/* ------------------------------------------------------------------------- */

/* affineLTS1.h:39-75: trilinear warp of `in` by the affine X (no field), zero
 * where the 2x2x2 neighbourhood leaves the volume; the weights are double,
 * except dx*dy (terms 5, 8) and dx*dy*dz*in (term 8), which are float. */
void orc_lin_warp(float* out, const float* in, const float* X, int m, int n, int o) {
    for (int k = 0; k < o; k++)
        for (int j = 0; j < n; j++)
            for (int i = 0; i < m; i++) {
                float y1 = (float)i * X[0] + (float)j * X[1] + (float)k * X[2] + X[3];
                float x1 = (float)i * X[4] + (float)j * X[5] + (float)k * X[6] + X[7];
                float z1 = (float)i * X[8] + (float)j * X[9] + (float)k * X[10] + X[11];
                int x = (int)floorf(x1), y = (int)floorf(y1), z = (int)floorf(z1);
                float dx = x1 - (float)x, dy = y1 - (float)y, dz = z1 - (float)z;
                long long p = IDX(i, j, k, m, n);
                if (y < 0 || y >= m - 1 || x < 0 || x >= n - 1 || z < 0 || z >= o - 1) {
                    out[p] = 0.0f;
                    continue;
                }
                double ddx = dx, ddy = dy, ddz = dz;
                double s = (1.0 - ddx) * (1.0 - ddy) * (1.0 - ddz) * in[IDX(y, x, z, m, n)];
                s += (1.0 - ddx) * ddy * (1.0 - ddz) * in[IDX(y + 1, x, z, m, n)];
                s += ddx * (1.0 - ddy) * (1.0 - ddz) * in[IDX(y, x + 1, z, m, n)];
                s += (1.0 - ddx) * (1.0 - ddy) * ddz * in[IDX(y, x, z + 1, m, n)];
                s += (double)(dx * dy) * (1.0 - ddz) * in[IDX(y + 1, x + 1, z, m, n)];
                s += (1.0 - ddx) * ddy * ddz * in[IDX(y + 1, x, z + 1, m, n)];
                s += ddx * (1.0 - ddy) * ddz * in[IDX(y, x + 1, z + 1, m, n)];
                s += (double)(dx * dy * dz * in[IDX(y + 1, x + 1, z + 1, m, n)]);
                out[p] = (float)s;
            }
}

/* QRsolve.h:1-7: float accumulation of the exact double squares (pow(float,2)). */
static float lin_norm(const float* v, int len) {
    float s = 0;
    for (int i = 0; i < len; i++) s = (float)((double)s + (double)v[i] * (double)v[i]);
    return sqrtf(s);
}

/* QRsolve.h:9-15 */
static float lin_dot(const float* a, const float* b, int len) {
    float d = 0;
    for (int i = 0; i < len; i++) d = d + a[i] * b[i];
    return d;
}

/* QRsolve.h:17-70: Gram-Schmidt QR of the len x 4 matrix A (columns A + c*len)
 * and back substitution for the len2 right-hand sides b + j*len -> X + j*4. */
void orc_qrsolve(float* X, const float* A, const float* b, int len, int len2) {
    float* Q = (float*)xmalloc(sizeof(float) * 4 * (size_t)len);
    float r11 = lin_norm(A, len);
    for (int i = 0; i < len; i++) Q[i] = A[i] / r11;
    float r12 = lin_dot(Q, A + len, len), r13 = lin_dot(Q, A + 2 * len, len), r14 = lin_dot(Q, A + 3 * len, len);
    for (int i = 0; i < len; i++) {
        Q[i + len] = A[i + len] - Q[i] * r12;
        Q[i + 2 * len] = A[i + 2 * len] - Q[i] * r13;
        Q[i + 3 * len] = A[i + 3 * len] - Q[i] * r14;
    }
    float r22 = lin_norm(Q + len, len);
    for (int i = 0; i < len; i++) Q[i + len] = Q[i + len] / r22;
    float r23 = lin_dot(Q + len, A + 2 * len, len), r24 = lin_dot(Q + len, A + 3 * len, len);
    for (int i = 0; i < len; i++) {
        Q[i + 2 * len] = A[i + 2 * len] - Q[i] * r13 - Q[i + len] * r23;
        Q[i + 3 * len] = A[i + 3 * len] - Q[i] * r14 - Q[i + len] * r24;
    }
    float r33 = lin_norm(Q + 2 * len, len);
    for (int i = 0; i < len; i++) Q[i + 2 * len] = Q[i + 2 * len] / r33;
    float r34 = lin_dot(Q + 2 * len, A + 3 * len, len);
    for (int i = 0; i < len; i++) Q[i + 3 * len] = A[i + 3 * len] - Q[i] * r14 - Q[i + len] * r24 - Q[i + 2 * len] * r34;
    float r44 = lin_norm(Q + 3 * len, len);
    for (int i = 0; i < len; i++) Q[i + 3 * len] = Q[i + 3 * len] / r44;
    for (int j = 0; j < len2; j++) {
        const float* bj = b + (size_t)j * len;
        float d1 = lin_dot(Q, bj, len), d2 = lin_dot(Q + len, bj, len);
        float d3 = lin_dot(Q + 2 * len, bj, len), d4 = lin_dot(Q + 3 * len, bj, len);
        float x4 = d4 / r44;
        float x3 = (d3 - r34 * x4) / r33;
        float x2 = (d2 - r23 * x3 - r24 * x4) / r22;
        float x1 = (d1 - r12 * x2 - r13 * x3 - r14 * x4) / r11;
        X[0 + j * 4] = x1; X[1 + j * 4] = x2; X[2 + j * 4] = x3; X[3 + j * 4] = x4;
    }
    free(Q);
}

/* QRsolve.h:72-122: four sweeps of one-sided Jacobi rotations on the columns
 * of the 3x3 A (U <- A rotated, V <- the rotations), then unit columns of U. */
static void lin_jacobi_svd3(const float* A, float* U, float* V) {
    for (int i = 0; i < 9; i++) { U[i] = A[i]; V[i] = 0; }
    V[0] = 1; V[4] = 1; V[8] = 1;
    float T[3];
    for (int iter = 0; iter < 4; iter++)
        for (int j = 1; j < 3; j++)
            for (int i = 0; i < j; i++) {
                float alpha = lin_dot(U + i * 3, U + i * 3, 3);
                float beta = lin_dot(U + j * 3, U + j * 3, 3);
                float gamma = lin_dot(U + i * 3, U + j * 3, 3);
                float zeta = (float)((double)(beta - alpha) / (2.0 * (double)gamma));
                float t = (float)((zeta > 0 ? 1.0 : -1.0) / ((double)fabsf(zeta) + sqrt(1.0 + (double)(zeta * zeta))));
                float c = (float)(1.0 / (double)sqrtf(1.0f + t * t));
                float s = c * t;
                for (int k = 0; k < 3; k++) T[k] = U[k + i * 3];
                for (int k = 0; k < 3; k++) U[k + i * 3] = c * T[k] - s * U[k + j * 3];
                for (int k = 0; k < 3; k++) U[k + j * 3] = s * T[k] + c * U[k + j * 3];
                for (int k = 0; k < 3; k++) T[k] = V[k + i * 3];
                for (int k = 0; k < 3; k++) V[k + i * 3] = c * T[k] - s * V[k + j * 3];
                for (int k = 0; k < 3; k++) V[k + j * 3] = s * T[k] + c * V[k + j * 3];
            }
    for (int j = 0; j < 3; j++) {
        float sv = lin_norm(U + j * 3, 3);
        for (int i = 0; i < 3; i++) U[i + j * 3] /= sv;
    }
}

/* QRsolve.h:124-254: rigid fit (rotation from the SVD of the centred
 * cross-covariance, translation from the mean residual). */
void orc_find_rigid(float* RT, const float* pts1t, const float* pts2t, int len) {
    float* p1 = (float*)xmalloc(sizeof(float) * 4 * (size_t)len);
    float* p2 = (float*)xmalloc(sizeof(float) * 4 * (size_t)len);
    for (int i = 0; i < 4; i++)
        for (int k = 0; k < len; k++) {
            p1[i + k * 4] = pts1t[k + (size_t)i * len];
            p2[i + k * 4] = pts2t[k + (size_t)i * len];
        }
    float m1[3] = {0, 0, 0}, m2[3] = {0, 0, 0};
    for (int k = 0; k < len; k++)
        for (int i = 0; i < 3; i++) { m1[i] += p1[i + k * 4]; m2[i] += p2[i + k * 4]; }
    for (int i = 0; i < 3; i++) { m1[i] /= (float)len; m2[i] /= (float)len; }
    float K[9] = {0}, U[9], V[9], R[16];
    for (int k = 0; k < len; k++)
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) K[i + j * 3] += (p1[j + k * 4] - m1[j]) * (p2[i + k * 4] - m2[i]);
    lin_jacobi_svd3(K, U, V);
    for (int i = 0; i < 16; i++) { R[i] = 0.0f; RT[i] = 0.0f; }
    R[15] = 1.0f;
    RT[15] = 1.0f;
    for (int k = 0; k < 3; k++)
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) R[j + i * 4] += U[j + k * 3] * V[i + k * 3];
    float* Rp2 = (float*)xmalloc(sizeof(float) * 4 * (size_t)len);
    orc_qrsolve(Rp2, R, p2, 4, len);
    float t1[3] = {0, 0, 0};
    for (int i = 0; i < len; i++) { t1[0] += Rp2[i * 4]; t1[1] += Rp2[1 + i * 4]; t1[2] += Rp2[2 + i * 4]; }
    for (int i = 0; i < 3; i++) t1[i] = t1[i] / (float)len - m1[i];
    float t0[3] = {0, 0, 0};
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) t0[i] += R[i + j * 4] * t1[j];
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++) RT[i + j * 4] += R[j + i * 4];
        RT[3 + i * 4] = t0[i];
    }
    free(p1); free(p2); free(Rp2);
}

static int lin_cmp_float(const void* a, const void* b) {
    float x = *(const float*)a, y = *(const float*)b;
    return x < y ? -1 : (x > y ? 1 : 0);
}

/* the k-th smallest of v[0..len) (the value std::nth_element places at k) */
static float lin_kth(const float* v, int len, int k) {
    float* c = (float*)xmalloc(sizeof(float) * (size_t)len);
    memcpy(c, v, sizeof(float) * (size_t)len);
    qsort(c, (size_t)len, sizeof(float), lin_cmp_float);
    float r = c[k];
    free(c);
    return r;
}

/* QRsolve.h:258-296: least trimmed squares -- fit to all points, then 15
 * refits to the half of the points whose residual is at most the median. */
void orc_affine_robust(float* RT, const float* pts1, const float* pts2, int len, int rigid) {
    if (rigid) orc_find_rigid(RT, pts1, pts2, len);
    else orc_qrsolve(RT, pts1, pts2, len, 4);
    int lenh = len / 2;
    float* err = (float*)xmalloc(sizeof(float) * (size_t)len);
    float* b1 = (float*)xmalloc(sizeof(float) * 4 * (size_t)(lenh > 0 ? lenh : 1));
    float* b2 = (float*)xmalloc(sizeof(float) * 4 * (size_t)(lenh > 0 ? lenh : 1));
    for (int i = 0; i < 4 * lenh; i++) { b1[i] = 1; b2[i] = 1; }
    for (int iter = 0; iter < 15; iter++) {
        for (int l = 0; l < len; l++) {
            float x = 0, y = 0, z = 0;
            for (int i = 0; i < 4; i++) {
                y += pts1[l + i * len] * RT[i];
                x += pts1[l + i * len] * RT[i + 4];
                z += pts1[l + i * len] * RT[i + 8];
            }
            double ey = y - pts2[l], ex = x - pts2[l + len], ez = z - pts2[l + 2 * len];
            err[l] = (float)(ey * ey + ex * ex + ez * ez);
        }
        float median = lin_kth(err, len, len / 2);
        int count = 0;
        for (int i = 0; i < len; i++) {
            int l = count < lenh - 1 ? count : lenh - 1;
            if (err[i] <= median) {
                for (int d = 0; d < 3; d++) {
                    b1[l + d * lenh] = pts1[i + d * len];
                    b2[l + d * lenh] = pts2[i + d * len];
                }
                count++;
            }
        }
        if (rigid) orc_find_rigid(RT, b1, b2, lenh);
        else orc_qrsolve(RT, b1, b2, lenh, 4);
    }
    free(err); free(b1); free(b2);
}

/* affineLTS1.h:1-12: C = A * B (4x4, column-major, float accumulation). */
static void lin_matmult(const float* A, const float* B, float* C) {
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            C[i + j * 4] = 0.0f;
            for (int k = 0; k < 4; k++) C[i + j * 4] += A[i + k * 4] * B[k + j * 4];
        }
}

/* affineLTS1.h:122-148: in-place smoothing along the node axis with index
 * stride `st` (1: y, m1: x, m1*n1: z): c[p] += 0.5*c[p+1] + 0.5*c[p-1] in
 * double, nodes visited in index order (p-1 already smoothed), clamped. */
static void lin_smooth(float* c, int len3, int m1, int n1, int o1, int axis) {
    int sz1 = m1 * n1 * o1;
    for (int i = 0; i < sz1; i++) {
        int z1 = i / (m1 * n1), x1 = (i - z1 * m1 * n1) / m1, y1 = i - z1 * m1 * n1 - x1 * m1;
        int i1, i2;
        if (axis == 0) { i1 = (y1 + 1 < m1 - 1 ? y1 + 1 : m1 - 1) + x1 * m1 + z1 * m1 * n1; i2 = (y1 - 1 > 0 ? y1 - 1 : 0) + x1 * m1 + z1 * m1 * n1; }
        else if (axis == 1) { i1 = y1 + (x1 + 1 < n1 - 1 ? x1 + 1 : n1 - 1) * m1 + z1 * m1 * n1; i2 = y1 + (x1 - 1 > 0 ? x1 - 1 : 0) * m1 + z1 * m1 * n1; }
        else { i1 = y1 + x1 * m1 + (z1 + 1 < o1 - 1 ? z1 + 1 : o1 - 1) * m1 * n1; i2 = y1 + x1 * m1 + (z1 - 1 > 0 ? z1 - 1 : 0) * m1 * n1; }
        for (int l = 0; l < len3; l++) {
            double add = 0.5 * (double)c[(size_t)i1 * len3 + l] + 0.5 * (double)c[(size_t)i2 * len3 + l];
            c[(size_t)i * len3 + l] = (float)((double)c[(size_t)i * len3 + l] + add);
        }
    }
}

/* affineLTS1.h:81-236 (estimateAffine2): smooth both cost volumes, take each
 * node's first minimum, drop background (block mean < 5) and border nodes,
 * keep the nodes whose minimum is below the quantile, and fit the affine (or
 * rigid) map between node centres and their best displacements robustly;
 * X = fit * Xprev.  cost / cost2 are smoothed in place. */
void orc_estimate_affine(float* X, const float* Xprev, const float* im1, const float* im2, float* cost, float* cost2,
                         int m, int n, int o, int step, float quant, int hw, int rigid, int* used) {
    int m1 = m / step, n1 = n / step, o1 = o / step, sz1 = m1 * n1 * o1;
    int len = 2 * hw + 1, len3 = len * len * len, stephw = (step - 1) / 2;
    for (int axis = 0; axis < 3; axis++) {
        lin_smooth(cost, len3, m1, n1, o1, axis);
        lin_smooth(cost2, len3, m1, n1, o1, axis);
    }
    float* mv1 = (float*)xmalloc(sizeof(float) * (size_t)sz1);
    float* mv2 = (float*)xmalloc(sizeof(float) * (size_t)sz1);
    int* mi1 = (int*)xmalloc(sizeof(int) * (size_t)sz1);
    int* mi2 = (int*)xmalloc(sizeof(int) * (size_t)sz1);
    int allcount = 0;
    for (int i = 0; i < sz1; i++) {
        const float* c1 = cost + (size_t)i * len3;
        const float* c2 = cost2 + (size_t)i * len3;
        int b1 = 0, b2 = 0;
        for (int l = 1; l < len3; l++) {
            if (c1[l] < c1[b1]) b1 = l;
            if (c2[l] < c2[b2]) b2 = l;
        }
        mv1[i] = c1[b1]; mi1[i] = b1;
        mv2[i] = c2[b2]; mi2[i] = b2;
        int z1 = i / (m1 * n1), x1 = (i - z1 * m1 * n1) / m1, y1 = i - z1 * m1 * n1 - x1 * m1;
        float mean1 = 0.0f, mean2 = 0.0f;
        for (int k1 = 0; k1 < step; k1++)
            for (int j1 = 0; j1 < step; j1++)
                for (int i1 = 0; i1 < step; i1++) {
                    long long p = IDX(y1 * step + i1, x1 * step + j1, z1 * step + k1, m, n);
                    mean1 += im1[p];
                    mean2 += im2[p];
                }
        mean1 /= (float)(step * step * step);
        mean2 /= (float)(step * step * step);
        if (mean1 < 5.0f) mv1[i] = 1e20f;
        if (mean2 < 5.0f) mv2[i] = 1e20f;
        if (z1 < 2 || z1 > o1 - 2 || x1 < 2 || x1 > n1 - 2 || y1 < 2 || y1 > m1 - 2) { mv1[i] = 1e20f; mv2[i] = 1e20f; }
        if ((double)mv1[i] < 1e19) allcount++;
        if ((double)mv2[i] < 1e19) allcount++;
    }
    /* quantile(values, sz1, 0.5*allcount/(float)sz1/2.0f), linearBCV.cpp:36-46 */
    float q = (float)(0.5 * allcount / (double)(float)sz1 / 2.0);
    float qc = q < 0.01f ? 0.01f : (q > 0.99f ? 0.99f : q);
    int qi = (int)((float)sz1 * qc);
    float med1 = lin_kth(mv1, sz1, qi), med2 = lin_kth(mv2, sz1, qi);
    int valsz = 0;
    for (int i = 0; i < sz1; i++) valsz += (mv1[i] < med1) + (mv2[i] < med2);
    if (used) *used = valsz;
    float* p1 = (float*)xmalloc(sizeof(float) * 4 * (size_t)(valsz > 0 ? valsz : 1));
    float* p2 = (float*)xmalloc(sizeof(float) * 4 * (size_t)(valsz > 0 ? valsz : 1));
    int cur = 0;
    for (int i = 0; i < sz1; i++) {
        int k1 = i / (m1 * n1), j1 = (i - k1 * m1 * n1) / m1, i1 = i - j1 * m1 - k1 * m1 * n1;
        int y1 = i1 * step + stephw, x1 = j1 * step + stephw, z1 = k1 * step + stephw;
        for (int dir = 0; dir < 2; dir++) {
            if (!(dir ? mv2[i] < med2 : mv1[i] < med1)) continue;
            int ind = dir ? mi2[i] : mi1[i];
            int li = ind % len, lj = (ind / len) % len, lk = ind / (len * len);
            float ys = (float)(li - hw) * quant, xs = (float)(lj - hw) * quant, zs = (float)(lk - hw) * quant;
            float* a = dir ? p2 : p1; /* the node centre */
            float* b = dir ? p1 : p2; /* the centre displaced by the best label */
            a[cur] = (float)y1; a[cur + valsz] = (float)x1; a[cur + 2 * valsz] = (float)z1;
            b[cur] = (float)y1 + ys; b[cur + valsz] = (float)x1 + xs; b[cur + 2 * valsz] = (float)z1 + zs;
            p1[cur + 3 * valsz] = 1.0f; p2[cur + 3 * valsz] = 1.0f;
            cur++;
        }
    }
    float X1[16];
    orc_affine_robust(X1, p1, p2, valsz, rigid);
    lin_matmult(X1, Xprev, X);
    free(mv1); free(mv2); free(mi1); free(mi2); free(p1); free(p2);
}

/* linearBCV.cpp:144-263: the multi-level affine (rigid) estimation on HU
 * volumes (+1024 inside); X_out receives the 16 floats column-wise. */
void orc_linear_register(const float* fixed, const float* moving, int m, int n, int o, int levels, const int* grid,
                         const int* radius, const int* quant, int rigid, float* X_out) {
    long long sz = (long long)m * n * o;
    float* im1b = (float*)xmalloc(sizeof(float) * sz);
    float* im1 = (float*)xmalloc(sizeof(float) * sz);
    float* w1 = (float*)xmalloc(sizeof(float) * sz);
    float* w2 = (float*)xmalloc(sizeof(float) * sz);
    uint64_t* mind_m = (uint64_t*)xmalloc(sizeof(uint64_t) * sz);
    uint64_t* mind_f = (uint64_t*)xmalloc(sizeof(uint64_t) * sz);
    uint64_t* mind_w = (uint64_t*)xmalloc(sizeof(uint64_t) * sz);
    for (long long p = 0; p < sz; p++) { im1b[p] = fixed[p] - (-1024.0f); im1[p] = moving[p] - (-1024.0f); }
    float Xprev[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1}, X[16];
    const float ident[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    memcpy(X, Xprev, sizeof(X));
    for (int level = 0; level < levels; level++) {
        float q1 = (float)quant[level];
        int step = grid[level], hw = radius[level];
        float Xinv[16];
        orc_qrsolve(Xinv, Xprev, ident, 4, 4);
        orc_lin_warp(w2, im1b, Xinv, m, n, o);
        orc_lin_warp(w1, im1, Xprev, m, n, o);
        int prev = (int)floorf(0.5f * (float)quant[level > 0 ? level - 1 : 0] + 1.0f);
        int curr = (int)floorf(0.5f * (float)quant[level] + 1.0f);
        if (level == 0 || prev != curr) {
            orc_descriptor(mind_m, im1, m, n, o, curr);
            orc_descriptor(mind_f, im1b, m, n, o, curr);
        }
        int len = 2 * hw + 1, len3 = len * len * len;
        int sz1 = (m / step) * (n / step) * (o / step);
        float* c1 = (float*)xmalloc(sizeof(float) * (size_t)sz1 * len3);
        float* c2 = (float*)xmalloc(sizeof(float) * (size_t)sz1 * len3);
        orc_descriptor(mind_w, w1, m, n, o, curr);
        orc_datacost(mind_f, mind_w, c1, m, n, o, step, hw, q1, 1.0f, 1);
        orc_descriptor(mind_w, w2, m, n, o, curr);
        orc_datacost(mind_m, mind_w, c2, m, n, o, step, hw, q1, 1.0f, 1);
        orc_estimate_affine(X, Xprev, im1b, im1, c1, c2, m, n, o, step, q1, hw, rigid, NULL);
        memcpy(Xprev, X, sizeof(X));
        free(c1); free(c2);
    }
    memcpy(X_out, X, sizeof(X));
    free(im1b); free(im1); free(w1); free(w2); free(mind_m); free(mind_f); free(mind_w);
}
