// Test infrastructure only: never linked into the product library.
//
// Thin extern "C" shim over the UNMODIFIED deedsBCV reference headers, which
// are compiled where they lie (under the reference checkout given by
// REF_SRC in oracle/Makefile) with the reference's own Makefile flags
// (Makefile:6-27: -O3 -fopenmp -mavx2 -msse4.2 -std=c++11).  The shim defines
// the same globals the reference's main program defines (deedsBCV0.cpp:24-44)
// and forwards to each hot-path function so that tests/golden/make_golden.py
// can dump input/output vectors and tests can compare the C restatement in
// oracle/deeds_oracle.c against the real thing.
#include <iostream>
#include <fstream>
#include <math.h>
#include <sys/time.h>
#include <vector>
#include <algorithm>
#include <numeric>
#include <map>
#include <functional>
#include <string.h>
#include <sstream>
#include <x86intrin.h>
#include <pthread.h>
#include <thread>
#include <cstddef>
#include <climits>
#include <stdint.h>
#include "zlib.h"
#include <sys/stat.h>

using namespace std;

// globals of deedsBCV0.cpp:24-44
int RAND_SAMPLES;
int image_m; int image_n; int image_o; int image_d = 1;
float SSD0 = 0.0; float SSD1 = 0.0; float SSD2 = 0.0; float distfx_global; float beta = 1;
int qc = 1;
struct mind_data { float* im1; float* d1; uint64_t* mindq; int qs; int ind_d1; };
struct parameters {
    float alpha; int levels; bool segment, affine, rigid;
    vector<int> grid_spacing; vector<int> search_radius; vector<int> quantisation;
    string fixed_file, moving_file, output_stem, moving_seg_file, affine_file, deformed_file;
};

#include "imageIOgzType.h"
#include "transformations.h"
#include "primsMST.h"
#include "regularisation.h"
#include "MINDSSCbox.h"
#include "dataCostD.h"

static void set_dims(int m, int n, int o) { image_m = m; image_n = n; image_o = o; }

extern "C" {

// MINDSSCbox.h:2 (in place on `input`)
void ref_boxfilter(float* input, int hw, int m, int n, int o) {
    size_t sz = (size_t)m * n * o;
    float* t1 = new float[sz];
    float* t2 = new float[sz];
    boxfilter(input, t1, t2, hw, m, n, o);
    delete[] t1; delete[] t2;
}

// MINDSSCbox.h:104, one channel l of the six
void ref_distances(const float* im, float* d1_all, int m, int n, int o, int qs, int l) {
    distances(const_cast<float*>(im), d1_all, m, n, o, qs, l);
}

// MINDSSCbox.h:127
void ref_descriptor(uint64_t* mindq, const float* im, int m, int n, int o, int qs) {
    descriptor(mindq, const_cast<float*>(im), m, n, o, qs);
}

// dataCostD.h:108
void ref_datacost(const uint64_t* F, const uint64_t* M, float* cost, int m, int n, int o,
                  int step, int hw, float quant, float alpha, int randnum) {
    int len3 = (2 * hw + 1) * (2 * hw + 1) * (2 * hw + 1);
    dataCostCL((unsigned long*)F, (unsigned long*)M, cost, m, n, o, len3, step, hw, quant, alpha, randnum);
}

// primsMST.h:52
void ref_prims(const float* im, int* ordered, int* parents, float* edgemst, int step, int m, int n, int o) {
    primsGraph(const_cast<float*>(im), ordered, parents, edgemst, step, m, n, o);
}

// regularisation.h:6 (single node, in place on data[ind*len^3 ...])
void ref_message_dt(int ind, float* data, short* indout, int len1, float offx, float offy, float offz) {
    messageDT(ind, data, indout, len1, offx, offy, offz);
}

// regularisation.h:115 (image dims passed explicitly instead of through globals)
void ref_regularisation(float* cost, const float* u0, const float* v0, const float* w0,
                        float* u1, float* v1, float* w1, int hw, int step, float quant,
                        int* ordered, int* parents, float* edgemst, int m, int n, int o) {
    set_dims(m, n, o);
    regularisationCL(cost, const_cast<float*>(u0), const_cast<float*>(v0), const_cast<float*>(w0),
                     u1, v1, w1, hw, step, quant, ordered, parents, edgemst);
}

// transformations.h:5
void ref_interp3(float* out, const float* in, const float* x1, const float* y1, const float* z1,
                 int m, int n, int o, int m2, int n2, int o2, int flag) {
    interp3(out, const_cast<float*>(in), const_cast<float*>(x1), const_cast<float*>(y1),
            const_cast<float*>(z1), m, n, o, m2, n2, o2, flag != 0);
}

// transformations.h:245
void ref_upsample(float* u1, float* v1, float* w1, const float* u0, const float* v0, const float* w0,
                  int m, int n, int o, int m2, int n2, int o2) {
    upsampleDeformationsCL(u1, v1, w1, const_cast<float*>(u0), const_cast<float*>(v0),
                           const_cast<float*>(w0), m, n, o, m2, n2, o2);
}

// transformations.h:189
void ref_consistent(float* u, float* v, float* w, float* u2, float* v2, float* w2,
                    int m, int n, int o, int factor) {
    consistentMappingCL(u, v, w, u2, v2, w2, m, n, o, factor);
}

// transformations.h:84
float ref_jacobian(const float* u, const float* v, const float* w, int m, int n, int o, int factor) {
    return jacobian(const_cast<float*>(u), const_cast<float*>(v), const_cast<float*>(w), m, n, o, factor);
}

// dataCostD.h:267
void ref_warp_affine(float* out, const float* in, const float* im1b, const float* X,
                     const float* u, const float* v, const float* w, int m, int n, int o, float* ssd01) {
    set_dims(m, n, o);
    warpAffine(out, const_cast<float*>(in), const_cast<float*>(im1b), const_cast<float*>(X),
               const_cast<float*>(u), const_cast<float*>(v), const_cast<float*>(w));
    if (ssd01) { ssd01[0] = SSD0; ssd01[1] = SSD1; }
}

// dataCostD.h:213
void ref_warp_image(float* out, const float* im1, const float* im1b,
                    const float* u, const float* v, const float* w, int m, int n, int o, float* ssd01) {
    set_dims(m, n, o);
    warpImageCL(out, const_cast<float*>(im1), const_cast<float*>(im1b),
                const_cast<float*>(u), const_cast<float*>(v), const_cast<float*>(w));
    if (ssd01) { ssd01[0] = SSD0; ssd01[1] = SSD1; }
}

// dataCostD.h:241
void ref_warp_affine_s(short* out, const short* in, const float* X,
                       const float* u, const float* v, const float* w, int m, int n, int o) {
    set_dims(m, n, o);
    warpAffineS(out, const_cast<short*>(in), const_cast<float*>(X),
                const_cast<float*>(u), const_cast<float*>(v), const_cast<float*>(w));
}

}  // extern "C"
