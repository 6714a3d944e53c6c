// Test infrastructure only: never linked into the product library.
//
// extern "C" shim over the UNMODIFIED reference linearBCV.cpp (with the
// headers it includes: affineLTS1.h, QRsolve.h, MINDSSCbox.h, dataCostD.h,
// ...), compiled where it lies (REF_SRC in oracle/Makefile) with the
// reference's Makefile flags.  The reference's main() is renamed by the
// preprocessor so that its functions can be called one at a time; nothing
// else is changed.  Tests compare the C restatement in oracle/deeds_oracle.c
// against these functions bit for bit.
#define main linearbcv_reference_main
#include "linearBCV.cpp"
#undef main

extern "C" {

// affineLTS1.h:39 warpAffine(warped, input, X, m, n, o)
void ref_lin_warp(float* out, const float* in, const float* X, int m, int n, int o) {
    warpAffine(out, const_cast<float*>(in), const_cast<float*>(X), m, n, o);
}

// QRsolve.h:17 qrsolve(X, A, b, len, len2)
void ref_qrsolve(float* X, const float* A, const float* b, int len, int len2) {
    qrsolve(X, const_cast<float*>(A), const_cast<float*>(b), len, len2);
}

// QRsolve.h:258 affineRobust (RIGID selects findRigid, linearBCV.cpp:48)
void ref_affine_robust(float* RT, const float* pts1, const float* pts2, int len, int rigid) {
    RIGID = rigid != 0;
    affineRobust(RT, const_cast<float*>(pts1), const_cast<float*>(pts2), len);
}

// affineLTS1.h:81 estimateAffine2 (smooths cost / cost2 in place)
void ref_estimate_affine(float* X, const float* Xprev, const float* im1, const float* im2, float* cost, float* cost2,
                         int m, int n, int o, int step, float quant, int hw, int rigid) {
    image_m = m;
    image_n = n;
    image_o = o;
    RIGID = rigid != 0;
    estimateAffine2(X, const_cast<float*>(Xprev), const_cast<float*>(im1), const_cast<float*>(im2), cost, cost2,
                    step, quant, hw);
}

}  // extern "C"
