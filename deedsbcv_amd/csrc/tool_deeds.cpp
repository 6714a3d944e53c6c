// deedsBCV drop-in: same command line, same output files, registration on the
// GPU through libdeeds_hip (reference: deedsBCV0.cpp:55-385).
//
//   deedsBCV -F fixed.nii.gz -M moving.nii.gz -O output [-a 1.6] [-l 5]
//            [-G 8x7x6x5x4] [-L 8x7x6x5x4] [-Q 5x4x3x2x1] [-S moving_seg.nii.gz] [-A affine.txt]
//
// Writes output_displacements.dat, output_deformed.nii.gz and, with -S,
// output_deformed_seg.nii.gz.
#include <sys/time.h>

#include <algorithm>

#include "tool_common.h"

static double now_s() {
    timeval t;
    gettimeofday(&t, nullptr);
    return t.tv_sec + t.tv_usec * 1e-6;
}

int main(int argc, char* const argv[]) {
    if (argc < 4 || argv[1][1] == 'h') {
        printf("=============================================================\n");
        printf("Usage (required input arguments):\n");
        printf("./deedsBCV -F fixed.nii.gz -M moving.nii.gz -O output\n");
        printf("optional parameters:\n");
        printf(" -a <regularisation parameter alpha> (default 1.6)\n");
        printf(" -l <number of levels> (default 5)\n");
        printf(" -G <grid spacing for each level> (default 8x7x6x5x4)\n");
        printf(" -L <maximum search radius - each level> (default 8x7x6x5x4)\n");
        printf(" -Q <quantisation of search step size> (default 5x4x3x2x1)\n");
        printf(" -S <moving_segmentation.nii> (short int)\n");
        printf(" -A <affine_matrix.txt> \n");
        printf("=============================================================\n");
        return 1;
    }
    tool::Args a = tool::parse_args(argc, argv);
    if (!a.ok) return -1;
    if (a.required < 3) printf("Missing argmuents, use -h for help.\n");
    if (!tool::ends_with_gz(a.fixed) || !tool::ends_with_gz(a.moving)) {
        printf("images must have nii.gz format\n");
        return -1;
    }
    printf("Starting registration of %s and %s\n", tool::base_name(a.fixed).c_str(), tool::base_name(a.moving).c_str());
    printf("=============================================================\n");
    tool::Volume fix, mov;
    if (!tool::read_nifti(a.fixed, fix, false) || !tool::read_nifti(a.moving, mov, false)) return -1;
    if (fix.m != mov.m || fix.n != mov.n || fix.o != mov.o) {
        printf("Inconsistent image sizes (must have same dimensions)\n");
        return -1;
    }
    const int m = fix.m, n = fix.n, o = fix.o;
    const size_t sz = (size_t)m * n * o;
    float X[16];
    if (a.has_affine) {
        printf("Reading affine matrix file: %s\n", tool::base_name(a.affine).c_str());
        if (!tool::read_affine(a.affine, X)) return -1;
    } else {
        printf("Starting with identity transform.\n");
        tool::identity(X);
    }
    tool::print_matrix(X);
    printf("MIND STEPS");
    for (int l = 0; l < a.levels; l++) printf(", %d", (int)std::floor(0.5f * (float)a.quant[l] + 1.0f));
    printf(" \n");

    deeds_schedule S{};
    S.levels = a.levels;
    for (int l = 0; l < a.levels; l++) {
        S.grid[l] = a.grid[l];
        S.radius[l] = a.radius[l];
        S.quant[l] = a.quant[l];
    }
    S.alpha = a.alpha;
    int g = S.grid[S.levels - 1];
    size_t sz1 = (size_t)(m / g) * (n / g) * (o / g);
    std::vector<float> flow(3 * sz1), warped(sz);
    deeds_timing t{};
    double t0 = now_s();
    int rc = deeds_register(fix.f.data(), mov.f.data(), X, m, n, o, &S, flow.data(), warped.data(), &t);
    double t1 = now_s();
    if (rc != DEEDS_OK) {
        printf("registration failed: %s\n", deeds_last_error());
        return -1;
    }
    for (int l = 0; l < t.levels; l++) {
        int step = S.grid[l], hw = S.radius[l];
        int len = 2 * hw + 1;
        double sz1l = (double)(m / step) * (n / step) * (o / step);
        printf("==========================================================\n");
        printf("Level %d grid=%d with sizes: %dx%dx%d hw=%d quant=%d\n", l, step, m / step, n / step, o / step, hw,
               S.quant[l]);
        printf("==========================================================\n");
        printf("TMdDSTMdDS\nTime: MIND=%g, data=%g, MST-reg=%g, transf.=%g\n speed=%g dof/s\n", t.mind[l], t.data[l],
               t.mst_reg[l], t.transf[l], 2.0 * sz1l * len * len * len / std::max(1e-9, t.data[l] + t.mst_reg[l]));
        printf("std(J)=%.2f (J<0)=%.2fe-7  \n", std::round(t.jac_std[l] * 100) / 100.0,
               std::round(t.jac_neg[l] * 1e7) / 100.0);
    }
    std::string stem = a.output;
    FILE* fd = fopen((stem + "_displacements.dat").c_str(), "wb");
    if (!fd || fwrite(flow.data(), sizeof(float), flow.size(), fd) != flow.size()) {
        printf("could not write %s_displacements.dat\n", stem.c_str());
        if (fd) fclose(fd);
        return -1;
    }
    fclose(fd);
    // the reference reuses one header buffer: the moving image's (deedsBCV0.cpp:125-128)
    if (!tool::write_nifti(stem + "_deformed.nii.gz", mov.header, warped.data(), m, n, o, 16, 32, 4)) return -1;
    printf("SSD before registration: %g and after %g\n", t.ssd_before, t.ssd_after);
    if (a.has_seg) {
        tool::Volume seg;
        if (!tool::read_nifti(a.seg, seg, true)) return -1;
        if (seg.m != m || seg.n != n || seg.o != o) {
            printf("Inconsistent segmentation size\n");
            return -1;
        }
        std::vector<int16_t> segw(sz, 0);
        rc = deeds_apply_seg(segw.data(), seg.s.data(), X, flow.data(), g, m, n, o);
        if (rc != DEEDS_OK) {
            printf("segmentation warp failed: %s\n", deeds_last_error());
            return -1;
        }
        std::string outseg = stem + "_deformed_seg.nii.gz";
        printf("outputseg %s\n", outseg.c_str());
        if (!tool::write_nifti(outseg, seg.header, segw.data(), m, n, o, 4, 16, 2)) return -1;
    }
    printf("Finished. Total time: %g sec. (%g sec. for MIND+data+reg+trans)\n", t.total, t1 - t0);
    return 0;
}
