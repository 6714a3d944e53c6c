// linearBCV drop-in: affine (or rigid) pre-alignment on the GPU with the
// reference's flags and output files (reference: linearBCV.cpp:63-329).
//
//   linearBCV -F fixed.nii.gz -M moving.nii.gz -O output [-R 1] [-l 4] [-G 7x6x5x4] [-L 5x4x3x2] [-Q 4x3x2x1]
//             [-S moving_seg.nii.gz]
// writes output_matrix.txt (the -A input of deedsBCV) and, with -S,
// output_deformed_seg.nii.gz.
#include <chrono>

#include "tool_common.h"

int main(int argc, char* const argv[]) {
    if (argc < 4 || argv[1][1] == 'h') {
        printf("=============================================================\n");
        printf("Usage (required input arguments):\n");
        printf("./linearBCV -F fixed.nii.gz -M moving.nii.gz -O output\n");
        printf("optional parameters:\n");
        printf(" -R <find rigid instead of affine transform> (default 0)\n");
        printf(" -l <number of levels> (default 4)\n");
        printf(" -G <grid spacing for each level> (default 7x6x5x4)\n");
        printf(" -L <maximum search radius - each level> (default 5x4x3x2)\n");
        printf(" -Q <quantisation of search step size> (default 4x3x2x1)\n");
        printf("=============================================================\n");
        return 1;
    }
    tool::Args defaults;
    defaults.levels = 4;
    const int g[4] = {7, 6, 5, 4}, r[4] = {5, 4, 3, 2}, q[4] = {4, 3, 2, 1};  // linearBCV.cpp:94-97
    for (int i = 0; i < 4; i++) {
        defaults.grid[i] = g[i];
        defaults.radius[i] = r[i];
        defaults.quant[i] = q[i];
    }
    tool::Args a = tool::parse_args(argc, argv, defaults);
    if (!a.ok) return -1;
    if (!tool::ends_with_gz(a.fixed) || !tool::ends_with_gz(a.moving)) {
        printf("images must have nii.gz format\n");
        return -1;
    }
    printf("Starting linear reg. of %s and %s\n", tool::base_name(a.fixed).c_str(), tool::base_name(a.moving).c_str());
    printf("=============================================================\n");
    if (a.rigid) printf("Searching only for rigid (6 parameter) transform\n");
    auto t0 = std::chrono::steady_clock::now();
    tool::Volume fix, mov;
    if (!tool::read_nifti(a.fixed, fix, false) || !tool::read_nifti(a.moving, mov, false)) return -1;
    if (fix.m != mov.m || fix.n != mov.n || fix.o != mov.o) {
        printf("Inconsistent image sizes (must have same dimensions)\n");
        return -1;
    }
    const int m = fix.m, n = fix.n, o = fix.o;
    deeds_schedule s{};
    s.levels = a.levels;
    for (int l = 0; l < a.levels; l++) {
        s.grid[l] = a.grid[l];
        s.radius[l] = a.radius[l];
        s.quant[l] = a.quant[l];
    }
    s.alpha = 1.0f;
    float X[16];
    std::vector<int> used(a.levels, 0);
    if (deeds_linear_register(fix.f.data(), mov.f.data(), m, n, o, &s, a.rigid, X, used.data()) != DEEDS_OK) {
        printf("linearBCV failed: %s\n", deeds_last_error());
        return -1;
    }
    for (int l = 0; l < a.levels; l++)
        printf("level %d: grid %d, radius %d, quantisation %d, # points used: %d\n", l, s.grid[l], s.radius[l],
               s.quant[l], used[l]);
    // linearBCV.cpp:271-279: the matrix rows, stream-formatted
    std::ofstream mat(a.output + "_matrix.txt");
    for (int i = 0; i < 4; i++) mat << X[i] << "  " << X[i + 4] << "  " << X[i + 8] << "  " << X[i + 12] << "\n";
    mat.close();
    tool::print_matrix(X);
    if (a.has_seg) {
        // linearBCV.cpp:283-303: the moving labels through the same matrix (zero field)
        tool::Volume seg;
        if (!tool::read_nifti(a.seg, seg, true)) return -1;
        if (seg.m != m || seg.n != n || seg.o != o) {
            printf("Inconsistent image sizes (must have same dimensions)\n");
            return -1;
        }
        const size_t sz = (size_t)m * n * o;
        std::vector<float> zero(sz, 0.0f);
        std::vector<int16_t> segw(sz, 0);
        if (deeds_warp_affine_s(segw.data(), seg.s.data(), X, zero.data(), zero.data(), zero.data(), m, n, o) !=
            DEEDS_OK) {
            printf("linearBCV failed: %s\n", deeds_last_error());
            return -1;
        }
        if (!tool::write_nifti(a.output + "_deformed_seg.nii.gz", seg.header, segw.data(), m, n, o, 4, 16, 2))
            return -1;
    }
    const double total = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    printf("Finished. Total time: %f sec.\n", total);
    return 0;
}
