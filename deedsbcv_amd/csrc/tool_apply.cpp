// applyBCV / applyBCVfloat drop-ins: apply a deeds _displacements.dat field to
// a label volume (nearest neighbour) or an image (trilinear) on the GPU
// (reference: applyBCV.cpp:54-204, applyBCVfloat.cpp:54-204).
//
//   applyBCV      -M moving_seg.nii.gz -O output -D deformed_seg.nii.gz [-A affine.txt]
//   applyBCVfloat -M moving.nii.gz     -O output -D deformed.nii.gz     [-A affine.txt]
#include "tool_common.h"

#ifndef DEEDS_APPLY_FLOAT
#define DEEDS_APPLY_FLOAT 0
#endif

int main(int argc, char* const argv[]) {
    const char* prog = DEEDS_APPLY_FLOAT ? "applyBCVfloat" : "applyBCV";
    if (argc < 4 || argv[1][1] == 'h') {
        printf("=============================================================\n");
        printf("Usage (required input arguments):\n");
        printf("./%s -M moving.nii.gz -O output -D deformed.nii.gz \n", prog);
        printf("optional parameters:\n");
        printf(" -A <affine_matrix.txt> \n");
        printf("=============================================================\n");
        return 1;
    }
    tool::Args a = tool::parse_args(argc, argv);
    if (!a.ok) return -1;
    if (!tool::ends_with_gz(a.deformed) || !tool::ends_with_gz(a.moving)) {
        printf("images must have nii.gz format\n");
        return -1;
    }
    printf("Transforming %s into %s\n", tool::base_name(a.moving).c_str(), tool::base_name(a.deformed).c_str());
    tool::Volume mov;
    if (!tool::read_nifti(a.moving, mov, !DEEDS_APPLY_FLOAT)) return -1;
    const int m = mov.m, n = mov.n, o = mov.o;
    const size_t sz = (size_t)m * n * o;
    float X[16];
    if (a.has_affine) {
        printf("Reading affine matrix file: %s\n", tool::base_name(a.affine).c_str());
        if (!tool::read_affine(a.affine, X)) return -1;
    } else {
        printf("Using identity transform.\n");
        tool::identity(X);
    }
    tool::print_matrix(X);
    std::string inputflow = a.output + "_displacements.dat";
    printf("Reading displacements from:\n%s\n", inputflow.c_str());
    printf("=============================================================\n");
    FILE* fd = fopen(inputflow.c_str(), "rb");
    if (!fd) {
        printf("readFile error. Did not find %s\n", inputflow.c_str());
        return -1;
    }
    fseek(fd, 0, SEEK_END);
    long bytes = ftell(fd);
    fseek(fd, 0, SEEK_SET);
    std::vector<float> flow(bytes / sizeof(float));
    size_t got = fread(flow.data(), sizeof(float), flow.size(), fd);
    fclose(fd);
    if (got != flow.size() || flow.size() < 3) {
        printf("could not read %s\n", inputflow.c_str());
        return -1;
    }
    // grid step from the file size (applyBCV.cpp:163-164)
    int sz3 = (int)(flow.size() / 3);
    int grid_step = (int)std::round(std::pow((double)((float)sz / (float)sz3), 0.3333333));
    printf("grid step %d\n", grid_step);
    if (grid_step < 1 || (size_t)(m / grid_step) * (n / grid_step) * (o / grid_step) != (size_t)sz3) {
        printf("displacement field does not match the image size\n");
        return -1;
    }
    int rc;
    if (DEEDS_APPLY_FLOAT) {
        std::vector<float> out(sz);
        rc = deeds_apply_float(out.data(), mov.f.data(), X, flow.data(), grid_step, m, n, o);
        if (rc == DEEDS_OK && !tool::write_nifti(a.deformed, mov.header, out.data(), m, n, o, 16, 32, 4)) return -1;
    } else {
        std::vector<int16_t> out(sz, 0);
        rc = deeds_apply_seg(out.data(), mov.s.data(), X, flow.data(), grid_step, m, n, o);
        if (rc == DEEDS_OK && !tool::write_nifti(a.deformed, mov.header, out.data(), m, n, o, 4, 16, 2)) return -1;
    }
    if (rc != DEEDS_OK) {
        printf("%s failed: %s\n", prog, deeds_last_error());
        return -1;
    }
    return 0;
}
