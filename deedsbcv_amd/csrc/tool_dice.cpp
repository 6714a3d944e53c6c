// diceMulti drop-in: per-label Dice overlap of two label volumes on the GPU
// (reference: diceMulti.cpp:17-79), with the reference's output lines.
//
//   diceMulti seg1.nii.gz seg2.nii.gz
#include <numeric>

#include "tool_common.h"

int main(int argc, char* const argv[]) {
    if (argc < 3) {
        printf("Not enough input arguments. Example syntax:\n");
        printf("./diceMulti seg1.nii seg2.nii\n");
        return 1;
    }
    tool::Volume a, b;
    if (!tool::read_nifti(argv[1], a, true) || !tool::read_nifti(argv[2], b, true)) return -1;
    if (a.m != b.m || a.n != b.n || a.o != b.o) {
        printf("Inconsistent image sizes (must have same dimensions)\n");
        return -1;
    }
    const long long sz = (long long)a.m * a.n * a.o;
    int count = 0;
    if (deeds_dice(a.s.data(), b.s.data(), sz, 0, nullptr, nullptr, &count) != DEEDS_OK) {
        printf("diceMulti failed: %s\n", deeds_last_error());
        return -1;
    }
    std::vector<int16_t> labels(count);
    std::vector<float> dice(count);
    if (deeds_dice(a.s.data(), b.s.data(), sz, count, labels.data(), dice.data(), &count) != DEEDS_OK) {
        printf("diceMulti failed: %s\n", deeds_last_error());
        return -1;
    }
    printf("Dice (for each label): ");
    for (float d : dice) printf(" %4.3f, ", d);
    // float sum in label order divided by the count (diceMulti.cpp:62)
    const float mean = std::accumulate(dice.begin(), dice.end(), 0.0f) / (float)dice.size();
    printf("\nAVERAGE DICE: %f\n", mean);
    return 0;
}
