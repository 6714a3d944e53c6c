// Exact image statistics with the per-chunk decomposition on the GPU
// (stats.hip) and the ordered application on the host (stats_host.cpp).
#pragma once

#include <hip/hip_runtime.h>

namespace deeds {

constexpr int kStatsChunk = 8192;       // voxels per chunk (one workgroup)
constexpr int kStatsMaxHalves = 512;    // exact halfway points kept per chunk (more: the host loop runs it)
constexpr int kStatsNoPlan = -100000;   // chunk exponent meaning "no decomposition"
constexpr int kStatsPrefetch = 128;     // chunks copied to the host in one go for the sequential loop

// One chunk's decomposition for ulp 2^E: increment sum (halves rounded down),
// min of (P_k + B_k) and max of (P_k + B_k + 1), number of halves, planned flag.
struct StatsChunk {
    long long sum, lo, hi;
    int nh, planned;
};

// mode 0: addends x; mode 1: (double)(x - mu)^2
void launch_stats_sums(int mode, const float* x, long long n, float mu, double* sums, hipStream_t s);
void launch_stats_plan(int mode, const float* x, long long n, float mu, const int* E, StatsChunk* out,
                       long long* halves, hipStream_t s);
// dst[i * kStatsChunk ...] = chunk list[i] of x (the last chunk of x may be short), i < count
void launch_gather_chunks(float* dst, const float* x, long long n, const int* list, int count, hipStream_t s);

// Device and pinned host scratch of one image_stats_device call: the same
// layout (sums | out | halves | E | prefetch list | kStatsPrefetch chunks of
// voxels | one chunk of voxels) in both memories, bytes(count) each.
struct StatsScratch {
    void* dev = nullptr;
    void* host = nullptr;
    static size_t chunks(long long count) { return (size_t)((count + kStatsChunk - 1) / kStatsChunk); }
    static size_t off_out(long long count) { return chunks(count) * sizeof(double); }
    static size_t off_halves(long long count) { return off_out(count) + chunks(count) * sizeof(StatsChunk); }
    static size_t off_E(long long count) {
        return off_halves(count) + chunks(count) * kStatsMaxHalves * sizeof(long long);
    }
    static size_t off_list(long long count) { return off_E(count) + chunks(count) * sizeof(int); }
    static size_t off_stage(long long count) {
        return (off_list(count) + kStatsPrefetch * sizeof(int) + 255) & ~(size_t)255;
    }
    static size_t off_chunk(long long count) {
        return off_stage(count) + (size_t)kStatsPrefetch * kStatsChunk * sizeof(float);
    }
    static size_t bytes(long long count) { return off_chunk(count) + kStatsChunk * sizeof(float); }
};

// Mean and standard deviation of the device image d_im (count floats),
// bit-identical to primsMST.h:101-110.  The chunks no decomposition covers
// (the first one, binade crossings) are copied to the host and run there as
// the sequential loop: every chunk whose running value may cross a binade is
// gathered and copied in one go before the plans are applied (a chunk that
// fails outside that set is copied on its own).  Synchronises stream s.
void image_stats_device(const float* d_im, long long count, const StatsScratch& sc, hipStream_t s, float* mean,
                        float* sd);

}  // namespace deeds
