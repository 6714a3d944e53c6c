// GPU half of the exact image statistics (primsMST.h:101-110): the per-chunk
// work of the decomposition described in stats_host.cpp -- the double sums
// of the addends that give each chunk's binade guess, and each chunk's
// increment sum, bounds and exact halfway points for that guess.  All
// arithmetic is exact (IEEE double/float operations whose results are
// representable, floor, round-to-even), so the GPU produces the same plans as
// the host code would.  One workgroup per chunk of kStatsChunk voxels, 32
// consecutive voxels per thread, the thread results combined in voxel order.
#include <climits>

#include "kernels.h"
#include "stats.h"

namespace deeds {

namespace {

constexpr int kThreads = kStatsChunk / 32;

template <int MODE>
__device__ __forceinline__ double addend(const float* __restrict__ x, long long t, float mu) {
    if (MODE == 0) return (double)x[t];
    const double d = (double)(x[t] - mu);
    return d * d;
}

// t/u = B + r; one rounded step adds B, B + 1, or B + parity at an exact half
template <int MODE>
__device__ __forceinline__ long long split(double a, double inv_u, long long& inc, bool& half) {
    const double q = a * inv_u;
    const double Bd = floor(q);
    const long long B = (long long)Bd;
    if (MODE == 0) {
        const double r = q - Bd;
        inc = B + (r > 0.5 ? 1 : 0);
        half = r == 0.5;
    } else {
        const double r29 = rint((q - Bd) * 536870912.0);  // the double sum's rounding 29 bits below the float ulp
        inc = B + (r29 > 268435456.0 ? 1 : 0);
        half = r29 == 268435456.0;
    }
    return B;
}

template <int MODE>
__global__ __launch_bounds__(kThreads) void k_stats_sums(const float* __restrict__ x, long long n, float mu,
                                                         double* __restrict__ sums) {
    __shared__ double part[kThreads];
    const long long k0 = (long long)blockIdx.x * kStatsChunk + 32LL * threadIdx.x;
    double s = 0;
    for (long long t = k0; t < k0 + 32 && t < n; t++) s += addend<MODE>(x, t, mu);
    part[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double tot = 0;
        for (int w = 0; w < kThreads; w++) tot += part[w];
        sums[blockIdx.x] = tot;
    }
}

template <int MODE>
__global__ __launch_bounds__(kThreads) void k_stats_plan(const float* __restrict__ x, long long n, float mu,
                                                         const int* __restrict__ E, StatsChunk* __restrict__ out,
                                                         long long* __restrict__ halves) {
    __shared__ long long sP[kThreads], sLo[kThreads], sHi[kThreads];
    __shared__ int sN[kThreads];
    __shared__ int sH[kStatsChunk];  // thread-relative halves (an accepted plan keeps them below 2^30)
    const int c = blockIdx.x, tid = threadIdx.x;
    const int e = E[c];
    if (e == kStatsNoPlan) {
        if (tid == 0) out[c] = StatsChunk{0, 0, 0, 0, 0};
        return;
    }
    const double inv_u = ldexp(1.0, -e);
    const long long k0 = (long long)c * kStatsChunk + 32LL * tid;
    long long P = 0, lo = LLONG_MAX, hi = LLONG_MIN;
    int nh = 0;
    bool wide = false;
    for (long long t = k0; t < k0 + 32 && t < n; t++) {
        long long inc;
        bool half;
        const long long B = split<MODE>(addend<MODE>(x, t, mu), inv_u, inc, half);
        const long long a = P + B;
        lo = a < lo ? a : lo;
        hi = a + 1 > hi ? a + 1 : hi;
        if (half) {
            wide |= a >= (1LL << 30) || a <= -(1LL << 30);
            sH[32 * tid + nh++] = (int)a;
        }
        P += inc;
    }
    if (wide) nh = kStatsMaxHalves + 1;  // cannot be an accepted plan: leave the chunk to the host loop
    sP[tid] = P;
    sLo[tid] = lo;
    sHi[tid] = hi;
    sN[tid] = nh;
    __syncthreads();
    if (tid == 0) {
        // combine the threads in voxel order: offsets are the prefix of their increment sums
        long long off = 0, clo = LLONG_MAX, chi = LLONG_MIN;
        int hn = 0;
        for (int w = 0; w < kThreads; w++) {
            if (sLo[w] != LLONG_MAX) {
                clo = off + sLo[w] < clo ? off + sLo[w] : clo;
                chi = off + sHi[w] > chi ? off + sHi[w] : chi;
            }
            const int cnt = sN[w];
            if (hn + cnt <= kStatsMaxHalves && cnt <= 32)
                for (int i = 0; i < cnt; i++) halves[(long long)c * kStatsMaxHalves + hn + i] = off + sH[32 * w + i];
            hn += cnt;
            off += sP[w];
        }
        out[c] = StatsChunk{off, clo, chi, hn, hn <= kStatsMaxHalves ? 1 : 0};
    }
}

__global__ __launch_bounds__(256) void k_gather_chunks(float* __restrict__ dst, const float* __restrict__ x,
                                                       long long n, const int* __restrict__ list) {
    const long long k0 = (long long)list[blockIdx.x] * kStatsChunk;
    float* d = dst + (long long)blockIdx.x * kStatsChunk;
    for (int t = threadIdx.x; t < kStatsChunk && k0 + t < n; t += blockDim.x) d[t] = x[k0 + t];
}

}  // namespace

void launch_gather_chunks(float* dst, const float* x, long long n, const int* list, int count, hipStream_t s) {
    if (count <= 0) return;
    k_gather_chunks<<<count, 256, 0, s>>>(dst, x, n, list);
    DEEDS_LAUNCH_CHECK();
}

void launch_stats_sums(int mode, const float* x, long long n, float mu, double* sums, hipStream_t s) {
    const unsigned chunks = (unsigned)((n + kStatsChunk - 1) / kStatsChunk);
    if (mode == 0)
        k_stats_sums<0><<<chunks, kThreads, 0, s>>>(x, n, mu, sums);
    else
        k_stats_sums<1><<<chunks, kThreads, 0, s>>>(x, n, mu, sums);
    DEEDS_LAUNCH_CHECK();
}

void launch_stats_plan(int mode, const float* x, long long n, float mu, const int* E, StatsChunk* out,
                       long long* halves, hipStream_t s) {
    const unsigned chunks = (unsigned)((n + kStatsChunk - 1) / kStatsChunk);
    if (mode == 0)
        k_stats_plan<0><<<chunks, kThreads, 0, s>>>(x, n, mu, E, out, halves);
    else
        k_stats_plan<1><<<chunks, kThreads, 0, s>>>(x, n, mu, E, out, halves);
    DEEDS_LAUNCH_CHECK();
}

}  // namespace deeds
