// Image mean and standard deviation exactly as primsMST.h:101-110 computes
// them -- one running float, v <- RN32(v + x) for the mean and
// v <- RN32(RN64(v + d*d)), d = (double)(x - mean), for the variance -- but
// without a dependent chain of eight million additions.
//
// While |v| stays in one binade [2^(E+23), 2^(E+24)) its ulp is u = 2^E and
// v = A*u with A an integer, 2^23 <= |A| < 2^24.  For an addend t write
// t/u = B + r, B = floor(t/u), 0 <= r < 1.  If A + B and A + B + 1 stay in
// that range (of A's sign) the exact sum keeps the ulp, and one rounded step
// gives
//   A + B            r below one half,
//   A + B + 1        r above one half,
//   A + B + (A+B)&1  r exactly one half (round to even),
// for either sign of A (round to nearest even is symmetric under negation).
// (For the variance the sum is first rounded to double, i.e. r to a multiple
// of 2^-29 with ties to even, before the comparison.)  Only the exact halves
// depend on the running value.  So a chunk of the image decomposes in
// parallel, for a given u, into the sum of its increments (halves rounded
// down), the relative positions of its halves and bounds on every
// intermediate A + B; a sequential pass over the chunks then applies them to
// the exact running A, resolving each half with its true parity.  u is
// guessed per chunk from a double prefix sum of the addends and checked in
// the sequential pass; a chunk whose guess or bounds fail there (the first
// chunks, binade crossings, sign changes) runs as the plain sequential loop.
// Either way the result is bit-identical to the sequential loops.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <thread>
#include <vector>

#include "common.h"
#include "mst_host.h"
#include "stats.h"

namespace deeds {

namespace {

constexpr long long kChunk = 8192;
constexpr long long kA0 = 1LL << 23, kA1 = 1LL << 24;

struct Chunk {
    int E = 0;            // log2 of the ulp the chunk was decomposed with
    bool planned = false;  // decomposed (the guessed running value was a normal float)
    double addend_sum = 0;  // double sum of the chunk's addends (for the next chunks' guesses)
    long long sum = 0;     // sum of the increments, every half rounded down
    long long lo = 0, hi = 0;  // min of (P_k + B_k), max of (P_k + B_k + 1) over the steps, P_k = prefix of increments
    std::vector<long long> halves;  // P_k + B_k at each exact half
};

// v = A * 2^E with 2^23 <= |A| < 2^24: E for a normal float v of either sign.
inline int ulp_exp(float v) {
    int x;
    std::frexp(v, &x);  // |v| in [2^(x-1), 2^x)
    return x - 24;
}

// floor(q) for |q| < 2^52 without a library call.
inline long long floor_ll(double q) {
    long long b = (long long)q;  // truncates toward zero
    return b - ((double)b > q ? 1 : 0);
}

// y rounded to an integer, ties to even (0 <= y < 2^52, default rounding mode).
inline double round_even(double y) {
    volatile double big = 4503599627370496.0;  // 2^52: y + 2^52 rounds y to an integer
    return (y + big) - big;
}

// Addend of step t and its decomposition for ulp 2^E.
struct MeanStep {
    const float* x;
    double addend(long long t) const { return (double)x[t]; }
    float step(float v, long long t) const { return v + x[t]; }
    // t/u = B + r; returns B, sets inc and half
    static long long split(double a, double inv_u, long long& inc, bool& half) {
        const double q = a * inv_u;
        const long long B = floor_ll(q);
        const double r = q - (double)B;
        inc = B + (r > 0.5 ? 1 : 0);
        half = r == 0.5;
        return B;
    }
};

struct VarStep {
    const float* x;
    float mu;
    double addend(long long t) const {
        const double d = (double)(x[t] - mu);
        return d * d;
    }
    float step(float v, long long t) const { return (float)((double)v + addend(t)); }
    static long long split(double a, double inv_u, long long& inc, bool& half) {
        const double q = a * inv_u;
        const long long B = floor_ll(q);
        // the double sum keeps 29 bits below the float ulp: round r there (ties to even) first
        const double r29 = round_even((q - (double)B) * 536870912.0);
        inc = B + (r29 > 268435456.0 ? 1 : 0);
        half = r29 == 268435456.0;
        return B;
    }
};

template <class Step>
void plan_chunk(const Step& st, long long k0, long long k1, float guess, Chunk& c) {
    c.planned = std::isnormal(guess);
    if (!c.planned) return;
    c.E = ulp_exp(guess);
    const double inv_u = std::ldexp(1.0, -c.E);
    long long P = 0, lo = LLONG_MAX, hi = LLONG_MIN;
    for (long long t = k0; t < k1; t++) {
        long long inc;
        bool half;
        const long long B = Step::split(st.addend(t), inv_u, inc, half);
        const long long a = P + B;
        lo = std::min(lo, a);
        hi = std::max(hi, a + 1);
        if (half) c.halves.push_back(a);
        P += inc;
    }
    c.sum = P;
    c.lo = lo;
    c.hi = hi;
}

// Applies a chunk decomposed for ulp 2^E to v; false when the decomposition
// does not hold for this v (then the chunk runs as the sequential loop).
bool apply_plan(int E, long long sum, long long lo, long long hi, const long long* halves, long long nh, float& v) {
    if (!std::isnormal(v) || ulp_exp(v) != E) return false;
    const long long A = (long long)std::ldexp((double)v, -E);
    // every step's integer part (rounded halves included, at most nh of them)
    // and its successor stay in A's binade: [2^23, 2^24 - 1] or [-(2^24 - 1), -2^23]
    if (A > 0 ? (A + lo < kA0 || A + nh + hi > kA1 - 1) : (A + lo < -(kA1 - 1) || A + nh + hi > -kA0))
        return false;
    long long T = 0;
    for (long long h = 0; h < nh; h++) T += (A + T + halves[h]) & 1;
    v = (float)std::ldexp((double)(A + sum + T), E);
    return true;
}

bool apply_chunk(const Chunk& c, float& v) {
    return c.planned && apply_plan(c.E, c.sum, c.lo, c.hi, c.halves.data(), (long long)c.halves.size(), v);
}

template <class Step>
float accumulate(const Step& st, long long count) {
    const long long nch = (count + kChunk - 1) / kChunk;
    std::vector<Chunk> ch(nch);
    // guesses: running double sums of the addends at each chunk start
    auto sums = [&](long long c0, long long c1) {
        for (long long c = c0; c < c1; c++) {
            double s = 0;
            const long long k1 = std::min(count, (c + 1) * kChunk);
            for (long long t = c * kChunk; t < k1; t++) s += st.addend(t);
            ch[c].addend_sum = s;
        }
    };
    const int threads = (int)std::max(1u, std::min(8u, std::thread::hardware_concurrency()));
    auto parallel = [&](auto fn) {
        if (nch < 4 * threads) {
            fn(0LL, nch);
            return;
        }
        std::vector<std::thread> pool;
        for (int w = 0; w < threads; w++)
            pool.emplace_back([&, w] { fn(nch * w / threads, nch * (w + 1) / threads); });
        for (auto& th : pool) th.join();
    };
    parallel(sums);
    std::vector<float> guess(nch);
    double run = 0;
    for (long long c = 0; c < nch; c++) {
        guess[c] = (float)run;
        run += ch[c].addend_sum;
    }
    parallel([&](long long c0, long long c1) {
        for (long long c = c0; c < c1; c++)
            plan_chunk(st, c * kChunk, std::min(count, (c + 1) * kChunk), guess[c], ch[c]);
    });
    float v = 0.0f;
    for (long long c = 0; c < nch; c++) {
        if (apply_chunk(ch[c], v)) continue;
        const long long k1 = std::min(count, (c + 1) * kChunk);
        for (long long t = c * kChunk; t < k1; t++) v = st.step(v, t);
    }
    return v;
}

// The same accumulation with the chunk sums and decompositions computed on the GPU.
template <class Step>
float accumulate_device(const Step& st, int mode, const float* d_im, long long count, float mu,
                        const StatsScratch& sc, hipStream_t s) {
    const size_t nch = StatsScratch::chunks(count);
    char* hb = static_cast<char*>(sc.host);
    char* db = static_cast<char*>(sc.dev);
    double* h_sums = reinterpret_cast<double*>(hb);
    StatsChunk* h_out = reinterpret_cast<StatsChunk*>(hb + StatsScratch::off_out(count));
    long long* h_halves = reinterpret_cast<long long*>(hb + StatsScratch::off_halves(count));
    int* h_E = reinterpret_cast<int*>(hb + StatsScratch::off_E(count));
    double* d_sums = reinterpret_cast<double*>(db);
    StatsChunk* d_out = reinterpret_cast<StatsChunk*>(db + StatsScratch::off_out(count));
    long long* d_halves = reinterpret_cast<long long*>(db + StatsScratch::off_halves(count));
    int* d_E = reinterpret_cast<int*>(db + StatsScratch::off_E(count));
    launch_stats_sums(mode, d_im, count, mu, d_sums, s);
    DEEDS_HIP(hipMemcpyAsync(h_sums, d_sums, nch * sizeof(double), hipMemcpyDeviceToHost, s));
    DEEDS_HIP(hipStreamSynchronize(s));
    std::vector<double> at(nch + 1);  // the double prefix of the addends at each chunk boundary
    double run = 0;
    for (size_t c = 0; c < nch; c++) {
        at[c] = run;
        const float g = (float)run;
        h_E[c] = std::isnormal(g) ? ulp_exp(g) : kStatsNoPlan;
        run += h_sums[c];
    }
    at[nch] = run;
    DEEDS_HIP(hipMemcpyAsync(d_E, h_E, nch * sizeof(int), hipMemcpyHostToDevice, s));
    launch_stats_plan(mode, d_im, count, mu, d_E, d_out, d_halves, s);
    DEEDS_HIP(hipMemcpyAsync(h_out, d_out, nch * sizeof(StatsChunk), hipMemcpyDeviceToHost, s));
    DEEDS_HIP(hipMemcpyAsync(h_halves, d_halves, nch * kStatsMaxHalves * sizeof(long long), hipMemcpyDeviceToHost,
                             s));
    DEEDS_HIP(hipStreamSynchronize(s));
    // Chunks likely to run as the sequential loop: no plan, or a running
    // value near a power of two (within 1/128 of the double prefix, which the
    // float sum tracks far closer than that): the binade the plan assumed may
    // not hold.  They are gathered on the device and copied in one go.
    int* h_list = reinterpret_cast<int*>(hb + StatsScratch::off_list(count));
    int* d_list = reinterpret_cast<int*>(db + StatsScratch::off_list(count));
    float* h_stage = reinterpret_cast<float*>(hb + StatsScratch::off_stage(count));
    float* d_stage = reinterpret_cast<float*>(db + StatsScratch::off_stage(count));
    std::vector<int> slot(nch, -1);
    int ns = 0;
    for (size_t c = 0; c < nch && ns < kStatsPrefetch; c++) {
        bool suspect = !h_out[c].planned;
        if (!suspect) {
            const double a = at[c], b = at[c + 1];
            if (!(a > 0 && b > 0) && !(a < 0 && b < 0)) {
                suspect = true;  // a sign change or zero
            } else {
                const double lo = std::min(std::fabs(a), std::fabs(b)) * (1.0 - 1.0 / 128);
                const double hi = std::max(std::fabs(a), std::fabs(b)) * (1.0 + 1.0 / 128);
                int el, eh;
                std::frexp(lo, &el);
                std::frexp(hi, &eh);
                suspect = el != eh;
            }
        }
        if (suspect) {
            slot[c] = ns;
            h_list[ns++] = (int)c;
        }
    }
    if (ns > 0) {
        DEEDS_HIP(hipMemcpyAsync(d_list, h_list, ns * sizeof(int), hipMemcpyHostToDevice, s));
        launch_gather_chunks(d_stage, d_im, count, d_list, ns, s);
        DEEDS_HIP(hipMemcpyAsync(h_stage, d_stage, (size_t)ns * kStatsChunk * sizeof(float), hipMemcpyDeviceToHost, s));
        DEEDS_HIP(hipStreamSynchronize(s));
    }
    float* h_chunk = reinterpret_cast<float*>(hb + StatsScratch::off_chunk(count));
    Step local = st;
    float v = 0.0f;
    for (size_t c = 0; c < nch; c++) {
        const StatsChunk& p = h_out[c];
        if (p.planned && apply_plan(h_E[c], p.sum, p.lo, p.hi, h_halves + c * kStatsMaxHalves, p.nh, v)) continue;
        const long long k0 = (long long)c * kStatsChunk, k1 = std::min(count, k0 + kStatsChunk);
        if (slot[c] >= 0) {
            local.x = h_stage + (size_t)slot[c] * kStatsChunk;
        } else {
            DEEDS_HIP(hipMemcpyAsync(h_chunk, d_im + k0, (k1 - k0) * sizeof(float), hipMemcpyDeviceToHost, s));
            DEEDS_HIP(hipStreamSynchronize(s));
            local.x = h_chunk;
        }
        for (long long t = 0; t < k1 - k0; t++) v = local.step(v, t);
    }
    return v;
}

}  // namespace

void image_stats_device(const float* d_im, long long count, const StatsScratch& sc, hipStream_t s, float* mean,
                        float* sd) {
    float mu = accumulate_device(MeanStep{nullptr}, 0, d_im, count, 0.0f, sc, s);
    mu = mu / (float)count;
    const float var = accumulate_device(VarStep{nullptr, mu}, 1, d_im, count, mu, sc, s);
    *mean = mu;
    *sd = std::sqrt(var / (float)count);
}

void image_stats(const float* im, long long count, float* mean, float* sd) {
    float mu = accumulate(MeanStep{im}, count);
    mu = mu / (float)count;
    const float var = accumulate(VarStep{im, mu}, count);
    *mean = mu;
    *sd = std::sqrt(var / (float)count);
}

}  // namespace deeds
