// Hamming data-cost volume on gfx950 (reference: dataCostD.h:108-210).
//
// cost[node*len3 + l] = scale * sum_{block samples s} popcount(F[s] ^ M[s + d(l)])
// with the moving codes clamp-padded by quant*hw (:126-138) so the inner loop
// needs no bounds logic.  Two forms: per control point (a workgroup per node,
// its fixed codes wave-uniform, each lane owning labels l, l+256, ...), and
// row-tiled (a workgroup per row of nodes and chunk of label planes, moving
// sample rows staged in LDS), chosen by the launcher.  Both accumulate
// exact integer sums with v_xor + v_bcnt (two 32-bit halves).
#include "kernels.h"

#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <string>
#include <tuple>

namespace deeds {

namespace {

__global__ void k_pad_codes(uint64_t* __restrict__ Mp, const uint64_t* __restrict__ M, int m, int n, int o,
                            int pad) {
    int mp = m + 2 * pad, np = n + 2 * pad, op = o + 2 * pad;
    long long total = (long long)mp * np * op;
    long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= total) return;
    int i = (int)(p % mp);
    int j = (int)((p / mp) % np);
    int k = (int)(p / ((long long)mp * np));
    i = min(max(i - pad, 0), m - 1);
    j = min(max(j - pad, 0), n - 1);
    k = min(max(k - pad, 0), o - 1);
    Mp[p] = M[i + (long long)j * m + (long long)k * m * n];
}

// Blocks are dispatched round-robin over the 8 XCDs; hand each XCD a
// contiguous run of control points so neighbouring points, whose moving-code
// neighbourhoods overlap almost entirely, share that XCD's L2.
__device__ __forceinline__ int xcd_node(int b, int nodes) {
    const int per = (nodes + 7) / 8;
    return (b & 7) * per + (b >> 3);
}

// Sample loops unrolled at compile time (STEP and the skips are template
// arguments), so every wave keeps the moving-code loads of a whole block in
// flight instead of one at a time.
template <int STEP, int SI, int SJ, int SK>
__global__ __launch_bounds__(256) void k_datacost_t(float* __restrict__ cost, const uint64_t* __restrict__ F,
                                                    const uint64_t* __restrict__ Mp, int m, int n, int m1, int n1,
                                                    int nodes, int len, int q, float scale, int mp, int np) {
    const int node = xcd_node(blockIdx.x, nodes);
    if (node >= nodes) return;
    const int y = node % m1;
    const int x = (node / m1) % n1;
    const int z = node / (m1 * n1);
    const int len3 = len * len * len;
    const long long mn = (long long)m * n, mpnp = (long long)mp * np;
    const uint64_t* fblk = F + (long long)y * STEP + (long long)x * STEP * m + (long long)z * STEP * mn;
    const uint64_t* mblk = Mp + (long long)y * STEP + (long long)x * STEP * mp + (long long)z * STEP * mpnp;
    for (int l = threadIdx.x; l < len3; l += blockDim.x) {
        const int ls = l % len, lx = (l / len) % len, lz = l / (len * len);
        const uint64_t* mb = mblk + ls * q + (long long)(lx * q) * mp + (long long)(lz * q) * mpnp;
        unsigned sum = 0;
#pragma unroll
        for (int k = 0; k < STEP; k += SK) {
#pragma unroll
            for (int j = 0; j < STEP; j += SJ) {
                const uint64_t* fr = fblk + k * mn + (long long)j * m;
                const uint64_t* mr = mb + k * mpnp + (long long)j * mp;
#pragma unroll
                for (int i = 0; i < STEP; i += SI) sum += __builtin_popcountll(fr[i] ^ mr[i]);
            }
        }
        cost[(long long)node * len3 + l] = (float)sum * scale;
    }
}

__global__ __launch_bounds__(256) void k_datacost(float* __restrict__ cost, const uint64_t* __restrict__ F,
                                                  const uint64_t* __restrict__ Mp, int m, int n, int m1, int n1,
                                                  int step, int len, int q, int si, int sj, int sk, float scale,
                                                  int mp, int np) {
    const int node = blockIdx.x;
    const int y = node % m1;
    const int x = (node / m1) % n1;
    const int z = node / (m1 * n1);
    const int len3 = len * len * len;
    const long long mn = (long long)m * n, mpnp = (long long)mp * np;
    const uint64_t* fblk = F + (long long)y * step + (long long)x * step * m + (long long)z * step * mn;
    // padded coordinate of (block origin + label displacement) = y*step + ls*q
    const uint64_t* mblk = Mp + (long long)y * step + (long long)x * step * mp + (long long)z * step * mpnp;
    for (int l = threadIdx.x; l < len3; l += blockDim.x) {
        int ls = l % len, lx = (l / len) % len, lz = l / (len * len);
        const uint64_t* mb = mblk + ls * q + (long long)(lx * q) * mp + (long long)(lz * q) * mpnp;
        unsigned sum = 0;
        for (int k = 0; k < step; k += sk) {
            for (int j = 0; j < step; j += sj) {
                const uint64_t* fr = fblk + k * mn + (long long)j * m;
                const uint64_t* mr = mb + k * mpnp + (long long)j * mp;
                for (int i = 0; i < step; i += si) {
                    uint64_t t = fr[i] ^ mr[i];
                    sum += __builtin_popcountll(t);
                }
            }
        }
        cost[(long long)node * len3 + l] = (float)sum * scale;
    }
}

// Row-tiled form for full sampling along i (skip_i == 1, every schedule of
// deedsBCV).  A workgroup owns one row of nodes along i -- fixed (x, z), all
// m1 nodes -- and a chunk of label planes (lx, lz).  The fixed codes of the
// row's R = ceil(STEP/SJ) * ceil(STEP/SK) sample rows are loaded once (each
// thread keeps its position's R codes in registers); for every plane the
// matching moving rows (the same sample rows displaced by
// (lx*q, lz*q), over the i-range the row's blocks reach under every ys) are
// staged with coalesced loads.  Thread t owns position t of the row: it keeps
// the fixed code in a register and accumulates popcount(F ^ M[t + ys*q]) for
// every ys over the R rows, so a wave reads consecutive LDS words (no bank
// conflicts for any q) and one fixed read serves all ys; with the label count
// and quantisation compile-time, every read has an immediate offset and an
// evaluation is xor, xor, bcnt, bcnt.  The per-position sums are then added
// over each node's STEP positions through LDS (reusing the moving-row buffer).
// Sums are exact integers: the result equals the per-node kernels'.
//
// Each evaluation accumulates with v_bcnt_u32_b32's own addend: xor, xor,
// bcnt, bcnt (4 operations; the compiler's popcount adds a fifth).
// PIPE (rows of at most 256 positions, 15 or more labels per axis): the
// workgroup is at most 256 threads and the kernel is compiled for three waves
// per SIMD (what the LDS allows), which leaves registers for the next sample
// row's LEN moving codes to be read while the current row is evaluated.
// Measured alone: 4.4 -> 4.8 T evals/s at 17 labels, 4.6 -> 5.0 at 15; at
// 13 and 11 labels the smaller LDS footprint allows more waves and the plain
// loop is faster.
__device__ __forceinline__ unsigned bcnt_acc(unsigned x, unsigned acc) {
    unsigned r;
    asm("v_bcnt_u32_b32 %0, %1, %2" : "=v"(r) : "v"(x), "v"(acc));
    return r;
}

template <int STEP, int SJ, int SK, int LEN, int Q, bool PIPE, bool RING>
__global__ __attribute__((amdgpu_flat_work_group_size(1, PIPE ? 256 : 512), amdgpu_waves_per_eu(PIPE ? 3 : 1)))
void k_datacost_rows(float* __restrict__ cost, const uint64_t* __restrict__ F, const uint64_t* __restrict__ Mp,
                     int m, int n, int m1, int n1, float scale, int mp, int np, int chunk_planes, int chunks,
                     int work) {
    constexpr int RJ = (STEP + SJ - 1) / SJ, RK = (STEP + SK - 1) / SK, R = RJ * RK;
    constexpr int len = LEN, q = Q;
    extern __shared__ uint64_t lds_rows[];
    // contiguous work ranges per XCD (blocks are dispatched round-robin over the 8 XCDs),
    // so the planes of one row and the rows next to it share an L2
    const int per = (work + 7) / 8;
    const int b = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    if (b >= work) return;
    const int row = b / chunks, chunk = b - row * chunks;
    const int x = row % n1, z = row / n1;
    const int W = m1 * STEP, WM = W + (len - 1) * q;
    // RING: consecutive planes of one z-slice (lx -> lx + 1) shift the j
    // sample rows by Q; when Q is a multiple of SJ, D = Q / SJ of the RJ
    // j-rows are new and the others were staged for the previous plane.  The
    // j-row of index D*lx + jr then lives in ring slot (D*lx + jr) mod RJ, and
    // only the new rows are loaded; the sums get their own buffer so the rows
    // survive (more LDS per workgroup: the autotuner weighs the two).
    constexpr int D = Q / SJ;
    constexpr bool REUSE = RING && Q % SJ == 0 && D < RJ;
    uint64_t* sM = lds_rows;  // moving rows (then, unless REUSE, the per-position sums)
    unsigned* sD = REUSE ? reinterpret_cast<unsigned*>(sM + R * WM) : reinterpret_cast<unsigned*>(sM);
    const long long mn = (long long)m * n, mpnp = (long long)mp * np;
    const int t = threadIdx.x, T = blockDim.x;  // T >= W, and WM < W + T
    // this thread's fixed codes of the R sample rows stay in registers for every plane
    uint64_t fr[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const uint64_t* src = F + (long long)(x * STEP + (r % RJ) * SJ) * m + (long long)(z * STEP + (r / RJ) * SK) * mn;
        fr[r] = t < W ? src[t] : 0;
    }
    const int len2 = len * len, len3 = len2 * len;
    const int p0 = chunk * chunk_planes, p1 = min(len2, p0 + chunk_planes);
    const int pairs = m1 * len;
    for (int p = p0; p < p1; p++) {
        const int lx = p % len, lz = p / len;
        __syncthreads();  // the previous plane's sums (and rows) have been read
        const uint64_t* mbase = Mp + (long long)(x * STEP + lx * q) * mp + (long long)(z * STEP + lz * q) * mpnp;
        // LDS row of sample row r = (jr, kr): ring slot of jr within kr's RJ rows
        int rb[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            if constexpr (REUSE) {
                const int jr = r % RJ, kr = r / RJ;
                rb[r] = (kr * RJ + (D * lx + jr) % RJ) * WM;
            } else {
                rb[r] = r * WM;
            }
        }
        const bool all = !REUSE || p == p0 || lx == 0;
        // every row's loads are issued before the first LDS store: the indices
        // are clamped into the row (lanes past it reload its last code and do
        // not store it), so no load sits under a branch and none waits for
        // the previous row's
        const int i0 = min(t, WM - 1), i1 = min(t + T, WM - 1);
        uint64_t ga[R], gb[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            if (!all && r % RJ < RJ - D) continue;  // staged for the previous plane
            const uint64_t* src = mbase + (long long)((r % RJ) * SJ) * mp + (long long)((r / RJ) * SK) * mpnp;
            ga[r] = src[i0];
            gb[r] = src[i1];
        }
#pragma unroll
        for (int r = 0; r < R; r++) {
            if (!all && r % RJ < RJ - D) continue;
            if (t < WM) sM[rb[r] + t] = ga[r];
            if (t + T < WM) sM[rb[r] + t + T] = gb[r];
        }
        __syncthreads();
        unsigned acc[LEN];
#pragma unroll
        for (int ys = 0; ys < LEN; ys++) acc[ys] = 0;
        if (t < W && PIPE) {
            uint64_t mv[2][LEN];  // this row's and the next row's moving codes
#pragma unroll
            for (int ys = 0; ys < LEN; ys++) mv[0][ys] = sM[rb[0] + t + ys * Q];
#pragma unroll
            for (int r = 0; r < R; r++) {
                const int cur = r & 1;
                if (r + 1 < R) {
#pragma unroll
                    for (int ys = 0; ys < LEN; ys++) mv[cur ^ 1][ys] = sM[rb[r + 1] + t + ys * Q];
                }
                const uint64_t f = fr[r];
#pragma unroll
                for (int ys = 0; ys < LEN; ys++) {
                    const uint64_t d = f ^ mv[cur][ys];
                    acc[ys] = bcnt_acc((unsigned)d, bcnt_acc((unsigned)(d >> 32), acc[ys]));
                }
            }
        } else if (t < W) {
#pragma unroll
            for (int r = 0; r < R; r++) {
                const uint64_t f = fr[r];
                const uint64_t* mr = sM + rb[r] + t;
#pragma unroll
                for (int ys = 0; ys < LEN; ys++) {
                    const uint64_t d = f ^ mr[ys * Q];
                    acc[ys] = bcnt_acc((unsigned)d, bcnt_acc((unsigned)(d >> 32), acc[ys]));
                }
            }
        }
        __syncthreads();  // every moving row has been read (unless REUSE, the buffer takes the sums)
        if (t < W) {
#pragma unroll
            for (int ys = 0; ys < LEN; ys++) sD[t * LEN + ys] = acc[ys];
        }
        __syncthreads();
        for (int pr = t; pr < pairs; pr += blockDim.x) {
            const int yi = pr / len, ys = pr - yi * len;
            const unsigned* d = sD + yi * STEP * LEN + ys;  // position-major: lanes read consecutive words
            unsigned sum = 0;
#pragma unroll
            for (int i = 0; i < STEP; i++) sum += d[i * LEN];
            const long long node = yi + (long long)x * m1 + (long long)z * m1 * n1;
            cost[node * len3 + (long long)p * len + ys] = (float)sum * scale;
        }
    }
}

}  // namespace

CostPlan datacost_plan(int m, int n, int o, int step, int hw, float quant, float alpha, int randnum) {
    // dataCostD.h:141-163 with RAND_SAMPLES = randnum
    CostPlan p{};
    int sz = 1, sx = 1, sy = 1;
    if (step > 4) {
        if (randnum > 0) { sz = 2; sx = 2; }
        if (randnum > 1) sy = 2;
    }
    if (randnum > 1 && step > 7) { sz = 3; sx = 3; sy = 3; }
    if (step == 4 && randnum > 1) sz = 2;
    p.skip_i = sy; p.skip_j = sx; p.skip_k = sz;
    float maxsamp = std::ceil((float)step / (float)sx) * std::ceil((float)step / (float)sz) *
                    std::ceil((float)step / (float)sy);
    float alphai = (float)step / (alpha * quant);
    p.scale = (float)(0.5 * (double)alphai / (double)maxsamp);
    int q = (int)quant;
    p.pad = q * hw;
    p.mp = m + 2 * p.pad; p.np = n + 2 * p.pad; p.op = o + 2 * p.pad;
    p.samples = (long long)((step + sy - 1) / sy) * ((step + sx - 1) / sx) * ((step + sz - 1) / sz);
    return p;
}

void launch_pad_codes(uint64_t* Mp, const uint64_t* M, int m, int n, int o, int pad, hipStream_t s) {
    long long total = (long long)(m + 2 * pad) * (n + 2 * pad) * (o + 2 * pad);
    k_pad_codes<<<ceil_div(total, 256), 256, 0, s>>>(Mp, M, m, n, o, pad);
    DEEDS_LAUNCH_CHECK();
}

namespace {

// ---- autotuner of the row-tiled kernel -------------------------------------
// Per shape (step, sample skips, labels, quantisation, control grid) the
// launcher picks the label planes per workgroup (the label slab one workgroup
// evaluates for its row of control points: half a z-slice of planes, one, two
// or three), at 15 and more labels per axis the pipelined or the plain inner
// loop, and, where the quantisation is a multiple of the j skip, whether
// sample rows are kept from plane to plane (the ring), by timing every
// candidate once on the call's own data with HIP events (all candidates
// compute the same exact integer sums, so any of them leaves the right cost
// volume).  The choice is kept for the process.
// DEEDS_TUNE=0 turns tuning off (one slice of planes, pipelined where allowed).
struct RowsChoice {
    int chunk_planes;
    int mode;  // bit 0: pipelined inner loop, bit 1: sample-row ring
    float ms;
};
struct TuneKey {
    int step, sj, sk, len, q, m1, n1, o1;
    bool operator<(const TuneKey& b) const {
        return std::tie(step, sj, sk, len, q, m1, n1, o1) < std::tie(b.step, b.sj, b.sk, b.len, b.q, b.m1, b.n1, b.o1);
    }
};
std::mutex g_tune_mu;
std::map<TuneKey, RowsChoice> g_tuned;
// a caller-pinned tile shape (deeds_set_datacost_tiles): planes per workgroup (0: tune) and pipelining
std::atomic<int> g_pin_planes{0}, g_pin_pipe{1};

bool tuning_enabled() {
    static const bool on = [] {
        const char* e = getenv("DEEDS_TUNE");
        return !(e && e[0] == '0');
    }();
    return on;
}

}  // namespace

void pin_datacost_tiles(int planes, int pipe) {
    g_pin_planes = planes;
    g_pin_pipe = pipe;
}

std::string datacost_tuning_report() {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    std::string out = "[";
    for (const auto& kv : g_tuned) {
        const TuneKey& k = kv.first;
        char buf[256];
        snprintf(buf, sizeof buf,
                 "%s{\"step\": %d, \"labels_per_axis\": %d, \"quant\": %d, \"grid\": [%d, %d, %d], "
                 "\"planes_per_workgroup\": %d, \"pipelined\": %s, \"row_ring\": %s, \"ms\": %.4f}",
                 out.size() > 1 ? ", " : "", k.step, k.len, k.q, k.m1, k.n1, k.o1, kv.second.chunk_planes,
                 (kv.second.mode & 1) ? "true" : "false", (kv.second.mode & 2) ? "true" : "false", kv.second.ms);
        out += buf;
    }
    return out + "]";
}

void launch_datacost(float* cost, const uint64_t* F, const uint64_t* Mp, int m, int n, int o, int step, int hw,
                     int q, const CostPlan& plan, hipStream_t s) {
    int m1 = m / step, n1 = n / step, o1 = o / step;
    long long nodes = (long long)m1 * n1 * o1;
    if (nodes == 0) return;
    int len = 2 * hw + 1;
    const unsigned grid = (unsigned)(8 * ((nodes + 7) / 8));
    const int N = (int)nodes;
    // row-tiled kernel when one thread per row position suffices and the row's
    // moving sample rows fit in LDS
    {
        const int W = m1 * step, WM = W + (len - 1) * q;
        const int RJ = (step + plan.skip_j - 1) / plan.skip_j, RK = (step + plan.skip_k - 1) / plan.skip_k;
        // with the sample-row ring the sums have their own buffer
        const size_t rows_b = (size_t)RJ * RK * WM * sizeof(uint64_t), sums_b = (size_t)len * W * sizeof(unsigned);
        const size_t smem = std::max(rows_b, sums_b), smem_ring = rows_b + sums_b;
        const bool ring_ok = q % plan.skip_j == 0 && q / plan.skip_j < RJ && smem_ring <= 150 * 1024;
        const int threads = std::min(1024, ((W + 63) / 64) * 64);
        const int len2 = len * len;
        // (two staging loads per thread and row cover WM = W + (len-1)q when (len-1)q <= threads)
        const bool rows_ok = plan.skip_i == 1 && W <= 512 && (len - 1) * q <= threads && smem <= 150 * 1024 &&
                             (long long)n1 * o1 * len2 < (1LL << 31) - 8;
        // launches the row kernel with `chunk_planes` label planes per workgroup in form `mode`
        // (bit 0: pipelined, bit 1: sample-row ring, each where the shape allows it); false when
        // no instantiation matches the shape
        auto launch_rows = [&](int chunk_planes, int mode) -> bool {
            const int chunks = (len2 + chunk_planes - 1) / chunk_planes;
            const long long work = (long long)n1 * o1 * chunks;
            const unsigned blocks = (unsigned)(8 * ((work + 7) / 8));
#define DEEDS_DCR_LAUNCH(ST, SJ, SK, LEN, QQ, PIPE, RING)                                                      \
    {                                                                                                         \
        static bool attr_set = false;                                                                         \
        const size_t sm = RING ? smem_ring : smem;                                                            \
        if (sm > 65536 && !attr_set) {                                                                        \
            DEEDS_HIP(hipFuncSetAttribute((const void*)k_datacost_rows<ST, SJ, SK, LEN, QQ, PIPE, RING>,      \
                                          hipFuncAttributeMaxDynamicSharedMemorySize, 150 * 1024));           \
            attr_set = true;                                                                                  \
        }                                                                                                     \
        k_datacost_rows<ST, SJ, SK, LEN, QQ, PIPE, RING><<<blocks, threads, sm, s>>>(                         \
            cost, F, Mp, m, n, m1, n1, plan.scale, plan.mp, plan.np, chunk_planes, chunks, (int)work);         \
        DEEDS_LAUNCH_CHECK();                                                                                 \
        return true;                                                                                          \
    }
#define DEEDS_DCR(ST, SJ, SK, LEN, QQ)                                                                        \
    if (step == ST && plan.skip_j == SJ && plan.skip_k == SK && len == LEN && q == QQ) {                      \
        const bool ring = mode >= 2 && ring_ok;                                                               \
        if constexpr (LEN >= 15) {                                                                            \
            if ((mode & 1) && threads <= 256) {                                                               \
                if (ring) DEEDS_DCR_LAUNCH(ST, SJ, SK, LEN, QQ, true, true)                                   \
                DEEDS_DCR_LAUNCH(ST, SJ, SK, LEN, QQ, true, false)                                            \
            }                                                                                                 \
        }                                                                                                     \
        if (ring) DEEDS_DCR_LAUNCH(ST, SJ, SK, LEN, QQ, false, true)                                          \
        DEEDS_DCR_LAUNCH(ST, SJ, SK, LEN, QQ, false, false)                                                   \
    }
            // the five levels of deedsBCV's default schedule (grid 8..4, radius 8..4,
            // quantisation 5..1); other schedules take the per-node kernels below
            DEEDS_DCR(8, 2, 2, 17, 5) DEEDS_DCR(7, 2, 2, 15, 4) DEEDS_DCR(6, 2, 2, 13, 3)
            DEEDS_DCR(5, 2, 2, 11, 2) DEEDS_DCR(4, 1, 1, 9, 1)
#undef DEEDS_DCR
#undef DEEDS_DCR_LAUNCH
            return false;
        };
        if (rows_ok) {
            const TuneKey key{step, plan.skip_j, plan.skip_k, len, q, m1, n1, o1};
            RowsChoice choice{len, 1, 0.0f};
            bool have = false;
            {
                std::lock_guard<std::mutex> lk(g_tune_mu);
                auto it = g_tuned.find(key);
                if (it != g_tuned.end()) choice = it->second, have = true;
            }
            hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
            DEEDS_HIP(hipStreamIsCapturing(s, &cap));
            if (!have && g_pin_planes.load() == 0 && tuning_enabled() && cap == hipStreamCaptureStatusNone) {
                const int cands[4] = {(len + 1) / 2, len, 2 * len, 3 * len};
                hipEvent_t e0, e1;
                DEEDS_HIP(hipEventCreate(&e0));
                DEEDS_HIP(hipEventCreate(&e1));
                bool any = false;
                for (int mode = 3; mode >= 0; mode--) {
                    if ((mode & 1) && !(len >= 15 && threads <= 256)) continue;
                    if ((mode & 2) && !ring_ok) continue;
                    for (int c : cands) {
                        if (c > len2) continue;
                        // the better of two launches (the first of a shape also pays for cold caches)
                        float best = 0.0f;
                        bool ran = true;
                        for (int rep = 0; rep < 2 && ran; rep++) {
                            DEEDS_HIP(hipEventRecord(e0, s));
                            ran = launch_rows(c, mode);
                            DEEDS_HIP(hipEventRecord(e1, s));
                            DEEDS_HIP(hipEventSynchronize(e1));
                            float ms = 0.0f;
                            DEEDS_HIP(hipEventElapsedTime(&ms, e0, e1));
                            best = rep == 0 ? ms : std::min(best, ms);
                        }
                        if (!ran) break;
                        if (!any || best < choice.ms) choice = RowsChoice{c, mode, best};
                        any = true;
                    }
                }
                (void)hipEventDestroy(e0);
                (void)hipEventDestroy(e1);
                if (any) {
                    std::lock_guard<std::mutex> lk(g_tune_mu);
                    g_tuned[key] = choice;
                }
            }
            const int pin = g_pin_planes.load();
            if (pin > 0 && launch_rows(std::min(pin, len2), g_pin_pipe.load())) return;
            if (launch_rows(choice.chunk_planes, choice.mode)) return;
        }
    }
#define DEEDS_DC(ST, SI, SJ, SK)                                                                              \
    if (step == ST && plan.skip_i == SI && plan.skip_j == SJ && plan.skip_k == SK) {                        \
        k_datacost_t<ST, SI, SJ, SK><<<grid, 256, 0, s>>>(cost, F, Mp, m, n, m1, n1, N, len, q, plan.scale, \
                                                          plan.mp, plan.np);                                \
        DEEDS_LAUNCH_CHECK();                                                                               \
        return;                                                                                             \
    }
    // the schedules of deedsBCV (RAND_SAMPLES = 1: skip 2 in j and k above step 4)
    DEEDS_DC(8, 1, 2, 2) DEEDS_DC(7, 1, 2, 2) DEEDS_DC(6, 1, 2, 2) DEEDS_DC(5, 1, 2, 2)
    DEEDS_DC(4, 1, 1, 1) DEEDS_DC(3, 1, 1, 1) DEEDS_DC(2, 1, 1, 1)
#undef DEEDS_DC
    k_datacost<<<(unsigned)nodes, 256, 0, s>>>(cost, F, Mp, m, n, m1, n1, step, len, q, plan.skip_i, plan.skip_j,
                                               plan.skip_k, plan.scale, plan.mp, plan.np);
    DEEDS_LAUNCH_CHECK();
}

}  // namespace deeds
