// Per-label Dice overlap of two label volumes on gfx950 (reference:
// diceMulti.cpp:17-79).
//
// Two HBM-bound passes over the int16 volumes:
//   k_label_presence  which values occur in the first volume (a 2^16-bit
//                     bitmap, built per workgroup in LDS, merged with
//                     atomicOr of the non-zero words)
//   k_dice_counts     |A_l|, |B_l| and |A_l & B_l| for the K labels of the
//                     first volume (the host maps each int16 value to its
//                     label slot), counted per workgroup in LDS and merged
//                     with one atomicAdd per slot
// The host sorts nothing: the bitmap is already in value order.  Dice and the
// mean are formed on the host in float, as diceMulti.cpp:50-62 does.
#include "kernels.h"

#include <mutex>
#include <vector>

namespace deeds {

namespace {

constexpr int kValues = 1 << 16;         // int16 values, offset by 32768
constexpr int kWords = kValues / 32;     // presence bitmap words
constexpr int kLdsLabels = 4096;         // labels counted in LDS (3 counters each)
constexpr int kThreads = 256;

__global__ __launch_bounds__(kThreads) void k_label_presence(const int16_t* __restrict__ seg, long long n,
                                                             unsigned* __restrict__ bits) {
    __shared__ unsigned sb[kWords];
    for (int w = threadIdx.x; w < kWords; w += kThreads) sb[w] = 0u;
    __syncthreads();
    const long long stride = (long long)gridDim.x * kThreads;
    for (long long v = (long long)blockIdx.x * kThreads + threadIdx.x; v < n; v += stride) {
        const int u = (int)seg[v] + 32768;
        atomicOr(&sb[u >> 5], 1u << (u & 31));
    }
    __syncthreads();
    for (int w = threadIdx.x; w < kWords; w += kThreads)
        if (sb[w]) atomicOr(&bits[w], sb[w]);
}

// counts[3*k + 0] = |A_k & B_k|, [3*k + 1] = |A_k|, [3*k + 2] = |B_k| for the
// label in slot k = slot[value + 32768] (-1: not a label of the first volume).
template <bool LDS>
__global__ __launch_bounds__(kThreads) void k_dice_counts(const int16_t* __restrict__ a,
                                                          const int16_t* __restrict__ b, long long n,
                                                          const int* __restrict__ slot, int labels,
                                                          unsigned* __restrict__ counts) {
    __shared__ unsigned sc[LDS ? 3 * kLdsLabels : 1];
    unsigned* c = LDS ? sc : counts;
    if (LDS) {
        for (int k = threadIdx.x; k < 3 * labels; k += kThreads) sc[k] = 0u;
        __syncthreads();
    }
    const long long stride = (long long)gridDim.x * kThreads;
    for (long long v = (long long)blockIdx.x * kThreads + threadIdx.x; v < n; v += stride) {
        const int va = a[v], vb = b[v];
        const int ka = slot[va + 32768], kb = slot[vb + 32768];
        if (ka >= 0) {
            atomicAdd(&c[3 * ka + 1], 1u);
            if (va == vb) atomicAdd(&c[3 * ka + 0], 1u);
        }
        if (kb >= 0) atomicAdd(&c[3 * kb + 2], 1u);
    }
    if (LDS) {
        __syncthreads();
        for (int k = threadIdx.x; k < 3 * labels; k += kThreads)
            if (sc[k]) atomicAdd(&counts[k], sc[k]);
    }
}

unsigned grid_for(long long n) {
    const long long want = (n + 8LL * kThreads - 1) / (8LL * kThreads);  // about 8 voxels per thread
    return (unsigned)std::max(1LL, std::min(want, 4096LL));
}

// ---- the allocation-free form: every intermediate in `scratch` -------------
//
// Layout of the file-scope scratch array (one per GPU, kept until the process
// ends): the presence bitmap, the value -> label slot table, three counters
// per label and the label count.  Every int16 label set fits, so the array
// has a fixed size of about 1 MiB.
constexpr size_t kScrBits = 0;                                      // kWords unsigned
constexpr size_t kScrSlot = kScrBits + kWords * sizeof(unsigned);   // kValues int
constexpr size_t kScrCounts = kScrSlot + kValues * sizeof(int);      // 3 * kValues unsigned
constexpr size_t kScrK = kScrCounts + 3 * (size_t)kValues * sizeof(unsigned);  // one int
constexpr size_t kScratchBytes = (kScrK + sizeof(int) + 255) & ~(size_t)255;

__global__ void k_fill_u32(unsigned* __restrict__ p, long long n, unsigned v) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

// One workgroup of 1024 threads, two bitmap words each: the rank of every
// present value among the present values (its label slot, in increasing value
// order as diceMulti.cpp:38-40 sorts them), -1 for absent values; the labels
// after the smallest go to labels_out[0 .. capacity); *K = their number + 1.
__global__ __launch_bounds__(1024) void k_label_slots(const unsigned* __restrict__ bits, int* __restrict__ slot,
                                                      int* __restrict__ K, int16_t* __restrict__ labels_out,
                                                      int capacity) {
    __shared__ int wsum[16];
    const int t = threadIdx.x;
    const unsigned w0 = bits[2 * t], w1 = bits[2 * t + 1];
    const int c = __popc(w0) + __popc(w1);
    // inclusive scan of c over the block: within each wave, then over the 16 waves
    int x = c;
    const int lane = t & 63, wid = t >> 6;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) wsum[wid] = x;
    __syncthreads();
    if (t == 0) {
        int acc = 0;
        for (int w = 0; w < 16; w++) {
            const int v = wsum[w];
            wsum[w] = acc;
            acc += v;
        }
        *K = acc;
    }
    __syncthreads();
    int r = wsum[wid] + x - c;  // present values below this thread's first value
    for (int h = 0; h < 2; h++) {
        const unsigned w = h ? w1 : w0;
        for (int b = 0; b < 32; b++) {
            const int u = 64 * t + 32 * h + b;
            if ((w >> b) & 1u) {
                slot[u] = r;
                if (r >= 1 && r - 1 < capacity) labels_out[r - 1] = (int16_t)(u - 32768);
                r++;
            } else {
                slot[u] = -1;
            }
        }
    }
}

// As k_dice_counts, with the label count read from the device: LDS counters
// when it fits, global atomics otherwise (a uniform branch).
__global__ __launch_bounds__(kThreads) void k_dice_counts_dev(const int16_t* __restrict__ a,
                                                              const int16_t* __restrict__ b, long long n,
                                                              const int* __restrict__ slot, const int* __restrict__ K,
                                                              unsigned* __restrict__ counts) {
    __shared__ unsigned sc[3 * kLdsLabels];
    const int labels = *K;
    const bool lds = labels <= kLdsLabels;
    unsigned* c = lds ? sc : counts;
    if (lds) {
        for (int k = threadIdx.x; k < 3 * labels; k += kThreads) sc[k] = 0u;
        __syncthreads();
    }
    const long long stride = (long long)gridDim.x * kThreads;
    for (long long v = (long long)blockIdx.x * kThreads + threadIdx.x; v < n; v += stride) {
        const int va = a[v], vb = b[v];
        const int ka = slot[va + 32768], kb = slot[vb + 32768];
        if (ka >= 0) {
            atomicAdd(&c[3 * ka + 1], 1u);
            if (va == vb) atomicAdd(&c[3 * ka + 0], 1u);
        }
        if (kb >= 0) atomicAdd(&c[3 * kb + 2], 1u);
    }
    if (lds) {
        __syncthreads();
        for (int k = threadIdx.x; k < 3 * labels; k += kThreads)
            if (sc[k]) atomicAdd(&counts[k], sc[k]);
    }
}

// dice_out[k-1] = 2|A_k & B_k| / (|A_k| + |B_k|) in float for the labels after
// the smallest (diceMulti.cpp:44-55), and *count_out = their number.
__global__ void k_dice_final(const unsigned* __restrict__ counts, const int* __restrict__ K,
                             float* __restrict__ dice_out, int* __restrict__ count_out, int capacity) {
    const int labels = *K;
    const int k = 1 + blockIdx.x * blockDim.x + threadIdx.x;
    if (blockIdx.x == 0 && threadIdx.x == 0) *count_out = labels > 0 ? labels - 1 : 0;
    if (k >= labels || k - 1 >= capacity) return;
    const float inter = (float)counts[3 * k], c1 = (float)counts[3 * k + 1], c2 = (float)counts[3 * k + 2];
    dice_out[k - 1] = __fdiv_rn(2.0f * inter, c1 + c2);
}

// The scratch array and its address on each device (hipGetSymbolAddress once
// per device, with that device current; host code never uses the array's name
// as an address).
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::once_flag g_scratch_once[kMaxDevices];
void* g_scratch_ptr[kMaxDevices] = {};
hipError_t g_scratch_err[kMaxDevices] = {};

unsigned char* scratch_on(int device) {
    if (device < 0 || device >= kMaxDevices) fail(DEEDS_ERR_ARG, "device index beyond the scratch table");
    std::call_once(g_scratch_once[device], [device] {
        g_scratch_err[device] = hipSetDevice(device);
        if (g_scratch_err[device] == hipSuccess)
            g_scratch_err[device] = hipGetSymbolAddress(&g_scratch_ptr[device], HIP_SYMBOL(scratch));
    });
    if (g_scratch_err[device] != hipSuccess || !g_scratch_ptr[device])
        fail(DEEDS_ERR_HIP, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(g_scratch_err[device]));
    return static_cast<unsigned char*>(g_scratch_ptr[device]);
}

}  // namespace

void dice_device(deeds_ctx* ctx, const int16_t* d_a, const int16_t* d_b, long long n, std::vector<int16_t>& labels,
                 std::vector<float>& dice) {
    hipStream_t s = ctx->stream;
    // presence of every value of the first volume
    unsigned* d_bits = ctx->get<unsigned>("dice.bits", kWords);
    int* d_slot = ctx->get<int>("dice.slot", kValues);
    unsigned* h_bits = static_cast<unsigned*>(ctx->pinned("dice.bits", kWords * sizeof(unsigned)));
    int* h_slot = static_cast<int*>(ctx->pinned("dice.slot", kValues * sizeof(int)));
    // cleared by copying zeros from the host (8 KiB)
    for (int w = 0; w < kWords; w++) h_bits[w] = 0u;
    DEEDS_HIP(hipMemcpyAsync(d_bits, h_bits, kWords * sizeof(unsigned), hipMemcpyHostToDevice, s));
    if (n > 0) {
        k_label_presence<<<grid_for(n), kThreads, 0, s>>>(d_a, n, d_bits);
        DEEDS_LAUNCH_CHECK();
    }
    DEEDS_HIP(hipMemcpyAsync(h_bits, d_bits, kWords * sizeof(unsigned), hipMemcpyDeviceToHost, s));
    DEEDS_HIP(hipStreamSynchronize(s));
    // labels in increasing value order (diceMulti.cpp:38-40: sort + unique)
    std::vector<int16_t> found;
    for (int u = 0; u < kValues; u++) {
        const bool present = (h_bits[u >> 5] >> (u & 31)) & 1u;
        h_slot[u] = present ? (int)found.size() : -1;
        if (present) found.push_back((int16_t)(u - 32768));
    }
    const int K = (int)found.size();
    labels.clear();
    dice.clear();
    if (K == 0) return;
    unsigned* d_counts = ctx->get<unsigned>("dice.counts", 3 * (size_t)K);
    unsigned* h_counts = static_cast<unsigned*>(ctx->pinned("dice.counts", 3 * (size_t)K * sizeof(unsigned)));
    for (int k = 0; k < 3 * K; k++) h_counts[k] = 0u;
    DEEDS_HIP(hipMemcpyAsync(d_counts, h_counts, 3 * (size_t)K * sizeof(unsigned), hipMemcpyHostToDevice, s));
    DEEDS_HIP(hipMemcpyAsync(d_slot, h_slot, kValues * sizeof(int), hipMemcpyHostToDevice, s));
    if (K <= kLdsLabels)
        k_dice_counts<true><<<grid_for(n), kThreads, 0, s>>>(d_a, d_b, n, d_slot, K, d_counts);
    else
        k_dice_counts<false><<<grid_for(n), kThreads, 0, s>>>(d_a, d_b, n, d_slot, K, d_counts);
    DEEDS_LAUNCH_CHECK();
    DEEDS_HIP(hipMemcpyAsync(h_counts, d_counts, 3 * (size_t)K * sizeof(unsigned), hipMemcpyDeviceToHost, s));
    DEEDS_HIP(hipStreamSynchronize(s));
    // every label but the smallest (diceMulti.cpp:44: the loop starts at 1)
    for (int k = 1; k < K; k++) {
        const float inter = (float)h_counts[3 * k], c1 = (float)h_counts[3 * k + 1], c2 = (float)h_counts[3 * k + 2];
        labels.push_back(found[k]);
        dice.push_back(2.0f * inter / (c1 + c2));  // diceMulti.cpp:50-55
    }
}

}  // namespace deeds

using namespace deeds;

namespace {

int copy_out(const std::vector<int16_t>& labels, const std::vector<float>& dice, int capacity, int16_t* lab_out,
             float* dice_out, int* count) {
    *count = (int)labels.size();
    const int k = std::min(capacity, (int)labels.size());
    for (int i = 0; i < k; i++) {
        if (lab_out) lab_out[i] = labels[i];
        if (dice_out) dice_out[i] = dice[i];
    }
    return DEEDS_OK;
}

}  // namespace

extern "C" {

int deeds_ctx_dice(deeds_ctx* ctx, const int16_t* d_seg1, const int16_t* d_seg2, long long voxels, int capacity,
                   int16_t* labels, float* dice, int* count) {
    if (!ctx) return DEEDS_ERR_ARG;
    return guarded(ctx, [&] {
        DEEDS_REQUIRE(count && capacity >= 0 && voxels >= 0, "bad arguments");
        DEEDS_REQUIRE(voxels == 0 || (d_seg1 && d_seg2), "null pointer");
        DEEDS_REQUIRE(voxels < (1LL << 32), "volume too large (counts are 32-bit)");
        std::vector<int16_t> lab;
        std::vector<float> d;
        dice_device(ctx, d_seg1, d_seg2, voxels, lab, d);
        copy_out(lab, d, capacity, labels, dice, count);
    });
}

int deeds_ctx_dice_noalloc(deeds_ctx* ctx, const int16_t* d_seg1, const int16_t* d_seg2, long long voxels,
                           int capacity, int16_t* d_labels, float* d_dice, int* d_count) {
    if (!ctx) return DEEDS_ERR_ARG;
    return guarded(ctx, [&] {
        DEEDS_REQUIRE(d_count && capacity >= 0 && voxels >= 0, "bad arguments");
        DEEDS_REQUIRE(capacity == 0 || (d_labels && d_dice), "null output pointer");
        DEEDS_REQUIRE(voxels == 0 || (d_seg1 && d_seg2), "null pointer");
        DEEDS_REQUIRE(voxels < (1LL << 32), "volume too large: the scratch counters are 32-bit (at most 2^32-1 voxels)");
        DEEDS_REQUIRE(capacity <= kValues, "capacity beyond the 65536 int16 labels");
        unsigned char* scr = scratch_on(ctx->device);
        hipStream_t s = ctx->stream;
        unsigned* bits = reinterpret_cast<unsigned*>(scr + kScrBits);
        int* slot = reinterpret_cast<int*>(scr + kScrSlot);
        unsigned* counts = reinterpret_cast<unsigned*>(scr + kScrCounts);
        int* K = reinterpret_cast<int*>(scr + kScrK);
        // the array keeps the last call's contents: clear the bitmap and the counters
        k_fill_u32<<<1, 1024, 0, s>>>(bits, kWords, 0u);
        k_fill_u32<<<192, 256, 0, s>>>(counts, 3LL * kValues, 0u);
        DEEDS_LAUNCH_CHECK();
        if (voxels > 0) {
            k_label_presence<<<grid_for(voxels), kThreads, 0, s>>>(d_seg1, voxels, bits);
            DEEDS_LAUNCH_CHECK();
        }
        k_label_slots<<<1, 1024, 0, s>>>(bits, slot, K, d_labels, capacity);
        DEEDS_LAUNCH_CHECK();
        if (voxels > 0) {
            k_dice_counts_dev<<<grid_for(voxels), kThreads, 0, s>>>(d_seg1, d_seg2, voxels, slot, K, counts);
            DEEDS_LAUNCH_CHECK();
        }
        k_dice_final<<<std::max(1, (int)ceil_div(std::max(capacity, 1), 256)), 256, 0, s>>>(counts, K, d_dice,
                                                                                              d_count, capacity);
        DEEDS_LAUNCH_CHECK();
    });
}

int deeds_dice(const int16_t* seg1, const int16_t* seg2, long long voxels, int capacity, int16_t* labels, float* dice,
               int* count) {
    return guarded(nullptr, [&] {
        DEEDS_REQUIRE(count && capacity >= 0 && voxels >= 0, "bad arguments");
        DEEDS_REQUIRE(voxels == 0 || (seg1 && seg2), "null pointer");
        deeds_ctx* c = default_ctx();
        DEEDS_HIP(hipSetDevice(c->device));
        int16_t* d = c->get<int16_t>("h.dice", 2 * (size_t)std::max(1LL, voxels));
        if (voxels > 0) {
            h2d(d, seg1, (size_t)voxels, c->stream);
            h2d(d + voxels, seg2, (size_t)voxels, c->stream);
        }
        const int rc = deeds_ctx_dice(c, d, d + voxels, voxels, capacity, labels, dice, count);
        if (rc) fail(rc, c->err);
    });
}

}  // extern "C"
