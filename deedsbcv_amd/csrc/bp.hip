// Tree belief propagation on gfx950 (reference: regularisation.h:6-303).
//
// Bottom-up pass: the leaves in one launch, then one launch per "round" of
// the tree schedule -- a window of 16 node steps -- with one workgroup per
// segment, a path v_0 .. v_L up the tree whose nodes' other children finished
// in earlier rounds.  The workgroup, for each segment node in turn,
//   - accumulates the node's cost plus its children's messages in `ordered`
//     order (:221-232): the previous segment node's from LDS, the others
//     stored in their cost rows;
//   - stages acc * edgemst (:197-202), runs the three separable
//     min-convolution passes y, x, z with the quadratic ((t - len) + offset)^2
//     in label units and the reference's lowest-index tie rule, writing the
//     node's composite argmin table (:6-113), and keeps msg - min(msg) in LDS
//     for the parent;
//   - at the segment's last node stores that message in the node's cost row.
// Rounds replace the reference's loop over BFS depths (:190-235) with fewer
// global synchronisation points (the kernel boundary makes the messages stored
// in one round visible to the next on every XCD), and the segments follow the
// tree's critical path, so the bottom-up pass takes little more than the
// tree's height in node steps.  The arithmetic per node is the reference's,
// in its order, so messages and argmin tables are bit-identical.
//
// Each 1-D pass keeps the reference's brute-force O(len) loop per output, but
// register-blocked: a thread loads a whole line of the label cube and a window
// of the quadratic table into VGPRs and produces its outputs from registers
// (LEN is a template parameter).
//
// Top-down: one workgroup walks the BFS depths selecting labels (:259-293).
//
// Batched form (launch_bp_batch): the same per-workgroup bodies for several
// trees of one label radius -- the two directions' trees of many pairs -- with
// one launch per round covering every tree's items of that round and one
// selection workgroup per tree; a workgroup finds its tree and item in a work
// list, so each tree's arithmetic is exactly its single-tree arithmetic.
#include "kernels.h"
#include "mst_host.h"

namespace deeds {

namespace {

// Outputs per thread.  The five label counts of the default schedule take
// the value measured fastest on a 2047-node chain (a single workgroup, so the
// message latency is all there is): fewer, fuller waves for 9..15 labels per
// axis (8-12% less time per node at 9, 11 and 13) and two threads per line at
// 17.  Other counts take the smallest OUT that keeps the block within 1024.
constexpr int out_per_thread(int len) {
    if (len == 17) return 9;
    if (len == 15 || len == 13) return 5;
    if (len == 11 || len == 9) return 3;
    int o = 1;
    while (len * len * ((len + o - 1) / o) > 1024) o++;
    return o;
}

template <int LEN>
struct BPShape {
    static constexpr int L2 = LEN * LEN;
    static constexpr int L3 = L2 * LEN;
    static constexpr int L3E = (L3 + 1) & ~1;
    // S = ceil(LEN / OUT) threads share one line, each producing OUT
    // consecutive outputs (L2 * S <= 1024 threads)
    static constexpr int OUT = out_per_thread(LEN);
    static constexpr int S = (LEN + OUT - 1) / OUT;
    static constexpr int BLOCK = ((L2 * S + 63) / 64) * 64;
    static constexpr int ZT = 2 * LEN + 1 + 2 * S;  // quadratic table, padded for the last window
    static constexpr size_t SMEM = 2 * (size_t)L3 * 4 + 2 * (size_t)L3E * 2 + 3 * (size_t)ZT * 4;
};

// One pass over all lines: line `ln` has elements base(ln) + j*STRIDE,
// j = 0..LEN-1; output i of the line goes to base(ln) + i*STRIDE.
// MODE 0: y pass (I1 = j + line*LEN), 1: x pass (I2 = I1[src of j]),
// 2: z pass (I1 = I2[src of j], the composite argmin table).  Returns the
// thread's minimum output.
template <int LEN, int MODE>
__device__ __forceinline__ float dt_pass(const float* __restrict__ src, float* __restrict__ dst,
                                         const float* __restrict__ zt, const int16_t* __restrict__ iin,
                                         int16_t* __restrict__ iout) {
    using B = BPShape<LEN>;
    constexpr int STRIDE = MODE == 0 ? 1 : MODE == 1 ? LEN : B::L2;
    const int t = threadIdx.x;
    const int line = t / B::S;
    const int sub = t - line * B::S;
    float mn = 3.402823466e38f;
    if (line >= B::L2) return mn;
    int base;
    if (MODE == 0) base = line * LEN;
    else if (MODE == 1) base = (line % LEN) + (line / LEN) * B::L2;
    else base = line;
    float v[LEN];
#pragma unroll
    for (int j = 0; j < LEN; j++) v[j] = src[base + j * STRIDE];
    const int i0 = sub * B::OUT;
    // window of the table: zw[q] = zt[i0 + q], so zt[i - j + LEN] = zw[r - j + LEN] for i = i0 + r
    float zw[B::OUT + LEN];
#pragma unroll
    for (int q = 0; q < B::OUT + LEN; q++) zw[q] = zt[i0 + q];
#pragma unroll
    for (int r = 0; r < B::OUT; r++) {
        const int i = i0 + r;
        if (i < LEN) {
            // The strict-< scan in j order keeps the first minimum: the same
            // value as the minimum of all candidates and the lowest j holding
            // it.  Minimum by three-input min, then a reverse scan for the
            // index (candidates are non-negative and finite: no signed zeros
            // or NaNs to tell apart).
            float c[LEN];
#pragma unroll
            for (int j = 0; j < LEN; j++) c[j] = v[j] + zw[r - j + LEN];
            float best = c[0];
#pragma unroll
            for (int j = 1; j + 1 < LEN; j += 2) best = __builtin_fminf(__builtin_fminf(best, c[j]), c[j + 1]);
            if ((LEN & 1) == 0) best = __builtin_fminf(best, c[LEN - 1]);
            int bj = LEN - 1;
#pragma unroll
            for (int j = LEN - 2; j >= 0; j--) bj = c[j] == best ? j : bj;
            const int o = base + i * STRIDE;
            dst[o] = best;
            mn = best < mn ? best : mn;
            if (MODE == 0) iout[o] = (int16_t)(bj + line * LEN);
            else iout[o] = iin[base + bj * STRIDE];
        }
    }
    return mn;
}

// One DPP step of a wave minimum: v = min(v, v of the lane CTRL selects);
// lanes of rows outside ROWS (or without a source lane) keep v.
template <int CTRL, int ROWS>
__device__ __forceinline__ float dpp_min(float v) {
    const float o = __int_as_float(
        __builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), CTRL, ROWS, 0xf, false));
    return o < v ? o : v;
}

// Minimum over the workgroup.  Within a wave by DPP (each lane of a row of 16
// takes its row's minimum by quad swaps and mirrors, then the row broadcasts
// carry rows 0-2 into lane 63) instead of a chain of LDS shuffles; across
// waves through `red` and one barrier.  Every lane of the workgroup must call it.
__device__ __forceinline__ float block_min(float v, float* red) {
    v = dpp_min<0xb1, 0xf>(v);   // quad_perm [1,0,3,2]
    v = dpp_min<0x4e, 0xf>(v);   // quad_perm [2,3,0,1]
    v = dpp_min<0x141, 0xf>(v);  // row_half_mirror
    v = dpp_min<0x140, 0xf>(v);  // row_mirror
    v = dpp_min<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
    v = dpp_min<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3
    v = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) red[wid] = v;
    __syncthreads();
    float r = red[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) r = red[w] < r ? red[w] : r;
    // no second barrier: `red` is written again only after the next message's
    // table and pass barriers, which every thread reaches after this read
    return r;
}

// The schedule (programs, leaves) and the parent offsets are written before
// a BP launch (host copies, k_parent_offsets, k_record_offsets) and only read inside it.  Read
// through the constant address space, their uniform loads become scalar
// loads: a vector load would wait (vmcnt) behind the node's outstanding
// argmin and message stores before the next node's addresses are known.
using cint = const int __attribute__((address_space(4)));
using cfloat = const float __attribute__((address_space(4)));
__device__ __forceinline__ cint* ro(const int* p) { return (cint*)p; }
__device__ __forceinline__ cfloat* ro(const float* p) { return (cfloat*)p; }

struct BPArgs {
    float* cost;
    int16_t* allinds;
    const int* prog_start;
    const int* prog;
    const int* leaves;
    const int* leaf_w;
    const float* off3;  // per node: label-unit offset (x, y, z) of its parent's field
    int root;
};

// Offsets of every node's parent field in label units (regularisation.h:208-210):
// off3[c] = ((u0[p]-u0[c])/q, (v0[p]-v0[c])/q, (w0[p]-w0[c])/q), p = parents[c].
__global__ void k_parent_offsets(float* __restrict__ off3, const int* __restrict__ parents,
                                 const float* __restrict__ u0, const float* __restrict__ v0,
                                 const float* __restrict__ w0, float quant, int N) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    const int p = parents[c];
    if (p < 0) return;
    off3[3 * c + 0] = (u0[p] - u0[c]) / quant;
    off3[3 * c + 1] = (v0[p] - v0[c]) / quant;
    off3[3 * c + 2] = (w0[p] - w0[c]) / quant;
}

// The same offsets written into record i of program block h (fields 4-6),
// where the BP reads them with the rest of the record; one thread per
// (block, record) slot, kBPWindow slots per block.
__device__ __forceinline__ void record_offsets(int* __restrict__ prog, const int* __restrict__ prog_start, int slot,
                                               const int* __restrict__ parents, const float* __restrict__ u0,
                                               const float* __restrict__ v0, const float* __restrict__ w0,
                                               float quant) {
    int* blk = prog + prog_start[slot / kBPWindow];
    const int i = slot % kBPWindow;
    if (i >= blk[0]) return;
    int* r = blk + kProgRec * (1 + i);
    const int c = r[0], p = parents[c];
    if (p < 0) return;
    r[4] = __float_as_int((u0[p] - u0[c]) / quant);
    r[5] = __float_as_int((v0[p] - v0[c]) / quant);
    r[6] = __float_as_int((w0[p] - w0[c]) / quant);
}

__global__ void k_record_offsets(int* __restrict__ prog, const int* __restrict__ prog_start, int slots,
                                 const int* __restrict__ parents, const float* __restrict__ u0,
                                 const float* __restrict__ v0, const float* __restrict__ w0, float quant) {
    const int slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot < slots) record_offsets(prog, prog_start, slot, parents, u0, v0, w0, quant);
}

// The composite argmin table, built in I1 by the z pass, to the node's row of
// allinds: contiguous 16-bit stores, none under a branch (lanes past the row
// store its last entry again, the same value), so a later wait for a load
// counts them exactly instead of waiting for every outstanding store.
template <int LEN>
__device__ __forceinline__ void store_table(int16_t* __restrict__ allinds_c, const int16_t* __restrict__ I1) {
    using B = BPShape<LEN>;
    constexpr int ELEMS = (B::L3 + B::BLOCK - 1) / B::BLOCK;
    const int t = threadIdx.x;
#pragma unroll
    for (int e = 0; e < ELEMS; e++) {
        const int l = (e + 1) * B::BLOCK <= B::L3 ? t + e * B::BLOCK : min(t + e * B::BLOCK, B::L3 - 1);
        allinds_c[l] = I1[l];
    }
}

// One message: A holds the sender's cost scaled by its edge weight.  Runs the
// three passes (the argmin table goes to allinds_c, the message to Bf) and
// returns the message's minimum; every thread may read Bf afterwards, and I1
// holds the table until the next message's first barrier.
template <int LEN>
__device__ __forceinline__ float message(int16_t* allinds_c, float ox, float oy, float oz, float* A, float* Bf,
                                         int16_t* I1, int16_t* I2, float* zy, float* zx, float* zz, float* red) {
    using B = BPShape<LEN>;
    const int t = threadIdx.x;
    for (int q = t; q < B::ZT; q += B::BLOCK) {
        float dy = (float)(q - LEN) + oy;
        float dx = (float)(q - LEN) + ox;
        float dz = (float)(q - LEN) + oz;
        zy[q] = dy * dy;
        zx[q] = dx * dx;
        zz[q] = dz * dz;
    }
    __syncthreads();
    dt_pass<LEN, 0>(A, Bf, zy, nullptr, I1);
    __syncthreads();
    dt_pass<LEN, 1>(Bf, A, zx, I1, I2);
    __syncthreads();
    const float mn = block_min(dt_pass<LEN, 2>(A, Bf, zz, I2, I1), red);  // its barrier also completes Bf and I1
    store_table<LEN>(allinds_c, I1);
    return mn;
}

template <int LEN>
struct BPLds {
    float *A, *Bf, *zy, *zx, *zz;
    int16_t *I1, *I2;
    __device__ explicit BPLds(float* smem) {
        using B = BPShape<LEN>;
        A = smem;
        Bf = A + B::L3;
        I1 = (int16_t*)(Bf + B::L3);
        I2 = I1 + B::L3E;
        zy = (float*)(I2 + B::L3E);
        zx = zy + B::ZT;
        zz = zx + B::ZT;
    }
};

// Round 0: every leaf's message to its parent, stored in the leaf's cost row
// as msg - min(msg), the term its parent folds (regularisation.h:197-232).
// One workgroup per leaf; `idx` is the leaf's position in a.leaves.
template <int LEN>
__device__ __forceinline__ void leaf_body(const BPArgs& a, int idx, float* smem, float* red) {
    using B = BPShape<LEN>;
    const BPLds<LEN> s(smem);
    const int t = threadIdx.x;
    const int v = ro(a.leaves)[idx];
    const float w = __int_as_float(ro(a.leaf_w)[idx]);
    float* row = a.cost + (long long)v * B::L3;
    for (int l = t; l < B::L3; l += B::BLOCK) s.A[l] = row[l] * w;  // costall[child] *= edgemst (:197-202)
    cfloat* off3 = ro(a.off3);
    const float mn = message<LEN>(a.allinds + (long long)v * B::L3, off3[3 * v], off3[3 * v + 1], off3[3 * v + 2],
                                  s.A, s.Bf, s.I1, s.I2, s.zy, s.zx, s.zz, red);
    for (int l = t; l < B::L3; l += B::BLOCK) row[l] = s.Bf[l] - mn;
}

// Element e of thread t's labels, clamped into the row: lanes past the row
// repeat its last entry, so loads, folds and stores need no branch (and the
// repeated stores write the same value as the entry's own lane).
template <int LEN>
__device__ __forceinline__ int lab_of(int e) {
    using B = BPShape<LEN>;
    const int l = (int)threadIdx.x + e * B::BLOCK;
    return (e + 1) * B::BLOCK <= B::L3 ? l : min(l, B::L3 - 1);
}

// One workgroup runs one segment's program block (at a.prog + `off`, layout
// in mst_host.h): the first node's accumulator is its cost plus its
// children's stored messages in `ordered` order (the reference's fold order,
// :221-232); then each segment node messages its parent, the next segment
// node, whose accumulator is its cost plus its children's messages in the
// same order -- the one just computed taken from LDS, the others (finished in
// earlier rounds) from their cost rows -- until the segment's last node,
// which stores its message in its own row -- or, at the root, its
// accumulator.  The accumulator lives in registers (each thread owns fixed
// labels); the next node's cost row and up to PF of its stored child messages
// are loaded while the current message computes.  The block is staged in LDS
// (`blk`, kProgStage ints) once, so a node's record costs an LDS read instead
// of a chain of dependent scalar loads from memory the host wrote.
template <int LEN>
__device__ __forceinline__ void prog_body(const BPArgs& a, int off, float* smem, float* red, int* blk) {
    using B = BPShape<LEN>;
    constexpr int ELEMS = (B::L3 + B::BLOCK - 1) / B::BLOCK;
    // stored child messages prefetched per node: what the registers left
    // beside the passes hold without spilling (none from 17 labels on)
    constexpr int PF = ELEMS >= 8 ? 0 : 2;
    constexpr int SN = (kProgStage + B::BLOCK - 1) / B::BLOCK;
    const BPLds<LEN> s(smem);
    const int t = threadIdx.x;
    auto row = [&](int v) { return a.cost + (long long)v * B::L3; };

    // stage the block (the padding after the last block keeps these in bounds)
    int stage[SN];
#pragma unroll
    for (int q = 0; q < SN; q++)
        if (q * B::BLOCK + t < kProgStage) stage[q] = a.prog[off + q * B::BLOCK + t];
    cint* g = ro(a.prog) + off;
    // entry q of a stored-children list, which may lie past the staged part
    auto child_at = [&](int q) {
        if (q >= kProgStage) return g[q];
        return __builtin_amdgcn_readfirstlane(blk[q]);
    };
    const int L = g[0] - 1;
    int node = g[kProgRec];
    float C[ELEMS];
    {
        // the first node's cost and stored child messages (all its children):
        // all loads issued before the first add (a grid node has at most six)
        constexpr int KMAX = 6;
        const int k0 = g[kProgRec + 2];
        cint* st = g + g[kProgRec + 7];
        float kv[KMAX][ELEMS];
#pragma unroll
        for (int e = 0; e < ELEMS; e++) C[e] = row(node)[lab_of<LEN>(e)];
#pragma unroll
        for (int j = 0; j < KMAX; j++) {
            if (j < k0) {
                const float* kr = row(st[j]);
#pragma unroll
                for (int e = 0; e < ELEMS; e++) kv[j][e] = kr[lab_of<LEN>(e)];
            }
        }
#pragma unroll
        for (int j = 0; j < KMAX; j++) {
            if (j < k0) {
#pragma unroll
                for (int e = 0; e < ELEMS; e++) C[e] += kv[j][e];  // parent += (msg - min), in `ordered` order
            }
        }
        for (int j = KMAX; j < k0; j++) {
            const float* kr = row(st[j]);
#pragma unroll
            for (int e = 0; e < ELEMS; e++) C[e] += kr[lab_of<LEN>(e)];
        }
    }
#pragma unroll
    for (int q = 0; q < SN; q++)
        if (q * B::BLOCK + t < kProgStage) blk[q * B::BLOCK + t] = stage[q];
    __syncthreads();
    // the current record's fields: node, edgemst, parent offsets
    auto rd4 = [&](int q) { return *reinterpret_cast<const int4*>(blk + q); };  // q: a multiple of 4
    int4 h1 = rd4(kProgRec + 4);
    float w = __int_as_float(__builtin_amdgcn_readfirstlane(rd4(kProgRec).y));
    int r = kProgRec;
    for (int i = 0; i <= L; i++) {
        if (node == a.root) break;  // the root keeps its accumulator
        const float ox = __int_as_float(__builtin_amdgcn_readfirstlane(h1.x));
        const float oy = __int_as_float(__builtin_amdgcn_readfirstlane(h1.y));
        const float oz = __int_as_float(__builtin_amdgcn_readfirstlane(h1.z));
        // the next record (read whole: past the last record it is unused), then
        // the next node's cost row and its first PF stored child messages
        const int nr = r + kProgRec;
        const int4 n0 = rd4(nr), n1 = rd4(nr + 4);
        const int next = i < L ? __builtin_amdgcn_readfirstlane(n0.x) : -1;
        const int nk = __builtin_amdgcn_readfirstlane(n0.z), npp = __builtin_amdgcn_readfirstlane(n0.w);
        const int ns = __builtin_amdgcn_readfirstlane(n1.w);
        const int nst = nk - (npp >= 0);  // stored children of the next node
        float nrow[ELEMS], p0[ELEMS], p1[ELEMS];
        if (next >= 0) {
#pragma unroll
            for (int e = 0; e < ELEMS; e++) nrow[e] = row(next)[lab_of<LEN>(e)];
            if (PF > 0 && nst > 0) {
                int c0, c1 = 0;
                if (ns + 1 < kProgStage) {
                    c0 = __builtin_amdgcn_readfirstlane(blk[ns]);
                    c1 = __builtin_amdgcn_readfirstlane(blk[ns + 1]);
                } else {
                    c0 = child_at(ns);
                    if (nst > 1) c1 = child_at(ns + 1);
                }
                const float* kr = row(c0);
#pragma unroll
                for (int e = 0; e < ELEMS; e++) p0[e] = kr[lab_of<LEN>(e)];
                if (PF > 1 && nst > 1) {
                    const float* kr1 = row(c1);
#pragma unroll
                    for (int e = 0; e < ELEMS; e++) p1[e] = kr1[lab_of<LEN>(e)];
                }
            }
        }
#pragma unroll
        for (int e = 0; e < ELEMS; e++) {
            const int l = t + e * B::BLOCK;
            if (l < B::L3) s.A[l] = C[e] * w;  // costall[child] *= edgemst (:197-202)
        }
        const float mn = message<LEN>(a.allinds + (long long)node * B::L3, ox, oy, oz, s.A, s.Bf, s.I1, s.I2,
                                      s.zy, s.zx, s.zz, red);
        if (next < 0) {
            // the segment's last node: store its message for the parent's fold
#pragma unroll
            for (int e = 0; e < ELEMS; e++) {
                const int l = lab_of<LEN>(e);
                row(node)[l] = s.Bf[l] - mn;
            }
            return;
        }
        // the next node's accumulator: its cost, then its children in order
#pragma unroll
        for (int e = 0; e < ELEMS; e++) C[e] = nrow[e];
        for (int j = 0, q = 0; j < nk; j++) {
            if (j == npp) {
#pragma unroll
                for (int e = 0; e < ELEMS; e++) {
                    const int l = lab_of<LEN>(e);
                    C[e] += s.Bf[l] - mn;
                }
                continue;
            }
            if (PF > 0 && q == 0) {
#pragma unroll
                for (int e = 0; e < ELEMS; e++) C[e] += p0[e];
            } else if (PF > 1 && q == 1) {
#pragma unroll
                for (int e = 0; e < ELEMS; e++) C[e] += p1[e];
            } else {
                const float* kr = row(child_at(ns + q));
#pragma unroll
                for (int e = 0; e < ELEMS; e++) C[e] += kr[lab_of<LEN>(e)];
            }
            q++;
        }
        r = nr;
        node = next;
        w = __int_as_float(__builtin_amdgcn_readfirstlane(n0.y));
        h1 = n1;
    }
    // reached the root: its accumulator is what the selection reads
#pragma unroll
    for (int e = 0; e < ELEMS; e++) row(node)[lab_of<LEN>(e)] = C[e];
}

template <int LEN>
__global__ __launch_bounds__(BPShape<LEN>::BLOCK) void k_bp_leaves(BPArgs a) {
    extern __shared__ float smem[];
    __shared__ float red[16];
    leaf_body<LEN>(a, blockIdx.x, smem, red);
}

template <int LEN>
__global__ __launch_bounds__(BPShape<LEN>::BLOCK) void k_bp_prog(BPArgs a, int head_begin) {
    extern __shared__ float smem[];
    __shared__ float red[16];
    __shared__ __align__(16) int blk[kProgStage];
    prog_body<LEN>(a, ro(a.prog_start)[head_begin + blockIdx.x], smem, red, blk);
}

// The spine: heads [head_begin, head_end), one per round, run in order by one
// workgroup (each program reads messages the previous ones stored: the
// barrier makes them visible inside the workgroup).
template <int LEN>
__global__ __launch_bounds__(BPShape<LEN>::BLOCK) void k_bp_tail(BPArgs a, int head_begin, int head_end) {
    extern __shared__ float smem[];
    __shared__ float red[16];
    __shared__ __align__(16) int blk[kProgStage];
    for (int h = head_begin; h < head_end; h++) {
        prog_body<LEN>(a, ro(a.prog_start)[h], smem, red, blk);
        __syncthreads();
    }
}

__device__ __forceinline__ BPArgs args_of(const BPBatchTree& T) {
    return BPArgs{T.cost, T.allinds, T.prog_start, T.prog, T.leaves, T.leaf_w, T.off3, T.root};
}

// Batched forms: workgroup b runs item work[b] = (tree, leaf position or
// program offset) of one round of several trees that share the label count.
template <int LEN>
__global__ __launch_bounds__(BPShape<LEN>::BLOCK) void k_bp_leaves_b(const BPBatchTree* __restrict__ trees,
                                                                     const int2* __restrict__ work) {
    extern __shared__ float smem[];
    __shared__ float red[16];
    const int2 w = work[blockIdx.x];
    leaf_body<LEN>(args_of(trees[w.x]), w.y, smem, red);
}

template <int LEN>
__global__ __launch_bounds__(BPShape<LEN>::BLOCK) void k_bp_prog_b(const BPBatchTree* __restrict__ trees,
                                                                   const int2* __restrict__ work) {
    extern __shared__ float smem[];
    __shared__ float red[16];
    __shared__ __align__(16) int blk[kProgStage];
    const int2 w = work[blockIdx.x];
    prog_body<LEN>(args_of(trees[w.x]), w.y, smem, red, blk);
}

// Spines of several trees: workgroup b runs tree tails[2b].x's programs at the
// offsets prog[tails[2b].y .. tails[2b+1].x) of its head list, in order.
template <int LEN>
__global__ __launch_bounds__(BPShape<LEN>::BLOCK) void k_bp_tail_b(const BPBatchTree* __restrict__ trees,
                                                                   const int2* __restrict__ tails) {
    extern __shared__ float smem[];
    __shared__ float red[16];
    __shared__ __align__(16) int blk[kProgStage];
    const int2 a0 = tails[2 * blockIdx.x], a1 = tails[2 * blockIdx.x + 1];
    const BPArgs a = args_of(trees[a0.x]);
    for (int h = a0.y; h < a1.x; h++) {
        prog_body<LEN>(a, ro(a.prog_start)[h], smem, red, blk);
        __syncthreads();
    }
}

template <int LEN>
void set_round_attrs() {
    constexpr int smem = (int)BPShape<LEN>::SMEM;
    static bool attr_set = false;
    if (attr_set) return;
    const void* fns[6] = {(const void*)k_bp_prog<LEN>,   (const void*)k_bp_leaves<LEN>,
                          (const void*)k_bp_prog_b<LEN>, (const void*)k_bp_leaves_b<LEN>,
                          (const void*)k_bp_tail<LEN>,   (const void*)k_bp_tail_b<LEN>};
    for (const void* f : fns) DEEDS_HIP(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, smem));
    attr_set = true;
}

// kind 0: the leaves; 1: heads [head_begin, head_begin + count), one
// workgroup each; 2: the same heads run in order by one workgroup (the spine)
template <int LEN>
void launch_round_len(const BPArgs& a, int head_begin, int count, int kind, hipStream_t s) {
    using B = BPShape<LEN>;
    constexpr size_t smem = B::SMEM;
    set_round_attrs<LEN>();
    if (kind == 0)
        k_bp_leaves<LEN><<<count, B::BLOCK, smem, s>>>(a);
    else if (kind == 1)
        k_bp_prog<LEN><<<count, B::BLOCK, smem, s>>>(a, head_begin);
    else
        k_bp_tail<LEN><<<1, B::BLOCK, smem, s>>>(a, head_begin, head_begin + count);
    DEEDS_LAUNCH_CHECK();
}

// One batched round: `count` work items starting at `work` (kind 0: leaves,
// 1: heads, 2: `count` tree spines described by pairs of items).
template <int LEN>
void launch_round_len_b(const BPBatchTree* trees, const int2* work, int count, int kind, hipStream_t s) {
    using B = BPShape<LEN>;
    constexpr size_t smem = B::SMEM;
    set_round_attrs<LEN>();
    if (kind == 0)
        k_bp_leaves_b<LEN><<<count, B::BLOCK, smem, s>>>(trees, work);
    else if (kind == 1)
        k_bp_prog_b<LEN><<<count, B::BLOCK, smem, s>>>(trees, work);
    else
        k_bp_tail_b<LEN><<<count, B::BLOCK, smem, s>>>(trees, work);
    DEEDS_LAUNCH_CHECK();
}

// Root argmin (first strict minimum) of the root's accumulated cost, then
// the label selection depth by depth, sel[c] = allinds[c][sel[parent]], with
// the displacement update u1 = label displacement + u0 (regularisation.h:
// 259-293).  The previous depth's labels stay in LDS (indexed by position
// within the depth), so each depth costs one dependent load from HBM, and
// the next depth's (node, parent index) pairs are loaded ahead.
// labs: two LDS buffers of `width` labels; rv, ri: 16 entries each.
__device__ __forceinline__ void select_body(const float* __restrict__ cost, const int16_t* __restrict__ allinds,
                                            const int* __restrict__ ordered, const int* __restrict__ depth_start,
                                            const int* __restrict__ sel_info, int maxd, const float* __restrict__ u0,
                                            const float* __restrict__ v0, const float* __restrict__ w0,
                                            float* __restrict__ u1, float* __restrict__ v1, float* __restrict__ w1,
                                            int width, int len, int hw, float quant, int* labs, float* rv, int* ri) {
    const int len3 = len * len * len;
    const int root = ordered[0];
    const float* rc = cost + (long long)root * len3;
    float bv = 3.402823466e38f;
    int bi = 0x7fffffff;
    for (int l = threadIdx.x; l < len3; l += blockDim.x) {
        float v = rc[l];
        if (v < bv || (v == bv && l < bi)) { bv = v; bi = l; }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        float ov = __shfl_xor(bv, off, 64);
        int oi = __shfl_xor(bi, off, 64);
        if (ov < bv || (ov == bv && oi < bi)) { bv = ov; bi = oi; }
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) { rv[wid] = bv; ri[wid] = bi; }
    __syncthreads();
    int* prev = labs;
    int* cur = labs + width;
    auto emit = [&](int c, int lab) {
        const int li = lab % len, lj = (lab / len) % len, lk = lab / (len * len);
        u1[c] = (float)(lj - hw) * quant + u0[c];
        v1[c] = (float)(li - hw) * quant + v0[c];
        w1[c] = (float)(lk - hw) * quant + w0[c];
    };
    if (threadIdx.x == 0) {
        float v = rv[0];
        int id = ri[0];
        for (int w = 1; w < (int)(blockDim.x >> 6); w++)
            if (rv[w] < v || (rv[w] == v && ri[w] < id)) { v = rv[w]; id = ri[w]; }
        prev[0] = id;
        emit(root, id);
    }
    // Thread t owns position t of every depth (wider depths loop below).  Per
    // depth the chain is one dependent load, allinds[c][label of parent]:
    // the next depth's (node, parent index) and the depth bound after it are
    // loaded a depth ahead, the node's u0/v0/w0 beside the label, and the
    // three field stores wait until the next depth has issued its loads, so
    // no wait for a load also waits for them.
    const int t = threadIdx.x;
    int b = 0, e = 0, E = 0;  // depth_start[d], [d + 1], [d + 2]
    int nc = -1, npp = 0;
    if (maxd >= 1) {
        b = depth_start[1];
        e = depth_start[2];
        E = depth_start[maxd >= 2 ? 3 : 2];
        if (b + t < e) {
            nc = sel_info[2 * (b + t)];
            npp = sel_info[2 * (b + t) + 1];
        }
    }
    int pc = -1;
    float pu = 0.0f, pv = 0.0f, pw = 0.0f;  // outputs of the previous depth's node, stored late
    __syncthreads();
    for (int d = 1; d <= maxd; d++) {
        const int c = nc, pp = npp;
        if (pc >= 0) {
            u1[pc] = pu;
            v1[pc] = pv;
            w1[pc] = pw;
        }
        nc = -1;
        if (d < maxd && e + t < E) {
            nc = sel_info[2 * (e + t)];
            npp = sel_info[2 * (e + t) + 1];
        }
        const int E2 = depth_start[d + 3 <= maxd + 1 ? d + 3 : maxd + 1];
        if (c >= 0) {
            const float a0 = u0[c], a1 = v0[c], a2 = w0[c];
            const int lab = allinds[(long long)c * len3 + prev[pp]];
            cur[t] = lab;
            // 0 <= lab < len^3: unsigned quotients
            const unsigned ul = (unsigned)lab, ulen = (unsigned)len;
            const int lk = (int)(ul / (ulen * ulen)), r2 = (int)(ul - (unsigned)lk * ulen * ulen);
            const int lj = (int)((unsigned)r2 / ulen), li = r2 - lj * len;
            pu = (float)(lj - hw) * quant + a0;
            pv = (float)(li - hw) * quant + a1;
            pw = (float)(lk - hw) * quant + a2;
        }
        pc = c;
        for (int p = b + (int)blockDim.x + t; p < e; p += blockDim.x) {  // depths wider than the block
            const int c2 = sel_info[2 * p], pp2 = sel_info[2 * p + 1];
            const int lab = allinds[(long long)c2 * len3 + prev[pp2]];
            cur[p - b] = lab;
            emit(c2, lab);
        }
        __syncthreads();
        int* tmp = prev;
        prev = cur;
        cur = tmp;
        b = e;
        e = E;
        E = E2;
    }
    if (pc >= 0) {
        u1[pc] = pu;
        v1[pc] = pv;
        w1[pc] = pw;
    }
}

// GLOBAL: the two label rows live in the tree's scratch (sel_labs) instead
// of LDS, for trees with a depth wider than LDS holds.  Only this workgroup
// touches them, and the barrier between depths orders its global stores and
// loads as it does for LDS.
template <bool GLOBAL>
__global__ __launch_bounds__(1024) void k_bp_select(const float* __restrict__ cost,
                                                    const int16_t* __restrict__ allinds,
                                                    const int* __restrict__ ordered,
                                                    const int* __restrict__ depth_start,
                                                    const int* __restrict__ sel_info, int maxd,
                                                    const float* __restrict__ u0, const float* __restrict__ v0,
                                                    const float* __restrict__ w0, float* __restrict__ u1,
                                                    float* __restrict__ v1, float* __restrict__ w1, int width,
                                                    int len, int hw, float quant, int* sel_labs) {
    extern __shared__ int labs[];
    __shared__ float rv[16];
    __shared__ int ri[16];
    select_body(cost, allinds, ordered, depth_start, sel_info, maxd, u0, v0, w0, u1, v1, w1, width, len, hw, quant,
                GLOBAL ? sel_labs : labs, rv, ri);
}

// One workgroup per tree; `width` is the largest depth width over the trees.
template <bool GLOBAL>
__global__ __launch_bounds__(1024) void k_bp_select_b(const BPBatchTree* __restrict__ trees, int width, int len,
                                                      int hw, float quant) {
    extern __shared__ int labs[];
    __shared__ float rv[16];
    __shared__ int ri[16];
    const BPBatchTree& T = trees[blockIdx.x];
    // in the tree's own scratch the second row starts N ints in (no depth of this tree is wider)
    select_body(T.cost, T.allinds, T.ordered, T.depth_start, T.sel_info, T.maxd, T.u0, T.v0, T.w0, T.u1, T.v1,
                T.w1, GLOBAL ? max(1, T.N) : width, len, hw, quant, GLOBAL ? T.sel_labs : labs, rv, ri);
}

// ---- label selection over jumps of kSelJump depths -----------------------
// The selection's chain is one dependent load per BFS depth.  For a node c at
// an anchor depth d (a multiple of J = kSelJump, d >= J) with ancestors
// a_1 .. a_J, the composed table
//   comp[c][x] = allinds[c][allinds[a_1][ ... allinds[a_{J-1}][x] ... ]]
// is the label of c given the label x of a_J (regularisation.h:273-293
// applied J times), so one workgroup walks only the anchor depths; every
// other node then follows in one parallel pass from its anchor's label
// through the tables of the nodes between.  Exact: the same table entries,
// looked up ahead of time.

// label -> displacement field entries u1 = (lj-hw)q + u0, v1 = (li-hw)q + v0,
// w1 = (lk-hw)q + w0 (regularisation.h:281-288), 0 <= lab < len^3
__device__ __forceinline__ void emit_label(int c, int lab, int len, int hw, float quant, const float* __restrict__ u0,
                                           const float* __restrict__ v0, const float* __restrict__ w0,
                                           float* __restrict__ u1, float* __restrict__ v1, float* __restrict__ w1) {
    const unsigned ul = (unsigned)lab, ulen = (unsigned)len;
    const int lk = (int)(ul / (ulen * ulen)), r2 = (int)(ul - (unsigned)lk * ulen * ulen);
    const int lj = (int)((unsigned)r2 / ulen), li = r2 - lj * len;
    u1[c] = (float)(lj - hw) * quant + u0[c];
    v1[c] = (float)(li - hw) * quant + v0[c];
    w1[c] = (float)(lk - hw) * quant + w0[c];
}

// One workgroup per position p >= `first` (the first position at depth J);
// the tables of a_1 .. a_{J-1} are staged in LDS ((J-1) * len3 int16).
__global__ __launch_bounds__(256) void k_compose(int16_t* __restrict__ comp, const int16_t* __restrict__ allinds,
                                                 const int* __restrict__ ordered, const int* __restrict__ parents,
                                                 const int* __restrict__ sel_jump, int first, int len3) {
    extern __shared__ int16_t tabs[];
    const int p = first + blockIdx.x;
    if (sel_jump[p] < 0) return;  // between anchors
    const int c = ordered[p];
    int anc[kSelJump - 1];
    int a = c;
#pragma unroll
    for (int j = 0; j < kSelJump - 1; j++) anc[j] = a = parents[a];
#pragma unroll
    for (int j = 0; j < kSelJump - 1; j++) {
        const int16_t* src = allinds + (long long)anc[j] * len3;
        for (int x = threadIdx.x; x < len3; x += blockDim.x) tabs[j * len3 + x] = src[x];
    }
    __syncthreads();
    const int16_t* ac = allinds + (long long)c * len3;
    int16_t* dst = comp + (long long)c * len3;
    for (int x = threadIdx.x; x < len3; x += blockDim.x) {
        int y = x;
#pragma unroll
        for (int j = kSelJump - 2; j >= 0; j--) y = tabs[j * len3 + y];
        dst[x] = ac[y];
    }
}

// The anchor depths (one workgroup; the structure of select_body with the
// composed tables and the J-th ancestors' indices).  Each anchor's label also
// goes to lab[node] for the fill pass.  labs: two LDS rows of `width` labels.
__global__ __launch_bounds__(1024) void k_bp_select_jump(const float* __restrict__ cost,
                                                         const int16_t* __restrict__ comp,
                                                         const int* __restrict__ ordered,
                                                         const int* __restrict__ depth_start,
                                                         const int* __restrict__ sel_jump, int maxd,
                                                         const float* __restrict__ u0, const float* __restrict__ v0,
                                                         const float* __restrict__ w0, float* __restrict__ u1,
                                                         float* __restrict__ v1, float* __restrict__ w1, int width,
                                                         int len, int hw, float quant, int* __restrict__ lab) {
    extern __shared__ int labs[];
    __shared__ float rv[16];
    __shared__ int ri[16];
    constexpr int J = kSelJump;
    const int len3 = len * len * len;
    const int t = threadIdx.x;
    const int root = ordered[0];
    const float* rc = cost + (long long)root * len3;
    float bv = 3.402823466e38f;
    int bi = 0x7fffffff;
    for (int l = t; l < len3; l += blockDim.x) {
        const float v = rc[l];
        if (v < bv || (v == bv && l < bi)) { bv = v; bi = l; }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const float ov = __shfl_xor(bv, off, 64);
        const int oi = __shfl_xor(bi, off, 64);
        if (ov < bv || (ov == bv && oi < bi)) { bv = ov; bi = oi; }
    }
    if ((t & 63) == 0) { rv[t >> 6] = bv; ri[t >> 6] = bi; }
    __syncthreads();
    int* prev = labs;
    int* cur = labs + width;
    if (t == 0) {
        float v = rv[0];
        int id = ri[0];
        for (int w = 1; w < (int)(blockDim.x >> 6); w++)
            if (rv[w] < v || (rv[w] == v && ri[w] < id)) { v = rv[w]; id = ri[w]; }
        prev[0] = id;
        lab[root] = id;
        emit_label(root, id, len, hw, quant, u0, v0, w0, u1, v1, w1);
    }
    // anchor depth d: positions [b, e); its (node, ancestor index) pairs are loaded a depth ahead
    int b = 0, e = 0;
    int nc = -1, nga = 0;
    if (maxd >= J) {
        b = depth_start[J];
        e = depth_start[J + 1];
        if (b + t < e) {
            nc = ordered[b + t];
            nga = sel_jump[b + t];
        }
    }
    int pc = -1, plab = 0;
    __syncthreads();
    for (int d = J; d <= maxd; d += J) {
        const int c = nc, ga = nga;
        if (pc >= 0) emit_label(pc, plab, len, hw, quant, u0, v0, w0, u1, v1, w1);
        // the next anchor depth d + J: positions [nb, ne)
        int nb = 0, ne = 0;
        if (d + J <= maxd) {
            nb = depth_start[d + J];
            ne = depth_start[d + J + 1];
        }
        nc = -1;
        if (nb + t < ne) {
            nc = ordered[nb + t];
            nga = sel_jump[nb + t];
        }
        if (c >= 0) {
            const int l = comp[(long long)c * len3 + prev[ga]];
            cur[t] = l;
            lab[c] = l;
            plab = l;
        }
        pc = c;
        for (int q = b + (int)blockDim.x + t; q < e; q += blockDim.x) {  // depths wider than the block
            const int c2 = ordered[q];
            const int l = comp[(long long)c2 * len3 + prev[sel_jump[q]]];
            cur[q - b] = l;
            lab[c2] = l;
            emit_label(c2, l, len, hw, quant, u0, v0, w0, u1, v1, w1);
        }
        __syncthreads();
        int* tmp = prev;
        prev = cur;
        cur = tmp;
        b = nb;
        e = ne;
    }
    if (pc >= 0) emit_label(pc, plab, len, hw, quant, u0, v0, w0, u1, v1, w1);
}

// Every node between anchors: its label from its anchor's through the
// tables of the r = -sel_jump nodes from it up to the anchor.
__global__ __launch_bounds__(256) void k_select_fill(const int16_t* __restrict__ allinds,
                                                     const int* __restrict__ ordered,
                                                     const int* __restrict__ parents,
                                                     const int* __restrict__ sel_jump, const int* __restrict__ lab,
                                                     int N, int len, int hw, float quant,
                                                     const float* __restrict__ u0, const float* __restrict__ v0,
                                                     const float* __restrict__ w0, float* __restrict__ u1,
                                                     float* __restrict__ v1, float* __restrict__ w1) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= N) return;
    const int r = -sel_jump[p];
    if (r <= 0) return;  // anchors (the root included) are labelled by the walk
    const int len3 = len * len * len;
    const int c = ordered[p];
    int chain[kSelJump - 1];
    int a = c;
#pragma unroll
    for (int i = 0; i < kSelJump - 1; i++) {
        if (i < r) {
            chain[i] = a;
            a = parents[a];
        }
    }
    int l = lab[a];
#pragma unroll
    for (int i = kSelJump - 2; i >= 0; i--)
        if (i < r) l = allinds[(long long)chain[i] * len3 + l];
    emit_label(c, l, len, hw, quant, u0, v0, w0, u1, v1, w1);
}

// Reads n 16-byte words of argmin tables just before the label selection
// walks them, one dependent load per depth: its loads then come from the
// memory-side cache instead of HBM.  `sink` is written only when `flag` is
// set (never by the launchers), which keeps the loads.
__global__ void k_touch(const int4* __restrict__ p, long long n, int* __restrict__ sink, int flag) {
    int acc = 0;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const int4 v = p[i];
        acc ^= v.x ^ v.y ^ v.z ^ v.w;
    }
    if (flag) sink[0] = acc;
}

void launch_touch(const void* p, size_t bytes, hipStream_t s) {
    if (((uintptr_t)p & 15) != 0 || bytes < 16) return;
    const long long n = (long long)(bytes / 16);
    const int blocks = (int)std::min<long long>(2048, ceil_div(n, 256));
    k_touch<<<blocks, 256, 0, s>>>(static_cast<const int4*>(p), n, nullptr, 0);
    DEEDS_LAUNCH_CHECK();
}

__global__ void k_root_identity(int16_t* __restrict__ allinds, int root, int len3) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l < len3) allinds[(long long)root * len3 + l] = (int16_t)l;
}

// Parent offsets of every tree of a batch (blockIdx.y = tree), per node and
// into the program records.
__global__ void k_parent_offsets_b(const BPBatchTree* __restrict__ trees, float quant) {
    const BPBatchTree& T = trees[blockIdx.y];
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    for (int slot = c; slot < T.n_segments * kBPWindow; slot += gridDim.x * blockDim.x)
        record_offsets(T.prog, T.prog_start, slot, T.parents, T.u0, T.v0, T.w0, quant);
    if (c >= T.N) return;
    const int p = T.parents[c];
    if (p < 0) return;
    T.off3[3 * c + 0] = (T.u0[p] - T.u0[c]) / quant;
    T.off3[3 * c + 1] = (T.v0[p] - T.v0[c]) / quant;
    T.off3[3 * c + 2] = (T.w0[p] - T.w0[c]) / quant;
}

// LDS bytes of the selection's two label rows; 0 when they do not fit (the
// kernel then keeps them in the tree's global scratch).
size_t select_smem(int width) {
    const size_t smem = 2 * (size_t)std::max(1, width) * sizeof(int);
    // beyond the default 64 KiB the kernel needs the attribute (its static LDS takes the rest)
    return smem > (size_t)kMaxLds - 1024 ? 0 : smem;
}

}  // namespace

void launch_bp_rounds(float* cost, int16_t* allinds, float* off3, const BPDevTree& t, int N,
                      const int* round_start, int rounds, int tail_round, const float* u0, const float* v0,
                      const float* w0, int hw, float quant, hipStream_t s) {
    if (N > 0) {
        k_parent_offsets<<<ceil_div(N, 256), 256, 0, s>>>(off3, t.parents, u0, v0, w0, quant, N);
        DEEDS_LAUNCH_CHECK();
    }
    if (t.n_segments > 0) {
        const int slots = t.n_segments * kBPWindow;
        k_record_offsets<<<ceil_div(slots, 256), 256, 0, s>>>(t.prog, t.prog_start, slots, t.parents, u0, v0, w0,
                                                              quant);
        DEEDS_LAUNCH_CHECK();
    }
    BPArgs a{cost, allinds, t.prog_start, t.prog, t.leaves, t.leaf_w, off3, t.root};
    tail_round = std::max(0, std::min(tail_round, rounds));
    // r = -1: the leaves; r < tail_round: one launch per round; then the spine in one launch
    for (int r = -1; r <= tail_round; r++) {
        int kind, b, n;
        if (r < 0) {
            kind = 0, b = 0, n = t.n_leaves;
        } else if (r < tail_round) {
            kind = 1, b = round_start[r], n = round_start[r + 1] - b;
        } else {
            kind = 2, b = round_start[tail_round], n = round_start[rounds] - b;
        }
        if (n <= 0) continue;
        switch (2 * hw + 1) {
            case 1: launch_round_len<1>(a, b, n, kind, s); break;
            case 3: launch_round_len<3>(a, b, n, kind, s); break;
            case 5: launch_round_len<5>(a, b, n, kind, s); break;
            case 7: launch_round_len<7>(a, b, n, kind, s); break;
            case 9: launch_round_len<9>(a, b, n, kind, s); break;
            case 11: launch_round_len<11>(a, b, n, kind, s); break;
            case 13: launch_round_len<13>(a, b, n, kind, s); break;
            case 15: launch_round_len<15>(a, b, n, kind, s); break;
            case 17: launch_round_len<17>(a, b, n, kind, s); break;
            case 19: launch_round_len<19>(a, b, n, kind, s); break;
            case 21: launch_round_len<21>(a, b, n, kind, s); break;
            default: fail(DEEDS_ERR_ARG, "search radius outside [0, 10]");
        }
    }
}

void launch_bp_select(const float* cost, const int16_t* allinds, int16_t* comp, int first_anchor,
                      const BPDevTree& t, int N, int maxd, const float* u0, const float* v0, const float* w0,
                      float* u1, float* v1, float* w1, int hw, float quant, hipStream_t s) {
    int len = 2 * hw + 1;
    const int width = std::max(1, t.max_width);
    const size_t smem = select_smem(width);
    const int len3 = len * len * len;
    if (comp && smem > 0 && t.sel_labs && N > 0) {
        // walk the anchor depths on composed tables, then every other depth in one parallel pass
        if (first_anchor < N) {
            const size_t tsmem = (size_t)(kSelJump - 1) * len3 * sizeof(int16_t);
            if (tsmem > 65536)
                DEEDS_HIP(hipFuncSetAttribute((const void*)k_compose, hipFuncAttributeMaxDynamicSharedMemorySize,
                                              (int)tsmem));
            k_compose<<<N - first_anchor, 256, tsmem, s>>>(comp, allinds, t.ordered, t.parents, t.sel_jump,
                                                           first_anchor, len3);
            DEEDS_LAUNCH_CHECK();
        }
        if (smem > 65536)
            DEEDS_HIP(hipFuncSetAttribute((const void*)k_bp_select_jump, hipFuncAttributeMaxDynamicSharedMemorySize,
                                          (int)smem));
        k_bp_select_jump<<<1, 1024, smem, s>>>(cost, comp, t.ordered, t.depth_start, t.sel_jump, maxd, u0, v0, w0,
                                               u1, v1, w1, width, len, hw, quant, t.sel_labs);
        DEEDS_LAUNCH_CHECK();
        k_select_fill<<<ceil_div(N, 256), 256, 0, s>>>(allinds, t.ordered, t.parents, t.sel_jump, t.sel_labs, N,
                                                       len, hw, quant, u0, v0, w0, u1, v1, w1);
        DEEDS_LAUNCH_CHECK();
        return;
    }
    launch_touch(allinds, (size_t)N * len3 * sizeof(int16_t), s);
    if (smem == 0) {
        if (!t.sel_labs || 2LL * width > 2LL * std::max(1, N))
            fail(DEEDS_ERR_INTERNAL, "label selection: no scratch for a depth too wide for LDS");
        k_bp_select<true><<<1, 1024, 0, s>>>(cost, allinds, t.ordered, t.depth_start, t.sel_info, maxd, u0, v0, w0,
                                             u1, v1, w1, width, len, hw, quant, t.sel_labs);
    } else {
        if (smem > 65536)
            DEEDS_HIP(hipFuncSetAttribute((const void*)k_bp_select<false>,
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
        k_bp_select<false><<<1, 1024, smem, s>>>(cost, allinds, t.ordered, t.depth_start, t.sel_info, maxd, u0, v0,
                                                 w0, u1, v1, w1, width, len, hw, quant, nullptr);
    }
    DEEDS_LAUNCH_CHECK();
}

void launch_root_identity(int16_t* allinds, int root, int len3, hipStream_t s) {
    k_root_identity<<<ceil_div(len3, 256), 256, 0, s>>>(allinds, root, len3);
    DEEDS_LAUNCH_CHECK();
}

void launch_bp_batch(const BPBatchTree* d_trees, int trees, int max_n, const int2* work,
                     const std::vector<int>& round_off, int width, int hw, float quant, hipStream_t s) {
    if (trees <= 0) return;
    if (max_n > 0) {
        k_parent_offsets_b<<<dim3(ceil_div(max_n, 256), trees), 256, 0, s>>>(d_trees, quant);
        DEEDS_LAUNCH_CHECK();
    }
    for (size_t r = 0; r + 1 < round_off.size(); r++) {
        const int b = round_off[r];
        // the last range holds the spines as pairs of items (one workgroup per pair)
        const bool spines = r + 2 == round_off.size();
        const int n = spines ? (round_off[r + 1] - b) / 2 : round_off[r + 1] - b;
        if (n <= 0) continue;
        const int kind = r == 0 ? 0 : spines ? 2 : 1;
        switch (2 * hw + 1) {
            case 1: launch_round_len_b<1>(d_trees, work + b, n, kind, s); break;
            case 3: launch_round_len_b<3>(d_trees, work + b, n, kind, s); break;
            case 5: launch_round_len_b<5>(d_trees, work + b, n, kind, s); break;
            case 7: launch_round_len_b<7>(d_trees, work + b, n, kind, s); break;
            case 9: launch_round_len_b<9>(d_trees, work + b, n, kind, s); break;
            case 11: launch_round_len_b<11>(d_trees, work + b, n, kind, s); break;
            case 13: launch_round_len_b<13>(d_trees, work + b, n, kind, s); break;
            case 15: launch_round_len_b<15>(d_trees, work + b, n, kind, s); break;
            case 17: launch_round_len_b<17>(d_trees, work + b, n, kind, s); break;
            case 19: launch_round_len_b<19>(d_trees, work + b, n, kind, s); break;
            case 21: launch_round_len_b<21>(d_trees, work + b, n, kind, s); break;
            default: fail(DEEDS_ERR_ARG, "search radius outside [0, 10]");
        }
    }
    // each tree's scratch holds 2*N ints, and a depth is at most N wide
    const size_t smem = select_smem(width);
    if (smem == 0) {
        k_bp_select_b<true><<<trees, 1024, 0, s>>>(d_trees, std::max(1, width), 2 * hw + 1, hw, quant);
    } else {
        if (smem > 65536)
            DEEDS_HIP(hipFuncSetAttribute((const void*)k_bp_select_b<false>,
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
        k_bp_select_b<false><<<trees, 1024, smem, s>>>(d_trees, std::max(1, width), 2 * hw + 1, hw, quant);
    }
    DEEDS_LAUNCH_CHECK();
}

}  // namespace deeds
