// Host half of the MST stage (reference: primsMST.h:101-223).
//
// The edge weights need libm's expf and the tree needs the exact tie-breaking
// of the reference's heap: Prim's with a lazy binary heap driven by the C++
// standard library's std::push_heap / std::pop_heap and the reference's edge
// order (lighter weight = higher priority, primsMST.h:24-26).  Running the
// same standard algorithms with the same comparisons reproduces the same tree,
// including among equal weights, which are common (integer-valued SADs).
#include "mst_host.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <thread>

namespace deeds {

namespace {

// Lazy Prim heap with exactly the sift rules of the C++ standard library's
// std::push_heap / std::pop_heap (libstdc++ __push_heap / __adjust_heap), the
// containers the reference's Prim's uses, ordered by the reference's
// Edge::operator< (primsMST.h:24-26: a < b iff a.weight > b.weight, so the
// top is the lightest edge).  Ties among equal weights resolve exactly as
// in the reference only because the sequence of sifts is the same, and that
// sequence depends only on how keys compare: any key type whose order and
// ties are those of the reference's weights gives the same tree.  Keys and
// edge codes (from*8 + neighbour slot) are separate arrays, so the hole walk
// of a pop, which compares keys only, touches as little memory as possible:
// float keys, or 16-bit order-preserving ranks of them (about a third faster
// on grids of 10^5 nodes, whose heap then stays in L2).
template <typename Key>
class EdgeHeap {
   public:
    explicit EdgeHeap(size_t reserve) : k_(reserve + 16), c_(reserve + 16) {}
    bool empty() const { return n_ == 0; }
    void push(Key w, uint32_t code) {
        if (n_ == k_.size()) {
            k_.resize(2 * n_);
            c_.resize(2 * n_);
        }
        sift_up((long)n_++, w, code);
    }
    // removes the top (lightest) edge and returns its code
    uint32_t pop() {
        Key* K = k_.data();
        uint32_t* C = c_.data();
        const uint32_t top = C[0];
        const long n = (long)n_ - 1;
        const Key vk = K[n];
        const uint32_t vc = C[n];
        K[n] = K[0];
        C[n] = C[0];
        // __adjust_heap(first, 0, n, value): the hole walks down along the
        // child that is not "less" (comp(child, child-1) <=> child.w > (child-1).w)
        long hole = 0, child = 0;
        const long lim = (n - 1) / 2;
        // Two levels per step while both children of the hole are below lim
        // (so the walk takes both): the four grandchildren are loaded with
        // the two children, so one load latency covers two levels.  The same
        // comparisons in the same order as the one-level walk below.
        while (2 * (child + 1) < lim) {
            const long r = 2 * (child + 1);
            __builtin_prefetch(K + 16 * child + 16);
            const Key a = K[r], b = K[r - 1];
            const Key g0 = K[2 * r - 1], g1 = K[2 * r], g2 = K[2 * r + 1], g3 = K[2 * r + 2];
            const bool left = a > b;  // the hole takes the left child
            const long c1 = r - (long)left;
            const Key k1 = left ? b : a;
            const Key ra = left ? g1 : g3, rb = left ? g0 : g2;  // c1's right and left children
            const bool left2 = ra > rb;
            const long c2 = 2 * (c1 + 1) - (long)left2;
            K[hole] = k1;
            C[hole] = C[c1];
            K[c1] = left2 ? rb : ra;
            C[c1] = C[c2];
            hole = child = c2;
        }
        while (child < lim) {
            child = 2 * (child + 1);
            // the children pairs two, three and four levels below
            __builtin_prefetch(K + 4 * child - 1);
            __builtin_prefetch(K + 8 * child - 9);
            __builtin_prefetch(K + 16 * child - 25);
            const Key a = K[child], b = K[child - 1];
            child -= (long)(a > b);  // branch-free: the hole takes the child that is not "less"
            K[hole] = K[child];
            C[hole] = C[child];
            hole = child;
        }
        if ((n & 1) == 0 && child == (n - 2) / 2) {
            child = 2 * (child + 1);
            K[hole] = K[child - 1];
            C[hole] = C[child - 1];
            hole = child - 1;
        }
        sift_up(hole, vk, vc);
        n_--;
        return top;
    }

   private:
    // __push_heap(first, hole, 0, value): move parents down while comp(parent, value)
    void sift_up(long hole, Key w, uint32_t code) {
        Key* K = k_.data();
        uint32_t* C = c_.data();
        long parent = (hole - 1) / 2;
        while (hole > 0 && K[parent] > w) {
            K[hole] = K[parent];
            C[hole] = C[parent];
            hole = parent;
            parent = (hole - 1) / 2;
        }
        K[hole] = w;
        C[hole] = code;
    }
    std::vector<Key> k_;
    std::vector<uint32_t> c_;
    size_t n_ = 0;
};

const int kNbI[6] = {0, 0, -1, 1, 0, 0};
const int kNbJ[6] = {-1, 1, 0, 0, 0, 0};
const int kNbK[6] = {0, 0, 0, 0, -1, 1};

}  // namespace

namespace {

// edge cost per voxel pair, c = SAD / step^3 (primsMST.h:112-114)
std::vector<float> edge_costs(const float* ec, int step, size_t count) {
    const float vol = (float)((double)step * step * step);
    std::vector<float> c(count);
    for (size_t e = 0; e < count; e++) c[e] = ec[e] / vol;
    return c;
}

// the reference's weight of an edge of cost c (primsMST.h:48-50, :115-117)
inline float edge_weight(float c, float two_sd) { return -std::exp(-c / two_sd); }

// Prim's with the given keys (smaller key = lighter edge), then the BFS
// order; records the neighbour slot of every node's edge to its parent with
// the reference's last-match scan (primsMST.h:198-211).
template <typename Key>
void prim_tree(const Key* key, int m1, int n1, int o1, Tree& t);

// 16-bit keys with the order and ties of the costs c = ec / step^3, when the
// SADs are integers below 2^22 (integer-valued images: HU volumes, an identity
// affine) and there are at most 65536 distinct costs.  c is non-decreasing in
// the integer SAD, so equal costs come from runs of consecutive present SADs:
// walking the present SADs in increasing order and advancing the rank when c
// changes numbers the distinct costs densely.
bool cost_ranks(const float* ec, float vol, size_t count, std::vector<uint16_t>& rank) {
    // the scans over all edges split over a few threads on large grids
    const int T = count >= ((size_t)1 << 18) ? 4 : 1;
    auto split = [&](auto body) {
        std::vector<std::thread> th;
        for (int i = 1; i < T; i++) th.emplace_back(body, i);
        body(0);
        for (auto& x : th) x.join();
    };
    auto part = [&](int i, size_t& b, size_t& e) {
        b = count * (size_t)i / (size_t)T;
        e = count * (size_t)(i + 1) / (size_t)T;
    };
    std::vector<float> mxs(T, 0.0f);
    std::vector<char> okay(T, 1);
    split([&](int i) {
        size_t b, e;
        part(i, b, e);
        float m = 0.0f;
        for (size_t x = b; x < e; x++) {
            const float v = ec[x];
            if (!(v >= 0.0f && v < 4194304.0f) || v != std::floor(v)) {
                okay[i] = 0;
                return;
            }
            m = std::max(m, v);
        }
        mxs[i] = m;
    });
    for (int i = 0; i < T; i++)
        if (!okay[i]) return false;
    const float mx = *std::max_element(mxs.begin(), mxs.end());
    const size_t top = (size_t)mx + 1;
    // (one thread: scattered byte stores from several would share cache lines)
    std::vector<uint8_t> present(top, 0);
    for (size_t x = 0; x < count; x++) present[(size_t)ec[x]] = 1;
    std::vector<uint16_t> of_sad(top, 0);
    int r = -1;
    float prev = 0.0f;
    for (size_t v = 0; v < top; v++) {
        if (!present[v]) continue;
        const float c = (float)v / vol;
        if (r < 0 || c != prev) {
            if (++r > 65535) return false;
            prev = c;
        }
        of_sad[v] = (uint16_t)r;
    }
    rank.resize(count);
    split([&](int i) {
        size_t b, e;
        part(i, b, e);
        for (size_t x = b; x < e; x++) rank[x] = of_sad[(size_t)ec[x]];
    });
    return true;
}

}  // namespace

bool weights_distinguish_costs(const float* ec, int step, int m1, int n1, int o1, float sd) {
    // Prim's only compares weights.  w(c) = -exp(-c/(2 sd)) never decreases
    // with c, so ordering by c equals ordering by w (ties included) exactly
    // when distinct costs of the pushed (valid) edges get distinct weights.
    const int N = m1 * n1 * o1;
    const float two_sd = 2.0f * sd;
    {
        // Integer SADs below 2^22 (as cost_ranks takes them): the distinct
        // costs in increasing order come from the present SADs in increasing
        // order (c = SAD / step^3 is exact in the SAD and non-decreasing), so
        // a table of present SADs replaces the sort.
        bool integral = true;
        float mx = 0.0f;
        for (size_t e = 0; e < (size_t)6 * N && integral; e++) {
            const float x = ec[e];
            integral = x >= 0.0f && x < 4194304.0f && x == std::floor(x);
            mx = std::max(mx, x);
        }
        if (integral) {
            std::vector<uint8_t> present((size_t)mx + 1, 0);
            for (int k = 0, v = 0; k < o1; k++)
                for (int j = 0; j < n1; j++)
                    for (int i = 0; i < m1; i++, v++) {
                        const bool ok[6] = {j > 0, j < n1 - 1, i > 0, i < m1 - 1, k > 0, k < o1 - 1};
                        for (int nb = 0; nb < 6; nb++)
                            if (ok[nb]) present[(size_t)ec[v + (size_t)nb * N]] = 1;
                    }
            const float vol = (float)((double)step * step * step);
            bool first = true;
            float prev_c = 0.0f, prev_w = 0.0f;
            for (size_t s = 0; s < present.size(); s++) {
                if (!present[s]) continue;
                const float cv = (float)s / vol;
                if (!first && cv == prev_c) continue;
                const float wv = edge_weight(cv, two_sd);
                if (!first && !(wv > prev_w)) return false;  // two costs, one weight
                first = false;
                prev_c = cv;
                prev_w = wv;
            }
            return true;
        }
    }
    std::vector<float> c = edge_costs(ec, step, (size_t)6 * N);
    std::vector<uint32_t> bits;
    bits.reserve((size_t)6 * N);
    for (int v = 0; v < N; v++) {
        int k = v / (m1 * n1), j = (v / m1) % n1, i = v % m1;
        for (int nb = 0; nb < 6; nb++) {
            int ii = i + kNbI[nb], jj = j + kNbJ[nb], kk = k + kNbK[nb];
            if (ii < 0 || ii >= m1 || jj < 0 || jj >= n1 || kk < 0 || kk >= o1) continue;
            uint32_t b;
            std::memcpy(&b, &c[v + (size_t)nb * N], 4);
            bits.push_back(b);  // costs are >= 0: the bit patterns sort like the values
        }
    }
    // LSD radix sort on 11-bit digits
    std::vector<uint32_t> tmp(bits.size());
    for (int shift = 0; shift < 33; shift += 11) {
        size_t cnt[2049] = {0};
        for (uint32_t b : bits) cnt[((b >> shift) & 2047u) + 1]++;
        for (int d = 0; d < 2048; d++) cnt[d + 1] += cnt[d];
        for (uint32_t b : bits) tmp[cnt[(b >> shift) & 2047u]++] = b;
        bits.swap(tmp);
    }
    float prev_w = 0.0f;
    for (size_t e = 0; e < bits.size(); e++) {
        if (e > 0 && bits[e] == bits[e - 1]) continue;
        float cv;
        std::memcpy(&cv, &bits[e], 4);
        float wv = edge_weight(cv, two_sd);
        if (e > 0 && !(wv > prev_w)) return false;  // two costs, one weight
        prev_w = wv;
    }
    return true;
}

bool build_tree_shape(const float* ec, int step, int m1, int n1, int o1, Tree& t) {
    const size_t count = (size_t)6 * m1 * n1 * o1;
    std::vector<uint16_t> rank;
    if (!cost_ranks(ec, (float)((double)step * step * step), count, rank)) return false;
    prim_tree(rank.data(), m1, n1, o1, t);
    return true;
}

void finish_tree(const float* ec, float sd, int step, Tree& t, bool schedule) {
    const int N = t.N;
    const float vol = (float)((double)step * step * step);
    const float two_sd = 2.0f * sd;
    t.edgemst.assign(N, 0.0f);
    for (int v = 0; v < N; v++) {
        const int nb = t.parent_slot[v];
        if (nb >= 0) t.edgemst[v] = -edge_weight(ec[v + (size_t)nb * N] / vol, two_sd);
    }
    if (schedule) build_schedule(t);
}

void build_tree(const float* ec, float sd, int step, int m1, int n1, int o1, Tree& t, bool schedule) {
    // weights -exp(-(c / step^3) / (2 sd)) as the keys (primsMST.h:112-117)
    const int N = m1 * n1 * o1;
    std::vector<float> w = edge_costs(ec, step, (size_t)6 * N);
    const float two_sd = 2.0f * sd;
    for (float& x : w) x = edge_weight(x, two_sd);
    prim_tree(w.data(), m1, n1, o1, t);
    finish_tree(ec, sd, step, t, schedule);
}

namespace {

template <typename Key>
void prim_tree(const Key* w, int m1, int n1, int o1, Tree& t) {
    const int N = m1 * n1 * o1;
    t.m1 = m1; t.n1 = n1; t.o1 = o1; t.N = N;
    t.ordered.assign(N, 0);
    t.parents.assign(N, -1);
    t.parent_slot.assign(N, -1);
    if (N == 0) return;
    // Prim's from the grid centre (primsMST.h:123-164)
    const int root = m1 / 2 + (n1 / 2) * m1 + (o1 / 2) * m1 * n1;
    // membership as a bit set: the check after every pop then reads a 16 KB
    // array (at 131072 nodes) that stays in L1
    std::vector<uint64_t> in_tree(((size_t)N + 63) / 64, 0);
    auto member = [&](int v) { return (bool)((in_tree[(size_t)v >> 6] >> (v & 63)) & 1u); };
    auto add = [&](int v) { in_tree[(size_t)v >> 6] |= (uint64_t)1 << (v & 63); };
    std::vector<int> depth(N, 0);
    EdgeHeap<Key> heap((size_t)4 * N + 16);
    add(root);
    int last = root;
    // linear offset of each neighbour slot (valid wherever the neighbour exists)
    const int delta[6] = {-m1, m1, -1, 1, -m1 * n1, m1 * n1};
    if ((long long)N * 8 >= (1LL << 32)) throw std::runtime_error("prims: grid too large");
    for (int it = 0; it < N - 1; it++) {
        {
            // the six neighbour slots of `last` in slot order (kNbI/J/K)
            const int lk = last / (m1 * n1), rem = last - lk * m1 * n1, lj = rem / m1, li = rem - lj * m1;
            const bool ok[6] = {lj > 0, lj < n1 - 1, li > 0, li < m1 - 1, lk > 0, lk < o1 - 1};
            for (int nb = 0; nb < 6; nb++)
                if (ok[nb]) heap.push(w[last + (size_t)nb * N], (uint32_t)last * 8u + (uint32_t)nb);
        }
        int from = -1, to = -1;
        while (to < 0) {
            if (heap.empty()) throw std::runtime_error("prims: heap exhausted");
            const uint32_t code = heap.pop();
            const int ea = (int)(code >> 3), eb = ea + delta[code & 7u];
            const bool a = member(ea), b = member(eb);
            if (a != b) {
                if (a) { from = ea; to = eb; } else { from = eb; to = ea; }
            }
        }
        add(to);
        depth[to] = depth[from] + 1;
        t.parents[to] = from;
        last = to;
    }
    // BFS-depth order, stable in vertex id (primsMST.h:166-190)
    int maxd = 0;
    for (int v = 0; v < N; v++) maxd = std::max(maxd, depth[v]);
    std::vector<int> start(maxd + 2, 0);
    for (int v = 0; v < N; v++) start[depth[v] + 1]++;
    for (int d = 1; d <= maxd + 1; d++) start[d] += start[d - 1];
    std::vector<int> fill(start.begin(), start.end());
    for (int v = 0; v < N; v++) t.ordered[fill[depth[v]]++] = v;
    // slot of the edge to the parent; the reference scans all six slots and
    // keeps the last whose linear index equals the parent (primsMST.h:198-211).
    // The linear index of slot nb is c + delta[nb] without bounds checks, so
    // a node walk in index order needs no division.
    for (int c = 0; c < N; c++) {
        const int par = t.parents[c];
        if (par < 0) continue;
        int slot = -1;
        for (int nb = 0; nb < 6; nb++)
            if (c + delta[nb] == par) slot = nb;
        t.parent_slot[c] = slot;
    }
}

}  // namespace

void build_schedule(Tree& t, bool bp_programs, int window) {
    const int N = t.N;
    if ((int)t.ordered.size() != N || (int)t.parents.size() != N)
        throw std::runtime_error("tree: ordered/parents must hold one entry per node");
    // `ordered` must be a permutation in which every parent precedes its child
    std::vector<int> pos(N, -1);
    for (int p = 0; p < N; p++) {
        int v = t.ordered[p];
        if (v < 0 || v >= N || pos[v] >= 0) throw std::runtime_error("tree: `ordered` is not a permutation");
        pos[v] = p;
    }
    // The schedule works in positions of `ordered`: a node's parent and
    // children lie in the depths next to its own, so the passes below touch
    // memory near the depth they are at rather than node ids all over the grid.
    std::vector<int> ppos(N, -1), dpos(N, 0);  // parent position, depth
    for (int p = 1; p < N; p++) {
        const int par = t.parents[t.ordered[p]];
        if (par < 0 || par >= N) throw std::runtime_error("tree: node without a valid parent");
        const int q = pos[par];
        if (q >= p) throw std::runtime_error("tree: a parent does not precede its child in `ordered`");
        ppos[p] = q;
        dpos[p] = dpos[q] + 1;
        if (dpos[p] < dpos[p - 1]) throw std::runtime_error("tree: `ordered` is not sorted by depth");
    }
    const int maxd = N > 0 ? dpos[N - 1] : 0;
    t.maxd = maxd;
    t.depth_start.assign(maxd + 2, 0);
    for (int p = 0; p < N; p++) t.depth_start[dpos[p] + 1]++;
    for (int d = 1; d <= maxd + 1; d++) t.depth_start[d] += t.depth_start[d - 1];
    t.max_width = 0;
    for (int d = 0; d <= maxd; d++) t.max_width = std::max(t.max_width, t.depth_start[d + 1] - t.depth_start[d]);
    if (!bp_programs) {
        // children of every node in `ordered` order (CSR by node id)
        t.child_start.assign(N + 1, 0);
        for (int p = 1; p < N; p++) t.child_start[t.parents[t.ordered[p]] + 1]++;
        for (int v = 0; v < N; v++) t.child_start[v + 1] += t.child_start[v];
        t.child_list.assign(N > 0 ? N - 1 : 0, 0);
        std::vector<int> fill(t.child_start.begin(), t.child_start.end() - 1);
        for (int p = 1; p < N; p++) {
            int c = t.ordered[p];
            t.child_list[fill[t.parents[c]]++] = c;
        }
        return;
    }
    // top-down selection: per position, the node and its parent's index within the parent's depth
    t.sel_info.assign(2 * (size_t)N, 0);
    t.sel_jump.assign(N, 0);
    for (int p = 0; p < N; p++) {
        t.sel_info[2 * p] = t.ordered[p];
        if (p > 0) t.sel_info[2 * p + 1] = ppos[p] - t.depth_start[dpos[p] - 1];
        const int d = dpos[p];
        if (d % kSelJump != 0) {
            t.sel_jump[p] = -(d % kSelJump);
        } else if (d > 0) {
            int q = p;
            for (int j = 0; j < kSelJump; j++) q = ppos[q];
            t.sel_jump[p] = q - t.depth_start[d - kSelJump];
        }
    }
    // children of every position, as positions in `ordered` order
    std::vector<int> cs(N + 1, 0), cl(N > 0 ? N - 1 : 0);
    for (int p = 1; p < N; p++) cs[ppos[p] + 1]++;
    for (int p = 0; p < N; p++) cs[p + 1] += cs[p];
    {
        std::vector<int> fill(cs.begin(), cs.end() - 1);
        for (int p = 1; p < N; p++) cl[fill[ppos[p]]++] = p;
    }
    // Steps and segments (bottom-up, so every child is placed before its
    // parent).  For each choice of the child c that continues into v, v's
    // earliest step is max(t(c) + 1, start of the window after the latest
    // window among the other children); v takes the smallest, preferring a
    // choice that continues c's segment inside its window, then the child
    // latest in fold order (so fewer stored messages are folded after it).
    if (window < 1 || window > kBPWindow) throw std::runtime_error("tree: BP window must be in [1, 16]");
    if ((long long)N * (window + 1) >= (1LL << 31)) throw std::runtime_error("tree: too many nodes for the BP schedule");
    // per position: step, window (-1: leaf) and the child (position) whose
    // segment it continues (-1: none)
    struct Slot {
        int step = -1, wn = -1, pred = -1;
    };
    std::vector<Slot> sl(N);
    std::vector<int> rev_heads, rev_leaves;  // positions, descending
    std::vector<int> per_round;              // segment heads per round
    rev_leaves.reserve(N);
    int internal = 0;
    for (int p = N - 1; p >= 0; p--) {
        const int b = cs[p], k = cs[p + 1] - b;
        if (k == 0) {
            if (p > 0) rev_leaves.push_back(p);
            continue;
        }
        const int* ch = cl.data() + b;
        // the two latest child windows, for "latest among the others"
        int w1 = -2, w2 = -2, i1 = -1;
        for (int i = 0; i < k; i++) {
            const int w = sl[ch[i]].wn;
            if (w > w1) {
                w2 = w1, w1 = w, i1 = i;
            } else if (w > w2) {
                w2 = w;
            }
        }
        int best = -1, best_c = -1;
        bool best_same = false;
        for (int i = k - 1; i >= 0; i--) {
            const int c = ch[i], sc = sl[c].step;
            const int other = i == i1 ? w2 : w1;
            const int cand = std::max(sc + 1, std::max(0, (other + 1) * window));
            // continuing: the very next step, still inside c's window
            const bool same = sc >= 0 && cand == sc + 1 && cand % window != 0;
            if (best < 0 || cand < best || (cand == best && same && !best_same)) {
                best = cand, best_same = same, best_c = c;
            }
        }
        Slot& s = sl[p];
        s.step = best;
        s.wn = best / window;
        s.pred = best_same ? best_c : -1;
        internal++;
        if (s.pred < 0) {
            rev_heads.push_back(p);
            if ((int)per_round.size() <= s.wn) per_round.resize(s.wn + 1, 0);
            per_round[s.wn]++;
        }
    }
    const int root = N > 0 ? t.ordered[0] : -1;
    t.rounds = N > 0 ? sl[0].wn + 1 : 0;  // the root is the last node placed
    per_round.resize(t.rounds, 0);
    t.round_start.assign(t.rounds + 1, 0);
    for (int r = 0; r < t.rounds; r++) t.round_start[r + 1] = t.round_start[r] + per_round[r];
    t.tail_round = t.rounds;
    while (t.tail_round > 0 && t.round_start[t.tail_round] - t.round_start[t.tail_round - 1] == 1) t.tail_round--;
    // segment heads by round, in `ordered` order within one (as positions, then as nodes)
    std::vector<int> head_pos(t.round_start[t.rounds]);
    std::vector<int> fill_r(t.round_start.begin(), t.round_start.end() - 1);
    for (size_t i = rev_heads.size(); i-- > 0;) {
        const int p = rev_heads[i];
        head_pos[fill_r[sl[p].wn]++] = p;
    }
    t.heads.resize(head_pos.size());
    for (size_t h = 0; h < head_pos.size(); h++) t.heads[h] = t.ordered[head_pos[h]];
    // Leaves message their parents before round 0 (one workgroup each).  A
    // segment's last node (unless it is the root) stores its message in its
    // cost row for the parent that folds it in a later round.
    t.leaves.resize(rev_leaves.size());
    for (size_t i = 0; i < rev_leaves.size(); i++) t.leaves[i] = t.ordered[rev_leaves[rev_leaves.size() - 1 - i]];
    t.leaf_w.resize(t.leaves.size());
    for (size_t i = 0; i < t.leaves.size(); i++) std::memcpy(&t.leaf_w[i], &t.edgemst[t.leaves[i]], 4);
    // programs: per segment a block of kProgRec-int slots -- the header
    // [nodes], then per node v_i the record [node, edgemst bits, k, pp,
    // parent offset x, y, z (filled on the device), s] -- followed by each
    // node's stored children (all but pp, in `ordered` order) at block offset
    // s.  Every non-root node is a child once: as pp or stored.
    const size_t heads = t.heads.size();
    const size_t stored = (N > 0 ? (size_t)N - 1 : 0) - ((size_t)internal - heads);
    t.prog_start.assign(heads + 1, 0);
    t.prog.assign(kProgRec * (heads + (size_t)internal) + stored + kProgStage, 0);
    int* out = t.prog.data();
    size_t at = 0;
    std::vector<int> longest(t.rounds, 0), seg;
    for (size_t h = 0; h < heads; h++) {
        const int x = head_pos[h];
        seg.clear();
        for (int q = x;; q = ppos[q]) {
            seg.push_back(q);
            if (q == 0 || sl[ppos[q]].pred != q) break;
        }
        const int nodes = (int)seg.size();
        t.prog_start[h] = (int)at;
        int* blk = out + at;
        blk[0] = nodes;
        size_t s_at = (size_t)kProgRec * (1 + nodes);  // stored lists after the records
        for (int i = 0; i < nodes; i++) {
            const int q = seg[i], prev = i > 0 ? seg[i - 1] : -1;
            const int cur = t.ordered[q];
            const int b = cs[q], k = cs[q + 1] - b;
            const int* ch = cl.data() + b;
            int* r = blk + kProgRec * (1 + i);
            int pp = -1;
            r[0] = cur;
            if (cur != root) std::memcpy(&r[1], &t.edgemst[cur], 4);
            r[2] = k;
            r[7] = (int)s_at;
            for (int j = 0; j < k; j++) {
                if (ch[j] == prev) pp = j;
                else blk[s_at++] = t.ordered[ch[j]];
            }
            r[3] = pp;
        }
        at += s_at;
        longest[sl[x].wn] = std::max(longest[sl[x].wn], nodes);
    }
    if (at + kProgStage != t.prog.size()) throw std::runtime_error("tree: BP programs do not cover the tree");
    t.prog_start[heads] = (int)at;
    t.bp_steps = 0;
    for (int r = 0; r < t.rounds; r++) t.bp_steps += longest[r];
    // Within a launched round, the longest segments first: workgroups are
    // dispatched about in index order, so a round with more segments than
    // the GPU holds at once starts its critical ones in the first wave.
    // Segments of one round are independent, and a program is found through
    // prog_start, so only prog_start and heads are permuted (stably, by
    // counting sort on the length, at most `window` nodes).
    std::vector<int> cnt(window + 2), ps, hd;
    for (int r = 0; r < t.tail_round; r++) {
        const int b = t.round_start[r], e = t.round_start[r + 1];
        std::fill(cnt.begin(), cnt.end(), 0);
        for (int h = b; h < e; h++) {
            const int L = t.prog[t.prog_start[h]] - 1;  // L + 1 nodes, inside one window
            if (L < 0 || L >= window) throw std::runtime_error("tree: BP segment longer than its window");
            cnt[window - L]++;
        }
        for (int i = 1; i < window + 2; i++) cnt[i] += cnt[i - 1];
        ps.assign(t.prog_start.begin() + b, t.prog_start.begin() + e);
        hd.assign(t.heads.begin() + b, t.heads.begin() + e);
        std::vector<int> pos(window + 2, 0);
        for (int i = 1; i < window + 2; i++) pos[i] = cnt[i - 1];
        for (int i = 0; i < e - b; i++) {
            const int slot = b + pos[window - (t.prog[ps[i]] - 1)]++;
            t.prog_start[slot] = ps[i];
            t.heads[slot] = hd[i];
        }
    }
}

// ordered, parents, depth_start, prog_start: 4N + 2; prog: kProgRec per segment
// and per internal node plus N - 1 children, and the staging pad; leaves,
// leaf_w, sel_info, sel_jump: 5N.
size_t tree_device_ints(int N) { return (2 * kProgRec + 10) * (size_t)N + kProgStage + 16; }

// The soft sweeps' schedule (declared in mst_host.h with its rules).
void soft_schedule(const int* ordered, const int* parents, const float* edgemst, int N, int W, SoftSchedule& out) {
    // ---- validate the given order (positions p of `ordered`) ----
    std::vector<int> pos(N, -1);
    for (int p = 0; p < N; p++) {
        const int v = ordered[p];
        if (v < 0 || v >= N) throw std::runtime_error("ordered holds an out-of-range node");
        if (pos[v] >= 0) throw std::runtime_error("tree: `ordered` is not a permutation");
        pos[v] = p;
    }
    std::vector<int> pp(N, -1), dp(N, 0);  // parent position, depth
    for (int p = 1; p < N; p++) {
        const int par = parents[ordered[p]];
        if (par < 0 || par >= N) throw std::runtime_error("tree: node without a valid parent");
        const int q = pos[par];
        if (q >= p) throw std::runtime_error("tree: a parent does not precede its child in `ordered`");
        pp[p] = q;
        dp[p] = dp[q] + 1;
        if (dp[p] < dp[p - 1]) throw std::runtime_error("tree: `ordered` is not sorted by depth");
    }
    // ---- centre: heights (best and second-best child height + 1) and upward paths ----
    std::vector<int> b1(N, 0), b2(N, 0), a1(N, -1), upd(N, 0);
    for (int p = N - 1; p >= 1; p--) {
        const int q = pp[p], v = b1[p] + 1;
        if (v > b1[q]) {
            b2[q] = b1[q];
            b1[q] = v;
            a1[q] = p;
        } else if (v > b2[q]) {
            b2[q] = v;
        }
    }
    int pc = 0;  // the centre's position
    if (N > 2) {
        for (int p = 1; p < N; p++) {
            const int q = pp[p];
            upd[p] = 1 + std::max(upd[q], a1[q] == p ? b2[q] : b1[q]);
        }
        int best = std::max(b1[0], upd[0]);
        for (int p = 1; p < N; p++) {
            const int e = std::max(b1[p], upd[p]);
            if (e < best || (e == best && ordered[p] < ordered[pc])) {
                best = e;
                pc = p;
            }
        }
    }
    // ---- the tree rooted at the centre, still in the given positions ----
    // np_[p]: new parent position (-1 at the centre); w_[p]: weight of that edge
    std::vector<int> np_(pp), nd(N, 0);
    std::vector<float> w_(N, 0.0f);
    for (int p = 1; p < N; p++) w_[p] = edgemst[ordered[p]];
    if (pc != 0) {
        // the path centre -> old root reverses: q's new parent is the path node below it
        std::vector<char> onpath(N, 0);
        for (int q = pc; q != 0; q = pp[q]) onpath[q] = 1;
        onpath[0] = 1;
        for (int below = pc, q = pp[pc];; below = q, q = pp[q]) {
            np_[q] = below;
            w_[q] = edgemst[ordered[below]];
            if (q == 0) break;
        }
        np_[pc] = -1;
        w_[pc] = 0.0f;
        for (int p = 0; p < N; p++) nd[p] = onpath[p] ? dp[pc] - dp[p] : nd[pp[p]] + 1;
    } else {
        nd = dp;
    }
    // ---- BFS order from the centre, sorted by (depth, id): counting sort of the ids ----
    int maxd = 0;
    for (int p = 0; p < N; p++) maxd = std::max(maxd, nd[p]);
    std::vector<int> ndv(N);  // new depth by id
    for (int p = 0; p < N; p++) ndv[ordered[p]] = nd[p];
    out.maxd = maxd;
    out.depth_start.assign(maxd + 2, 0);
    for (int v = 0; v < N; v++) out.depth_start[ndv[v] + 1]++;
    for (int d = 0; d <= maxd; d++) out.depth_start[d + 1] += out.depth_start[d];
    out.ordered.resize(N);
    {
        std::vector<int> fill(out.depth_start.begin(), out.depth_start.end() - 1);
        for (int v = 0; v < N; v++) out.ordered[fill[ndv[v]]++] = v;
    }
    // r: positions of the new order; rp[r] = new parent position; old position of r = pos[ordered_new[r]]
    std::vector<int> npos(N);  // new position by old position
    for (int r = 0; r < N; r++) npos[pos[out.ordered[r]]] = r;
    std::vector<int> rp(N, -1);
    out.parents.assign(N, -1);
    out.edgemst.assign(N, 0.0f);
    for (int r = 0; r < N; r++) {
        const int p = pos[out.ordered[r]], q = np_[p];
        rp[r] = q < 0 ? -1 : npos[q];
        out.parents[out.ordered[r]] = q < 0 ? -1 : ordered[q];
        out.edgemst[out.ordered[r]] = w_[p];
    }
    // ---- children: CSR by position (children in BFS order) and by id for the device ----
    std::vector<int> cs(N + 1, 0), cl(N > 0 ? N - 1 : 0);
    for (int r = 1; r < N; r++) cs[rp[r] + 1]++;
    for (int r = 0; r < N; r++) cs[r + 1] += cs[r];
    {
        std::vector<int> fill(cs.begin(), cs.end() - 1);
        for (int r = 1; r < N; r++) cl[fill[rp[r]]++] = r;
    }
    out.child_start.assign(N + 1, 0);
    for (int r = 0; r < N; r++) out.child_start[out.ordered[r] + 1] = cs[r + 1] - cs[r];
    for (int v = 0; v < N; v++) out.child_start[v + 1] += out.child_start[v];
    out.child_list.assign(cl.size(), 0);
    for (int r = 0; r < N; r++) {
        const int at = out.child_start[out.ordered[r]];
        for (int e = cs[r]; e < cs[r + 1]; e++) out.child_list[at + e - cs[r]] = out.ordered[cl[e]];
    }
    // ---- rounds (positions r; heads in BFS order) ----
    auto next_window = [W](int x) { return (x / W + 1) * W; };
    auto pack = [&](const std::vector<int>& heads, const std::vector<int>& step, const std::vector<int>& next,
                    RoundsHost& o) {
        int rounds = 0;
        std::vector<int> len(heads.size());
        for (size_t i = 0; i < heads.size(); i++) {
            int n = 0;
            for (int r = heads[i]; r >= 0; r = next[r]) n++;
            len[i] = n;  // at most W: a segment stays in one window
            rounds = std::max(rounds, step[heads[i]] / W + 1);
        }
        std::vector<int> bstart((size_t)rounds * W + 1, 0);
        auto bucket = [&](size_t i) { return (size_t)(step[heads[i]] / W) * W + (W - len[i]); };
        for (size_t i = 0; i < heads.size(); i++) bstart[bucket(i) + 1]++;
        for (size_t k = 0; k + 1 < bstart.size(); k++) bstart[k + 1] += bstart[k];
        std::vector<int> sorted(heads.size());
        for (size_t i = 0; i < heads.size(); i++) sorted[bstart[bucket(i)]++] = (int)i;
        o.round_start.assign(rounds + 1, 0);
        o.seg_start.assign(1, 0);
        o.seg_start.reserve(heads.size() + 1);
        o.nodes.clear();
        o.nodes.reserve(N);
        for (int i : sorted) {
            o.round_start[step[heads[i]] / W + 1]++;
            for (int r = heads[i]; r >= 0; r = next[r]) o.nodes.push_back(out.ordered[r]);
            o.seg_start.push_back((int)o.nodes.size());
        }
        for (int k = 0; k < rounds; k++) o.round_start[k + 1] += o.round_start[k];
    };
    // upward
    std::vector<int> tu(N, 0), up_next(N, -1), heads;
    std::vector<char> continued(N, 0);
    for (int r = N - 1; r >= 0; r--) {
        const int c0 = cs[r], c1 = cs[r + 1];
        if (c0 == c1) continue;
        int best = cl[c0];
        for (int e = c0 + 1; e < c1; e++)
            if (tu[cl[e]] > tu[best]) best = cl[e];
        int tv = tu[best] + 1;
        for (int e = c0; e < c1; e++)
            if (cl[e] != best) tv = std::max(tv, next_window(tu[cl[e]]));
        tu[r] = tv;
        if (tv == tu[best] + 1 && tv / W == tu[best] / W) {
            up_next[best] = r;
            continued[r] = 1;
        }
    }
    for (int r = 0; r < N; r++)
        if (!continued[r]) heads.push_back(r);
    pack(heads, tu, up_next, out.up);
    // downward
    std::vector<int> height(N, 0), heavy(N, -1);
    for (int r = N - 1; r >= 1; r--) height[rp[r]] = std::max(height[rp[r]], height[r] + 1);
    for (int r = 0; r < N; r++)
        for (int e = cs[r]; e < cs[r + 1]; e++)
            if (heavy[r] < 0 || height[cl[e]] > height[heavy[r]]) heavy[r] = cl[e];
    std::vector<int> td(N, 0), down_next(N, -1);
    heads.clear();
    for (int r = 1; r < N; r++) {
        const int p = rp[r];
        if (p == 0) {
            td[r] = 0;
            heads.push_back(r);
        } else if (r == heavy[p] && (td[p] + 1) / W == td[p] / W) {
            td[r] = td[p] + 1;
            down_next[p] = r;
        } else {
            td[r] = next_window(td[p]);
            heads.push_back(r);
        }
    }
    pack(heads, td, down_next, out.down);
}

}  // namespace deeds
