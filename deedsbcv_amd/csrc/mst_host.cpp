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
#include <stdexcept>
#include <unordered_map>

namespace deeds {

namespace {

struct HeapEdge {
    double weight;  // the reference stores the float weight in a double
    int from;
    int to;
};
// Edge::operator< of primsMST.h:24-26
struct LighterFirst {
    bool operator()(const HeapEdge& a, const HeapEdge& b) const { return a.weight > b.weight; }
};

const int kNbI[6] = {0, 0, -1, 1, 0, 0};
const int kNbJ[6] = {-1, 1, 0, 0, 0, 0};
const int kNbK[6] = {0, 0, 0, 0, -1, 1};

}  // namespace

void image_stats(const float* im, long long count, float* mean, float* sd) {
    float mu = 0.0f;
    for (long long t = 0; t < count; t++) mu = mu + im[t];
    mu = mu / (float)count;
    float var = 0.0f;
    for (long long t = 0; t < count; t++) {
        double d = (double)(im[t] - mu);
        var = (float)((double)var + d * d);
    }
    *mean = mu;
    *sd = std::sqrt(var / (float)count);
}

void build_tree(const float* ec, float sd, int step, int m1, int n1, int o1, Tree& t) {
    const int N = m1 * n1 * o1;
    t.m1 = m1; t.n1 = n1; t.o1 = o1; t.N = N;
    t.ordered.assign(N, 0);
    t.parents.assign(N, -1);
    t.edgemst.assign(N, 0.0f);
    if (N == 0) return;
    // weights -exp(-(c / step^3) / (2 sd)) (primsMST.h:112-117)
    const float vol = (float)((double)step * step * step);
    const float two_sd = 2.0f * sd;
    std::vector<float> w((size_t)6 * N);
    for (size_t e = 0; e < w.size(); e++) {
        float c = ec[e] / vol;
        w[e] = -std::exp(-c / two_sd);
    }
    auto neighbour = [&](int v, int nb) -> int {
        int k = v / (m1 * n1), j = (v / m1) % n1, i = v % m1;
        int ii = i + kNbI[nb], jj = j + kNbJ[nb], kk = k + kNbK[nb];
        if (ii < 0 || ii >= m1 || jj < 0 || jj >= n1 || kk < 0 || kk >= o1) return -1;
        return ii + jj * m1 + kk * m1 * n1;
    };
    // Prim's from the grid centre (primsMST.h:123-164)
    const int root = m1 / 2 + (n1 / 2) * m1 + (o1 / 2) * m1 * n1;
    std::vector<uint8_t> in_tree(N, 0);
    std::vector<int> depth(N, 0);
    std::vector<HeapEdge> heap;
    heap.reserve((size_t)4 * N + 16);
    in_tree[root] = 1;
    int last = root;
    LighterFirst cmp;
    for (int it = 0; it < N - 1; it++) {
        for (int nb = 0; nb < 6; nb++) {
            int v = neighbour(last, nb);
            if (v >= 0) {
                heap.push_back(HeapEdge{(double)w[last + (size_t)nb * N], last, v});
                std::push_heap(heap.begin(), heap.end(), cmp);
            }
        }
        int from = -1, to = -1;
        while (to < 0) {
            if (heap.empty()) throw std::runtime_error("prims: heap exhausted");
            HeapEdge e = heap.front();
            std::pop_heap(heap.begin(), heap.end(), cmp);
            heap.pop_back();
            bool a = in_tree[e.from], b = in_tree[e.to];
            if (a != b) {
                if (a) { from = e.from; to = e.to; } else { from = e.to; to = e.from; }
            }
        }
        in_tree[to] = 1;
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
    // weight of the edge to the parent; the reference scans all six slots and
    // keeps the last whose linear index equals the parent (primsMST.h:198-211)
    for (int p = 1; p < N; p++) {
        int c = t.ordered[p], par = t.parents[c];
        int z = c / (m1 * n1), x = (c - z * m1 * n1) / m1, y = c - z * m1 * n1 - x * m1;
        for (int nb = 0; nb < 6; nb++) {
            int idx = y + kNbI[nb] + (x + kNbJ[nb]) * m1 + (z + kNbK[nb]) * m1 * n1;
            if (idx == par) t.edgemst[c] = -w[c + (size_t)nb * N];
        }
    }
    build_schedule(t);
}

void build_schedule(Tree& t) {
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
    std::vector<int> depth(N, 0);
    for (int p = 1; p < N; p++) {
        int c = t.ordered[p];
        int par = t.parents[c];
        if (par < 0 || par >= N) throw std::runtime_error("tree: node without a valid parent");
        if (pos[par] >= p) throw std::runtime_error("tree: a parent does not precede its child in `ordered`");
        depth[c] = depth[par] + 1;
    }
    int maxd = 0;
    for (int v = 0; v < N; v++) maxd = std::max(maxd, depth[v]);
    t.maxd = maxd;
    t.depth_start.assign(maxd + 2, 0);
    for (int v = 0; v < N; v++) t.depth_start[depth[v] + 1]++;
    for (int d = 1; d <= maxd + 1; d++) t.depth_start[d] += t.depth_start[d - 1];
    for (int p = 0; p < N; p++)
        if (depth[t.ordered[p]] < (p == 0 ? 0 : depth[t.ordered[p - 1]]))
            throw std::runtime_error("tree: `ordered` is not sorted by depth");
    // groups of children per parent, per depth, in order of first appearance
    t.grp_depth_start.assign(maxd + 1, 0);
    t.grp_parent.clear();
    t.grp_child.assign(1, 0);
    t.child_list.clear();
    t.child_list.reserve(N > 0 ? N - 1 : 0);
    std::unordered_map<int, int> slot;
    std::vector<int> counts;
    for (int d = 1; d <= maxd; d++) {
        int g0 = (int)t.grp_parent.size();
        t.grp_depth_start[d - 1] = g0;
        slot.clear();
        counts.clear();
        int s = t.depth_start[d], e = t.depth_start[d + 1];
        for (int p = s; p < e; p++) {
            int par = t.parents[t.ordered[p]];
            auto it = slot.find(par);
            if (it == slot.end()) {
                slot.emplace(par, (int)counts.size());
                t.grp_parent.push_back(par);
                counts.push_back(1);
            } else {
                counts[it->second]++;
            }
        }
        int base = (int)t.child_list.size();
        std::vector<int> pos(counts.size());
        for (size_t g = 0; g < counts.size(); g++) {
            pos[g] = base;
            base += counts[g];
            t.grp_child.push_back(base);
        }
        t.child_list.resize(base);
        for (int p = s; p < e; p++) {
            int c = t.ordered[p];
            int g = slot[t.parents[c]];
            t.child_list[pos[g]++] = c;
        }
    }
    if (maxd >= 1) t.grp_depth_start[maxd] = (int)t.grp_parent.size();
}

}  // namespace deeds
