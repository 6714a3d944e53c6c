// CPU checker of the soft sweeps' schedule (soft_schedule in
// deedsbcv_amd/csrc/mst_host.cpp) on random trees, against brute force:
//  - the new root has the least eccentricity (BFS from every node), the
//    smallest id among ties;
//  - the output is the same undirected tree with the same edge weights,
//    ordered by (BFS depth from the new root, id), with consistent depth
//    starts and child lists (children in that order);
//  - the upward rounds hold every node once, each segment a path that walks
//    to the parent, and every child of a node is the node before it in its
//    segment or lies in an earlier round; the downward rounds hold every node
//    but the root, each segment walks to a child, and a node's parent is the
//    node before it, the root, or in an earlier round; no segment is longer
//    than the window.
// Prints "<k> mismatches" and exits nonzero on any failure.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <queue>
#include <random>
#include <vector>

#include "mst_host.h"

using namespace deeds;

namespace {

std::vector<int> bfs(const std::vector<std::vector<int>>& adj, int s) {
    std::vector<int> d(adj.size(), -1);
    std::queue<int> q;
    d[s] = 0;
    q.push(s);
    while (!q.empty()) {
        const int x = q.front();
        q.pop();
        for (int y : adj[x])
            if (d[y] < 0) {
                d[y] = d[x] + 1;
                q.push(y);
            }
    }
    return d;
}

// rounds: every segment's node list, its round, and the node -> (round, segment, index) map
bool check_rounds(const RoundsHost& r, int N, int W, const std::vector<int>& par, int root, bool up) {
    const int R = (int)r.round_start.size() - 1;
    std::vector<int> rnd(N, -1), seg(N, -1), idx(N, -1);
    size_t count = 0;
    for (int k = 0; k < R; k++)
        for (int sgi = r.round_start[k]; sgi < r.round_start[k + 1]; sgi++) {
            const int a = r.seg_start[sgi], b = r.seg_start[sgi + 1];
            if (b - a < 1 || b - a > W) return false;
            for (int i = a; i < b; i++) {
                const int v = r.nodes[i];
                if (v < 0 || v >= N || rnd[v] >= 0) return false;
                rnd[v] = k;
                seg[v] = sgi;
                idx[v] = i - a;
                count++;
                if (i > a) {
                    const int prev = r.nodes[i - 1];
                    if (up ? par[prev] != v : par[v] != prev) return false;  // a path up / down the tree
                }
            }
        }
    if (count != (size_t)(up ? N : N - 1) || (!up && rnd[root] >= 0)) return false;
    for (int v = 0; v < N; v++) {
        if (v == root) continue;
        const int p = par[v];
        if (up) {  // child v of p: before p in p's segment, or an earlier round
            if (!((seg[v] == seg[p] && idx[v] + 1 == idx[p]) || rnd[v] < rnd[p])) return false;
        } else {  // parent p of v: before v in v's segment, the root, or an earlier round
            if (!(p == root || (seg[v] == seg[p] && idx[p] + 1 == idx[v]) || rnd[p] < rnd[v])) return false;
        }
    }
    return true;
}

bool check(const std::vector<int>& ord, const std::vector<int>& par, const std::vector<float>& w, int W) {
    const int N = (int)ord.size();
    SoftSchedule s;
    soft_schedule(ord.data(), par.data(), w.data(), N, W, s);
    std::vector<std::vector<int>> adj(N);
    for (int v = 0; v < N; v++)
        if (par[v] >= 0) {
            adj[v].push_back(par[v]);
            adj[par[v]].push_back(v);
        }
    // centre
    int centre = 0, best = N + 1;
    for (int v = 0; v < N; v++) {
        const std::vector<int> d = bfs(adj, v);
        const int e = *std::max_element(d.begin(), d.end());
        if (e < best) {
            best = e;
            centre = v;
        }
    }
    const int root = s.ordered.empty() ? -1 : s.ordered[0];
    if (N > 2 && root != centre) return false;
    // same tree, same weights
    for (int v = 0; v < N; v++) {
        const int p = s.parents[v];
        if (v == root) {
            if (p != -1) return false;
            continue;
        }
        float win;
        if (par[v] == p) win = w[v];
        else if (p >= 0 && par[p] == v) win = w[p];
        else return false;
        if (std::memcmp(&win, &s.edgemst[v], sizeof(float)) != 0) return false;
    }
    // (depth, id) order and depth starts
    const std::vector<int> d = bfs(adj, root);
    for (int i = 1; i < N; i++) {
        const int a = s.ordered[i - 1], b = s.ordered[i];
        if (d[a] > d[b] || (d[a] == d[b] && a >= b)) return false;
    }
    if (s.maxd != d[s.ordered[N - 1]] || (int)s.depth_start.size() != s.maxd + 2) return false;
    for (int k = 0; k <= s.maxd + 1; k++) {
        int c = 0;
        for (int v = 0; v < N; v++) c += d[v] < k;
        if (s.depth_start[k] != c) return false;
    }
    // child lists in `ordered` order
    std::vector<int> cnt(N, 0);
    for (int i = 1; i < N; i++) {
        const int c = s.ordered[i], p = s.parents[c];
        if (s.child_list[s.child_start[p] + cnt[p]++] != c) return false;
    }
    for (int v = 0; v < N; v++)
        if (s.child_start[v + 1] - s.child_start[v] != cnt[v]) return false;
    return check_rounds(s.up, N, W, s.parents, root, true) && check_rounds(s.down, N, W, s.parents, root, false);
}

}  // namespace

int main() {
    std::mt19937 rng(2024);
    int bad = 0, trees = 0;
    for (int it = 0; it < 1500; it++) {
        const int n = 1 + (int)(rng() % (it < 1000 ? 30 : 400));
        const int W = it % 3 == 0 ? 16 : 1 + (int)(rng() % 6);
        // random tree over a random labelling: bushy, chain-like or few-children shapes
        std::vector<int> lab(n);
        for (int i = 0; i < n; i++) lab[i] = i;
        std::shuffle(lab.begin(), lab.end(), rng);
        std::vector<int> par(n, -1), dep(n, 0);
        std::vector<float> w(n, 0.0f);
        const int mode = (int)(rng() % 3);
        for (int i = 1; i < n; i++) {
            const int j = mode == 0 ? (int)(rng() % i) : mode == 1 ? std::max(0, i - 1 - (int)(rng() % 3))
                                                                : (int)(rng() % std::min(i, 5));
            par[lab[i]] = lab[j];
            dep[lab[i]] = dep[lab[j]] + 1;
            w[lab[i]] = (float)(rng() % 1000) / 997.0f;
        }
        std::vector<int> ord(n);
        for (int i = 0; i < n; i++) ord[i] = i;
        std::sort(ord.begin(), ord.end(), [&](int a, int b) { return dep[a] != dep[b] ? dep[a] < dep[b] : a < b; });
        trees++;
        if (!check(ord, par, w, W)) {
            if (bad < 5) std::printf("mismatch: tree %d, %d nodes, window %d\n", it, n, W);
            bad++;
        }
    }
    std::printf("%d mismatches in %d trees\n", bad, trees);
    return bad ? 1 : 0;
}
