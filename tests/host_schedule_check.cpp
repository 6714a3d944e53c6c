// Host-side checks of the tree schedule and the weight-order check
// (deedsbcv_amd/csrc/mst_host.cpp), built by tests/test_host_schedule.py
// with g++ and run on the CPU.
//
// 1. BP programs: for random trees in BFS order (chains, random recursive
//    trees, bushy paths, binary trees, Prim's trees of random grids) and
//    several window lengths, the programs are interpreted the way the BP
//    kernel runs them, with a scalar stand-in for the message
//    (msg = f(acc * edgemst), a non-linear float function), and compared
//    bit for bit with the reference's depth-by-depth fold
//    (regularisation.h:190-235: per depth, messages of all nodes, then
//    cost[parent] += message in `ordered` order).  The interpreter also
//    checks that a program reads only messages stored in earlier rounds (or
//    earlier in the spine) and that every node is handled exactly once.
// 2. weights_distinguish_costs against a brute force over the distinct costs
//    of the valid edges (std::set), for integer SADs (the table path) and
//    fractional ones (the sort path).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <set>
#include <vector>

#include "mst_host.h"

using namespace deeds;

static float msgf(float x, int v) { return std::sqrt(std::fabs(x)) * 1.37f + (float)(v % 97) * 0.013f; }

static int check_programs(Tree& t, int window, const std::vector<float>& data) {
    build_schedule(t, true, window);
    const int N = t.N;
    const int root = t.ordered[0];
    std::vector<int> depth(N, 0);
    for (int p = 1; p < N; p++) depth[t.ordered[p]] = depth[t.parents[t.ordered[p]]] + 1;
    int maxd = 0;
    for (int v = 0; v < N; v++) maxd = std::max(maxd, depth[v]);
    // the reference's fold
    std::vector<float> cost(data), refmsg(N, 0.0f);
    for (int d = maxd; d >= 1; d--) {
        for (int p = 0; p < N; p++) {
            const int c = t.ordered[p];
            if (depth[c] == d) refmsg[c] = msgf(cost[c] * t.edgemst[c], c);
        }
        for (int p = 0; p < N; p++) {
            const int c = t.ordered[p];
            if (depth[c] == d) cost[t.parents[c]] += refmsg[c];
        }
    }
    // children of every node in `ordered` order
    std::vector<std::vector<int>> kids(N);
    for (int p = 1; p < N; p++) kids[t.parents[t.ordered[p]]].push_back(t.ordered[p]);
    // the programs
    std::vector<float> row(data), got(N, NAN);
    std::vector<int> wrote(N, -2), done(N, 0);
    int bad = 0;
    for (size_t i = 0; i < t.leaves.size(); i++) {
        const int v = t.leaves[i];
        float w;
        std::memcpy(&w, &t.leaf_w[i], 4);
        row[v] = msgf(row[v] * w, v);
        got[v] = row[v];
        wrote[v] = -1;
        done[v]++;
    }
    auto run = [&](int h, int rnd, bool tail) {
        const int* blk = t.prog.data() + t.prog_start[h];
        const int L = blk[0] - 1;
        auto stored = [&](int c) {
            if (wrote[c] == -2 || (!tail && wrote[c] >= rnd) || (tail && wrote[c] > rnd)) bad++;
            return row[c];
        };
        // a record's children in order: the stored ones from the list at s, pp in between
        auto child = [&](const int* r, int j) {
            const int* st = blk + r[7];
            return r[3] < 0 || j < r[3] ? st[j] : st[j - 1];
        };
        const int* rec = blk + kProgRec;
        int node = rec[0];
        float w;
        std::memcpy(&w, &rec[1], 4);
        if (rec[3] != -1) bad++;
        float C = row[node];
        if (rec[2] != (int)kids[node].size()) bad++;
        for (int j = 0; j < rec[2]; j++) {
            if (j < (int)kids[node].size() && child(rec, j) != kids[node][j]) bad++;
            C += stored(child(rec, j));
        }
        for (int i = 0; i <= L; i++) {
            if (node == root) break;
            const int* nrec = rec + kProgRec;
            const int next = i < L ? nrec[0] : -1;
            const float m = msgf(C * w, node);
            got[node] = m;
            done[node]++;
            if (next < 0) {
                row[node] = m;
                wrote[node] = rnd;
                return;
            }
            if (t.parents[node] != next) bad++;
            const int nk = nrec[2], npp = nrec[3];
            if (npp < 0 || npp >= nk) bad++;
            // the children as listed must be exactly next's children in `ordered` order
            if (nk != (int)kids[next].size()) bad++;
            for (int j = 0; j < nk && j < (int)kids[next].size(); j++) {
                const int c = j == npp ? node : child(nrec, j);
                if (kids[next][j] != c) bad++;
            }
            C = row[next];
            for (int j = 0; j < nk; j++) C += j == npp ? m : stored(child(nrec, j));
            rec = nrec;
            node = next;
            std::memcpy(&w, &rec[1], 4);
        }
        row[node] = C;  // the root's accumulator
        done[node]++;
    };
    for (int r = 0; r < t.tail_round; r++)  // a round's segments in reverse: they must not depend on each other
        for (int h = t.round_start[r + 1] - 1; h >= t.round_start[r]; h--) run(h, r, false);
    for (int h = t.round_start[t.tail_round]; h < t.round_start[t.rounds]; h++)
        run(h, t.tail_round + (h - t.round_start[t.tail_round]), true);
    for (int v = 0; v < N; v++) {
        if (v == root) {
            if (N > 1 && (std::memcmp(&row[v], &cost[v], 4) != 0 || done[v] != 1)) bad++;
            continue;
        }
        if (done[v] != 1 || std::memcmp(&got[v], &refmsg[v], 4) != 0) bad++;
    }
    // the critical path is at least the tree's height in internal nodes, and
    // exactly that with one step per round (every node as early as it can be)
    if (t.bp_steps < maxd || (window == 1 && t.bp_steps != maxd)) bad++;
    return bad;
}

static Tree bfs_tree(const std::vector<int>& par, std::mt19937& g) {
    const int N = (int)par.size();
    std::vector<int> depth(N, 0);
    for (int v = 1; v < N; v++) depth[v] = depth[par[v]] + 1;
    std::vector<int> lab(N);
    for (int v = 0; v < N; v++) lab[v] = v;
    std::shuffle(lab.begin() + 1, lab.end(), g);
    // BFS order by (depth, node id) after relabelling, as the reference sorts it
    Tree t;
    t.N = N;
    t.parents.assign(N, -1);
    for (int v = 1; v < N; v++) t.parents[lab[v]] = lab[par[v]];
    std::vector<int> nd(N);
    for (int v = 0; v < N; v++) nd[lab[v]] = depth[v];
    t.ordered.resize(N);
    for (int v = 0; v < N; v++) t.ordered[v] = v;
    std::stable_sort(t.ordered.begin(), t.ordered.end(), [&](int a, int b) { return nd[a] < nd[b]; });
    t.edgemst.resize(N);
    for (int v = 0; v < N; v++) t.edgemst[v] = std::uniform_real_distribution<float>(0.1f, 1.0f)(g);
    return t;
}

static bool brute_distinguish(const std::vector<float>& ec, int step, int m1, int n1, int o1, float sd) {
    const int N = m1 * n1 * o1;
    const float vol = (float)((double)step * step * step);
    const int di[6] = {0, 0, -1, 1, 0, 0}, dj[6] = {-1, 1, 0, 0, 0, 0}, dk[6] = {0, 0, 0, 0, -1, 1};
    std::set<float> costs;
    for (int v = 0; v < N; v++) {
        const int k = v / (m1 * n1), j = (v / m1) % n1, i = v % m1;
        for (int nb = 0; nb < 6; nb++) {
            const int ii = i + di[nb], jj = j + dj[nb], kk = k + dk[nb];
            if (ii < 0 || ii >= m1 || jj < 0 || jj >= n1 || kk < 0 || kk >= o1) continue;
            costs.insert(ec[v + (size_t)nb * N] / vol);
        }
    }
    bool first = true;
    float prev = 0.0f;
    for (float c : costs) {
        const float w = -std::exp(-c / (2.0f * sd));
        if (!first && !(w > prev)) return false;
        first = false;
        prev = w;
    }
    return true;
}

int main() {
    std::mt19937 g(7);
    int bad = 0, trees = 0;
    for (int it = 0; it < 240; it++) {
        const int N = 1 + (int)(g() % 2500);
        std::vector<int> par(N, -1);
        const int kind = it % 4;
        for (int v = 1; v < N; v++) {
            if (kind == 0) par[v] = v - 1;                                     // a chain
            else if (kind == 1) par[v] = (int)(g() % v);                       // random recursive tree
            else if (kind == 2) par[v] = std::max(0, v - 1 - (int)(g() % 4));  // long bushy paths
            else par[v] = (v - 1) / 2;                                         // binary tree
        }
        Tree t = bfs_tree(par, g);
        std::vector<float> data(N);
        for (float& x : data) x = std::uniform_real_distribution<float>(0.0f, 100.0f)(g);
        for (int W : {1, 3, 16}) bad += check_programs(t, W, data);
        trees++;
    }
    // Prim's trees of random grids (the shapes registration produces)
    for (int it = 0; it < 12; it++) {
        const int m1 = 4 + g() % 28, n1 = 4 + g() % 28, o1 = 3 + g() % 20, N = m1 * n1 * o1;
        std::vector<float> ec(6 * (size_t)N);
        for (float& x : ec) x = (float)(g() % 400);
        Tree t;
        build_tree(ec.data(), 50.0f, 3, m1, n1, o1, t);
        std::vector<float> data(N);
        for (float& x : data) x = std::uniform_real_distribution<float>(0.0f, 100.0f)(g);
        for (int W : {1, 16}) bad += check_programs(t, W, data);
        trees++;
    }
    printf("schedule: %d trees, %d mismatches\n", trees, bad);
    int wbad = 0, wtrue = 0, wfalse = 0;
    for (int it = 0; it < 300; it++) {
        const int m1 = 2 + g() % 16, n1 = 2 + g() % 16, o1 = 1 + g() % 10, N = m1 * n1 * o1;
        const int step = 1 + g() % 6;
        std::vector<float> ec(6 * (size_t)N);
        const int range = (int[]){5, 300, 30000, 300000}[it % 4];
        for (float& x : ec) x = (float)(g() % range);
        if (it % 5 == 1)
            for (float& x : ec) x += 0.25f;  // fractional SADs: the sort path
        if (it % 3 == 0)
            for (int v = 0; v < N; v++) ec[v] = (float)(range + 7 + g() % 50);  // values only invalid slots may hold
        const float sd = (float[]){0.01f, 0.5f, 3.0f, 40.0f, 900.0f}[g() % 5];
        const bool got = weights_distinguish_costs(ec.data(), step, m1, n1, o1, sd);
        const bool want = brute_distinguish(ec, step, m1, n1, o1, sd);
        wbad += got != want;
        (want ? wtrue : wfalse)++;
    }
    printf("weights: %d cases (%d true, %d false), %d mismatches\n", wtrue + wfalse, wtrue, wfalse, wbad);
    return bad + wbad != 0;
}
