// Host half of the MST stage: edge weights, Prim's algorithm and the BFS
// ordering (reference: primsMST.h:101-223), plus the per-depth schedule the
// device belief propagation walks.
#pragma once

#include <vector>

namespace deeds {

struct Tree {
    int m1 = 0, n1 = 0, o1 = 0, N = 0;
    std::vector<int> ordered;      // nodes sorted by (depth, id); ordered[0] is the root
    std::vector<int> parents;      // parent of each node, -1 for the root
    std::vector<float> edgemst;    // weight of the edge to the parent (0 for the root)
    // schedule: depth_start[d] = first position of depth d in `ordered` (d = 0..maxd+1)
    int maxd = 0;
    std::vector<int> depth_start;
    // bottom-up groups: for every depth d >= 1, the parents (depth d-1) with children
    // at depth d; grp_depth_start[d-1 .. d] spans the groups of depth d.
    std::vector<int> grp_depth_start;  // maxd + 1 entries (index 0 unused = 0)
    std::vector<int> grp_parent;
    std::vector<int> grp_child;        // CSR offsets into child_list (size G+1)
    std::vector<int> child_list;       // children of each group, in `ordered` order
};

// primsMST.h:101-110: float mean and std of the image, accumulated exactly as
// the reference does (sequential float sum; float += double square).
void image_stats(const float* im, long long count, float* mean, float* sd);

// primsMST.h:112-223 given the SAD edge costs ec[node + nb*N] of the (m1,n1,o1)
// control grid, the image std `sd` and the grid step.
void build_tree(const float* ec, float sd, int step, int m1, int n1, int o1, Tree& t);

// Fills depth_start and the bottom-up groups from ordered/parents (any tree
// in BFS order, e.g. one handed in by a caller).
void build_schedule(Tree& t);

}  // namespace deeds
