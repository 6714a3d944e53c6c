// Host half of the MST stage: edge weights, Prim's algorithm and the BFS
// ordering (reference: primsMST.h:101-223), plus the per-depth schedule the
// device belief propagation walks.
#pragma once

#include <cstddef>
#include <vector>

namespace deeds {

struct Tree {
    int m1 = 0, n1 = 0, o1 = 0, N = 0;
    std::vector<int> ordered;      // nodes sorted by (depth, id); ordered[0] is the root
    std::vector<int> parents;      // parent of each node, -1 for the root
    std::vector<float> edgemst;    // weight of the edge to the parent (0 for the root)
    std::vector<int> parent_slot;  // neighbour slot (0..5) of that edge, -1 for the root
    // schedule: depth_start[d] = first position of depth d in `ordered` (d = 0..maxd+1)
    int maxd = 0;
    std::vector<int> depth_start;
    // top-down selection: (node, index of its parent within the parent's
    // depth) per position of `ordered`, and the largest depth width
    std::vector<int> sel_info;
    // selection over jumps of kSelJump depths: per position at a depth d
    // that is a multiple of kSelJump (an anchor), the index of its
    // kSelJump-th ancestor within that ancestor's depth (0 at the root);
    // elsewhere -(d mod kSelJump), the steps up to its anchor
    std::vector<int> sel_jump;
    int max_width = 0;
    // children of every node in `ordered` order (CSR by node id); built
    // only without BP programs (the schedule of the programs uses positions)
    std::vector<int> child_start;  // N + 1
    std::vector<int> child_list;   // N - 1
    // Bottom-up rounds.  Every internal node gets a step t(v) >= 0 (leaves:
    // the launch before round 0); round r is the window of steps
    // [r*W, (r+1)*W), W = kBPWindow, one launch.  A segment is a path
    // v_0 .. v_L, v_{i+1} the parent of v_i, at consecutive steps of one
    // window, run by one workgroup with v_i's message kept on chip for
    // v_{i+1}; every other child of a segment node finished in an earlier
    // window (its message is stored in its cost row).  t(v) is the earliest
    // step this allows: the later of t(c)+1 for the child c that continues
    // into v and the start of the window after every other child's.
    // heads are the segments' first nodes, grouped by round; below the spine
    // the longest segments come first within a round (then `ordered` order).
    std::vector<int> heads;
    std::vector<int> round_start;  // rounds + 1 offsets into heads
    int rounds = 0;
    // Rounds tail_round .. rounds-1 hold one segment each (the tree's spine):
    // one workgroup runs them in order in a single launch.
    int tail_round = 0;
    // Bottom-up critical path in node steps: the longest segment of each
    // round, summed over the rounds.
    long long bp_steps = 0;
    // Before round 0: every leaf messages its parent (leaf_w = edgemst bits).
    std::vector<int> leaves;
    std::vector<int> leaf_w;
    // Per segment (in round order) the program block one workgroup executes,
    // in slots of kProgRec ints: the header [L + 1, 0...], then for v_0 .. v_L
    // the record [node, edgemst bits, k, pp, off x, off y, off z, s], then the
    // nodes' stored children.  pp is the position of v_{i-1} among the k
    // children of v_i in `ordered` order (-1 for v_0); the other children
    // ("stored": their messages are in their cost rows) are listed, in that
    // order, at offset s of the block.  The offsets (the parent's field minus
    // the node's, in label units) are written on the device before each pass.
    // prog ends with kProgStage ints of padding, so a block's first
    // kProgStage ints can always be read.
    std::vector<int> prog_start;  // heads + 1 offsets into prog
    std::vector<int> prog;
};

// Steps per bottom-up round (one launch each, below the spine).
constexpr int kBPWindow = 16;

// Depths per step of the label selection's walk (the tables of the depths
// between two anchors are composed beforehand).
constexpr int kSelJump = 8;

// Ints per program record, and the ints of a program block the device BP
// stages in LDS (the header and 16 records take 136).
constexpr int kProgRec = 8;
constexpr int kProgStage = 256;
static_assert(kProgRec * (1 + kBPWindow) <= kProgStage, "a block's header and records must fit the staged part");


// Upper bound on the ints a device copy of a tree's BP schedule needs
// (ordered, parents, depth_start, prog_start, prog, leaves, leaf_w, sel_info, sel_jump).
size_t tree_device_ints(int N);

// primsMST.h:101-110: float mean and std of the image, bit-identical to the
// reference's sequential accumulation (float sum; float += double square),
// computed in parallel chunks (stats_host.cpp).
void image_stats(const float* im, long long count, float* mean, float* sd);

// primsMST.h:112-223 given the SAD edge costs ec[node + nb*N] of the (m1,n1,o1)
// control grid, the image std `sd` and the grid step.  Without `schedule`
// only ordered, parents, parent_slot and edgemst are filled.
void build_tree(const float* ec, float sd, int step, int m1, int n1, int o1, Tree& t, bool schedule = true);

// The same tree in two steps, so the image statistics need not be known when
// Prim's runs: build_tree_shape orders edges by their cost c = SAD/step^3,
// which is the reference's order by weight whenever
// weights_distinguish_costs(...) holds for the final sd; finish_tree then
// computes edgemst (and the schedule).  When the check fails, use build_tree.
// build_tree_shape runs Prim's only when the SADs are integers below 2^22
// (16-bit cost ranks, see cost_ranks): distinct non-integer costs are nearly
// all distinct floats, which the weight map, flat to a few ulps over a grid
// of 10^5 edges, then merges in places, so the check would fail and Prim's
// would run twice.  It returns false, leaving t as it was, otherwise.
bool build_tree_shape(const float* ec, int step, int m1, int n1, int o1, Tree& t);
bool weights_distinguish_costs(const float* ec, int step, int m1, int n1, int o1, float sd);
void finish_tree(const float* ec, float sd, int step, Tree& t, bool schedule = true);

// Fills depth_start, the child lists and (with bp_programs) the selection
// info and the bottom-up rounds and programs (rounds of `window` steps) from
// ordered/parents (any tree in
// BFS order, e.g. one handed in by a caller); throws std::runtime_error when
// ordered/parents do not describe one.
void build_schedule(Tree& t, bool bp_programs = true, int window = kBPWindow);

// Windowed segment rounds of one sweep direction: one launch per round, one
// workgroup per segment (round_start: rounds + 1 offsets into the segments;
// seg_start: segments + 1 offsets into nodes).
struct RoundsHost {
    std::vector<int> round_start, seg_start, nodes;
};

// The soft sweeps' tree: the MST re-rooted at its centre, in BFS order
// sorted by (depth, id), with its child lists and the rounds of both sweeps.
struct SoftSchedule {
    int maxd = 0;
    std::vector<int> ordered, parents, depth_start, child_start, child_list;
    std::vector<float> edgemst;
    RoundsHost up, down;
};

// Builds the schedule from the caller's tree (ordered / parents / edgemst
// as primsGraph returns them, primsMST.h:166-211), working in positions of
// BFS orders: a node's parent and children sit in the depths next to its
// own, so every pass walks nearby memory.
//
// Re-rooting.  The sum-product marginals do not depend on the root -- the
// pairwise term |(u0_p + q l_p) - (u0_c + q l_c)|^2 / w is symmetric in the
// two ends of an edge -- while the sequential depth of the sweeps drops
// from the height to half the diameter.  The new root is the centre: the
// node of least eccentricity (the smallest id among ties), with
// eccentricity = max(height below the node, longest path leaving it
// upwards), both from one pass each way over the given order.  Distances
// from the centre follow from the path between it and the old root (nodes
// on it reverse their edge, each edge keeping its weight).
//
// Rounds, W steps each.  Upward, every node gets a step: leaves 0; a node
// the later of t(c*) + 1, c* its child of latest step (the first among
// ties), and the start of the window after every other child's step.  It
// continues c*'s segment when that is t(c*) + 1 in the same window.
// Downward (every node but the root): the root's children start at step 0;
// a node continues its parent's segment at t(p) + 1 when it is the
// parent's child of greatest height (the first among ties) and that step
// stays in the parent's window, else it starts at the next window.  A
// segment runs in the round of its first node's window; within a round the
// longest segments come first (then BFS order).
void soft_schedule(const int* ordered, const int* parents, const float* edgemst, int N, int W, SoftSchedule& out);



}  // namespace deeds
