// Host side of the device belief propagation: staging a tree's schedule into
// HBM and launching BP for one tree or for a batch of trees (api.cpp).
#pragma once

#include <vector>

#include "kernels.h"
#include "mst_host.h"

namespace deeds {

// Bytes of a tree's staging area (schedule ints, then 3*N floats of scratch).
size_t tree_stage_bytes(int N);
// Packs t's schedule into `pinned` (host work only) and returns the device
// view of it at d_stage; *bytes receives the size the copy must move.
BPDevTree pack_tree(const Tree& t, void* pinned, void* d_stage, size_t* bytes);
// pack_tree, then the copy into d_stage enqueued on s; `pinned` must stay
// untouched until the copy has executed.
BPDevTree stage_tree(const Tree& t, void* pinned, void* d_stage, hipStream_t s);
// Bottom-up rounds, then the top-down selection, of one tree; d_comp: null
// or N*len^3 int16 of scratch for the selection's composed tables.
void run_bp(const Tree& t, const BPDevTree& dt, float* d_cost, int16_t* d_inds, int16_t* d_comp, const float* u0,
            const float* v0, const float* w0, float* u1, float* v1, float* w1, int hw, float quant, hipStream_t s);

// One tree of a batched BP: its host schedule, its staged device copy, its
// cost volume / argmin tables and its fields.
struct BPBatchItem {
    const Tree* tree;
    BPDevTree dt;
    float* cost;
    int16_t* inds;
    int16_t* comp;  // null or scratch for composed tables (used when the item runs alone)
    const float *u0, *v0, *w0;
    float *u1, *v1, *w1;
};
// Staging bytes (pinned and device alike) of a batch of `trees` trees with
// `nodes` nodes in total.
size_t bp_batch_bytes(int trees, long long nodes);
// BP of every item (same hw and quant) in one launch sequence on s: the
// descriptors and the round work list are packed into `pinned` (untouched
// until the copy has executed) and copied into d_stage (`bytes` each).
void run_bp_batch(const std::vector<BPBatchItem>& items, int hw, float quant, void* pinned, void* d_stage,
                  size_t bytes, hipStream_t s);

}  // namespace deeds
