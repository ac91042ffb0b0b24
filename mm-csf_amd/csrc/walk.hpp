// Root-tree traversals over 64-nonzero chunks (gfx950, wave64): the root-mode
// MTTKRP, the CP-completion loss / residual / gradient passes and the
// leaf-mode TTM.
//
// A CSF tree whose root level is the output mode is cut into chunks of 64
// consecutive nonzeros; every wave owns a contiguous run of chunks whose
// traversal cost (a prefix over chunks) is an equal share of the total.
// Per chunk:
//   * lane p reads leaf[x], val[x] and the fiber-start bits of nonzero
//     x = chunk*64 + p (coalesced);
//   * the leaf fiber of every nonzero is the running count of fiber starts
//     (a wave ballot and a prefix popcount); 4-D trees count level-1 starts
//     the same way over the fibers; the level-1 entry gives the fiber index
//     and the slice;
//   * groups of G lanes gather whole factor rows (16 or 8 bytes per
//     lane-load); group g takes the run of nonzeros g*G .. g*G + G-1 of the
//     chunk, and the runs are joined by a segmented scan over the groups;
//   * MTTKRP: contribution v * U_leaf[leaf] (.) U_1[idx_1] (. U_2[idx_2]),
//     factored by fiber: inside a run, v * U_leaf[leaf] is summed per leaf
//     fiber and multiplied by the fiber's upper-level rows once, which are
//     read only where a fiber segment of the run begins (out-of-range buffer
//     loads for the other lanes, so every load stays unconditional);
//     factoring is off for 4-D trees with 8-lane groups (leaf fibers of
//     1-1.5 nonzeros); CP, per nonzero: the model value
//     <U_root[i], U_1[j] (. U_2[k]) (.) U_leaf[l]> is a
//     dot product reduced over the group, e = model - v, the loss sums e^2
//     in float64, the gradient contribution is alpha * e * (the MTTKRP
//     product); TTM: v * U_leaf[leaf] summed per leaf fiber;
//   * a slice that begins and ends inside the wave's run is stored; only
//     slices crossing a wave boundary are added with float atomics;
//   * reads are software-pipelined over chunks: the level-1 entries of chunk
//     c+1 and the words of chunk c+2 are read while chunk c's rows are in
//     flight.
//   * the wave order is XCD-aware: each of the 8 XCDs (which take workgroups
//     round-robin) runs 16 spans of consecutive blocks.
// Bound: HBM for the structure (8 bytes per nonzero, 8 per leaf fiber, 4 per
// level-1 entry) and the L2 / MALL gather rate for factor rows.
#pragma once
#include <cstdint>

#include "mmcsf_internal.hpp"

namespace mmcsf {

struct WalkArgs {
    const uint32_t *ptr[3];
    const uint32_t *idx[3];
    const uint32_t *slc;   // slice coordinate of each level-1 entry
    const uint32_t *leaf;
    const float *val;
    const uint32_t *fbit;  // leaf-fiber starts over nonzeros
    const uint32_t *sbit;  // 4-D: level-1 starts over leaf fibers
    const uint32_t *cost;  // exclusive cost prefix over chunks (chunks + 1 entries)
    const uint32_t *pos;   // tensor-order position of every stored nonzero
    const float *wts;      // weights in tensor order
    const void *U[4];      // factor of each level
    uint32_t tabBytes[4];  // bytes of each level's factor table (buffer range)
    float *out;            // root-mode rows (zeroed by the caller)
    float *resid;          // residuals in tensor order, or null
    double *sumsq;         // sum of e^2 (accumulated), or null
    float alpha;
    int accum;             // add every row with atomics (other writers share `out`)
    uint32_t nnz;
    uint32_t cnt[4];
    int R;
};

enum WalkOp : int {
    kMttkrp = 0,      // out += v * product of the other rows
    kLossGrad = 1,    // out += alpha * e * product; sumsq += e^2
    kLossResid = 2,   // sumsq += e^2; resid[pos] = e
    kWeightGrad = 3,  // out += alpha * wts[pos] * product
    kTtm = 4,         // out[leaf fiber] += v * U_leaf[leaf] (tensor-times-matrix)
};

// one traversal of the root tree of c; instantiated per operation in its
// own translation unit (walk_<op>.hip)
template <int OP>
void walk_launch(const Csf &c, const Factors &F, const WalkArgs &a, hipStream_t s);
template <int OP>
void walk_preload();

}  // namespace mmcsf
