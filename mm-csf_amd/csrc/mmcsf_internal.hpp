// Internal declarations shared by the library's translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "mmcsf.h"

namespace mmcsf {

// ---------------------------------------------------------------------------
// errors: internal code throws mmcsf::Error; the C ABI catches it, stores the
// message for mmcsf_last_error() and returns the status code.
// ---------------------------------------------------------------------------
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string &msg);

[[noreturn]] inline void fail(int code, const std::string &msg) { throw Error(code, msg); }

#define MMCSF_HIP(call)                                                              \
    do {                                                                             \
        hipError_t e_ = (call);                                                      \
        if (e_ != hipSuccess)                                                        \
            ::mmcsf::fail(e_ == hipErrorOutOfMemory ? MMCSF_ERR_NOMEM : MMCSF_ERR_HIP, \
                          std::string(#call) + ": " + hipGetErrorString(e_) + " at " + \
                              __FILE__ + ":" + std::to_string(__LINE__));            \
    } while (0)

#define MMCSF_REQUIRE(cond, code, msg)                                               \
    do {                                                                             \
        if (!(cond)) ::mmcsf::fail((code), (msg));                                   \
    } while (0)

// ---------------------------------------------------------------------------
// owning device buffer
// ---------------------------------------------------------------------------
template <typename T>
struct DevArray {
    T *p = nullptr;
    size_t n = 0;
    DevArray() = default;
    explicit DevArray(size_t count) { alloc(count); }
    DevArray(const DevArray &) = delete;
    DevArray &operator=(const DevArray &) = delete;
    DevArray(DevArray &&o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DevArray &operator=(DevArray &&o) noexcept {
        if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
        return *this;
    }
    ~DevArray() { release(); }
    void alloc(size_t count) {
        release();
        n = count;
        if (count) MMCSF_HIP(hipMalloc(reinterpret_cast<void **>(&p), count * sizeof(T) + 16));
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    T *get() const { return p; }
};

// ---------------------------------------------------------------------------
// tensors and formats (device resident)
// ---------------------------------------------------------------------------
constexpr int kMaxDims = 4;
constexpr int kNumBins = 10;             // Options::nBin, util.h:205
constexpr uint32_t kHotBit = 0x80000000u;  // target-array tag of a privatised row
constexpr uint32_t kMaxHot = 4096;         // ranked rows per scatter level

// COO tensor: class Tensor (util.h:80-108) without the derived members.
struct Tensor {
    int ndims = 0;
    uint64_t dims[kMaxDims] = {0, 0, 0, 0};
    uint64_t nnz = 0;
    DevArray<uint32_t> inds[kMaxDims];   // per tensor mode
    DevArray<float> vals;
};

// Top-k rows of a scatter output, privatised in LDS by the scatter kernels.
struct HotTable {
    uint32_t count = 0;                  // ranked rows (<= kMaxHot)
    DevArray<uint32_t> rows;             // rank -> row
    DevArray<uint32_t> tgt;              // per contribution: row, or kHotBit|rank
};

// One compressed-sparse-fiber tree with N-1 index levels (TiledTensor,
// util.h:110-139, for one tile / one partition).  Level l = 0..N-2 has cnt[l]
// entries; ptr[l] has cnt[l]+1 entries pointing into level l+1 (into the
// nonzeros for the leaf-fiber level N-2).
struct Csf {
    int ndims = 0;
    int order[kMaxDims] = {0, 1, 2, 3};  // modeOrder
    uint64_t rows[kMaxDims] = {0, 0, 0, 0};  // dims[order[l]]: factor rows of level l
    uint64_t nnz = 0;
    uint64_t cnt[kMaxDims] = {0, 0, 0, 0};
    DevArray<uint32_t> ptr[kMaxDims - 1];
    DevArray<uint32_t> idx[kMaxDims - 1];
    DevArray<uint32_t> leaf;             // coordinate of order[N-1] per nonzero
    DevArray<float> vals;
    DevArray<uint32_t> pos;              // tensor-order position of every stored nonzero
                                         // (B-CSF and MM-CSF trees; empty otherwise)
    DevArray<uint32_t> fbrSlc;           // fbrLikeSlcInds: slice coord per level-1 entry
    DevArray<uint32_t> fbit;             // bit x: nonzero x starts a leaf fiber
    DevArray<uint32_t> sbit;             // 4-D, bit f: leaf fiber f starts a level-1 entry
    DevArray<uint32_t> chunkCost;        // exclusive prefix of the cost of 64-nonzero chunks
    // slice bins (make_TiledBin / make_HybBin)
    bool hasBins = false;
    DevArray<uint32_t> binList;          // concatenated slcMapperBin[0..9]
    uint64_t binCount[kNumBins] = {0};
    uint64_t dropped = 0;
    // privatisation tables for the scatter roles, indexed by output level
    HotTable hot[kMaxDims];               // built for levels 1..N-1 of MM-CSF
};

// HB-CSF parts (HYBTensor, util.h:141-173)
struct Hyb {
    int ndims = 0;
    int order[kMaxDims] = {0, 1, 2, 3};
    uint64_t cooNnz = 0, cslNnz = 0, nCsl = 0;
    DevArray<uint32_t> cooInd[kMaxDims];  // by level
    DevArray<float> cooVal;
    DevArray<uint32_t> cslPtr, cslIdx, cslInd1, cslInd2;
    DevArray<float> cslVal;
    DevArray<uint32_t> cslBinList;
    uint64_t cslBinCount[kNumBins] = {0};
    Csf csf;                               // CSF part (hasBins = true)
    // 3-D: the CSL part as a tree for the chunked walk (derived from the CSL
    // arrays: slices, one level-1 entry per nonzero); empty otherwise
    Csf cslWalk;
};

struct Format {
    int type = 0;
    int ndims = 0;
    uint64_t dims[kMaxDims] = {0, 0, 0, 0};
    uint64_t nnz = 0;
    int redunMode = -1;
    uint64_t fbrThreshold = 0;
    // type 3: sorted COO (levels in modeOrder order)
    int cooOrder[kMaxDims] = {0, 1, 2, 3};
    DevArray<uint32_t> cooKeys[kMaxDims];
    DevArray<float> cooVals;
    // types 8 / 12
    Csf csf;
    // type 12: a root tree per mode other than the partition's root, so every
    // mode's MTTKRP is a root traversal (internal; never exported)
    Csf tree[kMaxDims];
    bool hasTree[kMaxDims] = {false, false, false, false};
    // type 12 with the mixed-mode partition (-M 1): partition p holds the
    // nonzeros the heuristic assigns to mode p, in a tree of order
    // order_tensormode(X, p); csf is then empty
    bool mixed = false;
    Csf part[kMaxDims];
    uint64_t partNnz[kMaxDims] = {0, 0, 0, 0};
    std::vector<uint32_t> partPerNnz;   // per nonzero in the partition loop's order
    std::vector<uint32_t> partOrder;    // that order as tensor positions
    // type 10
    Hyb hyb;
};

// ---------------------------------------------------------------------------
// builder (builder.hip)
// ---------------------------------------------------------------------------
void order_tensormode(int ndims, const uint64_t *dims, int mode, bool nat, int *order);

// Sorts the tensor's nonzeros by the full key in `order`; outputs level keys
// keys[l] = coordinate of order[l], and permuted values.
// start bitmasks of a built tree (called by build_csf and build_hyb)
void build_walk_bits(Csf &c, hipStream_t s);
// pos (optional) receives the tensor-order position of every sorted nonzero.
void sort_coo(const Tensor &t, const int *order, DevArray<uint32_t> *keys, DevArray<float> &vals,
              hipStream_t s, DevArray<uint32_t> *pos = nullptr);

// CSF scan with the -f split and (4-D) fbrS split; consumes sorted keys.
void build_csf(int ndims, const uint64_t *dims, uint64_t nnz, const int *order,
               DevArray<uint32_t> *keys, DevArray<float> &&vals, uint64_t fbrThreshold,
               uint64_t fbrSThreshold, Csf &out, hipStream_t s);

void build_bins_tiled(Csf &c, hipStream_t s);           // make_TiledBin
void build_hyb(int ndims, const uint64_t *dims, Csf &&src, Hyb &out, hipStream_t s);
// the CSL part of a built 3-D HB-CSF as a walkable tree (h.cslWalk)
void build_csl_walk(Hyb &h, hipStream_t s);
void build_hot_table(const Csf &c, int level, uint64_t rows, HotTable &h, hipStream_t s);
// rows of tensor mode m holding `frac` of the nonzeros, most used first
uint64_t hot_footprint(const Tensor &t, int m, double frac, hipStream_t s);
// mixed-mode partition: level-1 entry of every nonzero of an unsplit CSF (at
// its tensor position, resolved like the reference's search over
// denseSlcPtr; 0xFFFFFFFF where that search reads before the first entry)
// and the nonzeros under each level-1 entry (cnt1 + 1 entries, the last 0)
void level1_ids(const Csf &c, const uint32_t *pos, DevArray<uint32_t> &idT, DevArray<uint32_t> &cnt, hipStream_t s);
// the nonzeros at tensor positions sel[0..n) as a tensor
void gather_tensor(const Tensor &t, const uint32_t *sel, uint64_t n, Tensor &out, hipStream_t s);

// ---------------------------------------------------------------------------
// MTTKRP launchers (mttkrp.hip)
// ---------------------------------------------------------------------------
struct Factors {
    const void *U[kMaxDims];
    int dtype;
    int R;
};

// evStart / evEnd (optional) bracket the MTTKRP kernel only, after the
// output has been zeroed.
// accumulate: add to `out` (not zeroed here; every row written with atomics)
// zeroed: `out` already holds zeros (the caller cleared it); not zeroed again
void mttkrp_csf(const Csf &c, const uint64_t *dims, int outLevel, const Factors &F, float *out,
                hipStream_t s, hipEvent_t evStart = nullptr, hipEvent_t evEnd = nullptr, bool accumulate = false,
                bool zeroed = false);
void mttkrp_coo(const Format &f, const Factors &F, float *out, hipStream_t s);
// MTTKRP of tensor mode `mode` of a CSF-based format (types 8 and 12),
// through the mode's root tree when the format has one
void mttkrp_mode(const Format &f, int mode, const Factors &F, float *out, hipStream_t s,
                 hipEvent_t evStart = nullptr, hipEvent_t evEnd = nullptr, bool zeroed = false);
// root-level MTTKRP over 64-nonzero chunks (walk.hip); needs R % 4 == 0 and
// the start bitmasks of the tree
bool walk_supports(const Csf &c, int R);
void mttkrp_root(const Csf &c, const Factors &F, float *out, hipStream_t s, bool accumulate = false);
// the walk's row loads are 16 bytes wide (8 for bfloat16 with R % 8 != 0):
// factor tables of every level but skipLevel, and `out`, must be aligned to that
void check_walk_align(const Csf &c, const Factors &F, int skipLevel, const float *out);
void mttkrp_hyb(const Hyb &h, const uint64_t *dims, const Factors &F, float *out, hipStream_t s);
// loads the code objects of the MTTKRP kernels (and, cp, of the CP
// completion passes) on the current device, once per process
void preload_kernels(bool cp);
void preload_scatter_kernels();
void fill_zero(float *p, uint64_t n, hipStream_t s);

// TTM (walk.hip): Y[leaf fiber] = sum over its nonzeros of v * U_leaf[leaf]
void ttm_fibers(const Csf &c, const Factors &F, float *Y, hipStream_t s);

// CP completion (walk.hip): loss 1/2 sum e^2 with e = model - value, residuals
// in tensor order, gradients of the root mode of a tree
bool cp_supports(int R);
void cp_forward(const Csf &c, const Factors &F, float *resid, double *loss, hipStream_t s);
void cp_gradient(const Csf &c, const Factors &F, const float *weights, float alpha, float *grad, double *loss,
                 hipStream_t s);
void cp_half(double *loss, hipStream_t s);
void cp_zero(double *loss, hipStream_t s);

// binary cache of built formats (cache.cpp)
void format_save(const Format &f, const char *path);
void format_load(const char *path, Format &f);
bool is_format_cache(const char *path);

// CP-ALS (cpd.hip)
void gram(const float *U, uint64_t rows, int R, double *G, hipStream_t s);
void cpd_als(const Format &f, int R, int iters, double tol, int dtype, float *const *U, double *lambdaHost,
             mmcsf_cpd_stats *stats, hipStream_t s);
// the same loop on the library's fixed device scratch (no allocation)
void cpd_als_noalloc(const Format &f, int R, int iters, double tol, int dtype, float *const *U, double *lambdaHost,
                     mmcsf_cpd_stats *stats, hipStream_t s);
size_t cpd_scratch_limit();

// generator (generate.hip)
void generate_tensor(int ndims, const uint64_t *dims, uint64_t nnz, const int *law,
                     const double *alpha, uint64_t seed, Tensor &t, hipStream_t s);

int device_cu_count();

}  // namespace mmcsf

// C-ABI handle types wrap the internal structures.
struct mmcsf_tensor {
    mmcsf::Tensor t;
};
struct mmcsf_format {
    mmcsf::Format f;
};
