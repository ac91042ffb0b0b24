// Entry points of the chunked root walk (the kernel is in walk_kernel.hpp,
// instantiated per operation in walk_mttkrp.hip, walk_loss.hip,
// walk_lossgrad.hip, walk_wgrad.hip and walk_ttm.hip).
#include <mutex>

#include "walk.hpp"

namespace mmcsf {

namespace {

WalkArgs args_of(const Csf &c, const Factors &F) {
    const int N = c.ndims;
    WalkArgs a{};
    for (int l = 0; l < N - 1; ++l) {
        a.ptr[l] = c.ptr[l].p;
        a.idx[l] = c.idx[l].p;
        a.cnt[l] = (uint32_t)c.cnt[l];
    }
    a.cnt[N - 1] = (uint32_t)c.nnz;
    a.slc = c.fbrSlc.p;
    a.leaf = c.leaf.p;
    a.val = c.vals.p;
    a.fbit = c.fbit.p;
    a.sbit = c.sbit.p;
    a.cost = c.chunkCost.p;
    a.pos = c.pos.p;
    const uint64_t es = F.dtype == MMCSF_F32 ? 4 : 2;
    for (int l = 0; l < N; ++l) {
        a.U[l] = F.U[c.order[l]];
        a.tabBytes[l] = (uint32_t)(c.rows[l] * (uint64_t)F.R * es);
    }
    a.nnz = (uint32_t)c.nnz;
    a.R = F.R;
    a.alpha = 1.f;
    return a;
}

__global__ void k_zero_f64(double *p) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = 0.0;
}

__global__ void k_half(double *p) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p *= 0.5;
}

}  // namespace

// the walk addresses factor rows by 32-bit byte offsets and masks the rows it
// does not need with out-of-range buffer offsets: every table of the tree's
// levels must stay below 4 GiB - 4 KiB (counted at 4 bytes per element)
static bool tables_fit(const Csf &c, int R) {
    for (int l = 0; l < c.ndims; ++l)
        if (c.rows[l] * (uint64_t)R * 4u > 0xFFFFF000ull) return false;
    return true;
}

// row loads of the walk are VEC elements wide (16 bytes, or 8 bytes for
// bfloat16 rows with R % 8 != 0); rows of `out` are written as float4
void check_walk_align(const Csf &c, const Factors &F, int skipLevel, const float *out) {
    const uintptr_t need = (F.dtype == MMCSF_BF16 && F.R % 8 != 0) ? 8 : 16;
    for (int l = 0; l < c.ndims; ++l) {
        if (l == skipLevel) continue;
        const void *p = F.U[c.order[l]];
        MMCSF_REQUIRE(p != nullptr, MMCSF_ERR_ARG, "missing factor matrix");
        MMCSF_REQUIRE(reinterpret_cast<uintptr_t>(p) % need == 0, MMCSF_ERR_ARG,
                      "factor pointer is not aligned to " + std::to_string(need) + " bytes");
    }
    MMCSF_REQUIRE(!out || reinterpret_cast<uintptr_t>(out) % 16 == 0, MMCSF_ERR_ARG,
                  "output pointer is not aligned to 16 bytes");
}

bool walk_supports(const Csf &c, int R) {
    return R >= 1 && R <= 256 && R % 4 == 0 && tables_fit(c, R) &&
           (c.nnz == 0 || (c.fbit.p && c.chunkCost.p && (c.ndims == 3 || c.sbit.p)));
}

void mttkrp_root(const Csf &c, const Factors &F, float *out, hipStream_t s, bool accumulate) {
    if (c.nnz == 0) return;
    MMCSF_REQUIRE(walk_supports(c, F.R), MMCSF_ERR_UNSUPPORTED,
                  "root walk needs R % 4 == 0, start bitmasks and factor tables below 4 GiB");
    check_walk_align(c, F, 0, out);
    WalkArgs a = args_of(c, F);
    a.out = out;
    a.accum = accumulate ? 1 : 0;
    walk_launch<kMttkrp>(c, F, a, s);
}

bool cp_supports(int R) { return R >= 1 && R <= 128 && R % 4 == 0; }

void cp_forward(const Csf &c, const Factors &F, float *resid, double *loss, hipStream_t s) {
    MMCSF_REQUIRE(tables_fit(c, F.R), MMCSF_ERR_UNSUPPORTED, "CP completion needs factor tables below 4 GiB");
    hipLaunchKernelGGL(k_zero_f64, dim3(1), dim3(64), 0, s, loss);
    MMCSF_HIP(hipGetLastError());
    if (c.nnz > 0) {
        WalkArgs a = args_of(c, F);
        a.resid = resid;
        a.sumsq = loss;
        walk_launch<kLossResid>(c, F, a, s);
    }
    hipLaunchKernelGGL(k_half, dim3(1), dim3(64), 0, s, loss);
    MMCSF_HIP(hipGetLastError());
}

void cp_gradient(const Csf &c, const Factors &F, const float *weights, float alpha, float *grad, double *loss,
                 hipStream_t s) {
    if (c.nnz == 0) return;
    MMCSF_REQUIRE(tables_fit(c, F.R), MMCSF_ERR_UNSUPPORTED, "CP completion needs factor tables below 4 GiB");
    WalkArgs a = args_of(c, F);
    a.out = grad;
    a.alpha = alpha;
    if (weights) {
        a.wts = weights;
        walk_launch<kWeightGrad>(c, F, a, s);
    } else {
        a.sumsq = loss;
        walk_launch<kLossGrad>(c, F, a, s);
    }
}

void ttm_fibers(const Csf &c, const Factors &F, float *Y, hipStream_t s) {
    if (c.nnz == 0) return;
    MMCSF_REQUIRE(walk_supports(c, F.R), MMCSF_ERR_UNSUPPORTED, "TTM needs R % 4 == 0, R <= 256 and factor tables below 4 GiB");
    const int leaf = c.ndims - 1;
    const uintptr_t need = (F.dtype == MMCSF_BF16 && F.R % 8 != 0) ? 8 : 16;
    MMCSF_REQUIRE(F.U[c.order[leaf]] && reinterpret_cast<uintptr_t>(F.U[c.order[leaf]]) % need == 0, MMCSF_ERR_ARG,
                  "factor pointer is not aligned to " + std::to_string(need) + " bytes");
    WalkArgs a = args_of(c, F);
    a.out = Y;
    walk_launch<kTtm>(c, F, a, s);
}

void preload_kernels(bool cp) {
    // once per device: the code objects of the traversals a format runs
    static std::mutex lock;
    static bool done[64][2];
    int dev = 0;
    MMCSF_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return;
    std::lock_guard<std::mutex> g(lock);
    if (!done[dev][0]) {
        walk_preload<kMttkrp>();
        preload_scatter_kernels();
        done[dev][0] = true;
    }
    if (cp && !done[dev][1]) {
        walk_preload<kLossResid>();
        walk_preload<kLossGrad>();
        walk_preload<kWeightGrad>();
        done[dev][1] = true;
    }
}

void cp_half(double *loss, hipStream_t s) {
    hipLaunchKernelGGL(k_half, dim3(1), dim3(64), 0, s, loss);
    MMCSF_HIP(hipGetLastError());
}

void cp_zero(double *loss, hipStream_t s) {
    hipLaunchKernelGGL(k_zero_f64, dim3(1), dim3(64), 0, s, loss);
    MMCSF_HIP(hipGetLastError());
}

}  // namespace mmcsf
