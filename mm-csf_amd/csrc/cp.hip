// CP completion on the observed entries of a sparse tensor (no reference
// counterpart): the squared loss of a CP model
//     L = 1/2 sum_x e_x^2,   e_x = sum_r prod_n U_n[i_n(x)][r] - v_x,
// its per-entry residuals, and the factor gradients
//     dL/dU_n = MTTKRP_n(e)  (the MTTKRP of mttkrp_cpu.cpp:60-113 with the
//                             residuals in place of the values),
// each in one lane-walker traversal of a CSF tree whose root level is the
// output mode.  The walker keeps the running Hadamard product of the rows on
// the path from the root (p = U_root[i] (.) U_1[j] (. U_2[k])), so the model
// value of a nonzero is one dot product with its leaf row; the gradient
// accumulates like the root-mode MTTKRP.  Lanes of one walker group (G
// adjacent lanes, CPL columns each) combine their partial dot products with
// xor shuffles; every lane of a group runs the same iterations.
#include <algorithm>

#include "root_walk.hpp"

namespace mmcsf {
namespace {
using namespace walk;

enum : int {
    kLossGrad = 0,    // coefficient e_x recomputed; gradient (and optionally the loss)
    kLossResid = 1,   // loss and (optionally) residuals in tensor order; no gradient
    kWeightGrad = 2,  // coefficient = caller weights in tensor order; gradient only
};

struct CpArgs {
    const uint32_t *ptr[3];
    const uint32_t *idx[3];
    const uint32_t *slc;
    const uint32_t *leaf;
    const float *val;
    const uint32_t *pos;   // tensor-order position of every stored nonzero
    const float *wts;      // kWeightGrad: weights in tensor order
    const void *U[4];      // factor of each level
    float *out;            // gradient of the root mode's factor (zeroed by the caller)
    float *resid;          // kLossResid: residuals in tensor order, or null
    double *sumsq;         // sum of e^2 (accumulated), or null
    float alpha;           // scale applied to the gradient
    uint32_t nnz;
    uint32_t cnt[4];
    int R;
};

template <int G>
__device__ __forceinline__ float group_sum(float x) {
#pragma unroll
    for (int o = 1; o < G; o <<= 1) x += __shfl_xor(x, o, 64);
    return x;
}

template <int CPL>
__device__ __forceinline__ void had(float *d, const float *a, const float *b) {
#pragma unroll
    for (int c = 0; c < CPL; ++c) d[c] = a[c] * b[c];
}

template <int CPL>
__device__ __forceinline__ void fma_had(float *d, const float *a, const float *b) {
#pragma unroll
    for (int c = 0; c < CPL; ++c) d[c] = __builtin_fmaf(a[c], b[c], d[c]);
}

template <int N, int CPL, int G, typename T, int MODE>
__global__ __launch_bounds__(256) void k_cp(CpArgs a) {
    static_assert(G > 0 && (G & (G - 1)) == 0 && G <= 64, "walker groups stay inside one wave");
    using Row = LaneRow<CPL, T>;
    constexpr bool kModel = MODE != kWeightGrad;
    constexpr bool kGrad = MODE != kLossResid;
    constexpr int LF = N - 2;
    const int R = a.R;
    const int sub = threadIdx.x % G;
    const int c0 = sub * CPL;
    const uint64_t walkers = (uint64_t)gridDim.x * blockDim.x / G;
    const uint64_t w = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / G;
    const uint32_t x0 = (uint32_t)(w * a.nnz / walkers);
    const uint32_t x1 = (uint32_t)((w + 1) * a.nnz / walkers);
    const T *U0 = static_cast<const T *>(a.U[0]);
    const T *U1 = static_cast<const T *>(a.U[1]);
    const T *U2 = static_cast<const T *>(a.U[2]);
    const T *UL = static_cast<const T *>(a.U[N - 1]);
    const uint32_t *__restrict__ ptrL = a.ptr[LF];
    const uint32_t *__restrict__ leaf = a.leaf;

    float aS[CPL];
#pragma unroll
    for (int c = 0; c < CPL; ++c) aS[c] = 0.f;
    uint32_t slc = 0xffffffffu;
    double sq = 0.0;

    auto coef_in = [&](uint32_t x) -> float {
        if constexpr (MODE == kWeightGrad) return a.wts[a.pos[x]];
        else return a.val[x];
    };

    if (x0 < x1) {
        Row Q, W1, W2, pJ, p, aF, aJ, row;
        aF.zero();
        aJ.zero();
        uint32_t f = upper_idx(ptrL, a.cnt[LF], x0);
        uint32_t fend = ptrL[f + 1];
        uint32_t j = f, jend = 0;
        if constexpr (N == 4) {
            j = upper_idx(a.ptr[1], a.cnt[1], f);
            jend = a.ptr[1][j + 1];
        }
        slc = a.slc[j];
        if constexpr (kModel) Q.load(U0, slc, R, c0);
        W1.load(U1, a.idx[1][j], R, c0);
        if constexpr (N == 4) {
            W2.load(U2, a.idx[2][f], R, c0);
            if constexpr (kModel) {
                had<CPL>(pJ.v, Q.v, W1.v);
                had<CPL>(p.v, pJ.v, W2.v);
            }
        } else if constexpr (kModel) {
            had<CPL>(p.v, Q.v, W1.v);
        }
        uint32_t li = leaf[x0];
        float v = coef_in(x0);
        for (uint32_t x = x0; x < x1; ++x) {
            uint32_t liN = 0;
            float vN = 0.f;
            if (x + 1 < x1) {
                liN = leaf[x + 1];
                vN = coef_in(x + 1);
            }
            row.load(UL, li, R, c0);
            if (x == fend) {
                // close fiber f (and its fbrS), open fiber f + 1
                if constexpr (kGrad) {
                    if constexpr (N == 3) fma_had<CPL>(aS, aF.v, W1.v);
                    else fma_had<CPL>(aJ.v, aF.v, W2.v);
                }
                aF.zero();
                ++f;
                fend = ptrL[f + 1];
                bool newJ = true;
                uint32_t ns;
                if constexpr (N == 4) {
                    newJ = f == jend;
                    if (newJ) {
                        if constexpr (kGrad) fma_had<CPL>(aS, aJ.v, W1.v);
                        aJ.zero();
                        ++j;
                        jend = a.ptr[1][j + 1];
                    }
                    ns = a.slc[j];
                } else {
                    j = f;
                    ns = a.slc[f];
                }
                if (ns != slc) {
                    if constexpr (kGrad) {
                        atomic_row<CPL>(a.out, slc, R, c0, aS);
#pragma unroll
                        for (int c = 0; c < CPL; ++c) aS[c] = 0.f;
                    }
                    slc = ns;
                    if constexpr (kModel) Q.load(U0, slc, R, c0);
                }
                if (newJ) {
                    W1.load(U1, a.idx[1][j], R, c0);
                    if constexpr (kModel) {
                        if constexpr (N == 4) had<CPL>(pJ.v, Q.v, W1.v);
                        else had<CPL>(p.v, Q.v, W1.v);
                    }
                }
                if constexpr (N == 4) {
                    W2.load(U2, a.idx[2][f], R, c0);
                    if constexpr (kModel) had<CPL>(p.v, pJ.v, W2.v);
                }
            }
            float e = v;
            if constexpr (kModel) {
                float m = 0.f;
#pragma unroll
                for (int c = 0; c < CPL; ++c) m = __builtin_fmaf(p.v[c], row.v[c], m);
                m = group_sum<G>(m);
                e = m - v;
                if (sub == 0) {
                    sq += (double)e * (double)e;
                    if constexpr (MODE == kLossResid)
                        if (a.resid) a.resid[a.pos[x]] = e;
                }
            }
            if constexpr (kGrad) {
                const float ea = e * a.alpha;
#pragma unroll
                for (int c = 0; c < CPL; ++c) aF.v[c] = __builtin_fmaf(ea, row.v[c], aF.v[c]);
            }
            li = liN;
            v = vN;
        }
        if constexpr (kGrad) {
            if constexpr (N == 3) {
                fma_had<CPL>(aS, aF.v, W1.v);
            } else {
                fma_had<CPL>(aJ.v, aF.v, W2.v);
                fma_had<CPL>(aS, aJ.v, W1.v);
            }
        }
    }
    else if (x0 < a.nnz) {
        // a walker without nonzeros takes the slice at its position, so the
        // slice keys stay non-decreasing across the wave for the reduction
        uint32_t f = upper_idx(ptrL, a.cnt[LF], x0);
        if constexpr (N == 4) f = upper_idx(a.ptr[1], a.cnt[1], f);
        slc = a.slc[f];
    }
    const int lane = threadIdx.x & 63;
    if constexpr (kGrad) {
        // segmented reduction of the last slice across the walkers of the wave
#pragma unroll
        for (int off = G; off < 64; off <<= 1) {
            uint32_t ko = __shfl_down(slc, off, 64);
            bool take = (lane + off < 64) && ko == slc;
#pragma unroll
            for (int c = 0; c < CPL; ++c) {
                float o = __shfl_down(aS[c], off, 64);
                if (take) aS[c] += o;
            }
        }
        uint32_t kp = __shfl_up(slc, G, 64);
        bool head = (lane < G) || kp != slc;
        if (head && slc != 0xffffffffu) atomic_row<CPL>(a.out, slc, R, c0, aS);
    }
    if constexpr (kModel) {
        if (a.sumsq) {
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) sq += __shfl_down(sq, off, 64);
            __shared__ double part[4];
            if (lane == 0) part[threadIdx.x >> 6] = sq;
            __syncthreads();
            if (threadIdx.x == 0) {
                double t = 0.0;
                for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t += part[i];
                unsafeAtomicAdd(a.sumsq, t);
            }
        }
    }
}

__global__ void k_zero_f64(double *p) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = 0.0;
}

__global__ void k_half(double *p) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p *= 0.5;
}

struct Cfg {
    int CPL, G;
};

Cfg pick(int N, int R) {
    if (R <= 4) return {4, 1};
    if (R <= 8) return {8, 1};
    if (N == 3) {
        if (R <= 16) return {16, 1};
        if (R <= 32) return {16, 2};
        if (R <= 64) return {16, 4};
        return {16, 8};
    }
    if (R <= 16) return {8, 2};
    if (R <= 32) return {8, 4};
    if (R <= 64) return {8, 8};
    return {8, 16};
}

template <int N, int CPL, int G, typename T, int MODE>
void go(const CpArgs &a, hipStream_t s) {
    auto fn = k_cp<N, CPL, G, T, MODE>;
    int perCU = 0;
    MMCSF_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, fn, 256, 0));
    perCU = std::max(perCU, 1);
    uint64_t blocks = (uint64_t)device_cu_count() * perCU;
    uint64_t maxBlocks = std::max<uint64_t>(1, ((uint64_t)a.nnz * G + 255) / 256);
    blocks = std::min(blocks, maxBlocks);
    fn<<<dim3((unsigned)blocks), dim3(256), 0, s>>>(a);
    MMCSF_HIP(hipGetLastError());
}

template <int N, typename T, int MODE>
void dispatch(const Cfg &c, const CpArgs &a, hipStream_t s) {
#define MMCSF_CP(CC, GG)                  \
    if (c.CPL == CC && c.G == GG) {       \
        go<N, CC, GG, T, MODE>(a, s);     \
        return;                           \
    }
    MMCSF_CP(4, 1)
    MMCSF_CP(8, 1)
    if constexpr (N == 3) {
        MMCSF_CP(16, 1)
        MMCSF_CP(16, 2)
        MMCSF_CP(16, 4)
        MMCSF_CP(16, 8)
    } else {
        MMCSF_CP(8, 2)
        MMCSF_CP(8, 4)
        MMCSF_CP(8, 8)
        MMCSF_CP(8, 16)
    }
#undef MMCSF_CP
    fail(MMCSF_ERR_UNSUPPORTED, "no CP-completion kernel for this rank");
}

template <int MODE>
void launch(const Csf &c, const Factors &F, CpArgs a, hipStream_t s) {
    const int N = c.ndims;
    const Cfg cfg = pick(N, F.R);
    if (N == 3) {
        if (F.dtype == MMCSF_F32) dispatch<3, float, MODE>(cfg, a, s);
        else dispatch<3, uint16_t, MODE>(cfg, a, s);
    } else {
        if (F.dtype == MMCSF_F32) dispatch<4, float, MODE>(cfg, a, s);
        else dispatch<4, uint16_t, MODE>(cfg, a, s);
    }
}

CpArgs args_of(const Csf &c, const Factors &F) {
    const int N = c.ndims;
    CpArgs a{};
    for (int l = 0; l < N - 1; ++l) {
        a.ptr[l] = c.ptr[l].p;
        a.idx[l] = c.idx[l].p;
        a.cnt[l] = (uint32_t)c.cnt[l];
    }
    a.cnt[N - 1] = (uint32_t)c.nnz;
    a.slc = c.fbrSlc.p;
    a.leaf = c.leaf.p;
    a.val = c.vals.p;
    a.pos = c.pos.p;
    for (int l = 0; l < N; ++l) a.U[l] = F.U[c.order[l]];
    a.nnz = (uint32_t)c.nnz;
    a.R = F.R;
    a.alpha = 1.f;
    return a;
}

}  // namespace

bool cp_supports(int R) { return R >= 1 && R <= 128 && R % 4 == 0; }

void cp_forward(const Csf &c, const Factors &F, float *resid, double *loss, hipStream_t s) {
    hipLaunchKernelGGL(k_zero_f64, dim3(1), dim3(64), 0, s, loss);
    MMCSF_HIP(hipGetLastError());
    if (c.nnz == 0) return;
    CpArgs a = args_of(c, F);
    a.resid = resid;
    a.sumsq = loss;
    launch<kLossResid>(c, F, a, s);
    hipLaunchKernelGGL(k_half, dim3(1), dim3(64), 0, s, loss);
    MMCSF_HIP(hipGetLastError());
}

void cp_gradient(const Csf &c, const Factors &F, const float *weights, float alpha, float *grad, double *loss,
                 hipStream_t s) {
    if (c.nnz == 0) return;
    CpArgs a = args_of(c, F);
    a.out = grad;
    a.alpha = alpha;
    if (weights) {
        a.wts = weights;
        launch<kWeightGrad>(c, F, a, s);
    } else {
        a.sumsq = loss;
        launch<kLossGrad>(c, F, a, s);
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
