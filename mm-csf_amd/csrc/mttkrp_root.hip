// Root-mode MTTKRP with one walker per lane (gfx950).
//
// Every lane (or G adjacent lanes when R > 32) is an independent walker over
// a contiguous, equal share of the nonzeros of a CSF tree whose root level is
// the output mode.  A walker keeps CPL columns of its running sums in
// registers:
//     aF += v * U_leaf[leaf]                      per nonzero
//     aJ += aF (.) U_{N-2}[idx_{N-2}]             per leaf fiber   (4-D)
//     aS += a. (.) U_1[idx_1]                     per level-1 entry
// and writes aS to the slice's output row when the slice ends.  A slice that
// ends inside the walker's share is flushed with float atomics; the slice a
// walker ends in is shared with the next walkers, so those partial rows are
// first summed across the wave with a segmented shuffle reduction (slices are
// non-decreasing across lanes) and the segment head issues the atomics.
// Lanes never exchange data on the hot loop, so a wave holds 64/G fully
// independent gather streams.
#include <algorithm>
#include <cstdlib>
#include <type_traits>

#include "root_walk.hpp"

namespace mmcsf {

struct RootArgs {
    const uint32_t *ptr[3];
    const uint32_t *idx[3];
    const uint32_t *slc;   // slice coordinate of each level-1 entry
    const uint32_t *leaf;
    const float *val;
    const void *U[4];      // factor of each level
    float *out;
    uint32_t nnz;
    uint32_t cnt[4];
    int R;
};

namespace {
using namespace walk;

template <int N, int CPL, int G, typename T>
__global__ __launch_bounds__(256) void k_root(RootArgs a) {
    using Row = LaneRow<CPL, T>;
    constexpr int LF = N - 2;
    const int R = a.R;
    static_assert(G > 0 && (G & (G - 1)) == 0 && G <= 64, "walker groups stay inside one wave");
    const int sub = threadIdx.x % G;
    const int c0 = sub * CPL;
    const uint64_t walkers = (uint64_t)gridDim.x * blockDim.x / G;
    const uint64_t w = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / G;
    const uint32_t x0 = (uint32_t)(w * a.nnz / walkers);
    const uint32_t x1 = (uint32_t)((w + 1) * a.nnz / walkers);
    const T *U1 = static_cast<const T *>(a.U[1]);
    const T *U2 = static_cast<const T *>(a.U[2]);
    const T *UL = static_cast<const T *>(a.U[N - 1]);
    const uint32_t *__restrict__ ptrL = a.ptr[LF];
    const uint32_t *__restrict__ leaf = a.leaf;
    const float *__restrict__ val = a.val;
    float *__restrict__ out = a.out;

    float aS[CPL];
#pragma unroll
    for (int c = 0; c < CPL; ++c) aS[c] = 0.f;
    uint32_t slc = 0xffffffffu;  // key of lanes without work: sorts after every slice

    if (x0 < x1) {
        Row aF, aJ, W, row;
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
        uint32_t li = leaf[x0];
        float v = val[x0];
        for (uint32_t x = x0; x < x1; ++x) {
            // prefetch the next nonzero's index and value
            uint32_t liN = 0;
            float vN = 0.f;
            if (x + 1 < x1) {
                liN = leaf[x + 1];
                vN = val[x + 1];
            }
            row.load(UL, li, R, c0);
            if (x == fend) {
                if constexpr (N == 3) {
                    W.load(U1, a.idx[1][f], R, c0);
#pragma unroll
                    for (int c = 0; c < CPL; ++c) aS[c] = __builtin_fmaf(aF.v[c], W.v[c], aS[c]);
                } else {
                    W.load(U2, a.idx[2][f], R, c0);
#pragma unroll
                    for (int c = 0; c < CPL; ++c) aJ.v[c] = __builtin_fmaf(aF.v[c], W.v[c], aJ.v[c]);
                }
                aF.zero();
                ++f;
                fend = ptrL[f + 1];
                uint32_t ns = slc;
                if constexpr (N == 4) {
                    if (f == jend) {
                        W.load(U1, a.idx[1][j], R, c0);
#pragma unroll
                        for (int c = 0; c < CPL; ++c) aS[c] = __builtin_fmaf(aJ.v[c], W.v[c], aS[c]);
                        aJ.zero();
                        ++j;
                        jend = a.ptr[1][j + 1];
                        ns = a.slc[j];
                    }
                } else {
                    ns = a.slc[f];
                }
                if (ns != slc) {
                    atomic_row<CPL>(out, slc, R, c0, aS);
#pragma unroll
                    for (int c = 0; c < CPL; ++c) aS[c] = 0.f;
                    slc = ns;
                }
            }
#pragma unroll
            for (int c = 0; c < CPL; ++c) aF.v[c] = __builtin_fmaf(v, row.v[c], aF.v[c]);
            li = liN;
            v = vN;
        }
        // close the fiber (and fbrS) the share ends in
        if constexpr (N == 3) {
            W.load(U1, a.idx[1][f], R, c0);
#pragma unroll
            for (int c = 0; c < CPL; ++c) aS[c] = __builtin_fmaf(aF.v[c], W.v[c], aS[c]);
        } else {
            W.load(U2, a.idx[2][f], R, c0);
#pragma unroll
            for (int c = 0; c < CPL; ++c) aJ.v[c] = __builtin_fmaf(aF.v[c], W.v[c], aJ.v[c]);
            W.load(U1, a.idx[1][j], R, c0);
#pragma unroll
            for (int c = 0; c < CPL; ++c) aS[c] = __builtin_fmaf(aJ.v[c], W.v[c], aS[c]);
        }
    }
    else if (x0 < a.nnz) {
        // a walker without nonzeros takes the slice at its position, so the
        // slice keys stay non-decreasing across the wave for the reduction
        uint32_t f = upper_idx(ptrL, a.cnt[LF], x0);
        if constexpr (N == 4) f = upper_idx(a.ptr[1], a.cnt[1], f);
        slc = a.slc[f];
    }
    // segmented reduction of the last slice across the walkers of the wave
    const int lane = threadIdx.x & 63;
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
    if (head && slc != 0xffffffffu) atomic_row<CPL>(out, slc, R, c0, aS);
}

struct Cfg {
    int CPL, G;
};

Cfg pick(int R) {
    if (const char *e = std::getenv("MMCSF_ROOT_CPL")) {  // tuning experiment
        int cpl = std::atoi(e);
        if ((cpl == 8 || cpl == 16) && R % cpl == 0 && R / cpl <= 8) return {cpl, R / cpl};
    }
    if (R <= 4) return {4, 1};
    if (R <= 8) return {8, 1};
    if (R <= 16) return {16, 1};
    if (R <= 32) return {32, 1};
    if (R <= 64) return {32, 2};
    if (R <= 128) return {32, 4};
    return {32, 8};
}

template <int N, int CPL, int G, typename T>
void go(const RootArgs &a, hipStream_t s) {
    auto fn = k_root<N, CPL, G, T>;
    int perCU = 0;
    MMCSF_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, fn, 256, 0));
    perCU = std::max(perCU, 1);
    // enough walkers to fill the device, never more walkers than nonzeros
    uint64_t blocks = (uint64_t)device_cu_count() * perCU;
    uint64_t maxBlocks = std::max<uint64_t>(1, ((uint64_t)a.nnz * G + 255) / 256);
    blocks = std::min(blocks, maxBlocks);
    fn<<<dim3((unsigned)blocks), dim3(256), 0, s>>>(a);
    MMCSF_HIP(hipGetLastError());
}

template <int N, typename T>
void dispatch(const Cfg &c, const RootArgs &a, hipStream_t s) {
#define MMCSF_ROOT(CC, GG)                        \
    if (c.CPL == CC && c.G == GG) {               \
        go<N, CC, GG, T>(a, s);                   \
        return;                                   \
    }
    MMCSF_ROOT(4, 1)
    MMCSF_ROOT(8, 1)
    MMCSF_ROOT(16, 1)
    MMCSF_ROOT(8, 2)
    MMCSF_ROOT(8, 4)
    MMCSF_ROOT(16, 2)
    MMCSF_ROOT(16, 4)
    MMCSF_ROOT(32, 1)
    MMCSF_ROOT(32, 2)
    MMCSF_ROOT(32, 4)
    MMCSF_ROOT(32, 8)
#undef MMCSF_ROOT
    fail(MMCSF_ERR_UNSUPPORTED, "no root kernel for this rank");
}

}  // namespace

bool root_walker_supports(int R) { return R >= 1 && R <= 256 && R % 4 == 0; }

void mttkrp_root(const Csf &c, const Factors &F, float *out, hipStream_t s) {
    const char *e = std::getenv("MMCSF_ROOT_KERNEL");  // tuning experiment
    if (chunk_supports(c, F.R) && !(e && e[0] == 'l')) {
        mttkrp_chunk(c, F, out, s);
        return;
    }
    const int N = c.ndims;
    RootArgs a{};
    for (int l = 0; l < N - 1; ++l) {
        a.ptr[l] = c.ptr[l].p;
        a.idx[l] = c.idx[l].p;
        a.cnt[l] = (uint32_t)c.cnt[l];
    }
    a.cnt[N - 1] = (uint32_t)c.nnz;
    a.slc = c.fbrSlc.p;
    a.leaf = c.leaf.p;
    a.val = c.vals.p;
    for (int l = 0; l < N; ++l) a.U[l] = F.U[c.order[l]];
    a.out = out;
    a.nnz = (uint32_t)c.nnz;
    a.R = F.R;
    const Cfg cfg = pick(F.R);
    if (N == 3) {
        if (F.dtype == MMCSF_F32) dispatch<3, float>(cfg, a, s);
        else dispatch<3, uint16_t>(cfg, a, s);
    } else {
        if (F.dtype == MMCSF_F32) dispatch<4, float>(cfg, a, s);
        else dispatch<4, uint16_t>(cfg, a, s);
    }
}

}  // namespace mmcsf
