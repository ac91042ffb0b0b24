// MTTKRP kernels and launchers (gfx950).
//
// Reference kernels these replace (mttkrp_gpu.cu):
//   MM-CSF root / fbrS / fiber / leaf output: mttkrp_MIHCSR_kernel_slc_atomic_fbrLvlPar[_4D]
//     (716, 774), ..._fbrS_atomic_fbrLvlPar_4D (1134), ..._fbr_atomic_fbrLvlPar[_4D]
//     (1060, 1095), ..._all_atomic_fbrLvlPar[_4D] (1438, 1475); chosen by the output
//     mode's level in modeOrder as in MTTKRP_MIHCSR_GPU (3610-3719).
//   B-CSF: mttkrp_HCSR_kernel_smllBin / _hvyBin [_4D] (539-713): root output.
//   HB-CSF: mttkrp_HYB_COO_kernel[_4D] (196, 255), mttkrp_CSL_kernel_bin / _hvyBin
//     (346, 427) and the B-CSF kernels on the CSF part.
//   COO: mttkrp_COO_kernel_loop[_4D] (102, 161).
// The reference kernels accumulate every r of a lane into one scalar and are
// only correct for R <= 32 (e.g. mttkrp_gpu.cu:487-494); these keep one
// accumulator per column and are correct for any R up to 256.
#include <algorithm>

#include "mmcsf_internal.hpp"
#include "mttkrp_kernels.hpp"

namespace mmcsf {

namespace {

using dev::Row;
using dev::upper_index;

constexpr int kBlock = 256;

// --------------------------------------------------------------------------
// CSF walker: N = 3 or 4 levels of coordinates, OUT = level of the output mode.
// Persistent: every group of G lanes owns one contiguous, equal share of the
// nonzeros.  Scatter roles (OUT > 0) accumulate the top-ranked output rows in
// LDS and flush them once per block; other rows take global atomics.
// --------------------------------------------------------------------------
constexpr int kMaxBlock = 1024;

template <int N, int OUT, int G, int NV, int VEC, int B, typename T>
__global__ __launch_bounds__(kMaxBlock) void k_csf(CsfArgs a) {
    extern __shared__ float lds[];
    using RowT = Row<G, NV, VEC>;
    constexpr int LF = N - 2;  // leaf-fiber level
    constexpr bool SCATTER = OUT > 0;
    const int R = a.R;
    const uint32_t nH = SCATTER ? a.nHot : 0u;
    const uint32_t stride = a.ldsStride;
    if constexpr (SCATTER) {
        for (uint32_t i = threadIdx.x; i < nH * stride; i += blockDim.x) lds[i] = 0.f;
        __syncthreads();
    }
    const int lane = threadIdx.x % G;
    const uint64_t ngroups = (uint64_t)gridDim.x * blockDim.x / G;
    const uint64_t grp = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / G;
    const uint32_t x0 = (uint32_t)(grp * a.nnz / ngroups);
    const uint32_t x1 = (uint32_t)((grp + 1) * a.nnz / ngroups);
    float *__restrict__ out = a.out;

    // out[target] += v, through LDS for ranked rows
    auto scatter = [&](uint32_t tgt, const RowT &v) {
        if (tgt & kHotBit) {
            uint32_t r = tgt & ~kHotBit;
            if (r < nH) {
                v.lds_add(lds + (size_t)r * stride, R, lane);
                return;
            }
            tgt = a.hotRows[r];
        }
        v.atomic_add(out, tgt, R, lane);
    };

    if (x0 < x1) {
        const T *U0 = static_cast<const T *>(a.U[0]);
        const T *U1 = static_cast<const T *>(a.U[1]);
        const T *U2 = static_cast<const T *>(a.U[2]);
        const T *UL = static_cast<const T *>(a.U[N - 1]);
        const uint32_t *__restrict__ ptrL = a.ptr[LF];
        const uint32_t *__restrict__ leaf = (OUT == N - 1) ? a.tgt : a.leaf;
        const uint32_t *__restrict__ tgtF = a.tgt;
        const float *__restrict__ val = a.val;

        uint32_t f = upper_index(ptrL, a.cnt[LF], x0);
        uint32_t fend = ptrL[f + 1];
        uint32_t j = f, jend = 0;
        if constexpr (N == 4) {
            j = upper_index(a.ptr[1], a.cnt[1], f);
            jend = a.ptr[1][j + 1];
        }
        uint32_t slc = a.slc[j];

        RowT aF, aJ, aS, PS, PJ, P, t;
        aF.zero();
        aJ.zero();
        aS.zero();
        PS.zero();
        PJ.zero();
        P.zero();

        auto begin_slice = [&]() {
            if constexpr (OUT >= 1) PS.load(U0, slc, R, lane);
        };
        auto begin_j = [&]() {  // 4-D: product of slice and fbrS rows
            if constexpr (N == 4 && OUT >= 2) {
                t.load(U1, a.idx[1][j], R, lane);
                PJ.had(PS, t);
            }
        };
        auto begin_fiber = [&]() {
            if constexpr (OUT == N - 1) {
                if constexpr (N == 3) {
                    t.load(U1, a.idx[1][f], R, lane);
                    P.had(PS, t);
                } else {
                    t.load(U2, a.idx[2][f], R, lane);
                    P.had(PJ, t);
                }
            }
        };
        auto end_fiber = [&]() {
            if constexpr (N == 3 && OUT == 0) {
                t.load(U1, a.idx[1][f], R, lane);
                aS.fma_had(aF, t);
            } else if constexpr (N == 3 && OUT == 1) {
                aF.mul(PS);
                scatter(tgtF[f], aF);
            } else if constexpr (N == 4 && (OUT == 0 || OUT == 1)) {
                t.load(U2, a.idx[2][f], R, lane);
                aJ.fma_had(aF, t);
            } else if constexpr (N == 4 && OUT == 2) {
                aF.mul(PJ);
                scatter(tgtF[f], aF);
            }
            aF.zero();
        };
        auto end_j = [&]() {
            if constexpr (N == 4 && OUT == 0) {
                t.load(U1, a.idx[1][j], R, lane);
                aS.fma_had(aJ, t);
                aJ.zero();
            } else if constexpr (N == 4 && OUT == 1) {
                aJ.mul(PS);
                scatter(tgtF[j], aJ);
                aJ.zero();
            }
        };
        auto end_slice = [&]() {
            if constexpr (OUT == 0) {
                aS.atomic_add(out, slc, R, lane);
                aS.zero();
            }
        };

        begin_slice();
        begin_j();
        begin_fiber();
        for (uint32_t xb = x0; xb < x1; xb += B) {
            uint32_t li[B];
            float vv[B];
            RowT rows[B];
#pragma unroll
            for (int k = 0; k < B; ++k) {
                uint32_t x = xb + k;
                li[k] = 0;
                vv[k] = 0.f;
                if (x < x1) {
                    li[k] = leaf[x];
                    vv[k] = val[x];
                }
            }
            if constexpr (OUT != N - 1) {
#pragma unroll
                for (int k = 0; k < B; ++k)
                    if (xb + k < x1) rows[k].load(UL, li[k], R, lane);
            }
#pragma unroll
            for (int k = 0; k < B; ++k) {
                uint32_t x = xb + k;
                if (x >= x1) break;
                if (x == fend) {
                    end_fiber();
                    ++f;
                    fend = ptrL[f + 1];
                    if constexpr (N == 4) {
                        if (f == jend) {
                            end_j();
                            ++j;
                            jend = a.ptr[1][j + 1];
                            uint32_t ns = a.slc[j];
                            if (ns != slc) {
                                end_slice();
                                slc = ns;
                                begin_slice();
                            }
                            begin_j();
                        }
                    } else {
                        uint32_t ns = a.slc[f];
                        if (ns != slc) {
                            end_slice();
                            slc = ns;
                            begin_slice();
                        }
                    }
                    begin_fiber();
                }
                if constexpr (OUT == N - 1) {
                    RowT c;
                    c.zero();
                    c.fma(vv[k], P);
                    scatter(li[k], c);
                } else {
                    aF.fma(vv[k], rows[k]);
                }
            }
        }
        end_fiber();
        end_j();
        end_slice();
    }

    if constexpr (SCATTER) {
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < nH * (uint32_t)R; i += blockDim.x) {
            uint32_t r = i / (uint32_t)R, c = i - r * (uint32_t)R;
            float v = lds[(size_t)r * stride + c];
            if (v != 0.f) unsafeAtomicAdd(out + (size_t)a.hotRows[r] * R + c, v);
        }
    }
}

// --------------------------------------------------------------------------
// COO (sorted by modeOrder): keys[0] is the output coordinate
// --------------------------------------------------------------------------
struct CooArgs {
    const uint32_t *key[4];
    const float *val;
    const void *U[4];  // per level
    float *out;
    uint32_t nnz;
    uint32_t perGroup;
    int R;
    int store;         // 1: every output row occurs once (HB-CSF COO part): plain stores
};

template <int N, int G, int NV, int VEC, int B, typename T>
__global__ __launch_bounds__(kBlock) void k_coo(CooArgs a) {
    using RowT = Row<G, NV, VEC>;
    const int lane = threadIdx.x % G;
    const uint64_t grp = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / G;
    const uint64_t x0l = grp * (uint64_t)a.perGroup;
    if (x0l >= a.nnz) return;
    const uint32_t x0 = (uint32_t)x0l;
    const uint32_t x1 = (uint32_t)min((uint64_t)a.nnz, x0l + a.perGroup);
    const int R = a.R;
    uint32_t cur = a.key[0][x0];
    RowT acc;
    acc.zero();
    for (uint32_t xb = x0; xb < x1; xb += B) {
        uint32_t k0[B];
        float vv[B];
        RowT r1[B], r2[B], r3[B];
#pragma unroll
        for (int k = 0; k < B; ++k) {
            uint32_t x = xb + k;
            if (x < x1) {
                k0[k] = a.key[0][x];
                vv[k] = a.val[x];
                r1[k].load(static_cast<const T *>(a.U[1]), a.key[1][x], R, lane);
                r2[k].load(static_cast<const T *>(a.U[2]), a.key[2][x], R, lane);
                if constexpr (N == 4) r3[k].load(static_cast<const T *>(a.U[3]), a.key[3][x], R, lane);
            }
        }
#pragma unroll
        for (int k = 0; k < B; ++k) {
            if (xb + k >= x1) break;
            if (k0[k] != cur) {
                if (a.store) acc.store(a.out, cur, R, lane);
                else acc.atomic_add(a.out, cur, R, lane);
                acc.zero();
                cur = k0[k];
            }
            RowT p;
            p.had(r1[k], r2[k]);
            if constexpr (N == 4) p.mul(r3[k]);
            acc.fma(vv[k], p);
        }
    }
    if (a.store) acc.store(a.out, cur, R, lane);
    else acc.atomic_add(a.out, cur, R, lane);
}

// --------------------------------------------------------------------------
// HB-CSF CSL part: slices whose fibers all hold one nonzero
// --------------------------------------------------------------------------
struct CslArgs {
    const uint32_t *ptr, *idx, *ind1, *ind2;
    const float *val;
    const void *U1, *U2;
    float *out;
    uint32_t nnz, nslc, perGroup;
    int R;
};

template <int G, int NV, int VEC, int B, typename T>
__global__ __launch_bounds__(kBlock) void k_csl(CslArgs a) {
    using RowT = Row<G, NV, VEC>;
    const int lane = threadIdx.x % G;
    const uint64_t grp = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / G;
    const uint64_t x0l = grp * (uint64_t)a.perGroup;
    if (x0l >= a.nnz) return;
    const uint32_t x0 = (uint32_t)x0l;
    const uint32_t x1 = (uint32_t)min((uint64_t)a.nnz, x0l + a.perGroup);
    const int R = a.R;
    uint32_t s = upper_index(a.ptr, a.nslc, x0);
    uint32_t send = a.ptr[s + 1];
    RowT acc;
    acc.zero();
    for (uint32_t xb = x0; xb < x1; xb += B) {
        float vv[B];
        RowT r1[B], r2[B];
#pragma unroll
        for (int k = 0; k < B; ++k) {
            uint32_t x = xb + k;
            if (x < x1) {
                vv[k] = a.val[x];
                r1[k].load(static_cast<const T *>(a.U1), a.ind1[x], R, lane);
                r2[k].load(static_cast<const T *>(a.U2), a.ind2[x], R, lane);
            }
        }
#pragma unroll
        for (int k = 0; k < B; ++k) {
            uint32_t x = xb + k;
            if (x >= x1) break;
            if (x == send) {
                acc.atomic_add(a.out, a.idx[s], R, lane);
                acc.zero();
                ++s;
                send = a.ptr[s + 1];
            }
            RowT p;
            p.had(r1[k], r2[k]);
            acc.fma(vv[k], p);
        }
    }
    acc.atomic_add(a.out, a.idx[s], R, lane);
}

__global__ void k_zero(float *p, uint64_t n) {
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = 0.f;
}

// --------------------------------------------------------------------------
// dispatch over the row geometry
// --------------------------------------------------------------------------
struct Geo {
    int G, NV, VEC;
};

Geo pick_geometry(int R, int dtype) {
    if (R < 1 || R > 256) fail(MMCSF_ERR_UNSUPPORTED, "R must be in 1..256");
    if (R % 4 == 0) {
        int r4 = R / 4, g = 1;
        while (g < r4 && g < 16) g *= 2;
        return Geo{g, (r4 + g - 1) / g, 4};
    }
    if (dtype == MMCSF_BF16) fail(MMCSF_ERR_UNSUPPORTED, "bfloat16 factors need R % 4 == 0");
    return Geo{64, (R + 63) / 64, 1};
}

uint32_t per_group_for(uint64_t nnz, int G) {
    uint64_t target = (uint64_t)device_cu_count() * (2048 / G) * 2;
    uint64_t pg = (nnz + target - 1) / std::max<uint64_t>(target, 1);
    return (uint32_t)std::max<uint64_t>(pg, 16);
}

template <template <int, int, int, int, typename> class K, typename Args, typename T>
void run_geo(const Geo &g, Args &a, uint64_t groups, hipStream_t s) {
    uint64_t threads = groups * (uint64_t)g.G;
    dim3 grid((unsigned)std::max<uint64_t>(1, (threads + kBlock - 1) / kBlock)), block(kBlock);
#define MMCSF_GEO(GG, NN, VV)                                                                  \
    if (g.G == GG && g.NV == NN && g.VEC == VV) {                                              \
        constexpr int Bk = (NN >= 3) ? 2 : (NN == 2 ? 4 : 8);                                  \
        K<GG, NN, VV, Bk, T>::launch(grid, block, s, a);                                       \
        MMCSF_HIP(hipGetLastError());                                                          \
        return;                                                                                \
    }
    MMCSF_GEO(1, 1, 4)
    MMCSF_GEO(2, 1, 4)
    MMCSF_GEO(4, 1, 4)
    MMCSF_GEO(8, 1, 4)
    MMCSF_GEO(16, 1, 4)
    MMCSF_GEO(16, 2, 4)
    MMCSF_GEO(16, 3, 4)
    MMCSF_GEO(16, 4, 4)
    if constexpr (std::is_same<T, float>::value) {
        MMCSF_GEO(64, 1, 1)
        MMCSF_GEO(64, 2, 1)
        MMCSF_GEO(64, 3, 1)
        MMCSF_GEO(64, 4, 1)
    }
#undef MMCSF_GEO
    fail(MMCSF_ERR_UNSUPPORTED, "no kernel for this rank");
}

// kernel selectors with a uniform template signature
template <int N>
struct CooSel {
    template <int G, int NV, int VEC, int B, typename T>
    struct K {
        static void launch(dim3 g, dim3 b, hipStream_t s, const CooArgs &a) {
            k_coo<N, G, NV, VEC, B, T><<<g, b, 0, s>>>(a);
        }
    };
};
template <int G, int NV, int VEC, int B, typename T>
struct CslK {
    static void launch(dim3 g, dim3 b, hipStream_t s, const CslArgs &a) {
        k_csl<G, NV, VEC, B, T><<<g, b, 0, s>>>(a);
    }
};

template <int N, int OUT>
struct CsfLaunch {
    template <int G, int NV, int VEC, int B, typename T>
    struct K {
        static void launch(dim3, dim3, hipStream_t s, const CsfArgs &a) {
            auto fn = k_csf<N, OUT, G, NV, VEC, B, T>;
            const size_t lds = (size_t)a.nHot * a.ldsStride * sizeof(float);
            const int block = OUT > 0 ? kMaxBlock : 256;
            if (lds > 65536) MMCSF_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(fn),
                                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
            int perCU = 0;
            MMCSF_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, fn, block, lds));
            perCU = std::max(perCU, 1);
            dim3 grid((unsigned)(device_cu_count() * perCU));
            fn<<<grid, dim3(block), lds, s>>>(a);
        }
    };
};

template <int N, int OUT>
void launch_csf_t(const Geo &g, CsfArgs &a, int dtype, hipStream_t s) {
    if (dtype == MMCSF_F32)
        run_geo<CsfLaunch<N, OUT>::template K, CsfArgs, float>(g, a, 1, s);
    else
        run_geo<CsfLaunch<N, OUT>::template K, CsfArgs, uint16_t>(g, a, 1, s);
}

int max_lds_per_block() {
    static int bytes = 0;
    if (!bytes) {
        int dev = 0;
        MMCSF_HIP(hipGetDevice(&dev));
        MMCSF_HIP(hipDeviceGetAttribute(&bytes, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
        if (bytes <= 0) bytes = 65536;
    }
    return bytes;
}

void check_align(const void *p, int dtype, const Geo &g) {
    size_t need = g.VEC == 4 ? (dtype == MMCSF_F32 ? 16 : 8) : (dtype == MMCSF_F32 ? 4 : 2);
    if (reinterpret_cast<uintptr_t>(p) % need)
        fail(MMCSF_ERR_ARG, "factor / output pointer is not aligned to " + std::to_string(need) + " bytes");
}

}  // namespace

void preload_scatter_kernels() {
    // one kernel of this translation unit: its code object loads with it
    hipFuncAttributes attr;
    MMCSF_HIP(hipFuncGetAttributes(&attr, reinterpret_cast<const void *>(k_zero)));
}

int device_cu_count() {
    static int cus = 0;
    if (!cus) {
        int dev = 0;
        MMCSF_HIP(hipGetDevice(&dev));
        MMCSF_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        if (cus <= 0) cus = 256;
    }
    return cus;
}

void fill_zero(float *p, uint64_t n, hipStream_t s) {
    if (!n) return;
    uint64_t blocks = std::min<uint64_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(k_zero, dim3((unsigned)blocks), dim3(256), 0, s, p, n);
    MMCSF_HIP(hipGetLastError());
}

void mttkrp_csf(const Csf &c, const uint64_t *dims, int outLevel, const Factors &F, float *out,
                hipStream_t s, hipEvent_t evStart, hipEvent_t evEnd, bool accumulate, bool zeroed) {
    const int N = c.ndims;
    if (!accumulate && !zeroed) fill_zero(out, dims[c.order[outLevel]] * (uint64_t)F.R, s);
    if (evStart) MMCSF_HIP(hipEventRecord(evStart, s));
    struct EndMark {
        hipEvent_t e;
        hipStream_t s;
        ~EndMark() {
            if (e) (void)hipEventRecord(e, s);
        }
    } endMark{evEnd, s};
    if (c.nnz == 0) return;
    Geo g = pick_geometry(F.R, F.dtype);
    check_align(out, MMCSF_F32, g);
    CsfArgs a{};
    for (int l = 0; l < N - 1; ++l) {
        a.ptr[l] = c.ptr[l].p;
        a.idx[l] = c.idx[l].p;
        a.cnt[l] = (uint32_t)c.cnt[l];
    }
    a.cnt[N - 1] = (uint32_t)c.nnz;
    a.slc = c.fbrSlc.p;
    a.leaf = c.leaf.p;
    a.val = c.vals.p;
    for (int l = 0; l < N; ++l) {
        a.U[l] = F.U[c.order[l]];
        if (l != outLevel) {
            MMCSF_REQUIRE(a.U[l] != nullptr, MMCSF_ERR_ARG, "missing factor matrix");
            check_align(a.U[l], F.dtype, g);
        }
    }
    a.out = out;
    a.nnz = (uint32_t)c.nnz;
    a.R = F.R;
    a.perGroup = per_group_for(c.nnz, g.G);
    if (outLevel == 0 && walk_supports(c, F.R)) {
        mttkrp_root(c, F, out, s, accumulate);  // checks the alignment its row loads need
        return;
    }
    if (outLevel > 0) {
        const HotTable &h = c.hot[outLevel];
        MMCSF_REQUIRE(h.tgt.p != nullptr || c.nnz == 0, MMCSF_ERR_ARG,
                      "this format serves its root mode only");
        a.tgt = h.tgt.p;
        a.hotRows = h.rows.p;
        a.ldsStride = (uint32_t)F.R + 1;
        const uint64_t cap = (uint64_t)(max_lds_per_block() - 1024) / (a.ldsStride * sizeof(float));
        a.nHot = (uint32_t)std::min<uint64_t>(h.count, cap);
    }
    if (N == 3) {
        switch (outLevel) {
            case 0: launch_csf_t<3, 0>(g, a, F.dtype, s); break;
            case 1: launch_csf_t<3, 1>(g, a, F.dtype, s); break;
            default: launch_csf_t<3, 2>(g, a, F.dtype, s); break;
        }
    } else {
        switch (outLevel) {
            case 0: launch_csf_t<4, 0>(g, a, F.dtype, s); break;
            case 1: launch_csf_t<4, 1>(g, a, F.dtype, s); break;
            case 2: launch_csf_t<4, 2>(g, a, F.dtype, s); break;
            default: launch_csf_t<4, 3>(g, a, F.dtype, s); break;
        }
    }
}

void mttkrp_mode(const Format &f, int mode, const Factors &F, float *out, hipStream_t s, hipEvent_t evStart,
                 hipEvent_t evEnd, bool zeroed) {
    if (f.mixed) {
        // every partition adds its nonzeros' contributions, with the kernel of
        // the level the mode takes in its tree (mttkrp_gpu.cu:3610-3719)
        if (!zeroed) fill_zero(out, f.dims[mode] * (uint64_t)F.R, s);
        if (evStart) MMCSF_HIP(hipEventRecord(evStart, s));
        for (int p = 0; p < f.ndims; ++p) {
            if (!f.partNnz[p]) continue;
            int lvl = 0;
            for (int l = 0; l < f.ndims; ++l)
                if (f.part[p].order[l] == mode) lvl = l;
            mttkrp_csf(f.part[p], f.dims, lvl, F, out, s, nullptr, nullptr, true);
        }
        if (evEnd) MMCSF_HIP(hipEventRecord(evEnd, s));
        return;
    }
    if (f.hasTree[mode]) {
        mttkrp_csf(f.tree[mode], f.dims, 0, F, out, s, evStart, evEnd, false, zeroed);
        return;
    }
    int lvl = 0;
    for (int l = 0; l < f.ndims; ++l)
        if (f.csf.order[l] == mode) lvl = l;
    mttkrp_csf(f.csf, f.dims, lvl, F, out, s, evStart, evEnd, false, zeroed);
}

void mttkrp_coo(const Format &f, const Factors &F, float *out, hipStream_t s) {
    const int N = f.ndims;
    fill_zero(out, f.dims[f.cooOrder[0]] * (uint64_t)F.R, s);
    if (f.nnz == 0) return;
    Geo g = pick_geometry(F.R, F.dtype);
    check_align(out, MMCSF_F32, g);
    CooArgs a{};
    for (int l = 0; l < N; ++l) {
        a.key[l] = f.cooKeys[l].p;
        a.U[l] = F.U[f.cooOrder[l]];
        if (l) {
            MMCSF_REQUIRE(a.U[l] != nullptr, MMCSF_ERR_ARG, "missing factor matrix");
            check_align(a.U[l], F.dtype, g);
        }
    }
    a.val = f.cooVals.p;
    a.out = out;
    a.nnz = (uint32_t)f.nnz;
    a.R = F.R;
    a.store = 0;
    a.perGroup = per_group_for(f.nnz, g.G);
    uint64_t groups = (a.nnz + a.perGroup - 1) / a.perGroup;
    if (N == 3) {
        if (F.dtype == MMCSF_F32) run_geo<CooSel<3>::template K, CooArgs, float>(g, a, groups, s);
        else run_geo<CooSel<3>::template K, CooArgs, uint16_t>(g, a, groups, s);
    } else {
        if (F.dtype == MMCSF_F32) run_geo<CooSel<4>::template K, CooArgs, float>(g, a, groups, s);
        else run_geo<CooSel<4>::template K, CooArgs, uint16_t>(g, a, groups, s);
    }
}

void mttkrp_hyb(const Hyb &h, const uint64_t *dims, const Factors &F, float *out, hipStream_t s) {
    const int N = h.ndims;
    Geo g = pick_geometry(F.R, F.dtype);
    check_align(out, MMCSF_F32, g);
    for (int l = 1; l < N; ++l) {
        MMCSF_REQUIRE(F.U[h.order[l]] != nullptr, MMCSF_ERR_ARG, "missing factor matrix");
        check_align(F.U[h.order[l]], F.dtype, g);
    }
    // the CSF part's walk loads wider rows than the COO / CSL kernels: refuse
    // its misaligned factors before any part is launched
    if (h.csf.nnz && walk_supports(h.csf, F.R)) check_walk_align(h.csf, F, 0, out);
    if (h.cslWalk.nnz && walk_supports(h.cslWalk, F.R)) check_walk_align(h.cslWalk, F, 0, out);
    fill_zero(out, dims[h.order[0]] * (uint64_t)F.R, s);
    // COO part: slices with one nonzero; each output row is written once
    if (h.cooNnz) {
        CooArgs a{};
        for (int l = 0; l < N; ++l) {
            a.key[l] = h.cooInd[l].p;
            a.U[l] = F.U[h.order[l]];
        }
        a.val = h.cooVal.p;
        a.out = out;
        a.nnz = (uint32_t)h.cooNnz;
        a.R = F.R;
        a.store = 1;
        a.perGroup = per_group_for(h.cooNnz, g.G);
        uint64_t groups = (a.nnz + a.perGroup - 1) / a.perGroup;
        if (N == 3) {
            if (F.dtype == MMCSF_F32) run_geo<CooSel<3>::template K, CooArgs, float>(g, a, groups, s);
            else run_geo<CooSel<3>::template K, CooArgs, uint16_t>(g, a, groups, s);
        } else {
            if (F.dtype == MMCSF_F32) run_geo<CooSel<4>::template K, CooArgs, float>(g, a, groups, s);
            else run_geo<CooSel<4>::template K, CooArgs, uint16_t>(g, a, groups, s);
        }
    }
    // CSL part (3-D): walked as a tree of one-nonzero fibers; the CSL kernel
    // where the walk cannot run (R % 4 != 0, tables above 4 GiB)
    if (h.cslNnz && h.cslWalk.nnz && walk_supports(h.cslWalk, F.R)) {
        mttkrp_root(h.cslWalk, F, out, s);
    } else if (h.cslNnz) {
        CslArgs a{};
        a.ptr = h.cslPtr.p;
        a.idx = h.cslIdx.p;
        a.ind1 = h.cslInd1.p;
        a.ind2 = h.cslInd2.p;
        a.val = h.cslVal.p;
        a.U1 = F.U[h.order[1]];
        a.U2 = F.U[h.order[2]];
        a.out = out;
        a.nnz = (uint32_t)h.cslNnz;
        a.nslc = (uint32_t)h.nCsl;
        a.R = F.R;
        a.perGroup = per_group_for(h.cslNnz, g.G);
        uint64_t groups = (a.nnz + a.perGroup - 1) / a.perGroup;
        if (F.dtype == MMCSF_F32) run_geo<CslK, CslArgs, float>(g, a, groups, s);
        else run_geo<CslK, CslArgs, uint16_t>(g, a, groups, s);
    }
    // CSF part: root-mode walker (accumulates on top of the other parts' rows,
    // which are disjoint from its slices)
    if (h.csf.nnz) {
        const Csf &c = h.csf;
        CsfArgs a{};
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
        a.perGroup = per_group_for(c.nnz, g.G);
        if (walk_supports(c, F.R)) mttkrp_root(c, F, out, s);
        else if (N == 3) launch_csf_t<3, 0>(g, a, F.dtype, s);
        else launch_csf_t<4, 0>(g, a, F.dtype, s);
    }
}

}  // namespace mmcsf
