// MTTKRP device code for gfx950 (wave64).
//
// Work mapping.  A "row group" of G lanes owns one factor row of R columns:
// lane l of the group holds columns (v*G + l)*VEC + [0, VEC) for v < NV, so a
// row load is one coalesced G*VEC*sizeof(T)-byte access per chunk (128 B for
// R=32 float32 with G=8, VEC=4).  A wave64 therefore carries 64/G independent
// groups.  Every group walks a contiguous range of nonzeros of the CSF tree
// (equal nonzero counts per group, so heavy slices and long fibers are split
// across groups automatically) and keeps its partial sums in registers:
//
//   output level      per nonzero              per leaf fiber / fbrS / slice
//   root (0)          aF += v * U_leaf[k]     aS += aF (.) U_fbr[j]; flush aS at a slice end
//   fiber levels      aF += v * U_leaf[k]     flush aF (.) prod(ancestor rows) at the fiber end
//   leaf (N-1)        out[k] += v * P          P = prod(ancestor rows), once per fiber
//
// Flushes are float atomics (no return) and happen only where a group's range
// ends or a tree node ends, i.e. atomics appear only across split fibers and
// slices.  Nonzero indices and values are fetched B at a time per group and
// the B factor rows are requested before any of them is consumed, so each
// group keeps B gathers in flight.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace mmcsf {
namespace dev {

// ---- factor row access ----------------------------------------------------
__device__ __forceinline__ float bf16_lo(uint32_t u) { return __uint_as_float(u << 16); }
__device__ __forceinline__ float bf16_hi(uint32_t u) { return __uint_as_float(u & 0xffff0000u); }

template <int VEC>
__device__ __forceinline__ void ld(const float *__restrict__ p, float *d) {
    if constexpr (VEC == 4) {
        float4 q = *reinterpret_cast<const float4 *>(p);
        d[0] = q.x; d[1] = q.y; d[2] = q.z; d[3] = q.w;
    } else {
        d[0] = *p;
    }
}

template <int VEC>
__device__ __forceinline__ void ld(const uint16_t *__restrict__ p, float *d) {
    if constexpr (VEC == 4) {
        uint2 q = *reinterpret_cast<const uint2 *>(p);
        d[0] = bf16_lo(q.x); d[1] = bf16_hi(q.x); d[2] = bf16_lo(q.y); d[3] = bf16_hi(q.y);
    } else {
        d[0] = __uint_as_float((uint32_t)(*p) << 16);
    }
}

template <int G, int NV, int VEC>
struct Row {
    float v[NV][VEC];
    __device__ __forceinline__ void zero() {
#pragma unroll
        for (int a = 0; a < NV; ++a)
#pragma unroll
            for (int b = 0; b < VEC; ++b) v[a][b] = 0.f;
    }
    template <typename T>
    __device__ __forceinline__ void load(const T *__restrict__ U, uint32_t row, int R, int lane) {
        const T *base = U + (size_t)row * (size_t)R;
#pragma unroll
        for (int a = 0; a < NV; ++a) {
            int c = (a * G + lane) * VEC;
            if (c < R) {
                ld<VEC>(base + c, v[a]);
            } else {
#pragma unroll
                for (int b = 0; b < VEC; ++b) v[a][b] = 0.f;
            }
        }
    }
    // this += s * o
    __device__ __forceinline__ void fma(float s, const Row &o) {
#pragma unroll
        for (int a = 0; a < NV; ++a)
#pragma unroll
            for (int b = 0; b < VEC; ++b) v[a][b] = __builtin_fmaf(s, o.v[a][b], v[a][b]);
    }
    // this += x (.) y
    __device__ __forceinline__ void fma_had(const Row &x, const Row &y) {
#pragma unroll
        for (int a = 0; a < NV; ++a)
#pragma unroll
            for (int b = 0; b < VEC; ++b) v[a][b] = __builtin_fmaf(x.v[a][b], y.v[a][b], v[a][b]);
    }
    // this = x (.) y
    __device__ __forceinline__ void had(const Row &x, const Row &y) {
#pragma unroll
        for (int a = 0; a < NV; ++a)
#pragma unroll
            for (int b = 0; b < VEC; ++b) v[a][b] = x.v[a][b] * y.v[a][b];
    }
    __device__ __forceinline__ void mul(const Row &y) {
#pragma unroll
        for (int a = 0; a < NV; ++a)
#pragma unroll
            for (int b = 0; b < VEC; ++b) v[a][b] *= y.v[a][b];
    }
    __device__ __forceinline__ bool nonzero() const {
        bool nz = false;
#pragma unroll
        for (int a = 0; a < NV; ++a)
#pragma unroll
            for (int b = 0; b < VEC; ++b) nz |= (v[a][b] != 0.f);
        return nz;
    }
    // out[row] += this, one no-return float atomic per column
    __device__ __forceinline__ void atomic_add(float *__restrict__ out, uint32_t row, int R,
                                               int lane) const {
        float *base = out + (size_t)row * (size_t)R;
#pragma unroll
        for (int a = 0; a < NV; ++a) {
            int c = (a * G + lane) * VEC;
            if (c < R) {
#pragma unroll
                for (int b = 0; b < VEC; ++b) unsafeAtomicAdd(base + c + b, v[a][b]);
            }
        }
    }
    // lds[c] += this[c] for an LDS-resident row (ds_add_f32)
    __device__ __forceinline__ void lds_add(float *base, int R, int lane) const {
#pragma unroll
        for (int a = 0; a < NV; ++a) {
            int c = (a * G + lane) * VEC;
            if (c < R) {
#pragma unroll
                for (int b = 0; b < VEC; ++b) atomicAdd(base + c + b, v[a][b]);
            }
        }
    }
    // out[row] = this (rows owned by one group)
    __device__ __forceinline__ void store(float *__restrict__ out, uint32_t row, int R, int lane) const {
        float *base = out + (size_t)row * (size_t)R;
#pragma unroll
        for (int a = 0; a < NV; ++a) {
            int c = (a * G + lane) * VEC;
            if (c < R) {
                if constexpr (VEC == 4) {
                    *reinterpret_cast<float4 *>(base + c) = make_float4(v[a][0], v[a][1], v[a][2], v[a][3]);
                } else {
                    base[c] = v[a][0];
                }
            }
        }
    }
};

// largest i in [0, n) with a[i] <= key (a ascending, a[0] <= key)
__device__ __forceinline__ uint32_t upper_index(const uint32_t *__restrict__ a, uint32_t n, uint32_t key) {
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        uint32_t mid = lo + ((hi - lo) >> 1);
        if (a[mid] <= key) lo = mid;
        else hi = mid;
    }
    return lo;
}

}  // namespace dev

// Kernel argument block for one CSF tree.
struct CsfArgs {
    const uint32_t *ptr[3];   // ptr[l], l = 0..N-2
    const uint32_t *idx[3];   // idx[l]
    const uint32_t *slc;      // fbrLikeSlcInds (slice coordinate of each level-1 entry)
    const uint32_t *leaf;     // leaf coordinate per nonzero
    const float *val;
    const void *U[4];         // factor of each LEVEL (U[order[l]])
    float *out;
    uint32_t nnz;
    uint32_t cnt[4];
    uint32_t perGroup;        // nonzeros per row group (non-persistent launches)
    int R;
    // scatter roles: output-level index with ranked rows tagged kHotBit|rank
    const uint32_t *tgt;
    const uint32_t *hotRows;  // rank -> row
    uint32_t nHot;            // ranks accumulated in LDS by this launch
    uint32_t ldsStride;       // floats per LDS row
};

}  // namespace mmcsf
