// Per-lane row helpers shared by the lane-walker kernels (root-mode MTTKRP
// and the CP-completion loss / gradient traversal).
#pragma once
#include <type_traits>

#include "mmcsf_internal.hpp"

namespace mmcsf {
namespace walk {

__device__ __forceinline__ uint32_t upper_idx(const uint32_t *__restrict__ a, uint32_t n, uint32_t key) {
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        uint32_t mid = lo + ((hi - lo) >> 1);
        if (a[mid] <= key) lo = mid;
        else hi = mid;
    }
    return lo;
}

template <int CPL, typename T>
struct LaneRow {
    float v[CPL];
    __device__ __forceinline__ void zero() {
#pragma unroll
        for (int c = 0; c < CPL; ++c) v[c] = 0.f;
    }
    // columns [c0, c0 + CPL) of row `row`; R % 4 == 0 and c0 % 4 == 0
    __device__ __forceinline__ void load(const T *__restrict__ U, uint32_t row, int R, int c0) {
        const T *p = U + (size_t)row * R + c0;
        if constexpr (std::is_same<T, float>::value) {
#pragma unroll
            for (int c = 0; c < CPL; c += 4) {
                if (c0 + c < R) {
                    float4 q = *reinterpret_cast<const float4 *>(p + c);
                    v[c] = q.x; v[c + 1] = q.y; v[c + 2] = q.z; v[c + 3] = q.w;
                } else {
                    v[c] = v[c + 1] = v[c + 2] = v[c + 3] = 0.f;
                }
            }
        } else {
            const bool wide = (R & 7) == 0;  // 16-byte aligned rows
#pragma unroll
            for (int c = 0; c < CPL; c += 8) {
                if (c0 + c < R) {
                    if (wide) {
                        uint4 q = *reinterpret_cast<const uint4 *>(p + c);
                        v[c] = __uint_as_float(q.x << 16); v[c + 1] = __uint_as_float(q.x & 0xffff0000u);
                        v[c + 2] = __uint_as_float(q.y << 16); v[c + 3] = __uint_as_float(q.y & 0xffff0000u);
                        v[c + 4] = __uint_as_float(q.z << 16); v[c + 5] = __uint_as_float(q.z & 0xffff0000u);
                        v[c + 6] = __uint_as_float(q.w << 16); v[c + 7] = __uint_as_float(q.w & 0xffff0000u);
                    } else {  // 8-byte aligned rows: two 4-column loads
                        uint2 q = *reinterpret_cast<const uint2 *>(p + c);
                        v[c] = __uint_as_float(q.x << 16); v[c + 1] = __uint_as_float(q.x & 0xffff0000u);
                        v[c + 2] = __uint_as_float(q.y << 16); v[c + 3] = __uint_as_float(q.y & 0xffff0000u);
                        if (c0 + c + 4 < R) {
                            uint2 q2 = *reinterpret_cast<const uint2 *>(p + c + 4);
                            v[c + 4] = __uint_as_float(q2.x << 16); v[c + 5] = __uint_as_float(q2.x & 0xffff0000u);
                            v[c + 6] = __uint_as_float(q2.y << 16); v[c + 7] = __uint_as_float(q2.y & 0xffff0000u);
                        } else {
                            v[c + 4] = v[c + 5] = v[c + 6] = v[c + 7] = 0.f;
                        }
                    }
                } else {
#pragma unroll
                    for (int b = 0; b < 8; ++b) v[c + b] = 0.f;
                }
            }
        }
    }
};

template <int CPL>
__device__ __forceinline__ void atomic_row(float *__restrict__ out, uint32_t row, int R, int c0, const float *v) {
    float *p = out + (size_t)row * R + c0;
#pragma unroll
    for (int c = 0; c < CPL; ++c)
        if (c0 + c < R) unsafeAtomicAdd(p + c, v[c]);
}

}  // namespace walk
}  // namespace mmcsf
