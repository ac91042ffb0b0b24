// Root-mode MTTKRP over 64-nonzero chunks (gfx950, wave64).
//
// A CSF tree whose root level is the output mode is cut into chunks of 64
// consecutive nonzeros; every wave owns a contiguous run of chunks.  Per chunk:
//   * lane p loads leaf[x], val[x] of nonzero x = chunk*64 + p (coalesced);
//   * the leaf fiber of every nonzero is the running count of fiber starts,
//     from a bitmask (bit x = nonzero x opens a leaf fiber) with a wave
//     ballot and a prefix popcount; 4-D trees count level-1 starts the same
//     way over the fibers;
//   * groups of G lanes gather whole factor rows (16 bytes per lane-load):
//     group g takes nonzeros g, g + 64/G, g + 2*64/G, ... so one load
//     instruction reads 64/G consecutive nonzeros' rows, and the rows of
//     the fiber level, shared by consecutive nonzeros, coalesce;
//   * every load of the chunk is issued before the first use;
//   * contributions v * U_leaf[leaf] (.) U_1[idx_1] (. U_2[idx_2]) are summed
//     per lane while the slice stays the same; at a slice boundary (a wave
//     ballot over the per-nonzero slice coordinates) the partial sums are
//     reduced across the groups with xor shuffles and one group adds the row
//     to the output with float atomics.  The last slice of a chunk stays
//     open into the next chunk.
// Bound: HBM for the structure (8 bytes per nonzero, 8 per fiber, 4 per
// level-1 entry) and the L2/MALL gather rate for factor rows.
#include <algorithm>
#include <type_traits>

#include "root_walk.hpp"

namespace mmcsf {

struct ChunkArgs {
    const uint32_t *ptr[3];
    const uint32_t *idx[3];
    const uint32_t *slc;   // slice coordinate of each level-1 entry
    const uint32_t *leaf;
    const float *val;
    const uint32_t *fbit;  // leaf-fiber starts over nonzeros
    const uint32_t *sbit;  // 4-D: level-1 starts over leaf fibers
    const void *U[4];      // factor of each level (U[0] unused)
    float *out;
    uint32_t nnz;
    uint32_t cnt[4];
    int R;
};

namespace {
using walk::upper_idx;

// VEC consecutive elements of one row per lane-load, as floats
template <typename T, int VEC>
struct Raw;
template <>
struct Raw<float, 4> {
    float4 q;
    __device__ __forceinline__ void load(const float *U, uint32_t row, int R, int col) {
        q = *reinterpret_cast<const float4 *>(U + (size_t)row * R + col);
    }
    __device__ __forceinline__ void get(float *o) const { o[0] = q.x; o[1] = q.y; o[2] = q.z; o[3] = q.w; }
};
template <>
struct Raw<uint16_t, 8> {
    uint4 q;
    __device__ __forceinline__ void load(const uint16_t *U, uint32_t row, int R, int col) {
        q = *reinterpret_cast<const uint4 *>(U + (size_t)row * R + col);
    }
    __device__ __forceinline__ void get(float *o) const {
        o[0] = __uint_as_float(q.x << 16); o[1] = __uint_as_float(q.x & 0xffff0000u);
        o[2] = __uint_as_float(q.y << 16); o[3] = __uint_as_float(q.y & 0xffff0000u);
        o[4] = __uint_as_float(q.z << 16); o[5] = __uint_as_float(q.z & 0xffff0000u);
        o[6] = __uint_as_float(q.w << 16); o[7] = __uint_as_float(q.w & 0xffff0000u);
    }
};
template <>
struct Raw<uint16_t, 4> {
    uint2 q;
    __device__ __forceinline__ void load(const uint16_t *U, uint32_t row, int R, int col) {
        q = *reinterpret_cast<const uint2 *>(U + (size_t)row * R + col);
    }
    __device__ __forceinline__ void get(float *o) const {
        o[0] = __uint_as_float(q.x << 16); o[1] = __uint_as_float(q.x & 0xffff0000u);
        o[2] = __uint_as_float(q.y << 16); o[3] = __uint_as_float(q.y & 0xffff0000u);
    }
};

__device__ __forceinline__ uint64_t low_bits(uint32_t d) { return d >= 64 ? ~0ull : ((1ull << d) - 1ull); }

// bits [base, base + 64) of a word array
__device__ __forceinline__ uint64_t window64(const uint32_t *__restrict__ bits, uint32_t base) {
    const uint32_t w = base >> 5, sh = base & 31;
    uint64_t lo = (uint64_t)bits[w] | ((uint64_t)bits[w + 1] << 32);
    uint64_t win = lo >> sh;
    if (sh) win |= (uint64_t)bits[w + 2] << (64 - sh);
    return win;
}

template <int N, int G, int NB, typename T, int VEC>
__global__ __launch_bounds__(256) void k_chunk(ChunkArgs a) {
    static_assert(G == 1 || G == 2 || G == 4 || G == 8, "row groups of 1..8 lanes");
    constexpr int GPR = 64 / G;  // rows per load instruction
    constexpr int STEPS = G;     // load instructions per chunk and column block
    constexpr int BW = G * VEC;  // columns per block
    constexpr int LF = N - 2;
    const int lane = threadIdx.x & 63;
    const int g = lane / G, k = lane % G;
    const uint64_t wavesPerBlock = blockDim.x >> 6;
    const uint64_t waves = (uint64_t)gridDim.x * wavesPerBlock;
    const uint64_t wv = blockIdx.x * wavesPerBlock + (threadIdx.x >> 6);
    const uint64_t chunks = ((uint64_t)a.nnz + 63) / 64;
    const uint64_t cb = wv * chunks / waves, ce = (wv + 1) * chunks / waves;
    if (cb >= ce) return;  // the whole wave leaves together
    const int R = a.R;
    const T *U1 = static_cast<const T *>(a.U[1]);
    const T *U2 = static_cast<const T *>(a.U[2]);
    const T *UL = static_cast<const T *>(a.U[N - 1]);
    const uint32_t *__restrict__ ptrL = a.ptr[LF];

    // leaf fiber (and 4-D level-1 entry) of the nonzero before the run; the
    // unsigned wrap at x0 = 0 stands for -1
    const uint32_t x0 = (uint32_t)(cb * 64);
    uint32_t fcur = x0 ? upper_idx(ptrL, a.cnt[LF], x0 - 1) : 0xffffffffu;
    uint32_t jcur = 0xffffffffu;
    if constexpr (N == 4) jcur = x0 ? upper_idx(a.ptr[1], a.cnt[1], fcur) : 0xffffffffu;
    uint32_t scur;
    {
        uint32_t f0 = upper_idx(ptrL, a.cnt[LF], x0);
        if constexpr (N == 4) f0 = upper_idx(a.ptr[1], a.cnt[1], f0);
        scur = a.slc[f0];
    }

    float acc[NB][VEC];
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
        for (int e = 0; e < VEC; ++e) acc[b][e] = 0.f;

    // add the open partial row to slice s and clear it (all lanes take part)
    auto flush = [&](int b, uint32_t s) {
#pragma unroll
        for (int off = G; off < 64; off <<= 1)
#pragma unroll
            for (int e = 0; e < VEC; ++e) acc[b][e] += __shfl_xor(acc[b][e], off, 64);
        const int col = b * BW + k * VEC;
        if (g == 0 && s != 0xffffffffu && col < R) {
            float *o = a.out + (size_t)s * R + col;
#pragma unroll
            for (int e = 0; e < VEC; ++e) unsafeAtomicAdd(o + e, acc[b][e]);
        }
#pragma unroll
        for (int e = 0; e < VEC; ++e) acc[b][e] = 0.f;
    };

    for (uint64_t c = cb; c < ce; ++c) {
        const uint32_t x = (uint32_t)(c * 64) + lane;
        const bool valid = x < a.nnz;
        uint32_t lw = 0, word = 0;
        float v = 0.f;
        if (valid) {
            lw = a.leaf[x];
            v = a.val[x];
            word = a.fbit[x >> 5];
        }
        const bool start = valid && ((word >> (x & 31)) & 1u);
        const uint64_t fm = __ballot(start);
        const uint64_t le = low_bits((uint32_t)lane + 1);
        const uint32_t f = fcur + (uint32_t)__popcll(fm & le);
        const uint32_t fnext = fcur + (uint32_t)__popcll(fm);
        uint32_t j = f, jnext = fnext;
        if constexpr (N == 4) {
            const uint64_t win = window64(a.sbit, fcur + 1);
            j = jcur + (uint32_t)__popcll(win & low_bits(f - fcur));
            jnext = jcur + (uint32_t)__popcll(win & low_bits(fnext - fcur));
        }
        uint32_t i1 = 0, i2 = 0, s = 0xffffffffu;
        if (valid) {
            i1 = a.idx[1][j];
            s = a.slc[j];
            if constexpr (N == 4) i2 = a.idx[2][f];
        }
        uint32_t sprev = __shfl_up(s, 1, 64);
        if (lane == 0) sprev = scur;
        const uint64_t bm = __ballot(s != sprev);
        const uint32_t s0 = scur;

        // the entries of this lane's group, one per step
        uint32_t li[STEPS], l1[STEPS], l2[STEPS];
        float vv[STEPS];
#pragma unroll
        for (int t = 0; t < STEPS; ++t) {
            const int p = t * GPR + g;
            li[t] = __shfl(lw, p, 64);
            l1[t] = __shfl(i1, p, 64);
            if constexpr (N == 4) l2[t] = __shfl(i2, p, 64);
            vv[t] = __shfl(v, p, 64);
        }
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            const int col = b * BW + k * VEC;
            const bool cok = col < R;
            Raw<T, VEC> rw[STEPS], w1[STEPS], w2[STEPS];
#pragma unroll
            for (int t = 0; t < STEPS; ++t) {
                if (cok) {
                    rw[t].load(UL, li[t], R, col);
                    w1[t].load(U1, l1[t], R, col);
                    if constexpr (N == 4) w2[t].load(U2, l2[t], R, col);
                }
            }
            float cf[STEPS][VEC];
#pragma unroll
            for (int t = 0; t < STEPS; ++t) {
                float r[VEC], q[VEC];
                if (cok) {
                    rw[t].get(r);
                    w1[t].get(q);
                } else {
#pragma unroll
                    for (int e = 0; e < VEC; ++e) r[e] = q[e] = 0.f;
                }
#pragma unroll
                for (int e = 0; e < VEC; ++e) cf[t][e] = vv[t] * r[e] * q[e];
                if constexpr (N == 4) {
                    if (cok) {
                        w2[t].get(q);
#pragma unroll
                        for (int e = 0; e < VEC; ++e) cf[t][e] *= q[e];
                    }
                }
            }
            if (bm == 0) {
#pragma unroll
                for (int t = 0; t < STEPS; ++t)
#pragma unroll
                    for (int e = 0; e < VEC; ++e) acc[b][e] += cf[t][e];
            } else {
                // slice segments of the chunk in order; wave-uniform loop
                uint64_t rest = bm;
                uint32_t lo = 0, sc = s0;
                while (rest) {
                    const uint32_t hi = (uint32_t)__builtin_ctzll(rest);
                    rest &= rest - 1;
#pragma unroll
                    for (int t = 0; t < STEPS; ++t) {
                        const uint32_t p = (uint32_t)(t * GPR + g);
                        if (p >= lo && p < hi)
#pragma unroll
                            for (int e = 0; e < VEC; ++e) acc[b][e] += cf[t][e];
                    }
                    flush(b, sc);
                    sc = __shfl(s, (int)hi, 64);
                    lo = hi;
                }
#pragma unroll
                for (int t = 0; t < STEPS; ++t) {
                    const uint32_t p = (uint32_t)(t * GPR + g);
                    if (p >= lo)
#pragma unroll
                        for (int e = 0; e < VEC; ++e) acc[b][e] += cf[t][e];
                }
            }
        }
        if (bm) scur = __shfl(s, 63 - (int)__builtin_clzll(bm), 64);
        fcur = fnext;
        jcur = jnext;
    }
#pragma unroll
    for (int b = 0; b < NB; ++b) flush(b, scur);
}

struct Cfg {
    int G, NB, VEC;
};

Cfg pick(int R, int dtype) {
    const int vec = (dtype == MMCSF_BF16 && R % 8 == 0) ? 8 : 4;
    if (R <= vec) return {1, 1, vec};
    if (R <= 2 * vec) return {2, 1, vec};
    if (R <= 4 * vec) return {4, 1, vec};
    const int nb = (R + 8 * vec - 1) / (8 * vec);
    return {8, nb <= 1 ? 1 : nb <= 2 ? 2 : nb <= 4 ? 4 : 8, vec};
}

template <int N, int G, int NB, typename T, int VEC>
void go(const ChunkArgs &a, hipStream_t s) {
    auto fn = k_chunk<N, G, NB, T, VEC>;
    int perCU = 0;
    MMCSF_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, fn, 256, 0));
    perCU = std::max(perCU, 1);
    const uint64_t chunks = ((uint64_t)a.nnz + 63) / 64;
    uint64_t blocks = (uint64_t)device_cu_count() * perCU;
    blocks = std::min(blocks, std::max<uint64_t>(1, (chunks + 3) / 4));
    fn<<<dim3((unsigned)blocks), dim3(256), 0, s>>>(a);
    MMCSF_HIP(hipGetLastError());
}

template <int N, typename T, int VEC>
void dispatch(const Cfg &c, const ChunkArgs &a, hipStream_t s) {
#define MMCSF_CHUNK(GG, BB)                  \
    if (c.G == GG && c.NB == BB) {           \
        go<N, GG, BB, T, VEC>(a, s);         \
        return;                              \
    }
    MMCSF_CHUNK(1, 1)
    MMCSF_CHUNK(2, 1)
    MMCSF_CHUNK(4, 1)
    MMCSF_CHUNK(8, 1)
    MMCSF_CHUNK(8, 2)
    MMCSF_CHUNK(8, 4)
    MMCSF_CHUNK(8, 8)
#undef MMCSF_CHUNK
    fail(MMCSF_ERR_UNSUPPORTED, "no chunked root kernel for this rank");
}

}  // namespace

bool chunk_supports(const Csf &c, int R) {
    return R >= 1 && R <= 256 && R % 4 == 0 && c.fbit.p != nullptr && (c.ndims == 3 || c.sbit.p != nullptr);
}

void mttkrp_chunk(const Csf &c, const Factors &F, float *out, hipStream_t s) {
    if (c.nnz == 0) return;
    const int N = c.ndims;
    ChunkArgs a{};
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
    for (int l = 0; l < N; ++l) a.U[l] = F.U[c.order[l]];
    a.out = out;
    a.nnz = (uint32_t)c.nnz;
    a.R = F.R;
    const Cfg cfg = pick(F.R, F.dtype);
    if (N == 3) {
        if (F.dtype == MMCSF_F32) dispatch<3, float, 4>(cfg, a, s);
        else if (cfg.VEC == 8) dispatch<3, uint16_t, 8>(cfg, a, s);
        else dispatch<3, uint16_t, 4>(cfg, a, s);
    } else {
        if (F.dtype == MMCSF_F32) dispatch<4, float, 4>(cfg, a, s);
        else if (cfg.VEC == 8) dispatch<4, uint16_t, 8>(cfg, a, s);
        else dispatch<4, uint16_t, 4>(cfg, a, s);
    }
}

}  // namespace mmcsf
