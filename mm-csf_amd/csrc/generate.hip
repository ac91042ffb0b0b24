// Synthetic sparse tensors generated on the GPU.
//
// Draw d (d = 0, 1, 2, ...) takes coordinate m from an independent counter
// stream (seed, m, d) and its value from stream (seed, 15, d); coordinates are
// uniform, or Zipf(alpha) by inverse CDF with rank r -> index r-1 (hot rows at
// low indices).  Repeated coordinates are rejected: the tensor holds the first
// nnz distinct coordinates in draw order, stored in draw order.  Every number
// is a pure function of (seed, stream, d), so the result does not depend on
// how many extra draws were made.
#include <cstring>
#include <rocprim/rocprim.hpp>

#include <cmath>
#include <vector>

#include "mmcsf_internal.hpp"

namespace mmcsf {

namespace {

__host__ __device__ inline uint64_t mix64(uint64_t z) {
    // SplitMix64 finaliser (Stafford's variant 13)
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

__host__ __device__ inline uint64_t draw_bits(uint64_t seed, uint32_t stream, uint64_t d) {
    uint64_t key = mix64(seed ^ (0x632be59bd9b4e019ull * (uint64_t)(stream + 1)));
    return mix64(key + (d + 1) * 0x9e3779b97f4a7c15ull);
}

struct GenSpec {
    int ndims;
    uint64_t dims[kMaxDims];
    int law[kMaxDims];
    const double *cdf[kMaxDims];
    int bits[kMaxDims];
    uint64_t seed;
};

__device__ inline uint32_t coord(const GenSpec &g, int m, uint64_t d) {
    uint64_t u = draw_bits(g.seed, (uint32_t)m, d);
    if (g.law[m] == 0) return (uint32_t)__umul64hi(u, g.dims[m]);
    double p = (double)(u >> 11) * 0x1.0p-53;
    const double *c = g.cdf[m];
    uint64_t lo = 0, hi = g.dims[m] - 1;  // first r with c[r] > p
    while (lo < hi) {
        uint64_t mid = (lo + hi) >> 1;
        if (c[mid] > p) hi = mid;
        else lo = mid + 1;
    }
    return (uint32_t)lo;
}

__global__ void k_draw_keys(GenSpec g, uint64_t *__restrict__ key, uint32_t *__restrict__ id, uint64_t n) {
    uint64_t d = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (d >= n) return;
    uint64_t k = 0;
    int shift = 0;
    for (int m = g.ndims - 1; m >= 0; --m) {
        k |= (uint64_t)coord(g, m, d) << shift;
        shift += g.bits[m];
    }
    key[d] = k;
    id[d] = (uint32_t)d;
}

// keep[draw] = 1 for the first draw of every distinct key (sorted stably)
__global__ void k_first_occ(const uint64_t *__restrict__ key, const uint32_t *__restrict__ id,
                            uint32_t *__restrict__ keep, uint64_t n) {
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    keep[id[i]] = (i == 0 || key[i] != key[i - 1]) ? 1u : 0u;
}

__global__ void k_emit(GenSpec g, const uint32_t *__restrict__ keep, const uint32_t *__restrict__ rank,
                       uint64_t nnz, uint32_t *i0, uint32_t *i1, uint32_t *i2, uint32_t *i3,
                       float *__restrict__ val, uint64_t n) {
    uint64_t d = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (d >= n || !keep[d]) return;
    uint64_t r = rank[d];
    if (r >= nnz) return;
    uint32_t *dst[4] = {i0, i1, i2, i3};
    for (int m = 0; m < g.ndims; ++m) dst[m][r] = coord(g, m, d);
    // U(0,1) float32 with 24 random bits
    val[r] = (float)(draw_bits(g.seed, 15u, d) >> 40) * 0x1.0p-24f;
}

int bitwidth(uint64_t v) {
    int b = 1;
    while (b < 64 && (v >> b)) ++b;
    return b;
}

}  // namespace

void generate_tensor(int N, const uint64_t *dims, uint64_t nnz, const int *law, const double *alpha, uint64_t seed,
                     Tensor &t, hipStream_t s) {
    GenSpec g{};
    g.ndims = N;
    g.seed = seed;
    int totBits = 0;
    std::vector<DevArray<double>> cdf(N);
    for (int m = 0; m < N; ++m) {
        g.dims[m] = dims[m];
        g.law[m] = law[m];
        g.bits[m] = bitwidth(dims[m] - 1);
        totBits += g.bits[m];
        if (law[m] == 1) {
            // Zipf(alpha): P(rank r) proportional to r^-alpha, r = 1..dims[m]
            MMCSF_REQUIRE(alpha[m] > 0, MMCSF_ERR_ARG, "Zipf exponent must be positive");
            std::vector<double> h(dims[m]);
            double acc = 0;
            for (uint64_t r = 0; r < dims[m]; ++r) {
                acc += std::pow((double)(r + 1), -alpha[m]);
                h[r] = acc;
            }
            for (uint64_t r = 0; r < dims[m]; ++r) h[r] /= acc;
            h[dims[m] - 1] = 2.0;  // every p < 1 maps inside the table
            cdf[m].alloc(dims[m]);
            MMCSF_HIP(hipMemcpy(cdf[m].p, h.data(), dims[m] * 8, hipMemcpyHostToDevice));
            g.cdf[m] = cdf[m].p;
        } else {
            MMCSF_REQUIRE(law[m] == 0, MMCSF_ERR_ARG, "law must be 0 (uniform) or 1 (Zipf)");
        }
    }
    MMCSF_REQUIRE(totBits <= 64, MMCSF_ERR_UNSUPPORTED, "coordinates wider than 64 bits");
    t.ndims = N;
    t.nnz = nnz;
    for (int m = 0; m < N; ++m) {
        t.dims[m] = dims[m];
        t.inds[m].alloc(nnz);
    }
    t.vals.alloc(nnz);
    if (nnz == 0) return;

    uint64_t draws = nnz + nnz / 8 + 1024;
    for (int attempt = 0; attempt < 12; ++attempt) {
        DevArray<uint64_t> k1(draws), k2(draws);
        DevArray<uint32_t> id1(draws), id2(draws);
        const unsigned blocks = (unsigned)((draws + 255) / 256);
        hipLaunchKernelGGL(k_draw_keys, dim3(blocks), dim3(256), 0, s, g, k1.p, id1.p, draws);
        size_t bytes = 0;
        MMCSF_HIP(rocprim::radix_sort_pairs(nullptr, bytes, k1.p, k2.p, id1.p, id2.p, (size_t)draws, 0u,
                                            (unsigned)totBits, s));
        DevArray<uint8_t> tmp(bytes + 16);
        MMCSF_HIP(rocprim::radix_sort_pairs(tmp.p, bytes, k1.p, k2.p, id1.p, id2.p, (size_t)draws, 0u,
                                            (unsigned)totBits, s));
        k1.release();
        id1.release();
        DevArray<uint32_t> keep(draws), rank(draws);
        hipLaunchKernelGGL(k_first_occ, dim3(blocks), dim3(256), 0, s, k2.p, id2.p, keep.p, draws);
        k2.release();
        id2.release();
        bytes = 0;
        MMCSF_HIP(rocprim::exclusive_scan(nullptr, bytes, keep.p, rank.p, 0u, (size_t)draws,
                                          rocprim::plus<uint32_t>(), s));
        tmp.alloc(bytes + 16);
        MMCSF_HIP(rocprim::exclusive_scan(tmp.p, bytes, keep.p, rank.p, 0u, (size_t)draws,
                                          rocprim::plus<uint32_t>(), s));
        uint32_t lastKeep = 0, lastRank = 0;
        MMCSF_HIP(hipMemcpyAsync(&lastKeep, keep.p + draws - 1, 4, hipMemcpyDeviceToHost, s));
        MMCSF_HIP(hipMemcpyAsync(&lastRank, rank.p + draws - 1, 4, hipMemcpyDeviceToHost, s));
        MMCSF_HIP(hipStreamSynchronize(s));
        uint64_t distinct = (uint64_t)lastKeep + lastRank;
        if (distinct >= nnz) {
            hipLaunchKernelGGL(k_emit, dim3(blocks), dim3(256), 0, s, g, keep.p, rank.p, nnz, t.inds[0].p,
                               t.inds[1].p, t.inds[2].p, N == 4 ? t.inds[3].p : t.inds[2].p, t.vals.p, draws);
            MMCSF_HIP(hipGetLastError());
            MMCSF_HIP(hipStreamSynchronize(s));
            return;
        }
        // not enough distinct coordinates yet: draw more (same prefix)
        double ratio = (double)draws / (double)std::max<uint64_t>(distinct, 1);
        uint64_t want = (uint64_t)((double)nnz * ratio * 1.08) + 4096;
        draws = std::max<uint64_t>(want, draws + draws / 2);
        MMCSF_REQUIRE(draws < (1ull << 32), MMCSF_ERR_UNSUPPORTED, "too many draws needed for distinct coordinates");
    }
    fail(MMCSF_ERR_UNSUPPORTED, "could not draw enough distinct coordinates");
}

}  // namespace mmcsf
