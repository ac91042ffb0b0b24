// GPU format builder: sort / scan / segment passes that produce the
// reference's COO, B-CSF, HB-CSF and MM-CSF arrays bit for bit.
//
// Reference flows restated here (util.h):
//   order_tensormode        255-311
//   sort_COOtensor / sort_MI_CSF   412-527 (full-key sort: one legal order)
//   create_HCSR / create_TiledHCSR 760-850 / 1320-1436
//   create_fbrLikeSlcInds   1770-1783
//   make_TiledBin           1887-1952
//   create_HYB / _4D, make_HybBin  852-1142, 1786-1885
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <rocprim/rocprim.hpp>

#include <algorithm>
#include <numeric>

#include "mmcsf_internal.hpp"

namespace mmcsf {

namespace {

constexpr unsigned kBS = 256;

inline unsigned nblk(uint64_t n) { return (unsigned)std::max<uint64_t>(1, (n + kBS - 1) / kBS); }

int bitwidth(uint64_t v) {  // bits needed for values 0..v
    int b = 1;
    while (b < 64 && (v >> b)) ++b;
    return b;
}

template <typename T>
T read_back(const T *dev, hipStream_t s) {
    T h{};
    MMCSF_HIP(hipMemcpyAsync(&h, dev, sizeof(T), hipMemcpyDeviceToHost, s));
    MMCSF_HIP(hipStreamSynchronize(s));
    return h;
}

struct Temp {
    DevArray<uint8_t> buf;
    void *get(size_t bytes) {
        if (buf.n < bytes) buf.alloc(std::max<size_t>(bytes, 256));
        return buf.p;
    }
};

// out[i] = sum_{j<i} in[j]; returns the total.
uint64_t exclusive_sum(const uint32_t *in, uint32_t *out, uint64_t n, hipStream_t s, Temp &tmp) {
    if (n == 0) return 0;
    size_t bytes = 0;
    MMCSF_HIP(rocprim::exclusive_scan(nullptr, bytes, in, out, 0u, (size_t)n, rocprim::plus<uint32_t>(), s));
    MMCSF_HIP(rocprim::exclusive_scan(tmp.get(bytes), bytes, in, out, 0u, (size_t)n,
                                      rocprim::plus<uint32_t>(), s));
    uint32_t last_in = read_back(in + n - 1, s), last_out = read_back(out + n - 1, s);
    return (uint64_t)last_in + last_out;
}

void inclusive_max(const uint32_t *in, uint32_t *out, uint64_t n, hipStream_t s, Temp &tmp) {
    if (n == 0) return;
    size_t bytes = 0;
    MMCSF_HIP(rocprim::inclusive_scan(nullptr, bytes, in, out, (size_t)n, rocprim::maximum<uint32_t>(), s));
    MMCSF_HIP(rocprim::inclusive_scan(tmp.get(bytes), bytes, in, out, (size_t)n,
                                      rocprim::maximum<uint32_t>(), s));
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
__global__ void k_iota(uint32_t *p, uint64_t n) {
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i < n) p[i] = (uint32_t)i;
}

struct PackSpec {
    const uint32_t *src[kMaxDims];
    int shift[kMaxDims];
    int nlev;
};

// key[p] = packed coordinates of element perm[p]
__global__ void k_pack(PackSpec ps, const uint32_t *__restrict__ perm, uint64_t *__restrict__ key,
                       uint64_t n) {
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t e = perm[i];
    uint64_t k = 0;
    for (int l = 0; l < ps.nlev; ++l) k |= (uint64_t)ps.src[l][e] << ps.shift[l];
    key[i] = k;
}

__global__ void k_gather_u32(const uint32_t *__restrict__ src, const uint32_t *__restrict__ perm,
                             uint32_t *__restrict__ dst, uint64_t n) {
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[perm[i]];
}

__global__ void k_gather_f32(const float *__restrict__ src, const uint32_t *__restrict__ perm,
                             float *__restrict__ dst, uint64_t n) {
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[perm[i]];
}

struct Keys {
    const uint32_t *k[kMaxDims];
};

// d[p] = first level whose coordinate differs from position p-1 (0 at p=0,
// N when the whole coordinate repeats); runPos[p] = p at a leaf-fiber key
// change (d <= N-2), else 0 -- the max-scan input for the -f chunk rule.
__global__ void k_first_diff(Keys K, int N, uint8_t *__restrict__ d, uint32_t *__restrict__ runPos,
                             uint64_t n) {
    uint64_t p = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (p >= n) return;
    int fd = 0;
    if (p > 0) {
        fd = N;
        for (int l = N - 1; l >= 0; --l)
            if (K.k[l][p] != K.k[l][p - 1]) fd = l;
    }
    d[p] = (uint8_t)fd;
    runPos[p] = (fd <= N - 2) ? (uint32_t)p : 0u;
}

// Leaf-fiber start flag: a key change at a fiber level, or the run of equal
// fiber keys reaching a multiple of the -f threshold (util.h:1364-1382).
__global__ void k_leaf_flag(const uint8_t *__restrict__ d, const uint32_t *__restrict__ runStart,
                            int N, uint64_t fthr, uint32_t *__restrict__ flag, uint64_t n) {
    uint64_t p = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (p >= n) return;
    bool f = d[p] <= N - 2;
    if (!f && runStart) f = ((p - runStart[p]) % fthr) == 0;
    flag[p] = f ? 1u : 0u;
}

// Scatter of the leaf-fiber level: pointer, index, first-diff level and
// start position of every leaf fiber.
__global__ void k_scatter_leaf(const uint32_t *__restrict__ flag, const uint32_t *__restrict__ fid,
                               const uint8_t *__restrict__ d, const uint32_t *__restrict__ key2,
                               uint32_t *__restrict__ ptr, uint32_t *__restrict__ idx,
                               uint8_t *__restrict__ dF, uint64_t n) {
    uint64_t p = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (p >= n || !flag[p]) return;
    uint32_t f = fid[p];
    ptr[f] = (uint32_t)p;
    idx[f] = key2[p];
    dF[f] = d[p];
}

// flag[f] = (dF[f] <= lev) for leaf fibers
__global__ void k_level_flag(const uint8_t *__restrict__ dF, int lev, uint32_t *__restrict__ flag,
                             uint64_t nf) {
    uint64_t f = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (f < nf) flag[f] = dF[f] <= lev ? 1u : 0u;
}

// 4-D fbrS split (util.h:1388-1404).  One thread walks one run of equal
// (k0,k1) keys over its leaf fibers.  A new fbrS starts at leaf fiber q when
// the current fbrS (started at nonzero s) already counts more than
// fsthr nonzeros by the reference's counter, which reads q - s for the very
// first fbrS of the tensor and q - s + 1 after every reset.
__global__ void k_fbrs_split(const uint32_t *__restrict__ runs, uint64_t nruns, uint64_t nf,
                             const uint32_t *__restrict__ leafPtr, uint64_t fsthr,
                             uint32_t *__restrict__ jflag) {
    uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (r >= nruns) return;
    uint32_t f0 = runs[r];
    uint32_t f1 = (r + 1 < nruns) ? runs[r + 1] : (uint32_t)nf;
    uint64_t s = leafPtr[f0];
    jflag[f0] = 1u;
    for (uint32_t f = f0 + 1; f < f1; ++f) {
        uint64_t q = leafPtr[f];
        uint64_t count = (q - s) + (s > 0 ? 1u : 0u);
        if (count > fsthr) {
            jflag[f] = 1u;
            s = q;
        } else {
            jflag[f] = 0u;
        }
    }
}

// Compaction of flagged leaf fibers into an upper level:
//   ptr[ord] = value_of(f) (index into the level below), idx[ord] = key[pos(f)]
__global__ void k_scatter_upper(const uint32_t *__restrict__ flag, const uint32_t *__restrict__ ord,
                                const uint32_t *__restrict__ below, const uint32_t *__restrict__ leafPtr,
                                const uint32_t *__restrict__ key, uint32_t *__restrict__ ptr,
                                uint32_t *__restrict__ idx, const uint32_t *__restrict__ key0,
                                uint32_t *__restrict__ slc, uint64_t nf) {
    uint64_t f = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (f >= nf || !flag[f]) return;
    uint32_t o = ord[f];
    ptr[o] = below ? below[f] : (uint32_t)f;
    uint32_t pos = leafPtr[f];
    idx[o] = key[pos];
    if (slc) slc[o] = key0[pos];
}

__global__ void k_set(uint32_t *p, uint64_t i, uint32_t v) { p[i] = v; }

__global__ void k_zero_words(uint32_t *p, uint64_t n) {
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// per 64-nonzero chunk: the number of slices starting in it
__global__ void k_chunk_starts(int N, const uint32_t *__restrict__ p0, const uint32_t *__restrict__ p1,
                               const uint32_t *__restrict__ p2, uint64_t nS, uint32_t *cnt) {
    uint64_t s = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (s >= nS) return;
    uint32_t x = p1[p0[s]];
    if (N == 4) x = p2[x];
    atomicAdd(cnt + (x >> 6), 1u);
}

// traversal cost of a chunk: 64 units (its leaf rows), wf per leaf fiber
// starting in it (the fiber-level rows, read once per fiber segment) and ws
// per slice start (row writes and the join of the group runs)
__global__ void k_chunk_cost(const uint32_t *__restrict__ cnt, const uint32_t *__restrict__ fbit, uint64_t n,
                             bool plain, uint32_t wf, uint32_t ws, uint32_t *cost) {
    uint64_t c = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (c < n) {
        const uint32_t fs = (uint32_t)(__popc(fbit[2 * c]) + __popc(fbit[2 * c + 1]));
        cost[c] = plain ? 1u : 64u + wf * fs + ws * cnt[c];
    }
}

// bit ptr[i] of `bits` for every entry i < n (first child of every entry)
__global__ void k_mark_starts(const uint32_t *__restrict__ ptr, uint64_t n, uint32_t *bits) {
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i < n) {
        uint32_t p = ptr[i];
        atomicOr(bits + (p >> 5), 1u << (p & 31));
    }
}

// ---------------------------------------------------------------------------
// bins
// ---------------------------------------------------------------------------
struct BinBounds {
    int64_t LB[kNumBins], UB[kNumBins];
};

BinBounds bin_bounds(uint64_t totNnz) {
    // make_TiledBin / make_HybBin bin table (util.h:1796-1812, 1897-1916)
    const int64_t TB = 512;
    BinBounds b;
    int64_t lb[kNumBins] = {0, 2, 4, 8, 16, 1024, 4 * TB, 8 * TB, 16 * TB, 32 * TB};
    int64_t ub[kNumBins] = {3, 5, 9, 17, 1025, 4 * TB + 1, 8 * TB + 1, 16 * TB + 1, 32 * TB + 1,
                            (int64_t)totNnz + 1};
    for (int i = 0; i < kNumBins; ++i) { b.LB[i] = lb[i]; b.UB[i] = ub[i]; }
    b.UB[0] = 1025;  // bins 0..4 merged
    return b;
}

// weight of slice s: (level-1 entries if withFibers) + nonzeros
__global__ void k_slice_bin(int N, const uint32_t *__restrict__ p0, const uint32_t *__restrict__ p1,
                            const uint32_t *__restrict__ p2, bool withFibers, BinBounds bb,
                            int8_t *__restrict__ bin, uint64_t nslc) {
    uint64_t s = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (s >= nslc) return;
    uint32_t a = p0[s], b = p0[s + 1];
    uint64_t nz;
    if (N == 3) nz = (uint64_t)p1[b] - p1[a];
    else nz = (uint64_t)p2[p1[b]] - p2[p1[a]];
    // the reference accumulates the weight in an int (util.h:1921)
    int w = (int)((withFibers ? (uint64_t)(b - a) : 0) + nz);
    int sel = -1;
    for (int k = 0; k < kNumBins; ++k)
        if ((int64_t)w > bb.LB[k] && (int64_t)w < bb.UB[k]) { sel = k; break; }
    bin[s] = (int8_t)sel;
}

__global__ void k_csl_bin(const uint32_t *__restrict__ ptr, BinBounds bb, int8_t *__restrict__ bin,
                          uint64_t n) {
    uint64_t s = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (s >= n) return;
    int w = (int)(ptr[s + 1] - ptr[s]);
    int sel = -1;
    for (int k = 0; k < kNumBins; ++k)
        if ((int64_t)w > bb.LB[k] && (int64_t)w < bb.UB[k]) { sel = k; break; }
    bin[s] = (int8_t)sel;
}

__global__ void k_eq_flag(const int8_t *__restrict__ v, int8_t want, uint32_t *__restrict__ flag,
                          uint64_t n) {
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i < n) flag[i] = v[i] == want ? 1u : 0u;
}

__global__ void k_compact_idx(const uint32_t *__restrict__ flag, const uint32_t *__restrict__ pos,
                              uint32_t *__restrict__ out, uint64_t n) {
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i < n && flag[i]) out[pos[i]] = (uint32_t)i;
}

// stable per-bin lists of slice ordinals (slcMapperBin push order)
void bin_lists(const int8_t *bin, uint64_t n, DevArray<uint32_t> &list, uint64_t *counts,
               uint64_t &dropped, hipStream_t s, Temp &tmp) {
    DevArray<uint32_t> flag(n), pos(n);
    uint64_t tot = 0, cnt[kNumBins];
    std::vector<DevArray<uint32_t>> parts(kNumBins);
    for (int b = 0; b < kNumBins; ++b) {
        hipLaunchKernelGGL(k_eq_flag, dim3(nblk(n)), dim3(kBS), 0, s, bin, (int8_t)b, flag.p, n);
        cnt[b] = exclusive_sum(flag.p, pos.p, n, s, tmp);
        parts[b].alloc(cnt[b]);
        if (cnt[b])
            hipLaunchKernelGGL(k_compact_idx, dim3(nblk(n)), dim3(kBS), 0, s, flag.p, pos.p,
                               parts[b].p, n);
        tot += cnt[b];
    }
    list.alloc(tot);
    uint64_t off = 0;
    for (int b = 0; b < kNumBins; ++b) {
        counts[b] = cnt[b];
        if (cnt[b])
            MMCSF_HIP(hipMemcpyAsync(list.p + off, parts[b].p, cnt[b] * 4, hipMemcpyDeviceToDevice, s));
        off += cnt[b];
    }
    MMCSF_HIP(hipStreamSynchronize(s));
    dropped = n - tot;
}

// seg[i] = segment containing item i for a pointer array ptr[0..nseg]
__global__ void k_mark_seg(const uint32_t *__restrict__ ptr, uint32_t *__restrict__ mark, uint64_t nseg) {
    uint64_t s = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (s >= 1 && s < nseg) atomicAdd(&mark[ptr[s]], 1u);  // segments are never empty
}

void segment_ids(const uint32_t *ptr, uint64_t nseg, uint64_t nitems, DevArray<uint32_t> &seg,
                 hipStream_t s, Temp &tmp) {
    seg.alloc(nitems);
    if (nitems == 0) return;
    DevArray<uint32_t> mark(nitems);
    MMCSF_HIP(hipMemsetAsync(mark.p, 0, nitems * 4, s));
    hipLaunchKernelGGL(k_mark_seg, dim3(nblk(nseg)), dim3(kBS), 0, s, ptr, mark.p, nseg);
    size_t bytes = 0;
    MMCSF_HIP(rocprim::inclusive_scan(nullptr, bytes, mark.p, seg.p, (size_t)nitems,
                                      rocprim::plus<uint32_t>(), s));
    MMCSF_HIP(rocprim::inclusive_scan(tmp.get(bytes), bytes, mark.p, seg.p, (size_t)nitems,
                                      rocprim::plus<uint32_t>(), s));
}

}  // namespace

// ---------------------------------------------------------------------------
// order_tensormode, util.h:255-311
// ---------------------------------------------------------------------------
void order_tensormode(int ndims, const uint64_t *dims, int mode, bool nat, int *order) {
    int natMode[kMaxDims], len[kMaxDims];
    bool taken[kMaxDims];
    for (int m = 0; m < ndims; ++m) {
        natMode[m] = (m + mode) % ndims;
        len[m] = (int)dims[natMode[m]];
        taken[m] = false;
    }
    if (nat) {
        for (int i = 0; i < ndims; ++i) order[i] = natMode[i];
        return;
    }
    // ascending lengths of positions 1..N-1; position 0 stays `mode`
    for (int i = 1; i < ndims; i++)
        for (int j = i + 1; j < ndims; j++)
            if (len[i] > len[j]) std::swap(len[i], len[j]);
    order[0] = mode;
    taken[mode] = true;
    for (int i = 1; i < ndims; i++)
        for (int j = 0; j < ndims; j++)
            if ((uint64_t)(int64_t)len[i] == dims[j] && !taken[j]) {
                order[i] = j;
                taken[j] = true;
                break;
            }
}

// ---------------------------------------------------------------------------
// full-key LSD radix sort of the nonzeros
// ---------------------------------------------------------------------------
void sort_coo(const Tensor &t, const int *order, DevArray<uint32_t> *keys, DevArray<float> &vals,
              hipStream_t s, DevArray<uint32_t> *pos) {
    const int N = t.ndims;
    const uint64_t n = t.nnz;
    if (n == 0) {
        for (int l = 0; l < N; ++l) keys[l].alloc(0);
        vals.alloc(0);
        if (pos) pos->alloc(0);
        return;
    }
    Temp tmp;
    int bits[kMaxDims];
    for (int l = 0; l < N; ++l) bits[l] = bitwidth(t.dims[order[l]] ? t.dims[order[l]] - 1 : 0);
    // group levels into <=64-bit passes, least significant group first
    std::vector<std::pair<int, int>> groups;  // [lo, hi) level ranges
    int hi = N;
    while (hi > 0) {
        int lo = hi, tot = 0;
        while (lo > 0 && tot + bits[lo - 1] <= 64) { tot += bits[lo - 1]; --lo; }
        if (lo == hi) fail(MMCSF_ERR_UNSUPPORTED, "coordinate wider than 64 bits");
        groups.push_back({lo, hi});
        hi = lo;
    }
    DevArray<uint32_t> perm(n), perm2(n);
    DevArray<uint64_t> k1(n), k2(n);
    hipLaunchKernelGGL(k_iota, dim3(nblk(n)), dim3(kBS), 0, s, perm.p, n);
    for (auto [lo, hi2] : groups) {
        PackSpec ps{};
        int shift = 0;
        ps.nlev = hi2 - lo;
        for (int l = hi2 - 1; l >= lo; --l) {
            ps.src[l - lo] = t.inds[order[l]].p;
            ps.shift[l - lo] = shift;
            shift += bits[l];
        }
        hipLaunchKernelGGL(k_pack, dim3(nblk(n)), dim3(kBS), 0, s, ps, perm.p, k1.p, n);
        size_t bytes = 0;
        MMCSF_HIP(rocprim::radix_sort_pairs(nullptr, bytes, k1.p, k2.p, perm.p, perm2.p, (size_t)n, 0u,
                                            (unsigned)shift, s));
        MMCSF_HIP(rocprim::radix_sort_pairs(tmp.get(bytes), bytes, k1.p, k2.p, perm.p, perm2.p,
                                            (size_t)n, 0u, (unsigned)shift, s));
        std::swap(perm, perm2);
    }
    k1.release();
    k2.release();
    for (int l = 0; l < N; ++l) {
        keys[l].alloc(n);
        hipLaunchKernelGGL(k_gather_u32, dim3(nblk(n)), dim3(kBS), 0, s, t.inds[order[l]].p, perm.p,
                           keys[l].p, n);
    }
    vals.alloc(n);
    hipLaunchKernelGGL(k_gather_f32, dim3(nblk(n)), dim3(kBS), 0, s, t.vals.p, perm.p, vals.p, n);
    MMCSF_HIP(hipStreamSynchronize(s));
    if (pos) *pos = std::move(perm);
}

// ---------------------------------------------------------------------------
// CSF scan (create_TiledHCSR semantics)
// ---------------------------------------------------------------------------
void build_csf(int N, const uint64_t *dims, uint64_t n, const int *order, DevArray<uint32_t> *keys,
               DevArray<float> &&vals, uint64_t fthr, uint64_t fsthr, Csf &c, hipStream_t s) {
    Temp tmp;
    c.ndims = N;
    for (int l = 0; l < N; ++l) {
        c.order[l] = order[l];
        c.rows[l] = dims[order[l]];
    }
    c.nnz = n;
    if (n == 0) {
        for (int l = 0; l < N - 1; ++l) {
            c.cnt[l] = 0;
            c.ptr[l].alloc(1);
            hipLaunchKernelGGL(k_set, dim3(1), dim3(1), 0, s, c.ptr[l].p, 0, 0u);
        }
        c.cnt[N - 1] = 0;
        MMCSF_HIP(hipStreamSynchronize(s));
        return;
    }
    if (fthr < 1) fthr = 1;  // -f 0 behaves like -f 1 (util.h:1364: fiberNnz starts at 1)
    Keys K{};
    for (int l = 0; l < N; ++l) K.k[l] = keys[l].p;
    DevArray<uint8_t> d(n);
    DevArray<uint32_t> a(n), b(n);  // a: run positions -> run starts; b: flags
    hipLaunchKernelGGL(k_first_diff, dim3(nblk(n)), dim3(kBS), 0, s, K, N, d.p, a.p, n);
    const uint32_t *runStart = nullptr;
    DevArray<uint32_t> rs;
    if (fthr < n) {
        rs.alloc(n);
        inclusive_max(a.p, rs.p, n, s, tmp);
        runStart = rs.p;
    }
    hipLaunchKernelGGL(k_leaf_flag, dim3(nblk(n)), dim3(kBS), 0, s, d.p, runStart, N, fthr, b.p, n);
    rs.release();
    // leaf-fiber ordinals
    uint64_t nLeaf = exclusive_sum(b.p, a.p, n, s, tmp);
    const int LL = N - 2;  // leaf-fiber level
    c.ptr[LL].alloc(nLeaf + 1);
    c.idx[LL].alloc(nLeaf);
    DevArray<uint8_t> dF(nLeaf);
    hipLaunchKernelGGL(k_scatter_leaf, dim3(nblk(n)), dim3(kBS), 0, s, b.p, a.p, d.p, keys[LL].p,
                       c.ptr[LL].p, c.idx[LL].p, dF.p, n);
    hipLaunchKernelGGL(k_set, dim3(1), dim3(1), 0, s, c.ptr[LL].p, nLeaf, (uint32_t)n);
    c.cnt[LL] = nLeaf;
    d.release();
    a.release();
    b.release();

    DevArray<uint32_t> flag(nLeaf), ord(nLeaf);
    if (N == 3) {
        // slices over leaf fibers
        hipLaunchKernelGGL(k_level_flag, dim3(nblk(nLeaf)), dim3(kBS), 0, s, dF.p, 0, flag.p, nLeaf);
        uint64_t nS = exclusive_sum(flag.p, ord.p, nLeaf, s, tmp);
        c.ptr[0].alloc(nS + 1);
        c.idx[0].alloc(nS);
        hipLaunchKernelGGL(k_scatter_upper, dim3(nblk(nLeaf)), dim3(kBS), 0, s, flag.p, ord.p,
                           (const uint32_t *)nullptr, c.ptr[LL].p, keys[0].p, c.ptr[0].p, c.idx[0].p,
                           (const uint32_t *)nullptr, (uint32_t *)nullptr, nLeaf);
        hipLaunchKernelGGL(k_set, dim3(1), dim3(1), 0, s, c.ptr[0].p, nS, (uint32_t)nLeaf);
        c.cnt[0] = nS;
        // fbrLikeSlcInds: slice coordinate of every fiber = key0 at its start
        c.fbrSlc.alloc(nLeaf);
        hipLaunchKernelGGL(k_gather_u32, dim3(nblk(nLeaf)), dim3(kBS), 0, s, keys[0].p, c.ptr[LL].p,
                           c.fbrSlc.p, nLeaf);
    } else {
        // level 1 (fbrS): key change at level <= 1, plus the 128-nnz split
        DevArray<uint32_t> jflag(nLeaf);
        hipLaunchKernelGGL(k_level_flag, dim3(nblk(nLeaf)), dim3(kBS), 0, s, dF.p, 1, flag.p, nLeaf);
        if (fsthr < n) {
            uint64_t nRuns = exclusive_sum(flag.p, ord.p, nLeaf, s, tmp);
            DevArray<uint32_t> runs(nRuns);
            hipLaunchKernelGGL(k_compact_idx, dim3(nblk(nLeaf)), dim3(kBS), 0, s, flag.p, ord.p, runs.p,
                               nLeaf);
            hipLaunchKernelGGL(k_fbrs_split, dim3(nblk(nRuns)), dim3(kBS), 0, s, runs.p, nRuns, nLeaf,
                               c.ptr[LL].p, fsthr, jflag.p);
        } else {
            MMCSF_HIP(hipMemcpyAsync(jflag.p, flag.p, nLeaf * 4, hipMemcpyDeviceToDevice, s));
        }
        uint64_t nJ = exclusive_sum(jflag.p, ord.p, nLeaf, s, tmp);
        c.ptr[1].alloc(nJ + 1);
        c.idx[1].alloc(nJ);
        c.fbrSlc.alloc(nJ);
        hipLaunchKernelGGL(k_scatter_upper, dim3(nblk(nLeaf)), dim3(kBS), 0, s, jflag.p, ord.p,
                           (const uint32_t *)nullptr, c.ptr[LL].p, keys[1].p, c.ptr[1].p, c.idx[1].p,
                           keys[0].p, c.fbrSlc.p, nLeaf);
        hipLaunchKernelGGL(k_set, dim3(1), dim3(1), 0, s, c.ptr[1].p, nJ, (uint32_t)nLeaf);
        c.cnt[1] = nJ;
        // level 0 (slices): ptr0 = fbrS ordinal of the slice's first leaf fiber
        DevArray<uint32_t> sord(nLeaf);
        hipLaunchKernelGGL(k_level_flag, dim3(nblk(nLeaf)), dim3(kBS), 0, s, dF.p, 0, flag.p, nLeaf);
        uint64_t nS = exclusive_sum(flag.p, sord.p, nLeaf, s, tmp);
        c.ptr[0].alloc(nS + 1);
        c.idx[0].alloc(nS);
        hipLaunchKernelGGL(k_scatter_upper, dim3(nblk(nLeaf)), dim3(kBS), 0, s, flag.p, sord.p, ord.p,
                           c.ptr[LL].p, keys[0].p, c.ptr[0].p, c.idx[0].p, (const uint32_t *)nullptr,
                           (uint32_t *)nullptr, nLeaf);
        hipLaunchKernelGGL(k_set, dim3(1), dim3(1), 0, s, c.ptr[0].p, nS, (uint32_t)nJ);
        c.cnt[0] = nS;
    }
    c.cnt[N - 1] = n;
    c.leaf = std::move(keys[N - 1]);
    c.vals = std::move(vals);
    for (int l = 0; l < N - 1; ++l) keys[l].release();
    MMCSF_HIP(hipStreamSynchronize(s));
    build_walk_bits(c, s);
}

// ---------------------------------------------------------------------------
// start bitmasks for the chunked root traversal: bit x of fbit = nonzero x
// starts a leaf fiber; 4-D: bit f of sbit = leaf fiber f starts a level-1
// entry.  Word arrays carry two spare words so 64-entry windows never read
// past the end.
// ---------------------------------------------------------------------------
void build_walk_bits(Csf &c, hipStream_t s) {
    const int N = c.ndims;
    const uint64_t n = c.nnz, nL = c.cnt[N - 2];
    auto mark = [&](DevArray<uint32_t> &bits, uint64_t items, const uint32_t *ptr, uint64_t nptr) {
        const uint64_t words = items / 32 + 3;
        bits.alloc(words);
        hipLaunchKernelGGL(k_zero_words, dim3(nblk(words)), dim3(kBS), 0, s, bits.p, words);
        if (nptr) hipLaunchKernelGGL(k_mark_starts, dim3(nblk(nptr)), dim3(kBS), 0, s, ptr, nptr, bits.p);
    };
    mark(c.fbit, n, c.ptr[N - 2].p, nL);
    if (N == 4) mark(c.sbit, nL, c.ptr[1].p, c.cnt[1]);
    // cost prefix over chunks: waves split the traversal by cost, not by
    // nonzeros (chunks where many slices start take longer)
    const uint64_t chunks = (n + 63) / 64;
    c.chunkCost.alloc(chunks + 1);
    if (chunks) {
        Temp tmp;
        DevArray<uint32_t> cnt(chunks), cost(chunks);
        hipLaunchKernelGGL(k_zero_words, dim3(nblk(chunks)), dim3(kBS), 0, s, cnt.p, chunks);
        if (c.cnt[0])
            hipLaunchKernelGGL(k_chunk_starts, dim3(nblk(c.cnt[0])), dim3(kBS), 0, s, N, c.ptr[0].p, c.ptr[1].p,
                               N == 4 ? c.ptr[2].p : (const uint32_t *)nullptr, c.cnt[0], cnt.p);
        // the weighted total (at most chunks * 64 (1 + wf) + ws * slices, every
        // slice starting in one chunk) must fit the 32-bit prefix; else split
        // by chunks
        const uint32_t wf = 1, ws = 24;
        const bool plain = chunks * 64ull * (1 + wf) + ws * (uint64_t)c.cnt[0] >= 0xF0000000ull;
        hipLaunchKernelGGL(k_chunk_cost, dim3(nblk(chunks)), dim3(kBS), 0, s, cnt.p, c.fbit.p, chunks, plain, wf, ws,
                           cost.p);
        uint64_t total = exclusive_sum(cost.p, c.chunkCost.p, chunks, s, tmp);
        hipLaunchKernelGGL(k_set, dim3(1), dim3(1), 0, s, c.chunkCost.p, chunks, (uint32_t)total);
    } else {
        hipLaunchKernelGGL(k_set, dim3(1), dim3(1), 0, s, c.chunkCost.p, 0, 0u);
    }
    MMCSF_HIP(hipGetLastError());
    MMCSF_HIP(hipStreamSynchronize(s));
}

// ---------------------------------------------------------------------------
// make_TiledBin (util.h:1887-1952)
// ---------------------------------------------------------------------------
void build_bins_tiled(Csf &c, hipStream_t s) {
    Temp tmp;
    uint64_t nS = c.cnt[0];
    DevArray<int8_t> bin(nS);
    if (nS)
        hipLaunchKernelGGL(k_slice_bin, dim3(nblk(nS)), dim3(kBS), 0, s, c.ndims, c.ptr[0].p, c.ptr[1].p,
                           c.ndims == 4 ? c.ptr[2].p : (const uint32_t *)nullptr, true, bin_bounds(c.nnz),
                           bin.p, nS);
    bin_lists(bin.p, nS, c.binList, c.binCount, c.dropped, s, tmp);
    c.hasBins = true;
}

// ---------------------------------------------------------------------------
// create_HYB / create_HYB_4D + make_HybBin
// ---------------------------------------------------------------------------
namespace {

enum : uint8_t { kCOO = 0, kCSL = 1, kCSF = 2 };

// per slice: class, nnz, non-CSF nnz (for the shift), COO/CSL/CSF flags
__global__ void k_hyb_class(int N, const uint32_t *__restrict__ p0, const uint32_t *__restrict__ p1,
                            const uint32_t *__restrict__ p2, uint8_t *__restrict__ cls,
                            uint32_t *__restrict__ fCoo, uint32_t *__restrict__ fCsl,
                            uint32_t *__restrict__ fCsf, uint32_t *__restrict__ nzCsl,
                            uint32_t *__restrict__ nzSub, uint64_t nslc) {
    uint64_t s = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (s >= nslc) return;
    uint32_t a = p0[s], b = p0[s + 1];
    uint64_t nLeafF, nz;
    if (N == 3) {
        nLeafF = b - a;
        nz = (uint64_t)p1[b] - p1[a];
    } else {
        nLeafF = (uint64_t)p1[b] - p1[a];
        nz = (uint64_t)p2[p1[b]] - p2[p1[a]];
    }
    uint8_t k;
    if (nz == 1) k = kCOO;                        // util.h:883 / 1026
    else if (N == 3 && nz == nLeafF) k = kCSL;    // every fiber of length 1 (util.h:885)
    else k = kCSF;
    cls[s] = k;
    fCoo[s] = k == kCOO;
    fCsl[s] = k == kCSL;
    fCsf[s] = k == kCSF;
    nzCsl[s] = k == kCSL ? (uint32_t)nz : 0u;
    // fibers (= nonzeros) consumed by COO / CSL slices (util.h:957, 1104)
    nzSub[s] = k == kCSF ? 0u : (uint32_t)nz;
}

struct HybOut {
    uint32_t *cooInd[kMaxDims];
    float *cooVal;
    uint32_t *cslPtr, *cslIdx, *cslInd1, *cslInd2;
    float *cslVal;
    uint32_t *hPtr[kMaxDims - 1], *hIdx[kMaxDims - 1];
    uint32_t *hLeaf;
    float *hVal;
};

struct CsfIn {
    const uint32_t *ptr[kMaxDims - 1], *idx[kMaxDims - 1];
    const uint32_t *leaf;
    const float *vals;
};

// slice-level outputs: COO entries, CSL slice pointers, CSF slice pointers
__global__ void k_hyb_slices(int N, CsfIn in, const uint8_t *__restrict__ cls,
                             const uint32_t *__restrict__ oCoo, const uint32_t *__restrict__ oCsl,
                             const uint32_t *__restrict__ oCsf, const uint32_t *__restrict__ cslOff,
                             const uint32_t *__restrict__ sub, HybOut o, uint64_t nslc) {
    uint64_t s = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (s >= nslc) return;
    uint32_t a = in.ptr[0][s];
    uint8_t k = cls[s];
    if (k == kCOO) {
        uint32_t c = oCoo[s];
        uint32_t leafF = (N == 3) ? a : in.ptr[1][a];
        uint32_t x = in.ptr[N - 2][leafF];
        o.cooInd[0][c] = in.idx[0][s];
        o.cooInd[1][c] = in.idx[1][a];
        if (N == 4) o.cooInd[2][c] = in.idx[2][leafF];
        o.cooInd[N - 1][c] = in.leaf[x];
        o.cooVal[c] = in.vals[x];
    } else if (k == kCSL) {
        uint32_t c = oCsl[s];
        o.cslPtr[c] = cslOff[s];
        o.cslIdx[c] = in.idx[0][s];
    } else {
        uint32_t c = oCsf[s];
        o.hPtr[0][c] = a - sub[s];
        o.hIdx[0][c] = in.idx[0][s];
    }
}

// level-1 entries (3-D: fibers; 4-D: fbrS) of CSL / CSF slices
__global__ void k_hyb_lvl1(int N, CsfIn in, const uint32_t *__restrict__ slcOf,
                           const uint8_t *__restrict__ cls, const uint32_t *__restrict__ cslOff,
                           const uint32_t *__restrict__ sub, HybOut o, uint64_t n1) {
    uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (j >= n1) return;
    uint32_t s = slcOf[j];
    uint8_t k = cls[s];
    if (k == kCSL) {  // 3-D only: one nonzero per fiber
        uint32_t pos = cslOff[s] + (uint32_t)(j - in.ptr[0][s]);
        uint32_t x = in.ptr[1][j];
        o.cslInd1[pos] = in.idx[1][j];
        o.cslInd2[pos] = in.leaf[x];
        o.cslVal[pos] = in.vals[x];
    } else if (k == kCSF) {
        uint32_t nj = (uint32_t)j - sub[s];
        o.hPtr[1][nj] = in.ptr[1][j] - sub[s];
        o.hIdx[1][nj] = in.idx[1][j];
    }
}

// 4-D leaf fibers of CSF slices
__global__ void k_hyb_lvl2(CsfIn in, const uint32_t *__restrict__ slcOf2,
                           const uint8_t *__restrict__ cls, const uint32_t *__restrict__ sub, HybOut o,
                           uint64_t n2) {
    uint64_t f = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (f >= n2) return;
    uint32_t s = slcOf2[f];
    if (cls[s] != kCSF) return;
    uint32_t nf = (uint32_t)f - sub[s];
    o.hPtr[2][nf] = in.ptr[2][f] - sub[s];
    o.hIdx[2][nf] = in.idx[2][f];
}

__global__ void k_hyb_nnz(CsfIn in, const uint32_t *__restrict__ slcOfNnz, const uint8_t *__restrict__ cls,
                          const uint32_t *__restrict__ sub, HybOut o, uint64_t n) {
    uint64_t x = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (x >= n) return;
    uint32_t s = slcOfNnz[x];
    if (cls[s] != kCSF) return;
    uint32_t nx = (uint32_t)x - sub[s];
    o.hLeaf[nx] = in.leaf[x];
    o.hVal[nx] = in.vals[x];
}

// compose two segment maps: out[i] = outer[inner[i]]
__global__ void k_compose(const uint32_t *__restrict__ inner, const uint32_t *__restrict__ outer,
                          uint32_t *__restrict__ out, uint64_t n) {
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i < n) out[i] = outer[inner[i]];
}

}  // namespace

// level-1 entry x of the CSL walk tree (one per nonzero) takes its slice's
// coordinate
__global__ void k_csl_slice_of(const uint32_t *__restrict__ ptr, const uint32_t *__restrict__ idx, uint64_t nS,
                               uint32_t *slc) {
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i >= nS) return;
    for (uint32_t x = ptr[i]; x < ptr[i + 1]; ++x) slc[x] = idx[i];
}

// The CSL part (slices whose leaf fibers all hold one nonzero) is a CSF tree
// with one level-1 entry per nonzero: ptr[0] = cslPtr, idx[0] = cslIdx,
// ptr[1] = 0, 1, ..., idx[1] = cslInd1, leaf = cslInd2.  Walked by the chunked
// root walk like the CSF part; the CSL arrays themselves stay as built.
void build_csl_walk(Hyb &h, hipStream_t s) {
    Csf &t = h.cslWalk;
    t = Csf{};
    if (h.ndims != 3 || h.cslNnz == 0) return;
    const uint64_t n = h.cslNnz, nS = h.nCsl;
    t.ndims = 3;
    for (int l = 0; l < 3; ++l) {
        t.order[l] = h.order[l];
        t.rows[l] = h.csf.rows[l];
    }
    t.nnz = n;
    t.cnt[0] = nS;
    t.cnt[1] = n;
    t.cnt[2] = n;
    auto copy = [&](DevArray<uint32_t> &dst, const uint32_t *src, uint64_t count) {
        dst.alloc(count);
        MMCSF_HIP(hipMemcpyAsync(dst.p, src, count * 4, hipMemcpyDeviceToDevice, s));
    };
    copy(t.ptr[0], h.cslPtr.p, nS + 1);
    copy(t.idx[0], h.cslIdx.p, nS);
    t.ptr[1].alloc(n + 1);
    hipLaunchKernelGGL(k_iota, dim3(nblk(n + 1)), dim3(kBS), 0, s, t.ptr[1].p, n + 1);
    copy(t.idx[1], h.cslInd1.p, n);
    copy(t.leaf, h.cslInd2.p, n);
    t.vals.alloc(n);
    MMCSF_HIP(hipMemcpyAsync(t.vals.p, h.cslVal.p, n * 4, hipMemcpyDeviceToDevice, s));
    t.fbrSlc.alloc(n);
    hipLaunchKernelGGL(k_csl_slice_of, dim3(nblk(nS)), dim3(kBS), 0, s, h.cslPtr.p, h.cslIdx.p, nS, t.fbrSlc.p);
    MMCSF_HIP(hipGetLastError());
    MMCSF_HIP(hipStreamSynchronize(s));
    build_walk_bits(t, s);
}

void build_hyb(int N, const uint64_t *dims, Csf &&src, Hyb &h, hipStream_t s) {
    (void)dims;
    Temp tmp;
    h.ndims = N;
    for (int l = 0; l < N; ++l) h.order[l] = src.order[l];
    const uint64_t n = src.nnz;
    const uint64_t nS = src.cnt[0];
    const uint64_t n1 = src.cnt[1];
    const uint64_t nL = src.cnt[N - 2];
    DevArray<uint8_t> cls(nS);
    DevArray<uint32_t> fCoo(nS), fCsl(nS), fCsf(nS), nzCsl(nS), nzSub(nS);
    DevArray<uint32_t> oCoo(nS), oCsl(nS), oCsf(nS), cslOff(nS), sub(nS);
    if (nS)
        hipLaunchKernelGGL(k_hyb_class, dim3(nblk(nS)), dim3(kBS), 0, s, N, src.ptr[0].p, src.ptr[1].p,
                           N == 4 ? src.ptr[2].p : (const uint32_t *)nullptr, cls.p, fCoo.p, fCsl.p,
                           fCsf.p, nzCsl.p, nzSub.p, nS);
    uint64_t nCoo = exclusive_sum(fCoo.p, oCoo.p, nS, s, tmp);
    uint64_t nCsl = exclusive_sum(fCsl.p, oCsl.p, nS, s, tmp);
    uint64_t nCsf = exclusive_sum(fCsf.p, oCsf.p, nS, s, tmp);
    uint64_t cslNnz = exclusive_sum(nzCsl.p, cslOff.p, nS, s, tmp);
    uint64_t subTot = exclusive_sum(nzSub.p, sub.p, nS, s, tmp);
    h.cooNnz = nCoo;
    h.cslNnz = cslNnz;
    h.nCsl = nCsl;
    for (int l = 0; l < N; ++l) h.cooInd[l].alloc(nCoo);
    h.cooVal.alloc(nCoo);
    h.cslPtr.alloc(nCsl + 1);
    h.cslIdx.alloc(nCsl);
    h.cslInd1.alloc(cslNnz);
    h.cslInd2.alloc(cslNnz);
    h.cslVal.alloc(cslNnz);
    // CSF part sizes: every non-CSF slice consumes exactly one entry per level
    // for COO (4-D) and one fiber per nonzero for CSL (3-D)
    Csf &c = h.csf;
    c.ndims = N;
    for (int l = 0; l < N; ++l) {
        c.order[l] = src.order[l];
        c.rows[l] = src.rows[l];
    }
    c.nnz = n - subTot;
    c.cnt[0] = nCsf;
    if (N == 3) {
        c.cnt[1] = nL - subTot;
    } else {
        c.cnt[1] = n1 - subTot;  // a 4-D non-CSF slice is a COO slice: one fbrS
        c.cnt[2] = nL - subTot;
    }
    c.cnt[N - 1] = c.nnz;
    for (int l = 0; l < N - 1; ++l) {
        c.ptr[l].alloc(c.cnt[l] + 1);
        c.idx[l].alloc(c.cnt[l]);
    }
    c.leaf.alloc(c.nnz);
    c.vals.alloc(c.nnz);

    HybOut o{};
    for (int l = 0; l < N; ++l) o.cooInd[l] = h.cooInd[l].p;
    o.cooVal = h.cooVal.p;
    o.cslPtr = h.cslPtr.p;
    o.cslIdx = h.cslIdx.p;
    o.cslInd1 = h.cslInd1.p;
    o.cslInd2 = h.cslInd2.p;
    o.cslVal = h.cslVal.p;
    for (int l = 0; l < N - 1; ++l) { o.hPtr[l] = c.ptr[l].p; o.hIdx[l] = c.idx[l].p; }
    o.hLeaf = c.leaf.p;
    o.hVal = c.vals.p;
    CsfIn in{};
    for (int l = 0; l < N - 1; ++l) { in.ptr[l] = src.ptr[l].p; in.idx[l] = src.idx[l].p; }
    in.leaf = src.leaf.p;
    in.vals = src.vals.p;

    if (nS)
        hipLaunchKernelGGL(k_hyb_slices, dim3(nblk(nS)), dim3(kBS), 0, s, N, in, cls.p, oCoo.p, oCsl.p,
                           oCsf.p, cslOff.p, sub.p, o, nS);
    // level-1 entries
    DevArray<uint32_t> slcOf1;
    segment_ids(src.ptr[0].p, nS, n1, slcOf1, s, tmp);
    if (n1)
        hipLaunchKernelGGL(k_hyb_lvl1, dim3(nblk(n1)), dim3(kBS), 0, s, N, in, slcOf1.p, cls.p, cslOff.p,
                           sub.p, o, n1);
    DevArray<uint32_t> slcOfLeafF;
    if (N == 4) {
        DevArray<uint32_t> jOf2;
        segment_ids(src.ptr[1].p, n1, nL, jOf2, s, tmp);
        slcOfLeafF.alloc(nL);
        if (nL)
            hipLaunchKernelGGL(k_compose, dim3(nblk(nL)), dim3(kBS), 0, s, jOf2.p, slcOf1.p, slcOfLeafF.p,
                               nL);
        if (nL)
            hipLaunchKernelGGL(k_hyb_lvl2, dim3(nblk(nL)), dim3(kBS), 0, s, in, slcOfLeafF.p, cls.p, sub.p,
                               o, nL);
    } else {
        slcOfLeafF = std::move(slcOf1);
    }
    // nonzeros
    {
        DevArray<uint32_t> fOfNnz, slcOfNnz(n);
        segment_ids(src.ptr[N - 2].p, nL, n, fOfNnz, s, tmp);
        if (n)
            hipLaunchKernelGGL(k_compose, dim3(nblk(n)), dim3(kBS), 0, s, fOfNnz.p, slcOfLeafF.p,
                               slcOfNnz.p, n);
        if (n)
            hipLaunchKernelGGL(k_hyb_nnz, dim3(nblk(n)), dim3(kBS), 0, s, in, slcOfNnz.p, cls.p, sub.p, o,
                               n);
    }
    // closing entries (util.h:976-978, 1125-1127)
    hipLaunchKernelGGL(k_set, dim3(1), dim3(1), 0, s, c.ptr[N - 2].p, c.cnt[N - 2], (uint32_t)c.nnz);
    for (int l = N - 3; l >= 0; --l)
        hipLaunchKernelGGL(k_set, dim3(1), dim3(1), 0, s, c.ptr[l].p, c.cnt[l], (uint32_t)c.cnt[l + 1]);
    hipLaunchKernelGGL(k_set, dim3(1), dim3(1), 0, s, h.cslPtr.p, nCsl, (uint32_t)cslNnz);
    // fbrLikeSlcInds of the CSF part (used by the kernels)
    {
        DevArray<uint32_t> so;
        segment_ids(c.ptr[0].p, c.cnt[0], c.cnt[1], so, s, tmp);
        c.fbrSlc.alloc(c.cnt[1]);
        if (c.cnt[1])
            hipLaunchKernelGGL(k_compose, dim3(nblk(c.cnt[1])), dim3(kBS), 0, s, so.p, c.idx[0].p,
                               c.fbrSlc.p, c.cnt[1]);
    }
    // bins: CSF slices by nnz (make_HybBin), CSL slices by nnz
    {
        DevArray<int8_t> bin(c.cnt[0]);
        if (c.cnt[0])
            hipLaunchKernelGGL(k_slice_bin, dim3(nblk(c.cnt[0])), dim3(kBS), 0, s, N, c.ptr[0].p,
                               c.ptr[1].p, N == 4 ? c.ptr[2].p : (const uint32_t *)nullptr, false,
                               bin_bounds(n), bin.p, c.cnt[0]);
        bin_lists(bin.p, c.cnt[0], c.binList, c.binCount, c.dropped, s, tmp);
        c.hasBins = true;
        DevArray<int8_t> cb(nCsl);
        if (nCsl)
            hipLaunchKernelGGL(k_csl_bin, dim3(nblk(nCsl)), dim3(kBS), 0, s, h.cslPtr.p, bin_bounds(n), cb.p,
                               nCsl);
        uint64_t cdrop = 0;
        bin_lists(cb.p, nCsl, h.cslBinList, h.cslBinCount, cdrop, s, tmp);
    }
    MMCSF_HIP(hipStreamSynchronize(s));
    build_walk_bits(c, s);
    // the source CSF is consumed
    Csf dead = std::move(src);
    (void)dead;
}


// ---------------------------------------------------------------------------
// Hot-row tables of a scatter level: the rows of level `level` used most
// often (by nonzeros for the leaf level, by tree entries for fiber levels)
// are ranked; the level's index array is re-encoded so that a row of rank
// r < kMaxHot reads kHotBit | r and every other row reads its coordinate.
// The scatter kernels accumulate ranked rows in LDS.
// ---------------------------------------------------------------------------
namespace {

__global__ void k_hist(const uint32_t *__restrict__ idx, uint64_t n, uint32_t *__restrict__ cnt) {
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i < n) atomicAdd(&cnt[idx[i]], 1u);
}

__global__ void k_rank_table(const uint32_t *__restrict__ rowsByCount, uint32_t nhot,
                             uint32_t *__restrict__ rank) {
    uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < nhot) rank[rowsByCount[r]] = kHotBit | r;
}

__global__ void k_encode(const uint32_t *__restrict__ idx, const uint32_t *__restrict__ rank, uint64_t n,
                         uint32_t *__restrict__ enc) {
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t v = idx[i];
    uint32_t r = rank[v];
    enc[i] = r != 0xffffffffu ? r : v;
}

}  // namespace

void build_hot_table(const Csf &c, int level, uint64_t rows, HotTable &h, hipStream_t s) {
    const uint32_t *idx = level == c.ndims - 1 ? c.leaf.p : c.idx[level].p;
    const uint64_t n = level == c.ndims - 1 ? c.nnz : c.cnt[level];
    Temp tmp;
    h.count = 0;
    h.tgt.alloc(n);
    if (n == 0 || rows == 0) return;
    DevArray<uint32_t> cnt(rows), cnt2(rows), ids(rows), ids2(rows), rank(rows);
    MMCSF_HIP(hipMemsetAsync(cnt.p, 0, rows * 4, s));
    hipLaunchKernelGGL(k_hist, dim3(nblk(n)), dim3(kBS), 0, s, idx, n, cnt.p);
    hipLaunchKernelGGL(k_iota, dim3(nblk(rows)), dim3(kBS), 0, s, ids.p, rows);
    size_t bytes = 0;
    MMCSF_HIP(rocprim::radix_sort_pairs_desc(nullptr, bytes, cnt.p, cnt2.p, ids.p, ids2.p, (size_t)rows, 0u,
                                             32u, s));
    MMCSF_HIP(rocprim::radix_sort_pairs_desc(tmp.get(bytes), bytes, cnt.p, cnt2.p, ids.p, ids2.p,
                                             (size_t)rows, 0u, 32u, s));
    // rows with at least one use, capped at kMaxHot
    std::vector<uint32_t> top(std::min<uint64_t>(rows, kMaxHot));
    MMCSF_HIP(hipMemcpyAsync(top.data(), cnt2.p, top.size() * 4, hipMemcpyDeviceToHost, s));
    MMCSF_HIP(hipStreamSynchronize(s));
    uint32_t nhot = 0;
    while (nhot < top.size() && top[nhot] > 0) ++nhot;
    h.count = nhot;
    h.rows.alloc(std::max<uint32_t>(nhot, 1));
    if (nhot) MMCSF_HIP(hipMemcpyAsync(h.rows.p, ids2.p, nhot * 4, hipMemcpyDeviceToDevice, s));
    MMCSF_HIP(hipMemsetAsync(rank.p, 0xff, rows * 4, s));
    if (nhot) hipLaunchKernelGGL(k_rank_table, dim3((nhot + 255) / 256), dim3(256), 0, s, ids2.p, nhot, rank.p);
    hipLaunchKernelGGL(k_encode, dim3(nblk(n)), dim3(kBS), 0, s, idx, rank.p, n, h.tgt.p);
    MMCSF_HIP(hipStreamSynchronize(s));
}

// Rows of tensor mode m that hold `frac` of the nonzeros, the most used rows
// first: the working set of gathers from that mode's factor table (small for
// skewed modes, most of the rows for uniform ones).
uint64_t hot_footprint(const Tensor &t, int m, double frac, hipStream_t s) {
    const uint64_t rows = t.dims[m], n = t.nnz;
    if (n == 0 || rows == 0) return 0;
    DevArray<uint32_t> cnt(rows);
    MMCSF_HIP(hipMemsetAsync(cnt.p, 0, rows * 4, s));
    hipLaunchKernelGGL(k_hist, dim3(nblk(n)), dim3(kBS), 0, s, t.inds[m].p, n, cnt.p);
    MMCSF_HIP(hipGetLastError());
    std::vector<uint32_t> h(rows);
    MMCSF_HIP(hipMemcpyAsync(h.data(), cnt.p, rows * 4, hipMemcpyDeviceToHost, s));
    MMCSF_HIP(hipStreamSynchronize(s));
    std::sort(h.begin(), h.end(), [](uint32_t a, uint32_t b) { return a > b; });
    const double want = frac * (double)n;
    double acc = 0;
    uint64_t k = 0;
    while (k < rows && acc < want) acc += h[k++];
    return k;
}


// ---------------------------------------------------------------------------
// MM-CSF mixed-mode partition support (mm_partition_reuseBased,
// util.h:2167-2512): the level-1 entry of every nonzero in a mode's CSF and
// the nonzeros under each entry (nnzPerFiber)
// ---------------------------------------------------------------------------
// first nonzero of level-1 entry j (j = cnt1: nnz), and a mark at it
__global__ void k_l1_start(int N, const uint32_t *__restrict__ p1, const uint32_t *__restrict__ p2, uint64_t cnt1,
                           uint64_t nnz, uint32_t *start, uint32_t *mark) {
    uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (j > cnt1) return;
    if (j == cnt1) {
        start[j] = (uint32_t)nnz;
        return;
    }
    const uint32_t x = N == 3 ? p1[j] : p2[p1[j]];
    start[j] = x;
    mark[x] = 1u;
}

__global__ void k_l1_count(const uint32_t *__restrict__ start, uint64_t cnt1, uint32_t *cnt) {
    uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (j < cnt1) cnt[j] = start[j + 1] - start[j];
}

// entry of sorted nonzero k = (marks before k) + mark[k] - 1, written at its
// tensor position; nonzeros of the last slice (from lastStart on) outside
// its first entry e0 resolve to e0 - 1 (kNoEntry when e0 = 0)
__global__ void k_l1_ids(const uint32_t *__restrict__ excl, const uint32_t *__restrict__ mark,
                         const uint32_t *__restrict__ pos, uint64_t nnz, uint32_t lastStart, uint32_t e0,
                         uint32_t *idT) {
    uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (k >= nnz) return;
    uint32_t id = excl[k] + mark[k] - 1u;
    if (k >= lastStart && id != e0) id = e0 == 0 ? 0xFFFFFFFFu : e0 - 1u;
    idT[pos[k]] = id;
}

void level1_ids(const Csf &c, const uint32_t *pos, DevArray<uint32_t> &idT, DevArray<uint32_t> &cnt, hipStream_t s) {
    const int N = c.ndims;
    const uint64_t n = c.nnz, c1 = c.cnt[1];
    idT.alloc(n);
    cnt.alloc(c1 + 1);
    if (n == 0) return;
    Temp tmp;
    DevArray<uint32_t> start(c1 + 1), mark(n), excl(n);
    hipLaunchKernelGGL(k_zero_words, dim3(nblk(n)), dim3(kBS), 0, s, mark.p, n);
    hipLaunchKernelGGL(k_l1_start, dim3(nblk(c1 + 1)), dim3(kBS), 0, s, N, c.ptr[1].p,
                       N == 4 ? c.ptr[2].p : (const uint32_t *)nullptr, c1, n, start.p, mark.p);
    hipLaunchKernelGGL(k_zero_words, dim3(1), dim3(kBS), 0, s, cnt.p + c1, (uint64_t)1);
    hipLaunchKernelGGL(k_l1_count, dim3(nblk(c1)), dim3(kBS), 0, s, start.p, c1, cnt.p);
    exclusive_sum(mark.p, excl.p, n, s, tmp);
    // the last slice's first level-1 entry and its first nonzero
    uint32_t e0 = 0, lastStart = 0;
    MMCSF_HIP(hipMemcpyAsync(&e0, c.ptr[0].p + (c.cnt[0] - 1), 4, hipMemcpyDeviceToHost, s));
    MMCSF_HIP(hipStreamSynchronize(s));
    MMCSF_HIP(hipMemcpyAsync(&lastStart, start.p + e0, 4, hipMemcpyDeviceToHost, s));
    MMCSF_HIP(hipStreamSynchronize(s));
    hipLaunchKernelGGL(k_l1_ids, dim3(nblk(n)), dim3(kBS), 0, s, excl.p, mark.p, pos, n, lastStart, e0, idT.p);
    MMCSF_HIP(hipGetLastError());
    MMCSF_HIP(hipStreamSynchronize(s));
}

__global__ void k_gather_sel_u32(const uint32_t *__restrict__ src, const uint32_t *__restrict__ sel, uint64_t n,
                                 uint32_t *dst) {
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[sel[i]];
}
__global__ void k_gather_sel_f32(const float *__restrict__ src, const uint32_t *__restrict__ sel, uint64_t n,
                                 float *dst) {
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[sel[i]];
}

// the nonzeros at tensor positions sel[0..n) as a tensor of their own
void gather_tensor(const Tensor &t, const uint32_t *sel, uint64_t n, Tensor &out, hipStream_t s) {
    out.ndims = t.ndims;
    out.nnz = n;
    for (int m = 0; m < t.ndims; ++m) {
        out.dims[m] = t.dims[m];
        out.inds[m].alloc(n);
        if (n) hipLaunchKernelGGL(k_gather_sel_u32, dim3(nblk(n)), dim3(kBS), 0, s, t.inds[m].p, sel, n, out.inds[m].p);
    }
    out.vals.alloc(n);
    if (n) hipLaunchKernelGGL(k_gather_sel_f32, dim3(nblk(n)), dim3(kBS), 0, s, t.vals.p, sel, n, out.vals.p);
    MMCSF_HIP(hipGetLastError());
}

}  // namespace mmcsf
