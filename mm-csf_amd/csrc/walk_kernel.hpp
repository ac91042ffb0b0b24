// The chunked root-walk kernel template, its launch-shape choice and the
// dispatch over (ndims, factor type, rank); included by one translation unit
// per operation.
#pragma once
#include <algorithm>
#include <type_traits>

#include "walk.hpp"

namespace mmcsf {
namespace {

// Largest i < n with a[i] <= key (a non-decreasing, a[0] <= key), searched by
// the whole wave: each round probes 64 evenly spaced entries and keeps the
// slot of the last probe <= key, so 2^26 entries take 5 rounds of one load.
// Every lane must call it with the same arguments.
__device__ __forceinline__ uint32_t wave_upper(const uint32_t *__restrict__ a, uint32_t n, uint32_t key) {
    const uint32_t lane = threadIdx.x & 63;
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        const uint32_t step = (hi - lo + 63) / 64;
        const uint32_t probe = lo + lane * step;
        const bool le = lane == 0 || (probe < hi && a[probe] <= key);
        const uint64_t m = __ballot(le);
        const uint32_t last = 63u - (uint32_t)__builtin_clzll(m);
        lo += last * step;
        hi = std::min(hi, lo + step);
    }
    return lo;
}

// first i in [0, n] with a[i] >= key (a non-decreasing over n + 1 entries)
__device__ __forceinline__ uint32_t wave_lower(const uint32_t *__restrict__ a, uint32_t n, uint32_t key) {
    const uint32_t lane = threadIdx.x & 63;
    uint32_t lo = 0, hi = n + 1;  // answer in [lo, hi)
    while (hi - lo > 1) {
        // probes at lo + 1 + lane * step: entries >= key bound the answer above
        const uint32_t step = (hi - lo - 1 + 63) / 64;
        const uint32_t probe = lo + 1 + lane * step;
        const bool lt = probe < hi && a[probe - 1] < key;  // answer > probe - 1
        const uint64_t m = __ballot(lt);
        if (m == 0) {
            hi = lo + 1;
        } else {
            const uint32_t last = 63u - (uint32_t)__builtin_clzll(m);
            const uint32_t nlo = lo + 1 + last * step;
            hi = std::min(hi, nlo + step);
            lo = nlo;
        }
    }
    return lo;
}

// VEC consecutive elements of one row per lane-load (at a byte offset from
// the table base: tables are below 4 GiB), converted to float
template <typename T, int VEC>
struct Raw;
template <>
struct Raw<float, 4> {
    float4 q;
    __device__ __forceinline__ void load(const float *U, uint32_t off) {
        q = *reinterpret_cast<const float4 *>(reinterpret_cast<const char *>(U) + off);
    }
    __device__ __forceinline__ void get(float *o) const { o[0] = q.x; o[1] = q.y; o[2] = q.z; o[3] = q.w; }
};
template <>
struct Raw<uint16_t, 8> {
    uint4 q;
    __device__ __forceinline__ void load(const uint16_t *U, uint32_t off) {
        q = *reinterpret_cast<const uint4 *>(reinterpret_cast<const char *>(U) + off);
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
    __device__ __forceinline__ void load(const uint16_t *U, uint32_t off) {
        q = *reinterpret_cast<const uint2 *>(reinterpret_cast<const char *>(U) + off);
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

template <int N, int G, int NB, int TB, typename T, int VEC, int OP>
__global__ __launch_bounds__(256, (OP == kMttkrp && N == 3 && NB == 1 && G * VEC <= 32) ? 3 : 2) void k_walk(WalkArgs a) {
    static_assert(G == 1 || G == 2 || G == 4 || G == 8, "row groups of 1..8 lanes");
    static_assert(G % TB == 0, "steps per batch divide the steps of a chunk");
    static_assert(VEC % 4 == 0, "rows are written as float4");
    constexpr bool kModel = OP == kLossGrad || OP == kLossResid;
    constexpr bool kGrad = OP != kLossResid;  // writes rows of `out`
    // multiplies the rows of the upper levels; rows of `out` are keyed by
    // slice (by leaf fiber for TTM, which has no upper levels)
    constexpr bool kUpper = OP != kTtm;
    // residuals / weights are addressed by tensor-order position
    constexpr bool kPos = OP == kLossResid || OP == kWeightGrad;
    constexpr int GPR = 64 / G;  // rows per load instruction
    constexpr int STEPS = G;     // load instructions per chunk and column block
    constexpr int BW = G * VEC;  // columns per block
    constexpr int LF = N - 2;
    constexpr uint32_t kNone = 0xffffffffu;
    const int lane = threadIdx.x & 63;
    const int g = lane / G, k = lane % G;
    const uint64_t wavesPerBlock = blockDim.x >> 6;
    const uint64_t waves = (uint64_t)gridDim.x * wavesPerBlock;
    // Workgroups are dispatched round-robin over the 8 XCDs (block b on XCD
    // b % 8).  Each XCD takes 16 spans of consecutive blocks spread over the
    // tree, so neighbouring waves (which share slices and fiber rows) sit on
    // one L2, while every XCD still sees all parts of the tree.
    constexpr uint32_t kSpans = 16;
    const uint32_t spanBlocks = gridDim.x / (8u * kSpans);
    const uint32_t lb = gridDim.x % (8u * kSpans)
                            ? blockIdx.x
                            : ((blockIdx.x & 7u) + 8u * ((blockIdx.x >> 3) / spanBlocks)) * spanBlocks +
                                  (blockIdx.x >> 3) % spanBlocks;
    const uint64_t wv = (uint64_t)lb * wavesPerBlock + (threadIdx.x >> 6);
    const uint32_t chunks = (uint32_t)(((uint64_t)a.nnz + 63) / 64);
    // this wave's chunks: an equal share of the total traversal cost
    const uint64_t total = a.cost[chunks];
    const uint32_t cb = wave_lower(a.cost, chunks, (uint32_t)(wv * total / waves));
    const uint32_t ce = wv + 1 == waves ? chunks : wave_lower(a.cost, chunks, (uint32_t)((wv + 1) * total / waves));
    if (cb >= ce) return;  // the whole wave leaves together
    const int R = a.R;
    const uint32_t rowBytes = (uint32_t)R * (uint32_t)sizeof(T);
    // byte offset of this lane's columns in block b; blocks past R read columns 0..
    uint32_t colBytes[NB];
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        const int col = b * BW + k * VEC;
        colBytes[b] = col < R ? (uint32_t)col * (uint32_t)sizeof(T) : 0u;
    }
    const T *U0 = static_cast<const T *>(a.U[0]);
    const T *U1 = static_cast<const T *>(a.U[1]);
    const T *U2 = static_cast<const T *>(a.U[2]);
    const T *UL = static_cast<const T *>(a.U[N - 1]);
    const uint32_t *__restrict__ ptrL = a.ptr[LF];

    // leaf fiber (and 4-D level-1 entry) of the nonzero before the run; the
    // unsigned wrap at x0 = 0 stands for -1
    const uint32_t x0 = (uint32_t)(cb * 64);
    const uint32_t f0 = wave_upper(ptrL, a.cnt[LF], x0);  // leaf fiber of x0
    const uint32_t fcur = x0 == 0 ? kNone : (a.ptr[LF][f0] == x0 ? f0 - 1 : f0);
    uint32_t jcur = kNone, j0 = f0;
    if constexpr (N == 4) {
        j0 = wave_upper(a.ptr[1], a.cnt[1], f0);
        jcur = fcur == kNone ? kNone : (a.ptr[1][j0] == f0 && fcur != f0 ? j0 - 1 : j0);
    }
    // the open slice (leaf fiber for TTM): its partial row is spread over acc of all groups
    uint32_t scur = OP == kTtm ? f0 : a.slc[j0];
    // whether the open slice began inside this wave's run: then every
    // contribution to it passes through this wave, and its row is stored
    // instead of added (the output starts zeroed)
    bool inside = false;

    float acc[NB][VEC];
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
        for (int e = 0; e < VEC; ++e) acc[b][e] = 0.f;
    double sq = 0.0;

    // row writers: VEC columns of block b from this lane
    auto put = [&](int b, uint32_t row, const float *v, bool store) {
        const int col = b * BW + k * VEC;
        if (row == kNone || col >= R) return;
        float *o = a.out + (size_t)row * R + col;
        if (store) {
#pragma unroll
            for (int e = 0; e < VEC; e += 4) *reinterpret_cast<float4 *>(o + e) = make_float4(v[e], v[e + 1], v[e + 2], v[e + 3]);
        } else {
#pragma unroll
            for (int e = 0; e < VEC; ++e) unsafeAtomicAdd(o + e, v[e]);
        }
    };
    // sum of acc over the groups, left in every lane
    auto reduce = [&](int b) {
#pragma unroll
        for (int off = G; off < 64; off <<= 1)
#pragma unroll
            for (int e = 0; e < VEC; ++e) acc[b][e] += __shfl_xor(acc[b][e], off, 64);
    };
    // write the open partial row of block b to slice s and clear it (all lanes)
    auto flush = [&](int b, uint32_t s, bool store) {
        reduce(b);
        if (g == 0) put(b, s, acc[b], store);
#pragma unroll
        for (int e = 0; e < VEC; ++e) acc[b][e] = 0.f;
    };

    // Software pipeline over the chunks of the run.  A chunk's entries come
    // from three dependent reads: its words (leaf index, value, fiber-start
    // bits), then the level-1 entries those bits select (fiber index, slice),
    // then the factor rows.  The level-1 entries of chunk c+1 and the words of
    // chunk c+2 are read while the rows of chunk c are in flight, so a chunk
    // waits for one read latency instead of three.  Every read in the loop is
    // unconditional (indices clamped into their arrays, validity applied
    // where the values are used): the in-order load counter then lets each
    // wait cover exactly the reads it needs.
    struct Words {
        uint32_t lw, word, ps;
        float v;
    };
    struct Entries {
        uint32_t s, i1, i2;
    };
    const uint32_t lastX = a.nnz - 1;  // nnz >= 1: the run is not empty
    auto readWords = [&](uint32_t c, Words &w) {
        const uint32_t x = std::min(c * 64 + (uint32_t)lane, lastX);
        w.lw = a.leaf[x];
        w.word = a.fbit[x >> 5];
        w.ps = 0;
        if constexpr (kPos) w.ps = a.pos[x];
        w.v = 0.f;
        if constexpr (OP != kWeightGrad) w.v = a.val[x];
    };
    // level-1 entries of chunk c; fB / jB: the leaf fiber and level-1 entry
    // of the nonzero before chunk c, advanced past it.  Chunks past the run
    // (or past the tensor) read clamped entries that are never used.
    uint32_t fB = fcur, jB = jcur;
    [[maybe_unused]] const uint32_t lastF = a.cnt[LF] - 1, lastJ = a.cnt[1] - 1;
    auto readEntries = [&](uint32_t c, Words &w, Entries &q) {
        const uint32_t x = c * 64 + (uint32_t)lane;
        const bool start = x < a.nnz && ((w.word >> (x & 31)) & 1u);
        const uint64_t fm = __ballot(start);
        const uint32_t f = fB + (uint32_t)__popcll(fm & low_bits((uint32_t)lane + 1));
        const uint32_t fnext = fB + (uint32_t)__popcll(fm);
        uint32_t j = f, jnext = fnext;
        if constexpr (N == 4) {
            const uint64_t win = window64(a.sbit, fB + 1);
            j = jB + (uint32_t)__popcll(win & low_bits(f - fB));
            jnext = jB + (uint32_t)__popcll(win & low_bits(fnext - fB));
        }
        q.i1 = q.i2 = 0;
        if constexpr (kUpper) {
            const uint32_t jc = std::min(j, lastJ);
            q.i1 = a.idx[1][jc];
            q.s = a.slc[jc];
            if constexpr (N == 4) q.i2 = a.idx[2][std::min(f, lastF)];
        } else {
            q.s = f;
        }
        if constexpr (OP == kWeightGrad) w.v = a.wts[w.ps];
        fB = fnext;
        jB = jnext;
    };
    Words w0, w1, w2;
    Entries e0, e1, e2;
    readWords(cb, w0);
    readEntries(cb, w0, e0);
    readWords(cb + 1, w1);

    // one chunk: wc / ec hold chunk c; the entries of chunk c+1 (words wn)
    // go to en, the words of chunk c+2 to wf
    auto step = [&](uint32_t c, Words &wc, Entries &ec, Words &wn, Entries &en,
                    Words &wf) __attribute__((always_inline)) {
        const uint32_t x = c * 64 + (uint32_t)lane;
        const bool valid = x < a.nnz;
        // lanes past the tensor: zero value, no slice; their (clamped) row
        // indices stay valid
        const uint32_t lw = wc.lw, ps = wc.ps, i1 = ec.i1, i2 = ec.i2;
        const uint32_t s = valid ? ec.s : kNone;
        const float v = valid ? wc.v : 0.f;
        uint32_t sprev = __shfl_up(s, 1, 64);
        if (lane == 0) sprev = scur;
        const uint64_t bm = kGrad ? __ballot(s != sprev) : 0ull;  // slice starts in the chunk
        const uint32_t nvalid = std::min<uint32_t>(64, a.nnz - c * 64);
        // Chunks with several slice starts (short slices) give every group a
        // contiguous run of entries instead of every GPR-th entry; slices are
        // then closed group by group and joined across groups by a segmented
        // scan (below) instead of a cross-group reduction per slice.
        const bool dense = kGrad && __popcll(bm) >= 3;

        // the entries of this lane's group, one per step
        // (byte offsets of their rows, computed once by the entry's own lane)
        uint32_t li[STEPS], l1[STEPS], l2[STEPS], l0[STEPS], pp[STEPS];
        // entry of step t for this lane's group
        const uint32_t peB = dense ? (uint32_t)(g * STEPS) : (uint32_t)g;
        const uint32_t peS = dense ? 1u : (uint32_t)GPR;
        auto pe = [&](int t) -> uint32_t { return peB + (uint32_t)t * peS; };
        // slice and value of step t, taken from their lanes where used
        auto svAt = [&](int t) -> uint32_t { return __shfl(s, (int)pe(t), 64); };
        auto vvAt = [&](int t) -> float { return __shfl(v, (int)pe(t), 64); };
#pragma unroll
        for (int t = 0; t < STEPS; ++t) {
            const int p = (int)pe(t);
            li[t] = __shfl(lw * rowBytes, p, 64);
            if constexpr (kUpper) l1[t] = __shfl(i1 * rowBytes, p, 64);
            if constexpr (N == 4 && kUpper) l2[t] = __shfl(i2 * rowBytes, p, 64);
            if constexpr (kModel) l0[t] = __shfl((valid ? s : 0u) * rowBytes, p, 64);
            if constexpr (OP == kLossResid) pp[t] = __shfl(ps, p, 64);
        }

        // dense chunks: per group, the partial of the slice open at the run
        // start (hd, until the first slice start inside the run) and of the
        // current slice (la)
        float hd[NB][VEC], la[NB][VEC];
        bool seen = false;  // a slice started inside this group's run
        uint32_t ls = kNone;
        if constexpr (kGrad) {
            if (dense) {
                // acc becomes the carry: the partial row of scur from earlier
                // chunks, summed over the groups (left in every lane)
#pragma unroll
                for (int b = 0; b < NB; ++b) {
                    reduce(b);
#pragma unroll
                    for (int e = 0; e < VEC; ++e) hd[b][e] = la[b][e] = 0.f;
                }
                ls = svAt(0);
            }
        }
        uint32_t sc0 = scur;
#pragma unroll
        for (int t0 = 0; t0 < STEPS; t0 += TB) {
            Raw<T, VEC> rw[TB][NB], w1[TB][NB], w2[TB][NB], w0[TB][NB];
#pragma unroll
            for (int t = 0; t < TB; ++t)
#pragma unroll
                for (int b = 0; b < NB; ++b) {
                    // blocks past R read columns 0.. and are dropped where used
                    const uint32_t cb4 = colBytes[b];
                    rw[t][b].load(UL, li[t0 + t] + cb4);
                    if constexpr (kUpper) w1[t][b].load(U1, l1[t0 + t] + cb4);
                    if constexpr (N == 4 && kUpper) w2[t][b].load(U2, l2[t0 + t] + cb4);
                    if constexpr (kModel) w0[t][b].load(U0, l0[t0 + t] + cb4);
                }
            if (t0 == 0) {
                // issued after this batch's rows and before their use, so the
                // rows can be waited for without waiting for these reads
                __builtin_amdgcn_sched_barrier(0);
                readEntries(c + 1, wn, en);
                readWords(c + 2, wf);
                __builtin_amdgcn_sched_barrier(0);
            }
            // coefficient of every entry of the batch
            float co[TB];
#pragma unroll
            for (int t = 0; t < TB; ++t) {
                const uint32_t p = pe(t0 + t);
                if constexpr (kModel) {
                    float d = 0.f;
#pragma unroll
                    for (int b = 0; b < NB; ++b) {
                        const int col = b * BW + k * VEC;
                        if (col < R) {
                            float r[VEC], q[VEC], z[VEC];
                            rw[t][b].get(r);
                            w1[t][b].get(q);
                            w0[t][b].get(z);
#pragma unroll
                            for (int e = 0; e < VEC; ++e) r[e] *= q[e] * z[e];
                            if constexpr (N == 4) {
                                w2[t][b].get(q);
#pragma unroll
                                for (int e = 0; e < VEC; ++e) r[e] *= q[e];
                            }
#pragma unroll
                            for (int e = 0; e < VEC; ++e) d += r[e];
                        }
                    }
#pragma unroll
                    for (int off = 1; off < G; off <<= 1) d += __shfl_xor(d, off, 64);
                    const bool ok = p < nvalid;
                    const float vt = vvAt(t0 + t);  // every lane takes part
                    const float e = ok ? d - vt : 0.f;
                    if (k == 0 && ok) {
                        sq += (double)e * (double)e;
                        if constexpr (OP == kLossResid)
                            if (a.resid) a.resid[pp[t0 + t]] = e;
                    }
                    co[t] = e * a.alpha;
                } else if constexpr (OP == kWeightGrad) {
                    co[t] = vvAt(t0 + t) * a.alpha;
                } else {
                    co[t] = vvAt(t0 + t);
                }
            }
            if constexpr (kGrad) {
                const uint32_t plo = (uint32_t)(t0 * GPR), phi = (uint32_t)((t0 + TB) * GPR);
                const uint64_t bmb = dense ? 0ull : bm & low_bits(phi) & ~low_bits(plo);
                const uint32_t ls0 = ls;
                const bool seen0 = seen;
#pragma unroll
                for (int b = 0; b < NB; ++b) {
                    // product of the entry's rows; the coefficient co is
                    // applied by the fused multiply-add that sums it.  Lanes
                    // of columns past R compute on column-0 data and are
                    // never written.
                    float cf[TB][VEC];
#pragma unroll
                    for (int t = 0; t < TB; ++t) {
                        rw[t][b].get(cf[t]);
                        if constexpr (kUpper) {
                            float q[VEC];
                            w1[t][b].get(q);
#pragma unroll
                            for (int e = 0; e < VEC; ++e) cf[t][e] *= q[e];
                            if constexpr (N == 4) {
                                w2[t][b].get(q);
#pragma unroll
                                for (int e = 0; e < VEC; ++e) cf[t][e] *= q[e];
                            }
                        }
                    }
                    if (dense) {
                        uint32_t lsb = ls0;
                        bool sb = seen0;
#pragma unroll
                        for (int t = 0; t < TB; ++t) {
                            const uint32_t sp = svAt(t0 + t);
                            if (sp != lsb) {
                                if (!sb) {
#pragma unroll
                                    for (int e = 0; e < VEC; ++e) hd[b][e] = la[b][e];
                                } else {
                                    put(b, lsb, la[b], true);  // began and ended in this run
                                }
#pragma unroll
                                for (int e = 0; e < VEC; ++e) la[b][e] = 0.f;
                                lsb = sp;
                                sb = true;
                            }
#pragma unroll
                            for (int e = 0; e < VEC; ++e) la[b][e] = fmaf(co[t], cf[t][e], la[b][e]);
                        }
                        ls = lsb;
                        seen = sb;
                    } else if (bmb == 0) {
#pragma unroll
                        for (int t = 0; t < TB; ++t)
#pragma unroll
                            for (int e = 0; e < VEC; ++e) acc[b][e] = fmaf(co[t], cf[t][e], acc[b][e]);
                    } else {
                        // slice segments of the batch in order; wave-uniform loop
                        uint64_t rest = bmb;
                        uint32_t lo = plo, sc = sc0;
                        bool in = inside;
                        while (rest) {
                            const uint32_t hi = (uint32_t)__builtin_ctzll(rest);
                            rest &= rest - 1;
#pragma unroll
                            for (int t = 0; t < TB; ++t) {
                                const uint32_t p = pe(t0 + t);
                                if (p >= lo && p < hi)
#pragma unroll
                                    for (int e = 0; e < VEC; ++e) acc[b][e] = fmaf(co[t], cf[t][e], acc[b][e]);
                            }
                            flush(b, sc, in);
                            sc = __shfl(s, (int)hi, 64);
                            lo = hi;
                            in = true;
                        }
#pragma unroll
                        for (int t = 0; t < TB; ++t) {
                            const uint32_t p = pe(t0 + t);
                            if (p >= lo)
#pragma unroll
                                for (int e = 0; e < VEC; ++e) acc[b][e] = fmaf(co[t], cf[t][e], acc[b][e]);
                        }
                    }
                }
                if (bmb) {
                    sc0 = __shfl(s, 63 - (int)__builtin_clzll(bmb), 64);
                    inside = true;
                }
            }
        }
        if constexpr (kGrad) {
            if (dense) {
                // Join the slices across the groups.  Group g links to group
                // g-1 when no slice starts at its first entry; a group without
                // a slice start inside its run passes the chain on.  The
                // segmented inclusive scan over groups gives every group the
                // total of the slice open at its run end (op); rt: that chain
                // runs back to the chunk start (and then carries the partial
                // row of scur from earlier chunks).
                const uint32_t gs = (uint32_t)(g * STEPS);
                const bool cont = !((bm >> gs) & 1ull);  // no slice starts at the run start
                const bool single = !seen;
                bool rt = single && cont;
                const uint32_t sLast = ls, sFirst = svAt(0);
                float op[NB][VEC];
#pragma unroll
                for (int b = 0; b < NB; ++b)
#pragma unroll
                    for (int e = 0; e < VEC; ++e) op[b][e] = la[b][e];
#pragma unroll
                for (int off = G; off < 64; off <<= 1) {
                    const bool rto = __shfl_up(rt ? 1 : 0, off, 64) != 0;
                    const bool has = lane >= off;
#pragma unroll
                    for (int b = 0; b < NB; ++b)
#pragma unroll
                        for (int e = 0; e < VEC; ++e) {
                            const float o = __shfl_up(op[b][e], off, 64);
                            if (rt && has) op[b][e] += o;
                        }
                    if (has) rt = rt && rto;
                }
                // rt now: the open chain of this group reaches the chunk start
#pragma unroll
                for (int b = 0; b < NB; ++b)
#pragma unroll
                    for (int e = 0; e < VEC; ++e)
                        if (rt) op[b][e] += acc[b][e];
                // a slice starting at the chunk start closes the slice left open
                // by the previous chunk
                if (g == 0 && !cont) {
#pragma unroll
                    for (int b = 0; b < NB; ++b) put(b, scur, acc[b], inside);
                }
                // a group with a slice start inside its run closes the slice
                // open at its run start: its head partial plus the chain ending
                // at the previous group's run end (the carry for group 0)
                const int rtUp = __shfl_up(rt ? 1 : 0, G, 64);  // every lane takes part
                const bool back = cont && (g == 0 || rtUp != 0);  // the slice is scur
#pragma unroll
                for (int b = 0; b < NB; ++b) {
                    float h[VEC];
#pragma unroll
                    for (int e = 0; e < VEC; ++e) {
                        const float o = __shfl_up(op[b][e], G, 64);
                        h[e] = hd[b][e] + (cont ? (g == 0 ? acc[b][e] : o) : 0.f);
                    }
                    if (!single) put(b, sFirst, h, !back || inside);
                }
                // the slice open at a run end closes there when the next run
                // starts a slice; the last group's stays open
                const bool nextStarts = g + 1 < GPR && ((bm >> (gs + STEPS)) & 1ull);
                if (nextStarts) {
#pragma unroll
                    for (int b = 0; b < NB; ++b) put(b, sLast, op[b], !rt || inside);
                }
                const bool rtLast = __shfl(rt ? 1 : 0, 63, 64) != 0;
#pragma unroll
                for (int b = 0; b < NB; ++b)
#pragma unroll
                    for (int e = 0; e < VEC; ++e) acc[b][e] = g == GPR - 1 ? op[b][e] : 0.f;
                inside = rtLast ? inside : true;
                sc0 = __shfl(s, 63, 64);
            }
        }
        scur = sc0;
    };
    // unrolled by the pipeline depth: the buffers rotate by name, so no
    // register with a read in flight is ever copied (a copy would wait for it)
    for (uint32_t c = cb; c < ce; c += 3) {
        step(c, w0, e0, w1, e1, w2);
        if (c + 1 >= ce) break;
        step(c + 1, w1, e1, w2, e2, w0);
        if (c + 2 >= ce) break;
        step(c + 2, w2, e2, w0, e0, w1);
    }
    if constexpr (kGrad) {
        // the last slice may continue in the next wave
#pragma unroll
        for (int b = 0; b < NB; ++b) flush(b, scur, false);
    }
    if constexpr (kModel) {
        if (a.sumsq) {
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) sq += __shfl_down(sq, off, 64);
            if (lane == 0) unsafeAtomicAdd(a.sumsq, sq);
        }
    }
}

struct Cfg {
    int G, NB, TB, VEC;
};

// row groups of G lanes x VEC columns per load; NB column blocks; TB steps
// of loads in flight (at most 32 row elements per lane and row kind)
Cfg pick(int R, int dtype) {
    const int vec = (dtype == MMCSF_BF16 && R % 8 == 0) ? 8 : 4;
    Cfg c{8, 1, 8, vec};
    if (R <= vec) c = {1, 1, 1, vec};
    else if (R <= 2 * vec) c = {2, 1, 2, vec};
    else if (R <= 4 * vec) c = {4, 1, 4, vec};
    else {
        const int nb = (R + 8 * vec - 1) / (8 * vec);
        c.NB = nb <= 1 ? 1 : nb <= 2 ? 2 : nb <= 4 ? 4 : 8;
    }
    c.TB = std::max(1, std::min(c.G, 32 / (c.NB * vec)));
    return c;
}

template <int N, int G, int NB, int TB, typename T, int VEC, int OP>
void go(const WalkArgs &a, hipStream_t s) {
    auto fn = k_walk<N, G, NB, TB, T, VEC, OP>;
    int perCU = 0;
    MMCSF_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, fn, 256, 0));
    perCU = std::max(perCU, 1);
    const uint64_t chunks = ((uint64_t)a.nnz + 63) / 64;
    uint64_t blocks = (uint64_t)device_cu_count() * perCU;
    blocks = std::min(blocks, std::max<uint64_t>(1, (chunks + 3) / 4));
    // a wave start is found by a search over the cost prefix; more waves
    // than chunks would give empty waves only
    fn<<<dim3((unsigned)blocks), dim3(256), 0, s>>>(a);
    MMCSF_HIP(hipGetLastError());
}

template <int N, typename T, int VEC, int OP>
void dispatch(const Cfg &c, const WalkArgs &a, hipStream_t s) {
#define MMCSF_WALK(GG, BB, TT)                       \
    if (c.G == GG && c.NB == BB && c.TB == TT) {     \
        go<N, GG, BB, TT, T, VEC, OP>(a, s);         \
        return;                                      \
    }
    MMCSF_WALK(1, 1, 1)
    MMCSF_WALK(2, 1, 2)
    MMCSF_WALK(4, 1, 4)
    if constexpr (VEC == 4) {
        MMCSF_WALK(8, 1, 8)
        MMCSF_WALK(8, 2, 4)
        MMCSF_WALK(8, 4, 2)
        MMCSF_WALK(8, 8, 1)
    } else {
        MMCSF_WALK(8, 1, 4)
        MMCSF_WALK(8, 2, 2)
        MMCSF_WALK(8, 4, 1)
    }
#undef MMCSF_WALK
    fail(MMCSF_ERR_UNSUPPORTED, "no root-walk kernel for this rank");
}

}  // namespace

template <int OP>
void walk_launch(const Csf &c, const Factors &F, const WalkArgs &a, hipStream_t s) {
    const Cfg cfg = pick(F.R, F.dtype);
    if (c.ndims == 3) {
        if (F.dtype == MMCSF_F32) dispatch<3, float, 4, OP>(cfg, a, s);
        else if (cfg.VEC == 8) dispatch<3, uint16_t, 8, OP>(cfg, a, s);
        else dispatch<3, uint16_t, 4, OP>(cfg, a, s);
    } else {
        if (F.dtype == MMCSF_F32) dispatch<4, float, 4, OP>(cfg, a, s);
        else if (cfg.VEC == 8) dispatch<4, uint16_t, 8, OP>(cfg, a, s);
        else dispatch<4, uint16_t, 4, OP>(cfg, a, s);
    }
}

}  // namespace mmcsf
