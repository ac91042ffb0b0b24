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

// Buffer resource of one factor table (gfx9 raw buffer: stride 0, byte
// range checked): lanes whose offset is at or past `bytes` read zeros and
// access no memory.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t table_rsrc(const void *p, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes, 0x00020000);
}

// VEC consecutive elements of one row per lane-load (at a byte offset from
// the table base: tables are below 4 GiB), converted to float.  load() is a
// plain global load; loadb() a buffer load whose out-of-range lanes read
// zeros without touching memory.
template <typename T, int VEC>
struct Raw;
template <>
struct Raw<float, 4> {
    float4 q;
    __device__ __forceinline__ void load(const float *U, uint32_t off) {
        q = *reinterpret_cast<const float4 *>(reinterpret_cast<const char *>(U) + off);
    }
    __device__ __forceinline__ void loadb(__amdgpu_buffer_rsrc_t r, uint32_t off) {
        q = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
    }
    __device__ __forceinline__ void get(float *o) const { o[0] = q.x; o[1] = q.y; o[2] = q.z; o[3] = q.w; }
    __device__ __forceinline__ void from_lane(const Raw &o, int src) {
        q.x = __shfl(o.q.x, src, 64); q.y = __shfl(o.q.y, src, 64); q.z = __shfl(o.q.z, src, 64); q.w = __shfl(o.q.w, src, 64);
    }
};
template <>
struct Raw<uint16_t, 8> {
    uint4 q;
    __device__ __forceinline__ void load(const uint16_t *U, uint32_t off) {
        q = *reinterpret_cast<const uint4 *>(reinterpret_cast<const char *>(U) + off);
    }
    __device__ __forceinline__ void loadb(__amdgpu_buffer_rsrc_t r, uint32_t off) {
        q = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
    }
    __device__ __forceinline__ void get(float *o) const {
        o[0] = __uint_as_float(q.x << 16); o[1] = __uint_as_float(q.x & 0xffff0000u);
        o[2] = __uint_as_float(q.y << 16); o[3] = __uint_as_float(q.y & 0xffff0000u);
        o[4] = __uint_as_float(q.z << 16); o[5] = __uint_as_float(q.z & 0xffff0000u);
        o[6] = __uint_as_float(q.w << 16); o[7] = __uint_as_float(q.w & 0xffff0000u);
    }
    __device__ __forceinline__ void from_lane(const Raw &o, int src) {
        q.x = __shfl(o.q.x, src, 64); q.y = __shfl(o.q.y, src, 64); q.z = __shfl(o.q.z, src, 64); q.w = __shfl(o.q.w, src, 64);
    }
};
template <>
struct Raw<uint16_t, 4> {
    uint2 q;
    __device__ __forceinline__ void load(const uint16_t *U, uint32_t off) {
        q = *reinterpret_cast<const uint2 *>(reinterpret_cast<const char *>(U) + off);
    }
    __device__ __forceinline__ void loadb(__amdgpu_buffer_rsrc_t r, uint32_t off) {
        q = __builtin_bit_cast(uint2, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
    }
    __device__ __forceinline__ void get(float *o) const {
        o[0] = __uint_as_float(q.x << 16); o[1] = __uint_as_float(q.x & 0xffff0000u);
        o[2] = __uint_as_float(q.y << 16); o[3] = __uint_as_float(q.y & 0xffff0000u);
    }
    __device__ __forceinline__ void from_lane(const Raw &o, int src) {
        q.x = __shfl(o.q.x, src, 64); q.y = __shfl(o.q.y, src, 64);
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

// value of lane t of this lane's group of G lanes (t a constant after
// unrolling): ds_swizzle in bit mode, lane' = (lane & ~(G-1)) | t within
// each half-wave
template <int G>
__device__ __forceinline__ uint32_t grp_lane(uint32_t v, int t) {
    if constexpr (G == 1) {
        return v;
    } else {
        constexpr int am = 0x1F & ~(G - 1);
        switch (t) {
            case 0: return (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, am | (0 << 5));
            case 1: return (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, am | (1 << 5));
            case 2: return (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, am | (2 << 5));
            case 3: return (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, am | (3 << 5));
            case 4: return (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, am | (4 << 5));
            case 5: return (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, am | (5 << 5));
            case 6: return (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, am | (6 << 5));
            default: return (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, am | (7 << 5));
        }
    }
}
__device__ __forceinline__ float as_float(uint32_t u) { return __uint_as_float(u); }

// Offset given to a masked buffer load: past every table (tables hold at most
// kOobOffset bytes, and a column offset below 4 KiB keeps it past them).
constexpr uint32_t kOobOffset = 0xFFFFF000u;

// Fiber-level factoring (MTTKRP): the rows of the levels between the root and
// the leaf are read once per fiber segment of a group run instead of once per
// nonzero.  Off for wide ranks (the per-block partials would not fit the
// registers), for one-lane groups (a run of one nonzero) and for 4-D trees
// with 8-lane groups: on a skewed 4-D tensor (100000 x 50000 x 20000 x 1000,
// 50M nonzeros) the leaf fibers hold 1-1.5 nonzeros, so factoring saves few
// loads while its per-step bookkeeping costs more than at narrower groups
// (all modes at R=64: bfloat16 3.79 ms unfactored against 4.15 factored,
// float32 7.62 against 7.71; at R=16 factoring still wins).  The CP passes
// multiply per nonzero: they also read the root row and reduce a dot product
// per nonzero, and on a 1000000 x 500000 x 100000 tensor with 50M nonzeros
// factoring them gained nothing at R=16 and cost 2-13 % at R=64.
template <int N, int G, int NB, int OP>
constexpr bool kFactorOf = OP == kMttkrp && NB <= 2 && G >= 2 && !(N == 4 && G == 8);

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
    constexpr bool kFactor = kFactorOf<N, G, NB, OP>;
    static_assert(!(kFactor && (OP == kLossGrad || OP == kLossResid)), "the CP passes multiply per nonzero");
    // residuals / weights are addressed by tensor-order position
    constexpr bool kPos = OP == kLossResid || OP == kWeightGrad;
    constexpr int GPR = 64 / G;  // groups (runs) per chunk
    constexpr int STEPS = G;     // nonzeros per run = load instructions per chunk and column block
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
    [[maybe_unused]] const __amdgpu_buffer_rsrc_t rs1 = table_rsrc(a.U[1], a.tabBytes[1]);
    [[maybe_unused]] const __amdgpu_buffer_rsrc_t rs2 = table_rsrc(a.U[2], a.tabBytes[2]);
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
    // the open slice (leaf fiber for TTM): its partial row from earlier
    // chunks is the carry `acc`, summed over the groups' shares
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

    // row writer: VEC columns of block b from this lane
    auto put = [&](int b, uint32_t row, const float *v, bool store) {
        const int col = b * BW + k * VEC;
        if (row == kNone || col >= R) return;
        float *o = a.out + (size_t)row * R + col;
        if (store && !a.accum) {
#pragma unroll
            for (int e = 0; e < VEC; e += 4) *reinterpret_cast<float4 *>(o + e) = make_float4(v[e], v[e + 1], v[e + 2], v[e + 3]);
        } else {
#pragma unroll
            for (int e = 0; e < VEC; ++e) unsafeAtomicAdd(o + e, v[e]);
        }
    };

    // Software pipeline over the chunks of the run.  A chunk's entries come
    // from three dependent reads: its words (leaf index, value, fiber-start
    // bits), then the level-1 entries those bits select (fiber index, slice),
    // then the factor rows.  The level-1 entries of chunk c+1 and the words of
    // chunk c+2 are read while the rows of chunk c are in flight, so a chunk
    // waits for one read latency instead of three.  Every read in the loop is
    // unconditional (indices clamped into their arrays, validity applied
    // where the values are used; rows that a lane does not need are read as
    // out-of-range buffer loads): the in-order load counter then lets each
    // wait cover exactly the reads it needs.
    struct Words {
        uint32_t lw, ps;
        uint64_t word;  // fiber-start bits of the chunk's 64 nonzeros (wave-uniform)
        float v;
    };
    struct Entries {
        uint32_t s, i1, i2;
    };
    const uint32_t lastX = a.nnz - 1;  // nnz >= 1: the run is not empty
    const uint32_t lastC = chunks - 1;
    auto readWords = [&](uint32_t c, Words &w) {
        const uint32_t x = std::min(c * 64 + (uint32_t)lane, lastX);
        w.lw = a.leaf[x];
        // one 8-byte word per chunk, read through the scalar cache (the chunk
        // index is wave-uniform; chunks past the tensor read the last word)
        const uint32_t cw = __builtin_amdgcn_readfirstlane(std::min(c, lastC));
        using ConstU64 = const __attribute__((address_space(4))) uint64_t;
        w.word = *reinterpret_cast<ConstU64 *>(reinterpret_cast<uintptr_t>(a.fbit + 2 * cw));
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
        const bool start = x < a.nnz && ((w.word >> lane) & 1ull);
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

    // Carried from the previous chunk (every lane alike): the level-1 /
    // level-2 indices of its last nonzero, and (fiber-level factoring) the
    // level-1 / level-2 / root rows in effect there.
    uint32_t i1c = kNone, i2c = kNone;
    [[maybe_unused]] Raw<T, VEC> cw1[NB], cw2[NB];

    // one chunk: wc / ec hold chunk c; the entries of chunk c+1 (words wn)
    // go to en, the words of chunk c+2 to wf.  Group g takes the run of
    // nonzeros g*STEPS .. g*STEPS + STEPS-1 of the chunk, one per step; lane
    // g*G + t holds the entry of step t (STEPS == G).
    auto step = [&](uint32_t c, Words &wc, Entries &ec, Words &wn, Entries &en,
                    Words &wf) __attribute__((always_inline)) {
        const uint32_t x = c * 64 + (uint32_t)lane;
        const bool valid = x < a.nnz;
        // lanes past the tensor: zero value, no slice; their (clamped) row
        // indices stay valid
        const uint32_t s = valid ? ec.s : kNone;
        const float v = valid ? wc.v : 0.f;
        uint32_t sprev = __shfl_up(s, 1, 64);
        if (lane == 0) sprev = scur;
        const bool sstart = s != sprev;
        const uint64_t bm = __ballot(sstart);  // slice starts in the chunk
        // Segments: maximal runs of nonzeros with the same slice and level-1
        // index (newJ) and, 4-D, the same level-2 index (newF).  A fiber cut by
        // the -f split continues its segment.  The rows of a level are read
        // only where one of its segments opens; lanes of a group whose run
        // begins inside the segment carried over from the previous chunk take
        // the carried row.
        uint64_t FM = 0, JM = 0;
        if constexpr (kFactor) {
            uint32_t p1 = __shfl_up(ec.i1, 1, 64);
            if (lane == 0) p1 = i1c;
            const bool newJ = sstart || ec.i1 != p1;
            bool newF = newJ;
            if constexpr (N == 4) {
                uint32_t p2 = __shfl_up(ec.i2, 1, 64);
                if (lane == 0) p2 = i2c;
                newF = newJ || ec.i2 != p2;
            }
            FM = __ballot(newF);
            JM = N == 4 ? __ballot(newJ) : FM;
        }
        const uint32_t gs = (uint32_t)(g * STEPS);
        // start bits of the group's steps (bit t: step t)
        const uint32_t gF = (uint32_t)(FM >> gs);
        const uint32_t gJ = (uint32_t)(JM >> gs);
        const uint32_t gS = (uint32_t)(bm >> gs);
        // the group's first entry still lies in the segment carried over
        const bool carryF = (FM & low_bits(gs + 1)) == 0;
        const bool carryJ = (JM & low_bits(gs + 1)) == 0;
        // steps at which some group reads rows (wave-uniform; bit t: step t)
        auto fold = [&](uint64_t m, uint64_t m0) -> uint32_t {
            uint32_t r = 0;
#pragma unroll
            for (int q = 0; q < GPR; ++q) r |= (uint32_t)(m >> (q * STEPS));
            // step 0: every group whose run does not begin in the carried segment
            return (r & ~1u & (uint32_t)low_bits(STEPS)) | ((m0 & low_bits((GPR - 1) * STEPS + 1)) ? 1u : 0u);
        };
        const uint32_t need1 = kFactor ? fold(N == 3 ? FM : JM, N == 3 ? FM : JM) : 0u;
        const uint32_t need2 = kFactor && N == 4 ? fold(FM, FM) : 0u;
        if constexpr (kFactor) {
            // the carried rows sit in the last group's lanes: spread them when
            // some run begins inside a carried segment
            constexpr int last = (GPR - 1) * G;
            if (!(((N == 3 ? FM : JM)) & 1ull)) {
#pragma unroll
                for (int b = 0; b < NB; ++b) cw1[b].from_lane(cw1[b], last + k);
            }
            if constexpr (N == 4) {
                if (!(FM & 1ull)) {
#pragma unroll
                    for (int b = 0; b < NB; ++b) cw2[b].from_lane(cw2[b], last + k);
                }
            }
        }
        // per-lane offsets and values, taken by the group's lanes step by step;
        // rows a lane does not need read as out-of-range (zero) buffer loads
        const bool runStart = (lane % STEPS) == 0;
        const uint32_t oL = wc.lw * rowBytes;
        const uint32_t o1 = !kFactor ? ec.i1 * rowBytes
                            : ((runStart && !(N == 3 ? carryF : carryJ)) || (!runStart && ((N == 3 ? FM : JM) >> lane) & 1ull))
                                ? ec.i1 * rowBytes : kOobOffset;
        const uint32_t o2 = !kFactor ? ec.i2 * rowBytes
                            : ((runStart && !carryF) || (!runStart && (FM >> lane) & 1ull)) ? ec.i2 * rowBytes : kOobOffset;
        const uint32_t o0 = (valid ? s : 0u) * rowBytes;  // root row (CP model value)

        // Partial sums of the group run (per column block):
        //   fa  leaf-fiber partial  sum co * U_leaf[leaf] over the open leaf fiber
        //   ga  (4-D) level-1 partial  sum fa (.) U_2[fiber] over its closed fibers
        //   la  slice partial  (products of closed segments with their rows)
        //   hd  partial of the slice open at the run start, once a later slice starts
        // fr1 / fr2: the level-1 / level-2 rows of the open segments
        // (fiber-level factoring).
        float hd[NB][VEC], la[NB][VEC], fa[NB][VEC], ga[NB][VEC];
        Raw<T, VEC> fr1[NB], fr2[NB];
#pragma unroll
        for (int b = 0; b < NB; ++b)
#pragma unroll
            for (int e = 0; e < VEC; ++e) hd[b][e] = la[b][e] = fa[b][e] = ga[b][e] = 0.f;
        bool seen = false;  // a slice started inside this group's run
        const uint32_t sFirst = grp_lane<G>(s, 0);
        uint32_t ls = sFirst;
        // close the slice open before step tt (all lanes of a group agree)
        auto closeSlice = [&](int b) {
            if (!seen) {
#pragma unroll
                for (int e = 0; e < VEC; ++e) hd[b][e] = la[b][e];
            } else {
                put(b, ls, la[b], true);  // began and ended in this run
            }
#pragma unroll
            for (int e = 0; e < VEC; ++e) la[b][e] = 0.f;
        };
#pragma unroll
        for (int t0 = 0; t0 < STEPS; t0 += TB) {
            Raw<T, VEC> rw[TB][NB], w1[TB][NB], w2[TB][NB], wz[TB][NB];
#pragma unroll
            for (int t = 0; t < TB; ++t) {
                const int tt = t0 + t;
                const uint32_t li = grp_lane<G>(oL, tt);
                [[maybe_unused]] uint32_t l1 = 0, l2 = 0, l0 = 0;
                if constexpr (kUpper) l1 = grp_lane<G>(o1, tt);
                if constexpr (kUpper && N == 4) l2 = grp_lane<G>(o2, tt);
                if constexpr (kModel) l0 = grp_lane<G>(o0, tt);
#pragma unroll
                for (int b = 0; b < NB; ++b) {
                    // blocks past R read columns 0.. and are dropped where used
                    const uint32_t cb4 = colBytes[b];
                    rw[t][b].load(UL, li + cb4);
                    if constexpr (kUpper && !kFactor) {
                        w1[t][b].load(U1, l1 + cb4);
                        if constexpr (N == 4) w2[t][b].load(U2, l2 + cb4);
                    }
                    if constexpr (kModel) wz[t][b].load(U0, l0 + cb4);
                }
                if constexpr (kFactor) {
                    // upper-level rows only at steps where some group opens a
                    // segment (wave-uniform branches)
                    if ((need1 >> tt) & 1u) {
#pragma unroll
                        for (int b = 0; b < NB; ++b) w1[t][b].loadb(rs1, l1 + colBytes[b]);
                    }
                    if constexpr (N == 4) {
                        if ((need2 >> tt) & 1u) {
#pragma unroll
                            for (int b = 0; b < NB; ++b) w2[t][b].loadb(rs2, l2 + colBytes[b]);
                        }
                    }
                }
            }
            if (t0 == 0) {
                // issued after this batch's rows and before their use, so the
                // rows can be waited for without waiting for these reads
                __builtin_amdgcn_sched_barrier(0);
                readEntries(c + 1, wn, en);
                readWords(c + 2, wf);
                __builtin_amdgcn_sched_barrier(0);
            }
#pragma unroll
            for (int t = 0; t < TB; ++t) {
                const int tt = t0 + t;
                const float vt = as_float(grp_lane<G>(__float_as_uint(v), tt));
                [[maybe_unused]] uint32_t pt = 0;
                if constexpr (OP == kLossResid) pt = grp_lane<G>(wc.ps, tt);
                const bool sst = tt > 0 && ((gS >> tt) & 1u);  // a slice starts here
                [[maybe_unused]] uint32_t sp = 0;
                if constexpr (kGrad) sp = grp_lane<G>(s, tt);
                // segment starts: close the open segments with their rows, take
                // the rows of the new ones
                if constexpr (kFactor) {
                    if (tt == 0) {
                        // the run's first rows: loaded, or carried over
#pragma unroll
                        for (int b = 0; b < NB; ++b) {
                            fr1[b] = cw1[b];
                            if (need1 & 1u)
                                if (!(N == 3 ? carryF : carryJ)) fr1[b] = w1[t][b];
                            if constexpr (N == 4) {
                                fr2[b] = cw2[b];
                                if (need2 & 1u)
                                    if (!carryF) fr2[b] = w2[t][b];
                            }
                        }
                    } else if ((gF >> tt) & 1u) {
#pragma unroll
                        for (int b = 0; b < NB; ++b) {
                            float q1[VEC];
                            fr1[b].get(q1);
                            if constexpr (N == 3) {
                                if constexpr (kGrad) {
#pragma unroll
                                    for (int e = 0; e < VEC; ++e) {
                                        la[b][e] = fmaf(fa[b][e], q1[e], la[b][e]);
                                        fa[b][e] = 0.f;
                                    }
                                    if (sst) closeSlice(b);
                                }
                                fr1[b] = w1[t][b];
                            } else {
                                float q2[VEC];
                                fr2[b].get(q2);
                                if constexpr (kGrad) {
#pragma unroll
                                    for (int e = 0; e < VEC; ++e) {
                                        ga[b][e] = fmaf(fa[b][e], q2[e], ga[b][e]);
                                        fa[b][e] = 0.f;
                                    }
                                }
                                if ((gJ >> tt) & 1u) {
                                    if constexpr (kGrad) {
#pragma unroll
                                        for (int e = 0; e < VEC; ++e) {
                                            la[b][e] = fmaf(ga[b][e], q1[e], la[b][e]);
                                            ga[b][e] = 0.f;
                                        }
                                        if (sst) closeSlice(b);
                                    }
                                    fr1[b] = w1[t][b];
                                }
                                fr2[b] = w2[t][b];
                            }
                        }
                    }
                } else if constexpr (kGrad) {
                    if (sst) {
#pragma unroll
                        for (int b = 0; b < NB; ++b) closeSlice(b);
                    }
                }
                if constexpr (kGrad) {
                    if (sst) {
                        ls = sp;
                        seen = true;
                    }
                }
                // coefficient of the entry
                const bool ok = c * 64 + gs + (uint32_t)tt < a.nnz;
                float co;
                if constexpr (kModel) {
                    float d = 0.f;
#pragma unroll
                    for (int b = 0; b < NB; ++b) {
                        const int col = b * BW + k * VEC;
                        if (col < R) {
                            float r[VEC], q[VEC], z[VEC];
                            rw[t][b].get(r);
                            w1[t][b].get(q);
                            wz[t][b].get(z);
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
                    const float e = ok ? d - vt : 0.f;
                    if (k == 0 && ok) {
                        sq += (double)e * (double)e;
                        if constexpr (OP == kLossResid)
                            if (a.resid) a.resid[pt] = e;
                    }
                    co = e * a.alpha;
                } else if constexpr (OP == kWeightGrad) {
                    co = vt * a.alpha;
                } else {
                    co = vt;
                }
                if constexpr (kGrad) {
#pragma unroll
                    for (int b = 0; b < NB; ++b) {
                        float r[VEC];
                        rw[t][b].get(r);
                        if constexpr (kFactor) {
#pragma unroll
                            for (int e = 0; e < VEC; ++e) fa[b][e] = fmaf(co, r[e], fa[b][e]);
                        } else {
                            if constexpr (kUpper) {
                                float q[VEC];
                                w1[t][b].get(q);
#pragma unroll
                                for (int e = 0; e < VEC; ++e) r[e] *= q[e];
                                if constexpr (N == 4) {
                                    w2[t][b].get(q);
#pragma unroll
                                    for (int e = 0; e < VEC; ++e) r[e] *= q[e];
                                }
                            }
#pragma unroll
                            for (int e = 0; e < VEC; ++e) la[b][e] = fmaf(co, r[e], la[b][e]);
                        }
                    }
                }
            }
        }
        if constexpr (kGrad) {
            if constexpr (kFactor) {
                // close the segment open at the run end with its rows
#pragma unroll
                for (int b = 0; b < NB; ++b) {
                    float q1[VEC];
                    fr1[b].get(q1);
                    if constexpr (N == 4) {
                        float q2[VEC];
                        fr2[b].get(q2);
#pragma unroll
                        for (int e = 0; e < VEC; ++e) ga[b][e] = fmaf(fa[b][e], q2[e], ga[b][e]);
#pragma unroll
                        for (int e = 0; e < VEC; ++e) la[b][e] = fmaf(ga[b][e], q1[e], la[b][e]);
                    } else {
#pragma unroll
                        for (int e = 0; e < VEC; ++e) la[b][e] = fmaf(fa[b][e], q1[e], la[b][e]);
                    }
                }
            }
            // No slice starts in the chunk: every run belongs to the open
            // slice, and each group adds its partial to its share of the carry.
            if (bm == 0) {
#pragma unroll
                for (int b = 0; b < NB; ++b)
#pragma unroll
                    for (int e = 0; e < VEC; ++e) acc[b][e] += la[b][e];
            } else {
                // The carry (the open slice's partial row from earlier chunks),
                // summed over the groups' shares and left in every lane.
#pragma unroll
                for (int b = 0; b < NB; ++b)
#pragma unroll
                    for (int off = G; off < 64; off <<= 1)
#pragma unroll
                        for (int e = 0; e < VEC; ++e) acc[b][e] += __shfl_xor(acc[b][e], off, 64);
                // Join the slices across the groups.  Group g links to group g-1
                // when no slice starts at its first entry; a group without a
                // slice start inside its run passes the chain on.  The segmented
                // inclusive scan over groups gives every group the total of the
                // slice open at its run end (op); rt: that chain runs back to the
                // chunk start (and then carries the partial row of scur from
                // earlier chunks, acc).
                const bool cont = !(gS & 1u);  // no slice starts at the run start
                const bool single = !seen;
                bool rt = single && cont;
                const uint32_t sLast = ls;
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
                // a group with a slice start inside its run closes the slice open
                // at its run start: its head partial plus the chain ending at the
                // previous group's run end (the carry for group 0)
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
                // starts a slice; the last group's stays open and becomes the
                // carry (held by the last group)
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
            }
        }
        if constexpr (kFactor) {
            // the rows in effect at the run ends (the last group's: those at
            // the chunk's last nonzero)
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                cw1[b] = fr1[b];
                if constexpr (N == 4) cw2[b] = fr2[b];
            }
            i1c = __shfl(ec.i1, 63, 64);
            if constexpr (N == 4) i2c = __shfl(ec.i2, 63, 64);
        }
        scur = __shfl(s, 63, 64);
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
        for (int b = 0; b < NB; ++b) {
#pragma unroll
            for (int off = G; off < 64; off <<= 1)
#pragma unroll
                for (int e = 0; e < VEC; ++e) acc[b][e] += __shfl_xor(acc[b][e], off, 64);
            if (g == 0) put(b, scur, acc[b], false);
        }
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

// loads this translation unit's code object on the current device (a kernel
// attribute query), so that the first traversal does not pay for it
template <int OP>
void walk_preload() {
    hipFuncAttributes attr;
    MMCSF_HIP(hipFuncGetAttributes(&attr, reinterpret_cast<const void *>(k_walk<3, 8, 1, 8, float, 4, OP>)));
}

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
