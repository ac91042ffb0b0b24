// CP-ALS on the GPU (cpd() of the reference, src/cpd_cpu.cu:484-691).
//
// Everything stays on the device; the host reads the fits (one float64 per
// iteration) and the solve statuses, after each iteration when a tolerance can
// stop the loop, else in batches.  Arithmetic per step:
//   MTTKRP                 the MM-CSF kernels (float32 accumulation)
//   Gram  A^T A            MFMA v_mfma_f32_16x16x4f32 over row chunks, chunk sums
//                          reduced in float64
//   Hadamard of Grams, the R x R inverse (Cholesky, in LDS for R <= 36; LU
//   with partial pivoting when the matrix is not SPD),
//   the row solve U = M H^-1, column norms and the fit     float64
#include <algorithm>
#include <chrono>
#include <mutex>
#include <string>
#include <cmath>
#include <cstring>
#include <vector>

#include "mmcsf_internal.hpp"

namespace mmcsf {

namespace {

typedef float v4f __attribute__((ext_vector_type(4)));

// LDS bytes for k_normal_inverse's three R x R float64 matrices (R <= 36)
constexpr size_t kInverseLds = 32768;

// iterations run between two host reads of the fits and solve statuses when no
// tolerance can stop the loop early
constexpr int kFitBatch = 64;

// ---------------------------------------------------------------------------
// Gram: G += U^T U for U row-major rows x R.  blockIdx.y picks a 16x16 tile
// pair (bi <= bj); every wave reduces `rowsPerWave` rows with MFMA:
//   A operand (16x4)  lane l: U[k0 + l/16][16*bi + l%16]
//   B operand (4x16)  lane l: U[k0 + l/16][16*bj + l%16]
//   D (16x16)         lane l, reg r: D[4*(l/16) + r][l%16]
// Four steps' operands are loaded before their MFMAs; the block's four waves
// are summed in LDS before one wave adds the tile to G with float64 atomics.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_gram(const float *__restrict__ U, uint32_t rows, int R,
                                              uint32_t rowsPerWave, double *__restrict__ G) {
    __shared__ float part[4][256];
    const int nb = (R + 15) / 16;
    int p = blockIdx.y, bi = 0;
    while (p >= nb - bi) {  // tile pair p -> (bi, bj), bi <= bj
        p -= nb - bi;
        ++bi;
    }
    const int bj = bi + p;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const uint64_t wave = (uint64_t)blockIdx.x * 4 + wid;
    const uint64_t r0 = wave * rowsPerWave;
    const uint64_t r1 = min((uint64_t)rows, r0 + rowsPerWave);
    const int ci = 16 * bi + (lane & 15), cj = 16 * bj + (lane & 15);
    // columns past R read column R-1 and contribute zero
    const int cic = min(ci, R - 1), cjc = min(cj, R - 1);
    const float mi = ci < R ? 1.f : 0.f, mj = cj < R ? 1.f : 0.f;
    const int kk = lane >> 4;
    v4f acc = {0.f, 0.f, 0.f, 0.f};
    uint64_t k = r0;
    for (; k + 16 <= r1; k += 16) {
        float a[4], b[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const float *u = U + (k + 4 * q + kk) * (uint64_t)R;
            a[q] = u[cic];
            b[q] = u[cjc];
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a[q] * mi, b[q] * mj, acc, 0, 0, 0);
    }
    for (; k < r1; k += 4) {
        const uint64_t row = k + kk;
        float a = 0.f, b = 0.f;
        if (row < r1) {
            const float *u = U + row * (uint64_t)R;
            a = u[cic] * mi;
            b = u[cjc] * mj;
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) part[wid][lane * 4 + r] = acc[r];
    __syncthreads();
    if (wid == 0) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const float v = part[0][lane * 4 + r] + part[1][lane * 4 + r] + part[2][lane * 4 + r] + part[3][lane * 4 + r];
            const int i = 16 * bi + 4 * (lane >> 4) + r;
            const int j = cj;
            if (i < R && j < R) {
                unsafeAtomicAdd(G + (size_t)i * R + j, (double)v);
                if (bi != bj) unsafeAtomicAdd(G + (size_t)j * R + i, (double)v);
            }
        }
    }
}

__global__ void k_zero_d(double *p, uint64_t n) {
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

// ---------------------------------------------------------------------------
// H = Hadamard_{n != skip} G[n]; Hinv = H^-1, or by LU with partial pivoting
// when the Cholesky factorization fails (MatrixSolveNormals,
// cpd_cpu.cu:140-208).  One workgroup; W is an R*R*3 float64 workspace, and
// for R <= 36 the three matrices live in LDS instead.  Cholesky (lower), the
// inverse X of L, then H^-1 = X^T X.  Hpad (optional, R <= 32) receives H^-1
// zero-padded to 32 x 32.
// ---------------------------------------------------------------------------
__global__ void k_normal_inverse(const double *__restrict__ G, int N, int skip, int R, double *__restrict__ W,
                                 double *__restrict__ Hinv, double *__restrict__ Hpad, int *__restrict__ status) {
    const int tid = threadIdx.x, nt = blockDim.x;
    // H, L and X live in LDS when the launch provides 3 R*R doubles of it
    // (every elimination step then waits on LDS, not on L2), else in W
    extern __shared__ double lw[];
    const bool inLds = (size_t)3 * R * R * sizeof(double) <= kInverseLds;
    double *const base = inLds ? lw : W;
    double *H = base, *L = base + (size_t)R * R, *X = base + 2 * (size_t)R * R;
    __shared__ int fail;
    for (int e = tid; e < R * R; e += nt) {
        double h = 1.0;
        for (int n = 0; n < N; ++n)
            if (n != skip) h *= G[(size_t)n * R * R + e];
        H[e] = h;
        L[e] = 0.0;
    }
    if (tid == 0) fail = 0;
    __syncthreads();
    // Cholesky H = L L^T (lower; spotrf, cpd_cpu.cu:168-179), right-looking: per
    // column k the diagonal, the column below it, then the trailing lower
    // triangle, each step's entries updated in parallel.  A diagonal that is
    // not positive (the leading minor is not positive definite) sends the
    // solve to LU, as spotrf's failure sends the reference to sgesv.
    for (int k = 0; k < R; ++k) {
        const double d = H[(size_t)k * R + k];  // the same value in every thread
        if (!(d > 0.0)) {
            if (tid == 0) fail = 1;
            break;
        }
        const double lkk = sqrt(d);
        for (int i = k + 1 + tid; i < R; i += nt) L[(size_t)i * R + k] = H[(size_t)i * R + k] / lkk;
        if (tid == 0) L[(size_t)k * R + k] = lkk;
        __syncthreads();
        const int m = R - k - 1;
        for (int e = tid; e < m * m; e += nt) {
            const int i = k + 1 + e / m, j = k + 1 + e % m;
            if (j <= i) H[(size_t)i * R + j] -= L[(size_t)i * R + k] * L[(size_t)j * R + k];
        }
        __syncthreads();
    }
    __syncthreads();
    if (!fail) {
        // X = L^-1 (lower) by column-oriented forward substitution on X = I:
        // per row k, scale it by 1 / L[k][k], then eliminate column k from the
        // rows below; every step's entries are updated in parallel
        for (int e = tid; e < R * R; e += nt) X[e] = (e / R == e % R) ? 1.0 : 0.0;
        __syncthreads();
        for (int k = 0; k < R; ++k) {
            const double rk = 1.0 / L[(size_t)k * R + k];
            for (int c = tid; c <= k; c += nt) X[(size_t)k * R + c] *= rk;
            __syncthreads();
            const int m = R - k - 1;
            for (int e = tid; e < m * (k + 1); e += nt) {
                const int i = k + 1 + e / (k + 1), c = e % (k + 1);
                X[(size_t)i * R + c] -= L[(size_t)i * R + k] * X[(size_t)k * R + c];
            }
            __syncthreads();
        }
        // H^-1 = X^T X
        for (int e = tid; e < R * R; e += nt) {
            const int i = e / R, j = e % R;
            double s = 0.0;
            for (int k = max(i, j); k < R; ++k) s += X[(size_t)k * R + i] * X[(size_t)k * R + j];
            Hinv[e] = s;
            if (Hpad) Hpad[i * 32 + j] = s;
        }
        if (Hpad)
            for (int e = tid; e < 32 * 32; e += nt)
                if (e / 32 >= R || e % 32 >= R) Hpad[e] = 0.0;
        if (tid == 0) status[0] = 0;
        return;
    }
    // LU with partial pivoting on a fresh copy of H (one thread: R is small)
    if (tid == 0) {
        status[0] = 1;
        for (int e = 0; e < R * R; ++e) {
            double h = 1.0;
            for (int n = 0; n < N; ++n)
                if (n != skip) h *= G[(size_t)n * R * R + e];
            L[e] = h;                              // LU in place
            X[e] = (e / R == e % R) ? 1.0 : 0.0;  // right-hand sides: identity
        }
        for (int k = 0; k < R; ++k) {
            int piv = k;
            double best = fabs(L[(size_t)k * R + k]);
            for (int i = k + 1; i < R; ++i)
                if (fabs(L[(size_t)i * R + k]) > best) { best = fabs(L[(size_t)i * R + k]); piv = i; }
            if (best == 0.0) { status[0] = 2; break; }
            if (piv != k)
                for (int c = 0; c < R; ++c) {
                    double t = L[(size_t)k * R + c]; L[(size_t)k * R + c] = L[(size_t)piv * R + c]; L[(size_t)piv * R + c] = t;
                    t = X[(size_t)k * R + c]; X[(size_t)k * R + c] = X[(size_t)piv * R + c]; X[(size_t)piv * R + c] = t;
                }
            for (int i = k + 1; i < R; ++i) {
                double f = L[(size_t)i * R + k] / L[(size_t)k * R + k];
                for (int c = k; c < R; ++c) L[(size_t)i * R + c] -= f * L[(size_t)k * R + c];
                for (int c = 0; c < R; ++c) X[(size_t)i * R + c] -= f * X[(size_t)k * R + c];
            }
        }
        if (status[0] == 1)
            for (int c = 0; c < R; ++c)
                for (int i = R - 1; i >= 0; --i) {
                    double s = X[(size_t)i * R + c];
                    for (int k = i + 1; k < R; ++k) s -= L[(size_t)i * R + k] * Hinv[(size_t)k * R + c];
                    Hinv[(size_t)i * R + c] = s / L[(size_t)i * R + i];
                }
        if (Hpad)
            for (int e = 0; e < 32 * 32; ++e)
                Hpad[e] = (e / 32 < R && e % 32 < R) ? Hinv[(e / 32) * R + e % 32] : 0.0;
    }
}

// U[i][j] = sum_k M[i][k] Hinv[k][j] (float64 accumulation); each thread
// computes four columns of one row, reading M[i][k] once for all four
// The same product for R <= 32 with R % 4 == 0: one row per thread (the grid
// covers the rows once, so nothing is loop-invariant across rows and the
// Hinv reads stay in the k loop), read with eight 16-byte loads; Hinv
// zero-padded to 32 x 32 in LDS (every lane of a wave reads the same entry),
// 32 float64 sums per thread.
// U = M H^-1 for R <= 32 (R % 4 == 0), one row per thread, fused with the
// column norms of the new rows: kind 0 adds float64 sums of squares to sq,
// kind 1 the largest non-negative entry (order-preserving bits) to mx.  The
// block's rows are one contiguous run of M and U: they are read and written
// as consecutive float4 and pass through LDS (33 floats per row), where each
// thread solves its row and the block reduces the columns, so that each block
// adds one value per column.  H^-1 (Hpad: zero-padded to 32 x 32) is the same
// for every lane: it is read with scalar loads and enters the float64 FMAs as
// scalar operands.  Block 0 also clears Gz, the R x R Gram that
// k_scale_gram32 accumulates next.
constexpr int kSolveThreads = 512;
__global__ __launch_bounds__(kSolveThreads) void k_solve_rows32(const float *__restrict__ M,
                                                                const double *__restrict__ Hpad, uint32_t rows,
                                                                int R, float *__restrict__ U, int kind,
                                                                double *__restrict__ sq, uint32_t *__restrict__ mx,
                                                                double *__restrict__ Gz) {
    __shared__ float tile[kSolveThreads * 33];
    __shared__ double ps[kSolveThreads];
    using ConstD = const __attribute__((address_space(4))) double;
    const ConstD *H = reinterpret_cast<const ConstD *>(reinterpret_cast<uintptr_t>(Hpad));
    if (blockIdx.x == 0)
        for (int e = threadIdx.x; e < R * R; e += blockDim.x) Gz[e] = 0.0;
    const uint64_t row0 = (uint64_t)blockIdx.x * kSolveThreads;
    const uint32_t nrows = (uint32_t)min((uint64_t)kSolveThreads, rows - row0);
    const int q4 = R / 4;  // float4 per row
    const uint32_t n4 = nrows * (uint32_t)q4;
    // at most 8 float4 per thread (n4 <= 512 rows x 8): every read is issued
    // before the first is waited for
    constexpr int kPer = 8;
    const float4 *m4 = reinterpret_cast<const float4 *>(M + row0 * R);
    float4 v[kPer];
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
        const uint32_t e = threadIdx.x + q * kSolveThreads;
        v[q] = e < n4 ? m4[e] : make_float4(0.f, 0.f, 0.f, 0.f);
    }
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
        const uint32_t e = threadIdx.x + q * kSolveThreads;
        if (e < n4) {
            float *t = tile + (e / q4) * 33 + (e % q4) * 4;
            t[0] = v[q].x; t[1] = v[q].y; t[2] = v[q].z; t[3] = v[q].w;
        }
    }
    __syncthreads();
    const bool live = threadIdx.x < nrows;
    float *mine = tile + threadIdx.x * 33;
    // the row times H^-1, 16 columns at a time (16 float64 sums live)
    float out[32];
#pragma unroll
    for (int half = 0; half < 2; ++half) {
        double s[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) s[j] = 0.0;
#pragma unroll 4
        for (int k = 0; k < 32; ++k) {
            const double mk = k < R ? (double)mine[k] : 0.0;
#pragma unroll
            for (int j = 0; j < 16; ++j) s[j] = fma(mk, H[k * 32 + half * 16 + j], s[j]);
        }
#pragma unroll
        for (int j = 0; j < 16; ++j) out[half * 16 + j] = (float)s[j];
    }
    // the new row replaces the thread's own row (rows past the factor and
    // columns past R hold zero)
#pragma unroll
    for (int j = 0; j < 32; ++j) mine[j] = (live && j < R) ? out[j] : 0.f;
    __syncthreads();
    float4 *u4 = reinterpret_cast<float4 *>(U + row0 * R);
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
        const uint32_t e = threadIdx.x + q * kSolveThreads;
        if (e < n4) {
            const float *t = tile + (e / q4) * 33 + (e % q4) * 4;
            u4[e] = make_float4(t[0], t[1], t[2], t[3]);
        }
    }
    // thread t: column t % 32 over rows t / 32, t / 32 + 16, ...
    constexpr int kSubs = kSolveThreads / 32;
    const int j = threadIdx.x % 32, sub = threadIdx.x / 32;
    double ss = 0.0;
    float mm = 0.f;
    for (int r = sub; r < kSolveThreads; r += kSubs) {
        const float v = tile[r * 33 + j];
        ss += (double)v * v;
        mm = fmaxf(mm, v);
    }
    __syncthreads();  // rows no longer read: the maxima reuse the front of the tile
    float *pm = tile;
    ps[sub * 32 + j] = ss;
    pm[sub * 32 + j] = mm;
    __syncthreads();
    if (threadIdx.x < R) {  // sub 0, column j
        for (int q = 1; q < kSubs; ++q) {
            ss += ps[q * 32 + j];
            mm = fmaxf(mm, pm[q * 32 + j]);
        }
        if (kind == 0) unsafeAtomicAdd(sq + j, ss);
        else atomicMax(mx + j, __float_as_uint(mm));
    }
}

// lambda of column c from the reduced norms: the 2-norm (kind 0) or the
// largest entry clamped to at least 1 (kind 1, MatrixMaxNorm, cpd_cpu.cu:316-319)
__device__ inline double lambda_of(int kind, const double *sq, const uint32_t *mx, int c) {
    return kind == 0 ? sqrt(sq[c]) : fmax((double)__uint_as_float(mx[c]), 1.0);
}

__device__ inline uint16_t to_bf16(float x) {
    uint32_t u = __float_as_uint(x);
    u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
    return (uint16_t)(u >> 16);
}

// The normalisation and Gram of a factor (R <= 32): lambda from the column
// norms, U[:, c] /= lambda_c in place, Ub (optional) = its bfloat16 copy, and
// G += U^T U of the normalised rows.  Every wave owns kScaleRows rows and
// all their columns (the 16 x 16 tiles cols 0-15 x 0-15, 0-15 x 16-31 and
// 16-31 x 16-31, MFMA operands laid out as in k_gram), so each element is
// read, scaled and written by one lane.  The block's 16 waves are summed in
// LDS before the block adds its tiles to G.  Block 0 writes lambda and clears
// sqNext / mxNext, the norm accumulators of the next update.  Mz (optional):
// the first nz floats of the MTTKRP output, cleared for the next mode's
// MTTKRP (nothing reads them after the row solve).
constexpr int kScaleThreads = 1024;
constexpr int kScaleRows = 64;
__global__ __launch_bounds__(kScaleThreads) void k_scale_gram32(float *__restrict__ U, uint32_t rows, int R,
                                                      int kind, const double *__restrict__ sq,
                                                      const uint32_t *__restrict__ mx, double *__restrict__ lam,
                                                      double *__restrict__ sqNext, uint32_t *__restrict__ mxNext,
                                                      uint16_t *__restrict__ Ub, double *__restrict__ G,
                                                      float *__restrict__ Mz, uint64_t nz) {
    constexpr int kWaves = kScaleThreads / 64;
    if (Mz) {
        float4 *z4 = reinterpret_cast<float4 *>(Mz);
        const uint64_t n4 = nz / 4, stride = (uint64_t)gridDim.x * blockDim.x;
        for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n4; i += stride)
            z4[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        for (uint64_t i = n4 * 4 + blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < nz; i += stride) Mz[i] = 0.f;
    }
    __shared__ float part[3][kWaves][256];
    if (blockIdx.x == 0 && threadIdx.x < R) {
        lam[threadIdx.x] = lambda_of(kind, sq, mx, threadIdx.x);
        sqNext[threadIdx.x] = 0.0;
        mxNext[threadIdx.x] = 0u;
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const bool two = R > 16;
    const int c0 = lane & 15, c1 = 16 + (lane & 15);
    const bool ok0 = c0 < R, ok1 = c1 < R;
    double d0 = ok0 ? lambda_of(kind, sq, mx, c0) : 1.0, d1 = ok1 ? lambda_of(kind, sq, mx, c1) : 1.0;
    if (d0 == 0.0) d0 = 1.0;
    if (d1 == 0.0) d1 = 1.0;
    const uint64_t wave = (uint64_t)blockIdx.x * kWaves + wid;
    const uint64_t r0 = wave * kScaleRows;
    const uint64_t r1 = min((uint64_t)rows, r0 + kScaleRows);
    const int kk = lane >> 4;
    v4f a00 = {0.f, 0.f, 0.f, 0.f}, a01 = a00, a11 = a00;
    for (uint64_t k = r0; k < r1; k += 16) {
        float x0[4], x1[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const uint64_t row = k + 4 * q + kk;
            const bool in = row < r1;
            x0[q] = in && ok0 ? U[row * R + c0] : 0.f;
            x1[q] = in && ok1 ? U[row * R + c1] : 0.f;
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const uint64_t row = k + 4 * q + kk;
            const bool in = row < r1;
            x0[q] = (float)((double)x0[q] / d0);
            x1[q] = (float)((double)x1[q] / d1);
            if (in && ok0) {
                U[row * R + c0] = x0[q];
                if (Ub) Ub[row * R + c0] = to_bf16(x0[q]);
            }
            if (in && ok1) {
                U[row * R + c1] = x1[q];
                if (Ub) Ub[row * R + c1] = to_bf16(x1[q]);
            }
            a00 = __builtin_amdgcn_mfma_f32_16x16x4f32(x0[q], x0[q], a00, 0, 0, 0);
            if (two) {
                a01 = __builtin_amdgcn_mfma_f32_16x16x4f32(x0[q], x1[q], a01, 0, 0, 0);
                a11 = __builtin_amdgcn_mfma_f32_16x16x4f32(x1[q], x1[q], a11, 0, 0, 0);
            }
        }
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        part[0][wid][lane * 4 + r] = a00[r];
        part[1][wid][lane * 4 + r] = a01[r];
        part[2][wid][lane * 4 + r] = a11[r];
    }
    __syncthreads();
    // tile t of the wave: t = 0 (0,0), 1 (0,1), 2 (1,1)
    if (wid < (two ? 3 : 1)) {
        const int bi = wid == 2 ? 1 : 0, bj = wid == 0 ? 0 : 1;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            float v = 0.f;
            for (int w = 0; w < kWaves; ++w) v += part[wid][w][lane * 4 + r];
            const int i = 16 * bi + 4 * (lane >> 4) + r;
            const int j = 16 * bj + (lane & 15);
            if (i < R && j < R) {
                unsafeAtomicAdd(G + (size_t)i * R + j, (double)v);
                if (bi != bj) unsafeAtomicAdd(G + (size_t)j * R + i, (double)v);
            }
        }
    }
}

__global__ __launch_bounds__(256) void k_apply_inverse(const float *__restrict__ M, const double *__restrict__ Hinv,
                                                       uint32_t rows, int R, float *__restrict__ U) {
    extern __shared__ double hs[];
    const bool inLds = (size_t)R * R * sizeof(double) <= 32768;
    if (inLds) {
        for (int e = threadIdx.x; e < R * R; e += blockDim.x) hs[e] = Hinv[e];
        __syncthreads();
    }
    const double *Hp = inLds ? hs : Hinv;
    const int quads = (R + 3) / 4;
    const uint64_t total = (uint64_t)rows * quads;
    for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < total;
         e += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t i = e / quads;
        const int j0 = (int)(e % quads) * 4;
        const float *m = M + i * R;
        double s[4] = {0.0, 0.0, 0.0, 0.0};
        if (j0 + 4 <= R) {
            for (int k = 0; k < R; ++k) {
                const double mk = (double)m[k];
                const double *h = Hp + (size_t)k * R + j0;
#pragma unroll
                for (int q = 0; q < 4; ++q) s[q] += mk * h[q];
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) U[i * R + j0 + q] = (float)s[q];
        } else {
            for (int k = 0; k < R; ++k) {
                const double mk = (double)m[k];
                for (int q = 0; j0 + q < R; ++q) s[q] += mk * Hp[(size_t)k * R + j0 + q];
            }
            for (int q = 0; j0 + q < R; ++q) U[i * R + j0 + q] = (float)s[q];
        }
    }
}

// column 2-norms (kind 0, float64 sums of squares) or maxima (kind 1,
// largest non-negative entry as an order-preserving uint32 bit pattern).
// Threads t, t + R, ... of a block share column t % R; they are combined in
// LDS first, so each block adds one value per column to the global result.
__global__ __launch_bounds__(256) void k_colnorm(const float *__restrict__ U, uint32_t rows, int R, int kind,
                                                 double *__restrict__ sq, uint32_t *__restrict__ mx) {
    __shared__ double ps[256];
    __shared__ float pm[256];
    const int j = threadIdx.x % R;
    const int rowsPerPass = blockDim.x / R;
    const int sub = threadIdx.x / R;
    double s = 0.0;
    float m = 0.f;
    if (sub < rowsPerPass)
        for (uint64_t i = (uint64_t)blockIdx.x * rowsPerPass + sub; i < rows; i += (uint64_t)gridDim.x * rowsPerPass) {
            const float v = U[i * R + j];
            s += (double)v * v;
            m = fmaxf(m, v);
        }
    ps[threadIdx.x] = s;
    pm[threadIdx.x] = m;
    __syncthreads();
    if (sub == 0) {
        for (int q = 1; q < rowsPerPass; ++q) {
            s += ps[threadIdx.x + q * R];
            m = fmaxf(m, pm[threadIdx.x + q * R]);
        }
        if (kind == 0) unsafeAtomicAdd(sq + j, s);
        else atomicMax(mx + j, __float_as_uint(m));
    }
}

__global__ void k_lambda(int R, int kind, const double *__restrict__ sq, const uint32_t *__restrict__ mx,
                         double *__restrict__ lam) {
    const int j = threadIdx.x;
    if (j >= R) return;
    if (kind == 0) lam[j] = sqrt(sq[j]);
    else lam[j] = fmax((double)__uint_as_float(mx[j]), 1.0);  // MatrixMaxNorm clamp (cpd_cpu.cu:316-319)
}

__global__ void k_scale_cols(float *__restrict__ U, uint32_t rows, int R, const double *__restrict__ lam,
                             double *__restrict__ lamAcc) {
    const uint64_t total = (uint64_t)rows * R;
    for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < total;
         e += (uint64_t)gridDim.x * blockDim.x) {
        const double l = lam[e % R];
        U[e] = (float)((double)U[e] / (l != 0.0 ? l : 1.0));
    }
    if (lamAcc && blockIdx.x == 0 && threadIdx.x < R) lamAcc[threadIdx.x] *= lam[threadIdx.x];
}

__global__ __launch_bounds__(256) void k_dot_lambda(const float *__restrict__ A, const float *__restrict__ B,
                                                    uint32_t rows, int R, const double *__restrict__ lam,
                                                    double *__restrict__ acc) {
    double s = 0.0;
    const uint64_t total = (uint64_t)rows * R;
    if (total < 0x80000000ull) {  // 32-bit indices and column modulo (e + step stays below 2^32)
        const uint32_t n = (uint32_t)total, r = (uint32_t)R, step = gridDim.x * blockDim.x;
        for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += step)
            s += (double)A[e] * (double)B[e] * lam[e % r];
    } else {
        for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < total;
             e += (uint64_t)gridDim.x * blockDim.x)
            s += (double)A[e] * (double)B[e] * lam[e % R];
    }
    __shared__ double red[256];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) unsafeAtomicAdd(acc, red[0]);
}

__global__ __launch_bounds__(256) void k_sumsq(const float *__restrict__ v, uint64_t n, double *__restrict__ acc) {
    double s = 0.0;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        s += (double)v[i] * v[i];
    __shared__ double red[256];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) unsafeAtomicAdd(acc, red[0]);
}

// fit = 1 - sqrt(|X|^2 + lam^T (Hadamard_m G_m) lam - 2 inner) / |X|
// (KruskalTensorFit, cpd_cpu.cu:440-465, with the full symmetric product)
// fit of the iteration into fit[0] and hist[0]; acc[0] counts the modes whose
// solve fell back to LU (status 1), acc[1] is set when one was singular (2).
// inner[0] is cleared once read: k_dot_lambda of the next iteration adds to it.
__global__ void k_fit(const double *__restrict__ G, int N, int R, const double *__restrict__ lam,
                      double *__restrict__ inner, const double *__restrict__ xnormsq, double *__restrict__ fit,
                      double *__restrict__ hist, const int *__restrict__ status, int *__restrict__ acc) {
    __shared__ double red[256];
    double s = 0.0;
    for (int e = threadIdx.x; e < R * R; e += blockDim.x) {
        double h = 1.0;
        for (int n = 0; n < N; ++n) h *= G[(size_t)n * R * R + e];
        s += h * lam[e / R] * lam[e % R];
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double normMats = fabs(red[0]);
        double resid = xnormsq[0] + normMats - 2.0 * inner[0];
        inner[0] = 0.0;
        if (resid > 0.0) resid = sqrt(resid);
        fit[0] = 1.0 - resid / sqrt(xnormsq[0]);
        hist[0] = fit[0];
        for (int m = 0; m < N; ++m) {
            if (status[m] == 1) ++acc[0];
            if (status[m] == 2) acc[1] = 1;
        }
    }
}

__global__ void k_f32_to_bf16(const float *__restrict__ in, uint16_t *__restrict__ out, uint64_t n) {
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t u = __float_as_uint(in[i]);
    u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
    out[i] = (uint16_t)(u >> 16);
}

__global__ void k_fill_d(double *p, int n, double v) {
    int i = threadIdx.x;
    if (i < n) p[i] = v;
}

unsigned grid_for(uint64_t n, unsigned cap = 65536) {
    return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((n + 255) / 256, cap));
}

// blocks for k_colnorm: enough to fill the chip, few enough that the
// per-column atomics do not contend
unsigned colnorm_grid(uint64_t rows, int R) {
    const uint64_t perPass = std::max(1, 256 / R);
    return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((rows + perPass - 1) / perPass, 512));
}

}  // namespace

void gram(const float *U, uint64_t rows, int R, double *G, hipStream_t s) {
    hipLaunchKernelGGL(k_zero_d, dim3(grid_for((uint64_t)R * R)), dim3(256), 0, s, G, (uint64_t)R * R);
    if (rows == 0) return;
    const int nb = (R + 15) / 16;
    const uint32_t rowsPerWave = 256;
    const uint64_t waves = (rows + rowsPerWave - 1) / rowsPerWave;
    dim3 grid((unsigned)((waves + 3) / 4), (unsigned)(nb * (nb + 1) / 2));
    hipLaunchKernelGGL(k_gram, grid, dim3(256), 0, s, U, (uint32_t)rows, R, rowsPerWave, G);
    MMCSF_HIP(hipGetLastError());
}

__global__ void k_zero_u32(uint32_t *p, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = 0u;
}

// CP-ALS work space: the float64 Grams, normal-matrix work space and inverse,
// lambda / norm vectors and fit scalars, the MTTKRP output of the mode being
// updated, and (bfloat16 MTTKRP) the bfloat16 factor copies; every part
// 256-byte aligned.
struct CpdSpace {
    double *G, *W, *Hinv, *Hpad, *lam, *lamFinal, *sq, *scal, *hist;
    uint32_t *mx;
    int *status, *acc;
    float *M;
    uint16_t *Ub[kMaxDims];
};

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// bytes of the work space, and (base != null) the parts carved from it
static size_t cpd_space(const Format &f, int R, int dtype, char *base, CpdSpace *sp) {
    const int N = f.ndims;
    const size_t RR = (size_t)R * R;
    uint64_t maxRows = 0;
    for (int m = 0; m < N; ++m) maxRows = std::max<uint64_t>(maxRows, f.dims[m]);
    size_t off = 0;
    auto take = [&](size_t bytes) -> char * {
        char *p = base ? base + off : nullptr;
        off += align256(bytes);
        return p;
    };
    CpdSpace t{};
    t.G = (double *)take(N * RR * sizeof(double));
    t.W = (double *)take(3 * RR * sizeof(double));
    t.Hinv = (double *)take(RR * sizeof(double));
    t.Hpad = (double *)take(32 * 32 * sizeof(double));
    t.lam = (double *)take(R * sizeof(double));
    t.lamFinal = (double *)take(R * sizeof(double));
    t.sq = (double *)take(2 * R * sizeof(double));  // two halves: this update's and the next one's
    t.scal = (double *)take(4 * sizeof(double));  // [0] |X|^2, [1] inner, [2] fit
    t.mx = (uint32_t *)take(2 * R * sizeof(uint32_t));
    t.status = (int *)take(kMaxDims * sizeof(int));  // solve status of each mode
    t.hist = (double *)take(kFitBatch * sizeof(double));  // fits of the iterations since the last host read
    t.acc = (int *)take(2 * sizeof(int));                 // LU fallbacks, singular flag
    t.M = (float *)take(maxRows * R * sizeof(float));
    for (int m = 0; m < kMaxDims; ++m) t.Ub[m] = nullptr;
    if (dtype == MMCSF_BF16)
        for (int m = 0; m < N; ++m) t.Ub[m] = (uint16_t *)take(f.dims[m] * R * sizeof(uint16_t));
    if (sp) *sp = t;
    return off;
}

// The ALS loop on a work space `ws` of cpd_space(...) bytes.  Every part that
// must start at zero is cleared by a kernel here, so the work space may hold
// anything on entry.  timeMttkrp: time the MTTKRP launches with HIP events.
// side (optional, non-blocking): a second stream on which each mode's normal
// matrix is inverted while that mode's MTTKRP runs (the inverse needs only the
// other modes' Grams); null runs everything on s.  The host reads the fits
// and solve statuses after every iteration when tol > 0, else once per batch
// of up to kFitBatch iterations.
static void cpd_als_on(const Format &f, int R, int iters, double tol, int dtype, float *const *U, double *lambdaHost,
                       mmcsf_cpd_stats *stats, hipStream_t s, char *ws, bool timeMttkrp, hipStream_t side) {
    const int N = f.ndims;
    const size_t RR = (size_t)R * R;
    CpdSpace st{};
    cpd_space(f, R, dtype, ws, &st);
    auto refresh_bf16 = [&](int m) {
        if (dtype == MMCSF_BF16)
            hipLaunchKernelGGL(k_f32_to_bf16, dim3(grid_for(f.dims[m] * R, 1u << 30)), dim3(256), 0, s, U[m],
                               st.Ub[m], (uint64_t)f.dims[m] * R);
    };
    // both halves of the norm accumulators; update u of the loop accumulates
    // into half u % 2, which is zero when it starts
    auto clear_norms = [&]() {
        hipLaunchKernelGGL(k_zero_d, dim3(1), dim3(256), 0, s, st.sq, (uint64_t)2 * R);
        hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(256), 0, s, st.mx, (uint64_t)2 * R);
    };
    for (int m = 0; m < N; ++m) {
        gram(U[m], f.dims[m], R, st.G + m * RR, s);
        refresh_bf16(m);
    }
    hipLaunchKernelGGL(k_zero_d, dim3(1), dim3(256), 0, s, st.scal, (uint64_t)4);
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, s, (uint32_t *)st.status, (uint64_t)kMaxDims);
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, s, (uint32_t *)st.acc, (uint64_t)2);
    // |X|^2 over every stored value (the partitions of a mixed MM-CSF)
    if (f.mixed) {
        for (int p = 0; p < N; ++p)
            if (f.partNnz[p])
                hipLaunchKernelGGL(k_sumsq, dim3(grid_for(f.partNnz[p], 4096)), dim3(256), 0, s, f.part[p].vals.p,
                                   (uint64_t)f.partNnz[p], st.scal);
    } else {
        hipLaunchKernelGGL(k_sumsq, dim3(grid_for(f.nnz, 4096)), dim3(256), 0, s, f.csf.vals.p, (uint64_t)f.nnz,
                           st.scal);
    }
    hipLaunchKernelGGL(k_fill_d, dim3(1), dim3(256), 0, s, st.lam, R, 1.0);
    Factors F{};
    F.dtype = dtype;
    F.R = R;
    for (int m = 0; m < N; ++m) F.U[m] = dtype == MMCSF_BF16 ? (const void *)st.Ub[m] : (const void *)U[m];
    // The host reads the fits and solve statuses after every iteration when a
    // tolerance may stop the loop (cpd_cpu.cu:670-672), else once per batch of
    // up to kFitBatch iterations: a tolerance <= 0 never stops it early.
    const int batch = tol > 0.0 ? 1 : std::min(kFitBatch, std::max(iters, 1));
    // events: MTTKRP start / end per iteration of a batch and mode (timing),
    // and the two joins of the side stream (after the previous mode's Gram,
    // after the inverse)
    struct Events {
        std::vector<hipEvent_t> e;
        ~Events() {
            for (hipEvent_t x : e)
                if (x) (void)hipEventDestroy(x);
        }
    } ev;
    const int nTime = timeMttkrp ? 2 * N * batch : 0;
    ev.e.assign(nTime + 2, nullptr);
    for (int i = 0; i < nTime; ++i) MMCSF_HIP(hipEventCreate(&ev.e[i]));
    hipEvent_t evReady = nullptr, evInv = nullptr;
    if (side) {
        MMCSF_HIP(hipEventCreateWithFlags(&ev.e[nTime], hipEventDisableTiming));
        MMCSF_HIP(hipEventCreateWithFlags(&ev.e[nTime + 1], hipEventDisableTiming));
        evReady = ev.e[nTime];
        evInv = ev.e[nTime + 1];
    }
    double mttkrpMs = 0, oldfit = 0, fit = 0;
    int it = 0, lu = 0;
    uint64_t upd = 0;  // factor updates so far (selects the half of the norm accumulators)
    bool mZeroed = false;  // the MTTKRP output was cleared for the next mode by the previous update
    clear_norms();
    auto t0 = std::chrono::steady_clock::now();
    bool stop = false;
    while (it < iters && !stop) {
        const int nb = std::min(batch, iters - it);
        for (int b = 0; b < nb; ++b) {
            const int cur = it + b;
            for (int m = 0; m < N; ++m) {
                const size_t ishm = 3 * RR * sizeof(double) <= kInverseLds ? 3 * RR * sizeof(double) : 0;
                // the fused row solve (R <= 32, 16-byte rows) reads H^-1 zero-padded
                const bool fused = R <= 32 && R % 4 == 0 && reinterpret_cast<uintptr_t>(U[m]) % 16 == 0;
                double *hpad = fused ? st.Hpad : nullptr;
                if (side) {
                    // the Grams of the other modes are final once the work queued on
                    // s so far (the previous mode's update) is done; the inverse
                    // then runs beside this mode's MTTKRP
                    MMCSF_HIP(hipEventRecord(evReady, s));
                    MMCSF_HIP(hipStreamWaitEvent(side, evReady, 0));
                    hipLaunchKernelGGL(k_normal_inverse, dim3(1), dim3(256), ishm, side, st.G, N, m, R, st.W, st.Hinv,
                                       hpad, st.status + m);
                    MMCSF_HIP(hipEventRecord(evInv, side));
                }
                hipEvent_t *te = timeMttkrp ? &ev.e[2 * (b * N + m)] : nullptr;
                mttkrp_mode(f, m, F, st.M, s, te ? te[0] : nullptr, te ? te[1] : nullptr, mZeroed);
                if (side) {
                    MMCSF_HIP(hipStreamWaitEvent(s, evInv, 0));
                } else {
                    hipLaunchKernelGGL(k_normal_inverse, dim3(1), dim3(256), ishm, s, st.G, N, m, R, st.W, st.Hinv,
                                       hpad, st.status + m);
                }
                const int kind = cur == 0 ? 0 : 1;  // 2-norm at iteration 0, max-norm after
                if (fused) {
                    // two launches: the row solve with the column norms, then lambda,
                    // the normalisation, the bfloat16 copy and the Gram
                    const int h = (int)(upd & 1);
                    const unsigned solveBlocks = (unsigned)((f.dims[m] + kSolveThreads - 1) / kSolveThreads);
                    hipLaunchKernelGGL(k_solve_rows32, dim3(solveBlocks), dim3(kSolveThreads), 0, s, st.M, st.Hpad,
                                       (uint32_t)f.dims[m], R, U[m], kind, st.sq + h * R, st.mx + h * R,
                                       st.G + m * RR);
                    // modes before the last also clear the next mode's MTTKRP output
                    // (the last mode's output feeds the inner product of the fit)
                    const bool clearNext = m + 1 < N;
                    const uint64_t waves = (f.dims[m] + kScaleRows - 1) / kScaleRows;
                    const uint64_t wavesPerBlock = kScaleThreads / 64;
                    hipLaunchKernelGGL(k_scale_gram32, dim3((unsigned)((waves + wavesPerBlock - 1) / wavesPerBlock)),
                                       dim3(kScaleThreads), 0, s, U[m],
                                       (uint32_t)f.dims[m], R, kind, st.sq + h * R, st.mx + h * R, st.lam,
                                       st.sq + (1 - h) * R, st.mx + (1 - h) * R,
                                       dtype == MMCSF_BF16 ? st.Ub[m] : (uint16_t *)nullptr, st.G + m * RR,
                                       clearNext ? st.M : (float *)nullptr,
                                       clearNext ? f.dims[m + 1] * (uint64_t)R : 0ull);
                    mZeroed = clearNext;
                } else {
                    const size_t shm = RR * sizeof(double) <= 32768 ? RR * sizeof(double) : 0;
                    hipLaunchKernelGGL(k_apply_inverse, dim3(grid_for(f.dims[m] * ((R + 3) / 4), 8192)), dim3(256),
                                       shm, s, st.M, st.Hinv, (uint32_t)f.dims[m], R, U[m]);
                    const int h = (int)(upd & 1);
                    clear_norms();
                    hipLaunchKernelGGL(k_colnorm, dim3(colnorm_grid(f.dims[m], R)), dim3(256), 0, s, U[m],
                                       (uint32_t)f.dims[m], R, kind, st.sq + h * R, st.mx + h * R);
                    hipLaunchKernelGGL(k_lambda, dim3(1), dim3(std::max(R, 64)), 0, s, R, kind, st.sq + h * R,
                                       st.mx + h * R, st.lam);
                    hipLaunchKernelGGL(k_scale_cols, dim3(grid_for(f.dims[m] * R, 8192)), dim3(256), 0, s, U[m],
                                       (uint32_t)f.dims[m], R, st.lam, (double *)nullptr);
                    gram(U[m], f.dims[m], R, st.G + m * RR, s);
                    refresh_bf16(m);
                    mZeroed = false;
                }
                ++upd;
                if (m == N - 1) {
                    // inner product of the last MTTKRP with the updated last factor
                    // (its accumulator is cleared by the setup and by k_fit)
                    hipLaunchKernelGGL(k_dot_lambda, dim3(grid_for(f.dims[m] * R, 512)), dim3(256), 0, s, U[m], st.M,
                                       (uint32_t)f.dims[m], R, st.lam, st.scal + 1);
                }
                MMCSF_HIP(hipGetLastError());
            }
            hipLaunchKernelGGL(k_fit, dim3(1), dim3(256), 0, s, st.G, N, R, st.lam, st.scal + 1, st.scal, st.scal + 2,
                               st.hist + b, st.status, st.acc);
        }
        double fits[kFitBatch];
        int acc[2] = {0, 0};
        MMCSF_HIP(hipMemcpyAsync(fits, st.hist, nb * sizeof(double), hipMemcpyDeviceToHost, s));
        MMCSF_HIP(hipMemcpyAsync(acc, st.acc, sizeof(acc), hipMemcpyDeviceToHost, s));
        MMCSF_HIP(hipStreamSynchronize(s));
        // a singular solve leaves that mode's factor (and the later modes')
        // undefined: the call fails
        MMCSF_REQUIRE(acc[1] == 0, MMCSF_ERR_UNSUPPORTED, "normal equations are singular");
        lu = acc[0];
        for (int b = 0; b < nb && !stop; ++b) {
            const int cur = it + b;
            if (timeMttkrp)
                for (int m = 0; m < N; ++m) {
                    float ms = 0;
                    MMCSF_HIP(hipEventElapsedTime(&ms, ev.e[2 * (b * N + m)], ev.e[2 * (b * N + m) + 1]));
                    mttkrpMs += ms;
                }
            fit = fits[b];
            if (stats && cur < (int)(sizeof(stats->fit_history) / sizeof(double))) stats->fit_history[cur] = fit;
            if (cur > 0 && std::fabs(fit - oldfit) < tol) {  // cpd_cpu.cu:670-672
                it = cur + 1;
                stop = true;
            }
            oldfit = fit;
        }
        if (!stop) it += nb;
    }
    auto t1 = std::chrono::steady_clock::now();
    // GetFinalLambda (cpd_cpu.cu:333-349): fold the column 2-norms into lambda
    MMCSF_HIP(hipMemcpyAsync(st.lamFinal, st.lam, R * sizeof(double), hipMemcpyDeviceToDevice, s));
    for (int m = 0; m < N; ++m) {
        clear_norms();
        hipLaunchKernelGGL(k_colnorm, dim3(colnorm_grid(f.dims[m], R)), dim3(256), 0, s, U[m],
                           (uint32_t)f.dims[m], R, 0, st.sq, st.mx);
        hipLaunchKernelGGL(k_lambda, dim3(1), dim3(std::max(R, 64)), 0, s, R, 0, st.sq, st.mx, st.lam);
        hipLaunchKernelGGL(k_scale_cols, dim3(grid_for(f.dims[m] * R, 8192)), dim3(256), 0, s, U[m],
                           (uint32_t)f.dims[m], R, st.lam, st.lamFinal);
    }
    MMCSF_HIP(hipGetLastError());
    if (lambdaHost)
        MMCSF_HIP(hipMemcpyAsync(lambdaHost, st.lamFinal, R * sizeof(double), hipMemcpyDeviceToHost, s));
    MMCSF_HIP(hipStreamSynchronize(s));
    if (stats) {
        stats->iterations = it;
        stats->fit = fit;
        stats->ms_total = std::chrono::duration<double, std::milli>(t1 - t0).count();
        stats->ms_mttkrp = mttkrpMs;
        stats->lu_fallbacks = lu;
    }
}

void cpd_als(const Format &f, int R, int iters, double tol, int dtype, float *const *U, double *lambdaHost,
             mmcsf_cpd_stats *stats, hipStream_t s) {
    DevArray<char> ws(cpd_space(f, R, dtype, nullptr, nullptr));
    // the stream the normal-matrix inverses run on (default priority)
    struct Side {
        hipStream_t s = nullptr;
        ~Side() {
            if (s) (void)hipStreamDestroy(s);
        }
    } side;
    MMCSF_HIP(hipStreamCreateWithFlags(&side.s, hipStreamNonBlocking));
    cpd_als_on(f, R, iters, tol, dtype, U, lambdaHost, stats, s, ws.p, true, side.s);
}

// ---------------------------------------------------------------------------
// CP-ALS without allocation: the work space is this file-scope device array,
// one per GPU, held until the process ends.  Host code never passes the
// array's name to a kernel (in host code it only names the symbol): the
// device address of each GPU's copy is looked up once per device with
// hipGetSymbolAddress and kept in kScratchAddr.  Every call on one GPU uses
// the same memory, so calls that overlap on one GPU race.
// ---------------------------------------------------------------------------
constexpr size_t kCpdScratchBytes = size_t(64) << 20;
static __device__ __align__(256) char scratch[kCpdScratchBytes];

constexpr int kMaxDevices = 64;
static void *kScratchAddr[kMaxDevices];
static std::mutex kScratchLock;

static char *scratch_on_current_device() {
    int dev = 0;
    MMCSF_HIP(hipGetDevice(&dev));
    MMCSF_REQUIRE(dev >= 0 && dev < kMaxDevices, MMCSF_ERR_UNSUPPORTED, "device index above 63");
    std::lock_guard<std::mutex> g(kScratchLock);
    if (!kScratchAddr[dev]) {
        void *p = nullptr;
        MMCSF_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        MMCSF_REQUIRE(p != nullptr, MMCSF_ERR_HIP, "no device address for the CP-ALS scratch array");
        kScratchAddr[dev] = p;
    }
    return static_cast<char *>(kScratchAddr[dev]);
}

size_t cpd_scratch_limit() { return kCpdScratchBytes; }

void cpd_als_noalloc(const Format &f, int R, int iters, double tol, int dtype, float *const *U, double *lambdaHost,
                     mmcsf_cpd_stats *stats, hipStream_t s) {
    const size_t need = cpd_space(f, R, dtype, nullptr, nullptr);
    MMCSF_REQUIRE(need <= kCpdScratchBytes, MMCSF_ERR_UNSUPPORTED,
                  "CP-ALS work space of " + std::to_string(need) + " bytes exceeds the " +
                      std::to_string(kCpdScratchBytes) +
                      "-byte scratch of the allocation-free entry point (largest mode rows x R x 4 bytes, plus "
                      "the Grams and, for bfloat16 MTTKRP, the bfloat16 factor copies)");
    char *ws = scratch_on_current_device();
    cpd_als_on(f, R, iters, tol, dtype, U, lambdaHost, stats, s, ws, false, nullptr);
}

}  // namespace mmcsf
