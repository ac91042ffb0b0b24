/*
 * mmcsf_oracle.c -- CPU restatement of the MM-CSF reference's MTTKRP path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.  The
 * product (mm-csf_amd/) never links, loads or calls it.
 *
 * Every function follows a reference function and cites it (paths are
 * relative to the reference checkout, src/ directory).  Indices are 64-bit
 * like the reference's ITYPE (size_t, util.h:60), values are float like its
 * DTYPE (util.h:59).
 *
 * How each function is pinned (tests/test_oracle.py):
 *   orc_order_tensormode   closed-form mode orders for every configuration
 *                          plus hand cases (toy.tns, cubes, ties).
 *   orc_sort_perm          sortedness + permutation property (numpy lexsort).
 *   orc_build_csf          toy.tns hand-derived arrays (MM-CSF and B-CSF),
 *                          an independent run/chunk closed-form restatement in
 *                          numpy for the -f split and the 4-D fbrS split, and
 *                          structural invariants (monotone pointers, chunk
 *                          lengths <= f, multiset of nonzeros preserved).
 *   orc_fbr_like_slc_inds  toy.tns hand values + definition check.
 *   orc_tiled_bin /        closed-form bin interval table; every slice lands
 *   orc_hyb_bin            in exactly one bin except the documented drop.
 *   orc_create_hyb         reconstruction of the sorted COO from the three
 *                          HYB parts and the class rules (1 nnz / all fibers
 *                          of length 1 / rest).
 *   orc_mttkrp_coo_f32/64  toy.tns closed form (slice sums with factors
 *                          m+0.5), brute-force float64 einsum on tiny random
 *                          tensors, rank-1 closed form, the cross-mode
 *                          inner-product identity.
 *   orc_mttkrp_f64_mt      identical, bit for bit, to orc_mttkrp_coo_f64 for
 *                          every thread count (every row is summed in stored
 *                          order), on 3-D and 4-D skewed tensors; toy.tns.
 *   orc_cp_resid_f64_mt    numpy float64 model values (dense.cp_loss) and a
 *                          planted exact model with zero residual.
 *   orc_mm_partition       hand-derived partitions of toy.tns (including
 *                          two last-slice searches and a decremented count)
 *                          and of two 4-nonzero cases: one where a decrement
 *                          changes a later choice, one where the last mode's
 *                          long fiber wins; every nonzero lands in exactly
 *                          one partition.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_MAXD 8

/* ------------------------------------------------------------------------ */
/* order_tensormode, util.h:255-311.  mode first, then the other modes by    */
/* ascending length; equal lengths take the lowest untaken mode index.      */
/* ------------------------------------------------------------------------ */
void orc_order_tensormode(int ndims, const uint64_t *dims, int mode, int nat,
                          int *order)
{
    int natMode[ORC_MAXD], sortModeLen[ORC_MAXD], taken[ORC_MAXD];
    for (int m = 0; m < ndims; ++m) {
        natMode[m] = (m + mode) % ndims;
        sortModeLen[m] = (int)dims[natMode[m]];
        taken[m] = 0;
    }
    if (nat) {
        for (int i = 0; i < ndims; ++i) order[i] = natMode[i];
        return;
    }
    /* util.h:282-290: selection-style sort of positions 1..N-1 only */
    for (int i = 1; i < ndims; i++)
        for (int j = i + 1; j < ndims; j++)
            if (sortModeLen[i] > sortModeLen[j]) {
                int t = sortModeLen[i];
                sortModeLen[i] = sortModeLen[j];
                sortModeLen[j] = t;
            }
    order[0] = mode;
    taken[mode] = 1;
    /* util.h:295-305 */
    for (int i = 1; i < ndims; i++)
        for (int j = 0; j < ndims; j++)
            if ((uint64_t)(int64_t)sortModeLen[i] == dims[j] && !taken[j]) {
                order[i] = j;
                taken[j] = 1;
                break;
            }
}

/* ------------------------------------------------------------------------ */
/* Lexicographic sort permutation.  The reference sorts on the first 2 (3-D) */
/* or 3 (4-D) keys with an unstable sample sort (util.h:389-527), leaving    */
/* the order inside a fiber unspecified; sorting on ALL keys with the input  */
/* position as the final tie-break is one legal outcome and is canonical.   */
/* ------------------------------------------------------------------------ */
static int g_nd;
static const uint64_t *g_keys[ORC_MAXD];

static int cmp_perm(const void *a, const void *b)
{
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    for (int l = 0; l < g_nd; ++l) {
        uint64_t kx = g_keys[l][x], ky = g_keys[l][y];
        if (kx != ky) return kx < ky ? -1 : 1;
    }
    return x < y ? -1 : (x > y ? 1 : 0);
}

/* inds[m][p] per tensor mode m; order = modeOrder; nkeys = how many of the */
/* ordered modes participate (ndims for the canonical full-key sort).       */
void orc_sort_perm(int ndims, uint64_t nnz, const uint64_t *const *inds,
                   const int *order, int nkeys, uint64_t *perm)
{
    (void)ndims;
    g_nd = nkeys;
    for (int l = 0; l < nkeys; ++l) g_keys[l] = inds[order[l]];
    for (uint64_t p = 0; p < nnz; ++p) perm[p] = p;
    qsort(perm, (size_t)nnz, sizeof(uint64_t), cmp_perm);
}

/* ------------------------------------------------------------------------ */
/* create_TiledHCSR, util.h:1320-1436 (create_HCSR, util.h:760-850, is the   */
/* same scan without the 4-D fbrS rule: pass fbrSThreshold = ~0).            */
/*                                                                          */
/* keys[l][p]: coordinate of level l (l = 0..ndims-1, in modeOrder order) of */
/* the p-th nonzero of the SORTED tensor.  fbrPtr[l] / fbrIdx[l] must hold   */
/* nnz+2 entries each.  On return ptrLen[l] / idxLen[l] hold the reference's */
/* vector sizes: fbrPtr[l] has count_l+1 entries; fbrIdx[l] has count_l     */
/* entries except the leaf-fiber level ndims-2, which carries one trailing  */
/* entry (util.h:1418-1419; garbage or a duplicate in the reference, the    */
/* last leaf-fiber index here).  Returns nFibers = fbrPtr[1].size()         */
/* (util.h:1424).                                                            */
/* ------------------------------------------------------------------------ */
int64_t orc_build_csf(int ndims, uint64_t nnz, const uint64_t *const *keys,
                      uint64_t fbrThreshold, uint64_t fbrSThreshold,
                      uint64_t *const *fbrPtr, uint64_t *const *fbrIdx,
                      uint64_t *ptrLen, uint64_t *idxLen)
{
    const int L = ndims - 1;              /* number of index levels */
    uint64_t prev[ORC_MAXD], cur[ORC_MAXD];
    uint64_t np[ORC_MAXD], ni[ORC_MAXD];  /* current vector sizes */
    if (nnz == 0) {
        for (int l = 0; l < L; ++l) { ptrLen[l] = 0; idxLen[l] = 0; }
        return 0;
    }
    for (int i = 0; i < L; ++i) {
        prev[i] = keys[i][0];
        cur[i] = 0;                        /* util.h:1345 value-initialised */
        fbrPtr[i][0] = 0;
        fbrIdx[i][0] = prev[i];
        np[i] = 1;
        ni[i] = 1;
    }
    uint64_t idx = 1;
    uint64_t fiberSNnz = 1;
    while (idx < nnz) {
        for (int i = 0; i < L; ++i) cur[i] = keys[i][idx];
        uint64_t fiberNnz = 1;
        int sameFbr = 1;
        for (int i = 0; i < L; ++i)
            if (cur[i] != prev[i]) sameFbr = 0;
        /* util.h:1364-1375: absorb nonzeros of the same fiber up to -f */
        while (sameFbr && idx < nnz && fiberNnz < fbrThreshold) {
            ++idx;
            fiberNnz++;
            fiberSNnz++;
            if (idx < nnz) {   /* the reference reads one past the end here */
                for (int i = 0; i < L; ++i) {
                    cur[i] = keys[i][idx];
                    if (cur[i] != prev[i]) sameFbr = 0;
                }
            }
        }
        if (idx == nnz) break;
        /* util.h:1381-1382: new leaf fiber */
        fbrPtr[L - 1][np[L - 1]++] = idx;
        fbrIdx[L - 1][ni[L - 1]++] = cur[L - 1];
        /* util.h:1385-1408: higher levels */
        for (int i = L - 2; i > -1; --i) {
            int diffFbr = 0;
            for (int iDim = i; iDim > -1; --iDim) {
                if (cur[iDim] != prev[iDim])
                    diffFbr = 1;
                else if (ndims == 4 && iDim == 1 && fiberSNnz > fbrSThreshold)
                    diffFbr = 1;
            }
            if (diffFbr) {
                if (ndims == 4 && i == 1) fiberSNnz = 1;
                fbrIdx[i][ni[i]++] = cur[i];
                fbrPtr[i][np[i]] = np[i + 1] - 1;
                np[i]++;
            }
        }
        for (int i = 0; i < L; ++i) prev[i] = cur[i];
        ++idx;
        fiberSNnz++;
    }
    /* util.h:1418-1422: closing entries */
    fbrPtr[L - 1][np[L - 1]++] = idx;
    {
        /* trailing leaf-fiber index: duplicate of the last one */
        uint64_t last = fbrIdx[L - 1][ni[L - 1] - 1];
        fbrIdx[L - 1][ni[L - 1]++] = last;
    }
    for (int i = L - 2; i > -1; --i) {
        fbrPtr[i][np[i]] = np[i + 1] - 1;
        np[i]++;
    }
    for (int l = 0; l < L; ++l) { ptrLen[l] = np[l]; idxLen[l] = ni[l]; }
    return (int64_t)np[1];
}

/* ------------------------------------------------------------------------ */
/* create_fbrLikeSlcInds, util.h:1770-1783: slice coordinate of every        */
/* level-1 entry.  out has nFibers entries; the last one is left untouched  */
/* (uninitialised in the reference).                                        */
/* ------------------------------------------------------------------------ */
void orc_fbr_like_slc_inds(uint64_t nSlc, const uint64_t *fbrPtr0,
                           const uint64_t *fbrIdx0, uint64_t *out)
{
    for (uint64_t s = 0; s < nSlc; ++s)
        for (uint64_t f = fbrPtr0[s]; f < fbrPtr0[s + 1]; ++f)
            out[f] = fbrIdx0[s];
}

/* ------------------------------------------------------------------------ */
/* Bin boundaries shared by make_HybBin (util.h:1786-1813) and make_TiledBin */
/* (util.h:1887-1917).  nBin = 10 (util.h:205).                              */
/* ------------------------------------------------------------------------ */
static void bin_bounds(uint64_t totNnz, int64_t *LB, int64_t *UB)
{
    const int64_t TB = 512;
    LB[0] = 0;       UB[0] = 3;
    LB[1] = 2;       UB[1] = 5;
    LB[2] = 4;       UB[2] = 9;
    LB[3] = 8;       UB[3] = 17;
    LB[4] = 16;      UB[4] = 1025;
    LB[5] = 1024;    UB[5] = 4 * TB + 1;
    LB[6] = 4 * TB;  UB[6] = 8 * TB + 1;
    LB[7] = 8 * TB;  UB[7] = 16 * TB + 1;
    LB[8] = 16 * TB; UB[8] = 32 * TB + 1;
    LB[9] = 32 * TB; UB[9] = (int64_t)totNnz + 1;
    UB[0] = 1025;                           /* util.h:1916 / 1812: merged */
}

static int pick_bin(int64_t n, const int64_t *LB, const int64_t *UB)
{
    for (int b = 0; b < 10; ++b)
        if (n > LB[b] && n < UB[b]) return b;
    return -1;                              /* dropped by the reference */
}

/* make_TiledBin, util.h:1887-1952: weight = #level-1 entries + #nnz of the */
/* slice.  binOf[s] = bin or -1.  Returns the number of dropped slices.      */
int64_t orc_tiled_bin(int ndims, uint64_t totNnz, uint64_t nSlc,
                      const uint64_t *const *fbrPtr, int32_t *binOf,
                      int64_t *weight)
{
    int64_t LB[10], UB[10], dropped = 0;
    bin_bounds(totNnz, LB, UB);
    for (uint64_t s = 0; s < nSlc; ++s) {
        /* util.h:1921: the reference accumulates in an int */
        int nnzSlc = 0;
        nnzSlc += (int)(fbrPtr[0][s + 1] - fbrPtr[0][s]);
        for (uint64_t f = fbrPtr[0][s]; f < fbrPtr[0][s + 1]; ++f) {
            if (ndims == 3)
                nnzSlc += (int)(fbrPtr[1][f + 1] - fbrPtr[1][f]);
            else
                for (uint64_t g = fbrPtr[1][f]; g < fbrPtr[1][f + 1]; ++g)
                    nnzSlc += (int)(fbrPtr[2][g + 1] - fbrPtr[2][g]);
        }
        weight[s] = nnzSlc;
        binOf[s] = pick_bin(nnzSlc, LB, UB);
        if (binOf[s] < 0) dropped++;
    }
    return dropped;
}

/* make_HybBin, util.h:1786-1885: weight = #nnz for CSF slices; #fibers     */
/* (= #nnz) for CSL slices.                                                  */
int64_t orc_hyb_bin(int ndims, uint64_t totNnz, uint64_t nSlc,
                    const uint64_t *const *fbrPtr, int32_t *binOf)
{
    int64_t LB[10], UB[10], dropped = 0;
    bin_bounds(totNnz, LB, UB);
    for (uint64_t s = 0; s < nSlc; ++s) {
        int nnzSlc = 0;
        for (uint64_t f = fbrPtr[0][s]; f < fbrPtr[0][s + 1]; ++f) {
            if (ndims == 3)
                nnzSlc += (int)(fbrPtr[1][f + 1] - fbrPtr[1][f]);
            else
                for (uint64_t g = fbrPtr[1][f]; g < fbrPtr[1][f + 1]; ++g)
                    nnzSlc += (int)(fbrPtr[2][g + 1] - fbrPtr[2][g]);
        }
        binOf[s] = pick_bin(nnzSlc, LB, UB);
        if (binOf[s] < 0) dropped++;
    }
    return dropped;
}

int64_t orc_csl_bin(uint64_t totNnz, uint64_t nCsl, const uint64_t *cslPtr,
                    int32_t *binOf)
{
    int64_t LB[10], UB[10], dropped = 0;
    bin_bounds(totNnz, LB, UB);
    for (uint64_t s = 0; s < nCsl; ++s) {
        int n = (int)(cslPtr[s + 1] - cslPtr[s]);
        binOf[s] = pick_bin(n, LB, UB);
        if (binOf[s] < 0) dropped++;
    }
    return dropped;
}

/* ------------------------------------------------------------------------ */
/* create_HYB (3-D, util.h:852-988) and create_HYB_4D (util.h:990-1142).     */
/* Input: the CSF of create_HCSR (levels in modeOrder order), its leaf       */
/* coordinate array (leaf[p] = coordinate of the last mode) and values, all */
/* in sorted order.  Outputs (caller allocates nnz+2 per array):            */
/*   COO part : cooInd[l][c] (l in level order), cooVal[c]                  */
/*   CSL part : cslPtr (nCsl+1), cslIdx (nCsl), cslInd1 / cslInd2, cslVal    */
/*   CSF part : hPtr[l] / hIdx[l] (lengths in hPtrLen / hIdxLen), hLeaf,     */
/*              hVal                                                        */
/* counts[0..4] = COOnnz, CSLnnz, HCSRnnz, nCsl slices, nFibers (the size   */
/* of the leaf-fiber pointer vector, util.h:980 / 1129).                     */
/* ------------------------------------------------------------------------ */
void orc_create_hyb(int ndims, uint64_t nnz, uint64_t nSlc,
                    const uint64_t *const *fbrPtr, const uint64_t *const *fbrIdx,
                    const uint64_t *leaf, const float *vals,
                    uint64_t *const *cooInd, float *cooVal,
                    uint64_t *cslPtr, uint64_t *cslIdx, uint64_t *cslInd1,
                    uint64_t *cslInd2, float *cslVal,
                    uint64_t *const *hPtr, uint64_t *const *hIdx,
                    uint64_t *hPtrLen, uint64_t *hIdxLen,
                    uint64_t *hLeaf, float *hVal, uint64_t *counts)
{
    (void)nnz;
    const int L = ndims - 1;
    uint64_t coo = 0, csl = 0, ind = 0, nCsl = 0;
    uint64_t usedCOOFbr = 0, usedCSLFbr = 0, usedHCSRFbr = 0;
    uint64_t np[ORC_MAXD] = {0}, ni[ORC_MAXD] = {0};
    for (uint64_t s = 0; s < nSlc; ++s) {
        /* util.h:873-889 / 1011-1030: slice nnz and "all fibers length 1" */
        uint64_t slcNnz = 0;
        int lenOne = 1;
        uint64_t f0 = fbrPtr[0][s], f1 = fbrPtr[0][s + 1];
        uint64_t leafLo, leafHi;     /* leaf-fiber range of the slice */
        if (ndims == 3) { leafLo = f0; leafHi = f1; }
        else { leafLo = fbrPtr[1][f0]; leafHi = fbrPtr[1][f1]; }
        for (uint64_t g = leafLo; g < leafHi; ++g) {
            uint64_t len = fbrPtr[L - 1][g + 1] - fbrPtr[L - 1][g];
            slcNnz += len;
            if (len > 1) lenOne = 0;
        }
        uint64_t x0 = fbrPtr[L - 1][leafLo];
        if (slcNnz == 1) {
            /* util.h:929-942 / 1089-1100 */
            cooInd[0][coo] = fbrIdx[0][s];
            if (ndims == 3) {
                cooInd[1][coo] = fbrIdx[1][f0];
            } else {
                cooInd[1][coo] = fbrIdx[1][f0];
                cooInd[2][coo] = fbrIdx[2][leafLo];
            }
            cooInd[L][coo] = leaf[x0];
            cooVal[coo] = vals[x0];
            coo++;
            usedCOOFbr++;
        } else if (ndims == 3 && lenOne) {
            /* util.h:944-955 */
            cslPtr[nCsl] = fbrPtr[0][s] - (usedCOOFbr + usedHCSRFbr);
            cslIdx[nCsl] = fbrIdx[0][s];
            nCsl++;
            for (uint64_t k = 0; k < slcNnz; ++k) {
                cslInd1[csl + k] = fbrIdx[1][f0 + k];
                cslInd2[csl + k] = leaf[x0 + k];
                cslVal[csl + k] = vals[x0 + k];
            }
            csl += slcNnz;
            usedCSLFbr += f1 - f0;
        } else {
            /* util.h:957-972 / 1102-1122 */
            uint64_t sub = usedCOOFbr + usedCSLFbr;
            hPtr[0][np[0]++] = fbrPtr[0][s] - sub;
            hIdx[0][ni[0]++] = fbrIdx[0][s];
            for (uint64_t f = f0; f < f1; ++f) {
                hPtr[1][np[1]++] = fbrPtr[1][f] - sub;
                hIdx[1][ni[1]++] = fbrIdx[1][f];
                if (ndims == 4)
                    for (uint64_t g = fbrPtr[1][f]; g < fbrPtr[1][f + 1]; ++g) {
                        hPtr[2][np[2]++] = fbrPtr[2][g] - sub;
                        hIdx[2][ni[2]++] = fbrIdx[2][g];
                    }
            }
            for (uint64_t k = 0; k < slcNnz; ++k) {
                hLeaf[ind + k] = leaf[x0 + k];
                hVal[ind + k] = vals[x0 + k];
            }
            usedHCSRFbr += f1 - f0;
            ind += slcNnz;
        }
    }
    /* util.h:976-978 / 1125-1127: closing entries */
    hPtr[L - 1][np[L - 1]++] = ind;
    for (int i = L - 2; i > -1; --i) {
        hPtr[i][np[i]] = np[i + 1] - 1;
        np[i]++;
    }
    cslPtr[nCsl] = csl;
    for (int l = 0; l < L; ++l) { hPtrLen[l] = np[l]; hIdxLen[l] = ni[l]; }
    counts[0] = coo;
    counts[1] = csl;
    counts[2] = ind;
    counts[3] = nCsl;
    counts[4] = np[L - 1];   /* nFibers: leaf-fiber pointer count */
}

/* ------------------------------------------------------------------------ */
/* MTTKRP_COO_CPU / MTTKRP_COO_CPU_4D, mttkrp_cpu.cpp:60-113.                */
/* Cyclic role map curMode[m] = (m + mode) % N; the output U[mode] is       */
/* accumulated in place, sequentially, in float, in stored order:           */
/*   tmp = v * U1[j][r] * U2[k][r] (* U3[l][r]);  out[i][r] += tmp.         */
/* U[m] is row-major dims[m] x R.                                            */
/* ------------------------------------------------------------------------ */
void orc_mttkrp_coo_f32(int ndims, uint64_t nnz, const uint64_t *const *inds,
                        const float *vals, int mode, uint64_t R,
                        const float *const *U, float *out)
{
    int cm[ORC_MAXD];
    for (int m = 0; m < ndims; ++m) cm[m] = (m + mode) % ndims;
    for (uint64_t x = 0; x < nnz; ++x) {
        uint64_t i0 = inds[cm[0]][x], i1 = inds[cm[1]][x], i2 = inds[cm[2]][x];
        const float *a = U[cm[1]] + i1 * R, *b = U[cm[2]] + i2 * R;
        float *o = out + i0 * R;
        if (ndims == 3) {
            for (uint64_t r = 0; r < R; ++r) {
                float t = vals[x] * a[r] * b[r];
                o[r] += t;
            }
        } else {
            const float *c = U[cm[3]] + inds[cm[3]][x] * R;
            for (uint64_t r = 0; r < R; ++r) {
                float t = vals[x] * a[r] * b[r] * c[r];
                o[r] += t;
            }
        }
    }
}

/* Same contraction accumulated in double from double copies of the float */
/* inputs: the float64 oracle every result is scored against.             */
void orc_mttkrp_coo_f64(int ndims, uint64_t nnz, const uint64_t *const *inds,
                        const float *vals, int mode, uint64_t R,
                        const float *const *U, double *out)
{
    int cm[ORC_MAXD];
    for (int m = 0; m < ndims; ++m) cm[m] = (m + mode) % ndims;
    for (uint64_t x = 0; x < nnz; ++x) {
        const float *a = U[cm[1]] + inds[cm[1]][x] * R;
        const float *b = U[cm[2]] + inds[cm[2]][x] * R;
        const float *c = ndims == 4 ? U[cm[3]] + inds[cm[3]][x] * R : 0;
        double *o = out + inds[cm[0]][x] * R;
        double v = (double)vals[x];
        for (uint64_t r = 0; r < R; ++r) {
            double t = v * (double)a[r] * (double)b[r];
            if (c) t *= (double)c[r];
            o[r] += t;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Float64 MTTKRP split over host threads, for tensors of 1e7-1e8 nonzeros. */
/* Same contraction and role map as orc_mttkrp_coo_f64 (mttkrp_cpu.cpp:     */
/* 60-113 in float64): thread t owns a contiguous range of output rows with */
/* about nnz/T nonzeros and scans the nonzeros in stored order, keeping     */
/* those of its rows, so every row is summed in the same order as the       */
/* sequential function and the result is identical to it bit for bit.       */
/* Coordinates are uint32 (the device layout); values are float (vf) or     */
/* double (vd, e.g. residuals); factors are double row-major dims[m] x R.   */
/* ------------------------------------------------------------------------ */
#include <pthread.h>

typedef struct {
    int ndims, mode;
    uint64_t nnz, R, lo, hi;
    const uint32_t *const *inds;
    const float *vf;
    const double *vd;
    const double *const *U;
    double *out;
} orc_mt_job;

static void *orc_mttkrp_f64_rows(void *arg)
{
    const orc_mt_job *j = (const orc_mt_job *)arg;
    int cm[ORC_MAXD];
    for (int m = 0; m < j->ndims; ++m) cm[m] = (m + j->mode) % j->ndims;
    const uint32_t *row = j->inds[cm[0]];
    const uint64_t R = j->R;
    for (uint64_t x = 0; x < j->nnz; ++x) {
        const uint64_t i = row[x];
        if (i < j->lo || i >= j->hi) continue;
        const double *a = j->U[cm[1]] + (uint64_t)j->inds[cm[1]][x] * R;
        const double *b = j->U[cm[2]] + (uint64_t)j->inds[cm[2]][x] * R;
        const double *c = j->ndims == 4 ? j->U[cm[3]] + (uint64_t)j->inds[cm[3]][x] * R : 0;
        double *o = j->out + i * R;
        const double v = j->vf ? (double)j->vf[x] : j->vd[x];
        for (uint64_t r = 0; r < R; ++r) {
            double t = v * a[r] * b[r];
            if (c) t *= c[r];
            o[r] += t;
        }
    }
    return 0;
}

/* out (dims[mode] x R) is overwritten.  Returns 0, or -1 if a thread could */
/* not be started or memory ran out.                                        */
int orc_mttkrp_f64_mt(int ndims, uint64_t nnz, const uint32_t *const *inds,
                      const float *vf, const double *vd, int mode,
                      uint64_t rows, uint64_t R, const double *const *U,
                      double *out, int nthreads)
{
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 64) nthreads = 64;
    memset(out, 0, rows * R * sizeof(double));
    const uint32_t *row = inds[(0 + mode) % ndims];
    uint64_t *cnt = (uint64_t *)calloc(rows + 1, sizeof(uint64_t));
    if (!cnt) return -1;
    for (uint64_t x = 0; x < nnz; ++x) cnt[row[x]]++;
    /* split points: thread t takes rows up to the first prefix >= (t+1)/T */
    uint64_t bound[65];
    bound[0] = 0;
    int t = 0;
    uint64_t acc = 0;
    for (uint64_t i = 0; i < rows && t < nthreads - 1; ++i) {
        acc += cnt[i];
        while (t < nthreads - 1 && acc * (uint64_t)nthreads >= (uint64_t)(t + 1) * nnz) bound[++t] = i + 1;
    }
    while (t < nthreads - 1) { bound[t + 1] = bound[t]; ++t; }
    bound[nthreads] = rows;
    free(cnt);
    orc_mt_job job[64];
    pthread_t th[64];
    int started[64] = {0};
    int rc = 0;
    for (int k = 0; k < nthreads; ++k) {
        job[k] = (orc_mt_job){ndims, mode, nnz, R, bound[k], bound[k + 1], inds, vf, vd, U, out};
        if (bound[k] == bound[k + 1]) continue;
        if (pthread_create(&th[k], 0, orc_mttkrp_f64_rows, &job[k]) != 0) { rc = -1; break; }
        started[k] = 1;
    }
    for (int k = 0; k < nthreads; ++k)
        if (started[k]) pthread_join(th[k], 0);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* CP-completion residuals in float64 over host threads (no reference       */
/* counterpart; the loss is the least-squares CP loss over the observed     */
/* entries):  e_x = sum_r prod_n U_n[i_n(x)][r] - v_x,  returns 1/2 sum e^2 */
/* summed per thread over contiguous nonzero ranges, then in thread order.  */
/* ------------------------------------------------------------------------ */
typedef struct {
    int ndims;
    uint64_t lo, hi, R;
    const uint32_t *const *inds;
    const float *v;
    const double *const *U;
    double *e;
    double sq;
} orc_res_job;

static void *orc_resid_range(void *arg)
{
    orc_res_job *j = (orc_res_job *)arg;
    double sq = 0.0;
    for (uint64_t x = j->lo; x < j->hi; ++x) {
        double m = 0.0;
        for (uint64_t r = 0; r < j->R; ++r) {
            double p = 1.0;
            for (int n = 0; n < j->ndims; ++n) p *= j->U[n][(uint64_t)j->inds[n][x] * j->R + r];
            m += p;
        }
        const double e = m - (double)j->v[x];
        if (j->e) j->e[x] = e;
        sq += e * e;
    }
    j->sq = sq;
    return 0;
}

int orc_cp_resid_f64_mt(int ndims, uint64_t nnz, const uint32_t *const *inds,
                        const float *vals, uint64_t R, const double *const *U,
                        double *e, double *loss, int nthreads)
{
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 64) nthreads = 64;
    orc_res_job job[64];
    pthread_t th[64];
    int started[64] = {0};
    int rc = 0;
    for (int k = 0; k < nthreads; ++k) {
        job[k] = (orc_res_job){ndims, nnz * (uint64_t)k / (uint64_t)nthreads,
                               nnz * (uint64_t)(k + 1) / (uint64_t)nthreads, R, inds, vals, U, e, 0.0};
        if (pthread_create(&th[k], 0, orc_resid_range, &job[k]) != 0) { rc = -1; break; }
        started[k] = 1;
    }
    double sq = 0.0;
    for (int k = 0; k < nthreads; ++k)
        if (started[k]) { pthread_join(th[k], 0); sq += job[k].sq; }
    *loss = 0.5 * sq;
    return rc;
}

/* ------------------------------------------------------------------------ */
/* MM-CSF mixed-mode partition: mm_partition_reuseBased, util.h:2167-2512,  */
/* with the heuristic's own choice instead of the override `mode = 2` of    */
/* util.h:2517, executed sequentially in index order (the reference runs the */
/* loop under OpenMP with unsynchronised decrements, util.h:2500-2512).     */
/*                                                                          */
/* inds: the tensor in the order the reference's loop walks it (after the   */
/* sorts of every non-redundant mode, the last one's order).  For every      */
/* mode m != redunMode: sp[m] (dims[order_m[0]] + 1 entries) = denseSlcPtr   */
/* (util.h:2203-2209, including its end pointer of the last slice, which is  */
/* the slice's start), f1[m] = fbrIdx[1] (level-1 coordinates), cnt[m] =     */
/* nnzPerFiber (counts per level-1 entry; decremented here, util.h:2500),    */
/* order[m] = modeOrder of arrX[m].  fbTh = -h.  Out: part[x].               */
/* ------------------------------------------------------------------------ */
static int64_t orc_bsearch1(const uint64_t *arr, int64_t left, int64_t right, uint64_t value)
{
    /* binarySearch1, util.h:2128-2141 (inclusive right bound) */
    while (left <= right) {
        int64_t middle = (int64_t)(((uint32_t)left + (uint32_t)right) >> 1);
        if (arr[middle] == value) return middle;
        else if (arr[middle] > value) right = middle - 1;
        else left = middle + 1;
    }
    return -1;
}

int orc_mm_partition(int ndims, uint64_t nnz, const uint64_t *const *inds, const uint64_t *dims,
                     int redunMode, int fbTh, int loadMode, const int *const *order, const uint64_t *const *sp,
                     const uint64_t *const *f1, uint64_t *const *cnt, uint64_t *part)
{
    int N = ndims;
    /* same-fiber pairs, shortMode / longMode (util.h:2261-2310): the last
     * detected pair wins */
    int same[4][4] = {{0}};
    int shortMode = 0, longMode = -1;
    for (int m = 1; m < N; ++m) {
        for (int q = 0; q < m; ++q) {
            if (order[m][1] == order[q][0] && order[m][0] == order[q][1]) {
                same[q][m] = 1;
                shortMode = dims[q] <= dims[m] ? q : m;
                longMode = dims[q] <= dims[m] ? m : q;
                break;
            }
        }
    }
    /* the reference tests m-2 before m-1 (m = 2) and m-3, m-2, m-1 (m = 3):
     * the loop above scans q ascending and stops at the first match, which
     * is that order */
    /* shortestMode (util.h:2253): X.modeOrder of the loaded tensor is
     * order_tensormode(X, -m) (load_tensor, util.h:367) */
    int o0[ORC_MAXD];
    orc_order_tensormode(N, dims, loadMode, 0, o0);
    int shortestMode = dims[o0[0]] <= dims[o0[1]] ? o0[0] : o0[1];
    uint64_t fbrNnz[ORC_MAXD], fbrNo[ORC_MAXD];
    int copied[ORC_MAXD];
    for (uint64_t x = 0; x < nnz; ++x) {
        for (int m = 0; m < N; ++m) {
            int copyFrom = -1;
            copied[m] = 0;
            for (int q = 0; q < m; ++q)
                if (same[q][m]) { copyFrom = q; break; }
            if (copyFrom >= 0) {
                /* util.h:2372-2383: the second mode of a pair copies the first */
                fbrNnz[m] = fbrNnz[copyFrom];
                fbrNo[m] = 99999;
                copied[m] = 1;
            } else {
                uint64_t slc = inds[m][x];
                uint64_t j = inds[order[m][1]][x];
                int64_t n = (int64_t)(sp[m][slc + 1] - sp[m][slc]);
                uint64_t fbr = sp[m][slc];
                int64_t r = orc_bsearch1(f1[m] + fbr, 0, n, j);
                uint64_t at = (uint64_t)((int64_t)fbr + r);
                fbrNo[m] = at;
                fbrNnz[m] = cnt[m][at];
            }
        }
        int modeDone = 0, mode = 0;
        uint64_t th = (uint64_t)(int64_t)fbTh;
        for (int m = 0; m < N && !modeDone; ++m) {
            uint64_t mx = 0;
            for (int q = 0; q < N; ++q)
                if (q != m && fbrNnz[q] > mx) mx = fbrNnz[q];
            if (fbrNnz[m] >= th * mx) {
                modeDone = 1;
                int pairAbove = 0;
                for (int q = m + 1; q < N; ++q) pairAbove |= same[m][q];
                /* util.h:2399-2435: a mode starting a same-fiber pair takes shortMode */
                mode = pairAbove ? shortMode : m;
            }
        }
        if (!modeDone) mode = shortestMode;
        /* util.h:2493-2512; a mode whose count was copied from its pair has
         * no fiber of its own (fbrNo 99999): no decrement (the reference's
         * write there is out of range unless the mode is skipped anyway) */
        for (int i = 0; i < N; ++i) {
            if (copied[i]) continue;
            if (mode == shortMode) {
                if (i != shortMode && i != longMode) cnt[i][fbrNo[i]]--;
            } else {
                if (i != redunMode && i != mode) cnt[i][fbrNo[i]]--;
            }
        }
        part[x] = (uint64_t)mode;
    }
    return 0;
}
