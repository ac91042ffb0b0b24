/*
 * mmcsf.h -- C ABI of the MI355X (gfx950) sparse-MTTKRP library.
 *
 * Drop-in boundary for the MM-CSF reference's MTTKRP path.  Every entry point
 * names the reference function it replaces (paths relative to the reference
 * checkout).  Conventions:
 *   - every function returns int: 0 = ok, negative = error; the message of the
 *     last error on the calling thread is mmcsf_last_error();
 *   - nothing calls exit() and nothing mutates caller-owned inputs;
 *   - device pointers are HIP device addresses, streams are hipStream_t passed
 *     as void* (NULL = the null stream); every device call is stream-ordered;
 *   - a handle may be used from one host thread at a time;
 *   - tensor coordinates are 0-based; device index arrays are uint32 (every
 *     count and coordinate of a supported tensor is < 2^31); host-side
 *     exports are widened to uint64 like the reference's ITYPE (size_t).
 */
#ifndef MMCSF_H
#define MMCSF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes --------------------------------------------------------- */
#define MMCSF_OK 0
#define MMCSF_ERR_ARG (-1)        /* bad argument / option */
#define MMCSF_ERR_HIP (-2)        /* HIP runtime error */
#define MMCSF_ERR_IO (-3)         /* file could not be read / written */
#define MMCSF_ERR_UNSUPPORTED (-4)/* valid request outside the supported range */
#define MMCSF_ERR_NOMEM (-5)      /* device or host allocation failed */

/* ---- format types: the reference's -t values (mttkrp.cu:104-417) ---------- */
#define MMCSF_TYPE_COO_CPU 1      /* -t 1  COO on the host (sequential float) */
#define MMCSF_TYPE_COO 3          /* -t 3  COO on the GPU */
#define MMCSF_TYPE_BCSF 8         /* -t 8  B-CSF (tiled HCSR + slice bins) */
#define MMCSF_TYPE_BCSF_ANYMODE 9 /* -t 9  the B-CSF of -t 8, every mode from it
                                     (MTTKRP_B_HCSR_GPU_ANYMODE, mttkrp.cu:223-231) */
#define MMCSF_TYPE_HBCSF 10       /* -t 10 HB-CSF (COO + CSL + CSF parts) */
#define MMCSF_TYPE_MMCSF 12       /* -t 12 MM-CSF, every mode from one build */
#define MMCSF_TYPE_SINGLECSF 13   /* -t 13 single untiled CSF, every mode from it
                                     (mttkrp.cu:234-262; the reference computes it on
                                     the host with MTTKRP_HCSR_CPU[_4D] / _mode1 /
                                     _mode2, mttkrp_cpu.cpp:160-348, here on the GPU) */
#define MMCSF_TYPE_ONECSF 14      /* -t 14 single CSF on the GPU, every mode from it
                                     (MTTKRP_ONE_HCSR_GPU, mttkrp.cu:234-299) */

/* ---- factor storage --------------------------------------------------------- */
#define MMCSF_F32 0               /* float32 rows */
#define MMCSF_BF16 1              /* bfloat16 rows (uint16 bit patterns), FP32 math */

/*
 * Options -- the reference's Options class (util.h:189-236), same defaults.
 * Parsed by mmcsf_options_parse with the grammar of parse_cmd_options
 * (util.h:3341-3442).
 */
typedef struct mmcsf_options {
    uint64_t R;               /* -R rank (32) */
    uint64_t mode;            /* -m MTTKRP mode (0) */
    uint64_t impType;         /* -t implementation type (1) */
    uint64_t warpPerSlice;    /* -w (4) */
    uint64_t nTile;           /* -l (1; tiling is forced off, util.h:724-727) */
    uint64_t gridSize;        /* -g (512) */
    uint64_t TBsize;          /* -b (128) */
    uint64_t MIfbTh;          /* -h (1) */
    uint64_t fiberPerWarp;    /* -s (1) */
    int verbose;              /* -v 1 */
    int correctness;          /* -c 1 */
    uint64_t nBin;            /* 10 (no flag) */
    uint64_t cpdIters;        /* -n CP-ALS iterations (10; no flag in the reference) */
    int natOrdering;          /* 0 (no flag) */
    uint64_t fbrThreashold;   /* -f fiber split threshold (99999999) */
    int factorDtype;          /* -d 0|1: FP32 / BF16 factor storage (new flag) */
    uint64_t seed;            /* -x seed of generated inputs (new flag, 1) */
    int modeTrees;            /* -a 1|0 (new flag, 1): MM-CSF also keeps a root tree per
                                 other mode, so every mode's MTTKRP is a root traversal;
                                 0 = single tree, non-root modes scatter (the reference) */
    int mixedPart;            /* -M 1|0 (new flag, 0): MM-CSF partitions the nonzeros by the
                                 reference's heuristic (mm_partition_reuseBased,
                                 util.h:2386-2512, threshold -h) instead of its override
                                 to partition 2 (util.h:2517); no extra mode trees */
    char inFileName[4096];    /* -i: a .tns file or a format cache (mmcsf_format_save) */
    char outFileName[4096];   /* -o */
    char cacheFileName[4096]; /* -k (new flag): write the built format to this cache file */
    char m0[16], m1[16], m2[16];  /* -p/-q/-r, unused like the reference */
} mmcsf_options;

/* Opaque handles. */
typedef struct mmcsf_tensor mmcsf_tensor;  /* device COO: class Tensor, util.h:80-108 */
typedef struct mmcsf_format mmcsf_format;  /* device COO / B-CSF / HB-CSF / MM-CSF:
                                              TiledTensor + HYBTensor, util.h:110-173 */

/* Last error message of the calling thread ("" if none). */
const char *mmcsf_last_error(void);

/* Options() defaults, util.h:189-211. */
int mmcsf_options_default(mmcsf_options *opt);

/* parse_cmd_options, util.h:3341-3442: "-X value" pairs, parsing stops at the
 * first token not starting with '-'.  Returns MMCSF_ERR_ARG (instead of the
 * reference's help-and-exit(1)) on an unknown flag or a flag without value. */
int mmcsf_options_parse(int argc, char **argv, mmcsf_options *opt);

/* ---- tensors ---------------------------------------------------------------- */

/* load_tensor, util.h:313-370: .tns text (ndims, dims, then 1-based
 * coordinates and a value per line) -> device COO with 0-based coordinates,
 * in file order.  The text is parsed by up to 16 host threads.  Errors
 * (instead of exit(0)) on a missing file, a malformed line, a coordinate
 * outside 1..dims[m] or ndims outside 3..4. */
int mmcsf_tensor_load_tns(const char *path, mmcsf_tensor **out);

/* Writes a device COO as .tns text (1-based coordinates, the shortest decimal
 * that reads back to the same float).  No reference counterpart (the
 * reference only reads .tns). */
int mmcsf_tensor_save_tns(const mmcsf_tensor *t, const char *path);

/* Binary COO cache (no reference counterpart): a 56-byte header ("MMCSFCOO",
 * version 1, ndims, nnz, dims[4], little-endian), then ndims x nnz uint32
 * 0-based coordinates (mode-major), then nnz float values.  Loading skips
 * text parsing entirely; the tensor equals the one that was saved. */
int mmcsf_tensor_save_bin(const mmcsf_tensor *t, const char *path);
int mmcsf_tensor_load_bin(const char *path, mmcsf_tensor **out);

/* Device COO from host arrays in the reference layout: inds[m][x] (size_t
 * per mode, as Tensor::inds, util.h:89) and vals[x] (float). */
int mmcsf_tensor_from_host(int ndims, const uint64_t *dims, uint64_t nnz,
                           const uint64_t *const *inds, const float *vals,
                           mmcsf_tensor **out);

/* Device COO from device arrays (copied; uint32 coordinates). */
int mmcsf_tensor_from_device(int ndims, const uint64_t *dims, uint64_t nnz,
                             const uint32_t *const *inds_dev,
                             const float *vals_dev, void *stream,
                             mmcsf_tensor **out);

/* Synthetic tensor generated on the GPU (no reference counterpart; the
 * reference reads FROSTT files).  law[m]: 0 = uniform, 1 = Zipf(alpha[m])
 * with hot rows at low indices.  Coordinates are distinct (accepted in draw
 * order, repeats rejected).  Values U(0,1) float32.  Deterministic in seed. */
int mmcsf_tensor_generate(int ndims, const uint64_t *dims, uint64_t nnz,
                          const int *law, const double *alpha, uint64_t seed,
                          void *stream, mmcsf_tensor **out);

int mmcsf_tensor_info(const mmcsf_tensor *t, int *ndims, uint64_t *dims,
                      uint64_t *nnz);
/* Device views of the COO arrays (owned by the handle). */
int mmcsf_tensor_device_arrays(const mmcsf_tensor *t, const uint32_t **inds,
                               const float **vals);
int mmcsf_tensor_copy_to_host(const mmcsf_tensor *t, uint64_t *const *inds,
                              float *vals);
/* Stream-ordered copy of the COO into caller device buffers (uint32
 * coordinates per mode, float values); NULL entries are skipped. */
int mmcsf_tensor_copy_to_device(const mmcsf_tensor *t, uint32_t *const *inds,
                                float *vals, void *stream);
void mmcsf_tensor_free(mmcsf_tensor *t);

/* ---- formats (GPU builder) -------------------------------------------------- */

/*
 * Builds a device format with GPU sort / scan / segment passes.  Reference
 * flows (mttkrp.cu):
 *   MMCSF_TYPE_COO   sort_COOtensor (util.h:412-475)
 *   MMCSF_TYPE_BCSF  create_HCSR + make_KTiling + create_TiledHCSR +
 *                    make_TiledBin + create_fbrLikeSlcInds (mttkrp.cu:153-219);
 *                    MMCSF_TYPE_BCSF_ANYMODE and MMCSF_TYPE_ONECSF build the
 *                    same arrays (mttkrp.cu:153-231, 234-299)
 *   MMCSF_TYPE_SINGLECSF sort_COOtensor + create_HCSR + create_fbrLikeSlcInds
 *                    (mttkrp.cu:234-262): the arrays of MMCSF_TYPE_BCSF
 *                    without the slice bins
 *   MMCSF_TYPE_HBCSF create_HCSR + create_HYB[_4D] + make_HybBin
 *                    (mttkrp.cu:138-150, util.h:852-1142, 1786-1885)
 *   MMCSF_TYPE_MMCSF mm_partition_reuseBased + populate_paritions +
 *                    sort_MI_CSF + create_TiledHCSR + create_fbrLikeSlcInds
 *                    (mttkrp.cu:289-395, util.h:2088-2526)
 * Uses opt->mode (root mode of COO/B-CSF/HB-CSF) and opt->fbrThreashold.
 * Index/pointer arrays equal the reference's bit for bit (see
 * mmcsf_format_export); the order of nonzeros inside a fiber is the
 * canonical full-key order, one of the orders the reference may produce.
 */
int mmcsf_format_build(const mmcsf_tensor *t, int type,
                       const mmcsf_options *opt, void *stream,
                       mmcsf_format **out);

/* Summary of a built format. */
typedef struct mmcsf_format_info {
    int type;
    int ndims;
    uint64_t dims[4];
    uint64_t totNnz;
    int modeOrder[4];          /* of the (non-empty) partition / the format */
    uint64_t nLevel[4];        /* entries per CSF level (level ndims-1 = nnz) */
    uint64_t nFibers;          /* reference convention: fbrPtr[1].size() */
    uint64_t droppedSlices;    /* slices the reference's binning drops */
    uint64_t COOnnz, CSLnnz, HCSRnnz;  /* HB-CSF parts */
    int redunMode;             /* mttkrp.cu:315-347 */
    int modeTrees;             /* internal root trees held besides the MM-CSF tree */
    uint64_t modeLevel[4][4];  /* MM-CSF: entries per level of the tree the MTTKRP of
                                  mode m traverses (its root tree, else the MM-CSF tree) */
    uint64_t deviceBytes;      /* device memory held by the format (every tree, its
                                  pointer / index / value arrays, start bitmasks, chunk
                                  costs, bins and hot-row tables) */
    uint64_t partNnz[4];       /* MM-CSF: nonzeros per partition (X.totnnzPerPart) */
    int mixedPart;             /* MM-CSF built with -M 1 */
    int modeTreeOrder[4][4];   /* MM-CSF and one-tree types: the mode order of the tree
                                  the MTTKRP of mode m traverses (level l holds mode
                                  modeTreeOrder[m][l]; its root is m when that tree is
                                  a root tree); -M 1: of partition m, as modeLevel */
} mmcsf_format_info;

int mmcsf_format_info_get(const mmcsf_format *f, mmcsf_format_info *info);

/*
 * Copies one reference field to host memory, widened to uint64 (float fields
 * are written as float).  `name` is the reference's member name; `k` selects
 * the level / bin / mode / partition where the member is a vector of
 * vectors.  Fields (sizes follow the reference's vectors exactly):
 *   "fbrPtr" k, "fbrIdx" k (level k; the leaf-fiber level carries the
 *   reference's trailing entry), "inds" k (coordinates of tensor mode k in
 *   stored order), "vals" (float), "fbrLikeSlcInds" (nFibers entries, the
 *   last one unspecified), "slcMapperBin" k, "partPerNnz",
 *   HB-CSF: "COOinds" k (tensor mode k), "COOvals", "CSLslicePtr",
 *   "CSLsliceIdx", "CSLinds" k (k = mode1 / mode2 of modeOrder),
 *   "CSLvals", "CSLslcMapperBin" k.
 *   MM-CSF built with -M 1 (k = 8 * partition + level / mode):
 *   "partPerNnz" (the partition of every nonzero, in the order the
 *   reference's partition loop walks the tensor), "partOrder" (that order as
 *   positions of the tensor the format was built from), "P.modeOrder",
 *   "P.fbrPtr", "P.fbrIdx", "P.inds", "P.vals", "P.fbrLikeSlcInds".
 * *len receives the element count; nothing is written when cap < *len.
 */
int mmcsf_format_export(const mmcsf_format *f, const char *name, int k,
                        void *dst, uint64_t cap, uint64_t *len);

void mmcsf_format_free(mmcsf_format *f);

/*
 * Binary cache of a built format (no reference counterpart; the reference
 * rebuilds every format from the .tns text on each run, util.h:313-370 and
 * mttkrp.cu:104-395).  mmcsf_format_save writes every array of every tree the
 * format holds (the reference's arrays, the per-mode root trees, the
 * mixed-mode partitions, start bitmasks, chunk costs, bins, hot-row tables)
 * after an 8-byte magic "MMCSFFMT", a uint32 version (1) and a byte-order
 * mark; mmcsf_format_load rebuilds the format from it, equal bit for bit to
 * the one saved, without reading text, sorting or building.  Errors
 * (MMCSF_ERR_IO) on a missing file, another version or byte order, a
 * truncated file or a bad end mark.
 */
int mmcsf_format_save(const mmcsf_format *f, const char *path);
int mmcsf_format_load(const char *path, mmcsf_format **out);
/* 1 if the file at path starts with the format-cache magic, else 0. */
int mmcsf_format_is_cache(const char *path);
/* The nonzeros a format stores, as a new device COO tensor (in the format's
 * stored order; HB-CSF parts and mixed-mode partitions one after another):
 * the tensor a cached format was built from, up to the order of nonzeros.
 * No reference counterpart (the reference keeps its Tensor beside the
 * formats). */
int mmcsf_format_to_tensor(const mmcsf_format *f, mmcsf_tensor **out);

/* ---- MTTKRP ------------------------------------------------------------------ */

/*
 * out = MTTKRP of the format on tensor mode `mode`:
 *   out[i][r] = sum over nonzeros x with coordinate i in `mode` of
 *               v_x * prod_{m != mode} U[m][i_m(x)][r].
 * factors[m]: device pointer to U[m], row-major dims[m] x R, float32 or
 * bfloat16 per `dtype`; factors[mode] is not read.  out: device float32
 * dims[mode] x R, overwritten.  Accumulation is float32.
 * Replaces (by format type):
 *   COO    MTTKRP_COO_GPU            (mttkrp_gpu.h:61, mttkrp_gpu.cu:1562)
 *   B-CSF  MTTKRP_B_HCSR_GPU         (mttkrp_gpu.h:67, mttkrp_gpu.cu:1850)
 *   HB-CSF MTTKRP_HYB_GPU            (mttkrp_gpu.h:73, mttkrp_gpu.cu:2530)
 *   MM-CSF MTTKRP_MIHCSR_GPU[_oneMode_forCPD]
 *                                    (mttkrp_gpu.h:75-81, mttkrp_gpu.cu:3380, 3860)
 * COO, B-CSF and HB-CSF serve their root mode (modeOrder[0]) only, like
 * the reference; MM-CSF and the one-tree types 9 / 13 / 14 serve every mode
 * (types 9 / 13 / 14: the root mode walks the tree, the other modes scatter
 * from their level, mttkrp_gpu.cu:2131-2528, 2853-3378, and for type 13 the
 * host routines MTTKRP_HCSR_CPU[_4D] / _mode1 / _mode2, mttkrp_cpu.cpp:160-348,
 * whose 3-D-only mode1 / mode2 roles this generalises to 4-D: every mode of a
 * 4-D type-13 format is its exact MTTKRP).
 */
int mmcsf_mttkrp(const mmcsf_format *f, int mode, int R, int dtype,
                 const void *const *factors, float *out, void *stream);

/*
 * All-mode MTTKRP of an MM-CSF or one-tree (types 9 / 13 / 14) format
 * (MTTKRP_MIHCSR_GPU's mode loop, mttkrp_gpu.cu:3548-3744; the mode loops of
 * mttkrp.cu:225-230, 244-298): outs[m] receives the MTTKRP of mode m for every
 * m in 0..ndims-1, all from the same factors (type 13: mttkrp.cu:244-262).  Optional ms_per_mode (host,
 * ndims entries) receives each mode's MTTKRP kernel time (after the output
 * has been zeroed), measured with HIP events on `stream`; the call then
 * synchronises on the last event (NULL: no timing, no synchronisation).
 */
int mmcsf_mttkrp_allmode(const mmcsf_format *f, int R, int dtype,
                         const void *const *factors, float *const *outs,
                         void *stream, float *ms_per_mode);

/* Sequential float32 COO MTTKRP on the host (-t 1, MTTKRP_COO_CPU[_4D],
 * mttkrp_cpu.cpp:60-113) over a format of type MMCSF_TYPE_COO_CPU or
 * MMCSF_TYPE_COO (the nonzeros sorted as sort_COOtensor leaves them,
 * mttkrp.cu:79): cyclic mode roles, stored order, accumulated into out (not
 * zeroed).  Host factor and output pointers.  This is the reference's CPU
 * type, selected explicitly; no GPU entry point ever falls back to it. */
int mmcsf_mttkrp_coo_host(const mmcsf_format *f, int mode, int R,
                          const float *const *factors_host, float *out_host);

/* ---- TTM ---------------------------------------------------------------------- */

/*
 * Tensor-times-matrix on the leaf mode of a B-CSF-family format (types 8, 9,
 * 13, 14: the tree of create_HCSR with the -f split, ttm.cu:76-78) -- TTM_CPU
 * (ttm_cpu.cpp:60-86) and TTM_GPU (ttm_gpu.cu:140-259):
 *     Y[f][r] = sum over the nonzeros x of leaf fiber f of
 *               v_x * U[modeOrder[ndims-1]][leaf_x][r].
 * factor: device U of the leaf mode (dims[modeOrder[ndims-1]] x R, float32 or
 * bfloat16 per dtype, R % 4 == 0, 16-byte aligned, the table below 4 GiB);
 * Y: device float32 with
 * one row per leaf fiber (format_info.nLevel[ndims-2] rows; the reference's
 * semiSpTensor has one more row, util.h:3175-3180, that no fiber writes),
 * overwritten. */
int mmcsf_ttm(const mmcsf_format *f, int R, int dtype, const void *factor,
              float *Y, void *stream);

/* TTM_CPU (ttm_cpu.cpp:60-86): the same product on the host, sequential
 * float32 per fiber, for the CLI's -t 1.  Host factor and output. */
int mmcsf_ttm_host(const mmcsf_format *f, int R, const float *factor_host,
                   float *Y_host);

/* ---- CP-ALS ------------------------------------------------------------------- */

typedef struct mmcsf_cpd_stats {
    int iterations;            /* ALS iterations run */
    double fit;                /* fit after the last iteration */
    double fit_history[64];    /* fit of iterations 0..63 */
    double ms_total;           /* ALS loop, host clock */
    double ms_mttkrp;          /* MTTKRP kernels, HIP events */
    int lu_fallbacks;          /* solves that fell back from Cholesky to LU */
} mmcsf_cpd_stats;

/*
 * CP-ALS on an MM-CSF format, entirely on the device (cpd(), cpd_cpu.h:62,
 * cpd_cpu.cu:484-691, with init_GPU and MTTKRP_MIHCSR_GPU_oneMode_forCPD,
 * mttkrp_gpu.h:77-81).  factors[m]: device float32 dims[m] x R, read as the
 * initial guess and overwritten with the final column-normalised factors;
 * lambda (host, R entries) receives the weights (GetFinalLambda).  Per
 * iteration and mode: MTTKRP (float32 or bfloat16 factor copies per
 * mttkrp_dtype), float64 Hadamard of the other Grams, float64 Cholesky
 * inverse (LU fallback), float64 row solve, column 2-norm (iteration 0) or
 * max-norm, MFMA Gram of the new factor.  The fit (float64) is
 * 1 - |X - [[lambda; U]]|_F / |X|_F; the loop stops after max_iters or when
 * the fit changes by less than tol.  A normal matrix that neither Cholesky nor
 * LU can factor fails the call (MMCSF_ERR_UNSUPPORTED, "singular").  The host
 * reads the fits after every iteration when tol > 0; with tol <= 0 (no early
 * stop) it reads them once per 64 iterations, and a singular solve is then
 * reported at the end of its batch.
 */
int mmcsf_cpd_als(const mmcsf_format *f, int R, int max_iters, double tol,
                  int mttkrp_dtype, float *const *factors, double *lambda,
                  mmcsf_cpd_stats *stats, void *stream);

/*
 * CP-ALS exactly as mmcsf_cpd_als, allocating no device memory.  No
 * counterpart in the reference: for callers that cannot allocate (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The work space -- the float64 Grams, normal-matrix work space and inverse,
 * lambda / norm vectors, the MTTKRP output of the largest mode (rows x R
 * float32) and, for bfloat16 MTTKRP, the bfloat16 factor copies, each part
 * rounded up to 256 bytes -- lives in a device array of the library of
 * MMCSF_CPD_NOALLOC_SCRATCH_BYTES, one per GPU, held until the process ends;
 * a larger work space is refused with MMCSF_ERR_UNSUPPORTED before any work.
 * Calls that overlap on one GPU share that array and race.  No timing
 * events are created: stats->ms_mttkrp is 0.
 */
#define MMCSF_CPD_NOALLOC_SCRATCH_BYTES (64ull << 20)
int mmcsf_cpd_als_noalloc(const mmcsf_format *f, int R, int max_iters,
                          double tol, int mttkrp_dtype, float *const *factors,
                          double *lambda, mmcsf_cpd_stats *stats,
                          void *stream);

/* G = U^T U (float64 R x R, device) of a device float32 rows x R matrix, with
 * MFMA (v_mfma_f32_16x16x4f32) over row chunks and float64 chunk sums
 * (the ssyrk_ calls of cpd_cpu.cu:543-546, 649-652). */
int mmcsf_gram(const float *U, uint64_t rows, int R, double *G, void *stream);

/* ---- CP completion (no reference counterpart) ---------------------------------
 *
 * Squared loss of a CP model on the observed entries and its gradients, for
 * tensor completion with a gradient optimiser:
 *     e_x = sum_r prod_n U_n[i_n(x)][r] - v_x,    L = 1/2 sum_x e_x^2,
 *     dL/dU_n = MTTKRP_n(e)  (mttkrp_cpu.cpp:60-113 with e in place of v).
 * Each call is one chunked-walk pass over a CSF tree of the format, computing
 * e_x on the fly (float32 arithmetic, float64 loss sum).  The format must be
 * B-CSF (its root mode only) or MM-CSF (every mode when built with mode
 * trees, MMCSF_TYPE_MMCSF with modeTrees = 1).  R % 4 == 0 and R <= 128;
 * factor rows 16-byte aligned (8-byte for bfloat16 with R % 8 != 0); every
 * factor table below 4 GiB (rows x R x 4 bytes), else MMCSF_ERR_UNSUPPORTED.
 * Residuals and weights are in the order of the tensor the format was built
 * from.  `loss` is a DEVICE float64 scalar, written stream-ordered. */

/* loss = 1/2 sum e^2; residual (device float[nnz], may be NULL) = e. */
int mmcsf_cp_forward(const mmcsf_format *f, int R, int dtype,
                     const void *const *factors, float *residual,
                     double *loss, void *stream);

/* grad (device float dims[mode] x R, overwritten) =
 *     alpha * sum_x w_x * prod_{m != mode} U_m[i_m(x)]  (row i_mode(x)),
 * with w = weights (device float[nnz]) -- the vector-Jacobian product of
 * the model values -- or, when weights is NULL, w = e recomputed from the
 * factors, so that grad = alpha * dL/dU_mode. */
int mmcsf_cp_adjoint(const mmcsf_format *f, int mode, int R, int dtype,
                     const void *const *factors, const float *weights,
                     float alpha, float *grad, void *stream);

/* loss = 1/2 sum e^2 and grads[m] = alpha * dL/dU_m for every m with
 * grads[m] != NULL, one pass per mode (the loss is summed in the first). */
int mmcsf_cp_loss_grad(const mmcsf_format *f, int R, int dtype,
                       const void *const *factors, float alpha,
                       float *const *grads, double *loss, void *stream);

/* ---- version ---------------------------------------------------------------- */
const char *mmcsf_version(void);

#ifdef __cplusplus
}
#endif

#endif /* MMCSF_H */
