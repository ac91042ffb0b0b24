// mttkrp -- command-line driver with the reference's flags and output lines
// (src/mttkrp.cu:71-472), running every GPU type on gfx950 through the C ABI.
//
//   mttkrp -i X.tns|X.fmt [-m mode] [-R rank] [-t 1|3|8|9|10|12|13|14] [-f fbrThreshold]
//          [-w w] [-b b] [-s s] [-c 1] [-o out.txt] [-v 1] [-d 0|1] [-M 0|1] [-k X.fmt]
//
// -d 1 (new) stores the factors in bfloat16.  -M 1 (new) builds MM-CSF with
// the reference's mixed-mode partition heuristic instead of its override.
// -k (new) writes the built format to a binary cache; -i then accepts that
// cache in place of the .tns text (no text parsing, sorting or building).  Factors are the reference's
// constants U[m] = m + 0.5 (util.h:3220-3223).  Errors print a message and
// exit with a non-zero status (the reference exits 0 on most of them).
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "mmcsf.h"

namespace {

[[noreturn]] void die(const char *what) {
    std::fprintf(stderr, "%s: %s\n", what, mmcsf_last_error());
    std::exit(1);
}

#define CK(call) \
    do {                                       \
        if ((call) != MMCSF_OK) die(#call);    \
    } while (0)

#define HK(call)                                                                  \
    do {                                                                          \
        hipError_t e_ = (call);                                                   \
        if (e_ != hipSuccess) {                                                   \
            std::fprintf(stderr, "%s: %s\n", #call, hipGetErrorString(e_));       \
            std::exit(1);                                                         \
        }                                                                         \
    } while (0)

double seconds() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

uint16_t to_bf16(float x) {  // round to nearest even
    uint32_t u;
    std::memcpy(&u, &x, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

void print_help() {
    std::printf(
        "options:\n"
        "        -R rank/feature : set the rank (default 32)\n"
        "        -m mode : set the mode of MTTKRP (default 0, MMCSF evaluates all modes)\n"
        "        -t implementation type: 1: COO CPU, 3: COO GPU 8: B-CSF 10: HB-CSF 12: MM-CSF on GPU (default 1)\n"
        "           9: B-CSF, every mode; 13: single CSF, every mode; 14: single CSF with bins, every mode\n"
        "        -f fiber-splitting threshold: maximum nonzeros per fiber (default inf)\n"
        "        -w warp per slice (default 4); -w, -s, -b and -g are checked against the reference's limits\n"
        "           and do not shape the launch (the chunked walk splits work by traversal cost)\n"
        "        -h MM-CSF fiber threshold of the mixed-mode partition (default 1)\n"
        "        -M 1: MM-CSF with the mixed-mode partition heuristic (default 0: every nonzero in partition 2)\n"
        "        -k format cache file to write after building; -i accepts such a file\n"
        "        -d factor storage: 0 float32, 1 bfloat16 (default 0)\n"
        "        -c correctness check against COO CPU on the last mode: 1 to enable\n"
        "        -i input file name: e.g., ../dataset/delicious.tns\n"
        "        -o output file name: if not set no output file will be written\n");
}

// write_output, util.h:3270-3283
void write_output(const std::vector<float> &U, uint64_t rows, uint64_t R, const char *path) {
    FILE *fp = std::fopen(path, "w");
    if (!fp) {
        std::fprintf(stderr, "cannot write %s\n", path);
        std::exit(1);
    }
    std::fprintf(fp, "%llu x %llu matrix\n", (unsigned long long)rows, (unsigned long long)R);
    for (uint64_t i = 0; i < rows; ++i) {
        for (uint64_t j = 0; j < R; ++j) std::fprintf(fp, "%.2f\t", U[i * R + j]);
        std::fprintf(fp, "\n");
    }
    std::fclose(fp);
}

// correctness_check, util.h:3285-3311 (absolute threshold 0.1)
void correctness_check(const std::vector<float> &out, const std::vector<float> &ref, uint64_t nr, uint64_t nc) {
    long mismatch = 0;
    float maxDiff = 0;
    const float precision = 0.1f;
    for (uint64_t i = 0; i < nr; ++i)
        for (uint64_t j = 0; j < nc; ++j) {
            float diff = std::fabs(out[i * nc + j] - ref[i * nc + j]);
            if (diff > precision) {
                if (diff > maxDiff) maxDiff = diff;
                if (mismatch < 5 && j == 0)
                    std::printf("mismatch at (%llu,%llu) got: %f exp: %f. \n", (unsigned long long)i,
                                (unsigned long long)j, out[i * nc + j], ref[i * nc + j]);
                mismatch++;
            }
        }
    if (mismatch == 0) std::printf("Correctness pass!\n");
    else std::printf("%ld mismatches found at %.6f precision. \nMaximum diff %f. \n", mismatch, precision, maxDiff);
}

}  // namespace

int main(int argc, char **argv) {
    for (int i = 1; i < argc; ++i)
        if (!std::strcmp(argv[i], "--help") || !std::strcmp(argv[i], "-help")) {
            print_help();
            return 0;
        }
    mmcsf_options opt;
    if (mmcsf_options_parse(argc, argv, &opt) != MMCSF_OK) {
        std::fprintf(stderr, "%s\n", mmcsf_last_error());
        print_help();
        return 1;
    }
    if (!opt.inFileName[0]) {
        std::printf("Provide input file path. Program will exit.\n");
        return 1;
    }
    mmcsf_tensor *X = nullptr;
    mmcsf_format *cachedF = nullptr;
    if (mmcsf_format_is_cache(opt.inFileName)) {
        // a format cache: the format as built, and the nonzeros it holds
        CK(mmcsf_format_load(opt.inFileName, &cachedF));
        CK(mmcsf_format_to_tensor(cachedF, &X));
    } else {
        CK(mmcsf_tensor_load_tns(opt.inFileName, &X));
    }
    int ndims = 0;
    uint64_t dims[4] = {0, 0, 0, 0}, nnz = 0;
    CK(mmcsf_tensor_info(X, &ndims, dims, &nnz));
    if (opt.mode >= (uint64_t)ndims) {
        std::printf("Mode cannot be larger than tensor dimension.\n");
        return 1;
    }
    const int type = (int)opt.impType;
    if (type == 11 || type == 12) opt.mode = 0;  // check_opt, util.h:250-251
    // the format of -t: the cached one, or built from X (and cached with -k)
    auto make_format = [&](const mmcsf_options &o) -> mmcsf_format * {
        mmcsf_format *F = nullptr;
        if (cachedF) {
            mmcsf_format_info ci;
            CK(mmcsf_format_info_get(cachedF, &ci));
            if (ci.type != type) {
                std::fprintf(stderr, "the format cache holds a -t %d format, not -t %d\n", ci.type, type);
                std::exit(1);
            }
            F = cachedF;
            cachedF = nullptr;
            return F;
        }
        CK(mmcsf_format_build(X, type, &o, nullptr, &F));
        if (opt.cacheFileName[0]) CK(mmcsf_format_save(F, opt.cacheFileName));
        return F;
    };
    const uint64_t R = opt.R;
    const int dtype = opt.factorDtype;

    // factors U[m] = m + 0.5 (randomize_mats, util.h:3209-3237)
    std::vector<std::vector<float>> Uh(ndims);
    for (int m = 0; m < ndims; ++m) Uh[m].assign(dims[m] * R, (float)m + 0.5f);
    std::vector<void *> Ud(ndims, nullptr);
    std::vector<float *> Od(ndims, nullptr);
    for (int m = 0; m < ndims; ++m) {
        size_t es = dtype == MMCSF_BF16 ? 2 : 4;
        HK(hipMalloc(&Ud[m], dims[m] * R * es + 16));
        if (dtype == MMCSF_BF16) {
            std::vector<uint16_t> b(dims[m] * R);
            for (size_t q = 0; q < b.size(); ++q) b[q] = to_bf16(Uh[m][q]);
            HK(hipMemcpy(Ud[m], b.data(), b.size() * 2, hipMemcpyHostToDevice));
        } else {
            HK(hipMemcpy(Ud[m], Uh[m].data(), Uh[m].size() * 4, hipMemcpyHostToDevice));
        }
        HK(hipMalloc(reinterpret_cast<void **>(&Od[m]), dims[m] * R * 4 + 16));
    }
    hipEvent_t e0, e1;
    HK(hipEventCreate(&e0));
    HK(hipEventCreate(&e1));
    std::vector<float> result;  // MTTKRP of the last computed mode
    int resultMode = (int)opt.mode;

    if (type == MMCSF_TYPE_COO_CPU) {
        mmcsf_format *F = make_format(opt);
        std::vector<const float *> Up(ndims);
        for (int m = 0; m < ndims; ++m) Up[m] = Uh[m].data();
        std::vector<float> out = Uh[opt.mode];  // not zeroed (mttkrp.cu:88-89)
        double t0 = seconds();
        CK(mmcsf_mttkrp_coo_host(F, (int)opt.mode, (int)R, Up.data(), out.data()));
        std::printf("COO CPU - time: %.3f sec \n", seconds() - t0);
        result = out;
        mmcsf_format_free(F);
    } else if (type == MMCSF_TYPE_COO || type == MMCSF_TYPE_BCSF || type == MMCSF_TYPE_HBCSF) {
        double tb = seconds();
        mmcsf_format *F = make_format(opt);
        if (opt.verbose) std::printf("GPU format build - time: %.3f sec \n", seconds() - tb);
        mmcsf_format_info info;
        CK(mmcsf_format_info_get(F, &info));
        int mode = info.modeOrder[0];
        HK(hipEventRecord(e0, nullptr));
        CK(mmcsf_mttkrp(F, mode, (int)R, dtype, Ud.data(), Od[mode], nullptr));
        HK(hipEventRecord(e1, nullptr));
        HK(hipEventSynchronize(e1));
        float ms = 0;
        HK(hipEventElapsedTime(&ms, e0, e1));
        if (type == MMCSF_TYPE_COO) std::printf("Loop on. COO GPU using loop - time %gms\n", ms);
        else if (type == MMCSF_TYPE_BCSF) std::printf("B-CSF-GPU-mode %d :%g,\n", mode, ms);
        else std::printf("HYB GPU: %g\n", ms);
        if (info.droppedSlices)
            std::printf("note: %llu slice(s) fall outside every bin of the reference's binning "
                        "(the reference skips them; they are computed here)\n",
                        (unsigned long long)info.droppedSlices);
        result.resize(dims[mode] * R);
        HK(hipMemcpy(result.data(), Od[mode], result.size() * 4, hipMemcpyDeviceToHost));
        resultMode = mode;
        mmcsf_format_free(F);
    } else if (type == MMCSF_TYPE_MMCSF) {
        double t0 = seconds();
        mmcsf_format *F = make_format(opt);
        std::printf("mm_partition & populate - time: %.3f sec \n", seconds() - t0);
        mmcsf_format_info info;
        CK(mmcsf_format_info_get(F, &info));
        std::vector<float> ms(ndims, 0.f);
        CK(mmcsf_mttkrp_allmode(F, (int)R, dtype, Ud.data(), Od.data(), nullptr, ms.data()));
        float tot = 0;
        for (int m = 0; m < ndims; ++m) {
            if (opt.verbose) std::printf("MI-HCSR-GPU-mode %d : %g,\n", m, ms[m]);
            tot += ms[m];
        }
        std::printf("Total GPU time: %g", tot);
        if (ndims == 3)
            std::printf(" nSlc:%llu, nFibers:%llu, nnz:%llu\n", (unsigned long long)info.nLevel[0],
                        (unsigned long long)info.nFibers, (unsigned long long)nnz);
        else
            std::printf(" nSlc:%llu, nSFibers:%llu, nFibers:%llu, nnz:%llu\n", (unsigned long long)info.nLevel[0],
                        (unsigned long long)info.nFibers, (unsigned long long)info.nLevel[2] + 1,
                        (unsigned long long)nnz);
        // the reference copies back the last mode (mttkrp_gpu.cu:3771-3774)
        resultMode = ndims - 1;
        result.resize(dims[resultMode] * R);
        HK(hipMemcpy(result.data(), Od[resultMode], result.size() * 4, hipMemcpyDeviceToHost));
        mmcsf_format_free(F);
    } else if (type == MMCSF_TYPE_BCSF_ANYMODE || type == MMCSF_TYPE_ONECSF || type == MMCSF_TYPE_SINGLECSF) {
        // one tree, every mode (mttkrp.cu:223-231, 234-299)
        if (type == MMCSF_TYPE_SINGLECSF && opt.verbose)
            std::printf("Starting sameCSF: MTTKRP on all modes using same CSF\n");
        double tb = seconds();
        mmcsf_format *F = make_format(opt);
        if (opt.verbose) std::printf("GPU format build - time: %.3f sec \n", seconds() - tb);
        std::vector<float> ms(ndims, 0.f);
        CK(mmcsf_mttkrp_allmode(F, (int)R, dtype, Ud.data(), Od.data(), nullptr, ms.data()));
        if (type == MMCSF_TYPE_SINGLECSF) {
            // the role lines of MTTKRP_HCSR_CPU / _mode1 / _mode2 (mttkrp_cpu.cpp:171, 235, 289),
            // then the time of each mode
            mmcsf_format_info info;
            CK(mmcsf_format_info_get(F, &info));
            const int *o = info.modeOrder;
            for (int m = 0; m < ndims; ++m) {
                if (ndims == 3 && m == o[0]) std::printf("mode 0: %d %d %d\n", o[0], o[1], o[2]);
                else if (ndims == 3 && m == o[1]) std::printf("mode 1: %d %d %d\n", m, o[2], o[0]);
                else if (ndims == 3 && m == o[2]) std::printf("mode 2: %d %d %d\n", m, o[0], o[1]);
                std::printf("sameCSF-GPU-mode %d :%g,\n", m, ms[m]);
            }
        } else {
            for (int m = 0; m < ndims; ++m)
                std::printf("%s-GPU-mode %d :%g,\n", type == MMCSF_TYPE_BCSF_ANYMODE ? "ONE-B-CSF" : "singleCSF", m,
                            ms[m]);
        }
        resultMode = ndims - 1;
        result.resize(dims[resultMode] * R);
        HK(hipMemcpy(result.data(), Od[resultMode], result.size() * 4, hipMemcpyDeviceToHost));
        mmcsf_format_free(F);
    } else {
        std::printf("no MTTKRP\n");
    }

    if (opt.outFileName[0] && !result.empty()) write_output(result, dims[resultMode], R, opt.outFileName);

    if (opt.correctness && !result.empty()) {
        if (type == MMCSF_TYPE_COO_CPU) {
            std::printf("Already running COO seq on CPU!\n");
        } else {
            // mttkrp.cu:446-471: the last mode against COO CPU with a zeroed output
            int mode = ndims - 1;
            if (resultMode != mode) {
                // single-mode types computed their root mode; check that one
                mode = resultMode;
            }
            mmcsf_options o2 = opt;
            o2.mode = (uint64_t)mode;
            mmcsf_format *F = nullptr;
            CK(mmcsf_format_build(X, MMCSF_TYPE_COO_CPU, &o2, nullptr, &F));
            std::vector<const float *> Up(ndims);
            for (int m = 0; m < ndims; ++m) Up[m] = Uh[m].data();
            std::vector<float> ref(dims[mode] * R, 0.f);
            std::printf("correctness with COO on mode %d. \n", mode);
            CK(mmcsf_mttkrp_coo_host(F, mode, (int)R, Up.data(), ref.data()));
            correctness_check(result, ref, dims[mode], R);
            mmcsf_format_free(F);
        }
    }
    for (int m = 0; m < ndims; ++m) {
        (void)hipFree(Ud[m]);
        (void)hipFree(Od[m]);
    }
    if (cachedF) mmcsf_format_free(cachedF);
    mmcsf_tensor_free(X);
    return 0;
}
