// ttm -- command-line driver of the tensor-times-matrix path with the
// reference's flags and output lines (src/ttm.cu:70-230), on gfx950 through
// the C ABI.
//
//   ttm -i X.tns [-m mode] [-R rank] [-t 1|2|3] [-f fbrThreshold] [-c 1]
//       [-o out.txt] [-v 1] [-d 0|1]
//
// The tree is the B-CSF of create_HCSR with the -f split (ttm.cu:76-78) and
// mode order order_tensormode(X, -m); Y has one row per leaf fiber.  -t 1 runs
// TTM_CPU on the host, -t 2 / 3 the GPU kernel (the reference's slice- and
// fiber-parallel launches; one kernel serves both here).  Factors are the
// reference's constants U[m] = m + 0.5.  The reference's check that the input
// file name contains the root mode (ttm.cu:82-89) is not repeated.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "mmcsf.h"

namespace {

[[noreturn]] void die(const char *what) {
    std::fprintf(stderr, "%s: %s\n", what, mmcsf_last_error());
    std::exit(1);
}

#define CK(call)                               \
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

// write_output_ttmY, util.h:3239-3252
void write_y(const std::vector<float> &Y, uint64_t rows, uint64_t R, const char *path) {
    FILE *fp = std::fopen(path, "w");
    if (!fp) {
        std::fprintf(stderr, "cannot write %s\n", path);
        std::exit(1);
    }
    std::fprintf(fp, "%llu x %llu semiSpTensor\n", (unsigned long long)rows, (unsigned long long)R);
    for (uint64_t i = 0; i < rows; ++i) {
        for (uint64_t j = 0; j < R; ++j) std::fprintf(fp, "%.2f\t", Y[i * R + j]);
        std::fprintf(fp, "\n");
    }
    std::fclose(fp);
}

// correctness_check, util.h:3285-3311 (absolute threshold 0.1)
void correctness_check(const std::vector<float> &out, const std::vector<float> &ref, uint64_t nr, uint64_t nc) {
    long mismatch = 0;
    float maxDiff = 0;
    for (uint64_t i = 0; i < nr * nc; ++i) {
        float d = std::fabs(out[i] - ref[i]);
        if (d > 0.1f) {
            maxDiff = std::fmax(maxDiff, d);
            ++mismatch;
        }
    }
    if (mismatch == 0) std::printf("Correctness pass!\n");
    else std::printf("%ld mismatches found at 0.100000 precision. \nMaximum diff %f. \n", mismatch, maxDiff);
}

}  // namespace

int main(int argc, char **argv) {
    mmcsf_options opt;
    if (mmcsf_options_parse(argc, argv, &opt) != MMCSF_OK) {
        std::fprintf(stderr, "%s\n", mmcsf_last_error());
        return 1;
    }
    if (!opt.inFileName[0]) {
        std::fprintf(stderr, "no input file (-i)\n");
        return 1;
    }
    mmcsf_tensor *X = nullptr;
    CK(mmcsf_tensor_load_tns(opt.inFileName, &X));
    int ndims = 0;
    uint64_t dims[4] = {0, 0, 0, 0}, nnz = 0;
    CK(mmcsf_tensor_info(X, &ndims, dims, &nnz));
    if (opt.mode >= (uint64_t)ndims) {
        std::printf("Mode cannot be larger than tensor dimension.\n");
        return 1;
    }
    const int type = (int)opt.impType;
    const int R = (int)opt.R;
    mmcsf_format *F = nullptr;
    CK(mmcsf_format_build(X, MMCSF_TYPE_BCSF, &opt, nullptr, &F));
    mmcsf_format_info info;
    CK(mmcsf_format_info_get(F, &info));
    const int leafMode = info.modeOrder[ndims - 1];
    const uint64_t rows = info.nLevel[ndims - 2];
    std::vector<float> Uh(dims[leafMode] * (uint64_t)R, (float)leafMode + 0.5f);
    if (opt.verbose) std::printf("\nStarting TTM...\n");

    std::vector<float> Y;
    if (type == 1) {
        Y.assign(rows * R, 0.f);
        double t0 = seconds();
        CK(mmcsf_ttm_host(F, R, Uh.data(), Y.data()));
        std::printf("TTM - COO CPU time: %.3f sec \n", seconds() - t0);
    } else if (type == 2 || type == 3) {
        void *Ud = nullptr;
        float *Yd = nullptr;
        const int dtype = opt.factorDtype;
        const size_t es = dtype == MMCSF_BF16 ? 2 : 4;
        HK(hipMalloc(&Ud, Uh.size() * es + 16));
        HK(hipMalloc(reinterpret_cast<void **>(&Yd), rows * R * 4 + 16));
        if (dtype == MMCSF_BF16) {
            std::vector<uint16_t> b(Uh.size());
            for (size_t q = 0; q < b.size(); ++q) b[q] = to_bf16(Uh[q]);
            HK(hipMemcpy(Ud, b.data(), b.size() * 2, hipMemcpyHostToDevice));
        } else {
            HK(hipMemcpy(Ud, Uh.data(), Uh.size() * 4, hipMemcpyHostToDevice));
        }
        hipEvent_t e0, e1;
        HK(hipEventCreate(&e0));
        HK(hipEventCreate(&e1));
        HK(hipEventRecord(e0, nullptr));
        CK(mmcsf_ttm(F, R, dtype, Ud, Yd, nullptr));
        HK(hipEventRecord(e1, nullptr));
        HK(hipEventSynchronize(e1));
        float ms = 0;
        HK(hipEventElapsedTime(&ms, e0, e1));
        std::printf("TTM GPU - time %gms\n", ms);
        Y.resize(rows * R);
        HK(hipMemcpy(Y.data(), Yd, Y.size() * 4, hipMemcpyDeviceToHost));
        (void)hipFree(Ud);
        (void)hipFree(Yd);
    } else {
        std::printf("no TTM\n");
    }
    if (opt.outFileName[0] && !Y.empty()) {
        std::printf("Writing Y to %s\n", opt.outFileName);
        write_y(Y, rows, R, opt.outFileName);
    }
    if (opt.correctness && !Y.empty()) {
        std::printf("correctness with CPU \n");
        if (type == 1) {
            std::printf("Already running COO seq on CPU!\n");
        } else {
            std::vector<float> ref(rows * R, 0.f);
            CK(mmcsf_ttm_host(F, R, Uh.data(), ref.data()));
            correctness_check(Y, ref, rows, R);
        }
    }
    mmcsf_format_free(F);
    mmcsf_tensor_free(X);
    return 0;
}
