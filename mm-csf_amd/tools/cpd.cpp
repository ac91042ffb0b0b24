// cpd -- CP-ALS driver with the reference's flags and output lines
// (src/cpd.cu:70-151, cpd_cpu.cu:484-691), fully on gfx950 through the C ABI.
//
//   cpd -i X.tns [-R rank] [-f fbrThreshold] [-n iterations] [-x seed] [-d 0|1] [-v 1]
//
// -n (new; the reference fixes 10 iterations, util.h:210), -x (new) seeds the
// initial factors, drawn like RandomValue (util.h:3199-3206): uniform
// magnitude in [0, 3) with a random sign.  -d 1 runs the MTTKRP on bfloat16
// factor copies.
#include <hip/hip_runtime.h>

#include <chrono>
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

#define CK(call) \
    do {                                       \
        if ((call) != MMCSF_OK) die(#call);    \
    } while (0)

uint64_t splitmix(uint64_t &s) {
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

}  // namespace

int main(int argc, char **argv) {
    mmcsf_options opt;
    if (mmcsf_options_parse(argc, argv, &opt) != MMCSF_OK) {
        std::fprintf(stderr, "%s\n", mmcsf_last_error());
        return 1;
    }
    if (!opt.inFileName[0]) {
        std::printf("Provide input file path. Program will exit.\n");
        return 1;
    }
    opt.impType = MMCSF_TYPE_MMCSF;  // cpd.cu:74-75
    const int R = (int)opt.R;
    mmcsf_tensor *X = nullptr;
    CK(mmcsf_tensor_load_tns(opt.inFileName, &X));
    int ndims = 0;
    uint64_t dims[4] = {0, 0, 0, 0}, nnz = 0;
    CK(mmcsf_tensor_info(X, &ndims, dims, &nnz));
    mmcsf_format *F = nullptr;
    CK(mmcsf_format_build(X, MMCSF_TYPE_MMCSF, &opt, nullptr, &F));
    std::vector<float *> U(ndims, nullptr);
    uint64_t seed = opt.seed;
    for (int m = 0; m < ndims; ++m) {
        std::vector<float> h(dims[m] * R);
        for (auto &v : h) {
            float mag = 3.0f * (float)((splitmix(seed) >> 40) * 0x1.0p-24);
            v = (splitmix(seed) & 1) ? mag : -mag;
        }
        if (hipMalloc(reinterpret_cast<void **>(&U[m]), h.size() * 4 + 16) != hipSuccess) die("hipMalloc");
        if (hipMemcpy(U[m], h.data(), h.size() * 4, hipMemcpyHostToDevice) != hipSuccess) die("hipMemcpy");
    }
    std::vector<double> lambda(R);
    mmcsf_cpd_stats st;
    auto t0 = std::chrono::steady_clock::now();
    CK(mmcsf_cpd_als(F, R, (int)opt.cpdIters, 1e-5, opt.factorDtype, U.data(), lambda.data(), &st, nullptr));
    double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (opt.verbose)
        for (int it = 0; it < st.iterations && it < 64; ++it)
            std::printf("  its = %d, fit = %0.5f\n", it + 1, st.fit_history[it]);
    double prev = st.iterations > 1 ? st.fit_history[st.iterations - 2] : 0.0;
    std::printf("  its = %u (%.3f), fit = %0.5lf  delta = %+0.4e\n", (unsigned)st.iterations,
                st.ms_total / 1e3 / std::max(1, st.iterations), st.fit, st.fit - prev);
    std::printf("\tTotal MTTKRP time: %.3f sec \n", st.ms_mttkrp / 1e3);
    std::printf("CPD time (the same measure with SPLATT) = %.3f \n", st.ms_total / 1e3);
    std::printf("GPU CPD time: %.3f sec \n", sec);
    for (int m = 0; m < ndims; ++m) (void)hipFree(U[m]);
    mmcsf_format_free(F);
    mmcsf_tensor_free(X);
    return 0;
}
