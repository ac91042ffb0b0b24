// gather_ceiling: the random factor-row gather rate of this GPU, measured by
// kernels that do nothing but gather rows.  It is the ceiling the MTTKRP walk
// is held against: an MTTKRP over a sparse tensor reads a factor row per
// nonzero (and per fiber), at random, and those row reads, not the streamed
// structure, set its time.
//
//   gather_ceiling [-b rowBytes,...] [-t tableMB,...]
//   gather_ceiling -s [-t tableMB,...]
//
// Prints one JSON line per measurement:
//   {"kind": "stream", ...}           coalesced read of 256 MiB (HBM ceiling)
//   {"kind": "gather", ...}           rows of rowBytes (16 B per lane, rowBytes/16
//                                     lanes per row, 64/(rowBytes/16) rows per
//                                     load instruction), uniformly at random
//                                     over a table of tableMB
//   {"kind": "gather_one_row", ...}   every group of an instruction reads the
//                                     same row (one line per instruction: the
//                                     per-instruction issue ceiling)
//   {"kind": "gather_masked", ...}    128-byte rows with only one group of 8
//                                     lanes in range per instruction (raw
//                                     buffer loads; out-of-range lanes read
//                                     zeros without touching memory)
//   -s: {"kind": "shape", ...}       the cost of one load instruction in CU
//                                     clocks (chip clocks x CUs / instructions
//                                     per second) for each load shape a row
//                                     gather can take: lanes per row x bytes per
//                                     lane, groups enabled by the exec mask,
//                                     coalesced words, and 128-byte rows read
//                                     from an LDS table
// Defaults: row bytes 128, 64, 32; tables 1, 3.2, 6.4, 25.6, 64 MB (-s: 1,
// 6.4, 25.6).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#define CK(x)                                                                                   \
    do {                                                                                        \
        hipError_t e_ = (x);                                                                    \
        if (e_ != hipSuccess) {                                                                 \
            std::fprintf(stderr, "HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__); \
            std::exit(1);                                                                       \
        }                                                                                       \
    } while (0)

namespace {

constexpr int kBlock = 256;
constexpr int kInFlight = 8;     // load instructions in flight per iteration
constexpr uint32_t kIters = 256; // iterations per wave

__device__ __forceinline__ uint32_t mix(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

__device__ __forceinline__ void fold(uint4 &acc, const uint4 &v) {
    acc.x ^= v.x;
    acc.y += v.y;
    acc.z ^= v.z;
    acc.w += v.w;
}

// G lanes per row; ONE: every group of an instruction reads the same row.
// Row r < rows and lane k < G, so every read stays inside rows * G uint4.
template <int G, bool ONE>
__global__ __launch_bounds__(kBlock) void k_gather(const uint4 *__restrict__ tab, uint32_t rows, uint4 *sink) {
    const int lane = threadIdx.x & 63;
    const int g = lane / G, k = lane % G;
    const uint32_t wave = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
    const uint32_t seed = mix(wave * 7919u + 17u);
    uint4 acc = make_uint4(0, 0, 0, 0);
    for (uint32_t it = 0; it < kIters; ++it) {
        uint4 v[kInFlight];
#pragma unroll
        for (int u = 0; u < kInFlight; ++u) {
            const uint32_t key = seed + it * kInFlight + u;
            const uint32_t r = mix(ONE ? key : key * 64u + g) % rows;
            v[u] = tab[(size_t)r * G + k];
        }
#pragma unroll
        for (int u = 0; u < kInFlight; ++u) fold(acc, v[u]);
    }
    if ((acc.x ^ acc.y ^ acc.z ^ acc.w) == 0x12345678u) sink[threadIdx.x] = acc;  // keeps the loads live
}

// 128-byte rows by raw buffer loads; only group 0 of each instruction is in
// range, the other lanes get an offset past the table's byte count.
__global__ __launch_bounds__(kBlock) void k_gather_masked(const uint4 *__restrict__ tab, uint32_t rows, uint4 *sink) {
    constexpr int G = 8;
    const int lane = threadIdx.x & 63;
    const int g = lane / G, k = lane % G;
    const uint32_t wave = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
    const uint32_t seed = mix(wave * 7919u + 17u);
    const __amdgpu_buffer_rsrc_t rs =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<uint4 *>(tab), (short)0, (int)(rows * G * 16u), 0x00020000);
    uint4 acc = make_uint4(0, 0, 0, 0);
    for (uint32_t it = 0; it < kIters; ++it) {
        uint4 v[kInFlight];
#pragma unroll
        for (int u = 0; u < kInFlight; ++u) {
            const uint32_t r = mix((seed + it * kInFlight + u) * 64u + g) % rows;
            const uint32_t off = g == 0 ? (r * G + k) * 16u : 0xFFFFF000u;
            v[u] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(rs, off, 0, 0));
        }
#pragma unroll
        for (int u = 0; u < kInFlight; ++u) fold(acc, v[u]);
    }
    if ((acc.x ^ acc.y ^ acc.z ^ acc.w) == 0x12345678u) sink[threadIdx.x] = acc;
}

template <typename V>
__device__ __forceinline__ uint32_t fold1(const V &v);
template <>
__device__ __forceinline__ uint32_t fold1<uint4>(const uint4 &v) { return v.x ^ v.y ^ v.z ^ v.w; }
template <>
__device__ __forceinline__ uint32_t fold1<uint2>(const uint2 &v) { return v.x ^ v.y; }
template <>
__device__ __forceinline__ uint32_t fold1<uint32_t>(const uint32_t &v) { return v; }

// Rows of L * sizeof(V) bytes, L lanes per row (64 / L rows per load
// instruction); only groups g < A load (the others are off in the exec mask).
// Row r < rows and lane k < L, so every read stays inside rows * L elements.
template <typename V, int L, int A>
__global__ __launch_bounds__(kBlock) void k_shape(const V *__restrict__ tab, uint32_t rows, uint32_t *sink) {
    const int lane = threadIdx.x & 63;
    const int g = lane / L, k = lane % L;
    const uint32_t wave = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
    const uint32_t seed = mix(wave * 7919u + 17u);
    uint32_t acc = 0;
    for (uint32_t it = 0; it < kIters; ++it) {
        V v[kInFlight];
#pragma unroll
        for (int u = 0; u < kInFlight; ++u) {
            const uint32_t r = mix((seed + it * kInFlight + u) * 64u + g) % rows;
            if (A >= 64 / L || g < A) v[u] = tab[(size_t)r * L + k];
            else v[u] = V{};
        }
#pragma unroll
        for (int u = 0; u < kInFlight; ++u) acc += fold1(v[u]);
    }
    if (acc == 0x12345678u) sink[threadIdx.x] = acc;  // keeps the loads live
}

// 64 consecutive words per instruction at a random 256-byte block b < blocks
__global__ __launch_bounds__(kBlock) void k_coalesced(const uint32_t *__restrict__ a, uint32_t blocks, uint32_t *sink) {
    const int lane = threadIdx.x & 63;
    const uint32_t wave = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
    const uint32_t seed = mix(wave * 7919u + 17u);
    uint32_t acc = 0;
    for (uint32_t it = 0; it < kIters; ++it) {
        uint32_t v[kInFlight];
#pragma unroll
        for (int u = 0; u < kInFlight; ++u) {
            const uint32_t b = __builtin_amdgcn_readfirstlane(mix(seed + it * kInFlight + u) % blocks);
            v[u] = a[(size_t)b * 64 + lane];
        }
#pragma unroll
        for (int u = 0; u < kInFlight; ++u) acc += v[u];
    }
    if (acc == 0x12345678u) sink[threadIdx.x] = acc;
}

// 128-byte rows (8 lanes x 16 bytes) at random from an LDS table of H rows
// copied from the first H rows of tab
template <int H>
__global__ __launch_bounds__(kBlock) void k_lds_rows(const uint4 *__restrict__ tab, uint32_t *sink) {
    __shared__ uint4 sm[H * 8];
    for (int i = threadIdx.x; i < H * 8; i += kBlock) sm[i] = tab[i];
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int g = lane / 8, k = lane % 8;
    const uint32_t wave = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
    const uint32_t seed = mix(wave * 7919u + 17u);
    uint32_t acc = 0;
    for (uint32_t it = 0; it < kIters; ++it) {
        uint4 v[kInFlight];
#pragma unroll
        for (int u = 0; u < kInFlight; ++u) v[u] = sm[(mix((seed + it * kInFlight + u) * 64u + g) % H) * 8 + k];
#pragma unroll
        for (int u = 0; u < kInFlight; ++u) acc += fold1(v[u]);
    }
    if (acc == 0x12345678u) sink[threadIdx.x] = acc;
}

__global__ __launch_bounds__(kBlock) void k_stream(const uint4 *__restrict__ a, size_t n, uint4 *sink) {
    uint4 acc = make_uint4(0, 0, 0, 0);
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) fold(acc, a[i]);
    if ((acc.x ^ acc.y ^ acc.z ^ acc.w) == 0x12345678u) sink[threadIdx.x] = acc;
}

std::vector<double> parse_list(const char *s) {
    std::vector<double> v;
    std::string t(s);
    size_t p = 0;
    while (p <= t.size()) {
        size_t q = t.find(',', p);
        if (q == std::string::npos) q = t.size();
        if (q > p) v.push_back(std::atof(t.substr(p, q - p).c_str()));
        p = q + 1;
    }
    return v;
}

}  // namespace

int main(int argc, char **argv) {
    std::vector<double> rowBytes = {128, 64, 32};
    std::vector<double> tableMB = {1.0, 3.2, 6.4, 25.6, 64.0};
    bool shapes = false, tablesGiven = false;
    for (int i = 1; i < argc; ++i) {
        if (!std::strcmp(argv[i], "-s")) {
            shapes = true;
        } else if (i + 1 < argc && !std::strcmp(argv[i], "-b")) {
            rowBytes = parse_list(argv[++i]);
        } else if (i + 1 < argc && !std::strcmp(argv[i], "-t")) {
            tableMB = parse_list(argv[++i]);
            tablesGiven = true;
        } else {
            std::fprintf(stderr, "usage: gather_ceiling [-b rowBytes,...] [-t tableMB,...] | -s [-t tableMB,...]\n");
            return 2;
        }
    }
    if (shapes && !tablesGiven) tableMB = {1.0, 6.4, 25.6};
    for (double b : rowBytes)
        if (b != 128 && b != 64 && b != 32) {
            std::fprintf(stderr, "row bytes must be 128, 64 or 32\n");
            return 2;
        }
    for (double t : tableMB)
        if (!(t > 0 && t <= 2048)) {
            std::fprintf(stderr, "table sizes must lie in (0, 2048] MB\n");
            return 2;
        }
    CK(hipSetDevice(0));
    hipDeviceProp_t prop;
    CK(hipGetDeviceProperties(&prop, 0));
    const int cus = prop.multiProcessorCount;
    const size_t streamBytes = 256ull << 20;
    size_t tabBytes = streamBytes;
    for (double t : tableMB) tabBytes = std::max(tabBytes, (size_t)(t * 1e6) + 4096);
    uint4 *tab = nullptr, *sink = nullptr;
    CK(hipMalloc(&tab, tabBytes));
    CK(hipMalloc(&sink, kBlock * sizeof(uint4)));
    CK(hipMemset(tab, 1, tabBytes));
    hipEvent_t e0, e1;
    CK(hipEventCreate(&e0));
    CK(hipEventCreate(&e1));
    const int blocks = cus * 8;
    const double waves = (double)blocks * (kBlock / 64);
    auto time_ms = [&](auto launch) {
        launch();
        CK(hipGetLastError());
        CK(hipDeviceSynchronize());
        CK(hipEventRecord(e0));
        for (int r = 0; r < 5; ++r) launch();
        CK(hipEventRecord(e1));
        CK(hipEventSynchronize(e1));
        float ms = 0;
        CK(hipEventElapsedTime(&ms, e0, e1));
        return ms / 5.0;
    };
    std::printf("{\"kind\": \"device\", \"name\": \"%s\", \"cus\": %d, \"l2_bytes\": %d}\n", prop.gcnArchName, cus,
                prop.l2CacheSize);
    {
        const size_t n = streamBytes / 16;
        const double ms = time_ms([&] { k_stream<<<cus * 16, kBlock>>>(tab, n, sink); });
        std::printf("{\"kind\": \"stream\", \"bytes\": %zu, \"ms\": %.4f, \"gb_per_s\": %.1f}\n", streamBytes, ms,
                    streamBytes / (ms * 1e-3) / 1e9);
    }
    const double instr = waves * kIters * kInFlight;
    if (shapes) {
        // CU clocks per load instruction: (clock x CUs) / (instructions per second)
        const double cuClk = (double)cus * prop.clockRate * 1e3;
        auto shape = [&](const char *what, int laneBytes, int lanesPerRow, int groups, double mb, double ms) {
            const double ips = instr / (ms * 1e-3);
            const int rowsPerInstr = groups;
            std::printf("{\"kind\": \"shape\", \"what\": \"%s\", \"lane_bytes\": %d, \"lanes_per_row\": %d, "
                        "\"rows_per_instr\": %d, \"table_mb\": %.3f, \"ms\": %.4f, \"g_instr_per_s\": %.2f, "
                        "\"g_rows_per_s\": %.1f, \"cu_clk_per_instr\": %.1f}\n",
                        what, laneBytes, lanesPerRow, rowsPerInstr, mb, ms, ips / 1e9, ips * rowsPerInstr / 1e9,
                        cuClk / ips);
            std::fflush(stdout);
        };
        for (double mb : tableMB) {
            const uint32_t r128 = std::max<uint32_t>(1, (uint32_t)(mb * 1e6 / 128));
            const uint32_t r64 = std::max<uint32_t>(1, (uint32_t)(mb * 1e6 / 64));
            const uint32_t r32 = std::max<uint32_t>(1, (uint32_t)(mb * 1e6 / 32));
            const uint32_t r4 = std::max<uint32_t>(1, (uint32_t)(mb * 1e6 / 4));
            const uint32_t *w = reinterpret_cast<const uint32_t *>(tab);
            const uint2 *d = reinterpret_cast<const uint2 *>(tab);
            shape("128 B rows, 8 lanes x 16 B", 16, 8, 8, mb,
                  time_ms([&] { k_shape<uint4, 8, 8><<<blocks, kBlock>>>(tab, r128, (uint32_t *)sink); }));
            shape("128 B rows, 8 lanes x 16 B, 4 groups in exec", 16, 8, 4, mb,
                  time_ms([&] { k_shape<uint4, 8, 4><<<blocks, kBlock>>>(tab, r128, (uint32_t *)sink); }));
            shape("128 B rows, 8 lanes x 16 B, 1 group in exec", 16, 8, 1, mb,
                  time_ms([&] { k_shape<uint4, 8, 1><<<blocks, kBlock>>>(tab, r128, (uint32_t *)sink); }));
            shape("128 B rows, 16 lanes x 8 B", 8, 16, 4, mb,
                  time_ms([&] { k_shape<uint2, 16, 4><<<blocks, kBlock>>>(d, r128, (uint32_t *)sink); }));
            shape("128 B rows, 32 lanes x 4 B", 4, 32, 2, mb,
                  time_ms([&] { k_shape<uint32_t, 32, 2><<<blocks, kBlock>>>(w, r128, (uint32_t *)sink); }));
            shape("64 B rows, 4 lanes x 16 B", 16, 4, 16, mb,
                  time_ms([&] { k_shape<uint4, 4, 16><<<blocks, kBlock>>>(tab, r64, (uint32_t *)sink); }));
            shape("64 B rows, 8 lanes x 8 B", 8, 8, 8, mb,
                  time_ms([&] { k_shape<uint2, 8, 8><<<blocks, kBlock>>>(d, r64, (uint32_t *)sink); }));
            shape("32 B rows, 2 lanes x 16 B", 16, 2, 32, mb,
                  time_ms([&] { k_shape<uint4, 2, 32><<<blocks, kBlock>>>(tab, r32, (uint32_t *)sink); }));
            shape("4 B rows, 1 lane x 4 B", 4, 1, 64, mb,
                  time_ms([&] { k_shape<uint32_t, 1, 64><<<blocks, kBlock>>>(w, r4, (uint32_t *)sink); }));
            shape("coalesced: 64 consecutive words", 4, 64, 1, mb,
                  time_ms([&] { k_coalesced<<<blocks, kBlock>>>(w, std::max<uint32_t>(1, r4 / 64), (uint32_t *)sink); }));
        }
        shape("LDS: 128 B rows of a 256-row table, 8 lanes x 16 B", 16, 8, 8, 0.032768,
              time_ms([&] { k_lds_rows<256><<<blocks, kBlock>>>(tab, (uint32_t *)sink); }));
        CK(hipFree(tab));
        CK(hipFree(sink));
        return 0;
    }
    auto report = [&](const char *kind, int rb, double mb, uint32_t rows, double ms, double rowsPerInstr) {
        const double rowsTot = instr * rowsPerInstr;
        std::printf("{\"kind\": \"%s\", \"row_bytes\": %d, \"table_mb\": %.3f, \"rows\": %u, \"ms\": %.4f, "
                    "\"g_rows_per_s\": %.1f, \"g_instr_per_s\": %.2f}\n",
                    kind, rb, mb, rows, ms, rowsTot / (ms * 1e-3) / 1e9, instr / (ms * 1e-3) / 1e9);
        std::fflush(stdout);
    };
    for (double mb : tableMB) {
        for (double b : rowBytes) {
            const int rb = (int)b;
            const uint32_t rows = std::max<uint32_t>(1, (uint32_t)(mb * 1e6 / rb));
            double ms = 0;
            if (rb == 128) ms = time_ms([&] { k_gather<8, false><<<blocks, kBlock>>>(tab, rows, sink); });
            else if (rb == 64) ms = time_ms([&] { k_gather<4, false><<<blocks, kBlock>>>(tab, rows, sink); });
            else ms = time_ms([&] { k_gather<2, false><<<blocks, kBlock>>>(tab, rows, sink); });
            report("gather", rb, mb, rows, ms, 64.0 / (rb / 16));
            if (rb == 128) {
                ms = time_ms([&] { k_gather<8, true><<<blocks, kBlock>>>(tab, rows, sink); });
                report("gather_one_row", rb, mb, rows, ms, 8.0);
                ms = time_ms([&] { k_gather_masked<<<blocks, kBlock>>>(tab, rows, sink); });
                report("gather_masked", rb, mb, rows, ms, 1.0);
            }
        }
    }
    CK(hipFree(tab));
    CK(hipFree(sink));
    return 0;
}
