// Binary cache of built formats: every array of every tree of a format
// (pointer / index / value arrays, positions, start bitmasks, chunk cost
// prefix, bins, hot-row tables, the per-mode root trees and the mixed-mode
// partitions) with its sizes, so a reload skips reading, sorting and building.
//
// File layout (little-endian): the 8-byte magic "MMCSFFMT", a uint32 version
// (kVersion) and a uint32 byte-order mark 0x01020304, then the format's
// fields in a fixed order (scalars as their bytes, each array as a uint64
// element count and its elements), then the 8-byte end mark "MMCSFEND".
// Loading checks the magic, version, byte order, every count against the
// remaining file size and the end mark.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "mmcsf_internal.hpp"

namespace mmcsf {
namespace {

constexpr char kMagic[8] = {'M', 'M', 'C', 'S', 'F', 'F', 'M', 'T'};
constexpr char kEnd[8] = {'M', 'M', 'C', 'S', 'F', 'E', 'N', 'D'};
constexpr uint32_t kVersion = 1;
constexpr uint32_t kByteOrder = 0x01020304u;

// Two pinned host buffers between the file and the device: the file read or
// write of one chunk overlaps the copy of the next on a stream of its own.
struct Stage {
    static constexpr size_t kChunk = 32ull << 20;
    char *buf[2] = {nullptr, nullptr};
    hipEvent_t done[2] = {nullptr, nullptr};
    bool busy[2] = {false, false};
    hipStream_t s = nullptr;
    int next = 0;
    void init() {
        MMCSF_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        for (int i = 0; i < 2; ++i) {
            MMCSF_HIP(hipHostMalloc(reinterpret_cast<void **>(&buf[i]), kChunk));
            MMCSF_HIP(hipEventCreateWithFlags(&done[i], hipEventDisableTiming));
        }
    }
    // the next buffer, once the copy that last used it has finished
    int take() {
        const int i = next;
        next ^= 1;
        if (busy[i]) MMCSF_HIP(hipEventSynchronize(done[i]));
        busy[i] = false;
        return i;
    }
    void drain() {
        if (s) MMCSF_HIP(hipStreamSynchronize(s));
        busy[0] = busy[1] = false;
    }
    ~Stage() {
        if (s) (void)hipStreamSynchronize(s);
        for (int i = 0; i < 2; ++i) {
            if (buf[i]) (void)hipHostFree(buf[i]);
            if (done[i]) (void)hipEventDestroy(done[i]);
        }
        if (s) (void)hipStreamDestroy(s);
    }
};

struct Writer {
    FILE *f;
    Stage &st;
    void bytes(const void *p, size_t n) {
        if (n && std::fwrite(p, 1, n, f) != n) fail(MMCSF_ERR_IO, "format cache: write failed");
    }
    template <typename T>
    void pod(const T &v) { bytes(&v, sizeof(T)); }
    template <typename T>
    void dev(const DevArray<T> &a) {
        pod<uint64_t>(a.n);
        const size_t total = a.n * sizeof(T);
        const char *src = reinterpret_cast<const char *>(a.p);
        const size_t chunks = (total + Stage::kChunk - 1) / Stage::kChunk;
        int slot[2] = {0, 0};
        auto issue = [&](size_t c) {
            const size_t o = c * Stage::kChunk;
            const int i = st.take();
            slot[c & 1] = i;
            MMCSF_HIP(hipMemcpyAsync(st.buf[i], src + o, std::min(Stage::kChunk, total - o), hipMemcpyDeviceToHost,
                                     st.s));
            MMCSF_HIP(hipEventRecord(st.done[i], st.s));
            st.busy[i] = true;
        };
        if (chunks) issue(0);
        for (size_t c = 0; c < chunks; ++c) {
            const int i = slot[c & 1];
            MMCSF_HIP(hipEventSynchronize(st.done[i]));
            st.busy[i] = false;
            if (c + 1 < chunks) issue(c + 1);  // copies while this chunk is written
            const size_t o = c * Stage::kChunk;
            bytes(st.buf[i], std::min(Stage::kChunk, total - o));
        }
    }
    template <typename T>
    void vec(const std::vector<T> &v) {
        pod<uint64_t>(v.size());
        bytes(v.data(), v.size() * sizeof(T));
    }
};

struct Reader {
    FILE *f;
    uint64_t left;  // bytes not yet read
    Stage &st;
    void bytes(void *p, size_t n) {
        MMCSF_REQUIRE(n <= left, MMCSF_ERR_IO, "format cache: file is truncated");
        if (n && std::fread(p, 1, n, f) != n) fail(MMCSF_ERR_IO, "format cache: read failed");
        left -= n;
    }
    template <typename T>
    T pod() {
        T v;
        bytes(&v, sizeof(T));
        return v;
    }
    uint64_t count(size_t elem) {
        const uint64_t n = pod<uint64_t>();
        MMCSF_REQUIRE(n <= left / elem, MMCSF_ERR_IO, "format cache: array size past the end of the file");
        return n;
    }
    template <typename T>
    void dev(DevArray<T> &a) {
        const uint64_t n = count(sizeof(T));
        a.release();
        if (!n) return;
        a.alloc(n);
        const size_t total = n * sizeof(T);
        char *dst = reinterpret_cast<char *>(a.p);
        for (size_t o = 0; o < total; o += Stage::kChunk) {
            const size_t m = std::min(Stage::kChunk, total - o);
            const int i = st.take();
            bytes(st.buf[i], m);  // reads while the previous chunk copies
            MMCSF_HIP(hipMemcpyAsync(dst + o, st.buf[i], m, hipMemcpyHostToDevice, st.s));
            MMCSF_HIP(hipEventRecord(st.done[i], st.s));
            st.busy[i] = true;
        }
    }
    template <typename T>
    void vec(std::vector<T> &v) {
        const uint64_t n = count(sizeof(T));
        v.resize(n);
        bytes(v.data(), n * sizeof(T));
    }
};

// one traversal of a format's fields, shared by writing (W = Writer, const
// objects) and reading (W = Reader)
template <typename W, typename C>
void csf_fields(W &w, C &c) {
    auto &cc = const_cast<Csf &>(c);
    if constexpr (std::is_same_v<W, Writer>) {
        w.pod(c.ndims);
        w.bytes(c.order, sizeof(c.order));
        w.bytes(c.rows, sizeof(c.rows));
        w.pod(c.nnz);
        w.bytes(c.cnt, sizeof(c.cnt));
        w.pod(c.hasBins);
        w.bytes(c.binCount, sizeof(c.binCount));
        w.pod(c.dropped);
        for (int l = 0; l < kMaxDims; ++l) w.pod(c.hot[l].count);
    } else {
        cc.ndims = w.template pod<int>();
        w.bytes(cc.order, sizeof(cc.order));
        w.bytes(cc.rows, sizeof(cc.rows));
        cc.nnz = w.template pod<uint64_t>();
        w.bytes(cc.cnt, sizeof(cc.cnt));
        cc.hasBins = w.template pod<bool>();
        w.bytes(cc.binCount, sizeof(cc.binCount));
        cc.dropped = w.template pod<uint64_t>();
        for (int l = 0; l < kMaxDims; ++l) cc.hot[l].count = w.template pod<uint32_t>();
        MMCSF_REQUIRE(cc.ndims == 0 || cc.ndims == 3 || cc.ndims == 4, MMCSF_ERR_IO, "format cache: bad tree");
    }
    for (int l = 0; l < kMaxDims - 1; ++l) {
        w.dev(cc.ptr[l]);
        w.dev(cc.idx[l]);
    }
    w.dev(cc.leaf);
    w.dev(cc.vals);
    w.dev(cc.pos);
    w.dev(cc.fbrSlc);
    w.dev(cc.fbit);
    w.dev(cc.sbit);
    w.dev(cc.chunkCost);
    w.dev(cc.binList);
    for (int l = 0; l < kMaxDims; ++l) {
        w.dev(cc.hot[l].rows);
        w.dev(cc.hot[l].tgt);
    }
}

template <typename W, typename F>
void format_fields(W &w, F &f0) {
    auto &f = const_cast<Format &>(f0);
    constexpr bool wr = std::is_same_v<W, Writer>;
    auto scalar = [&](auto &v) {
        if constexpr (wr) w.pod(v);
        else v = w.template pod<std::remove_reference_t<decltype(v)>>();
    };
    scalar(f.type);
    scalar(f.ndims);
    w.bytes(f.dims, sizeof(f.dims));
    scalar(f.nnz);
    scalar(f.redunMode);
    scalar(f.fbrThreshold);
    w.bytes(f.cooOrder, sizeof(f.cooOrder));
    for (int l = 0; l < kMaxDims; ++l) w.dev(f.cooKeys[l]);
    w.dev(f.cooVals);
    csf_fields(w, f.csf);
    for (int m = 0; m < kMaxDims; ++m) {
        scalar(f.hasTree[m]);
        if (f.hasTree[m]) csf_fields(w, f.tree[m]);
    }
    scalar(f.mixed);
    w.bytes(f.partNnz, sizeof(f.partNnz));
    if (f.mixed) {
        for (int p = 0; p < kMaxDims; ++p)
            if (f.partNnz[p]) csf_fields(w, f.part[p]);
        w.vec(f.partPerNnz);
        w.vec(f.partOrder);
    }
    // HB-CSF parts
    Hyb &h = f.hyb;
    scalar(h.ndims);
    w.bytes(h.order, sizeof(h.order));
    scalar(h.cooNnz);
    scalar(h.cslNnz);
    scalar(h.nCsl);
    for (int l = 0; l < kMaxDims; ++l) w.dev(h.cooInd[l]);
    w.dev(h.cooVal);
    w.dev(h.cslPtr);
    w.dev(h.cslIdx);
    w.dev(h.cslInd1);
    w.dev(h.cslInd2);
    w.dev(h.cslVal);
    w.dev(h.cslBinList);
    w.bytes(h.cslBinCount, sizeof(h.cslBinCount));
    csf_fields(w, h.csf);
}

}  // namespace

void format_save(const Format &f, const char *path) {
    FILE *fp = std::fopen(path, "wb");
    MMCSF_REQUIRE(fp != nullptr, MMCSF_ERR_IO, std::string("cannot open ") + path + " for writing");
    struct Closer {
        FILE *f;
        ~Closer() { std::fclose(f); }
    } closer{fp};
    MMCSF_HIP(hipDeviceSynchronize());
    Stage st;
    st.init();
    Writer w{fp, st};
    w.bytes(kMagic, 8);
    w.pod(kVersion);
    w.pod(kByteOrder);
    format_fields(w, f);
    w.bytes(kEnd, 8);
    MMCSF_REQUIRE(std::fflush(fp) == 0, MMCSF_ERR_IO, "format cache: write failed");
}

bool is_format_cache(const char *path) {
    FILE *fp = std::fopen(path, "rb");
    if (!fp) return false;
    char m[8] = {0};
    const bool ok = std::fread(m, 1, 8, fp) == 8 && std::memcmp(m, kMagic, 8) == 0;
    std::fclose(fp);
    return ok;
}

void format_load(const char *path, Format &f) {
    FILE *fp = std::fopen(path, "rb");
    MMCSF_REQUIRE(fp != nullptr, MMCSF_ERR_IO, std::string("cannot open ") + path);
    struct Closer {
        FILE *f;
        ~Closer() { std::fclose(f); }
    } closer{fp};
    MMCSF_REQUIRE(std::fseek(fp, 0, SEEK_END) == 0, MMCSF_ERR_IO, "format cache: cannot seek");
    const long size = std::ftell(fp);
    MMCSF_REQUIRE(size >= 0 && std::fseek(fp, 0, SEEK_SET) == 0, MMCSF_ERR_IO, "format cache: cannot seek");
    Stage st;
    st.init();
    Reader r{fp, (uint64_t)size, st};
    char m[8];
    r.bytes(m, 8);
    MMCSF_REQUIRE(std::memcmp(m, kMagic, 8) == 0, MMCSF_ERR_IO, "not a format cache file");
    const uint32_t version = r.pod<uint32_t>();
    MMCSF_REQUIRE(version == kVersion, MMCSF_ERR_IO,
                  "format cache version " + std::to_string(version) + ", this library reads version " +
                      std::to_string(kVersion));
    MMCSF_REQUIRE(r.pod<uint32_t>() == kByteOrder, MMCSF_ERR_IO, "format cache written with another byte order");
    format_fields(r, f);
    st.drain();
    // derived from the stored CSL arrays: rebuilt, not stored
    build_csl_walk(f.hyb, st.s);
    r.bytes(m, 8);
    MMCSF_REQUIRE(std::memcmp(m, kEnd, 8) == 0 && r.left == 0, MMCSF_ERR_IO, "format cache: bad end of file");
    MMCSF_REQUIRE(f.ndims == 3 || f.ndims == 4, MMCSF_ERR_IO, "format cache: bad tensor order");
}

}  // namespace mmcsf
