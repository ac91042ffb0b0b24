// C ABI of the library (include/mmcsf.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cerrno>
#include <charconv>
#include <thread>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "mmcsf_internal.hpp"

namespace mmcsf {

static thread_local std::string g_last_error;

void set_last_error(const std::string &msg) { g_last_error = msg; }

namespace {

template <typename F>
int guarded(F &&fn) {
    try {
        fn();
        return MMCSF_OK;
    } catch (const Error &e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::bad_alloc &) {
        set_last_error("host allocation failed");
        return MMCSF_ERR_NOMEM;
    } catch (const std::exception &e) {
        set_last_error(e.what());
        return MMCSF_ERR_ARG;
    }
}

constexpr uint64_t kMaxIndex = 0x7fffffffull;  // device coordinates and counts are uint32 < 2^31

void check_shape(int ndims, const uint64_t *dims, uint64_t nnz) {
    // check_opt, util.h:238-243 (the reference accepts 3 or 4 modes)
    MMCSF_REQUIRE(ndims == 3 || ndims == 4, MMCSF_ERR_UNSUPPORTED, "Supported tensor dimension is 3 or 4.");
    for (int m = 0; m < ndims; ++m)
        MMCSF_REQUIRE(dims[m] >= 1 && dims[m] <= kMaxIndex, MMCSF_ERR_UNSUPPORTED,
                      "mode length must be in 1..2^31-1");
    MMCSF_REQUIRE(nnz <= kMaxIndex, MMCSF_ERR_UNSUPPORTED, "nnz must be below 2^31");
}

__global__ void k_u32_check(const uint32_t *in, uint64_t n, uint64_t lim, uint32_t *bad) {
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i < n && in[i] >= lim) atomicOr(bad, 1u);
}

// ---- .tns text and binary COO files ---------------------------------------

void read_file(const char *path, std::string &text) {
    FILE *fp = std::fopen(path, "rb");
    MMCSF_REQUIRE(fp, MMCSF_ERR_IO, std::string(path) + " does not exist!");
    std::unique_ptr<FILE, int (*)(FILE *)> guard(fp, std::fclose);
    MMCSF_REQUIRE(std::fseek(fp, 0, SEEK_END) == 0, MMCSF_ERR_IO, "cannot seek");
    long n = std::ftell(fp);
    MMCSF_REQUIRE(n >= 0, MMCSF_ERR_IO, "cannot size the file");
    std::rewind(fp);
    text.resize((size_t)n);
    MMCSF_REQUIRE(std::fread(text.data(), 1, (size_t)n, fp) == (size_t)n, MMCSF_ERR_IO, "read failed");
}

inline bool is_space(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\v' || c == '\f'; }

// unsigned decimal after optional white space; false at end of text or on a non-digit
inline bool parse_u(const char *&p, const char *e, uint64_t &v) {
    while (p < e && is_space(*p)) ++p;
    auto r = std::from_chars(p, e, v);
    if (r.ec != std::errc()) return false;
    p = r.ptr;
    return true;
}

inline bool parse_f(const char *&p, const char *e, float &v) {
    while (p < e && is_space(*p)) ++p;
    if (p < e && *p == '+') ++p;  // strtof accepts a leading '+', from_chars does not
    auto r = std::from_chars(p, e, v);
    if (r.ec == std::errc::result_out_of_range) {  // strtof semantics: +-inf / denormal / 0
        char tmp[64];
        size_t n = std::min<size_t>(sizeof tmp - 1, (size_t)(e - p));
        std::memcpy(tmp, p, n);
        tmp[n] = 0;
        char *end = nullptr;
        v = std::strtof(tmp, &end);
        p += end - tmp;
        return end != tmp;
    }
    if (r.ec != std::errc()) return false;
    p = r.ptr;
    return true;
}

// load_tensor's header (util.h:313-329): number of modes, then the mode lengths
size_t parse_tns_header(const std::string &text, int &ndims, uint64_t *dims) {
    const char *p = text.data(), *e = p + text.size();
    uint64_t nd = 0;
    MMCSF_REQUIRE(parse_u(p, e, nd), MMCSF_ERR_IO, "cannot read the number of modes");
    MMCSF_REQUIRE(nd == 3 || nd == 4, MMCSF_ERR_UNSUPPORTED, "Supported tensor dimension is 3 or 4.");
    ndims = (int)nd;
    for (int m = 0; m < ndims; ++m) MMCSF_REQUIRE(parse_u(p, e, dims[m]), MMCSF_ERR_IO, "cannot read mode lengths");
    return (size_t)(p - text.data());
}

// Nonzero lines (util.h:335-365): coordinates 1-based -> 0-based, value as
// float.  The text is cut at line ends into one chunk per host thread; the
// chunks are parsed concurrently and concatenated in file order.
void parse_tns_body(const std::string &text, size_t begin, int ndims, const uint64_t *dims,
                    std::vector<std::vector<uint32_t>> &inds, std::vector<float> &vals) {
    const size_t n = text.size();
    unsigned T = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (n - begin < (1u << 20)) T = 1;
    std::vector<size_t> cut(T + 1, n);
    cut[0] = begin;
    for (unsigned k = 1; k < T; ++k) {
        size_t c = std::max(cut[k - 1], begin + (n - begin) * k / T);
        while (c < n && text[c] != '\n') ++c;
        cut[k] = c;
    }
    struct Part {
        std::vector<uint32_t> ind[kMaxDims];
        std::vector<float> val;
        std::string err;
    };
    std::vector<Part> parts(T);
    auto work = [&](unsigned k) {
        Part &P = parts[k];
        const char *p = text.data() + cut[k], *e = text.data() + cut[k + 1];
        const size_t guess = (size_t)(e - p) / (8 * ndims + 8) + 16;
        for (int m = 0; m < ndims; ++m) P.ind[m].reserve(guess);
        P.val.reserve(guess);
        for (;;) {
            uint64_t c[kMaxDims];
            if (!parse_u(p, e, c[0])) {
                while (p < e && is_space(*p)) ++p;
                if (p != e) P.err = "malformed nonzero line";
                return;
            }
            for (int m = 1; m < ndims; ++m)
                if (!parse_u(p, e, c[m])) { P.err = "truncated nonzero line"; return; }
            float v;
            if (!parse_f(p, e, v)) { P.err = "missing value"; return; }
            for (int m = 0; m < ndims; ++m) {
                if (c[m] < 1 || c[m] > dims[m]) { P.err = "coordinate out of range"; return; }
                P.ind[m].push_back((uint32_t)(c[m] - 1));
            }
            P.val.push_back(v);
        }
    };
    std::vector<std::thread> th;
    for (unsigned k = 1; k < T; ++k) th.emplace_back(work, k);
    work(0);
    for (auto &t : th) t.join();
    size_t total = 0;
    for (auto &P : parts) {
        MMCSF_REQUIRE(P.err.empty(), MMCSF_ERR_IO, P.err);
        total += P.val.size();
    }
    for (int m = 0; m < ndims; ++m) {
        inds[m].clear();
        inds[m].reserve(total);
        for (auto &P : parts) inds[m].insert(inds[m].end(), P.ind[m].begin(), P.ind[m].end());
    }
    vals.clear();
    vals.reserve(total);
    for (auto &P : parts) vals.insert(vals.end(), P.val.begin(), P.val.end());
}

constexpr char kBinMagic[8] = {'M', 'M', 'C', 'S', 'F', 'C', 'O', 'O'};
struct BinHeader {  // little-endian; followed by ndims x nnz uint32 coordinates, nnz float values
    char magic[8];
    uint32_t version;
    uint32_t ndims;
    uint64_t nnz;
    uint64_t dims[kMaxDims];
};

void upload_tensor(int ndims, const uint64_t *dims, uint64_t nnz, const std::vector<std::vector<uint32_t>> &inds,
                   const float *vals, Tensor &t) {
    t.ndims = ndims;
    t.nnz = nnz;
    for (int m = 0; m < ndims; ++m) {
        t.dims[m] = dims[m];
        t.inds[m].alloc(nnz);
        if (nnz) MMCSF_HIP(hipMemcpy(t.inds[m].p, inds[m].data(), nnz * 4, hipMemcpyHostToDevice));
    }
    t.vals.alloc(nnz);
    if (nnz) MMCSF_HIP(hipMemcpy(t.vals.p, vals, nnz * 4, hipMemcpyHostToDevice));
}

int level_of(const int *order, int ndims, int mode) {
    for (int l = 0; l < ndims; ++l)
        if (order[l] == mode) return l;
    return -1;
}

template <typename T>
std::vector<T> to_host(const DevArray<T> &d, uint64_t n) {
    std::vector<T> h(n);
    if (n) MMCSF_HIP(hipMemcpy(h.data(), d.p, n * sizeof(T), hipMemcpyDeviceToHost));
    return h;
}

void download_tensor(const Tensor &t, std::vector<std::vector<uint32_t>> &inds, std::vector<float> &vals) {
    inds.resize(t.ndims);
    for (int m = 0; m < t.ndims; ++m) inds[m] = to_host(t.inds[m], t.nnz);
    vals = to_host(t.vals, t.nnz);
}

// Coordinates of level l for every nonzero of a CSF tree (expands the tree).
std::vector<uint32_t> expand_level(const Csf &c, int l) {
    const int N = c.ndims;
    if (l == N - 1) return to_host(c.leaf, c.nnz);
    std::vector<uint32_t> out(c.nnz);
    std::vector<std::vector<uint32_t>> ptr(N - 1);
    for (int q = 0; q < N - 1; ++q) ptr[q] = to_host(c.ptr[q], c.cnt[q] + 1);
    std::vector<uint32_t> idx = to_host(c.idx[l], c.cnt[l]);
    for (uint64_t e = 0; e < c.cnt[l]; ++e) {
        uint64_t lo = e, hi = e + 1;
        for (int q = l; q < N - 1; ++q) { lo = ptr[q][lo]; hi = ptr[q][hi]; }
        for (uint64_t x = lo; x < hi; ++x) out[x] = idx[e];
    }
    return out;
}

// the nonzeros a format stores, per tensor mode, in its stored order (HB-CSF
// parts and mixed-mode partitions one after another)
void format_nonzeros(const Format &f, std::vector<std::vector<uint32_t>> &inds, std::vector<float> &vals) {
    const int N = f.ndims;
    inds.assign(N, {});
    vals.clear();
    auto add_csf = [&](const Csf &c) {
        if (!c.nnz) return;
        for (int m = 0; m < N; ++m) {
            std::vector<uint32_t> v = expand_level(c, level_of(c.order, N, m));
            inds[m].insert(inds[m].end(), v.begin(), v.end());
        }
        std::vector<float> v = to_host(c.vals, c.nnz);
        vals.insert(vals.end(), v.begin(), v.end());
    };
    if (f.type == MMCSF_TYPE_COO || f.type == MMCSF_TYPE_COO_CPU) {
        for (int m = 0; m < N; ++m) inds[m] = to_host(f.cooKeys[level_of(f.cooOrder, N, m)], f.nnz);
        vals = to_host(f.cooVals, f.nnz);
    } else if (f.type == MMCSF_TYPE_HBCSF) {
        const Hyb &h = f.hyb;
        for (int m = 0; m < N; ++m) inds[m] = to_host(h.cooInd[level_of(h.order, N, m)], h.cooNnz);
        vals = to_host(h.cooVal, h.cooNnz);
        if (h.cslNnz) {
            const std::vector<uint32_t> ptr = to_host(h.cslPtr, h.nCsl + 1), idx = to_host(h.cslIdx, h.nCsl);
            std::vector<uint32_t> slc(h.cslNnz);
            for (uint64_t q = 0; q < h.nCsl; ++q)
                for (uint32_t x = ptr[q]; x < ptr[q + 1]; ++x) slc[x] = idx[q];
            const std::vector<uint32_t> i1 = to_host(h.cslInd1, h.cslNnz), i2 = to_host(h.cslInd2, h.cslNnz);
            const std::vector<uint32_t> *lv[3] = {&slc, &i1, &i2};
            for (int l = 0; l < 3; ++l) inds[h.order[l]].insert(inds[h.order[l]].end(), lv[l]->begin(), lv[l]->end());
            const std::vector<float> v = to_host(h.cslVal, h.cslNnz);
            vals.insert(vals.end(), v.begin(), v.end());
        }
        add_csf(h.csf);
    } else if (f.mixed) {
        for (int p = 0; p < N; ++p) add_csf(f.part[p]);
    } else {
        add_csf(f.csf);
    }
}

// one tree built by create_HCSR with the -f split, every mode served from it
bool one_tree_type(int type) {
    return type == MMCSF_TYPE_BCSF_ANYMODE || type == MMCSF_TYPE_SINGLECSF || type == MMCSF_TYPE_ONECSF;
}

// the tree of create_HCSR (-t 8 and the one-tree types)
bool bcsf_family(int type) { return type == MMCSF_TYPE_BCSF || one_tree_type(type); }

bool mmcsf_type_ok(int type) {
    return type == MMCSF_TYPE_COO_CPU || type == MMCSF_TYPE_COO || type == MMCSF_TYPE_BCSF ||
           type == MMCSF_TYPE_BCSF_ANYMODE || type == MMCSF_TYPE_ONECSF || type == MMCSF_TYPE_SINGLECSF ||
           type == MMCSF_TYPE_HBCSF || type == MMCSF_TYPE_MMCSF;
}

int redundant_mode(int ndims, const uint64_t *dims) {
    // mttkrp.cu:315-347: the last m whose (order[0], order[1]) swaps an
    // earlier mode's pair
    int orders[kMaxDims][kMaxDims];
    for (int m = 0; m < ndims; ++m) order_tensormode(ndims, dims, m, false, orders[m]);
    int red = -1;
    for (int m = 1; m < ndims; ++m)
        for (int q = 0; q < m; ++q)
            if (orders[m][1] == orders[q][0] && orders[m][0] == orders[q][1]) red = m;
    return red;
}

// MM-CSF with the mixed-mode partition: mm_partition_reuseBased
// (util.h:2167-2512) with the heuristic's own choice instead of the override
// to partition 2 (util.h:2517), then populate_paritions, sort_MI_CSF,
// create_TiledHCSR and create_fbrLikeSlcInds per partition
// (mttkrp.cu:306-395).  The level-1 entry of every nonzero in every mode's
// CSF comes from GPU sort / scan passes; the decisions then run on the host
// in the order the reference's loop walks the tensor, because each one
// reads the fiber counts that every earlier decision decremented.
void build_mixed(const Tensor &t, const mmcsf_options *opt, bool nat, Format &f, hipStream_t s) {
    const int N = t.ndims;
    const uint64_t n = t.nnz;
    const uint64_t kNoSplit = 999999999ull;
    int orders[kMaxDims][kMaxDims];
    for (int m = 0; m < N; ++m) order_tensormode(N, t.dims, m, nat, orders[m]);
    const int red = f.redunMode;
    // same-fiber pairs (util.h:2261-2310): mode m pairs with the first q < m
    // whose (order[0], order[1]) it swaps; the last pair sets shortMode / longMode
    bool same[kMaxDims][kMaxDims] = {};
    int shortMode = 0, longMode = -1;
    for (int m = 1; m < N; ++m)
        for (int q = 0; q < m; ++q)
            if (orders[m][1] == orders[q][0] && orders[m][0] == orders[q][1]) {
                same[q][m] = true;
                shortMode = t.dims[q] <= t.dims[m] ? q : m;
                longMode = t.dims[q] <= t.dims[m] ? m : q;
                break;
            }
    // shortestMode (util.h:2253) from the loaded tensor's order, which
    // load_tensor takes for -m (util.h:367)
    int o0[kMaxDims];
    order_tensormode(N, t.dims, (int)opt->mode, nat, o0);
    const int shortestMode = t.dims[o0[0]] <= t.dims[o0[1]] ? o0[0] : o0[1];
    // every non-redundant mode's CSF without the -f split (create_HCSR forces
    // 99999999 for type 12, util.h:763-764): the level-1 entry of each nonzero
    // and the nonzeros under each entry.  The arrays of the reference's
    // tensors are shared (util.h:380-381), so its loop walks the tensor in the
    // order of the last of these sorts.
    // nnzPerFiber is ITYPE (64-bit) in the reference: a fiber that the
    // last-slice search stands in for can be decremented below zero and wraps
    std::vector<uint32_t> id[kMaxDims], walk;
    std::vector<uint64_t> cnt[kMaxDims];
    for (int m = 0; m < N; ++m) {
        if (m == red) continue;
        DevArray<uint32_t> keys[kMaxDims], pos, idT, cntD;
        DevArray<float> vals;
        sort_coo(t, orders[m], keys, vals, s, &pos);
        Csf c;
        build_csf(N, t.dims, n, orders[m], keys, std::move(vals), 99999999ull, kNoSplit, c, s);
        level1_ids(c, pos.p, idT, cntD, s);
        id[m] = to_host(idT, n);
        const std::vector<uint32_t> c32 = to_host(cntD, c.cnt[1] + 1);
        cnt[m].assign(c32.begin(), c32.end());
        walk = to_host(pos, n);
    }
    // the decisions, in walk order (util.h:2330-2512)
    const uint64_t th = (uint64_t)(int64_t)(int)opt->MIfbTh;  // int fbTh = Opt.MIfbTh
    std::vector<uint8_t> partT(n);
    f.partPerNnz.assign(n, 0);
    for (uint64_t k = 0; k < n; ++k) {
        const uint32_t x = walk[k];
        uint64_t len[kMaxDims];
        uint32_t fno[kMaxDims] = {0, 0, 0, 0};
        bool copied[kMaxDims] = {false, false, false, false};
        for (int m = 0; m < N; ++m) {
            for (int q = 0; q < m && !copied[m]; ++q)
                if (same[q][m]) {
                    len[m] = len[q];  // the second mode of a pair takes the first's count
                    copied[m] = true;
                }
            if (copied[m]) continue;
            fno[m] = id[m][x];
            MMCSF_REQUIRE(fno[m] != 0xFFFFFFFFu, MMCSF_ERR_UNSUPPORTED,
                          "mixed-mode partition: a mode's CSF has one slice, and the reference's search for "
                          "a nonzero outside its first fiber reads before the fiber array");
            len[m] = cnt[m][fno[m]];
        }
        int mode = -1;
        for (int m = 0; m < N && mode < 0; ++m) {
            uint64_t mx = 0;
            for (int q = 0; q < N; ++q)
                if (q != m) mx = std::max(mx, len[q]);
            if (len[m] >= th * mx) {
                bool pairAbove = false;
                for (int q = m + 1; q < N; ++q) pairAbove = pairAbove || same[m][q];
                mode = pairAbove ? shortMode : m;
            }
        }
        if (mode < 0) mode = shortestMode;
        // the counts of the fibers this nonzero leaves (util.h:2493-2512);
        // a copied mode has no fiber of its own
        for (int i = 0; i < N; ++i) {
            if (copied[i]) continue;
            const bool dec = mode == shortMode ? (i != shortMode && i != longMode) : (i != red && i != mode);
            if (dec) cnt[i][fno[i]]--;
        }
        partT[x] = (uint8_t)mode;
        f.partPerNnz[k] = (uint32_t)mode;
    }
    for (int m = 0; m < N; ++m) {
        id[m].clear();
        id[m].shrink_to_fit();
    }
    // the partitions: their nonzeros in the canonical order of their own
    // mode order, a CSF with the -f split and, 4-D, the fbrS split at 128
    for (int p = 0; p < N; ++p) {
        std::vector<uint32_t> sel;
        for (uint64_t x = 0; x < n; ++x)
            if (partT[x] == p) sel.push_back((uint32_t)x);
        f.partNnz[p] = sel.size();
        if (sel.empty()) continue;
        DevArray<uint32_t> selD(sel.size());
        MMCSF_HIP(hipMemcpyAsync(selD.p, sel.data(), sel.size() * 4, hipMemcpyHostToDevice, s));
        Tensor sub;
        gather_tensor(t, selD.p, sel.size(), sub, s);
        DevArray<uint32_t> keys[kMaxDims];
        DevArray<float> vals;
        sort_coo(sub, orders[p], keys, vals, s);
        const uint64_t fs = N == 4 ? 128 : kNoSplit;  // util.h:1325-1326
        build_csf(N, t.dims, sub.nnz, orders[p], keys, std::move(vals), opt->fbrThreashold, fs, f.part[p], s);
        for (int l = 1; l < N; ++l) build_hot_table(f.part[p], l, t.dims[orders[p][l]], f.part[p].hot[l], s);
        MMCSF_HIP(hipStreamSynchronize(s));
    }
    f.partOrder = std::move(walk);
    f.mixed = true;
}

}  // namespace
}  // namespace mmcsf

using namespace mmcsf;

namespace {
template <typename T>
uint64_t bytes_of(const mmcsf::DevArray<T> &a) { return a.n * sizeof(T); }

uint64_t csf_bytes(const mmcsf::Csf &c) {
    uint64_t b = bytes_of(c.leaf) + bytes_of(c.vals) + bytes_of(c.pos) + bytes_of(c.fbrSlc) + bytes_of(c.fbit) +
                 bytes_of(c.sbit) + bytes_of(c.chunkCost) + bytes_of(c.binList);
    for (int l = 0; l < mmcsf::kMaxDims - 1; ++l) b += bytes_of(c.ptr[l]) + bytes_of(c.idx[l]);
    for (int l = 0; l < mmcsf::kMaxDims; ++l) b += bytes_of(c.hot[l].rows) + bytes_of(c.hot[l].tgt);
    return b;
}

uint64_t format_bytes(const mmcsf::Format &f) {
    uint64_t b = csf_bytes(f.csf) + bytes_of(f.cooVals) + csf_bytes(f.hyb.csf) + csf_bytes(f.hyb.cslWalk) + bytes_of(f.hyb.cooVal) +
                 bytes_of(f.hyb.cslPtr) + bytes_of(f.hyb.cslIdx) + bytes_of(f.hyb.cslInd1) +
                 bytes_of(f.hyb.cslInd2) + bytes_of(f.hyb.cslVal) + bytes_of(f.hyb.cslBinList);
    for (int l = 0; l < mmcsf::kMaxDims; ++l) {
        b += bytes_of(f.cooKeys[l]) + bytes_of(f.hyb.cooInd[l]);
        if (f.hasTree[l]) b += csf_bytes(f.tree[l]);
        if (f.mixed) b += csf_bytes(f.part[l]);
    }
    return b;
}
}  // namespace

extern "C" {

const char *mmcsf_last_error(void) { return g_last_error.c_str(); }

const char *mmcsf_version(void) { return "mmcsf-mi355x 0.1 (gfx950)"; }

int mmcsf_options_default(mmcsf_options *o) {
    if (!o) { set_last_error("null options"); return MMCSF_ERR_ARG; }
    std::memset(o, 0, sizeof(*o));
    // Options member initialisers, util.h:189-211
    o->R = 32;
    o->mode = 0;
    o->impType = 1;
    o->warpPerSlice = 4;
    o->nTile = 1;
    o->gridSize = 512;
    o->TBsize = 128;
    o->MIfbTh = 1;
    o->fiberPerWarp = 1;
    o->verbose = 0;
    o->correctness = 0;
    o->nBin = 10;
    o->cpdIters = 10;
    o->natOrdering = 0;
    o->fbrThreashold = 99999999;
    o->factorDtype = MMCSF_F32;
    o->seed = 1;
    o->modeTrees = 1;
    o->mixedPart = 0;
    std::strcpy(o->m0, "012");
    std::strcpy(o->m1, "120");
    std::strcpy(o->m2, "201");
    return MMCSF_OK;
}

int mmcsf_options_parse(int argc, char **argv, mmcsf_options *o) {
    return guarded([&] {
        MMCSF_REQUIRE(o != nullptr, MMCSF_ERR_ARG, "null options");
        mmcsf_options_default(o);
        // parse_cmd_options, util.h:3341-3442.  atoi semantics: ITYPE fields
        // receive the int converted to size_t.
        auto as_u64 = [](const char *s) { return (uint64_t)(int64_t)std::atoi(s); };
        int i;
        for (i = 1; i < argc; i++) {
            if (argv[i][0] != '-') break;
            if (++i >= argc) fail(MMCSF_ERR_ARG, std::string("option ") + argv[i - 1] + " requires a value");
            const char *v = argv[i];
            switch (argv[i - 1][1]) {
                case 'R': o->R = as_u64(v); break;
                case 'm': o->mode = as_u64(v); break;
                case 't': o->impType = as_u64(v); break;
                case 'w': o->warpPerSlice = as_u64(v); break;
                case 'l': o->nTile = as_u64(v); break;
                case 'f': o->fbrThreashold = as_u64(v); break;
                case 'b': o->TBsize = as_u64(v); break;
                case 's': o->fiberPerWarp = as_u64(v); break;
                case 'h': o->MIfbTh = as_u64(v); break;
                case 'g': o->gridSize = as_u64(v); break;
                case 'v': o->verbose = std::atoi(v) == 1; break;
                case 'c': o->correctness = std::atoi(v) == 1; break;
                case 'd': o->factorDtype = std::atoi(v) == 1 ? MMCSF_BF16 : MMCSF_F32; break;
                case 'n': o->cpdIters = as_u64(v); break;
                case 'a': o->modeTrees = std::atoi(v) != 0; break;
                case 'M': o->mixedPart = std::atoi(v) != 0; break;
                case 'x': o->seed = std::strtoull(v, nullptr, 0); break;
                case 'i':
                    MMCSF_REQUIRE(std::strlen(v) < sizeof(o->inFileName), MMCSF_ERR_ARG, "input path too long");
                    std::strcpy(o->inFileName, v);
                    break;
                case 'o':
                    MMCSF_REQUIRE(std::strlen(v) < sizeof(o->outFileName), MMCSF_ERR_ARG, "output path too long");
                    std::strcpy(o->outFileName, v);
                    break;
                case 'k':
                    MMCSF_REQUIRE(std::strlen(v) < sizeof(o->cacheFileName), MMCSF_ERR_ARG, "cache path too long");
                    std::strcpy(o->cacheFileName, v);
                    break;
                case 'p': std::snprintf(o->m0, sizeof(o->m0), "%s", v); break;
                case 'q': std::snprintf(o->m1, sizeof(o->m1), "%s", v); break;
                case 'r': std::snprintf(o->m2, sizeof(o->m2), "%s", v); break;
                default: fail(MMCSF_ERR_ARG, std::string("unknown option: ") + argv[i - 1]);
            }
        }
    });
}

int mmcsf_tensor_load_tns(const char *path, mmcsf_tensor **out) {
    return guarded([&] {
        MMCSF_REQUIRE(path && out, MMCSF_ERR_ARG, "null argument");
        std::string text;
        read_file(path, text);
        int ndims = 0;
        uint64_t dims[kMaxDims] = {0, 0, 0, 0};
        size_t body = parse_tns_header(text, ndims, dims);
        std::vector<std::vector<uint32_t>> inds(ndims);
        std::vector<float> vals;
        parse_tns_body(text, body, ndims, dims, inds, vals);
        check_shape(ndims, dims, vals.size());
        auto h = std::make_unique<mmcsf_tensor>();
        upload_tensor(ndims, dims, vals.size(), inds, vals.data(), h->t);
        *out = h.release();
    });
}

int mmcsf_tensor_save_tns(const mmcsf_tensor *th, const char *path) {
    return guarded([&] {
        MMCSF_REQUIRE(th && path, MMCSF_ERR_ARG, "null argument");
        const Tensor &t = th->t;
        std::vector<std::vector<uint32_t>> inds;
        std::vector<float> vals;
        download_tensor(t, inds, vals);
        FILE *fp = std::fopen(path, "wb");
        MMCSF_REQUIRE(fp, MMCSF_ERR_IO, std::string("cannot open ") + path);
        std::string line;
        line.reserve(128);
        bool okw = std::fprintf(fp, "%d\n", t.ndims) > 0;
        for (int m = 0; m < t.ndims; ++m) okw = okw && std::fprintf(fp, m ? " %llu" : "%llu", (unsigned long long)t.dims[m]) > 0;
        okw = okw && std::fputc('\n', fp) != EOF;
        char buf[64];
        for (uint64_t x = 0; x < t.nnz && okw; ++x) {
            line.clear();
            for (int m = 0; m < t.ndims; ++m) {
                auto r = std::to_chars(buf, buf + sizeof buf, (unsigned long long)inds[m][x] + 1);
                line.append(buf, r.ptr);
                line.push_back(' ');
            }
            // shortest representation that reads back to the same float
            auto r = std::to_chars(buf, buf + sizeof buf, vals[x]);
            line.append(buf, r.ptr);
            line.push_back('\n');
            okw = std::fwrite(line.data(), 1, line.size(), fp) == line.size();
        }
        okw = (std::fclose(fp) == 0) && okw;
        MMCSF_REQUIRE(okw, MMCSF_ERR_IO, std::string("write failed: ") + path);
    });
}

int mmcsf_tensor_save_bin(const mmcsf_tensor *th, const char *path) {
    return guarded([&] {
        MMCSF_REQUIRE(th && path, MMCSF_ERR_ARG, "null argument");
        const Tensor &t = th->t;
        std::vector<std::vector<uint32_t>> inds;
        std::vector<float> vals;
        download_tensor(t, inds, vals);
        BinHeader hd{};
        std::memcpy(hd.magic, kBinMagic, 8);
        hd.version = 1;
        hd.ndims = (uint32_t)t.ndims;
        hd.nnz = t.nnz;
        for (int m = 0; m < t.ndims; ++m) hd.dims[m] = t.dims[m];
        FILE *fp = std::fopen(path, "wb");
        MMCSF_REQUIRE(fp, MMCSF_ERR_IO, std::string("cannot open ") + path);
        bool okw = std::fwrite(&hd, sizeof hd, 1, fp) == 1;
        for (int m = 0; m < t.ndims && okw; ++m)
            okw = std::fwrite(inds[m].data(), 4, t.nnz, fp) == t.nnz;
        okw = okw && std::fwrite(vals.data(), 4, t.nnz, fp) == t.nnz;
        okw = (std::fclose(fp) == 0) && okw;
        MMCSF_REQUIRE(okw, MMCSF_ERR_IO, std::string("write failed: ") + path);
    });
}

int mmcsf_tensor_load_bin(const char *path, mmcsf_tensor **out) {
    return guarded([&] {
        MMCSF_REQUIRE(path && out, MMCSF_ERR_ARG, "null argument");
        FILE *fp = std::fopen(path, "rb");
        MMCSF_REQUIRE(fp, MMCSF_ERR_IO, std::string(path) + " does not exist!");
        std::unique_ptr<FILE, int (*)(FILE *)> guard(fp, std::fclose);
        BinHeader hd{};
        MMCSF_REQUIRE(std::fread(&hd, sizeof hd, 1, fp) == 1 && std::memcmp(hd.magic, kBinMagic, 8) == 0 &&
                          hd.version == 1,
                      MMCSF_ERR_IO, std::string(path) + " is not a binary COO tensor of this library");
        MMCSF_REQUIRE(hd.ndims == 3 || hd.ndims == 4, MMCSF_ERR_UNSUPPORTED, "Supported tensor dimension is 3 or 4.");
        const int ndims = (int)hd.ndims;
        check_shape(ndims, hd.dims, hd.nnz);
        std::vector<std::vector<uint32_t>> inds(ndims, std::vector<uint32_t>(hd.nnz));
        std::vector<float> vals(hd.nnz);
        for (int m = 0; m < ndims; ++m) {
            MMCSF_REQUIRE(std::fread(inds[m].data(), 4, hd.nnz, fp) == hd.nnz, MMCSF_ERR_IO, "truncated file");
            for (uint32_t c : inds[m]) MMCSF_REQUIRE(c < hd.dims[m], MMCSF_ERR_IO, "coordinate out of range");
        }
        MMCSF_REQUIRE(std::fread(vals.data(), 4, hd.nnz, fp) == hd.nnz, MMCSF_ERR_IO, "truncated file");
        auto h = std::make_unique<mmcsf_tensor>();
        upload_tensor(ndims, hd.dims, hd.nnz, inds, vals.data(), h->t);
        *out = h.release();
    });
}

int mmcsf_tensor_from_host(int ndims, const uint64_t *dims, uint64_t nnz, const uint64_t *const *inds,
                           const float *vals, mmcsf_tensor **out) {
    return guarded([&] {
        MMCSF_REQUIRE(dims && out && (nnz == 0 || (inds && vals)), MMCSF_ERR_ARG, "null argument");
        check_shape(ndims, dims, nnz);
        std::vector<std::vector<uint32_t>> h(ndims);
        for (int m = 0; m < ndims; ++m) {
            h[m].resize(nnz);
            for (uint64_t x = 0; x < nnz; ++x) {
                MMCSF_REQUIRE(inds[m][x] < dims[m], MMCSF_ERR_ARG, "coordinate out of range");
                h[m][x] = (uint32_t)inds[m][x];
            }
        }
        auto t = std::make_unique<mmcsf_tensor>();
        upload_tensor(ndims, dims, nnz, h, vals, t->t);
        *out = t.release();
    });
}

int mmcsf_tensor_from_device(int ndims, const uint64_t *dims, uint64_t nnz, const uint32_t *const *inds,
                             const float *vals, void *stream, mmcsf_tensor **out) {
    return guarded([&] {
        MMCSF_REQUIRE(dims && out && (nnz == 0 || (inds && vals)), MMCSF_ERR_ARG, "null argument");
        check_shape(ndims, dims, nnz);
        hipStream_t s = static_cast<hipStream_t>(stream);
        auto t = std::make_unique<mmcsf_tensor>();
        t->t.ndims = ndims;
        t->t.nnz = nnz;
        DevArray<uint32_t> bad(1);
        MMCSF_HIP(hipMemsetAsync(bad.p, 0, 4, s));
        for (int m = 0; m < ndims; ++m) {
            t->t.dims[m] = dims[m];
            t->t.inds[m].alloc(nnz);
            if (nnz) {
                MMCSF_HIP(hipMemcpyAsync(t->t.inds[m].p, inds[m], nnz * 4, hipMemcpyDeviceToDevice, s));
                hipLaunchKernelGGL(k_u32_check, dim3((unsigned)((nnz + 255) / 256)), dim3(256), 0, s,
                                   t->t.inds[m].p, nnz, dims[m], bad.p);
            }
        }
        t->t.vals.alloc(nnz);
        if (nnz) MMCSF_HIP(hipMemcpyAsync(t->t.vals.p, vals, nnz * 4, hipMemcpyDeviceToDevice, s));
        uint32_t hb = 0;
        MMCSF_HIP(hipMemcpyAsync(&hb, bad.p, 4, hipMemcpyDeviceToHost, s));
        MMCSF_HIP(hipStreamSynchronize(s));
        MMCSF_REQUIRE(hb == 0, MMCSF_ERR_ARG, "coordinate out of range");
        *out = t.release();
    });
}

int mmcsf_tensor_generate(int ndims, const uint64_t *dims, uint64_t nnz, const int *law, const double *alpha,
                          uint64_t seed, void *stream, mmcsf_tensor **out) {
    return guarded([&] {
        MMCSF_REQUIRE(dims && law && alpha && out, MMCSF_ERR_ARG, "null argument");
        check_shape(ndims, dims, nnz);
        long double space = 1;
        for (int m = 0; m < ndims; ++m) space *= (long double)dims[m];
        MMCSF_REQUIRE((long double)nnz <= space, MMCSF_ERR_ARG, "more nonzeros than coordinates");
        auto t = std::make_unique<mmcsf_tensor>();
        generate_tensor(ndims, dims, nnz, law, alpha, seed, t->t, static_cast<hipStream_t>(stream));
        *out = t.release();
    });
}

int mmcsf_tensor_info(const mmcsf_tensor *t, int *ndims, uint64_t *dims, uint64_t *nnz) {
    return guarded([&] {
        MMCSF_REQUIRE(t != nullptr, MMCSF_ERR_ARG, "null tensor");
        if (ndims) *ndims = t->t.ndims;
        if (dims)
            for (int m = 0; m < t->t.ndims; ++m) dims[m] = t->t.dims[m];
        if (nnz) *nnz = t->t.nnz;
    });
}

int mmcsf_tensor_device_arrays(const mmcsf_tensor *t, const uint32_t **inds, const float **vals) {
    return guarded([&] {
        MMCSF_REQUIRE(t != nullptr, MMCSF_ERR_ARG, "null tensor");
        if (inds)
            for (int m = 0; m < t->t.ndims; ++m) inds[m] = t->t.inds[m].p;
        if (vals) *vals = t->t.vals.p;
    });
}

int mmcsf_tensor_copy_to_host(const mmcsf_tensor *t, uint64_t *const *inds, float *vals) {
    return guarded([&] {
        MMCSF_REQUIRE(t != nullptr, MMCSF_ERR_ARG, "null tensor");
        for (int m = 0; m < t->t.ndims; ++m)
            if (inds && inds[m]) {
                auto h = to_host(t->t.inds[m], t->t.nnz);
                for (uint64_t x = 0; x < t->t.nnz; ++x) inds[m][x] = h[x];
            }
        if (vals && t->t.nnz) MMCSF_HIP(hipMemcpy(vals, t->t.vals.p, t->t.nnz * 4, hipMemcpyDeviceToHost));
    });
}

int mmcsf_tensor_copy_to_device(const mmcsf_tensor *t, uint32_t *const *inds, float *vals, void *stream) {
    return guarded([&] {
        MMCSF_REQUIRE(t != nullptr, MMCSF_ERR_ARG, "null tensor");
        hipStream_t s = static_cast<hipStream_t>(stream);
        const uint64_t n = t->t.nnz;
        for (int m = 0; m < t->t.ndims; ++m)
            if (inds && inds[m] && n)
                MMCSF_HIP(hipMemcpyAsync(inds[m], t->t.inds[m].p, n * 4, hipMemcpyDeviceToDevice, s));
        if (vals && n) MMCSF_HIP(hipMemcpyAsync(vals, t->t.vals.p, n * 4, hipMemcpyDeviceToDevice, s));
    });
}

void mmcsf_tensor_free(mmcsf_tensor *t) { delete t; }

int mmcsf_format_build(const mmcsf_tensor *th, int type, const mmcsf_options *opt, void *stream,
                       mmcsf_format **out) {
    return guarded([&] {
        MMCSF_REQUIRE(th && opt && out, MMCSF_ERR_ARG, "null argument");
        MMCSF_REQUIRE(mmcsf_type_ok(type), MMCSF_ERR_UNSUPPORTED,
                      "format type must be one of 1, 3, 8, 9, 10, 12, 13, 14");
        const Tensor &t = th->t;
        // check_opt, util.h:245-248
        MMCSF_REQUIRE(opt->mode < (uint64_t)t.ndims, MMCSF_ERR_ARG, "Mode cannot be larger than tensor dimension.");
        // The launch-shape flags (-w, -s, -b) do not shape this library's
        // launches (the walk splits work by traversal cost); the limits the
        // reference's launchers enforce on them are kept:
        //   B-CSF / HB-CSF: -w * 32 <= 512 (mttkrp_gpu.cu:1961-1964, 2240-2243, 2660-2663)
        //   MM-CSF / single CSF: -w, -s <= -b / 32 (mttkrp_gpu.cu:3624-3628, 3093-3097)
        if (type == MMCSF_TYPE_BCSF || type == MMCSF_TYPE_BCSF_ANYMODE || type == MMCSF_TYPE_HBCSF)
            MMCSF_REQUIRE(opt->warpPerSlice * 32 <= 512, MMCSF_ERR_ARG,
                          "BLOCKSIZE is smaller than work per slice! Increase BLOCKSIZE.");
        if (type == MMCSF_TYPE_MMCSF || type == MMCSF_TYPE_ONECSF)
            MMCSF_REQUIRE(opt->warpPerSlice <= opt->TBsize / 32 && opt->fiberPerWarp <= opt->TBsize / 32,
                          MMCSF_ERR_ARG,
                          "warpPerFbr (-w) or fbrPerWarp (-s) cannot be higher than threadblock size!\n"
                          "hint: increase -b!");
        hipStream_t s = static_cast<hipStream_t>(stream);
        auto fh = std::make_unique<mmcsf_format>();
        Format &f = fh->f;
        f.type = type;
        f.ndims = t.ndims;
        f.nnz = t.nnz;
        for (int m = 0; m < t.ndims; ++m) f.dims[m] = t.dims[m];
        f.redunMode = redundant_mode(t.ndims, t.dims);
        f.fbrThreshold = opt->fbrThreashold;
        const bool nat = opt->natOrdering != 0;
        int order[kMaxDims];
        DevArray<uint32_t> keys[kMaxDims];
        DevArray<float> vals;
        const uint64_t kNoSplit = 999999999ull;  // util.h:1323
        if (type == MMCSF_TYPE_MMCSF && opt->mixedPart) {
            build_mixed(t, opt, nat, f, s);
        } else if (type == MMCSF_TYPE_MMCSF) {
            // every nonzero lands in partition 2 (util.h:2517), whose mode
            // order is order_tensormode(X, 2) (init_tensor, util.h:372-387)
            order_tensormode(t.ndims, t.dims, 2, nat, order);
            DevArray<uint32_t> pos;
            sort_coo(t, order, keys, vals, s, &pos);
            uint64_t fs = t.ndims == 4 ? 128 : kNoSplit;  // util.h:1325-1326
            build_csf(t.ndims, t.dims, t.nnz, order, keys, std::move(vals), opt->fbrThreashold, fs, f.csf, s);
            f.csf.pos = std::move(pos);
            if (opt->modeTrees) {
                // a root tree for every other mode: each MTTKRP becomes a root
                // traversal.  These trees are internal, so their level order is
                // free.  The leaf level is gathered once per nonzero, the
                // level-1 rows once per fiber segment: a 3-D tree puts at its
                // leaf the mode whose used rows concentrate in a clearly
                // smaller working set (rows holding 90% of the nonzeros; those
                // stay in the L2), and otherwise takes the order with fewer
                // level-1 entries (fewer fiber-level gathers).
                auto build_tree = [&](const int *o2, Csf &c) {
                    DevArray<uint32_t> k2[kMaxDims];
                    DevArray<float> v2;
                    DevArray<uint32_t> p2;
                    sort_coo(t, o2, k2, v2, s, &p2);
                    build_csf(t.ndims, t.dims, t.nnz, o2, k2, std::move(v2), opt->fbrThreashold, kNoSplit, c, s);
                    c.pos = std::move(p2);
                };
                uint64_t hot[kMaxDims] = {0, 0, 0, 0};
                if (t.ndims == 3)
                    for (int m = 0; m < t.ndims; ++m) hot[m] = hot_footprint(t, m, 0.9, s);
                for (int m = 0; m < t.ndims; ++m) {
                    if (m == order[0]) continue;
                    int o2[kMaxDims];
                    order_tensormode(t.ndims, t.dims, m, nat, o2);
                    if (t.ndims == 3) {
                        int o3[kMaxDims] = {o2[0], o2[2], o2[1], 0};
                        if (2 * hot[o3[2]] < hot[o2[2]]) {
                            build_tree(o3, f.tree[m]);
                        } else if (2 * hot[o2[2]] < hot[o3[2]]) {
                            build_tree(o2, f.tree[m]);
                        } else {
                            build_tree(o2, f.tree[m]);
                            Csf alt;
                            build_tree(o3, alt);
                            if (alt.cnt[1] < f.tree[m].cnt[1]) f.tree[m] = std::move(alt);
                        }
                    } else {
                        build_tree(o2, f.tree[m]);
                    }
                    f.hasTree[m] = true;
                }
            } else {
                // privatisation tables of the scatter levels (every level but the root)
                for (int l = 1; l < t.ndims; ++l) build_hot_table(f.csf, l, t.dims[order[l]], f.csf.hot[l], s);
            }
        } else {
            order_tensormode(t.ndims, t.dims, (int)opt->mode, nat, order);
            if (type == MMCSF_TYPE_COO || type == MMCSF_TYPE_COO_CPU) {
                for (int l = 0; l < t.ndims; ++l) f.cooOrder[l] = order[l];
                sort_coo(t, order, f.cooKeys, f.cooVals, s);
            } else {
                const bool bcsf = type != MMCSF_TYPE_HBCSF;  // types 8, 9, 13, 14
                DevArray<uint32_t> pos;
                sort_coo(t, order, keys, vals, s, bcsf ? &pos : nullptr);
                build_csf(t.ndims, t.dims, t.nnz, order, keys, std::move(vals), opt->fbrThreashold, kNoSplit,
                          f.csf, s);
                if (bcsf) {
                    f.csf.pos = std::move(pos);
                    // type 13 is create_HCSR alone (mttkrp.cu:244-247): no slice bins
                    if (type != MMCSF_TYPE_SINGLECSF) build_bins_tiled(f.csf, s);
                    // one-tree all-mode types: the other modes scatter from their level
                    if (type != MMCSF_TYPE_BCSF)
                        for (int l = 1; l < t.ndims; ++l) build_hot_table(f.csf, l, t.dims[order[l]], f.csf.hot[l], s);
                } else {
                    build_hyb(t.ndims, t.dims, std::move(f.csf), f.hyb, s);
                    build_csl_walk(f.hyb, s);
                }
            }
        }
        // the first MTTKRP on this format runs without loading code objects
        if (type != MMCSF_TYPE_COO_CPU) preload_kernels(type == MMCSF_TYPE_MMCSF || type == MMCSF_TYPE_BCSF);
        *out = fh.release();
    });
}

int mmcsf_format_info_get(const mmcsf_format *fh, mmcsf_format_info *info) {
    return guarded([&] {
        MMCSF_REQUIRE(fh && info, MMCSF_ERR_ARG, "null argument");
        const Format &f = fh->f;
        std::memset(info, 0, sizeof(*info));
        info->type = f.type;
        info->ndims = f.ndims;
        info->totNnz = f.nnz;
        info->redunMode = f.redunMode;
        for (int m = 0; m < f.ndims; ++m) info->dims[m] = f.dims[m];
        const Csf *c = nullptr;
        if (f.type == MMCSF_TYPE_COO || f.type == MMCSF_TYPE_COO_CPU) {
            for (int l = 0; l < f.ndims; ++l) info->modeOrder[l] = f.cooOrder[l];
        } else if (f.type == MMCSF_TYPE_HBCSF) {
            c = &f.hyb.csf;
            for (int l = 0; l < f.ndims; ++l) info->modeOrder[l] = f.hyb.order[l];
            info->COOnnz = f.hyb.cooNnz;
            info->CSLnnz = f.hyb.cslNnz;
            info->HCSRnnz = f.hyb.csf.nnz;
        } else if (f.mixed) {
            int big = 0;
            for (int p = 0; p < f.ndims; ++p)
                if (f.partNnz[p] > f.partNnz[big]) big = p;
            c = &f.part[big];
        } else {
            c = &f.csf;
        }
        if (c) {
            for (int l = 0; l < f.ndims; ++l) {
                info->modeOrder[l] = c->order[l];
                info->nLevel[l] = c->cnt[l];
            }
            // TiledTensor::nFibers = fbrPtr[1].size() (util.h:1424); HYB: the
            // leaf-fiber pointer count (util.h:980, 1129)
            info->nFibers = (f.type == MMCSF_TYPE_HBCSF ? c->cnt[f.ndims - 2] : c->cnt[1]) + 1;
            info->droppedSlices = c->dropped;
        }
        for (int m = 0; m < f.ndims; ++m) info->modeTrees += f.hasTree[m] ? 1 : 0;
        if (f.type == MMCSF_TYPE_MMCSF || one_tree_type(f.type))
            for (int m = 0; m < f.ndims; ++m) {
                const Csf &t = f.hasTree[m] ? f.tree[m] : f.csf;
                for (int l = 0; l < f.ndims; ++l) {
                    info->modeLevel[m][l] = t.cnt[l];
                    info->modeTreeOrder[m][l] = t.order[l];
                }
            }
        info->deviceBytes = format_bytes(f);
        for (int p = 0; p < f.ndims; ++p) info->partNnz[p] = f.mixed ? f.partNnz[p] : (p == 2 ? f.nnz : 0);
        info->mixedPart = f.mixed ? 1 : 0;
        if (f.mixed)  // per partition m (an empty partition reports zeros)
            for (int m = 0; m < f.ndims; ++m)
                for (int l = 0; l < f.ndims; ++l) {
                    info->modeLevel[m][l] = f.part[m].cnt[l];
                    info->modeTreeOrder[m][l] = f.partNnz[m] ? f.part[m].order[l] : 0;
                }
    });
}

int mmcsf_format_export(const mmcsf_format *fh, const char *name, int k, void *dst, uint64_t cap, uint64_t *len) {
    return guarded([&] {
        MMCSF_REQUIRE(fh && name && len, MMCSF_ERR_ARG, "null argument");
        const Format &f = fh->f;
        const int N = f.ndims;
        std::string nm(name);
        std::vector<uint64_t> u;
        std::vector<float> fl;
        bool isFloat = false;
        const bool hyb = f.type == MMCSF_TYPE_HBCSF;
        const bool coo = f.type == MMCSF_TYPE_COO || f.type == MMCSF_TYPE_COO_CPU;
        const Csf &c = hyb ? f.hyb.csf : f.csf;
        MMCSF_REQUIRE(!f.mixed || nm == "partPerNnz" || nm == "partOrder" || nm.rfind("P.", 0) == 0, MMCSF_ERR_ARG,
                      "an MM-CSF built with -M 1 exports its partitions (P.<field>, k = 8 * partition + level)");
        auto widen = [&](const std::vector<uint32_t> &v) { u.assign(v.begin(), v.end()); };
        auto need_level = [&](int lvl) {
            MMCSF_REQUIRE(!coo && lvl >= 0 && lvl < N - 1, MMCSF_ERR_ARG, "bad level for " + nm);
        };
        if (nm == "fbrPtr") {
            need_level(k);
            widen(to_host(c.ptr[k], c.cnt[k] + 1));
        } else if (nm == "fbrIdx") {
            need_level(k);
            widen(to_host(c.idx[k], c.cnt[k]));
            // create_TiledHCSR appends a trailing leaf-fiber index
            // (util.h:1418-1419); create_HYB does not
            if (!hyb && k == N - 2 && c.cnt[k]) u.push_back(u.back());
        } else if (nm == "inds") {
            MMCSF_REQUIRE(k >= 0 && k < N, MMCSF_ERR_ARG, "bad mode for inds");
            if (coo) {
                int l = level_of(f.cooOrder, N, k);
                widen(to_host(f.cooKeys[l], f.nnz));
            } else {
                widen(expand_level(c, level_of(c.order, N, k)));
            }
        } else if (nm == "vals") {
            isFloat = true;
            fl = coo ? to_host(f.cooVals, f.nnz) : to_host(c.vals, c.nnz);
        } else if (nm == "fbrLikeSlcInds") {
            MMCSF_REQUIRE(!coo && !hyb, MMCSF_ERR_ARG, "fbrLikeSlcInds exists for B-CSF and MM-CSF");
            widen(to_host(c.fbrSlc, c.cnt[1]));
            // nFibers entries; the last one is never written by the reference
            u.push_back(u.empty() ? 0 : u.back());
        } else if (nm == "slcMapperBin") {
            MMCSF_REQUIRE(c.hasBins && k >= 0 && k < kNumBins, MMCSF_ERR_ARG, "no such bin");
            uint64_t off = 0;
            for (int b = 0; b < k; ++b) off += c.binCount[b];
            std::vector<uint32_t> all = to_host(c.binList, off + c.binCount[k]);
            u.assign(all.begin() + off, all.end());
        } else if (nm == "partPerNnz") {
            MMCSF_REQUIRE(f.type == MMCSF_TYPE_MMCSF, MMCSF_ERR_ARG, "partPerNnz exists for MM-CSF");
            if (f.mixed) widen(f.partPerNnz);
            else u.assign(f.nnz, 2);
        } else if (nm == "partOrder") {
            MMCSF_REQUIRE(f.mixed, MMCSF_ERR_ARG, "partOrder exists for MM-CSF built with -M 1");
            widen(f.partOrder);
        } else if (nm.rfind("P.", 0) == 0) {
            MMCSF_REQUIRE(f.mixed, MMCSF_ERR_ARG, nm + " exists for MM-CSF built with -M 1");
            const int p = k / 8, l = k % 8;
            MMCSF_REQUIRE(k >= 0 && p < N && f.partNnz[p] > 0, MMCSF_ERR_ARG, "no such partition");
            const Csf &pc = f.part[p];
            const std::string fld = nm.substr(2);
            if (fld == "modeOrder") {
                u.assign(pc.order, pc.order + N);
            } else if (fld == "fbrPtr") {
                MMCSF_REQUIRE(l < N - 1, MMCSF_ERR_ARG, "bad level");
                widen(to_host(pc.ptr[l], pc.cnt[l] + 1));
            } else if (fld == "fbrIdx") {
                MMCSF_REQUIRE(l < N - 1, MMCSF_ERR_ARG, "bad level");
                widen(to_host(pc.idx[l], pc.cnt[l]));
                if (l == N - 2 && pc.cnt[l]) u.push_back(u.back());
            } else if (fld == "inds") {
                MMCSF_REQUIRE(l < N, MMCSF_ERR_ARG, "bad mode");
                widen(expand_level(pc, level_of(pc.order, N, l)));
            } else if (fld == "vals") {
                isFloat = true;
                fl = to_host(pc.vals, pc.nnz);
            } else if (fld == "fbrLikeSlcInds") {
                widen(to_host(pc.fbrSlc, pc.cnt[1]));
                u.push_back(u.empty() ? 0 : u.back());
            } else {
                fail(MMCSF_ERR_ARG, "unknown field " + nm);
            }
        } else if (hyb && nm == "COOinds") {
            MMCSF_REQUIRE(k >= 0 && k < N, MMCSF_ERR_ARG, "bad mode");
            widen(to_host(f.hyb.cooInd[level_of(f.hyb.order, N, k)], f.hyb.cooNnz));
        } else if (hyb && nm == "COOvals") {
            isFloat = true;
            fl = to_host(f.hyb.cooVal, f.hyb.cooNnz);
        } else if (hyb && nm == "CSLslicePtr") {
            widen(to_host(f.hyb.cslPtr, f.hyb.nCsl + 1));
        } else if (hyb && nm == "CSLsliceIdx") {
            widen(to_host(f.hyb.cslIdx, f.hyb.nCsl));
        } else if (hyb && nm == "CSLinds") {
            int l = level_of(f.hyb.order, N, k);
            MMCSF_REQUIRE(N == 3 && (l == 1 || l == 2), MMCSF_ERR_ARG, "CSLinds exists for modeOrder[1..2] of 3-D");
            widen(to_host(l == 1 ? f.hyb.cslInd1 : f.hyb.cslInd2, f.hyb.cslNnz));
        } else if (hyb && nm == "CSLvals") {
            isFloat = true;
            fl = to_host(f.hyb.cslVal, f.hyb.cslNnz);
        } else if (hyb && nm == "CSLslcMapperBin") {
            MMCSF_REQUIRE(k >= 0 && k < kNumBins, MMCSF_ERR_ARG, "no such bin");
            uint64_t off = 0;
            for (int b = 0; b < k; ++b) off += f.hyb.cslBinCount[b];
            std::vector<uint32_t> all = to_host(f.hyb.cslBinList, off + f.hyb.cslBinCount[k]);
            u.assign(all.begin() + off, all.end());
        } else {
            fail(MMCSF_ERR_ARG, "unknown field " + nm);
        }
        *len = isFloat ? fl.size() : u.size();
        if (dst && cap >= *len) {
            if (isFloat) std::memcpy(dst, fl.data(), fl.size() * 4);
            else std::memcpy(dst, u.data(), u.size() * 8);
        }
    });
}

void mmcsf_format_free(mmcsf_format *f) { delete f; }

int mmcsf_format_save(const mmcsf_format *fh, const char *path) {
    return guarded([&] {
        MMCSF_REQUIRE(fh && path, MMCSF_ERR_ARG, "null argument");
        format_save(fh->f, path);
    });
}

int mmcsf_format_load(const char *path, mmcsf_format **out) {
    return guarded([&] {
        MMCSF_REQUIRE(path && out, MMCSF_ERR_ARG, "null argument");
        auto fh = std::make_unique<mmcsf_format>();
        format_load(path, fh->f);
        *out = fh.release();
    });
}

int mmcsf_format_is_cache(const char *path) { return path && is_format_cache(path) ? 1 : 0; }

int mmcsf_format_to_tensor(const mmcsf_format *fh, mmcsf_tensor **out) {
    return guarded([&] {
        MMCSF_REQUIRE(fh && out, MMCSF_ERR_ARG, "null argument");
        const Format &f = fh->f;
        std::vector<std::vector<uint32_t>> inds;
        std::vector<float> vals;
        format_nonzeros(f, inds, vals);
        MMCSF_REQUIRE(vals.size() == f.nnz, MMCSF_ERR_ARG, "format does not hold every nonzero");
        auto th = std::make_unique<mmcsf_tensor>();
        Tensor &t = th->t;
        t.ndims = f.ndims;
        t.nnz = f.nnz;
        for (int m = 0; m < f.ndims; ++m) {
            t.dims[m] = f.dims[m];
            t.inds[m].alloc(f.nnz);
            if (f.nnz) MMCSF_HIP(hipMemcpy(t.inds[m].p, inds[m].data(), f.nnz * 4, hipMemcpyHostToDevice));
        }
        t.vals.alloc(f.nnz);
        if (f.nnz) MMCSF_HIP(hipMemcpy(t.vals.p, vals.data(), f.nnz * 4, hipMemcpyHostToDevice));
        *out = th.release();
    });
}

int mmcsf_mttkrp(const mmcsf_format *fh, int mode, int R, int dtype, const void *const *factors, float *out,
                 void *stream) {
    return guarded([&] {
        MMCSF_REQUIRE(fh && factors && out, MMCSF_ERR_ARG, "null argument");
        MMCSF_REQUIRE(dtype == MMCSF_F32 || dtype == MMCSF_BF16, MMCSF_ERR_ARG, "dtype must be F32 or BF16");
        const Format &f = fh->f;
        MMCSF_REQUIRE(mode >= 0 && mode < f.ndims, MMCSF_ERR_ARG, "Mode cannot be larger than tensor dimension.");
        hipStream_t s = static_cast<hipStream_t>(stream);
        Factors F{};
        for (int m = 0; m < f.ndims; ++m) F.U[m] = factors[m];
        F.dtype = dtype;
        F.R = R;
        switch (f.type) {
            case MMCSF_TYPE_COO:
            case MMCSF_TYPE_COO_CPU:
                MMCSF_REQUIRE(mode == f.cooOrder[0], MMCSF_ERR_ARG, "COO format serves its root mode only");
                mttkrp_coo(f, F, out, s);
                break;
            case MMCSF_TYPE_BCSF:
                MMCSF_REQUIRE(mode == f.csf.order[0], MMCSF_ERR_ARG, "B-CSF serves its root mode only");
                mttkrp_csf(f.csf, f.dims, 0, F, out, s);
                break;
            case MMCSF_TYPE_HBCSF:
                MMCSF_REQUIRE(mode == f.hyb.order[0], MMCSF_ERR_ARG, "HB-CSF serves its root mode only");
                mttkrp_hyb(f.hyb, f.dims, F, out, s);
                break;
            default:
                mttkrp_mode(f, mode, F, out, s);
                break;
        }
    });
}

int mmcsf_mttkrp_allmode(const mmcsf_format *fh, int R, int dtype, const void *const *factors, float *const *outs,
                         void *stream, float *ms_per_mode) {
    return guarded([&] {
        MMCSF_REQUIRE(fh && factors && outs, MMCSF_ERR_ARG, "null argument");
        const Format &f = fh->f;
        MMCSF_REQUIRE(f.type == MMCSF_TYPE_MMCSF || one_tree_type(f.type), MMCSF_ERR_ARG,
                      "all-mode MTTKRP needs an MM-CSF or one-tree (-t 9 / 13 / 14) format");
        MMCSF_REQUIRE(dtype == MMCSF_F32 || dtype == MMCSF_BF16, MMCSF_ERR_ARG, "dtype must be F32 or BF16");
        hipStream_t s = static_cast<hipStream_t>(stream);
        // one event pair per mode, around the MTTKRP kernel only
        hipEvent_t ev[2 * kMaxDims] = {};
        struct Events {
            hipEvent_t *e;
            int n;
            ~Events() {
                for (int i = 0; i < n; ++i)
                    if (e[i]) (void)hipEventDestroy(e[i]);
            }
        } guard{ev, 2 * kMaxDims};
        if (ms_per_mode)
            for (int i = 0; i < 2 * f.ndims; ++i) MMCSF_HIP(hipEventCreate(&ev[i]));
        Factors F{};
        for (int m = 0; m < f.ndims; ++m) F.U[m] = factors[m];
        F.dtype = dtype;
        F.R = R;
        for (int m = 0; m < f.ndims; ++m) {
            MMCSF_REQUIRE(outs[m] != nullptr, MMCSF_ERR_ARG, "missing output");
            mttkrp_mode(f, m, F, outs[m], s, ms_per_mode ? ev[2 * m] : nullptr,
                        ms_per_mode ? ev[2 * m + 1] : nullptr);
        }
        if (ms_per_mode) {
            MMCSF_HIP(hipEventSynchronize(ev[2 * f.ndims - 1]));
            for (int m = 0; m < f.ndims; ++m)
                MMCSF_HIP(hipEventElapsedTime(&ms_per_mode[m], ev[2 * m], ev[2 * m + 1]));
        }
    });
}

int mmcsf_mttkrp_coo_host(const mmcsf_format *fh, int mode, int R, const float *const *U, float *out) {
    return guarded([&] {
        MMCSF_REQUIRE(fh && U && out && R >= 1, MMCSF_ERR_ARG, "bad argument");
        const Format &f = fh->f;
        MMCSF_REQUIRE(f.type == MMCSF_TYPE_COO_CPU || f.type == MMCSF_TYPE_COO, MMCSF_ERR_ARG,
                      "the host COO MTTKRP needs a COO format");
        MMCSF_REQUIRE(mode >= 0 && mode < f.ndims, MMCSF_ERR_ARG, "Mode cannot be larger than tensor dimension.");
        const int N = f.ndims;
        std::vector<std::vector<uint32_t>> ind(N);
        for (int m = 0; m < N; ++m) ind[m] = to_host(f.cooKeys[level_of(f.cooOrder, N, m)], f.nnz);
        std::vector<float> v = to_host(f.cooVals, f.nnz);
        // MTTKRP_COO_CPU[_4D] (mttkrp_cpu.cpp:60-113): cyclic roles, stored
        // order, sequential float accumulation into out
        int cm[kMaxDims];
        for (int m = 0; m < N; ++m) cm[m] = (m + mode) % N;
        const size_t r = (size_t)R;
        for (uint64_t x = 0; x < f.nnz; ++x) {
            float *o = out + ind[cm[0]][x] * r;
            const float *a = U[cm[1]] + ind[cm[1]][x] * r;
            const float *b = U[cm[2]] + ind[cm[2]][x] * r;
            if (N == 3) {
                for (size_t q = 0; q < r; ++q) {
                    float tmp = v[x] * a[q] * b[q];
                    o[q] += tmp;
                }
            } else {
                const float *c = U[cm[3]] + ind[cm[3]][x] * r;
                for (size_t q = 0; q < r; ++q) {
                    float tmp = v[x] * a[q] * b[q] * c[q];
                    o[q] += tmp;
                }
            }
        }
    });
}

namespace {
void cpd_checks(const mmcsf_format *fh, int R, int max_iters, int dtype, float *const *factors, mmcsf_cpd_stats *stats) {
    MMCSF_REQUIRE(fh && factors, MMCSF_ERR_ARG, "null argument");
    const Format &f = fh->f;
    MMCSF_REQUIRE(f.type == MMCSF_TYPE_MMCSF, MMCSF_ERR_ARG, "CP-ALS runs on an MM-CSF format");
    MMCSF_REQUIRE(R >= 1 && R <= 256, MMCSF_ERR_UNSUPPORTED, "R must be in 1..256");
    MMCSF_REQUIRE(dtype == MMCSF_F32 || dtype == MMCSF_BF16, MMCSF_ERR_ARG, "dtype must be F32 or BF16");
    MMCSF_REQUIRE(f.nnz > 0, MMCSF_ERR_ARG, "CP-ALS needs at least one nonzero");
    MMCSF_REQUIRE(max_iters >= 1, MMCSF_ERR_ARG, "max_iters must be positive");
    for (int m = 0; m < f.ndims; ++m) MMCSF_REQUIRE(factors[m] != nullptr, MMCSF_ERR_ARG, "missing factor");
    if (stats) std::memset(stats, 0, sizeof(*stats));
}
}  // namespace

int mmcsf_cpd_als(const mmcsf_format *fh, int R, int max_iters, double tol, int dtype, float *const *factors,
                  double *lambda, mmcsf_cpd_stats *stats, void *stream) {
    return guarded([&] {
        cpd_checks(fh, R, max_iters, dtype, factors, stats);
        cpd_als(fh->f, R, max_iters, tol, dtype, factors, lambda, stats, static_cast<hipStream_t>(stream));
    });
}

int mmcsf_cpd_als_noalloc(const mmcsf_format *fh, int R, int max_iters, double tol, int dtype,
                          float *const *factors, double *lambda, mmcsf_cpd_stats *stats, void *stream) {
    return guarded([&] {
        static_assert(MMCSF_CPD_NOALLOC_SCRATCH_BYTES == (64ull << 20), "scratch size stated in mmcsf.h");
        cpd_checks(fh, R, max_iters, dtype, factors, stats);
        cpd_als_noalloc(fh->f, R, max_iters, tol, dtype, factors, lambda, stats, static_cast<hipStream_t>(stream));
    });
}

int mmcsf_gram(const float *U, uint64_t rows, int R, double *G, void *stream) {
    return guarded([&] {
        MMCSF_REQUIRE(U && G, MMCSF_ERR_ARG, "null argument");
        MMCSF_REQUIRE(R >= 1 && R <= 256 && rows <= kMaxIndex, MMCSF_ERR_UNSUPPORTED, "unsupported shape");
        gram(U, rows, R, G, static_cast<hipStream_t>(stream));
    });
}

// ---- TTM ----------------------------------------------------------------------

namespace {
const Csf *ttm_tree(const Format &f) {
    MMCSF_REQUIRE(bcsf_family(f.type), MMCSF_ERR_ARG, "TTM needs a B-CSF format (-t 8, 9, 13 or 14)");
    return &f.csf;
}
}  // namespace

int mmcsf_ttm(const mmcsf_format *fh, int R, int dtype, const void *factor, float *Y, void *stream) {
    return guarded([&] {
        MMCSF_REQUIRE(fh && factor && Y, MMCSF_ERR_ARG, "null argument");
        MMCSF_REQUIRE(dtype == MMCSF_F32 || dtype == MMCSF_BF16, MMCSF_ERR_ARG, "dtype must be F32 or BF16");
        MMCSF_REQUIRE(R >= 1 && R <= 256 && R % 4 == 0, MMCSF_ERR_UNSUPPORTED, "TTM needs R % 4 == 0 and R <= 256");
        const Format &f = fh->f;
        const Csf &c = *ttm_tree(f);
        const size_t need = dtype == MMCSF_F32 || R % 8 == 0 ? 16 : 8;
        MMCSF_REQUIRE(reinterpret_cast<uintptr_t>(factor) % need == 0 && reinterpret_cast<uintptr_t>(Y) % 16 == 0,
                      MMCSF_ERR_ARG, "factor / output pointer is not aligned");
        hipStream_t s = static_cast<hipStream_t>(stream);
        const uint64_t rows = c.cnt[c.ndims - 2];
        fill_zero(Y, rows * (uint64_t)R, s);
        Factors F{};
        for (int m = 0; m < f.ndims; ++m) F.U[m] = factor;  // only the leaf factor is read
        F.dtype = dtype;
        F.R = R;
        ttm_fibers(c, F, Y, s);
    });
}

int mmcsf_ttm_host(const mmcsf_format *fh, int R, const float *U, float *Y) {
    return guarded([&] {
        MMCSF_REQUIRE(fh && U && Y && R >= 1, MMCSF_ERR_ARG, "bad argument");
        const Format &f = fh->f;
        const Csf &c = *ttm_tree(f);
        const int N = c.ndims;
        const uint64_t nF = c.cnt[N - 2];
        std::vector<uint32_t> ptr = to_host(c.ptr[N - 2], nF + 1);
        std::vector<uint32_t> leaf = to_host(c.leaf, c.nnz);
        std::vector<float> vals = to_host(c.vals, c.nnz);
        std::vector<float> tmp(R);
        // one fiber at a time, summed in float, then added (ttm_cpu.cpp:68-84)
        for (uint64_t fb = 0; fb < nF; ++fb) {
            std::fill(tmp.begin(), tmp.end(), 0.f);
            for (uint32_t x = ptr[fb]; x < ptr[fb + 1]; ++x) {
                const float *u = U + (size_t)leaf[x] * R;
                for (int r = 0; r < R; ++r) tmp[r] += vals[x] * u[r];
            }
            for (int r = 0; r < R; ++r) Y[fb * R + r] = tmp[r];
        }
    });
}

// ---- CP completion -----------------------------------------------------------

namespace {

// the tree whose root level is `mode` (with tensor-order positions), or null
const Csf *root_tree(const Format &f, int mode) {
    if (f.mixed) return nullptr;  // partitions keep no tensor-order positions
    if (f.type == MMCSF_TYPE_MMCSF) {
        if (f.csf.order[0] == mode) return &f.csf;
        if (f.hasTree[mode]) return &f.tree[mode];
        return nullptr;
    }
    if (bcsf_family(f.type) && f.csf.order[0] == mode) return &f.csf;
    return nullptr;
}

Factors cp_factors(const Format &f, int R, int dtype, const void *const *factors) {
    MMCSF_REQUIRE(factors, MMCSF_ERR_ARG, "null argument");
    MMCSF_REQUIRE(!f.mixed, MMCSF_ERR_UNSUPPORTED,
                  "CP completion needs a B-CSF or an MM-CSF without the mixed-mode partition (-M 0)");
    MMCSF_REQUIRE(dtype == MMCSF_F32 || dtype == MMCSF_BF16, MMCSF_ERR_ARG, "dtype must be F32 or BF16");
    MMCSF_REQUIRE(cp_supports(R), MMCSF_ERR_UNSUPPORTED, "CP completion needs R % 4 == 0 and R <= 128");
    const size_t need = dtype == MMCSF_F32 || R % 8 == 0 ? 16 : 8;
    Factors F{};
    for (int m = 0; m < f.ndims; ++m) {
        MMCSF_REQUIRE(factors[m], MMCSF_ERR_ARG, "missing factor matrix");
        MMCSF_REQUIRE(reinterpret_cast<uintptr_t>(factors[m]) % need == 0, MMCSF_ERR_ARG,
                      "factor pointer is not aligned to " + std::to_string(need) + " bytes");
        F.U[m] = factors[m];
    }
    F.dtype = dtype;
    F.R = R;
    return F;
}

}  // namespace

int mmcsf_cp_forward(const mmcsf_format *fh, int R, int dtype, const void *const *factors, float *residual,
                     double *loss, void *stream) {
    return guarded([&] {
        MMCSF_REQUIRE(fh && loss, MMCSF_ERR_ARG, "null argument");
        const Format &f = fh->f;
        MMCSF_REQUIRE(f.type == MMCSF_TYPE_MMCSF || bcsf_family(f.type), MMCSF_ERR_ARG,
                      "CP completion needs a B-CSF or MM-CSF format");
        const Factors F = cp_factors(f, R, dtype, factors);
        MMCSF_REQUIRE(f.csf.pos.p || f.nnz == 0, MMCSF_ERR_ARG, "format holds no nonzero positions");
        cp_forward(f.csf, F, residual, loss, static_cast<hipStream_t>(stream));
    });
}

int mmcsf_cp_adjoint(const mmcsf_format *fh, int mode, int R, int dtype, const void *const *factors,
                     const float *weights, float alpha, float *grad, void *stream) {
    return guarded([&] {
        MMCSF_REQUIRE(fh && grad, MMCSF_ERR_ARG, "null argument");
        const Format &f = fh->f;
        MMCSF_REQUIRE(mode >= 0 && mode < f.ndims, MMCSF_ERR_ARG, "Mode cannot be larger than tensor dimension.");
        const Factors F = cp_factors(f, R, dtype, factors);
        const Csf *c = root_tree(f, mode);
        MMCSF_REQUIRE(c, MMCSF_ERR_ARG, "the format holds no tree rooted at this mode");
        MMCSF_REQUIRE(c->pos.p || f.nnz == 0, MMCSF_ERR_ARG, "format holds no nonzero positions");
        MMCSF_REQUIRE(reinterpret_cast<uintptr_t>(grad) % 16 == 0, MMCSF_ERR_ARG,
                      "gradient pointer is not aligned to 16 bytes");
        hipStream_t s = static_cast<hipStream_t>(stream);
        fill_zero(grad, f.dims[mode] * (uint64_t)R, s);
        cp_gradient(*c, F, weights, alpha, grad, nullptr, s);
    });
}

int mmcsf_cp_loss_grad(const mmcsf_format *fh, int R, int dtype, const void *const *factors, float alpha,
                       float *const *grads, double *loss, void *stream) {
    return guarded([&] {
        MMCSF_REQUIRE(fh && grads && loss, MMCSF_ERR_ARG, "null argument");
        const Format &f = fh->f;
        const Factors F = cp_factors(f, R, dtype, factors);
        for (int m = 0; m < f.ndims; ++m) {
            MMCSF_REQUIRE(!grads[m] || root_tree(f, m), MMCSF_ERR_ARG,
                          "all-mode gradients need an MM-CSF format with a root tree per mode");
            MMCSF_REQUIRE(reinterpret_cast<uintptr_t>(grads[m]) % 16 == 0, MMCSF_ERR_ARG,
                          "gradient pointer is not aligned to 16 bytes");
        }
        MMCSF_REQUIRE(f.type == MMCSF_TYPE_MMCSF || bcsf_family(f.type), MMCSF_ERR_ARG,
                      "CP completion needs a B-CSF or MM-CSF format");
        MMCSF_REQUIRE(f.csf.pos.p || f.nnz == 0, MMCSF_ERR_ARG, "format holds no nonzero positions");
        hipStream_t s = static_cast<hipStream_t>(stream);
        cp_zero(loss, s);
        bool counted = false;
        for (int m = 0; m < f.ndims; ++m) {
            if (!grads[m]) continue;
            fill_zero(grads[m], f.dims[m] * (uint64_t)R, s);
            cp_gradient(*root_tree(f, m), F, nullptr, alpha, grads[m], counted ? nullptr : loss, s);
            counted = true;
        }
        if (counted) cp_half(loss, s);
        else cp_forward(f.csf, F, nullptr, loss, s);
    });
}

}  // extern "C"
