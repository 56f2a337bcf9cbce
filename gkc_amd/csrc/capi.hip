// C ABI and host-side driver of the k-mer counting path (gkc_hip.h).
//
// A context owns a HIP stream, the device partition bins (2-bit packed
// superkmers + headers per partition, the device counterpart of the
// reference's SKMStore, superkmers.hpp:138-187), the key buffers of the
// radix sort and the (k-mer, count) table.
//
// Two forms of the same kernels:
//   plain    count-then-allocate with host-synchronised sizes: phase 1 runs a
//            count pass, then the scatter; phase 2 counts one partition at a
//            time (extract -> passes -> run-length count -> sync -> emit).
//   batched  (GKC_BATCHED) every partition in one launch per stage from a
//            device work queue (persistent workgroups claim tiles from
//            counters), the launch sequence captured once as a hipGraph and
//            replayed with sizes read from device memory; with GKC_PRESIZED the
//            table is sized from predicted distinct counts and the count pass
//            is skipped, an on-device overflow flag triggering the plain
//            count-then-allocate emission as fallback.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>
#include <cmath>
#include <cerrno>
#include <fcntl.h>
#include <unistd.h>

#include "../../include/gkc_hip.h"
#include "kernels.hpp"

using namespace gkc;

static thread_local std::string g_err;

struct HipError {
    hipError_t e; std::string where;
};
#define HIPX(call)                                                                             \
    do {                                                                                       \
        hipError_t _e = (call);                                                                \
        if (_e != hipSuccess) throw HipError{_e, std::string(#call) + " @" + std::to_string(__LINE__)}; \
    } while (0)
struct ArgError { std::string msg; int code; };

namespace gkc {
// mlp.hip: learned table capacities on the device (queued on st)
int gkc_mlp_presize_device(gkc_mlp* c, const uint32_t* sketch, const uint32_t* total, uint32_t P, const float* norm,
                           float* x, float* y, uint32_t* caps, uint64_t* cap_sum, hipStream_t st);

// ingest.hip: text -> read CSR on the device (0, or a GKC_E* code with err set)
int parse_reads(const char* text, uint64_t n, int format, char* bases, uint64_t* offs, hipStream_t st,
                int64_t* n_reads, std::string& err);
int parse_reads_noalloc(const char* text, uint64_t n, int format, char* bases, uint64_t* offs, hipStream_t st,
                int64_t* n_reads, std::string& err);
}

namespace {

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    ~DevBuf() { if (p) (void)hipFree(p); }
    // grow (contents not kept)
    void ensure(size_t n) {
        if (n <= bytes) return;
        if (p) HIPX(hipFree(p));
        p = nullptr; bytes = 0;
        size_t nb = std::max<size_t>(n, 256);
        hipError_t e = hipMalloc(&p, nb);
        if (e != hipSuccess) { (void)hipGetLastError(); throw ArgError{"device allocation of " + std::to_string(nb) + " bytes failed", GKC_ENOMEM}; }
        bytes = nb;
    }
    template <typename T> T* as() const { return reinterpret_cast<T*>(p); }
};

inline uint32_t cdiv(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

}  // namespace

struct gkc_ctx {
    std::mutex mu;
    int device = 0;
    gkc_params prm{};
    int keyw = 8;                 // key bytes
    int npass = 8;
    int n_cu = 256;
    hipStream_t st = nullptr;
    hipEvent_t ev[8]{};
    hipEvent_t evp[2 * MAX_PASS]{};   // per-pass timing

    // bins
    DevBuf words, hdrs;
    std::vector<uint64_t> h_word_base, h_hdr_base;
    std::vector<uint32_t> h_word_cap, h_hdr_cap, h_word_cur, h_hdr_cur, h_kmer_cur;
    DevBuf d_word_base, d_hdr_base, d_word_cap, d_hdr_cap, d_word_cur, d_hdr_cur, d_kmer_cur, d_overflow;
    DevBuf d_hk_cur;                             // [P] headers | k-mers << 32 (the kernels' cursors; split
                                                 // into d_hdr_cur / d_kmer_cur after each scatter)
    DevBuf d_tmp_hk, d_tmp_word;                 // per-batch counts of the count pass
    DevBuf d_sketch;

    // batch staging
    DevBuf d_reads, d_offs, d_lens, d_hpc_reads, d_hpc_pos, d_chunk, d_ctr, d_mm;
    DevBuf d_win_mm, d_win_pos, d_win_str;     // per-window records of the last partition call
    DevBuf d_backup;                           // cursors before a presized scatter (3 x P)
    uint64_t bases_done = 0;                   // bases partitioned since the last reset
    uint64_t reads_done = 0;                   // reads partitioned since the last reset
    int pos_mode = 0;                          // since the last reset: 0 no batch, 1 with positions, 2 without
    DevBuf skmpos;                             // GKC_WITH_POS: {read, offset} beside each superkmer header
    uint32_t p1_fallbacks = 0;                 // presized scatters redone since the last reset
    gkc_tuning tune{};                         // kernel variants picked by the autotuner
    DevBuf d_wl;                               // work lists of the phase-2 kernels
    std::vector<uint32_t> h_wl;
    WorkLists wl{};
    hipGraphExec_t g_p1 = nullptr;             // captured presized scatter of one chunk
    std::vector<const void*> g_p1_key;
    uint64_t win_n = 0;
    uint32_t win_lead = 0;

    // counting
    DevBuf keysA, keysB, lookback, d_plan, d_segk, d_segh, d_keystart, d_itemstart, d_tilestart, d_rlstart, d_hist;
    DevBuf d_tdesc;                            // sort tile descriptors in work-list order (k_tile_desc)
    DevBuf d_iseg;                             // segment of each extraction item (k_tile_desc)
    DevBuf d_tstart, d_tcap, d_nent, d_ndist, d_toverflow;
    DevBuf d_tcap_total;              // total table capacity of a presized binned count (u64)
    uint64_t h_tcap_total = 0;
    DevBuf d_wdesc;                            // binned-count windows {segment, n, first key low, high}
    DevBuf d_tiler0;                           // read holding the first position of each phase-1 tile
    DevBuf d_chunk_s;                          // descriptor of a chunk's sample (first-batch presizing)
    DevBuf d_hlist, d_hres;                    // heavy tiles of a wave-level count and their {kept, distinct}
    DevBuf d_wkept;                            // kept entries per window, then window offsets
    DevBuf tkeys, tcnts;
    std::vector<uint64_t> part_off, part_distinct, part_kmers, part_start;   // part_start: physical table offsets
    std::vector<uint32_t> h_caps, h_nent;
    std::vector<uint64_t> h_tstart;
    // captured launch sequence of the batched form (re-captured when buffers move)
    hipGraphExec_t g_sort = nullptr, g_emit = nullptr, g_cnt = nullptr;
    const void* g_emit_tk = nullptr;           // table buffers g_emit was captured with
    const void* g_emit_tc = nullptr;
    hipGraphExec_t g_emit_p = nullptr;         // emit pass that also prepares (no count pass before it)
    const void* g_emit_p_tk = nullptr;
    const void* g_emit_p_tc = nullptr;
    std::vector<const void*> g_key;
    std::vector<uint32_t> presize_caps;
    // learned presizing on the device (gkc_set_predictor): the MLP, its
    // normalisation + margin, feature / output buffers, and the capacities'
    // sum read back into pinned memory while the sort runs
    gkc_mlp* mlp = nullptr;
    DevBuf d_pnorm, d_px, d_py, d_lcaps;
    uint64_t* h_lsum = nullptr;
    hipEvent_t ev_lsum = nullptr;
    // the batched partition's cursors and overflow flag packed on the device,
    // read back by one copy into pinned memory per chunk
    DevBuf d_curpack;
    uint32_t* h_curpack = nullptr;
    // a device-input batch's first and end read offsets, gathered on the
    // device and read back by one copy into pinned memory
    DevBuf d_oend;
    uint64_t* h_oend = nullptr;
    bool counted = false;
    gkc_stats stats{};
};

static int fail(gkc_ctx* c, int code, const std::string& m) {
    g_err = m;
    (void)c;
    return code;
}

template <typename F>
static int guarded(gkc_ctx* c, F&& f) {
    if (!c) { g_err = "null context"; return GKC_EINVAL; }
    std::lock_guard<std::mutex> lk(c->mu);
    try {
        HIPX(hipSetDevice(c->device));
        return f();
    } catch (const HipError& e) {
        (void)hipGetLastError();
        return fail(c, GKC_EHIP, std::string("HIP error ") + hipGetErrorString(e.e) + " in " + e.where);
    } catch (const ArgError& e) {
        return fail(c, e.code, e.msg);
    }
}

#include <map>
static std::mutex g_occ_mu;
static std::map<std::pair<const void*, size_t>, int> g_occ;
// Resident workgroups per CU for a kernel (cached, so no runtime query runs
// while a stream is being captured) times the CU count.
static int occupancy_blocks(const void* fn, int nt, int n_cu, size_t dyn_lds = 0) {
    std::lock_guard<std::mutex> lk(g_occ_mu);
    auto it = g_occ.find({fn, dyn_lds});
    int b = 0;
    if (it != g_occ.end()) b = it->second;
    else {
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, fn, nt, dyn_lds) != hipSuccess || b < 1) { (void)hipGetLastError(); b = 1; }
        g_occ[{fn, dyn_lds}] = b;
    }
    return b * n_cu;
}

static void zero(gkc_ctx* c, void* p, uint64_t n32);
template <typename F> hipGraphExec_t capture(gkc_ctx* c, F&& enqueue);
template <typename K> static int sort_items(gkc_ctx* c, int db);
template <typename K> static int sort_nt(gkc_ctx* c, int db);

// ---------------------------------------------------------------- bins

static void bins_init(gkc_ctx* c) {
    const uint32_t P = c->prm.n_parts;
    c->h_word_base.assign(P, 0); c->h_hdr_base.assign(P, 0);
    c->h_word_cap.assign(P, 0); c->h_hdr_cap.assign(P, 0);
    c->h_word_cur.assign(P, 0); c->h_hdr_cur.assign(P, 0); c->h_kmer_cur.assign(P, 0);
    c->d_word_base.ensure(P * 8); c->d_hdr_base.ensure(P * 8);
    c->d_word_cap.ensure(P * 4); c->d_hdr_cap.ensure(P * 4);
    c->d_word_cur.ensure(P * 4); c->d_hdr_cur.ensure(P * 4); c->d_kmer_cur.ensure(P * 4);
    c->d_hk_cur.ensure(P * 8); c->d_tmp_hk.ensure(P * 8); c->d_tmp_word.ensure(P * 4);
    c->d_overflow.ensure(16);
    c->d_sketch.ensure((size_t)P * 256 * 4);
    HIPX(hipMemsetAsync(c->d_word_cur.p, 0, P * 4, c->st));
    HIPX(hipMemsetAsync(c->d_hdr_cur.p, 0, P * 4, c->st));
    HIPX(hipMemsetAsync(c->d_kmer_cur.p, 0, P * 4, c->st));
    HIPX(hipMemsetAsync(c->d_hk_cur.p, 0, P * 8, c->st));
    HIPX(hipMemsetAsync(c->d_word_cap.p, 0, P * 4, c->st));
    HIPX(hipMemsetAsync(c->d_hdr_cap.p, 0, P * 4, c->st));
    HIPX(hipMemsetAsync(c->d_sketch.p, 0, (size_t)P * 256 * 4, c->st));
    HIPX(hipMemsetAsync(c->d_overflow.p, 0, 16, c->st));
    c->counted = false;
}

// Make room for `need_w` / `need_h` more words / headers per partition
// (cursor-relative).  Regrows all regions with slack and copies contents.
static void bins_reserve(gkc_ctx* c, const std::vector<uint32_t>& need_w, const std::vector<uint32_t>& need_h) {
    const uint32_t P = c->prm.n_parts;
    bool grow = false;
    for (uint32_t q = 0; q < P; q++)
        if ((uint64_t)c->h_word_cur[q] + need_w[q] > c->h_word_cap[q] || (uint64_t)c->h_hdr_cur[q] + need_h[q] > c->h_hdr_cap[q]) grow = true;
    if (!grow) return;
    std::vector<uint64_t> nwb(P), nhb(P);
    std::vector<uint32_t> nwc(P), nhc(P);
    uint64_t tw = 0, th = 0;
    for (uint32_t q = 0; q < P; q++) {
        uint64_t ww = (uint64_t)c->h_word_cur[q] + need_w[q], hh = (uint64_t)c->h_hdr_cur[q] + need_h[q];
        if (ww > c->h_word_cap[q]) ww = ww + ww / 4 + 64; else ww = c->h_word_cap[q];
        if (hh > c->h_hdr_cap[q]) hh = hh + hh / 4 + 16; else hh = c->h_hdr_cap[q];
        if (ww >= (1ull << 32) || hh >= (1ull << 32)) throw ArgError{"partition bin exceeds 2^32 words", GKC_EINVAL};
        nwb[q] = tw; nhb[q] = th; nwc[q] = (uint32_t)ww; nhc[q] = (uint32_t)hh;
        tw += ww; th += hh;
    }
    DevBuf nw, nh;
    nw.ensure((tw + 8) * 4);      // +8 words: extraction reads up to 7 words past a superkmer's last one
    nh.ensure(th * sizeof(SkmHdr) + 16);
    for (uint32_t q = 0; q < P; q++) {
        if (c->h_word_cur[q]) HIPX(hipMemcpyAsync(nw.as<uint32_t>() + nwb[q], c->words.as<uint32_t>() + c->h_word_base[q],
                                                  (size_t)c->h_word_cur[q] * 4, hipMemcpyDeviceToDevice, c->st));
        if (c->h_hdr_cur[q]) HIPX(hipMemcpyAsync(nh.as<SkmHdr>() + nhb[q], c->hdrs.as<SkmHdr>() + c->h_hdr_base[q],
                                                 (size_t)c->h_hdr_cur[q] * sizeof(SkmHdr), hipMemcpyDeviceToDevice, c->st));
    }
    DevBuf np;
    if (c->pos_mode == 1) {         // superkmer positions follow the header layout
        np.ensure(th * sizeof(uint2) + 16);
        for (uint32_t q = 0; q < P; q++)
            if (c->h_hdr_cur[q]) HIPX(hipMemcpyAsync(np.as<uint2>() + nhb[q], c->skmpos.as<uint2>() + c->h_hdr_base[q],
                                                     (size_t)c->h_hdr_cur[q] * sizeof(uint2), hipMemcpyDeviceToDevice, c->st));
    }
    HIPX(hipMemsetAsync(static_cast<char*>(nw.p) + tw * 4, 0, 32, c->st));
    HIPX(hipStreamSynchronize(c->st));
    std::swap(c->words.p, nw.p); std::swap(c->words.bytes, nw.bytes);
    std::swap(c->hdrs.p, nh.p); std::swap(c->hdrs.bytes, nh.bytes);
    if (c->pos_mode == 1) { std::swap(c->skmpos.p, np.p); std::swap(c->skmpos.bytes, np.bytes); }
    c->h_word_base = nwb; c->h_hdr_base = nhb; c->h_word_cap = nwc; c->h_hdr_cap = nhc;
    HIPX(hipMemcpyAsync(c->d_word_base.p, nwb.data(), P * 8, hipMemcpyHostToDevice, c->st));
    HIPX(hipMemcpyAsync(c->d_hdr_base.p, nhb.data(), P * 8, hipMemcpyHostToDevice, c->st));
    HIPX(hipMemcpyAsync(c->d_word_cap.p, nwc.data(), P * 4, hipMemcpyHostToDevice, c->st));
    HIPX(hipMemcpyAsync(c->d_hdr_cap.p, nhc.data(), P * 4, hipMemcpyHostToDevice, c->st));
    HIPX(hipStreamSynchronize(c->st));
}

static BinSet binset(gkc_ctx* c, bool counting) {
    BinSet b;
    b.words = c->words.as<uint32_t>(); b.hdrs = c->hdrs.as<SkmHdr>();
    b.word_base = c->d_word_base.as<uint64_t>(); b.hdr_base = c->d_hdr_base.as<uint64_t>();
    b.word_cap = c->d_word_cap.as<uint32_t>(); b.hdr_cap = c->d_hdr_cap.as<uint32_t>();
    if (counting) {
        b.word_cur = c->d_tmp_word.as<uint32_t>(); b.hk_cur = c->d_tmp_hk.as<unsigned long long>();
    } else {
        b.word_cur = c->d_word_cur.as<uint32_t>(); b.hk_cur = c->d_hk_cur.as<unsigned long long>();
    }
    b.overflow = c->d_overflow.as<uint32_t>();
    b.pos = c->pos_mode == 1 ? c->skmpos.as<uint2>() : nullptr;
    return b;
}

// Stage a batch on the device; returns the number of bases. Fills d_chunk.
// With HPC, hpc_pos also keeps each run's in-read origin (d_hpc_pos: only the
// reference-layout arrays of gkc_superkmer_arrays hand it out).
__global__ void k_gather_oend(const uint64_t* __restrict__ offs, uint64_t n, uint64_t* __restrict__ out) {
    out[0] = offs[0];
    out[1] = offs[n];
}

static uint64_t stage_batch(gkc_ctx* c, const char* reads, const uint64_t* offs, int64_t n, uint32_t flags,
                            ChunkDesc& cd, bool hpc_pos = false) {
    uint64_t o0, on;
    const uint8_t* dreads;
    const uint64_t* doffs;
    uint32_t lead = 0;
    if (flags & GKC_DEVICE_INPUT) {
        // on the context's own (non-blocking) stream: a null-stream copy would
        // also wait for the caller's work on other streams (its next upload)
        c->d_oend.ensure(16);
        if (!c->h_oend) HIPX(hipHostMalloc((void**)&c->h_oend, 16, hipHostMallocDefault));
        hipLaunchKernelGGL(k_gather_oend, dim3(1), dim3(1), 0, c->st, offs, (uint64_t)n, c->d_oend.as<uint64_t>());
        HIPX(hipGetLastError());
        HIPX(hipMemcpyAsync(c->h_oend, c->d_oend.p, 16, hipMemcpyDeviceToHost, c->st));
        HIPX(hipStreamSynchronize(c->st));
        o0 = c->h_oend[0];
        on = c->h_oend[1];
        // align the base pointer down to 16 bytes; positions [0, lead) belong to no read
        lead = (uint32_t)(reinterpret_cast<uintptr_t>(reads) & 15);
        dreads = reinterpret_cast<const uint8_t*>(reads) - lead;
        doffs = offs;
    } else {
        o0 = offs[0]; on = offs[n];
        const uint64_t B = on - o0;
        c->d_reads.ensure(B + 64);
        c->d_offs.ensure((n + 1) * 8);
        if (B) HIPX(hipMemcpyAsync(c->d_reads.p, reads, B, hipMemcpyHostToDevice, c->st));
        HIPX(hipMemcpyAsync(c->d_offs.p, offs, (n + 1) * 8, hipMemcpyHostToDevice, c->st));
        dreads = c->d_reads.as<uint8_t>();
        doffs = c->d_offs.as<uint64_t>();
    }
    const uint64_t B = on - o0;
    if (B + lead >= (1ull << 31)) throw ArgError{"a batch must hold fewer than 2^31 bases", GKC_EINVAL};
    cd.reads = dreads; cd.offs = doffs; cd.lens = nullptr; cd.base = o0 - lead;
    cd.n_reads = (uint32_t)n; cd.n_bases = (uint32_t)(B + lead); cd.n_tiles = cdiv(B + lead, P1_T); cd.lead = lead;
    if (c->reads_done + (uint64_t)n >= (1ull << 32)) throw ArgError{"more than 2^32 reads since the last reset", GKC_EINVAL};
    cd.read0 = (uint32_t)c->reads_done;
    if (c->prm.hpc && n > 0) {
        // compressed reads keep the input's layout, lead included: read r's runs
        // start at the same position, so the tiles index them like the input
        c->d_hpc_reads.ensure(B + lead + 64);
        c->d_lens.ensure((n + 1) * 4);
        if (hpc_pos) c->d_hpc_pos.ensure((B + 1) * 4);
        int blocks = std::min<int64_t>((n + 3) / 4, 65535);
        hipLaunchKernelGGL(k_hpc, dim3(blocks), dim3(256), 0, c->st, dreads + lead, doffs, (uint32_t)n,
                           c->d_hpc_reads.as<uint8_t>() + lead, c->d_lens.as<int32_t>(),
                           hpc_pos ? c->d_hpc_pos.as<int32_t>() : nullptr);
        HIPX(hipGetLastError());
        cd.reads = c->d_hpc_reads.as<uint8_t>();
        cd.lens = c->d_lens.as<int32_t>();
    }
    c->d_tiler0.ensure(((size_t)cd.n_tiles + 1) * 4);
    cd.tile_r0 = c->d_tiler0.as<uint32_t>();
    hipLaunchKernelGGL(k_tile_reads, dim3(std::max<uint64_t>(1, (n + 255) / 256)), dim3(256), 0, c->st, doffs,
                       (uint32_t)n, cd.base, cd.n_tiles, c->d_tiler0.as<uint32_t>());
    c->d_chunk.ensure(sizeof(ChunkDesc));
    hipLaunchKernelGGL(k_set_chunk, dim3(1), dim3(1), 0, c->st, c->d_chunk.as<ChunkDesc>(), cd);
    HIPX(hipGetLastError());
    return B;
}

static void launch_superkmers(gkc_ctx* c, const ChunkDesc& cd, P1Params prm, BinSet bs, bool persistent = false,
                              const ChunkDesc* dcd = nullptr) {
    if (!dcd) dcd = c->d_chunk.as<ChunkDesc>();
    if (cd.n_tiles == 0 && !persistent) return;
    const bool win = prm.win_mm != nullptr;
    const bool wide = c->prm.p > 16;
    const void* fn = win ? (const void*)k_superkmers<uint64_t, true>
                         : wide ? (const void*)k_superkmers<uint64_t, false> : (const void*)k_superkmers<uint32_t, false>;
    const size_t dyn = (size_t)prm.P * 12;          // per-partition counters
    int grid = occupancy_blocks(fn, P1_NT, c->n_cu, dyn);
    if (!persistent) grid = std::min<int>(std::max<uint32_t>(cd.n_tiles, 1), grid);
    if (win)
        hipLaunchKernelGGL((k_superkmers<uint64_t, true>), dim3(grid), dim3(P1_NT), dyn, c->st, dcd, prm, bs);
    else if (wide)
        hipLaunchKernelGGL((k_superkmers<uint64_t, false>), dim3(grid), dim3(P1_NT), dyn, c->st, dcd, prm, bs);
    else
        hipLaunchKernelGGL((k_superkmers<uint32_t, false>), dim3(grid), dim3(P1_NT), dyn, c->st, dcd, prm, bs);
    HIPX(hipGetLastError());
}

static P1Params p1params(gkc_ctx* c, int scatter) {
    P1Params p;
    p.k = c->prm.k; p.p = c->prm.p; p.fmode = c->prm.filter_mode; p.pmode = c->prm.part_mode;
    p.P = c->prm.n_parts; p.scatter = scatter; p.sketch = nullptr; p.mm_out = nullptr;
    p.win_mm = nullptr; p.win_pos = nullptr; p.win_strand = nullptr;
    return p;
}

// ---------------------------------------------------------------- counting

template <typename K>
static void count_impl(gkc_ctx* c, uint32_t flags);
template <typename K>
static void count_batched(gkc_ctx* c, uint32_t flags);

extern "C" {

const char* gkc_last_error(void) { return g_err.c_str(); }

void* gkc_host_alloc(uint64_t bytes) {
    void* p = nullptr;
    if (hipHostMalloc(&p, bytes ? bytes : 1, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        g_err = "pinned host allocation of " + std::to_string(bytes) + " bytes failed";
        return nullptr;
    }
    return p;
}

void gkc_host_free(void* p) {
    if (p) (void)hipHostFree(p);
}

int gkc_device_select(int device) {
    if (hipSetDevice(device) != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
        g_err = std::string("cannot select device ") + std::to_string(device) + ": " + hipGetErrorString(hipGetLastError());
        return GKC_EHIP;
    }
    return 0;
}
int gkc_version(void) { return 10000; }

int64_t gkc_parse_reads(const char* text, uint64_t n_bytes, int format, char* bases, uint64_t* offs, void* stream) {
    if (!offs || (n_bytes && (!text || !bases))) { g_err = "null buffer"; return GKC_EINVAL; }
    int64_t n = 0;
    std::string err;
    const int rc = gkc::parse_reads(text, n_bytes, format, bases, offs, static_cast<hipStream_t>(stream), &n, err);
    if (rc) { g_err = err; return rc; }
    return n;
}

int64_t gkc_parse_reads_noalloc(const char* text, uint64_t n_bytes, int format, char* bases, uint64_t* offs, void* stream) {
    if (!offs || (n_bytes && (!text || !bases))) { g_err = "null buffer"; return GKC_EINVAL; }
    int64_t n = 0;
    std::string err;
    const int rc = gkc::parse_reads_noalloc(text, n_bytes, format, bases, offs, static_cast<hipStream_t>(stream), &n, err);
    if (rc) { g_err = err; return rc; }
    return n;
}

gkc_ctx* gkc_create(int device, const gkc_params* prm) {
    if (!prm) { g_err = "null params"; return nullptr; }
    const gkc_params& q = *prm;
    if (q.k < 1 || q.k > 64 || q.p < 3 || q.p > 32 || q.p > q.k || q.n_parts < 1 || q.n_parts > 1024 ||
        (q.filter_mode != 0 && q.filter_mode != 1) || (q.part_mode != 0 && q.part_mode != 1)) {
        g_err = "invalid parameters (need 1<=k<=64, 3<=p<=min(k,32), 1<=n_parts<=1024, modes 0/1)";
        return nullptr;
    }
    gkc_ctx* c = new gkc_ctx();
    c->device = device;
    c->prm = q;
    if (c->prm.min_count == 0) c->prm.min_count = 1;
    c->keyw = q.k <= 32 ? 8 : 16;
    c->npass = (2 * q.k + 7) / 8;
    try {
        HIPX(hipSetDevice(device));
        hipDeviceProp_t pr;
        HIPX(hipGetDeviceProperties(&pr, device));
        c->n_cu = pr.multiProcessorCount;
        HIPX(hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking));
        for (auto& e : c->ev) HIPX(hipEventCreate(&e));
        for (auto& e : c->evp) HIPX(hipEventCreate(&e));
        bins_init(c);
        HIPX(hipStreamSynchronize(c->st));
    } catch (const HipError& e) {
        (void)hipGetLastError();
        g_err = std::string("HIP error ") + hipGetErrorString(e.e) + " in " + e.where;
        delete c;
        return nullptr;
    } catch (const ArgError& e) {
        g_err = e.msg;
        delete c;
        return nullptr;
    }
    return c;
}

void gkc_destroy(gkc_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->st) (void)hipStreamSynchronize(c->st);
    if (c->g_sort) (void)hipGraphExecDestroy(c->g_sort);
    if (c->g_emit) (void)hipGraphExecDestroy(c->g_emit);
    if (c->g_emit_p) (void)hipGraphExecDestroy(c->g_emit_p);
    if (c->g_cnt) (void)hipGraphExecDestroy(c->g_cnt);
    if (c->g_p1) (void)hipGraphExecDestroy(c->g_p1);
    for (auto& e : c->ev) if (e) (void)hipEventDestroy(e);
    for (auto& e : c->evp) if (e) (void)hipEventDestroy(e);
    if (c->ev_lsum) (void)hipEventDestroy(c->ev_lsum);
    if (c->h_lsum) (void)hipHostFree(c->h_lsum);
    if (c->h_curpack) (void)hipHostFree(c->h_curpack);
    if (c->h_oend) (void)hipHostFree(c->h_oend);
    if (c->st) (void)hipStreamDestroy(c->st);
    delete c;
}

void* gkc_stream(gkc_ctx* c) { return c ? (void*)c->st : nullptr; }

int gkc_synchronize(gkc_ctx* c) {
    return guarded(c, [&] { HIPX(hipStreamSynchronize(c->st)); return 0; });
}

int gkc_reset(gkc_ctx* c) {
    return guarded(c, [&] {
        // keep the bin regions and capacities, clear the cursors and the sketch
        const uint32_t P = c->prm.n_parts;
        std::fill(c->h_word_cur.begin(), c->h_word_cur.end(), 0u);
        std::fill(c->h_hdr_cur.begin(), c->h_hdr_cur.end(), 0u);
        std::fill(c->h_kmer_cur.begin(), c->h_kmer_cur.end(), 0u);
        zero(c, c->d_word_cur.p, P); zero(c, c->d_hdr_cur.p, P); zero(c, c->d_kmer_cur.p, P); zero(c, c->d_hk_cur.p, 2 * P);
        zero(c, c->d_sketch.p, (uint64_t)P * 256); zero(c, c->d_overflow.p, 4);
        c->bases_done = 0;
        c->reads_done = 0;
        c->pos_mode = 0;
        c->p1_fallbacks = 0;
        c->counted = false;
        HIPX(hipStreamSynchronize(c->st));
        return 0;
    });
}

int gkc_superkmer_arrays(gkc_ctx* c, const char* reads, const uint64_t* offs, int64_t n, uint32_t* minimizers,
                         int32_t* skm_offs, char* reads_out, int32_t* read_len, int32_t* hpc_orig_pos) {
    return guarded(c, [&] {
        if (n < 0 || !offs || (n > 0 && !reads)) throw ArgError{"bad batch", GKC_EINVAL};
        if (c->prm.p > 16) throw ArgError{"the reference's u32 minimizer arrays need p <= 16", GKC_EINVAL};
        if (n == 0) return 0;
        {   // the offset lists [0, i1, .., len-k+1, 0] cannot skip windows: reads must be A/C/G/T only
            static const auto acgt = [] {
                std::array<uint8_t, 256> t{};
                for (uint8_t ch : {'A', 'C', 'G', 'T', 'a', 'c', 'g', 't'}) t[ch] = 1;
                return t;
            }();
            const uint8_t* b = reinterpret_cast<const uint8_t*>(reads);
            const uint64_t nb = offs[n] - offs[0];
            uint8_t all = 1;
            for (uint64_t i = 0; i < nb; i++) all &= acgt[b[i]];
            if (!all) throw ArgError{"read bytes other than A/C/G/T: split the reads at them first (the offset "
                                     "lists cannot skip windows)", GKC_EINVAL};
        }
        ChunkDesc cd;
        const uint64_t B = stage_batch(c, reads, offs, n, 0, cd, true);
        c->d_mm.ensure((B + 1) * 4);
        HIPX(hipMemsetAsync(c->d_mm.p, 0, (B + 1) * 4, c->st));
        P1Params prm = p1params(c, 0);
        prm.mm_out = c->d_mm.as<uint32_t>();
        BinSet bs = binset(c, true);
        HIPX(hipMemsetAsync(c->d_tmp_hk.p, 0, c->prm.n_parts * 8, c->st));
        HIPX(hipMemsetAsync(c->d_tmp_word.p, 0, c->prm.n_parts * 4, c->st));
        launch_superkmers(c, cd, prm, bs);
        // superkmer lists (reuse d_hpc_pos-sized scratch: i32 per base)
        DevBuf d_skm;
        d_skm.ensure((B + 1) * 4);
        HIPX(hipMemsetAsync(d_skm.p, 0, (B + 1) * 4, c->st));
        int blocks = std::min<int64_t>((n + 3) / 4, 65535);
        hipLaunchKernelGGL(k_skm_lists, dim3(blocks), dim3(256), 0, c->st, cd.offs, cd.lens, (uint32_t)n, c->prm.k,
                           c->d_mm.as<uint32_t>(), d_skm.as<int32_t>());
        HIPX(hipGetLastError());
        if (minimizers) HIPX(hipMemcpyAsync(minimizers, c->d_mm.p, B * 4, hipMemcpyDeviceToHost, c->st));
        if (skm_offs) HIPX(hipMemcpyAsync(skm_offs, d_skm.p, B * 4, hipMemcpyDeviceToHost, c->st));
        if (reads_out) HIPX(hipMemcpyAsync(reads_out, cd.reads, B, hipMemcpyDeviceToHost, c->st));
        if (c->prm.hpc) {
            if (read_len) HIPX(hipMemcpyAsync(read_len, c->d_lens.p, n * 4, hipMemcpyDeviceToHost, c->st));
            if (hpc_orig_pos) HIPX(hipMemcpyAsync(hpc_orig_pos, c->d_hpc_pos.p, B * 4, hipMemcpyDeviceToHost, c->st));
        }
        HIPX(hipStreamSynchronize(c->st));
        return 0;
    });
}

int gkc_window_minimizers(gkc_ctx* c, const char* reads, const uint64_t* offs, int64_t n, uint32_t* mm, uint8_t* pos,
                          int8_t* strand) {
    return guarded(c, [&] {
        if (n < 0 || !offs || (n > 0 && !reads)) throw ArgError{"bad batch", GKC_EINVAL};
        if (c->prm.p > 16) throw ArgError{"per-window records hold u32 minimizers: p <= 16", GKC_EINVAL};
        if (n == 0) return 0;
        ChunkDesc cd;
        const uint64_t B = stage_batch(c, reads, offs, n, 0, cd);
        c->d_win_mm.ensure((B + 1) * 4); c->d_win_pos.ensure(B + 1); c->d_win_str.ensure(B + 1);
        P1Params prm = p1params(c, 0);
        prm.win_mm = c->d_win_mm.as<uint32_t>(); prm.win_pos = c->d_win_pos.as<uint8_t>();
        prm.win_strand = c->d_win_str.as<int8_t>();
        const uint32_t P = c->prm.n_parts;
        zero(c, c->d_tmp_hk.p, 2 * P); zero(c, c->d_tmp_word.p, P);
        launch_superkmers(c, cd, prm, binset(c, true));
        if (mm) HIPX(hipMemcpyAsync(mm, c->d_win_mm.p, B * 4, hipMemcpyDeviceToHost, c->st));
        if (pos) HIPX(hipMemcpyAsync(pos, c->d_win_pos.p, B, hipMemcpyDeviceToHost, c->st));
        if (strand) HIPX(hipMemcpyAsync(strand, c->d_win_str.p, B, hipMemcpyDeviceToHost, c->st));
        HIPX(hipStreamSynchronize(c->st));
        return 0;
    });
}

int gkc_windows_device(gkc_ctx* c, uint32_t** mm, uint8_t** pos, int8_t** strand, uint64_t* n) {
    return guarded(c, [&] {
        if (!c->win_n) throw ArgError{"no per-window records (gkc_partition without GKC_EMIT_WINDOWS)", GKC_ESTATE};
        // records are indexed by batch-local base position (+ the alignment lead)
        if (mm) *mm = c->d_win_mm.as<uint32_t>() + c->win_lead;
        if (pos) *pos = c->d_win_pos.as<uint8_t>() + c->win_lead;
        if (strand) *strand = c->d_win_str.as<int8_t>() + c->win_lead;
        if (n) *n = c->win_n - c->win_lead;
        return 0;
    });
}

// Plain phase 1 of one batch: count pass, host sync, exact reservation, scatter.
// The kernels' combined cursors into the per-partition header and k-mer
// counts the counting launch reads.
__global__ void k_split_hk(const unsigned long long* __restrict__ hk, uint32_t* __restrict__ hdr,
                           uint32_t* __restrict__ kmer, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const unsigned long long v = hk[i];
        hdr[i] = (uint32_t)v;
        kmer[i] = (uint32_t)(v >> 32);
    }
}
// The word, header and k-mer cursors and the overflow flag in one array
// [3P + 1], so one copy brings them to the host.
__global__ void k_pack_cursors(const uint32_t* __restrict__ wcur, const uint32_t* __restrict__ hcur,
                               const uint32_t* __restrict__ kcur, const uint32_t* __restrict__ ovf,
                               uint32_t* __restrict__ out, uint32_t P) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < 3 * P + 1; i += gridDim.x * blockDim.x)
        out[i] = i < P ? wcur[i] : i < 2 * P ? hcur[i - P] : i < 3 * P ? kcur[i - 2 * P] : *ovf;
}
static void split_cursors(gkc_ctx* c) {
    const uint32_t P = c->prm.n_parts;
    hipLaunchKernelGGL(k_split_hk, dim3((P + 255) / 256), dim3(256), 0, c->st, c->d_hk_cur.as<unsigned long long>(),
                       c->d_hdr_cur.as<uint32_t>(), c->d_kmer_cur.as<uint32_t>(), P);
    HIPX(hipGetLastError());
}

static void partition_plain(gkc_ctx* c, const ChunkDesc& cd, uint32_t flags) {
    const uint32_t P = c->prm.n_parts;
    zero(c, c->d_tmp_hk.p, 2 * P); zero(c, c->d_tmp_word.p, P);
    P1Params pc = p1params(c, 0);
    launch_superkmers(c, cd, pc, binset(c, true));
    std::vector<uint32_t> nh(P), nw(P), nk(P);
    std::vector<uint64_t> hk(P);
    HIPX(hipMemcpyAsync(hk.data(), c->d_tmp_hk.p, P * 8, hipMemcpyDeviceToHost, c->st));
    HIPX(hipMemcpyAsync(nw.data(), c->d_tmp_word.p, P * 4, hipMemcpyDeviceToHost, c->st));
    HIPX(hipStreamSynchronize(c->st));
    for (uint32_t q = 0; q < P; q++) { nh[q] = (uint32_t)hk[q]; nk[q] = (uint32_t)(hk[q] >> 32); }
    for (uint32_t q = 0; q < P; q++)
        if ((uint64_t)c->h_kmer_cur[q] + nk[q] >= (1ull << 30)) throw ArgError{"partition exceeds 2^30 k-mers; use more partitions", GKC_EINVAL};
    bins_reserve(c, nw, nh);
    P1Params ps = p1params(c, 1);
    ps.sketch = c->d_sketch.as<uint32_t>();
    if (flags & GKC_EMIT_WINDOWS) {
        if (c->prm.p > 16) throw ArgError{"per-window records hold u32 minimizers: p <= 16", GKC_EINVAL};
        const uint64_t B = cd.n_bases;
        c->d_win_mm.ensure((B + 1) * 4); c->d_win_pos.ensure(B + 1); c->d_win_str.ensure(B + 1);
        ps.win_mm = c->d_win_mm.as<uint32_t>(); ps.win_pos = c->d_win_pos.as<uint8_t>();
        ps.win_strand = c->d_win_str.as<int8_t>();
        c->win_n = B; c->win_lead = cd.lead;
    }
    launch_superkmers(c, cd, ps, binset(c, false));
    split_cursors(c);
    uint32_t ovf = 0;
    HIPX(hipMemcpyAsync(&ovf, c->d_overflow.p, 4, hipMemcpyDeviceToHost, c->st));
    HIPX(hipStreamSynchronize(c->st));
    if (ovf) throw ArgError{"internal: bin overflow after exact reservation", GKC_ESTATE};
    for (uint32_t q = 0; q < P; q++) { c->h_hdr_cur[q] += nh[q]; c->h_word_cur[q] += nw[q]; c->h_kmer_cur[q] += nk[q]; }
}

// Dense table layout of a presized binned count, from the exact kept
// entries per segment (counted by then): starts = exclusive scan, and each
// segment's capacity = what is left of the allocation after its start, so
// only the total has to fit.  One block of 1024 threads, n_segs <= 1024.
__global__ __launch_bounds__(1024) void k_table_layout(const uint32_t* __restrict__ n_entries, uint32_t n_segs,
                                                       const uint64_t* __restrict__ total_cap,
                                                       uint64_t* __restrict__ table_start, uint32_t* __restrict__ table_cap) {
    __shared__ uint64_t s_w[1024 / GKC_WAVE];
    const uint32_t t = threadIdx.x, lane = __lane_id(), wv = t / GKC_WAVE;
    const uint64_t v = t < n_segs ? n_entries[t] : 0u;
    uint64_t inc = v;
#pragma unroll
    for (int off = 1; off < GKC_WAVE; off <<= 1) {
        const uint64_t y = __shfl_up(inc, off);
        if ((int)lane >= off) inc += y;
    }
    if (lane == GKC_WAVE - 1) s_w[wv] = inc;
    __syncthreads();
    uint64_t before = 0;
    for (uint32_t w = 0; w < wv; w++) before += s_w[w];
    const uint64_t start = before + inc - v, tot = *total_cap;
    if (t < n_segs) {
        table_start[t] = start;
        table_cap[t] = start >= tot ? 0u : (tot - start > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)(tot - start));
    }
}
__global__ void k_copy_words(uint32_t* __restrict__ dst, const uint32_t* __restrict__ src, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// Batched phase 1 of one batch: bins presized from the per-partition shares of
// the batches already partitioned (for the first batch, from a count pass
// over its first 1/16: reads come in no genome order, so that is a random
// sample), a single scatter pass replayed from a captured graph (the chunk
// descriptor lives in device memory), one host sync.  Returns false when a
// bin overflowed (the cursors are then restored and the caller runs the
// plain form).
static bool partition_presized(gkc_ctx* c, const ChunkDesc& cd) {
    const uint32_t P = c->prm.n_parts;
    std::vector<uint32_t> nw(P), nh(P);
    if (c->bases_done == 0) {
        ChunkDesc sc = cd;
        sc.n_tiles = std::max<uint32_t>(1, cd.n_tiles / 16);
        sc.n_bases = std::min<uint32_t>(cd.n_bases, sc.n_tiles * (uint32_t)P1_T);
        c->d_chunk_s.ensure(sizeof(ChunkDesc));
        hipLaunchKernelGGL(k_set_chunk, dim3(1), dim3(1), 0, c->st, c->d_chunk_s.as<ChunkDesc>(), sc);
        zero(c, c->d_tmp_hk.p, 2 * P); zero(c, c->d_tmp_word.p, P);
        launch_superkmers(c, sc, p1params(c, 0), binset(c, true), false, c->d_chunk_s.as<ChunkDesc>());
        std::vector<uint32_t> sw(P), sh(P);
        std::vector<uint64_t> shk(P);
        HIPX(hipMemcpyAsync(sw.data(), c->d_tmp_word.p, P * 4, hipMemcpyDeviceToHost, c->st));
        HIPX(hipMemcpyAsync(shk.data(), c->d_tmp_hk.p, P * 8, hipMemcpyDeviceToHost, c->st));
        HIPX(hipStreamSynchronize(c->st));
        for (uint32_t q = 0; q < P; q++) sh[q] = (uint32_t)shk[q];
        const double f = (double)cd.n_bases / (double)sc.n_bases;
        for (uint32_t q = 0; q < P; q++) {
            nw[q] = (uint32_t)std::min<double>(std::ceil(sw[q] * f * 1.25) + 1024, 4.0e9);
            nh[q] = (uint32_t)std::min<double>(std::ceil(sh[q] * f * 1.25) + 256, 4.0e9);
        }
    } else {
        const double f = (double)cd.n_bases / (double)c->bases_done;
        for (uint32_t q = 0; q < P; q++) {
            nw[q] = (uint32_t)std::min<double>(std::ceil(c->h_word_cur[q] * f * 1.25) + 256, 4.0e9);
            nh[q] = (uint32_t)std::min<double>(std::ceil(c->h_hdr_cur[q] * f * 1.25) + 64, 4.0e9);
        }
    }
    bins_reserve(c, nw, nh);
    c->d_backup.ensure((size_t)(3 + 256) * P * 4);     // cursors, then the sketch
    c->d_curpack.ensure((3 * (size_t)P + 1) * 4);
    if (!c->h_curpack) HIPX(hipHostMalloc((void**)&c->h_curpack, (3 * (size_t)P + 1) * 4, hipHostMallocDefault));
    std::vector<const void*> key = {c->words.p, c->hdrs.p, c->d_chunk.p, c->d_backup.p, c->skmpos.p,
                                    (const void*)(uintptr_t)c->pos_mode, c->d_curpack.p};
    auto enqueue = [&] {
        hipLaunchKernelGGL(k_copy_words, dim3(4), dim3(256), 0, c->st, c->d_backup.as<uint32_t>(),
                           (const uint32_t*)c->d_word_cur.as<uint32_t>(), (uint64_t)P);
        hipLaunchKernelGGL(k_copy_words, dim3(8), dim3(256), 0, c->st, c->d_backup.as<uint32_t>() + P,
                           (const uint32_t*)c->d_hk_cur.as<uint32_t>(), (uint64_t)2 * P);
        hipLaunchKernelGGL(k_copy_words, dim3(256), dim3(256), 0, c->st, c->d_backup.as<uint32_t>() + 3 * P,
                           (const uint32_t*)c->d_sketch.as<uint32_t>(), (uint64_t)P * 256);
        zero(c, c->d_overflow.p, 1);
        P1Params ps = p1params(c, 1);
        ps.sketch = c->d_sketch.as<uint32_t>();
        launch_superkmers(c, cd, ps, binset(c, false), true);
        split_cursors(c);
        hipLaunchKernelGGL(k_pack_cursors, dim3((3 * P + 1 + 255) / 256), dim3(256), 0, c->st,
                           (const uint32_t*)c->d_word_cur.as<uint32_t>(), (const uint32_t*)c->d_hdr_cur.as<uint32_t>(),
                           (const uint32_t*)c->d_kmer_cur.as<uint32_t>(), (const uint32_t*)c->d_overflow.as<uint32_t>(),
                           c->d_curpack.as<uint32_t>(), P);
    };
    // occupancy is queried (and cached) outside the capture
    (void)occupancy_blocks(c->prm.p > 16 ? (const void*)k_superkmers<uint64_t, false> : (const void*)k_superkmers<uint32_t, false>,
                           P1_NT, c->n_cu, (size_t)c->prm.n_parts * 12);
    if (!c->g_p1 || c->g_p1_key != key) {
        if (c->g_p1) { HIPX(hipGraphExecDestroy(c->g_p1)); c->g_p1 = nullptr; }
        c->g_p1 = capture(c, enqueue);
        c->g_p1_key = key;
    }
    HIPX(hipGraphLaunch(c->g_p1, c->st));
    HIPX(hipMemcpyAsync(c->h_curpack, c->d_curpack.p, (3 * (size_t)P + 1) * 4, hipMemcpyDeviceToHost, c->st));
    HIPX(hipStreamSynchronize(c->st));
    const uint32_t* cur = c->h_curpack;
    const uint32_t ovf = cur[3 * P];
    bool too_big = false;
    for (uint32_t q = 0; q < P; q++) too_big |= cur[2 * P + q] >= (1u << 30);
    if (ovf || too_big) {
        hipLaunchKernelGGL(k_copy_words, dim3(4), dim3(256), 0, c->st, c->d_word_cur.as<uint32_t>(),
                           (const uint32_t*)c->d_backup.as<uint32_t>(), (uint64_t)P);
        hipLaunchKernelGGL(k_copy_words, dim3(8), dim3(256), 0, c->st, c->d_hk_cur.as<uint32_t>(),
                           (const uint32_t*)(c->d_backup.as<uint32_t>() + P), (uint64_t)2 * P);
        split_cursors(c);
        hipLaunchKernelGGL(k_copy_words, dim3(256), dim3(256), 0, c->st, c->d_sketch.as<uint32_t>(),
                           (const uint32_t*)(c->d_backup.as<uint32_t>() + 3 * P), (uint64_t)P * 256);
        zero(c, c->d_overflow.p, 1);
        HIPX(hipStreamSynchronize(c->st));
        c->p1_fallbacks++;
        return false;
    }
    for (uint32_t q = 0; q < P; q++) { c->h_word_cur[q] = cur[q]; c->h_hdr_cur[q] = cur[P + q]; c->h_kmer_cur[q] = cur[2 * P + q]; }
    return true;
}

int gkc_partition(gkc_ctx* c, const char* reads, const uint64_t* offs, int64_t n, uint32_t flags) {
    return guarded(c, [&] {
        if (n < 0 || !offs || (n > 0 && !reads)) throw ArgError{"bad batch", GKC_EINVAL};
        if (n == 0) return 0;
        const int mode = (flags & GKC_WITH_POS) ? 1 : 2;
        if (c->pos_mode && c->pos_mode != mode)
            throw ArgError{"every batch between resets must agree on GKC_WITH_POS", GKC_EINVAL};
        if (mode == 1 && c->pos_mode == 0) {        // no headers yet: a positions region for the current capacity
            uint64_t th = 0;
            for (uint32_t q = 0; q < c->prm.n_parts; q++) th += c->h_hdr_cap[q];
            c->skmpos.ensure(th * sizeof(uint2) + 16);
        }
        c->pos_mode = mode;
        ChunkDesc cd;
        HIPX(hipEventRecord(c->ev[0], c->st));
        stage_batch(c, reads, offs, n, flags, cd);
        bool done = false;
        if ((flags & GKC_BATCHED) && !(flags & GKC_EMIT_WINDOWS))
            done = partition_presized(c, cd);        // a failed attempt restores the cursors and the sketch
        if (!done) partition_plain(c, cd, flags);
        c->bases_done += cd.n_bases - cd.lead;
        c->reads_done += (uint64_t)n;
        HIPX(hipEventRecord(c->ev[1], c->st));
        HIPX(hipEventSynchronize(c->ev[1]));
        float ms = 0;
        HIPX(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
        c->stats.ms_partition = ms;
        c->counted = false;
        return 0;
    });
}

int gkc_count(gkc_ctx* c, uint32_t flags) {
    return guarded(c, [&] {
        if (flags & GKC_BATCHED) {
            if (c->keyw == 8) count_batched<uint64_t>(c, flags);
            else count_batched<u128k>(c, flags);
        } else {
            if (c->keyw == 8) count_impl<uint64_t>(c, flags);
            else count_impl<u128k>(c, flags);
        }
        c->counted = true;
        return 0;
    });
}

int gkc_table_info(gkc_ctx* c, uint64_t* part_off, uint64_t* part_distinct, uint64_t* part_kmers) {
    return guarded(c, [&] {
        if (!c->counted) throw ArgError{"gkc_count has not run", GKC_ESTATE};
        const uint32_t P = c->prm.n_parts;
        if (part_off) std::memcpy(part_off, c->part_off.data(), (P + 1) * 8);
        if (part_distinct) std::memcpy(part_distinct, c->part_distinct.data(), P * 8);
        if (part_kmers) std::memcpy(part_kmers, c->part_kmers.data(), P * 8);
        return 0;
    });
}

int64_t gkc_table_copy(gkc_ctx* c, int32_t part, uint64_t* key_lo, uint64_t* key_hi, uint32_t* cnt, uint64_t cap) {
    int64_t ret = 0;
    int rc = guarded(c, [&] {
        if (!c->counted) throw ArgError{"gkc_count has not run", GKC_ESTATE};
        const uint32_t P = c->prm.n_parts;
        if (part >= (int32_t)P) throw ArgError{"partition out of range", GKC_EINVAL};
        const uint32_t q0 = part < 0 ? 0 : part, q1 = part < 0 ? P : part + 1;
        const uint64_t a = c->part_off[q0], n = c->part_off[q1] - a;
        if (n > cap) throw ArgError{"table copy buffer too small", GKC_ECAP};
        std::vector<u128k> tmp;
        if (c->keyw == 16) tmp.resize(n);
        // runs of partitions that are physically contiguous are copied at once
        uint32_t q = q0;
        while (q < q1) {
            uint32_t e = q + 1;
            while (e < q1 && c->part_start[e] == c->part_start[q] + (c->part_off[e] - c->part_off[q])) e++;
            const uint64_t m = c->part_off[e] - c->part_off[q], src = c->part_start[q], dst = c->part_off[q] - a;
            if (m) {
                if (c->keyw == 8) HIPX(hipMemcpyAsync(key_lo + dst, c->tkeys.as<uint64_t>() + src, m * 8, hipMemcpyDeviceToHost, c->st));
                else HIPX(hipMemcpyAsync(tmp.data() + dst, c->tkeys.as<u128k>() + src, m * 16, hipMemcpyDeviceToHost, c->st));
                if (cnt) HIPX(hipMemcpyAsync(cnt + dst, c->tcnts.as<uint32_t>() + src, m * 4, hipMemcpyDeviceToHost, c->st));
            }
            q = e;
        }
        HIPX(hipStreamSynchronize(c->st));
        if (c->keyw == 16)
            for (uint64_t i = 0; i < n; i++) { key_lo[i] = tmp[i].lo; if (key_hi) key_hi[i] = tmp[i].hi; }
        ret = (int64_t)n;
        return 0;
    });
    return rc < 0 ? rc : ret;
}

namespace {
struct PinnedBuf {
    void* p = nullptr;
    ~PinnedBuf() { if (p) (void)hipHostFree(p); }
};
struct FileFd {
    int fd = -1;
    ~FileFd() { if (fd >= 0) ::close(fd); }
};
struct Events2 {
    hipEvent_t e[2]{};
    ~Events2() { for (auto& x : e) if (x) (void)hipEventDestroy(x); }
};
void pwrite_all(int fd, const void* b, size_t n, uint64_t off, const char* path) {
    const char* q = static_cast<const char*>(b);
    while (n) {
        const ssize_t w = ::pwrite(fd, q, n, (off_t)off);
        if (w < 0 && errno == EINTR) continue;
        if (w <= 0) throw ArgError{std::string("writing ") + path + " failed: " + std::strerror(w < 0 ? errno : EIO), GKC_EIO};
        q += w;
        n -= (size_t)w;
        off += (uint64_t)w;
    }
}
constexpr size_t TF_STAGE = 64u << 20;   // bytes per pinned staging buffer
}  // namespace

int64_t gkc_table_write(gkc_ctx* c, const char* path) {
    int64_t ret = 0;
    int rc = guarded(c, [&] {
        if (!c->counted) throw ArgError{"gkc_count has not run", GKC_ESTATE};
        if (!path) throw ArgError{"null path", GKC_EINVAL};
        const uint32_t P = c->prm.n_parts, kw = (uint32_t)c->keyw / 8;
        const uint64_t N = c->part_off[P];
        uint64_t nk = 0;
        for (uint32_t q = 0; q < P; q++) nk += c->part_kmers[q];
        unsigned char hdr[64] = {};
        std::memcpy(hdr, "GKCTAB1", 8);
        const uint32_t h32[9] = {1u, (uint32_t)c->prm.k, (uint32_t)c->prm.p, (uint32_t)c->prm.hpc, P, kw,
                                 c->prm.min_count, c->prm.max_count, 0u};
        std::memcpy(hdr + 8, h32, 32);
        std::memcpy(hdr + 40, &N, 8);
        std::memcpy(hdr + 48, &nk, 8);
        const uint64_t o_info = 64, o_keys = o_info + 8 * (3 * (uint64_t)P + 1), o_cnts = o_keys + N * 8 * kw;
        FileFd f;
        f.fd = ::open(path, O_WRONLY | O_CREAT | O_TRUNC, 0644);
        if (f.fd < 0) throw ArgError{std::string("cannot create ") + path + ": " + std::strerror(errno), GKC_EIO};
        pwrite_all(f.fd, hdr, 64, 0, path);
        pwrite_all(f.fd, c->part_off.data(), 8 * ((size_t)P + 1), o_info, path);
        pwrite_all(f.fd, c->part_distinct.data(), 8 * (size_t)P, o_info + 8 * ((uint64_t)P + 1), path);
        pwrite_all(f.fd, c->part_kmers.data(), 8 * (size_t)P, o_info + 8 * (2 * (uint64_t)P + 1), path);
        // chunks of the physically contiguous partition runs, at most `per` entries each
        struct Chunk { uint64_t src, dst, m; };
        std::vector<Chunk> ch;
        const uint64_t per = TF_STAGE / (8 * kw + 4);
        for (uint32_t q = 0; q < P;) {
            uint32_t e = q + 1;
            while (e < P && c->part_start[e] == c->part_start[q] + (c->part_off[e] - c->part_off[q])) e++;
            const uint64_t m = c->part_off[e] - c->part_off[q];
            for (uint64_t i = 0; i < m; i += per) ch.push_back({c->part_start[q] + i, c->part_off[q] + i, std::min(per, m - i)});
            q = e;
        }
        PinnedBuf stage[2];
        Events2 ev;
        struct Drain {                                 // no copy into the staging buffers outlives them
            hipStream_t s;
            ~Drain() { (void)hipStreamSynchronize(s); }
        } drain{c->st};
        for (int b = 0; b < 2 && b < (int)ch.size(); b++) {
            HIPX(hipHostMalloc(&stage[b].p, TF_STAGE, hipHostMallocDefault));
            HIPX(hipEventCreateWithFlags(&ev.e[b], hipEventDisableTiming));
        }
        const char* tk = c->tkeys.as<char>();
        const uint32_t* tc = c->tcnts.as<uint32_t>();
        auto issue = [&](size_t j) {
            const int b = (int)(j & 1);
            char* h = static_cast<char*>(stage[b].p);
            HIPX(hipMemcpyAsync(h, tk + ch[j].src * 8 * kw, ch[j].m * 8 * kw, hipMemcpyDeviceToHost, c->st));
            HIPX(hipMemcpyAsync(h + ch[j].m * 8 * kw, tc + ch[j].src, ch[j].m * 4, hipMemcpyDeviceToHost, c->st));
            HIPX(hipEventRecord(ev.e[b], c->st));
        };
        if (!ch.empty()) issue(0);
        for (size_t j = 0; j < ch.size(); j++) {
            if (j + 1 < ch.size()) issue(j + 1);       // its buffer's previous chunk (j - 1) is written
            const int b = (int)(j & 1);
            HIPX(hipEventSynchronize(ev.e[b]));
            const char* h = static_cast<const char*>(stage[b].p);
            pwrite_all(f.fd, h, ch[j].m * 8 * kw, o_keys + ch[j].dst * 8 * kw, path);
            pwrite_all(f.fd, h + ch[j].m * 8 * kw, ch[j].m * 4, o_cnts + ch[j].dst * 4, path);
        }
        if (::close(f.fd) != 0) { f.fd = -1; throw ArgError{std::string("closing ") + path + " failed: " + std::strerror(errno), GKC_EIO}; }
        f.fd = -1;
        ret = (int64_t)(o_cnts + 4 * N);
        return 0;
    });
    return rc < 0 ? rc : ret;
}

int gkc_table_starts(gkc_ctx* c, uint64_t* starts) {
    return guarded(c, [&] {
        if (!c->counted) throw ArgError{"gkc_count has not run", GKC_ESTATE};
        std::memcpy(starts, c->part_start.data(), c->prm.n_parts * 8);
        return 0;
    });
}

int gkc_table_device(gkc_ctx* c, void** keys, uint32_t** counts, uint64_t* n_entries) {
    return guarded(c, [&] {
        if (!c->counted) throw ArgError{"gkc_count has not run", GKC_ESTATE};
        if (keys) *keys = c->tkeys.p;
        if (counts) *counts = c->tcnts.as<uint32_t>();
        if (n_entries) *n_entries = c->part_off[c->prm.n_parts];
        return 0;
    });
}

int64_t gkc_superkmer_positions(gkc_ctx* c, int32_t part, uint32_t* read, uint32_t* offset, uint32_t* n_kmers,
                                uint64_t cap) {
    return guarded(c, [&]() -> int64_t {
        if (part < 0 || (uint32_t)part >= c->prm.n_parts) throw ArgError{"partition out of range", GKC_EINVAL};
        if (c->pos_mode != 1) throw ArgError{"the bins hold no positions (partition with GKC_WITH_POS)", GKC_ESTATE};
        const uint32_t n = c->h_hdr_cur[part];
        if (!read && !offset && !n_kmers) return (int64_t)n;
        if (n > cap) throw ArgError{"output buffers too small for the partition's superkmers", GKC_ECAP};
        if (n == 0) return 0;
        std::vector<SkmHdr> hd(n);
        std::vector<uint2> ps(n);
        HIPX(hipMemcpyAsync(hd.data(), c->hdrs.as<SkmHdr>() + c->h_hdr_base[part], n * sizeof(SkmHdr), hipMemcpyDeviceToHost, c->st));
        HIPX(hipMemcpyAsync(ps.data(), c->skmpos.as<uint2>() + c->h_hdr_base[part], n * sizeof(uint2), hipMemcpyDeviceToHost, c->st));
        HIPX(hipStreamSynchronize(c->st));
        for (uint32_t i = 0; i < n; i++) {
            if (read) read[i] = ps[i].x;
            if (offset) offset[i] = ps[i].y;
            if (n_kmers) n_kmers[i] = hd[i].n_kmers;
        }
        return (int64_t)n;
    });
}

int gkc_kmer_positions(gkc_ctx* c, int32_t part, uint32_t sample_mod, uint64_t* keys, uint32_t* counts, uint64_t* pos_off,
                       uint64_t* positions, uint64_t entry_cap, uint64_t pos_cap, uint64_t* n_entries, uint64_t* n_positions) {
    return guarded(c, [&] {
        if (part < 0 || (uint32_t)part >= c->prm.n_parts) throw ArgError{"partition out of range", GKC_EINVAL};
        if (sample_mod == 0) throw ArgError{"sample_mod must be at least 1", GKC_EINVAL};
        if (c->keyw != 8) throw ArgError{"k-mer positions are listed for k <= 32", GKC_EINVAL};
        if (c->pos_mode != 1) throw ArgError{"the bins hold no positions (partition with GKC_WITH_POS)", GKC_ESTATE};
        if (!c->counted) throw ArgError{"count the bins first (gkc_count)", GKC_ESTATE};
        // the partition's table: sampled entries and their ranges of positions
        const uint64_t a = c->part_start[part], ne = c->part_off[part + 1] - c->part_off[part];
        std::vector<uint64_t> tk(ne), eo(ne, ~0ull);
        std::vector<uint32_t> tc(ne);
        if (ne) {
            HIPX(hipMemcpyAsync(tk.data(), c->tkeys.as<uint64_t>() + a, ne * 8, hipMemcpyDeviceToHost, c->st));
            HIPX(hipMemcpyAsync(tc.data(), c->tcnts.as<uint32_t>() + a, ne * 4, hipMemcpyDeviceToHost, c->st));
            HIPX(hipStreamSynchronize(c->st));
        }
        uint64_t n_e = 0, n_p = 0;
        for (uint64_t i = 0; i < ne; i++)
            if (mix64(tk[i]) % sample_mod == 0) { eo[i] = n_p; n_p += tc[i]; n_e++; }
        *n_entries = n_e; *n_positions = n_p;
        if (n_e > entry_cap || n_p > pos_cap) throw ArgError{"output buffers too small (sizes returned)", GKC_ECAP};
        if (n_e == 0) { if (pos_off) pos_off[0] = 0; return 0; }
        DevBuf d_eo, d_cur, d_out;
        d_eo.ensure(ne * 8); d_cur.ensure(ne * 4); d_out.ensure(n_p * 8 + 8);
        HIPX(hipMemcpyAsync(d_eo.p, eo.data(), ne * 8, hipMemcpyHostToDevice, c->st));
        zero(c, d_cur.p, ne);
        const uint32_t nh = c->h_hdr_cur[part];
        hipLaunchKernelGGL(k_pos_scatter, dim3(std::max<uint32_t>(1, std::min<uint32_t>((nh + 255) / 256, 8 * c->n_cu))),
                           dim3(256), 0, c->st, (const SkmHdr*)(c->hdrs.as<SkmHdr>() + c->h_hdr_base[part]),
                           (const uint2*)(c->skmpos.as<uint2>() + c->h_hdr_base[part]), nh,
                           (const uint32_t*)(c->words.as<uint32_t>() + c->h_word_base[part]), c->prm.k, sample_mod,
                           (const uint64_t*)(c->tkeys.as<uint64_t>() + a), (const uint32_t*)(c->tcnts.as<uint32_t>() + a),
                           ne, (const uint64_t*)d_eo.as<uint64_t>(),
                           d_cur.as<uint32_t>(), d_out.as<uint64_t>());
        HIPX(hipGetLastError());
        std::vector<uint64_t> out(n_p);
        HIPX(hipMemcpyAsync(out.data(), d_out.p, n_p * 8, hipMemcpyDeviceToHost, c->st));
        HIPX(hipStreamSynchronize(c->st));
        // each k-mer's occurrences, claimed in no set order on the device, in ascending order
        uint64_t e = 0;
        for (uint64_t i = 0; i < ne; i++) {
            if (eo[i] == ~0ull) continue;
            std::sort(out.begin() + eo[i], out.begin() + eo[i] + tc[i]);
            if (keys) keys[e] = tk[i];
            if (counts) counts[e] = tc[i];
            if (pos_off) pos_off[e] = eo[i];
            e++;
        }
        if (pos_off) pos_off[n_e] = n_p;
        if (positions) std::copy(out.begin(), out.end(), positions);
        return 0;
    });
}

int gkc_get_stats(gkc_ctx* c, gkc_stats* out) {
    return guarded(c, [&] {
        c->stats.partition_fallbacks = c->p1_fallbacks;
        if (out) *out = c->stats;
        return 0;
    });
}

int gkc_sketch(gkc_ctx* c, float* out) {
    return guarded(c, [&] {
        const uint32_t P = c->prm.n_parts;
        std::vector<uint32_t> h((size_t)P * 256);
        HIPX(hipMemcpyAsync(h.data(), c->d_sketch.p, h.size() * 4, hipMemcpyDeviceToHost, c->st));
        HIPX(hipStreamSynchronize(c->st));
        for (size_t i = 0; i < h.size(); i++) out[i] = std::log1p((float)h[i]);
        return 0;
    });
}

int gkc_set_tuning(gkc_ctx* c, const gkc_tuning* t) {
    return guarded(c, [&] {
        if (!t) throw ArgError{"null tuning", GKC_EINVAL};
        if (t->count_mode < 0 || t->count_mode > 1) throw ArgError{"count_mode must be 0 or 1", GKC_EINVAL};
        if (t->bin_passes < 0 || t->bin_passes > 3) throw ArgError{"bin_passes must be 0..3", GKC_EINVAL};
        if (t->sort_waves != 0 && t->sort_waves != 4 && t->sort_waves != 8)
            throw ArgError{"sort_waves must be 0, 4 or 8", GKC_EINVAL};
        if (t->digit_bits != 0 && (t->digit_bits < 6 || t->digit_bits > 8))
            throw ArgError{"digit_bits must be 0 or 6..8", GKC_EINVAL};
        if (t->count_window != 0 && (t->count_window < 128 || t->count_window > 1024 || t->count_window % 64))
            throw ArgError{"count_window must be 0 or a multiple of 64 in 128..1024", GKC_EINVAL};
        c->tune = *t;
        return 0;
    });
}

int gkc_get_tuning(gkc_ctx* c, gkc_tuning* t) {
    return guarded(c, [&] {
        gkc_tuning r = c->tune;
        r.sort_items = c->keyw == 8 ? sort_items<uint64_t>(c, 8) : sort_items<u128k>(c, 8);
        r.sort_waves = (c->keyw == 8 ? sort_nt<uint64_t>(c, 8) : sort_nt<u128k>(c, 8)) / 64;
        r.digit_bits = c->tune.digit_bits ? c->tune.digit_bits : 8;
        if (t) *t = r;
        return 0;
    });
}

int gkc_set_table_caps(gkc_ctx* c, const uint32_t* caps) {
    return guarded(c, [&] {
        c->presize_caps.assign(caps, caps + c->prm.n_parts);
        return 0;
    });
}

int gkc_learned_caps(gkc_ctx* c, uint32_t* caps) {
    return guarded(c, [&] {
        if (!c->mlp || !caps) throw ArgError{"no predictor set", GKC_ESTATE};
        HIPX(hipMemcpyAsync(caps, c->d_lcaps.p, (size_t)c->prm.n_parts * 4, hipMemcpyDeviceToHost, c->st));
        HIPX(hipStreamSynchronize(c->st));
        return 0;
    });
}

int gkc_set_predictor(gkc_ctx* c, gkc_mlp* mlp, const float* x_mean, const float* x_std, const float* t_mean,
                      const float* t_std, float margin) {
    return guarded(c, [&] {
        if (!mlp) { c->mlp = nullptr; return 0; }
        if (!x_mean || !x_std || !t_mean || !t_std || !(margin >= 1.0f))
            throw ArgError{"predictor: normalisation arrays and a margin >= 1 are required", GKC_EINVAL};
        if (c->prm.n_parts > 1024) throw ArgError{"predictor: at most 1024 partitions", GKC_EINVAL};
        std::vector<float> nm(2 * 256 + 5);
        std::copy(x_mean, x_mean + 256, nm.begin());
        std::copy(x_std, x_std + 256, nm.begin() + 256);
        nm[512] = t_mean[0]; nm[513] = t_mean[1]; nm[514] = t_std[0]; nm[515] = t_std[1]; nm[516] = margin;
        c->d_pnorm.ensure(nm.size() * 4);
        HIPX(hipMemcpyAsync(c->d_pnorm.p, nm.data(), nm.size() * 4, hipMemcpyHostToDevice, c->st));
        const uint32_t P = c->prm.n_parts;
        c->d_px.ensure((size_t)P * 256 * 4); c->d_py.ensure((size_t)P * 2 * 4); c->d_lcaps.ensure((size_t)P * 4);
        if (!c->h_lsum) HIPX(hipHostMalloc((void**)&c->h_lsum, 8, hipHostMallocDefault));
        if (!c->ev_lsum) HIPX(hipEventCreateWithFlags(&c->ev_lsum, hipEventDisableTiming));
        HIPX(hipStreamSynchronize(c->st));
        c->mlp = mlp;
        return 0;
    });
}

int gkc_part_kmers(gkc_ctx* c, uint64_t* out) {
    return guarded(c, [&] {
        for (uint32_t q = 0; q < c->prm.n_parts; q++) out[q] = c->h_kmer_cur[q];
        return 0;
    });
}

}  // extern "C"

// ---------------------------------------------------------------- counting impl

template <typename K>
static void ensure_count_buffers(gkc_ctx* c, uint64_t max_keys, uint32_t n_segs) {
    const size_t kw = sizeof(K);
    c->keysA.ensure((max_keys + 1) * kw);
    c->keysB.ensure((max_keys + 1) * kw);
    const uint64_t tiles = max_keys / (SORT_NT * (kw == 8 ? 8 : 4)) + n_segs + 1;   // smallest tile variant
    c->lookback.ensure(tiles * RADIX * 4);
    c->d_tdesc.ensure(tiles * sizeof(TileDesc));
    c->d_plan.ensure(sizeof(Plan));
    c->d_keystart.ensure((n_segs + 1) * 8);
    c->d_itemstart.ensure((n_segs + 1) * 4);
    c->d_tilestart.ensure((n_segs + 1) * 4);
    c->d_rlstart.ensure((n_segs + 1) * 4);
    c->d_hist.ensure((size_t)n_segs * MAX_PASS * RADIX * 4);
    c->d_tstart.ensure(n_segs * 8 + 8);
    c->d_tcap.ensure(n_segs * 4 + 4);
    c->d_tcap_total.ensure(8);
    c->d_nent.ensure(n_segs * 4 + 4);
    c->d_ndist.ensure(n_segs * 4 + 4);
    c->d_toverflow.ensure(64);
}

static void zero(gkc_ctx* c, void* p, uint64_t n32) {
    if (!n32) return;
    int blocks = (int)std::min<uint64_t>((n32 + 255) / 256, 4096);
    hipLaunchKernelGGL(k_zero32, dim3(blocks), dim3(256), 0, c->st, static_cast<uint32_t*>(p), n32);
    HIPX(hipGetLastError());
}

// Run extraction + sort for a plan whose per-segment arrays are set; the
// sorted keys end in keysA or keysB (returned).  Host knows n_keys_total and
// the sort tile count upper bound.
// onesweep variants (keys per thread x threads per workgroup x digit bits),
// selected by gkc_tuning.sort_items, sort_waves and digit_bits (the last in
// the binned form only: the full LSD sort keeps 8-bit digits)
template <typename K>
using OnesweepFn = void (*)(Plan*, SegArrays, WorkLists, const TileDesc*, const K*, K*, uint32_t*, int);
template <typename K> static OnesweepFn<K> onesweep_variant(int items, int nt, int db);
template <> OnesweepFn<uint64_t> onesweep_variant<uint64_t>(int items, int nt, int db) {
    if (db == 7) return items == 16 ? k_onesweep<uint64_t, 16, 256, 7> : k_onesweep<uint64_t, 20, 256, 7>;
    if (db == 6) return items == 16 ? k_onesweep<uint64_t, 16, 256, 6> : k_onesweep<uint64_t, 20, 256, 6>;
    if (nt == 512) {
        switch (items) {
            case 8: return k_onesweep<uint64_t, 8, 512, 8>;
            case 12: return k_onesweep<uint64_t, 12, 512, 8>;
            case 16: return k_onesweep<uint64_t, 16, 512, 8>;
            case 18: return k_onesweep<uint64_t, 18, 512, 8>;
            default: return k_onesweep<uint64_t, 10, 512, 8>;
        }
    }
    switch (items) {
        case 8: return k_onesweep<uint64_t, 8, 256, 8>;
        case 12: return k_onesweep<uint64_t, 12, 256, 8>;
        case 20: return k_onesweep<uint64_t, 20, 256, 8>;
        case 24: return k_onesweep<uint64_t, 24, 256, 8>;
        default: return k_onesweep<uint64_t, 16, 256, 8>;
    }
}
template <> OnesweepFn<u128k> onesweep_variant<u128k>(int items, int, int db) {
    if (db == 7) return items == 12 ? k_onesweep<u128k, 12, 256, 7> : k_onesweep<u128k, 8, 256, 7>;
    if (db == 6) return items == 12 ? k_onesweep<u128k, 12, 256, 6> : k_onesweep<u128k, 8, 256, 6>;
    switch (items) {
        case 4: return k_onesweep<u128k, 4, 256, 8>;
        case 12: return k_onesweep<u128k, 12, 256, 8>;
        default: return k_onesweep<u128k, 8, 256, 8>;
    }
}
// threads per onesweep workgroup: 8 waves on request for 64-bit keys and
// 8-bit digits, else 4
template <typename K> static int sort_nt(gkc_ctx* c, int db) {
    return sizeof(K) == 8 && db == 8 && c->tune.sort_waves == 8 ? 512 : 256;
}
template <typename K> static int sort_items(gkc_ctx* c, int db) {
    const int it = c->tune.sort_items;
    if (db != 8) return sizeof(K) == 8 ? (it == 16 ? 16 : 20) : (it == 12 ? 12 : 8);
    if (sort_nt<K>(c, db) == 512) return (it == 8 || it == 10 || it == 12 || it == 16 || it == 18) ? it : 10;
    if (sizeof(K) == 8) return (it == 8 || it == 12 || it == 16 || it == 20 || it == 24) ? it : 16;
    return (it == 4 || it == 8 || it == 12) ? it : 8;
}
template <typename K> static int sort_tile_keys(gkc_ctx* c, int db) { return sort_nt<K>(c, db) * sort_items<K>(c, db); }
template <typename K> static int sort_grid(gkc_ctx* c, int db) {
    OnesweepFn<K> fn = onesweep_variant<K>(sort_items<K>(c, db), sort_nt<K>(c, db), db);
    int g = occupancy_blocks((const void*)fn, sort_nt<K>(c, db), c->n_cu);
    if (c->tune.sort_blocks_per_cu > 0) g = c->tune.sort_blocks_per_cu * c->n_cu;   // explicit (may exceed residency)
    return std::max(g, 1);
}
// Work lists of one counting launch: n_lists lists (8 = one per XCD in the
// batched form), segments assigned longest-first to the list with the fewest
// sort tiles so far, each list's segments in increasing id order.
static void set_lists(gkc_ctx* c, uint32_t n_segs, const uint32_t* keys, const uint32_t* hdrs, uint32_t n_lists,
                      int sort_tile_keys, int rl_tile_keys = RL_TILE) {
    const uint32_t n = n_segs;
    std::vector<uint32_t> st(n), rt(n), it(n), owner(n);
    for (uint32_t s = 0; s < n; s++) {
        st[s] = (uint32_t)(((uint64_t)keys[s] + sort_tile_keys - 1) / sort_tile_keys);
        rt[s] = (uint32_t)(((uint64_t)keys[s] + rl_tile_keys - 1) / rl_tile_keys);
        it[s] = (hdrs[s] + EX_HDRS - 1) / EX_HDRS;
    }
    std::vector<uint32_t> byload(n);
    for (uint32_t s = 0; s < n; s++) byload[s] = s;
    std::stable_sort(byload.begin(), byload.end(), [&](uint32_t a, uint32_t b) { return st[a] > st[b]; });
    std::vector<uint64_t> load(n_lists, 0);
    for (uint32_t s : byload) {
        uint32_t best = 0;
        for (uint32_t L = 1; L < n_lists; L++) if (load[L] < load[best]) best = L;
        owner[s] = best;
        load[best] += st[s] + 1;
    }
    std::vector<uint32_t>& w = c->h_wl;
    w.assign(33 + 4 * (size_t)n + 9, 0);
    uint32_t pos = 0;
    for (uint32_t L = 0; L < n_lists; L++) {
        w[24 + L] = pos;
        uint64_t ts = 0, tr = 0, ti = 0;
        for (uint32_t s = 0; s < n; s++) {
            if (owner[s] != L) continue;
            w[33 + pos] = s;
            w[33 + n + pos] = (uint32_t)ts;
            w[33 + 2 * n + pos] = (uint32_t)tr;
            w[33 + 3 * n + pos] = (uint32_t)ti;
            ts += st[s]; tr += rt[s]; ti += it[s];
            pos++;
        }
        w[L] = (uint32_t)ts; w[8 + L] = (uint32_t)tr; w[16 + L] = (uint32_t)ti;
    }
    for (uint32_t L = n_lists; L <= 8; L++) w[24 + L] = pos;
    uint32_t tb = 0;                     // each list's first sort tile in the lists' concatenation
    for (uint32_t L = 0; L <= 8; L++) { w[33 + 4 * (size_t)n + L] = tb; if (L < 8) tb += w[L]; }
    c->d_wl.ensure(w.size() * 4);
    HIPX(hipMemcpyAsync(c->d_wl.p, w.data(), w.size() * 4, hipMemcpyHostToDevice, c->st));
    c->wl.w = c->d_wl.as<uint32_t>();
    c->wl.n_lists = n_lists;
    c->wl.n_segs = n;
}

template <typename K>
static void launch_onesweep(gkc_ctx* c, int grid, Plan* plan, SegArrays sa, const K* src, K* dst, int pass, int db) {
    OnesweepFn<K> fn = onesweep_variant<K>(sort_items<K>(c, db), sort_nt<K>(c, db), db);
    uint32_t* lb = c->lookback.as<uint32_t>();
    WorkLists wl = c->wl;
    const TileDesc* td = c->d_tdesc.as<TileDesc>();
    void* args[] = {&plan, &sa, &wl, &td, &src, &dst, &lb, &pass};
    HIPX(hipLaunchKernel((const void*)fn, dim3(grid), dim3(sort_nt<K>(c, db)), args, 0, c->st));
}

template <typename K>
static K* sort_segments(gkc_ctx* c, SegArrays sa, const uint64_t* hdr_base, const uint64_t* word_base,
                        uint32_t n_segs, uint64_t n_keys, uint64_t n_hdrs, float* ms_ex, float* ms_sort, float* ms_pass_sum) {
    Plan* plan = c->d_plan.as<Plan>();
    const int TILE = sort_tile_keys<K>(c, 8);
    {
        const uint32_t nk = (uint32_t)n_keys, nh = (uint32_t)n_hdrs;
        set_lists(c, 1, &nk, &nh, 1, TILE);
    }
    hipLaunchKernelGGL(k_plan_init, dim3(1), dim3(64), 0, c->st, plan, n_segs, (uint32_t)c->npass, 0u, (uint32_t)RADIX_BITS);
    hipLaunchKernelGGL(k_plan, dim3(1), dim3(1024), 0, c->st, plan, sa, TILE, RL_TILE);
    c->d_iseg.ensure((n_hdrs / EX_HDRS + n_segs + 1) * 4);
    hipLaunchKernelGGL(k_tile_desc, dim3(n_segs), dim3(256), 0, c->st, sa, c->wl, TILE, c->d_tdesc.as<TileDesc>(),
                       c->d_iseg.as<uint32_t>());
    HIPX(hipGetLastError());
    zero(c, sa.hist, (uint64_t)n_segs * MAX_PASS * RADIX);
    HIPX(hipEventRecord(c->ev[2], c->st));
    const uint64_t items = n_hdrs / EX_HDRS + n_segs;
    const size_t ex_lds = (size_t)c->npass * RADIX * sizeof(uint32_t);     // digit histograms
    int g_ex = (int)std::min<uint64_t>(items, occupancy_blocks((const void*)k_extract<K>, SORT_NT, c->n_cu, ex_lds));
    hipLaunchKernelGGL(k_extract<K>, dim3(std::max(g_ex, 1)), dim3(SORT_NT), ex_lds, c->st, plan, sa, c->d_iseg.as<uint32_t>(), c->hdrs.as<SkmHdr>(), hdr_base,
                       c->words.as<uint32_t>(), word_base, c->keysA.as<K>(), c->prm.k);
    HIPX(hipGetLastError());
    hipLaunchKernelGGL(k_hist_scan, dim3(std::max<uint32_t>(1, (n_segs * MAX_PASS + 3) / 4)), dim3(256), 0, c->st, plan, sa.hist);
    HIPX(hipGetLastError());
    HIPX(hipEventRecord(c->ev[3], c->st));
    const uint64_t tiles = n_keys / TILE + n_segs;
    int g_s = (int)std::min<uint64_t>(tiles, sort_grid<K>(c, 8));
    K* src = c->keysA.as<K>();
    K* dst = c->keysB.as<K>();
    for (int ps = 0; ps < c->npass; ps++) {
        zero(c, c->lookback.p, tiles * RADIX);
        HIPX(hipEventRecord(c->evp[2 * ps], c->st));
        launch_onesweep<K>(c, std::max(g_s, 1), plan, sa, src, dst, ps, 8);
        HIPX(hipGetLastError());
        HIPX(hipEventRecord(c->evp[2 * ps + 1], c->st));
        std::swap(src, dst);
    }
    HIPX(hipEventRecord(c->ev[6], c->st));
    HIPX(hipEventSynchronize(c->ev[6]));
    float pass_sum = 0;
    for (int ps = 0; ps < c->npass; ps++) {
        float m = 0;
        HIPX(hipEventElapsedTime(&m, c->evp[2 * ps], c->evp[2 * ps + 1]));
        pass_sum += m;
    }
    float a = 0, b = 0;
    HIPX(hipEventElapsedTime(&a, c->ev[2], c->ev[3]));
    HIPX(hipEventElapsedTime(&b, c->ev[3], c->ev[6]));
    if (ms_ex) *ms_ex += a;
    if (ms_sort) *ms_sort += b;
    if (ms_pass_sum) *ms_pass_sum += pass_sum;
    return src;
}

template <typename K>
static void table_reserve(gkc_ctx* c, uint64_t n) {
    c->tkeys.ensure((n + 1) * sizeof(K));
    c->tcnts.ensure((n + 1) * 4);
}

// Count one segment whose keys are sorted in `sorted`: run-length count pass,
// host sync, table growth, emit pass at table offset `out_off`.
template <typename K>
static void table_grow_keep(gkc_ctx* c, uint64_t used, uint64_t need) {
    if (need + 1 <= c->tkeys.bytes / sizeof(K) && need + 1 <= c->tcnts.bytes / 4) return;
    DevBuf nk, nc;
    uint64_t ncap = need * 3 / 2 + 1024;
    nk.ensure((ncap + 1) * sizeof(K)); nc.ensure((ncap + 1) * 4);
    if (used) {
        HIPX(hipMemcpyAsync(nk.p, c->tkeys.p, used * sizeof(K), hipMemcpyDeviceToDevice, c->st));
        HIPX(hipMemcpyAsync(nc.p, c->tcnts.p, used * 4, hipMemcpyDeviceToDevice, c->st));
    }
    HIPX(hipStreamSynchronize(c->st));
    std::swap(c->tkeys.p, nk.p); std::swap(c->tkeys.bytes, nk.bytes);
    std::swap(c->tcnts.p, nc.p); std::swap(c->tcnts.bytes, nc.bytes);
}

template <typename K>
static void runlength_one(gkc_ctx* c, SegArrays sa, const K* sorted, uint64_t nkeys, uint64_t out_off,
                          uint32_t* ne_out, uint32_t* nd_out, float* ms_rl) {
    HIPX(hipEventRecord(c->ev[2], c->st));
    TableOut to;
    to.keys = nullptr; to.counts = nullptr;
    to.table_start = c->d_tstart.as<uint64_t>(); to.table_cap = c->d_tcap.as<uint32_t>();
    to.n_entries = c->d_nent.as<uint32_t>(); to.n_distinct = c->d_ndist.as<uint32_t>();
    to.overflow = c->d_toverflow.as<uint32_t>();
    to.min_cnt = c->prm.min_count; to.max_cnt = c->prm.max_count; to.emit = 0;
    const uint64_t rl_tiles = nkeys / RL_TILE + 1;
    int g_r = (int)std::min<uint64_t>(rl_tiles, occupancy_blocks((const void*)k_runlength<K>, SORT_NT, c->n_cu));
    Plan* plan = c->d_plan.as<Plan>();
    zero(c, c->d_nent.p, 1); zero(c, c->d_ndist.p, 1); zero(c, c->d_toverflow.p, 1);
    zero(c, c->lookback.p, rl_tiles);
    zero(c, &plan->ctr[ctr_rl()], NLIST);
    hipLaunchKernelGGL(k_runlength<K>, dim3(g_r), dim3(SORT_NT), 0, c->st, plan, sa, c->wl, sorted, to, c->lookback.as<uint32_t>());
    HIPX(hipGetLastError());
    uint32_t ne = 0, nd = 0;
    HIPX(hipMemcpyAsync(&ne, c->d_nent.p, 4, hipMemcpyDeviceToHost, c->st));
    HIPX(hipMemcpyAsync(&nd, c->d_ndist.p, 4, hipMemcpyDeviceToHost, c->st));
    HIPX(hipStreamSynchronize(c->st));
    c->stats.host_syncs++;
    table_grow_keep<K>(c, out_off, out_off + ne);
    hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(1), 0, c->st, c->d_tstart.as<uint64_t>(), (uint64_t)out_off);
    hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(1), 0, c->st, c->d_tcap.as<uint32_t>(), ne);
    to.emit = 1; to.keys = c->tkeys.p; to.counts = c->tcnts.as<uint32_t>();
    zero(c, c->d_nent.p, 1); zero(c, c->d_ndist.p, 1);
    zero(c, c->lookback.p, rl_tiles);
    zero(c, &plan->ctr[ctr_rl()], NLIST);
    hipLaunchKernelGGL(k_runlength<K>, dim3(g_r), dim3(SORT_NT), 0, c->st, plan, sa, c->wl, sorted, to, c->lookback.as<uint32_t>());
    HIPX(hipGetLastError());
    HIPX(hipEventRecord(c->ev[3], c->st));
    HIPX(hipEventSynchronize(c->ev[3]));
    float m = 0;
    HIPX(hipEventElapsedTime(&m, c->ev[2], c->ev[3]));
    if (ms_rl) *ms_rl += m;
    *ne_out = ne; *nd_out = nd;
}

static SegArrays seg_arrays_single(gkc_ctx* c, const uint32_t* seg_keys, const uint32_t* seg_hdrs) {
    SegArrays sa;
    sa.seg_keys = seg_keys; sa.seg_hdrs = seg_hdrs;
    sa.key_start = c->d_keystart.as<uint64_t>();
    sa.item_start = c->d_itemstart.as<uint32_t>();
    sa.sort_tile_start = c->d_tilestart.as<uint32_t>();
    sa.rl_tile_start = c->d_rlstart.as<uint32_t>();
    sa.hist = c->d_hist.as<uint32_t>();
    return sa;
}

template <typename K>
static void count_impl(gkc_ctx* c, uint32_t flags) {
    const uint32_t P = c->prm.n_parts;
    (void)flags;
    c->part_off.assign(P + 1, 0);
    c->part_distinct.assign(P, 0);
    c->part_kmers.assign(P, 0);
    uint64_t max_keys = 0, tot_keys = 0, tot_hdrs = 0, tot_words = 0;
    for (uint32_t q = 0; q < P; q++) {
        c->part_kmers[q] = c->h_kmer_cur[q];
        max_keys = std::max<uint64_t>(max_keys, c->h_kmer_cur[q]);
        tot_keys += c->h_kmer_cur[q];
        tot_hdrs += c->h_hdr_cur[q];
        tot_words += c->h_word_cur[q];
    }
    c->stats = gkc_stats{c->stats.ms_partition};   // the count's figures; the last partition's time is kept
    c->stats.n_kmers = tot_keys; c->stats.n_superkmers = tot_hdrs; c->stats.n_words = tot_words;
    c->stats.n_pass = c->npass;
    HIPX(hipEventRecord(c->ev[0], c->st));
    // ---- plain: one partition at a time, count-then-allocate ----
    ensure_count_buffers<K>(c, max_keys, 1);
    table_grow_keep<K>(c, 0, std::max<uint64_t>(tot_keys / 4, 1024));
    float ms_ex = 0, ms_sort = 0, ms_pass = 0, ms_rl = 0;
    uint64_t out_off = 0;
    for (uint32_t q = 0; q < P; q++) {
        c->part_off[q] = out_off;
        if (c->h_kmer_cur[q] == 0) continue;
        SegArrays sa = seg_arrays_single(c, c->d_kmer_cur.as<uint32_t>() + q, c->d_hdr_cur.as<uint32_t>() + q);
        K* sorted = sort_segments<K>(c, sa, c->d_hdr_base.as<uint64_t>() + q, c->d_word_base.as<uint64_t>() + q, 1,
                                     c->h_kmer_cur[q], c->h_hdr_cur[q], &ms_ex, &ms_sort, &ms_pass);
        uint32_t ne = 0, nd = 0;
        runlength_one<K>(c, sa, sorted, c->h_kmer_cur[q], out_off, &ne, &nd, &ms_rl);
        out_off += ne;
        c->part_distinct[q] = nd;
    }
    c->part_off[P] = out_off;
    c->part_start.assign(c->part_off.begin(), c->part_off.end() - 1);
    HIPX(hipEventRecord(c->ev[1], c->st));
    HIPX(hipEventSynchronize(c->ev[1]));
    float total = 0;
    HIPX(hipEventElapsedTime(&total, c->ev[0], c->ev[1]));
    c->stats.ms_count = total;
    c->stats.ms_extract = ms_ex; c->stats.ms_sort = ms_sort; c->stats.ms_runlength = ms_rl;
    uint64_t launched = 0;
    for (uint32_t q = 0; q < P; q++) if (c->h_kmer_cur[q]) launched++;
    c->stats.ms_sort_pass_avg = ms_pass / std::max<uint64_t>(1, (uint64_t)c->npass * launched);
    uint64_t nd = 0;
    for (uint32_t q = 0; q < P; q++) nd += c->part_distinct[q];
    c->stats.n_distinct = nd; c->stats.n_entries = out_off;
}

extern "C" int gkc_count_packed(gkc_ctx* c, const uint32_t* words, uint64_t n_words, const uint32_t* lens, uint64_t n_skm,
                                uint64_t* key_lo, uint64_t* key_hi, uint32_t* cnt, uint64_t cap, uint64_t* n_entries,
                                uint64_t* n_distinct) {
    return guarded(c, [&]() -> int {
        const int k = c->prm.k;
        std::vector<SkmHdr> hd;
        hd.reserve(n_skm);
        uint64_t wo = 0, ko = 0;
        for (uint64_t i = 0; i < n_skm; i++) {
            const uint32_t L = lens[i];
            if (L >= (uint32_t)k) {
                SkmHdr h{(uint32_t)wo, (uint32_t)ko, L - k + 1, 0};
                hd.push_back(h);
                ko += L - k + 1;
            }
            wo += (L + 15) / 16;
        }
        if (wo > n_words) throw ArgError{"superkmer lengths exceed the word buffer", GKC_EINVAL};
        if (ko >= (1ull << 30)) throw ArgError{"partition exceeds 2^30 k-mers", GKC_EINVAL};
        *n_entries = 0; *n_distinct = 0;
        if (ko == 0) return 0;
        DevBuf dw, dh, dmeta;
        dw.ensure((n_words + 8) * 4);
        dh.ensure(hd.size() * sizeof(SkmHdr));
        dmeta.ensure(32);
        HIPX(hipMemsetAsync(dw.p, 0, (n_words + 8) * 4, c->st));
        HIPX(hipMemcpyAsync(dw.p, words, n_words * 4, hipMemcpyHostToDevice, c->st));
        HIPX(hipMemcpyAsync(dh.p, hd.data(), hd.size() * sizeof(SkmHdr), hipMemcpyHostToDevice, c->st));
        hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(1), 0, c->st, dmeta.as<uint32_t>(), (uint32_t)ko);
        hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(1), 0, c->st, dmeta.as<uint32_t>() + 1, (uint32_t)hd.size());
        hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(1), 0, c->st, reinterpret_cast<uint64_t*>(static_cast<char*>(dmeta.p) + 16), (uint64_t)0);
        const uint32_t* segk = dmeta.as<uint32_t>();
        const uint32_t* segh = dmeta.as<uint32_t>() + 1;
        const uint64_t* base0 = reinterpret_cast<const uint64_t*>(static_cast<char*>(dmeta.p) + 16);
        // point the bins at the temporary storage for this call
        std::swap(c->words.p, dw.p); std::swap(c->words.bytes, dw.bytes);
        std::swap(c->hdrs.p, dh.p); std::swap(c->hdrs.bytes, dh.bytes);
        struct Restore { gkc_ctx* c; DevBuf* dw; DevBuf* dh;
            ~Restore() { std::swap(c->words.p, dw->p); std::swap(c->words.bytes, dw->bytes);
                         std::swap(c->hdrs.p, dh->p); std::swap(c->hdrs.bytes, dh->bytes); } } restore{c, &dw, &dh};
        uint32_t ne = 0, nd = 0;
        auto run = [&](auto tag) {
            using K = decltype(tag);
            ensure_count_buffers<K>(c, ko, 1);
            SegArrays sa = seg_arrays_single(c, segk, segh);
            K* sorted = sort_segments<K>(c, sa, base0, base0, 1, ko, hd.size(), nullptr, nullptr, nullptr);
            table_grow_keep<K>(c, 0, 1024);
            runlength_one<K>(c, sa, sorted, ko, 0, &ne, &nd, nullptr);
            if (ne > cap) throw ArgError{"output buffer too small", GKC_ECAP};
            if (ne) {
                if (sizeof(K) == 8) {
                    HIPX(hipMemcpyAsync(key_lo, c->tkeys.p, (size_t)ne * 8, hipMemcpyDeviceToHost, c->st));
                } else {
                    std::vector<u128k> tmp(ne);
                    HIPX(hipMemcpyAsync(tmp.data(), c->tkeys.p, (size_t)ne * 16, hipMemcpyDeviceToHost, c->st));
                    HIPX(hipStreamSynchronize(c->st));
                    for (uint32_t i = 0; i < ne; i++) { key_lo[i] = tmp[i].lo; if (key_hi) key_hi[i] = tmp[i].hi; }
                }
                HIPX(hipMemcpyAsync(cnt, c->tcnts.p, (size_t)ne * 4, hipMemcpyDeviceToHost, c->st));
                HIPX(hipStreamSynchronize(c->st));
            }
        };
        if (c->keyw == 8) run(uint64_t{}); else run(u128k{});
        c->counted = false;     // the context table now holds this call's entries only
        *n_entries = ne; *n_distinct = nd;
        return 0;
    });
}

// ---------------------------------------------------------------- batched form

// Shape of one batched counting launch.  Binned (count_mode 0, k >= 8): up
// to three onesweep passes over the top 24 key bits group each segment's
// keys by 24-bit bin; k_wavecount counts WC_TILE-key windows of whole bins
// per wave, after k_bincount has counted the ranges larger than
// WC_CAP into staging.  Otherwise (count_mode 1, k < 8): the full LSD sort
// (ceil(2k/8) passes) and k_runlength.
struct BatchShape {
    bool binned;
    bool wave;
    int npass;
    uint32_t sh0;
    int rl_tile;
    int bin_bits;
    int dbits;      // digit width of the passes
};
template <typename K>
static BatchShape batch_shape(const gkc_ctx* c) {
    const int kb = 2 * c->prm.k;
    if (c->tune.count_mode == 1 || c->prm.k < 8) return BatchShape{false, false, c->npass, 0u, RL_TILE, 0, RADIX_BITS};
    const int db = c->tune.digit_bits > 0 ? c->tune.digit_bits : RADIX_BITS;
    const int np = std::min(c->tune.bin_passes > 0 ? c->tune.bin_passes : 3, (kb + db - 1) / db);
    const int win = c->tune.count_window > 0 ? c->tune.count_window : WC_TILE;
    return BatchShape{true, true, np, (uint32_t)std::max(0, kb - db * np), win, db * np, db};
}

// Enqueue extraction + the onesweep passes over every partition at once.
// Sizes are read by the kernels from the device plan, so the sequence can be
// captured once and replayed.  With `timed`, per-pass events are recorded.
template <typename K>
static K* enqueue_sort_all(gkc_ctx* c, SegArrays sa, const BatchShape& bs, uint64_t tiles_cap, bool timed) {
    const uint32_t P = c->prm.n_parts;
    Plan* plan = c->d_plan.as<Plan>();
    const int TILE = sort_tile_keys<K>(c, bs.dbits);
    hipLaunchKernelGGL(k_plan_init, dim3(1), dim3(64), 0, c->st, plan, P, (uint32_t)bs.npass, bs.sh0, (uint32_t)bs.dbits);
    hipLaunchKernelGGL(k_plan, dim3(1), dim3(1024), 0, c->st, plan, sa, TILE, bs.rl_tile);
    hipLaunchKernelGGL(k_tile_desc, dim3(P), dim3(256), 0, c->st, sa, c->wl, TILE, c->d_tdesc.as<TileDesc>(),
                       c->d_iseg.as<uint32_t>());
    zero(c, sa.hist, (uint64_t)P * MAX_PASS * RADIX);
    if (timed) HIPX(hipEventRecord(c->ev[2], c->st));
    const size_t ex_lds = (size_t)bs.npass * RADIX * sizeof(uint32_t);     // digit histograms
    const int g_ex = occupancy_blocks((const void*)k_extract<K>, SORT_NT, c->n_cu, ex_lds);
    hipLaunchKernelGGL(k_extract<K>, dim3(g_ex), dim3(SORT_NT), ex_lds, c->st, plan, sa, c->d_iseg.as<uint32_t>(), c->hdrs.as<SkmHdr>(),
                       c->d_hdr_base.as<uint64_t>(), c->words.as<uint32_t>(), c->d_word_base.as<uint64_t>(),
                       c->keysA.as<K>(), c->prm.k);
    hipLaunchKernelGGL(k_hist_scan, dim3((P * MAX_PASS + 3) / 4), dim3(256), 0, c->st, plan, sa.hist);
    if (timed) HIPX(hipEventRecord(c->ev[3], c->st));
    const int g_s = sort_grid<K>(c, bs.dbits);
    K* src = c->keysA.as<K>();
    K* dst = c->keysB.as<K>();
    for (int ps = 0; ps < bs.npass; ps++) {
        zero(c, c->lookback.p, tiles_cap * RADIX);
        if (timed) HIPX(hipEventRecord(c->evp[2 * ps], c->st));
        launch_onesweep<K>(c, g_s, plan, sa, src, dst, ps, bs.dbits);
        if (timed) HIPX(hipEventRecord(c->evp[2 * ps + 1], c->st));
        std::swap(src, dst);
    }
    if (timed) HIPX(hipEventRecord(c->ev[6], c->st));
    HIPX(hipGetLastError());
    return src;
}

// Count (emit = 0) or count and write the table (emit = 1) from the keys the
// sort left in `sorted`.  d_toverflow[0]: table capacity exceeded;
// d_toverflow[1], [2]: count and claim counter of the heavy windows
// k_bincount counts ahead into staging (in the LSD form: ranges too large
// for the run-length kernel), [3] [4] and [5] [6]: the same for the two
// k_midcount lists.
// Wave-level count: `prep` (the first pass over the sorted keys of a
// gkc_count) counts every window into staging and scans the windows' kept
// counts into table offsets; an emit pass then only copies staging to the
// table (again with new capacities after an under-prediction).
template <typename K>
static void enqueue_runlength_all(gkc_ctx* c, SegArrays sa, const BatchShape& bs, K* sorted,
                                  uint64_t rl_tiles_cap, int emit, bool prep) {
    const uint32_t P = c->prm.n_parts;
    Plan* plan = c->d_plan.as<Plan>();
    TableOut to;
    to.keys = emit ? c->tkeys.p : nullptr; to.counts = emit ? c->tcnts.as<uint32_t>() : nullptr;
    to.table_start = c->d_tstart.as<uint64_t>(); to.table_cap = c->d_tcap.as<uint32_t>();
    to.n_entries = c->d_nent.as<uint32_t>(); to.n_distinct = c->d_ndist.as<uint32_t>();
    to.overflow = c->d_toverflow.as<uint32_t>();
    to.min_cnt = c->prm.min_count; to.max_cnt = c->prm.max_count; to.emit = emit;
    K* alt = sorted == c->keysA.as<K>() ? c->keysB.as<K>() : c->keysA.as<K>();
    // d_toverflow: [0] table overflow, [1] [2] k_bincount's list count and claim counter, [3] [4]
    // and [5] [6] the two k_midcount lists' (64-bit keys; the second and third thirds of d_hlist
    // hold those lists), [7] medium windows k_wavecount took, [8] [9] the second k_bincount
    // pass's list count and claim counter (its list: the fourth quarter of d_hlist)
    uint32_t* ctrs = c->d_toverflow.as<uint32_t>();
    uint32_t* hlist = c->d_hlist.as<uint32_t>();
    const bool mid = sizeof(K) == 8;
    HeavyList hl{hlist, ctrs + 1, ctrs + 2, c->d_hres.as<uint2>(), c->d_wkept.as<uint32_t>(),
                 {mid ? hlist + rl_tiles_cap : nullptr, mid ? hlist + 2 * rl_tiles_cap : nullptr},
                 {ctrs + 3, ctrs + 5}, {ctrs + 4, ctrs + 6}, mid ? ctrs + 7 : nullptr,
                 mid ? hlist + 3 * rl_tiles_cap : nullptr, ctrs + 8, ctrs + 9};
    if (bs.wave) {
        {
            if (prep) {
                zero(c, c->d_nent.p, P); zero(c, c->d_ndist.p, P);
                zero(c, ctrs + 1, 9);
                zero(c, &plan->ctr[ctr_rl()], NLIST);
                const int g_r = occupancy_blocks((const void*)k_binranges<K>, 256, c->n_cu);
                hipLaunchKernelGGL(k_binranges<K>, dim3(g_r), dim3(256), 0, c->st, plan, sa, (const K*)sorted,
                                   c->d_wdesc.as<uint4>(), bs.rl_tile, (int)bs.sh0, bs.bin_bits, (uint32_t)WC_CAP, hl);
                const int g_h = occupancy_blocks((const void*)k_bincount<K>, BC_NT, c->n_cu);
                hipLaunchKernelGGL(k_bincount<K>, dim3(g_h), dim3(BC_NT), 0, c->st, plan, sa, sorted, alt,
                                   (const uint4*)c->d_wdesc.as<uint4>(), to, hl, 0);
                const int g_w = occupancy_blocks((const void*)k_wavecount<K>, WC_NT, c->n_cu);
                hipLaunchKernelGGL(k_wavecount<K>, dim3(g_w), dim3(WC_NT), 0, c->st, plan, sa, sorted, alt,
                                   (const uint4*)c->d_wdesc.as<uint4>(), hl, to);
                if constexpr (sizeof(K) == 8) {     // the medium windows k_wavecount handed on
                    const int g_m0 = occupancy_blocks((const void*)k_midcount<0>, GKC_WAVE, c->n_cu);
                    hipLaunchKernelGGL(k_midcount<0>, dim3(g_m0), dim3(GKC_WAVE), 0, c->st, plan, sa, sorted, alt,
                                       (const uint4*)c->d_wdesc.as<uint4>(), to, hl);
                    const int g_m1 = occupancy_blocks((const void*)k_midcount<1>, GKC_WAVE, c->n_cu);
                    hipLaunchKernelGGL(k_midcount<1>, dim3(g_m1), dim3(GKC_WAVE), 0, c->st, plan, sa, sorted, alt,
                                       (const uint4*)c->d_wdesc.as<uint4>(), to, hl);
                    hipLaunchKernelGGL(k_bincount<K>, dim3(g_h), dim3(BC_NT), 0, c->st, plan, sa, sorted, alt,
                                       (const uint4*)c->d_wdesc.as<uint4>(), to, hl, 1);
                }
                hipLaunchKernelGGL(k_tilescan, dim3(std::min<uint32_t>(P, 4 * c->n_cu)), dim3(1024), 0, c->st, plan, sa,
                                   hl.kept);
            }
            if (emit && prep)       // presized: a dense layout from the exact counts, only the total must fit
                hipLaunchKernelGGL(k_table_layout, dim3(1), dim3(1024), 0, c->st, (const uint32_t*)c->d_nent.as<uint32_t>(),
                                   P, (const uint64_t*)c->d_tcap_total.as<uint64_t>(), c->d_tstart.as<uint64_t>(),
                                   c->d_tcap.as<uint32_t>());
            if (emit) {
                zero(c, c->d_toverflow.p, 1);
                const int g_c = occupancy_blocks((const void*)k_wavecopy<K>, 256, c->n_cu);
                hipLaunchKernelGGL(k_wavecopy<K>, dim3(g_c), dim3(256), 0, c->st, plan, sa, (const K*)sorted,
                                   (const K*)alt, (const uint4*)c->d_wdesc.as<uint4>(), hl, to);
            }
        }
        HIPX(hipGetLastError());
        return;
    }
    zero(c, c->d_nent.p, P); zero(c, c->d_ndist.p, P); zero(c, c->d_toverflow.p, 3);
    zero(c, c->lookback.p, rl_tiles_cap);
    zero(c, &plan->ctr[ctr_rl()], NLIST);
    {
        const int g_r = occupancy_blocks((const void*)k_runlength<K>, SORT_NT, c->n_cu);
        hipLaunchKernelGGL(k_runlength<K>, dim3(g_r), dim3(SORT_NT), 0, c->st, plan, sa, c->wl, sorted, to,
                           c->lookback.as<uint32_t>());
    }
    HIPX(hipGetLastError());
}

template <typename F>
hipGraphExec_t capture(gkc_ctx* c, F&& enqueue) {
    hipGraph_t g;
    HIPX(hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
    try {
        enqueue();
    } catch (...) {
        hipGraph_t dead;
        (void)hipStreamEndCapture(c->st, &dead);
        if (dead) (void)hipGraphDestroy(dead);
        throw;
    }
    HIPX(hipStreamEndCapture(c->st, &g));
    hipGraphExec_t ex;
    HIPX(hipGraphInstantiate(&ex, g, nullptr, nullptr, 0));
    HIPX(hipGraphDestroy(g));
    return ex;
}

template <typename K>
static void count_batched(gkc_ctx* c, uint32_t flags) {
    const uint32_t P = c->prm.n_parts;
    const bool timed = (flags & GKC_SYNC_ONLY) == 0 && getenv("GKC_NO_GRAPH") != nullptr;
    const BatchShape bs = batch_shape<K>(c);
    c->part_off.assign(P + 1, 0);
    c->part_distinct.assign(P, 0);
    c->part_kmers.assign(P, 0);
    uint64_t tot_keys = 0, tot_hdrs = 0, tot_words = 0;
    for (uint32_t q = 0; q < P; q++) {
        c->part_kmers[q] = c->h_kmer_cur[q];
        tot_keys += c->h_kmer_cur[q]; tot_hdrs += c->h_hdr_cur[q]; tot_words += c->h_word_cur[q];
    }
    c->stats = gkc_stats{c->stats.ms_partition};   // the count's figures; the last partition's time is kept
    c->stats.n_kmers = tot_keys; c->stats.n_superkmers = tot_hdrs; c->stats.n_words = tot_words;
    c->stats.n_pass = bs.npass;
    c->stats.binned = bs.binned ? 1 : 0;
    ensure_count_buffers<K>(c, tot_keys, P);
    c->d_iseg.ensure((tot_hdrs / EX_HDRS + P + 1) * 4);
    const uint64_t tiles_cap = tot_keys / (sort_tile_keys<K>(c, bs.dbits)) + P + 1;
    const uint64_t rl_tiles_cap = tot_keys / bs.rl_tile + P + 1;
    SegArrays sa;
    sa.seg_keys = c->d_kmer_cur.as<uint32_t>(); sa.seg_hdrs = c->d_hdr_cur.as<uint32_t>();
    sa.key_start = c->d_keystart.as<uint64_t>(); sa.item_start = c->d_itemstart.as<uint32_t>();
    sa.sort_tile_start = c->d_tilestart.as<uint32_t>(); sa.rl_tile_start = c->d_rlstart.as<uint32_t>();
    sa.hist = c->d_hist.as<uint32_t>();
    K* sorted = (bs.npass % 2 == 0) ? c->keysA.as<K>() : c->keysB.as<K>();
    if (bs.wave) c->d_wdesc.ensure(rl_tiles_cap * sizeof(uint4));
    c->d_hlist.ensure(4 * rl_tiles_cap * sizeof(uint32_t));
    c->d_hres.ensure(rl_tiles_cap * sizeof(uint2));
    c->d_wkept.ensure(rl_tiles_cap * sizeof(uint32_t));
    set_lists(c, P, c->h_kmer_cur.data(), c->h_hdr_cur.data(), c->tune.single_list == 1 ? 1 : NLIST,
              sort_tile_keys<K>(c, bs.dbits), bs.rl_tile);

    // table capacities: predicted on the device (learned), given by the caller
    // (presized) or from a count pass (count-then-allocate)
    const bool learned = (flags & GKC_PRESIZED) && (flags & GKC_LEARNED) && c->mlp;
    if ((flags & GKC_LEARNED) && !c->mlp) throw ArgError{"GKC_LEARNED without gkc_set_predictor", GKC_ESTATE};
    const bool presized = learned || ((flags & GKC_PRESIZED) && c->presize_caps.size() == P);
    (void)occupancy_blocks((const void*)k_extract<K>, SORT_NT, c->n_cu, (size_t)bs.npass * RADIX * sizeof(uint32_t));
    (void)sort_grid<K>(c, bs.dbits);
    (void)occupancy_blocks((const void*)k_runlength<K>, SORT_NT, c->n_cu);
    (void)occupancy_blocks((const void*)k_binranges<K>, 256, c->n_cu);
    (void)occupancy_blocks((const void*)k_bincount<K>, BC_NT, c->n_cu);
    (void)occupancy_blocks((const void*)k_wavecount<K>, WC_NT, c->n_cu);
    (void)occupancy_blocks((const void*)k_midcount<0>, GKC_WAVE, c->n_cu);
    (void)occupancy_blocks((const void*)k_midcount<1>, GKC_WAVE, c->n_cu);
    (void)occupancy_blocks((const void*)k_wavecopy<K>, 256, c->n_cu);
    HIPX(hipEventRecord(c->ev[0], c->st));
    if (learned) {      // capacities on the device, ahead of the sort; only their sum comes back, while it runs
        if (gkc_mlp_presize_device(c->mlp, c->d_sketch.as<uint32_t>(), c->d_kmer_cur.as<uint32_t>(), P,
                                   c->d_pnorm.as<float>(), c->d_px.as<float>(), c->d_py.as<float>(),
                                   c->d_lcaps.as<uint32_t>(), c->d_tcap_total.as<uint64_t>(), c->st) != 0)
            throw ArgError{std::string("learned presizing: ") + gkc_mlp_last_error(), GKC_EHIP};
        HIPX(hipMemcpyAsync(c->h_lsum, c->d_tcap_total.p, 8, hipMemcpyDeviceToHost, c->st));
        HIPX(hipEventRecord(c->ev_lsum, c->st));
    }
    std::vector<const void*> key = {c->keysA.p, c->keysB.p, c->lookback.p, c->d_tdesc.p, c->d_iseg.p, c->words.p, c->hdrs.p,
                                    c->d_wdesc.p,
                                    c->d_hlist.p, c->d_hres.p, c->d_wkept.p,
                                    (const void*)(uintptr_t)tiles_cap, (const void*)(uintptr_t)bs.npass,
                                    (const void*)(uintptr_t)sort_items<K>(c, bs.dbits), (const void*)(uintptr_t)sort_grid<K>(c, bs.dbits),
                                    (const void*)(uintptr_t)sort_nt<K>(c, bs.dbits), (const void*)(uintptr_t)bs.dbits,
                                    c->d_wl.p, (const void*)(uintptr_t)c->wl.n_lists,
                                    (const void*)(uintptr_t)(bs.binned ? 1 : 0), (const void*)(uintptr_t)(bs.wave ? 1 : 0),
                                    (const void*)(uintptr_t)bs.rl_tile};
    if (timed) {
        enqueue_sort_all<K>(c, sa, bs, tiles_cap, true);
    } else {
        if (!c->g_sort || c->g_key != key) {
            if (c->g_sort) { HIPX(hipGraphExecDestroy(c->g_sort)); c->g_sort = nullptr; }
            if (c->g_cnt) { HIPX(hipGraphExecDestroy(c->g_cnt)); c->g_cnt = nullptr; }
            if (c->g_emit) { HIPX(hipGraphExecDestroy(c->g_emit)); c->g_emit = nullptr; }
            if (c->g_emit_p) { HIPX(hipGraphExecDestroy(c->g_emit_p)); c->g_emit_p = nullptr; }
            c->g_sort = capture(c, [&] { enqueue_sort_all<K>(c, sa, bs, tiles_cap, false); });
            c->g_cnt = capture(c, [&] { enqueue_runlength_all<K>(c, sa, bs, sorted, rl_tiles_cap, 0, true); });
            c->g_key = key;
        }
        HIPX(hipGraphLaunch(c->g_sort, c->st));
    }
    HIPX(hipEventRecord(c->ev[4], c->st));
    std::vector<uint32_t>& caps = c->h_caps;
    caps.assign(P, 0);
    uint32_t flags_dev[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (learned) {
        HIPX(hipEventSynchronize(c->ev_lsum));      // long done: the sort graph is running behind it
        if (!bs.wave) {                             // per-partition capacities for the run-length count
            HIPX(hipMemcpyAsync(caps.data(), c->d_lcaps.p, P * 4, hipMemcpyDeviceToHost, c->st));
            HIPX(hipStreamSynchronize(c->st));
            c->stats.host_syncs++;
        }
    } else if (presized) {
        caps = c->presize_caps;
    } else {
        if (timed) enqueue_runlength_all<K>(c, sa, bs, sorted, rl_tiles_cap, 0, true);
        else HIPX(hipGraphLaunch(c->g_cnt, c->st));
        HIPX(hipMemcpyAsync(caps.data(), c->d_nent.p, P * 4, hipMemcpyDeviceToHost, c->st));
        HIPX(hipMemcpyAsync(flags_dev, c->d_toverflow.p, 32, hipMemcpyDeviceToHost, c->st));
        HIPX(hipStreamSynchronize(c->st));
        c->stats.host_syncs++;
    }
    // table placement (physical starts) and capacity.  Presized binned
    // count: the starts come from the exact counts on the device (dense
    // table), the predicted capacities only size the whole table; otherwise
    // each partition gets its own capacity
    const bool dense = presized && bs.wave;
    c->h_tstart.assign(P, 0);
    uint64_t tot_cap = 0;
    for (uint32_t q = 0; q < P; q++) { c->h_tstart[q] = tot_cap; tot_cap += caps[q]; }
    if (learned && dense) tot_cap = *c->h_lsum;     // the device holds the sum (and the layout reads it there)
    c->tkeys.ensure((tot_cap + 1) * sizeof(K));
    c->tcnts.ensure((tot_cap + 1) * 4);
    if (learned && dense) {
    } else if (dense) {
        c->h_tcap_total = tot_cap;
        HIPX(hipMemcpyAsync(c->d_tcap_total.p, &c->h_tcap_total, 8, hipMemcpyHostToDevice, c->st));
    } else {
        HIPX(hipMemcpyAsync(c->d_tstart.p, c->h_tstart.data(), P * 8, hipMemcpyHostToDevice, c->st));
        HIPX(hipMemcpyAsync(c->d_tcap.p, caps.data(), P * 4, hipMemcpyHostToDevice, c->st));
    }
    // emit pass: replay the captured graph only if it writes the current table
    // buffers (either may have moved when the table grew).  The first pass
    // over the sorted keys of this count prepares (see enqueue_runlength_all).
    bool prepared = !presized;
    auto emit_pass = [&]() {
        const bool prep = !prepared;
        prepared = true;
        if (timed) { enqueue_runlength_all<K>(c, sa, bs, sorted, rl_tiles_cap, 1, prep); return; }
        hipGraphExec_t& g = prep ? c->g_emit_p : c->g_emit;
        const void*& gk = prep ? c->g_emit_p_tk : c->g_emit_tk;
        const void*& gc = prep ? c->g_emit_p_tc : c->g_emit_tc;
        if (!g || gk != c->tkeys.p || gc != c->tcnts.p) {
            if (g) { HIPX(hipGraphExecDestroy(g)); g = nullptr; }
            g = capture(c, [&] { enqueue_runlength_all<K>(c, sa, bs, sorted, rl_tiles_cap, 1, prep); });
            gk = c->tkeys.p; gc = c->tcnts.p;
        }
        HIPX(hipGraphLaunch(g, c->st));
    };
    emit_pass();
    HIPX(hipEventRecord(c->ev[5], c->st));
    std::vector<uint32_t>& nent = c->h_nent;
    nent.assign(P, 0);
    std::vector<uint32_t> ndist(P);
    HIPX(hipMemcpyAsync(nent.data(), c->d_nent.p, P * 4, hipMemcpyDeviceToHost, c->st));
    HIPX(hipMemcpyAsync(ndist.data(), c->d_ndist.p, P * 4, hipMemcpyDeviceToHost, c->st));
    HIPX(hipMemcpyAsync(flags_dev, c->d_toverflow.p, 32, hipMemcpyDeviceToHost, c->st));
    HIPX(hipStreamSynchronize(c->st));
    c->stats.host_syncs++;
    c->stats.heavy_ranges = flags_dev[1] + flags_dev[7];
    if (flags_dev[0]) {
        // under-prediction caught on device: count-then-allocate emission with
        // the exact per-partition entry counts the emit pass just measured
        c->stats.fallbacks++;
        caps = nent;
        tot_cap = 0;
        for (uint32_t q = 0; q < P; q++) { c->h_tstart[q] = tot_cap; tot_cap += caps[q]; }
        c->tkeys.ensure((tot_cap + 1) * sizeof(K));
        c->tcnts.ensure((tot_cap + 1) * 4);
        HIPX(hipMemcpyAsync(c->d_tstart.p, c->h_tstart.data(), P * 8, hipMemcpyHostToDevice, c->st));
        HIPX(hipMemcpyAsync(c->d_tcap.p, caps.data(), P * 4, hipMemcpyHostToDevice, c->st));
        emit_pass();
        HIPX(hipMemcpyAsync(nent.data(), c->d_nent.p, P * 4, hipMemcpyDeviceToHost, c->st));
        HIPX(hipMemcpyAsync(flags_dev, c->d_toverflow.p, 32, hipMemcpyDeviceToHost, c->st));
        HIPX(hipStreamSynchronize(c->st));
        c->stats.host_syncs++;
        if (flags_dev[0]) throw ArgError{"internal: table overflow after exact sizing", GKC_ESTATE};
    } else if (dense) {
        uint64_t a = 0;
        for (uint32_t q = 0; q < P; q++) { c->h_tstart[q] = a; a += nent[q]; }
    }
    HIPX(hipEventRecord(c->ev[1], c->st));
    HIPX(hipEventSynchronize(c->ev[1]));
    c->part_start = c->h_tstart;
    uint64_t acc = 0, nd = 0;
    for (uint32_t q = 0; q < P; q++) {
        c->part_off[q] = acc; acc += nent[q];
        c->part_distinct[q] = ndist[q]; nd += ndist[q];
    }
    c->part_off[P] = acc;
    float total = 0, s1 = 0, s2 = 0;
    HIPX(hipEventElapsedTime(&total, c->ev[0], c->ev[1]));
    HIPX(hipEventElapsedTime(&s1, c->ev[0], c->ev[4]));
    HIPX(hipEventElapsedTime(&s2, c->ev[4], c->ev[5]));
    c->stats.ms_count = total;
    c->stats.ms_runlength = s2;
    if (timed) {
        float a = 0, b = 0, pass_sum = 0;
        HIPX(hipEventElapsedTime(&a, c->ev[2], c->ev[3]));
        HIPX(hipEventElapsedTime(&b, c->ev[3], c->ev[6]));
        for (int ps = 0; ps < bs.npass; ps++) {
            float m = 0;
            HIPX(hipEventElapsedTime(&m, c->evp[2 * ps], c->evp[2 * ps + 1]));
            pass_sum += m;
        }
        c->stats.ms_extract = a; c->stats.ms_sort = b;
        c->stats.ms_sort_pass_avg = pass_sum / bs.npass;
    } else {
        c->stats.ms_sort = s1;
    }
    c->stats.n_distinct = nd; c->stats.n_entries = acc;
}
