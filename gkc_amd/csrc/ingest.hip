// On-device read ingest: FASTQ or single-line FASTA text in HBM to the read
// CSR that gkc_partition takes (bases back to back, n + 1 absolute offsets).
//
// Reference semantics followed (file:line in the reference tree):
//   FASTQ records are four lines: '@' header, sequence, '+', quality; the
//   read is the sequence line              read_loader.hpp:44-121
//   FASTA records are a '>' header and one sequence line; the last line may
//   lack its newline                       read_loader.hpp:123-196
//   a '\r' before a sequence line's newline is not part of the read
//                                          read_loader.hpp:103, 177
// The text must be whole records (the reference's loader asserts the record
// starts and hands partial records between its threads); anything else is
// reported as malformed rather than guessed at.
//
// Steps, every one over all the text at once: newlines counted per 4 KiB
// block and the counts scanned; each newline ranked and placed in its record
// (sequence start and end, header / separator checks); the sequence lengths
// scanned into offsets; the sequence bytes gathered by output chunks.
#include <algorithm>
#include <mutex>
#include <string>

#include "gkc_hip.h"
#include "kernels.hpp"

namespace gkc {

constexpr int IN_NT = 256;                     // threads per block
// u64 words of gkc_parse_reads_noalloc's fixed scratch (GKC_PARSE_NOALLOC_SCRATCH_BYTES)
constexpr uint64_t NOALLOC_SCRATCH_WORDS = GKC_PARSE_NOALLOC_SCRATCH_BYTES / 8;
constexpr uint64_t IN_BLK = IN_NT * 16;        // bytes of the aligned view per block (16 per thread)

// The text viewed from the 16-byte boundary at or below it: text byte q is
// base[lead + q].  A thread's 16 bytes are base[a, a + 16), a = 16 * (thread).
struct TextView {
    const uint8_t* base;
    uint64_t lead, n;
    __device__ uint8_t at(uint64_t q) const { return base[lead + q]; }
};

// The view's bytes [a, a + 16) as four little-endian words: loaded whole
// where they end inside the text, byte by byte (0 past the text's end) at the
// text's end, not at all (0) where they hold no text byte.
__device__ __forceinline__ void load16(const TextView& tv, uint64_t a, uint32_t w[4]) {
    const uint64_t lo = tv.lead, hi = tv.lead + tv.n;
    w[0] = w[1] = w[2] = w[3] = 0u;
    if (a + 16 <= lo || a >= hi) return;
    if (a + 16 <= hi) {
        const uint4 v = *reinterpret_cast<const uint4*>(tv.base + a);
        w[0] = v.x; w[1] = v.y; w[2] = v.z; w[3] = v.w;
    } else {                                           // the text's last bytes: no load past its end
#pragma unroll
        for (int j = 0; j < 4; j++) {
            uint32_t x = 0;
#pragma unroll
            for (int b = 0; b < 4; b++) {
                const uint64_t p = a + 4 * j + b;
                if (p < hi) x |= (uint32_t)tv.base[p] << (8 * b);
            }
            w[j] = x;
        }
    }
}

// Newline mask of the view's bytes [a, a + 16) held in w: bit i for byte
// a + i, only for bytes of the text (lead <= a + i < lead + n).
__device__ __forceinline__ uint32_t nl_bits(const TextView& tv, uint64_t a, const uint32_t w[4]) {
    const uint64_t lo = tv.lead, hi = tv.lead + tv.n;
    if (a + 16 <= lo || a >= hi) return 0u;
    uint32_t m = 0;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const uint32_t x = w[j] ^ 0x0A0A0A0Au;                 // newline bytes -> 0
        uint32_t t = (x & 0x7F7F7F7Fu) + 0x7F7F7F7Fu;          // exact zero-byte test: bit 7 of each byte
        t = ~(t | x | 0x7F7F7F7Fu) >> 7;                       // bits 0, 8, 16, 24
        t |= t >> 7;                                           // bits 0, 1 | 9, 16 | 17, 24
        t |= t >> 14;                                          // bits 0, 1, 2, 3 (and above)
        m |= (t & 15u) << (4 * j);
    }
    // bytes outside the text
    if (a < lo) m &= ~0u << (uint32_t)(lo - a);
    if (a + 16 > hi) m &= (1u << (uint32_t)(hi - a)) - 1u;
    return m;
}

__device__ __forceinline__ uint32_t nl_mask16(const TextView& tv, uint64_t a) {
    uint32_t w[4];
    load16(tv, a, w);
    return nl_bits(tv, a, w);
}

// Block sum of one value per thread (IN_NT threads).
__device__ __forceinline__ uint64_t block_excl_scan(uint64_t v, uint64_t* s_w, uint64_t* total) {
    const uint32_t lane = __lane_id(), wv = threadIdx.x / GKC_WAVE;
    uint64_t inc = v;
#pragma unroll
    for (int off = 1; off < GKC_WAVE; off <<= 1) {
        const uint64_t y = __shfl_up(inc, off);
        if ((int)lane >= off) inc += y;
    }
    if (lane == GKC_WAVE - 1) s_w[wv] = inc;
    __syncthreads();
    uint64_t before = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < IN_NT / GKC_WAVE; w++) { const uint64_t x = s_w[w]; before += (uint32_t)w < wv ? x : 0ull; tot += x; }
    *total = tot;
    return before + inc - v;
}

// Newlines per block.  Block 0 also records the text's first and last bytes
// in misc[0] and clears the malformed-record flag misc[1].
__global__ __launch_bounds__(IN_NT) void k_nl_count(TextView tv, uint64_t nblk, uint64_t* __restrict__ cnt,
                                                    uint64_t* __restrict__ misc) {
    __shared__ uint64_t s_w[IN_NT / GKC_WAVE];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        misc[0] = (uint64_t)tv.at(0) | (uint64_t)tv.at(tv.n - 1) << 8;
        misc[1] = 0;
    }
    for (uint64_t b = blockIdx.x; b < nblk; b += gridDim.x) {
        const uint32_t m = nl_mask16(tv, b * IN_BLK + 16 * threadIdx.x);
        uint64_t tot;
        (void)block_excl_scan((uint64_t)__popc(m), s_w, &tot);
        if (threadIdx.x == 0) cnt[b] = tot;
        __syncthreads();
    }
}

// Every newline, from its rank g among the newlines (incl[b] = newlines of
// blocks 0..b), placed in its record: with L lines per record (FASTQ 4,
// FASTA 2), newline g ends line g of record r = g / L.
//   g % L == 0      header end: the sequence starts after it     -> src[r]
//   g % L == 1      sequence end, less a '\r' before it         -> end[r]
//                   (FASTQ: the next line must start with '+')
//   g % L == L - 1  record end: the next record's header must start with
//                   the format's marker
// A last FASTA line without newline is the last record's sequence; its end,
// tail_end, comes from the host (~0 when there is none).  The first header is
// checked on the host.
__global__ __launch_bounds__(IN_NT) void k_lines(TextView tv, uint64_t nblk, const uint64_t* __restrict__ incl,
                                                 int fastq, uint64_t nrec, uint64_t tail_end,
                                                 uint64_t* __restrict__ src, uint64_t* __restrict__ end,
                                                 uint32_t* __restrict__ bad) {
    __shared__ uint64_t s_w[IN_NT / GKC_WAVE];
    __shared__ uint32_t s_nl[IN_BLK];                           // at most one newline per byte
    const uint32_t lg = fastq ? 2 : 1, lm = (1u << lg) - 1;
    const uint8_t marker = fastq ? '@' : '>';
    if (blockIdx.x == 0 && threadIdx.x == 0 && tail_end != ~0ull) end[nrec - 1] = tail_end;
    // a thread's 16 bytes and the bytes just before and after them (from
    // addresses clamped into the text: they are used only where they are
    // text), loaded one block ahead, so that nothing loaded is waited for
    // before the end of the iteration
    const uint64_t lo = tv.lead, hi = tv.lead + tv.n - 1;
    auto load_chunk = [&](uint64_t a, uint32_t w[4], uint32_t& pb, uint32_t& nb) {
        load16(tv, a, w);
        pb = tv.base[std::min(std::max(a, lo + 1) - 1, hi)];
        nb = tv.base[std::min(std::max(a + 16, lo), hi)];
    };
    uint32_t w[4] = {0u, 0u, 0u, 0u}, prev = 0, next = 0;
    if (blockIdx.x < nblk) load_chunk((uint64_t)blockIdx.x * IN_BLK + 16 * threadIdx.x, w, prev, next);
    for (uint64_t b = blockIdx.x; b < nblk; b += gridDim.x) {
        const uint64_t a = b * IN_BLK + 16 * threadIdx.x;
        const uint64_t inc_b = incl[b];
        uint32_t nw[4] = {0u, 0u, 0u, 0u}, nprev = 0, nnext = 0;
        if (b + gridDim.x < nblk) load_chunk((b + gridDim.x) * IN_BLK + 16 * threadIdx.x, nw, nprev, nnext);
        uint32_t m = nl_bits(tv, a, w);
        auto byte_at = [&](int i) -> uint32_t {                 // view byte a + i, -1 <= i <= 16
            if (i < 0) return prev;
            if (i > 15) return next;
            return (w[i >> 2] >> (8 * (i & 3))) & 0xFFu;
        };
        uint64_t tot;
        uint32_t o = (uint32_t)block_excl_scan((uint64_t)__popc(m), s_w, &tot);
        // the block's newlines in order in LDS: offset in the block (12 bits),
        // the bytes before (bits 12-19) and after (20-27) it
        for (; m; m &= m - 1) {
            const int i = __builtin_ctz(m);
            s_nl[o++] = (16 * threadIdx.x + (uint32_t)i) | byte_at(i - 1) << 12 | byte_at(i + 1) << 20;
        }
        __syncthreads();
        // then placed one per thread, newline rank g = the block's first + t
        const uint64_t g0 = inc_b - tot, base = b * IN_BLK - tv.lead;
        bool wrong = false;
        for (uint32_t t = threadIdx.x; t < (uint32_t)tot; t += IN_NT) {
            const uint32_t e = s_nl[t], after = e >> 20, before = (e >> 12) & 0xFFu;
            const uint64_t g = g0 + t, p = base + (e & 0xFFFu);                 // the newline's text position
            const uint32_t j = (uint32_t)g & lm;
            const uint64_t r = g >> lg;
            if (j <= 1)                  // j == 1: p >= 1, a header line precedes
                (j == 0 ? src : end)[r] = j == 0 ? p + 1 : p - (before == '\r' ? 1 : 0);
            wrong |= (fastq && j == 1 && after != '+') || (j == lm && r + 1 < nrec && after != marker);
        }
        if (wrong) atomicOr(bad, 1u);
#pragma unroll
        for (int k = 0; k < 4; k++) w[k] = nw[k];
        prev = nprev;
        next = nnext;
        __syncthreads();
    }
}

// In-place inclusive scan of x[0, n) - sub[0, n) (u64; sub may be null): each
// 1024-element group scanned by a workgroup with its total to gsum[g]; gsum
// scanned by one workgroup; the groups' exclusive offsets added.
constexpr int SC_NT = 1024;
__global__ __launch_bounds__(SC_NT) void k_scan_groups(uint64_t* __restrict__ x, const uint64_t* __restrict__ sub,
                                                       uint64_t n, uint64_t* __restrict__ gsum) {
    __shared__ uint64_t s_w[SC_NT / GKC_WAVE];
    const uint32_t lane = __lane_id(), wv = threadIdx.x / GKC_WAVE;
    const uint64_t i = (uint64_t)blockIdx.x * SC_NT + threadIdx.x;
    const uint64_t v = i < n ? x[i] - (sub ? sub[i] : 0ull) : 0ull;
    uint64_t inc = v;
#pragma unroll
    for (int off = 1; off < GKC_WAVE; off <<= 1) {
        const uint64_t y = __shfl_up(inc, off);
        if ((int)lane >= off) inc += y;
    }
    if (lane == GKC_WAVE - 1) s_w[wv] = inc;
    __syncthreads();
    uint64_t before = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < SC_NT / GKC_WAVE; w++) { const uint64_t y = s_w[w]; before += (uint32_t)w < wv ? y : 0ull; tot += y; }
    if (i < n) x[i] = before + inc;
    if (threadIdx.x == 0) gsum[blockIdx.x] = tot;
}
__global__ __launch_bounds__(SC_NT) void k_scan_sums(uint64_t* __restrict__ gsum, uint64_t ng) {
    __shared__ uint64_t s_w[SC_NT / GKC_WAVE];
    const uint32_t lane = __lane_id(), wv = threadIdx.x / GKC_WAVE;
    uint64_t carry = 0;
    for (uint64_t b0 = 0; b0 < ng; b0 += SC_NT) {
        const uint64_t i = b0 + threadIdx.x;
        const uint64_t v = i < ng ? gsum[i] : 0ull;
        uint64_t inc = v;
#pragma unroll
        for (int off = 1; off < GKC_WAVE; off <<= 1) {
            const uint64_t y = __shfl_up(inc, off);
            if ((int)lane >= off) inc += y;
        }
        if (lane == GKC_WAVE - 1) s_w[wv] = inc;
        __syncthreads();
        uint64_t before = 0, tot = 0;
#pragma unroll
        for (int w = 0; w < SC_NT / GKC_WAVE; w++) { const uint64_t y = s_w[w]; before += (uint32_t)w < wv ? y : 0ull; tot += y; }
        if (i < ng) gsum[i] = carry + before + inc - v;          // exclusive
        carry += tot;
        __syncthreads();
    }
}
__global__ __launch_bounds__(SC_NT) void k_scan_add(uint64_t* __restrict__ x, uint64_t n, const uint64_t* __restrict__ gsum) {
    const uint64_t i = (uint64_t)blockIdx.x * SC_NT + threadIdx.x;
    if (i < n) x[i] += gsum[blockIdx.x];
}

// The bases output viewed from the 4-byte boundary at or below it: output
// byte q is base[lead + q].
struct OutView {
    uint8_t* base;
    uint64_t lead;
};

constexpr int GA_NT = 256;
constexpr int GA_U = 8;                                 // 64-dword steps per pass: 2 KiB of loads in flight per wave
constexpr uint64_t GA_CHUNK = 4 * GKC_WAVE * GA_U * 4;  // output-view bytes per chunk (8 KiB, 4 passes)

// Sequence bytes to the output, by output chunks: each wave takes 8 KiB
// chunks of the output view (chunk edges on 4-byte boundaries) and writes
// them as aligned dwords.  Record r's output byte q comes from text byte
// q + src[r] - offs[r].  The wave keeps a window of 64 records (their ends
// and that difference) in LDS, starting at the record that holds the
// chunk's first byte (a 64-way search over offs); each lane finds its
// dword's record in the window by stepping forward from the record that
// holds the pass's first byte, loads the two aligned text dwords that hold
// the dword's 4 bytes and shifts them together; a dword where one record
// ends and the next begins takes two such pairs and a byte mask.  A dword at
// a pass's edge or with bytes of three or more records is written byte by
// byte, its loads issued before its stores.
// A pass stops at the window's last record end; the window then moves on by
// 64 records (empty records take window slots but no bytes).
__global__ __launch_bounds__(GA_NT) void k_gather(TextView tv, const uint64_t* __restrict__ src,
                                                  const uint64_t* __restrict__ offs, uint64_t nrec, OutView ov) {
    __shared__ uint64_t s_e[GA_NT / GKC_WAVE][GKC_WAVE], s_d[GA_NT / GKC_WAVE][GKC_WAVE];
    const uint32_t lane = __lane_id(), wv = threadIdx.x / GKC_WAVE;
    uint64_t* se = s_e[wv];
    uint64_t* sd = s_d[wv];
    const GKC_GLOBAL uint32_t* tw = as_global(reinterpret_cast<const uint32_t*>(tv.base));
    uint32_t* ow = reinterpret_cast<uint32_t*>(ov.base);
    const uint64_t T = offs[nrec];
    const uint64_t vend = ov.lead + T;                          // output view end
    const uint64_t nchunk = (vend + GA_CHUNK - 1) / GA_CHUNK;
    const uint64_t safe = tv.lead >> 2;                          // a dword holding text byte 0
    for (uint64_t c = (uint64_t)blockIdx.x * (GA_NT / GKC_WAVE) + wv; c < nchunk;
         c += (uint64_t)gridDim.x * (GA_NT / GKC_WAVE)) {
        uint64_t q = std::max<uint64_t>(c * GA_CHUNK, ov.lead) - ov.lead;      // output cursor
        const uint64_t qc = std::min<uint64_t>((c + 1) * GA_CHUNK, vend) - ov.lead;
        if (q >= qc) continue;                                  // no bytes (every read empty)
        // the record holding byte q: offs[lo] <= q < offs[lo + 1]
        uint64_t lo = 0, hi = nrec;
        while (hi - lo > 1) {
            const uint64_t step = (hi - lo + GKC_WAVE - 1) / GKC_WAVE, pos = lo + lane * step;
            const uint64_t b = __ballot(pos < hi && offs[pos] <= q);            // lane 0 always
            lo += (uint64_t)(63 - __clzll(b)) * step;
            hi = std::min(hi, lo + step);
        }
        uint64_t wr = lo - GKC_WAVE, E = q, e = 0;
        while (q < qc) {
            if (q >= E) {                                        // move the window on
                wr += GKC_WAVE;
                const uint64_t r = wr + lane;
                e = r < nrec ? offs[r + 1] : ~0ull;
                se[lane] = e;
                sd[lane] = r < nrec ? src[r] - offs[r] : 0ull;
                E = __shfl(e, GKC_WAVE - 1);
                wave_lds_fence();
                continue;
            }
            const uint64_t D0 = (ov.lead + q) >> 2;
            const uint64_t qe = std::min(std::min(qc, E), 4 * (D0 + GKC_WAVE * GA_U) - ov.lead);
            const uint64_t va0 = ov.lead + q, vb0 = ov.lead + qe;           // view bytes of this pass
            // window slot of the record holding byte q; each lane's slot only
            // moves forward as its dwords do
            uint32_t idx = (uint32_t)__popcll(__ballot(e <= q));
            // per step: the aligned text dwords [a, a + 1] of the bytes from the
            // dword's first record and [b, b + 1] of those from the next, and in
            // `how` the two shifts (bits 0-7, 8-15), how many leading bytes come
            // from the first record (16-18; 4 = all), the form (20-21: 0 nothing,
            // 1 aligned dwords, 2 byte by byte) and the first record's slot (24-29)
            uint32_t wa0[GA_U], wa1[GA_U], wb0[GA_U], wb1[GA_U], how[GA_U];
#pragma unroll
            for (int u = 0; u < GA_U; u++) {
                const uint64_t D = D0 + GKC_WAVE * u + lane;
                const uint64_t va = std::max(4 * D, va0), vb = std::min(4 * D + 4, vb0);
                uint64_t A = safe, B = safe;
                uint32_t h = 0;
                if (va < vb) {
                    const uint64_t pa = va - ov.lead;
                    while (se[idx] <= pa) idx++;
                    h = 2u << 20 | idx << 24;
                    if (va == 4 * D && vb == 4 * D + 4) {
                        const uint64_t ea = se[idx];
                        const uint64_t Va = tv.lead + pa + sd[idx];
                        if (ea >= pa + 4) {                          // one record
                            A = B = Va >> 2;
                            h = 1u << 20 | 4u << 16 | (uint32_t)(Va & 3) * 8 | (uint32_t)(Va & 3) * 8 << 8;
                        } else if (se[idx + 1] >= pa + 4) {          // two: a record ends inside the dword
                            const uint64_t Vb = tv.lead + pa + sd[idx + 1];
                            A = Va >> 2;
                            B = Vb >> 2;
                            h = 1u << 20 | (uint32_t)(ea - pa) << 16 | (uint32_t)(Va & 3) * 8 | (uint32_t)(Vb & 3) * 8 << 8;
                        }
                    }
                }
                how[u] = h;
                const uint32_t sa = h & 0xFF, sb = (h >> 8) & 0xFF;
                wa0[u] = tw[A];
                wa1[u] = tw[sa ? A + 1 : A];
                wb0[u] = tw[B];
                wb1[u] = tw[sb ? B + 1 : B];
            }
#pragma unroll
            for (int u = 0; u < GA_U; u++) {
                const uint64_t D = D0 + GKC_WAVE * u + lane;
                const uint32_t h = how[u], form = (h >> 20) & 3;
                if (form == 1) {
                    const uint32_t sa = h & 0xFF, sb = (h >> 8) & 0xFF, na = (h >> 16) & 7;
                    const uint32_t xa = sa ? (wa0[u] >> sa) | (wa1[u] << (32 - sa)) : wa0[u];
                    const uint32_t xb = sb ? (wb0[u] >> sb) | (wb1[u] << (32 - sb)) : wb0[u];
                    const uint32_t m = na >= 4 ? ~0u : (1u << (8 * na)) - 1u;
                    ow[D] = (xa & m) | (xb & ~m);
                } else if (form == 2) {                              // pass edge, or 3+ records in the dword
                    const uint64_t va = std::max(4 * D, va0), vb = std::min(4 * D + 4, vb0);
                    uint32_t k = (h >> 24) & 63;
                    uint8_t bt[4];
#pragma unroll
                    for (int t = 0; t < 4; t++) {
                        const uint64_t v = 4 * D + t;
                        bt[t] = 0;
                        if (v >= va && v < vb) {
                            const uint64_t p = v - ov.lead;
                            while (se[k] <= p) k++;
                            bt[t] = tv.at(p + sd[k]);
                        }
                    }
#pragma unroll
                    for (int t = 0; t < 4; t++)
                        if (4 * D + t >= va && 4 * D + t < vb) ov.base[4 * D + t] = bt[t];
                }
            }
            q = qe;
        }
    }
}

namespace {
constexpr int MAX_DEV = 64;
// Scratch kept per device between calls (grown when an input needs more;
// held until the process ends); a call holds its device's lock throughout.
struct DevScratch {
    std::mutex mu;
    void* p[2] = {nullptr, nullptr};
    size_t bytes[2] = {0, 0};
};
DevScratch g_scratch[MAX_DEV];

// The fixed scratch of gkc_parse_reads_noalloc: one array per GPU in the
// library's device image.  Host code never passes the symbol itself (a
// host-side stand-in); it takes the device address from hipGetSymbolAddress,
// once per device, and passes that.
__device__ __align__(256) uint64_t scratch[NOALLOC_SCRATCH_WORDS];
struct FixedScratch {
    std::mutex mu;
    uint64_t* ptr = nullptr;       // device address of `scratch` on this device
};
FixedScratch g_fixed[MAX_DEV];

int hip_fail(hipError_t e, const char* where, std::string& err) {
    (void)hipGetLastError();
    err = std::string("HIP error ") + hipGetErrorString(e) + " in " + where;
    return GKC_EHIP;
}
#define IN_HIP(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) return hip_fail(e_, #x, err); } while (0)

uint64_t scan_groups(uint64_t n) { return (n + SC_NT - 1) / SC_NT; }

// Inclusive scan of x - sub over n elements, with scan_groups(n) words of
// group sums at gs; queued on st.
void scan_inplace(uint64_t* x, const uint64_t* sub, uint64_t n, uint64_t* gs, hipStream_t st) {
    if (n == 0) return;
    const uint64_t ng = scan_groups(n);
    hipLaunchKernelGGL(k_scan_groups, dim3((uint32_t)ng), dim3(SC_NT), 0, st, x, sub, n, gs);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(SC_NT), 0, st, gs, ng);
    hipLaunchKernelGGL(k_scan_add, dim3((uint32_t)ng), dim3(SC_NT), 0, st, x, n, (const uint64_t*)gs);
}

// Words of the two scratch regions: block counts (+ first/last bytes, the
// malformed flag and the scan's group sums), then record starts (+ group sums).
uint64_t words_blocks(uint64_t nblk) { return nblk + 2 + scan_groups(nblk); }
uint64_t words_records(uint64_t nrec) { return nrec + scan_groups(nrec); }

// The parse proper.  region(i, words, &p) hands out scratch region i (0:
// block counts, 1: record starts) of at least `words` u64 words, or fails
// with a GKC_E* code and err set.  Two device round trips: after the newline
// count (the host needs the record count and the text's first and last
// bytes) and at the end (the malformed flag).  The kernels write every
// scratch word they read, except the malformed flag, which k_nl_count clears.
template <typename Region>
int parse_core(const char* text, uint64_t n, int format, char* bases, uint64_t* offs, hipStream_t st,
               int64_t* n_reads, std::string& err, Region&& region) {
    TextView tv;
    tv.lead = reinterpret_cast<uintptr_t>(text) & 15u;
    tv.base = reinterpret_cast<const uint8_t*>(text) - tv.lead;
    tv.n = n;
    const uint64_t nblk = (tv.lead + n + IN_BLK - 1) / IN_BLK;
    const uint32_t grid = (uint32_t)std::min<uint64_t>(nblk, 65536);

    // blk[nblk] | misc[2] | group sums
    uint64_t* blk = nullptr;
    if (int rc = region(0, words_blocks(nblk), &blk)) return rc;
    uint64_t* misc = blk + nblk;
    hipLaunchKernelGGL(k_nl_count, dim3(grid), dim3(IN_NT), 0, st, tv, nblk, blk, misc);
    scan_inplace(blk, nullptr, nblk, misc + 2, st);
    IN_HIP(hipGetLastError());
    uint64_t h[2];
    IN_HIP(hipMemcpyAsync(h, blk + nblk - 1, 16, hipMemcpyDeviceToHost, st));
    IN_HIP(hipStreamSynchronize(st));
    const uint64_t n_nl = h[0];
    const uint8_t first = (uint8_t)h[1], last = (uint8_t)(h[1] >> 8);

    if (format == 0) format = first == '>' ? 1 : first == '@' ? 2 : 0;
    if (format == 0) { err = "text starts with neither '>' (FASTA) nor '@' (FASTQ)"; return GKC_EINVAL; }
    const int fastq = format == 2;
    if (first != (fastq ? '@' : '>')) {
        err = fastq ? "malformed FASTQ record (header without '@' or separator without '+')" : "malformed FASTA record (header without '>')";
        return GKC_EINVAL;
    }
    const uint64_t n_lines = n_nl + (last != '\n' ? 1 : 0);
    const uint64_t L = fastq ? 4 : 2;
    if (n_lines % L) { err = fastq ? "FASTQ text does not hold whole 4-line records" : "FASTA text does not hold whole 2-line records"; return GKC_EINVAL; }
    const uint64_t nrec = n_lines / L;
    const uint64_t tail_end = (!fastq && last != '\n') ? n - (last == '\r' ? 1 : 0) : ~0ull;

    // src[nrec] | group sums
    uint64_t* src = nullptr;
    if (int rc = region(1, words_records(nrec), &src)) return rc;
    uint32_t* bad = reinterpret_cast<uint32_t*>(misc + 1);
    IN_HIP(hipMemsetAsync(offs, 0, 8, st));
    hipLaunchKernelGGL(k_lines, dim3(grid), dim3(IN_NT), 0, st, tv, nblk, (const uint64_t*)blk, fastq, nrec, tail_end,
                       src, offs + 1, bad);
    scan_inplace(offs + 1, src, nrec, src + nrec, st);
    OutView ov;
    ov.lead = reinterpret_cast<uintptr_t>(bases) & 3u;
    ov.base = reinterpret_cast<uint8_t*>(bases) - ov.lead;
    // the output is at most n bytes: chunks of GA_CHUNK view bytes, 4 per workgroup
    const uint64_t max_chunks = (ov.lead + n + GA_CHUNK - 1) / GA_CHUNK;
    const uint32_t ggrid = (uint32_t)std::min<uint64_t>((max_chunks + 3) / 4, 2048);
    hipLaunchKernelGGL(k_gather, dim3(ggrid), dim3(GA_NT), 0, st, tv, (const uint64_t*)src, (const uint64_t*)offs, nrec, ov);
    IN_HIP(hipGetLastError());
    uint32_t hbad = 0;
    IN_HIP(hipMemcpyAsync(&hbad, bad, 4, hipMemcpyDeviceToHost, st));
    IN_HIP(hipStreamSynchronize(st));
    if (hbad) { err = fastq ? "malformed FASTQ record (header without '@' or separator without '+')" : "malformed FASTA record (header without '>')"; return GKC_EINVAL; }
    *n_reads = (int64_t)nrec;
    return 0;
}

// Checks shared by both entry points; 1 when the text is empty (done).
int parse_precheck(uint64_t n, int format, uint64_t* offs, hipStream_t st, std::string& err, int* rc) {
    *rc = 0;
    if (format < 0 || format > 2) { err = "format must be 0 (from the first byte), 1 (FASTA) or 2 (FASTQ)"; *rc = GKC_EINVAL; return 1; }
    if (n == 0) {
        hipError_t e = hipMemsetAsync(offs, 0, 8, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) *rc = hip_fail(e, "empty text", err);
        return 1;
    }
    return 0;
}

int current_device(std::string& err, int* dev) {
    IN_HIP(hipGetDevice(dev));
    if (*dev < 0 || *dev >= MAX_DEV) { err = "device id out of range"; return GKC_EINVAL; }
    return 0;
}
}  // namespace

int parse_reads(const char* text, uint64_t n, int format, char* bases, uint64_t* offs, hipStream_t st,
                int64_t* n_reads, std::string& err) {
    *n_reads = 0;
    int rc = 0;
    if (parse_precheck(n, format, offs, st, err, &rc)) return rc;
    int dev = 0;
    if ((rc = current_device(err, &dev))) return rc;
    DevScratch& ds = g_scratch[dev];
    std::lock_guard<std::mutex> lk(ds.mu);
    struct Drain {                               // no work on the scratch outlives the lock
        hipStream_t s;
        ~Drain() { (void)hipStreamSynchronize(s); }
    } drain{st};
    auto region = [&](int i, uint64_t words, uint64_t** p) -> int {
        const size_t need = words * 8;
        if (need > ds.bytes[i]) {
            if (ds.p[i]) {
                const hipError_t e = hipFree(ds.p[i]);
                ds.p[i] = nullptr;
                ds.bytes[i] = 0;
                if (e != hipSuccess) return hip_fail(e, "hipFree(scratch)", err);
            }
            const hipError_t e = hipMalloc(&ds.p[i], need);
            if (e != hipSuccess) { ds.p[i] = nullptr; return hip_fail(e, "hipMalloc(scratch)", err); }
            ds.bytes[i] = need;
        }
        *p = static_cast<uint64_t*>(ds.p[i]);
        return 0;
    };
    return parse_core(text, n, format, bases, offs, st, n_reads, err, region);
}

int parse_reads_noalloc(const char* text, uint64_t n, int format, char* bases, uint64_t* offs, hipStream_t st,
                        int64_t* n_reads, std::string& err) {
    *n_reads = 0;
    int rc = 0;
    if (parse_precheck(n, format, offs, st, err, &rc)) return rc;
    int dev = 0;
    if ((rc = current_device(err, &dev))) return rc;
    FixedScratch& fs = g_fixed[dev];
    std::lock_guard<std::mutex> lk(fs.mu);
    if (!fs.ptr) {                               // first call on this device (it is current)
        void* p = nullptr;
        IN_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        fs.ptr = static_cast<uint64_t*>(p);
    }
    struct Drain {
        hipStream_t s;
        ~Drain() { (void)hipStreamSynchronize(s); }
    } drain{st};
    // region 0 from the array's start, region 1 after it (256-byte aligned)
    uint64_t used = 0;
    auto region = [&](int i, uint64_t words, uint64_t** p) -> int {
        const uint64_t at = i == 0 ? 0 : (used + 31) / 32 * 32;
        if (at + words > NOALLOC_SCRATCH_WORDS) {
            err = "text too large for gkc_parse_reads_noalloc's fixed scratch: it needs " + std::to_string((at + words) * 8) +
                  " bytes, the scratch holds " + std::to_string(NOALLOC_SCRATCH_WORDS * 8) + " (use gkc_parse_reads)";
            return GKC_ECAP;
        }
        used = at + words;
        *p = fs.ptr + at;
        return 0;
    };
    return parse_core(text, n, format, bases, offs, st, n_reads, err, region);
}

}  // namespace gkc
