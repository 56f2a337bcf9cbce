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
// block, the counts scanned, the newline positions written in order; per
// record its header / separator checked and its sequence line located; the
// sequence lengths scanned into offsets; the sequence bytes gathered.
#include <algorithm>
#include <string>

#include "gkc_hip.h"
#include "kernels.hpp"

namespace gkc {

constexpr int IN_NT = 256;                     // threads per block
constexpr uint64_t IN_BLK = IN_NT * 16;        // bytes of the aligned view per block (16 per thread)

// The text viewed from the 16-byte boundary at or below it: text byte q is
// base[lead + q].  A thread's 16 bytes are base[a, a + 16), a = 16 * (thread).
struct TextView {
    const uint8_t* base;
    uint64_t lead, n;
    __device__ uint8_t at(uint64_t q) const { return base[lead + q]; }
};

// Newline mask of the view's bytes [a, a + 16): bit i for byte a + i, only
// for bytes of the text (lead <= a + i < lead + n).
__device__ __forceinline__ uint32_t nl_mask16(const TextView& tv, uint64_t a) {
    const uint64_t lo = tv.lead, hi = tv.lead + tv.n;
    if (a + 16 <= lo || a >= hi) return 0u;
    uint32_t w[4];
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
    uint32_t m = 0;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const uint32_t x = w[j] ^ 0x0A0A0A0Au;                 // newline bytes -> 0
        uint32_t t = (x & 0x7F7F7F7Fu) + 0x7F7F7F7Fu;          // exact zero-byte test: bit 7 of each byte
        t = ~(t | x | 0x7F7F7F7Fu);
#pragma unroll
        for (int b = 0; b < 4; b++) m |= ((t >> (8 * b + 7)) & 1u) << (4 * j + b);
    }
    // bytes outside the text
    if (a < lo) m &= ~0u << (uint32_t)(lo - a);
    if (a + 16 > hi) m &= (1u << (uint32_t)(hi - a)) - 1u;
    return m;
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

// Newlines per block.
__global__ __launch_bounds__(IN_NT) void k_nl_count(TextView tv, uint64_t nblk, uint64_t* __restrict__ cnt) {
    __shared__ uint64_t s_w[IN_NT / GKC_WAVE];
    for (uint64_t b = blockIdx.x; b < nblk; b += gridDim.x) {
        const uint32_t m = nl_mask16(tv, b * IN_BLK + 16 * threadIdx.x);
        uint64_t tot;
        (void)block_excl_scan((uint64_t)__popc(m), s_w, &tot);
        if (threadIdx.x == 0) cnt[b] = tot;
        __syncthreads();
    }
}

// Newline positions (text positions, ascending); incl[b] = newlines of blocks 0..b.
__global__ __launch_bounds__(IN_NT) void k_nl_write(TextView tv, uint64_t nblk, const uint64_t* __restrict__ incl,
                                                    uint64_t* __restrict__ nl) {
    __shared__ uint64_t s_w[IN_NT / GKC_WAVE];
    for (uint64_t b = blockIdx.x; b < nblk; b += gridDim.x) {
        const uint64_t a = b * IN_BLK + 16 * threadIdx.x;
        uint32_t m = nl_mask16(tv, a);
        uint64_t tot;
        const uint64_t before = block_excl_scan((uint64_t)__popc(m), s_w, &tot);
        uint64_t o = incl[b] - tot + before;
        while (m) {
            const uint32_t i = (uint32_t)__builtin_ctz(m);
            m &= m - 1;
            nl[o++] = a + i - tv.lead;
        }
        __syncthreads();
    }
}

// Per record: the header (and, FASTQ, the separator) checked, the sequence
// line's start and length (without a trailing '\r').  Line i of the text is
// [start(i), end(i)): start(0) = 0, start(i) = nl[i - 1] + 1, end(i) = nl[i]
// or the text's end for a last line without newline.
__global__ __launch_bounds__(IN_NT) void k_records(TextView tv, const uint64_t* __restrict__ nl, uint64_t n_nl,
                                                   uint64_t nrec, int fastq, uint64_t* __restrict__ src,
                                                   uint64_t* __restrict__ len, uint32_t* __restrict__ bad) {
    const uint64_t L = fastq ? 4 : 2;
    for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < nrec; r += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t h = L * r;
        auto lstart = [&](uint64_t i) { return i == 0 ? 0ull : nl[i - 1] + 1; };
        auto lend = [&](uint64_t i) { return i < n_nl ? nl[i] : tv.n; };
        const uint64_t hs = lstart(h);
        bool ok = hs < lend(h) && tv.at(hs) == (fastq ? '@' : '>');
        const uint64_t ss = lstart(h + 1);
        uint64_t se = lend(h + 1);
        if (se > ss && tv.at(se - 1) == '\r') se--;
        if (fastq) {
            const uint64_t ps = lstart(h + 2);
            ok = ok && ps < lend(h + 2) && tv.at(ps) == '+';
        }
        if (!ok) atomicOr(bad, 1u);
        src[r] = ss;
        len[r] = se - ss;
    }
}

// In-place inclusive scan of x[0, n) (u64): each 1024-element group scanned
// by a workgroup with its total to gsum[g]; gsum scanned by one workgroup;
// the groups' exclusive offsets added.
constexpr int SC_NT = 1024;
__global__ __launch_bounds__(SC_NT) void k_scan_groups(uint64_t* __restrict__ x, uint64_t n, uint64_t* __restrict__ gsum) {
    __shared__ uint64_t s_w[SC_NT / GKC_WAVE];
    const uint32_t lane = __lane_id(), wv = threadIdx.x / GKC_WAVE;
    const uint64_t i = (uint64_t)blockIdx.x * SC_NT + threadIdx.x;
    const uint64_t v = i < n ? x[i] : 0ull;
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

// The reads' bases back to back: one wave per read (grid-stride).
__global__ __launch_bounds__(IN_NT) void k_gather(TextView tv, const uint64_t* __restrict__ src,
                                                  const uint64_t* __restrict__ offs, uint64_t nrec,
                                                  uint8_t* __restrict__ bases) {
    const uint32_t lane = __lane_id();
    const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / GKC_WAVE;
    const uint64_t nw = (uint64_t)gridDim.x * blockDim.x / GKC_WAVE;
    for (uint64_t r = wave; r < nrec; r += nw) {
        const uint64_t s = src[r], o = offs[r], l = offs[r + 1] - o;
        for (uint64_t i = lane; i < l; i += GKC_WAVE) bases[o + i] = tv.at(s + i);
    }
}

namespace {
struct Scratch {
    void* p = nullptr;
    ~Scratch() { if (p) (void)hipFree(p); }
};
int hip_fail(hipError_t e, const char* where, std::string& err) {
    (void)hipGetLastError();
    err = std::string("HIP error ") + hipGetErrorString(e) + " in " + where;
    return GKC_EHIP;
}
#define IN_HIP(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) return hip_fail(e_, #x, err); } while (0)

int scan_inplace(uint64_t* x, uint64_t n, hipStream_t st, std::string& err) {
    if (n == 0) return 0;
    const uint64_t ng = (n + SC_NT - 1) / SC_NT;
    Scratch g;
    IN_HIP(hipMalloc(&g.p, ng * 8));
    uint64_t* gs = static_cast<uint64_t*>(g.p);
    hipLaunchKernelGGL(k_scan_groups, dim3((uint32_t)ng), dim3(SC_NT), 0, st, x, n, gs);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(SC_NT), 0, st, gs, ng);
    hipLaunchKernelGGL(k_scan_add, dim3((uint32_t)ng), dim3(SC_NT), 0, st, x, n, (const uint64_t*)gs);
    IN_HIP(hipGetLastError());
    IN_HIP(hipStreamSynchronize(st));                // before the group sums are freed
    return 0;
}
}  // namespace

int parse_reads(const char* text, uint64_t n, int format, char* bases, uint64_t* offs, hipStream_t st,
                int64_t* n_reads, std::string& err) {
    *n_reads = 0;
    if (n == 0) {
        IN_HIP(hipMemsetAsync(offs, 0, 8, st));
        IN_HIP(hipStreamSynchronize(st));
        return 0;
    }
    if (format < 0 || format > 2) { err = "format must be 0 (from the first byte), 1 (FASTA) or 2 (FASTQ)"; return GKC_EINVAL; }
    uint8_t ends[2];
    IN_HIP(hipMemcpyAsync(&ends[0], text, 1, hipMemcpyDeviceToHost, st));
    IN_HIP(hipMemcpyAsync(&ends[1], text + n - 1, 1, hipMemcpyDeviceToHost, st));
    IN_HIP(hipStreamSynchronize(st));
    if (format == 0) format = ends[0] == '>' ? 1 : ends[0] == '@' ? 2 : 0;
    if (format == 0) { err = "text starts with neither '>' (FASTA) nor '@' (FASTQ)"; return GKC_EINVAL; }
    const int fastq = format == 2;

    TextView tv;
    tv.lead = reinterpret_cast<uintptr_t>(text) & 15u;
    tv.base = reinterpret_cast<const uint8_t*>(text) - tv.lead;
    tv.n = n;
    const uint64_t nblk = (tv.lead + n + IN_BLK - 1) / IN_BLK;
    const uint32_t grid = (uint32_t)std::min<uint64_t>(nblk, 65536);
    Scratch sb;
    IN_HIP(hipMalloc(&sb.p, nblk * 8));
    uint64_t* blk = static_cast<uint64_t*>(sb.p);
    hipLaunchKernelGGL(k_nl_count, dim3(grid), dim3(IN_NT), 0, st, tv, nblk, blk);
    IN_HIP(hipGetLastError());
    if (int rc = scan_inplace(blk, nblk, st, err)) return rc;
    uint64_t n_nl = 0;
    IN_HIP(hipMemcpyAsync(&n_nl, blk + nblk - 1, 8, hipMemcpyDeviceToHost, st));
    IN_HIP(hipStreamSynchronize(st));

    const uint64_t n_lines = n_nl + (ends[1] != '\n' ? 1 : 0);
    const uint64_t L = fastq ? 4 : 2;
    if (n_lines % L) { err = fastq ? "FASTQ text does not hold whole 4-line records" : "FASTA text does not hold whole 2-line records"; return GKC_EINVAL; }
    const uint64_t nrec = n_lines / L;

    Scratch snl, ssrc, sbad;
    IN_HIP(hipMalloc(&snl.p, std::max<uint64_t>(n_nl, 1) * 8));
    IN_HIP(hipMalloc(&ssrc.p, std::max<uint64_t>(nrec, 1) * 8));
    IN_HIP(hipMalloc(&sbad.p, 4));
    uint64_t* nl = static_cast<uint64_t*>(snl.p);
    uint64_t* src = static_cast<uint64_t*>(ssrc.p);
    uint32_t* bad = static_cast<uint32_t*>(sbad.p);
    IN_HIP(hipMemsetAsync(bad, 0, 4, st));
    IN_HIP(hipMemsetAsync(offs, 0, 8, st));
    hipLaunchKernelGGL(k_nl_write, dim3(grid), dim3(IN_NT), 0, st, tv, nblk, (const uint64_t*)blk, nl);
    const uint32_t rgrid = (uint32_t)std::min<uint64_t>((nrec + IN_NT - 1) / IN_NT + 1, 65536);
    hipLaunchKernelGGL(k_records, dim3(rgrid), dim3(IN_NT), 0, st, tv, (const uint64_t*)nl, n_nl, nrec, fastq, src,
                       offs + 1, bad);
    IN_HIP(hipGetLastError());
    if (int rc = scan_inplace(offs + 1, nrec, st, err)) return rc;
    hipLaunchKernelGGL(k_gather, dim3(rgrid), dim3(IN_NT), 0, st, tv, (const uint64_t*)src, (const uint64_t*)offs, nrec,
                       reinterpret_cast<uint8_t*>(bases));
    IN_HIP(hipGetLastError());
    uint32_t hbad = 0;
    IN_HIP(hipMemcpyAsync(&hbad, bad, 4, hipMemcpyDeviceToHost, st));
    IN_HIP(hipStreamSynchronize(st));
    if (hbad) { err = fastq ? "malformed FASTQ record (header without '@' or separator without '+')" : "malformed FASTA record (header without '>')"; return GKC_EINVAL; }
    *n_reads = (int64_t)nrec;
    return 0;
}

}  // namespace gkc
