// Phase 1 of the k-mer counting path on gfx950: read encoding, rolling
// canonical minimizers, superkmer marking and the partition scatter, fused
// into one tiled kernel; plus the homopolymer-compression pre-pass and the
// kernels that produce the reference's per-sub-batch arrays (minimizer per
// window and superkmer offset lists) for the GenSuperkmerGPU drop-in.
//
// Reference semantics followed (file:line in the reference tree):
//   minimizer of window i = min({4^p-1} U {passing forward and reverse-
//   complement p-mers inside the window})   gkc_cuda.cu:253-286
//   superkmer boundaries where mm(i) != mm(i-1)   gkc_cuda.cu:308-322
//   partition = mm % P                      superkmers.hpp:180
//   HPC: collapse runs of identical bytes    gkc_cuda.cu:178-190
#include "kernels.hpp"

namespace gkc {

// Largest r with offs[r] <= pos (absolute), i.e. the read containing pos.
__device__ __forceinline__ uint32_t find_read(const uint64_t* offs, uint32_t n, uint64_t pos) {
    uint32_t lo = 0, hi = n;   // invariant: offs[lo] <= pos < offs[hi] (offs[n] > pos)
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (offs[mid] <= pos) lo = mid; else hi = mid;
    }
    return lo;
}

// Read holding the first position of each tile (one thread per read; a read
// covering positions [s, e) of the chunk owns the tiles whose first position
// lies in it).  Tile 0 may start in the alignment lead, before read 0.
__global__ __launch_bounds__(256) void k_tile_reads(const uint64_t* __restrict__ offs, uint32_t n_reads, uint64_t base,
                                                    uint32_t n_tiles, uint32_t* __restrict__ tile_r0) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r == 0) tile_r0[0] = 0;
    if (r >= n_reads) return;
    const uint64_t s = offs[r] - base, e = offs[r + 1] - base;
    for (uint64_t t = (s + P1_T - 1) / P1_T; t * P1_T < e && t < n_tiles; t++) tile_r0[t] = r;
}

constexpr uint32_t P1_RO = 512;     // read offsets a tile stages in LDS (more: global search)

// With WIN, every candidate carries its position: the window minimum runs
// over composites (value << 14 | no-candidate << 13 | 8191 - start), so ties
// go to the later p-mer start (the reference's update order, gkc_cuda.cu:281-284)
// and the chosen start and strand are recovered per window.
template <typename MM, bool WIN>
__global__ __launch_bounds__(P1_NT, 4) void k_superkmers(const ChunkDesc* __restrict__ cdp, P1Params prm,
                                                       BinSet bins, uint32_t* __restrict__ tile_ctr) {
    constexpr int CSH = WIN ? 14 : 0;
    __shared__ int8_t s_str[WIN ? P1_T + MAXK : 1];
    __shared__ __align__(16) uint8_t s_bytes[P1_T + MAXK + 32];
    __shared__ MM s_v[P1_T + MAXK];
    __shared__ MM s_last[P1_NT];
    __shared__ uint16_t s_start[P1_T];
    __shared__ uint16_t s_wend[P1_T];
    __shared__ uint16_t s_part[P1_T];
    __shared__ uint32_t s_scan[P1_NT / GKC_WAVE];
    __shared__ uint32_t s_ro[P1_RO];              // offsets (chunk positions) of reads r0 .. r0 + n_ro - 1
    // per partition: block-local [63:48] headers, [47:24] k-mers, [23:0] words;
    // then the block's global bases [63:32] words, [31:0] headers
    __shared__ unsigned long long s_loc[1024];
    __shared__ uint32_t s_gk[1024];
    __shared__ uint32_t s_tile, s_ns, s_ovf;

    const ChunkDesc cd = *cdp;
    const int tid = threadIdx.x;
    const int k = prm.k, p = prm.p, w = k - p + 1;
    const MM M = (MM)((p >= (int)(sizeof(MM) * 4)) ? ~(MM)0 : (((MM)1 << (2 * p)) - 1));
    const MM MC = WIN ? (MM)((M << CSH) | (MM)0x3FFF) : M;     // largest composite
    const uint32_t P = prm.P;

    if (tid == 0) s_tile = atomicAdd(tile_ctr, 1u);
    __syncthreads();
    for (;;) {
        const uint32_t tile = s_tile;
        if (tile >= cd.n_tiles) break;
        uint32_t next_tile = 0;
        if (tid == 0) next_tile = atomicAdd(tile_ctr, 1u);      // claimed now, read at the end of this tile
        const uint32_t t0 = tile * P1_T;
        const uint32_t tlen = min((uint32_t)P1_T, cd.n_bases - t0);      // window starts owned
        const uint32_t blen = min((uint32_t)(P1_T + MAXK + 32), cd.n_bases - t0);  // bytes staged
        // reads holding this tile's window starts: r0 .. r1 (r1 holds the next tile's first position)
        const uint32_t r0 = cd.tile_r0[tile];
        const uint32_t r1 = tile + 1 < cd.n_tiles ? cd.tile_r0[tile + 1] : cd.n_reads - 1;
        const uint32_t n_ro = r1 - r0 + 2;                                 // offsets r0 .. r1 + 1
        const bool ro_lds = n_ro <= P1_RO;

        // 1. stage bytes [t0, t0+blen) in LDS with 16-byte loads (t0 is 16-aligned)
        {
            const uint4* src = reinterpret_cast<const uint4*>(cd.reads + t0);
            uint4* dst = reinterpret_cast<uint4*>(s_bytes);
            const uint32_t nv = blen >> 4;
            for (uint32_t i = tid; i < nv; i += P1_NT) dst[i] = src[i];
            for (uint32_t i = (nv << 4) + tid; i < (uint32_t)(P1_T + MAXK + 32); i += P1_NT)
                s_bytes[i] = (i < blen) ? cd.reads[t0 + i] : (uint8_t)'A';
            for (int i = tid; i < 1024 && i < (int)P; i += P1_NT) s_loc[i] = 0;
            if (tid == 0) s_ovf = 0;
            if (ro_lds)
                for (uint32_t i = tid; i < n_ro; i += P1_NT) s_ro[i] = (uint32_t)(cd.offs[r0 + i] - cd.base);
        }
        __syncthreads();

        // 2. candidate value per p-mer start j: min over the passing forward and
        //    reverse-complement p-mer (M if neither passes), rolled over V starts.
        auto fill_v = [&](int j0, int cnt) {
            MM f = 0, r = 0;
            const int rs = 2 * (p - 1);
            for (int t = 0; t < p - 1; t++) {
                uint32_t c = base_code(s_bytes[j0 + t]);
                f = (f << 2) | c;
                r = (r >> 2) | ((MM)(3u - c) << rs);
            }
            for (int s = 0; s < cnt; s++) {
                uint32_t c = base_code(s_bytes[j0 + s + p - 1]);
                f = ((f << 2) | c) & M;
                r = (r >> 2) | ((MM)(3u - c) << rs);
                const bool fp = pmer_pass(f, p, prm.fmode), rp = pmer_pass(r, p, prm.fmode);
                if constexpr (WIN) {
                    MM val; int sent = 0; int8_t st;
                    if (fp && rp) { val = f < r ? f : r; st = f < r ? 0 : (r < f ? 1 : -1); }
                    else if (fp) { val = f; st = 0; }
                    else if (rp) { val = r; st = 1; }
                    else { val = M; sent = 1; st = -1; }
                    s_v[j0 + s] = (val << CSH) | ((MM)sent << 13) | (MM)(8191 - (j0 + s));
                    s_str[j0 + s] = st;
                } else {
                    MM a = fp ? f : M;
                    MM b = rp ? r : M;
                    s_v[j0 + s] = a < b ? a : b;
                }
            }
        };
        fill_v(tid * P1_V, P1_V);
        {   // the w-1 extra p-mer starts after the tile
            const int extra = w - 1;
            const int j0 = P1_T + tid * P1_V;
            if (tid * P1_V < extra) fill_v(j0, min(P1_V, extra - tid * P1_V));
        }
        __syncthreads();

        // 3. window minimum for this thread's V window starts
        MM mm[P1_V];
        {
            const int b = tid * P1_V;
            if (w >= P1_V) {
                MM vv[P1_V];
#pragma unroll
                for (int s = 0; s < P1_V; s++) vv[s] = s_v[b + s];
                MM mid = vv[P1_V - 1];
                for (int j = b + P1_V; j <= b + w - 1; j++) { MM x = s_v[j]; mid = x < mid ? x : mid; }
                // window s = [b+s, b+s+w-1] = left(s) U [b+V-1, b+w-1] U right(s), where
                // left(s) = vv[s..V-2] (suffix minima) and right(s) = s_v[b+w .. b+w+s-1]
                MM left[P1_V];
                left[P1_V - 1] = MC;
                MM suf = MC;
#pragma unroll
                for (int s = P1_V - 2; s >= 0; s--) { MM x = vv[s]; suf = x < suf ? x : suf; left[s] = suf; }
                MM pre = MC;
#pragma unroll
                for (int s = 0; s < P1_V; s++) {
                    MM a = left[s] < mid ? left[s] : mid;
                    mm[s] = a < pre ? a : pre;
                    if (s + 1 < P1_V) { MM x = s_v[b + w + s]; pre = x < pre ? x : pre; }
                }
            } else {
#pragma unroll
                for (int s = 0; s < P1_V; s++) {
                    MM a = MC;
                    for (int j = 0; j < w; j++) { MM x = s_v[b + s + j]; a = x < a ? x : a; }
                    mm[s] = a;
                }
            }
            s_last[tid] = mm[P1_V - 1] >> CSH;
        }
        __syncthreads();

        // 4. validity and boundaries
        uint32_t bmask = 0;
        uint32_t wend[P1_V];
        {
            const uint32_t g0 = t0 + tid * P1_V;
            uint32_t r = 0;
            uint64_t rbeg = 0, rnext = 0, rend = 0;
            if (tid * P1_V < (int)tlen) {
                const uint32_t gs = max(g0, cd.lead);
                if (ro_lds) {           // largest i with s_ro[i] <= gs, in LDS
                    uint32_t lo = 0, hi = n_ro - 1;
                    while (hi - lo > 1) {
                        const uint32_t mid = (lo + hi) >> 1;
                        if (s_ro[mid] <= gs) lo = mid; else hi = mid;
                    }
                    r = r0 + lo;
                } else {
                    r = find_read(cd.offs, cd.n_reads, cd.base + gs);
                }
                rbeg = ro_lds ? s_ro[r - r0] : cd.offs[r] - cd.base;
                rnext = ro_lds ? s_ro[r - r0 + 1] : cd.offs[r + 1] - cd.base;
                rend = cd.lens ? rbeg + (uint32_t)cd.lens[r] : rnext;
            }
            MM prev = tid > 0 ? s_last[tid - 1] : (MM)0;
            MM cmp[P1_V];
#pragma unroll
            for (int s = 0; s < P1_V; s++) cmp[s] = mm[s];
#pragma unroll
            for (int s = 0; s < P1_V; s++) mm[s] = cmp[s] >> CSH;      // minimizer values
#pragma unroll
            for (int s = 0; s < P1_V; s++) {
                const uint32_t g = g0 + s;
                wend[s] = 0;
                if (tid * P1_V + s < (int)tlen) {
                    while (g >= rnext) {            // advance (skips empty reads)
                        r++;
                        rbeg = rnext;
                        rnext = ro_lds ? s_ro[r - r0 + 1] : cd.offs[r + 1] - cd.base;
                        rend = cd.lens ? rbeg + (uint32_t)cd.lens[r] : rnext;
                    }
                    const bool valid = g >= rbeg && (uint64_t)g + k <= rend;
                    if (valid) {
                        const bool first = (g == rbeg) || (g == t0);
                        if (first || mm[s] != prev) bmask |= 1u << s;
                        uint64_t we = rend - k + 1 - t0;           // tile-local end of windows
                        wend[s] = (uint32_t)min(we, (uint64_t)tlen);
                        if (prm.mm_out) prm.mm_out[g] = (uint32_t)mm[s];
                        if constexpr (WIN) {
                            if (prm.win_mm) {
                                const bool none = (cmp[s] >> 13) & 1;
                                const uint32_t js = 8191u - (uint32_t)(cmp[s] & 8191u);   // tile-local p-mer start
                                prm.win_mm[g] = (uint32_t)mm[s];
                                prm.win_pos[g] = none ? (uint8_t)255 : (uint8_t)(js - (g - t0));
                                prm.win_strand[g] = none ? (int8_t)-1 : s_str[js];
                            }
                        }
                        if (prm.sketch && ((g - rbeg) & 15u) == 0) {
                            uint32_t q = 0;
                            const int jl = g - t0;
#pragma unroll
                            for (int t = 0; t < 4; t++) q = (q << 2) | base_code(s_bytes[jl + t]);
                            atomicAdd(&prm.sketch[(size_t)part_of(mm[s], P, prm.pmode) * 256 + q], 1u);
                        }
                    }
                }
                prev = mm[s];
            }
        }

        // 5. compact the superkmer starts in tile order (wave scans + wave sums)
        {
            const uint32_t c = __popc(bmask);
            const uint32_t lane = __lane_id(), wv = tid / GKC_WAVE;
            uint32_t inc = c;
#pragma unroll
            for (int off = 1; off < GKC_WAVE; off <<= 1) {
                const uint32_t y = __shfl_up(inc, off);
                if ((int)lane >= off) inc += y;
            }
            if (lane == GKC_WAVE - 1) s_scan[wv] = inc;
            __syncthreads();
            uint32_t o = inc - c, tot = 0;
#pragma unroll
            for (int w2 = 0; w2 < P1_NT / GKC_WAVE; w2++) { const uint32_t x = s_scan[w2]; o += (uint32_t)w2 < wv ? x : 0u; tot += x; }
            if (tid == 0) s_ns = tot;
#pragma unroll
            for (int s = 0; s < P1_V; s++) {
                if (bmask & (1u << s)) {
                    s_start[o] = (uint16_t)(tid * P1_V + s);
                    s_wend[o] = (uint16_t)wend[s];
                    s_part[o] = (uint16_t)part_of(mm[s], P, prm.pmode);
                    o++;
                }
            }
        }
        __syncthreads();

        // 6. sizes and per-partition block-local offsets.  One 64-bit LDS atomic
        //    per superkmer reserves its header, k-mer and word slots together, so
        //    the three run in the same order inside a partition.
        const uint32_t ns = s_ns;
        unsigned long long loc[P1_V];
#pragma unroll
        for (int q = 0; q < P1_V; q++) {
            const uint32_t i = tid + q * P1_NT;
            if (i < ns) {
                const uint32_t a = s_start[i];
                uint32_t e = (i + 1 < ns) ? (uint32_t)s_start[i + 1] : tlen;
                e = min(e, (uint32_t)s_wend[i]);
                const unsigned long long n = e - a;
                const unsigned long long nw = (n + k - 1 + 15) >> 4;
                loc[q] = atomicAdd(&s_loc[s_part[i]], (1ull << 48) | (n << 24) | nw);
            }
        }
        __syncthreads();
        // 7. one global reservation per (block, partition)
        for (uint32_t q = tid; q < P; q += P1_NT) {
            const unsigned long long v = s_loc[q];
            const uint32_t h = (uint32_t)(v >> 48), kk = (uint32_t)(v >> 24) & 0xFFFFFFu, ww = (uint32_t)v & 0xFFFFFFu;
            if (h) {
                const uint32_t gh = atomicAdd(&bins.hdr_cur[q], h);
                const uint32_t gw = atomicAdd(&bins.word_cur[q], ww);
                const uint32_t gk = atomicAdd(&bins.kmer_cur[q], kk);
                if (prm.scatter && (gh + h > bins.hdr_cap[q] || gw + ww > bins.word_cap[q])) {
                    s_ovf = 1;
                    atomicOr(bins.overflow, 1u);
                }
                s_loc[q] = ((unsigned long long)gw << 32) | gh;
                s_gk[q] = gk;
            }
        }
        __syncthreads();
        // 8. write headers and 2-bit packed bases
        if (prm.scatter && !s_ovf) {
#pragma unroll
            for (int q = 0; q < P1_V; q++) {
                const uint32_t i = tid + q * P1_NT;
                if (i < ns) {
                    const uint32_t a = s_start[i];
                    uint32_t e = (i + 1 < ns) ? (uint32_t)s_start[i + 1] : tlen;
                    e = min(e, (uint32_t)s_wend[i]);
                    const uint32_t n = e - a;
                    const uint32_t nb = n + k - 1;
                    const uint32_t nw = (nb + 15) >> 4;
                    const uint32_t part = s_part[i];
                    const unsigned long long gb = s_loc[part];
                    const uint32_t wo = (uint32_t)(gb >> 32) + (uint32_t)(loc[q] & 0xFFFFFFu);
                    SkmHdr h;
                    h.word_off = wo;
                    h.kmer_off = s_gk[part] + (uint32_t)((loc[q] >> 24) & 0xFFFFFFu);
                    h.n_kmers = n;
                    h.reserved = 0;
                    bins.hdrs[bins.hdr_base[part] + (uint32_t)gb + (uint32_t)(loc[q] >> 48)] = h;
                    uint32_t* dst = bins.words + bins.word_base[part] + wo;
                    for (uint32_t wi = 0; wi < nw; wi++) {
                        uint32_t x = 0;
                        const uint32_t b0 = a + wi * 16;
#pragma unroll
                        for (int t = 0; t < 16; t++) {
                            uint32_t c = (b0 + t < a + nb) ? base_code(s_bytes[b0 + t]) : 0u;
                            x |= c << (30 - 2 * t);
                        }
                        dst[wi] = x;
                    }
                }
            }
        }
        if (tid == 0) s_tile = next_tile;
        __syncthreads();
    }
}

// Homopolymer compression (gkc_cuda.cu:178-190), one wave per read: keep
// byte j when it differs from byte j-1; out[beg + h] = byte of run h,
// orig_pos[beg + h] = in-read index of the run's first byte, lens[r] = runs.
__global__ __launch_bounds__(256) void k_hpc(const uint8_t* __restrict__ reads, const uint64_t* __restrict__ offs,
                                             uint32_t n_reads, uint8_t* __restrict__ out,
                                             int32_t* __restrict__ lens, int32_t* __restrict__ orig_pos) {
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) / GKC_WAVE;
    const uint32_t lane = __lane_id();
    const uint32_t nwaves = gridDim.x * blockDim.x / GKC_WAVE;
    const uint64_t base = offs[0];
    for (uint32_t r = wave; r < n_reads; r += nwaves) {
        const uint64_t beg = offs[r] - base, L = offs[r + 1] - offs[r];
        uint32_t cnt = 0;
        for (uint64_t j0 = 0; j0 < L; j0 += GKC_WAVE) {
            const uint64_t j = j0 + lane;
            bool keep = false;
            uint8_t c = 0;
            if (j < L) {
                c = reads[beg + j];
                keep = (j == 0) || (c != reads[beg + j - 1]);
            }
            const uint64_t m = __ballot(keep);
            if (keep) {
                const uint32_t h = cnt + __popcll(m & lanemask_lt());
                out[beg + h] = c;
                orig_pos[beg + h] = (int32_t)j;
            }
            cnt += __popcll(m);
        }
        if (lane == 0) lens[r] = (int32_t)cnt;
    }
}

// Superkmer offset lists in the reference's layout (GPU_GenSKM,
// gkc_cuda.cu:308-322): at each read's batch-local start,
// [0, i1, i2, ..., len-k+1, 0] with the i where mm(i) != mm(i-1).
// One wave per read; reads shorter than k get [0, 0] (when they have room).
__global__ __launch_bounds__(256) void k_skm_lists(const uint64_t* __restrict__ offs, const int32_t* __restrict__ lens,
                                                   uint32_t n_reads, int k, const uint32_t* __restrict__ mm,
                                                   int32_t* __restrict__ skm) {
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) / GKC_WAVE;
    const uint32_t lane = __lane_id();
    const uint32_t nwaves = gridDim.x * blockDim.x / GKC_WAVE;
    const uint64_t base = offs[0];
    for (uint32_t r = wave; r < n_reads; r += nwaves) {
        const uint64_t beg = offs[r] - base;
        const int64_t L = lens ? (int64_t)lens[r] : (int64_t)(offs[r + 1] - offs[r]);
        if (L < k) {
            if (lane == 0 && L >= 1) skm[beg] = 0;
            if (lane == 1 && L >= 2) skm[beg + 1] = 0;
            continue;
        }
        const int64_t W = L - k + 1;
        uint32_t c = 0;
        if (lane == 0) skm[beg] = 0;
        for (int64_t i0 = 1; i0 < W; i0 += GKC_WAVE) {
            const int64_t i = i0 + lane;
            const bool b = i < W && mm[beg + i] != mm[beg + i - 1];
            const uint64_t m = __ballot(b);
            if (b) skm[beg + 1 + c + __popcll(m & lanemask_lt())] = (int32_t)i;
            c += __popcll(m);
        }
        if (lane == 0) {
            skm[beg + c + 1] = (int32_t)W;
            if ((int64_t)c + 2 < L) skm[beg + c + 2] = 0;
        }
    }
}

template __global__ void k_superkmers<uint32_t, false>(const ChunkDesc*, P1Params, BinSet, uint32_t*);
template __global__ void k_superkmers<uint64_t, false>(const ChunkDesc*, P1Params, BinSet, uint32_t*);
template __global__ void k_superkmers<uint64_t, true>(const ChunkDesc*, P1Params, BinSet, uint32_t*);

}  // namespace gkc
