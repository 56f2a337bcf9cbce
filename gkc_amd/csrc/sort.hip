// Phase 2 of the k-mer counting path on gfx950: k-mer extraction from the
// 2-bit superkmer bins (with the onesweep digit histograms of every pass
// fused in), a segmented onesweep LSD radix sort (LDS ranking by wave64
// ballots, decoupled look-back across workgroups), and a fused run-length
// count + threshold pass that writes the (k-mer, count) table.
//
// Reference semantics followed (file:line in the reference tree):
//   forward k-mers of each superkmer, rolling         kmer_counting.cu:95-113
//   canonical = min(x, revcomp(x))                    kmer_counting.cu:33-44
//   ascending sort                                    kmer_counting.cu:134
//   run-length count, clamp to 0xFFFFFFFF, [min,max]  kmer_counting.cu:148-191
// A segment is one partition: every kernel here works on all segments of a
// plan at once (the plain path passes a plan with a single segment).
#include <type_traits>

#include "kernels.hpp"

namespace gkc {

// Segment owning global work index t (largest s with start[s] <= t).
__device__ __forceinline__ uint32_t find_seg(const uint32_t* start, uint32_t n_segs, uint32_t t) {
    uint32_t lo = 0, hi = n_segs;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (start[mid] <= t) lo = mid; else hi = mid;
    }
    return lo;
}

// Claim the next unit of work of kind KIND (1: run-length tile of the plain
// count) from the work lists; run by one thread.  `list` is the
// workgroup's current list (advanced when a list is empty).  Returns false when
// every list is exhausted.
template <int KIND>
__device__ __forceinline__ bool claim(const WorkLists& wl, uint32_t* ctr, uint32_t& list, uint32_t& seg, uint32_t& idx) {
    const uint32_t n = wl.n_segs;
    for (uint32_t tries = 0; tries < wl.n_lists; tries++) {
        const uint32_t L = list;
        const uint32_t tot = wl.w[KIND * 8 + L];
        if (tot) {
            const uint32_t t = atomicAdd(&ctr[L], 1u);
            if (t < tot) {
                const uint32_t* off = wl.w + 33 + n + KIND * n;
                uint32_t lo = wl.w[24 + L], hi = wl.w[24 + L + 1];   // order indices of list L
                while (hi - lo > 1) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (off[mid] <= t) lo = mid; else hi = mid;
                }
                seg = wl.w[33 + lo];
                idx = t - off[lo];
                return true;
            }
        }
        list = (L + 1) % wl.n_lists;
    }
    return false;
}

// Claim the next sort tile (run by one thread): its index g in the lists'
// concatenation (the list's first tile + the list counter's value), which
// indexes its descriptor and its look-back slot.  The counter's atomic is
// issued together with the loads of the list's size and base.  Returns false
// when every list is exhausted.
__device__ __forceinline__ bool claim_tile(const WorkLists& wl, uint32_t* ctr, uint32_t& list, uint32_t& g) {
    const uint32_t* base = wl.w + 33 + 4 * wl.n_segs;
    for (uint32_t tries = 0; tries < wl.n_lists; tries++) {
        const uint32_t L = list;
        const uint32_t t = atomicAdd(&ctr[L], 1u);
        const uint32_t tot = wl.w[L], b = base[L];
        if (t < tot) { g = b + t; return true; }
        list = (L + 1) % wl.n_lists;
    }
    return false;
}

// Sort tile descriptors in work-list order (one workgroup per list-order
// segment): tile g = its list's first tile + the segment's tile offset in the
// list + its index in the segment.  Tiles hold tile_keys keys, the last one
// of a segment the rest.  Also the segment of each extraction item (segment
// order, item_start from k_plan).
__global__ __launch_bounds__(256) void k_tile_desc(SegArrays sa, WorkLists wl, int tile_keys, TileDesc* __restrict__ desc,
                                                   uint32_t* __restrict__ iseg) {
    const uint32_t n = wl.n_segs;
    for (uint32_t pos = blockIdx.x; pos < n; pos += gridDim.x) {
        uint32_t L = 0;
        while (L < 8 && wl.w[24 + L + 1] <= pos) L++;          // the list holding order index pos
        const uint32_t seg = wl.w[33 + pos];
        const uint32_t g0 = wl.w[33 + 4 * n + L] + wl.w[33 + n + pos];
        const uint32_t nk = sa.seg_keys[seg];
        const uint32_t nt = (uint32_t)(((uint64_t)nk + tile_keys - 1) / tile_keys);
        const uint64_t kb = sa.key_start[seg];
        for (uint32_t ti = threadIdx.x; ti < nt; ti += blockDim.x)
            desc[g0 + ti] = TileDesc{kb, ti, (seg << 16) | min((uint32_t)tile_keys, nk - ti * (uint32_t)tile_keys)};
        const uint32_t i0 = sa.item_start[seg], i1 = sa.item_start[seg + 1];
        for (uint32_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) iseg[i] = seg;
    }
}

// Build the segment layout from per-segment key and header counts (one block).
__global__ __launch_bounds__(1024) void k_plan(Plan* plan, SegArrays sa, int sort_tile_keys, int rl_tile_keys) {
    __shared__ uint64_t s_k[1024];
    __shared__ uint32_t s_i[1024], s_t[1024], s_r[1024];
    const uint32_t S = plan->n_segs;
    uint64_t accK = 0; uint32_t accI = 0, accT = 0, accR = 0;
    for (uint32_t base = 0; base < S; base += 1024) {
        const uint32_t s = base + threadIdx.x;
        uint64_t nk = 0; uint32_t ni = 0, nt = 0, nr = 0;
        if (s < S) {
            nk = sa.seg_keys[s];
            ni = (sa.seg_hdrs[s] + EX_HDRS - 1) / EX_HDRS;
            nt = (uint32_t)((nk + sort_tile_keys - 1) / sort_tile_keys);
            nr = (uint32_t)((nk + rl_tile_keys - 1) / rl_tile_keys);
        }
        s_k[threadIdx.x] = nk; s_i[threadIdx.x] = ni; s_t[threadIdx.x] = nt; s_r[threadIdx.x] = nr;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            uint64_t a = threadIdx.x >= off ? s_k[threadIdx.x - off] : 0;
            uint32_t b = threadIdx.x >= off ? s_i[threadIdx.x - off] : 0;
            uint32_t c = threadIdx.x >= off ? s_t[threadIdx.x - off] : 0;
            uint32_t d = threadIdx.x >= off ? s_r[threadIdx.x - off] : 0;
            __syncthreads();
            s_k[threadIdx.x] += a; s_i[threadIdx.x] += b; s_t[threadIdx.x] += c; s_r[threadIdx.x] += d;
            __syncthreads();
        }
        if (s < S) {
            sa.key_start[s] = accK + s_k[threadIdx.x] - nk;
            sa.item_start[s] = accI + s_i[threadIdx.x] - ni;
            sa.sort_tile_start[s] = accT + s_t[threadIdx.x] - nt;
            sa.rl_tile_start[s] = accR + s_r[threadIdx.x] - nr;
        }
        accK += s_k[1023]; accI += s_i[1023]; accT += s_t[1023]; accR += s_r[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        sa.key_start[S] = accK; sa.item_start[S] = accI;
        sa.sort_tile_start[S] = accT; sa.rl_tile_start[S] = accR;
        plan->tot_keys = accK; plan->tot_items = accI;
        plan->tot_sort_tiles = accT; plan->tot_rl_tiles = accR;
    }
}

// Small parameter stores done by kernels (kernel arguments are captured at
// launch, so no host buffer has to outlive an asynchronous copy).
__global__ void k_set_u64(uint64_t* p, uint64_t v) { *p = v; }
__global__ void k_set_u32(uint32_t* p, uint32_t v) { *p = v; }
__global__ void k_plan_init(Plan* plan, uint32_t n_segs, uint32_t n_pass, uint32_t sh0, uint32_t dbits) {
    const int t = threadIdx.x;
    if (t == 0) { plan->n_segs = n_segs; plan->n_pass = n_pass; plan->tot_items = 0; plan->tot_sort_tiles = 0;
                  plan->tot_rl_tiles = 0; plan->sh0 = sh0; plan->dbits = dbits; plan->tot_keys = 0; }
    for (int i = t; i < NCTR; i += blockDim.x) plan->ctr[i] = 0;
}
__global__ void k_set_chunk(ChunkDesc* dst, ChunkDesc v) { *dst = v; }

// Zero the work counters of a plan and the histograms (fill kernel, so the
// sequence stays capturable in a graph).
__global__ void k_zero32(uint32_t* p, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) p[i] = 0;
}

// Extract the canonical k-mers of each superkmer into the segment's key
// range and accumulate the digit histograms of all passes (pass p's digit is
// key bits [sh0 + d p, sh0 + d p + d), d = plan->dbits; each pass's histogram
// keeps RADIX slots, the first 2^d used).
// Bases are 2-bit, 16 per u32 word, MSB first.
//
// A lane takes a chunk of R consecutive k-mers of one superkmer.  The
// chunk's window of words is read once (4 words for 64-bit keys, 6 for
// 128-bit ones: every word a chunk may reach, whether or not this one reaches
// it; the word buffers end with 8 words of padding); each forward k-mer is a
// constant shift of the window, the first reverse complement is computed in
// full and each following one is the previous one shifted down one base with
// the complement of the base entering the forward k-mer put in at the top.
// k-mers per chunk.  128-bit keys take one: their extraction is bound by
// the key writes (16 bytes a k-mer), which the LDS staging of multi-k-mer
// chunks' keys only slows (more LDS per workgroup, fewer waves).
template <typename K> constexpr int ex_r() { return sizeof(K) == 8 ? 4 : 1; }
// blocks of 64 chunks per wave iteration
template <typename K> constexpr int ex_u() { return 2; }
constexpr int EX_SW = 512;   // staged words per wave
constexpr int EX_SWP = 8;    // zero words after the staged ones (window reads past the last superkmer)
constexpr uint32_t EX_B = 8; // items per claimed batch

template <typename K> struct ExChunk;

template <> struct ExChunk<uint64_t> {
    // window: bases [o0, o0 + 64 - (o0 & 15)) as hi:lo, base o0 at the top of
    // hi; k-mer r takes bases [o0 + r, o0 + r + k), within the window for
    // R + k <= 50
    template <int R, typename W>
    __device__ static __forceinline__ void kmers(const W* wds, uint32_t o0, int k, uint64_t (&c)[R]) {
        const uint32_t wi = o0 >> 4, sh = (o0 & 15) * 2;   // sh <= 30
        const uint32_t w0 = wds[wi], w1 = wds[wi + 1], w2 = wds[wi + 2], w3 = wds[wi + 3];
        const uint64_t a = ((uint64_t)w0 << 32) | w1, b = ((uint64_t)w2 << 32) | w3;
        const uint64_t hi = (a << sh) | ((b >> 1) >> (63 - sh));    // (v >> 1) >> (63 - sh) == v >> (64 - sh), 0 for sh = 0
        const uint64_t lo = b << sh;
        const int ks = 64 - 2 * k;
        uint64_t f = hi >> ks;
        uint64_t rc = revcomp64(f, k);
        c[0] = rc < f ? rc : f;
#pragma unroll
        for (int r = 1; r < R; r++) {
            f = ((hi << (2 * r)) | (lo >> (64 - 2 * r))) >> ks;
            rc = (rc >> 2) | ((uint64_t)(~(uint32_t)f & 3u) << (2 * k - 2));
            c[r] = rc < f ? rc : f;
        }
    }
};

template <> struct ExChunk<u128k> {
    // the smaller of the two, word by word (a select of whole structs is
    // lowered to an indexed private array)
    __device__ static __forceinline__ void pick(u128k& c, const u128k& rc, const u128k& f) {
        const bool lt = less128(rc, f);
        c.lo = lt ? rc.lo : f.lo;
        c.hi = lt ? rc.hi : f.hi;
    }
    // top 2k bits of t_hi:t_lo (s = 128 - 2k, wave-uniform)
    __device__ static __forceinline__ u128k top(uint64_t th, uint64_t tl, int s) {
        u128k x;
        if (s == 0) { x.hi = th; x.lo = tl; }
        else if (s < 64) { x.lo = (tl >> s) | (th << (64 - s)); x.hi = th >> s; }
        else { x.lo = th >> (s - 64); x.hi = 0; }
        return x;
    }
    // window: bases [o0, o0 + 96 - (o0 & 15)) as x0:x1:x2, base o0 at the top
    // of x0; within the window for R + k <= 82
    template <int R, typename W>
    __device__ static __forceinline__ void kmers(const W* wds, uint32_t o0, int k, u128k (&c)[R]) {
        const uint32_t wi = o0 >> 4, sh = (o0 & 15) * 2, rs = 63 - sh;
        uint64_t x0 = ((uint64_t)wds[wi] << 32) | wds[wi + 1];
        uint64_t x1 = ((uint64_t)wds[wi + 2] << 32) | wds[wi + 3];
        uint64_t x2 = ((uint64_t)wds[wi + 4] << 32) | wds[wi + 5];
        x0 = (x0 << sh) | ((x1 >> 1) >> rs);
        x1 = (x1 << sh) | ((x2 >> 1) >> rs);
        x2 = x2 << sh;
        const int s = 128 - 2 * k;
        const int ins = 2 * k - 2;                       // bit of the reverse complement's top base
        u128k f = top(x0, x1, s);
        u128k rc = revcomp128(f, k);
        pick(c[0], rc, f);
#pragma unroll
        for (int r = 1; r < R; r++) {
            f = top((x0 << (2 * r)) | (x1 >> (64 - 2 * r)), (x1 << (2 * r)) | (x2 >> (64 - 2 * r)), s);
            const uint64_t cb = ~f.lo & 3u;
            rc.lo = (rc.lo >> 2) | (rc.hi << 62);
            rc.hi >>= 2;
            if (ins >= 64) rc.hi |= cb << (ins - 64);
            else rc.lo |= cb << ins;
            pick(c[r], rc, f);
        }
    }
};

// Work item = EX_HDRS consecutive headers of one segment.  Workgroups claim
// batches of EX_B consecutive items (segment order) from one counter, and
// find a batch's segment in the per-item segment table (k_tile_desc); a
// batch's next claim is taken while it is extracted.  Each wave takes 64
// headers of an item, prefix-sums their chunk counts and strides its lanes
// over the combined chunks.  The wave first copies its superkmers' words into
// LDS (each lane its own superkmer's few words, all its loads issued before
// the first wait), so the chunk windows are LDS reads rather than gathers
// from L2; a wave whose superkmers hold more than EX_SW words (long
// superkmers) reads them from global memory instead.  The next item's headers
// are loaded while this one is extracted, and the waves of a workgroup meet
// only at batch ends and where the segment changes (the histogram flush): all
// other LDS is per wave.
template <typename K>
__global__ __launch_bounds__(SORT_NT) void k_extract(Plan* __restrict__ plan, SegArrays sa, const uint32_t* __restrict__ iseg, const SkmHdr* __restrict__ hdrs,
                                                     const uint64_t* __restrict__ hdr_base,
                                                     const uint32_t* __restrict__ words,
                                                     const uint64_t* __restrict__ word_base, K* __restrict__ keys, int k) {
    constexpr int NWAVE = SORT_NT / GKC_WAVE;
    constexpr int EX_U = ex_u<K>();
    constexpr int EX_R = ex_r<K>();
    extern __shared__ uint32_t s_hist[];                // [n_pass * RADIX] (dynamic LDS)
    __shared__ uint4 s_hi[NWAVE][GKC_WAVE];             // per header: k-mer offset, word base, first chunk, k-mers
    __shared__ uint32_t s_flag[NWAVE][EX_U * GKC_WAVE]; // 1 + header whose first chunk is at each position of a block
    __shared__ uint32_t s_words[NWAVE][EX_SW + EX_SWP]; // the wave's superkmer words (+ read-past padding)
    __shared__ K s_keys[NWAVE][EX_R * GKC_WAVE];        // a block's keys in chunk order
    __shared__ uint2 s_cd[NWAVE][GKC_WAVE];             // a block's chunks: first key (segment offset), k-mers
    const int npass = plan->n_pass;
    const int sh0 = plan->sh0;
    const int db = plan->dbits;
    const uint32_t lane = __lane_id(), wv = threadIdx.x / GKC_WAVE;
    // (the item and segment values are read as wave-uniform, so the control
    // flow around the workgroup barriers below is uniform)
    auto u32 = [](uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane(v); };
    auto u64 = [&](uint64_t v) { return ((uint64_t)u32((uint32_t)(v >> 32)) << 32) | u32((uint32_t)v); };
    // a segment's values, loaded once per segment: header count, first and
    // end item, first header, word and key
    struct SegV { uint32_t nh, start, end; uint64_t hb, wb, kb; };
    auto seg_vals = [&](uint32_t s) {
        return SegV{u32(sa.seg_hdrs[s]), u32(sa.item_start[s]), u32(sa.item_start[s + 1]), u64(hdr_base[s]), u64(word_base[s]),
                    u64(sa.key_start[s])};
    };
    // an item's headers; past the segment's last header a valid load (its
    // first), marked empty
    auto load_hdr = [&](const SegV& v, uint32_t ix, bool& valid) {
        const uint32_t h = ix * EX_HDRS + threadIdx.x;
        valid = h < v.nh;
        return hdrs[v.hb + (valid ? h : 0u)];
    };
    const uint32_t T = plan->tot_items;
    uint32_t* ctr = &plan->ctr[ctr_extract()];
    __shared__ uint32_t s_claim[2];                      // next batch: first item (>= T: none), its segment
    if (threadIdx.x == 0) {
        const uint32_t g0 = atomicAdd(ctr, 1u) * EX_B;
        s_claim[0] = g0;
        s_claim[1] = iseg[min(g0, T - 1)];
    }
    for (int i = threadIdx.x; i < npass * RADIX; i += SORT_NT) s_hist[i] = 0;
    __syncthreads();                                     // histograms cleared, first claim out
    uint32_t g = u32(s_claim[0]);
    uint32_t seg = u32(s_claim[1]);
    __syncthreads();                                     // the claim read before the batch end rewrites it
    if (g >= T) return;
    SegV cs = seg_vals(seg);
    uint32_t idx = g - cs.start;
    bool hv;
    SkmHdr hd = load_hdr(cs, idx, hv);
    for (;;) {                                           // batches
        const uint32_t gend = min(T, g + EX_B);
        uint32_t nb = 0, nbseg = 0;                      // the next batch's claim (thread 0)
        if (threadIdx.x == 0) nb = atomicAdd(ctr, 1u);
        bool asked = false;
        for (;;) {                                       // items [g, gend)
            // the next item of the batch (the current one again after its last, not used)
            const bool more = g + 1 < gend;
            uint32_t nseg = seg, nidx = idx + (more ? 1u : 0u);
            SegV ns = cs;
            if (more && g + 1 == cs.end) {
                nseg = seg + 1;
                while (u32(sa.item_start[nseg + 1]) <= g + 1) nseg++;
                nidx = 0;
                ns = seg_vals(nseg);
            }
            const uint32_t n = hv ? hd.n_kmers : 0u;
            const uint32_t* wseg = words + cs.wb;
            K* kseg = keys + cs.kb;
            // stage the wave's words: header h's words [0, nws) (the bases of all
            // its k-mers) at s_words[wv][wst]
            const uint32_t nws = n ? ((n + (uint32_t)k - 2) >> 4) + 1 : 0u;
            uint32_t winc = nws;
            const uint32_t nch = (n + EX_R - 1) / EX_R;
            uint32_t cinc = nch;
#pragma unroll
            for (int off = 1; off < GKC_WAVE; off <<= 1) {
                const uint32_t y = __shfl_up(winc, off), z = __shfl_up(cinc, off);
                if ((int)lane >= off) { winc += y; cinc += z; }
            }
            const bool staged = __shfl(winc, GKC_WAVE - 1) <= (uint32_t)EX_SW;       // wave-uniform
            const uint32_t Ct = __shfl(cinc, GKC_WAVE - 1);                            // the wave's chunks
            const uint32_t wst = winc - nws, cpre = cinc - nch;
            if (staged) {
                const uint32_t* src = wseg + hd.word_off;
                for (uint32_t j0 = 0; j0 < nws; j0 += 8) {
                    uint32_t v[8];
#pragma unroll
                    for (int t = 0; t < 8; t++) v[t] = src[j0 + t];      // past the superkmer: its segment's next words or the padding
#pragma unroll
                    for (int t = 0; t < 8; t++)
                        if (j0 + t < nws) s_words[wv][wst + j0 + t] = v[t];
                }
                if (lane == GKC_WAVE - 1)
#pragma unroll
                    for (int t = 0; t < EX_SWP; t++) s_words[wv][winc + t] = 0;
            }
            s_hi[wv][lane] = make_uint4(hd.kmer_off, staged ? wst : hd.word_off, cpre, n);
            wave_lds_fence();
            // the claim has returned by now (the staging waited on it): look up
            // its segment, read at the batch end
            if (threadIdx.x == 0 && !asked) {
                asked = true;
                nbseg = iseg[min(nb * EX_B, T - 1)];
            }
            bool hnv;
            const SkmHdr hn = load_hdr(ns, nidx, hnv);
            // blocks of EX_U * 64 chunks, chunk qb + 64u + lane on lane `lane`:
            // each header with chunks flags its first one in LDS; a chunk's header
            // is the last flag at or before it (a ballot and a lane shuffle), or
            // the header carried from the previous positions
            uint32_t cur = 0;
            for (uint32_t qb = 0; qb < Ct; qb += EX_U * GKC_WAVE) {
#pragma unroll
                for (int u = 0; u < EX_U; u++) s_flag[wv][u * GKC_WAVE + lane] = 0;
                wave_lds_fence();    // other lanes' clears
                if (nch && cpre >= qb && cpre < qb + EX_U * GKC_WAVE) s_flag[wv][cpre - qb] = lane + 1;
                wave_lds_fence();    // other lanes' flags
                uint32_t hj[EX_U];
#pragma unroll
                for (int u = 0; u < EX_U; u++) {
                    const uint32_t f = s_flag[wv][u * GKC_WAVE + lane];
                    const uint64_t M = __ballot(f != 0);
                    const uint64_t le = M & (lanemask_lt() | (1ull << lane));
                    const uint32_t fs = __shfl(f, le ? 63 - __builtin_clzll(le) : 0);
                    hj[u] = le ? fs - 1 : cur;
                    if (M) cur = __shfl(f, 63 - __builtin_clzll(M)) - 1;
                }
                K c[EX_U][EX_R];
                uint2 cd[EX_U];                               // chunk's first key (segment offset), k-mers
#pragma unroll
                for (int u = 0; u < EX_U; u++) {
                    const uint32_t q = qb + u * GKC_WAVE + lane;
                    const uint4 h = s_hi[wv][hj[u]];
                    const bool ok = q < Ct;                   // past Ct: the header's first chunk, read but not stored
                    const uint32_t o0 = ok ? (q - h.z) * EX_R : 0u;
                    cd[u] = make_uint2(h.x + o0, ok ? min((uint32_t)EX_R, h.w - o0) : 0u);
                    if (staged) ExChunk<K>::template kmers<EX_R>(&s_words[wv][h.y], o0, k, c[u]);
                    else ExChunk<K>::template kmers<EX_R>(wseg + h.y, o0, k, c[u]);
                }
                // stores through LDS: lane l of store s takes key l % EX_R of
                // chunk (64 s + l) / EX_R, so each store covers a run of
                // consecutive keys (chunk-strided lane stores would write each
                // line in EX_R partial pieces); one-k-mer chunks store directly
                if constexpr (EX_R == 1) {
#pragma unroll
                    for (int u = 0; u < EX_U; u++) {
                        if (cd[u].y) {
                            kseg[cd[u].x] = c[u][0];
                            for (int ps = 0; ps < npass; ps++)
                                atomicAdd(&s_hist[ps * RADIX + KeyTraits<K>::digit_n(c[u][0], sh0 + ps * db, db)], 1u);
                        }
                    }
                    continue;
                }
#pragma unroll
                for (int u = 0; u < EX_U; u++) {
                    s_cd[wv][lane] = cd[u];
#pragma unroll
                    for (int r = 0; r < EX_R; r++) s_keys[wv][lane * EX_R + r] = c[u][r];
                    wave_lds_fence();
#pragma unroll
                    for (int st = 0; st < EX_R; st++) {
                        const uint32_t slot = st * GKC_WAVE + lane, r = lane % EX_R;
                        const uint2 d = s_cd[wv][slot / EX_R];
                        if (r < d.y) {
                            const K key = s_keys[wv][slot];
                            kseg[d.x + r] = key;
                            for (int ps = 0; ps < npass; ps++)
                                atomicAdd(&s_hist[ps * RADIX + KeyTraits<K>::digit_n(key, sh0 + ps * db, db)], 1u);
                        }
                    }
                    wave_lds_fence();                         // read before the next block's writes
                }
            }
            wave_lds_fence();                                 // this item's LDS reads before the next item's writes
            if (!more) break;
            // the histograms go out when the segment changes
            if (nseg != seg) {
                __syncthreads();                             // every wave's k-mers of the segment counted
                uint32_t* gh = sa.hist + (size_t)seg * MAX_PASS * RADIX;
                for (int i = threadIdx.x; i < npass * RADIX; i += SORT_NT)
                    if (s_hist[i]) atomicAdd(&gh[i], s_hist[i]);
                __syncthreads();                             // flushed before the histograms are cleared
                for (int i = threadIdx.x; i < npass * RADIX; i += SORT_NT) s_hist[i] = 0;
                __syncthreads();                             // cleared before the next segment's counts
            }
            g++;
            seg = nseg;
            idx = nidx;
            cs = ns;
            hd = hn;
            hv = hnv;
        }
        // batch done: the next one, and the histograms out if its segment differs
        if (threadIdx.x == 0) {
            s_claim[0] = nb * EX_B;
            s_claim[1] = nbseg;
        }
        __syncthreads();                                 // every wave's k-mers of the batch counted; the claim out
        const uint32_t ng = u32(s_claim[0]), ngseg = u32(s_claim[1]);
        const bool last = ng >= T;
        if (last || ngseg != seg) {
            uint32_t* gh = sa.hist + (size_t)seg * MAX_PASS * RADIX;
            for (int i = threadIdx.x; i < npass * RADIX; i += SORT_NT)
                if (s_hist[i]) atomicAdd(&gh[i], s_hist[i]);
            if (!last) {
                __syncthreads();                         // flushed before the histograms are cleared
                for (int i = threadIdx.x; i < npass * RADIX; i += SORT_NT) s_hist[i] = 0;
            }
        }
        if (last) break;
        __syncthreads();                                 // the claim read (and the histograms cleared) before the next batch
        if (ngseg != seg) {
            seg = ngseg;
            cs = seg_vals(seg);
        }
        g = ng;
        idx = g - cs.start;
        hd = load_hdr(cs, idx, hv);
    }
}

// In-place exclusive scan of every (segment, pass) histogram; one wave each.
__global__ __launch_bounds__(256) void k_hist_scan(Plan* __restrict__ plan, uint32_t* __restrict__ hist) {
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) / GKC_WAVE;
    const uint32_t lane = __lane_id();
    const uint32_t nw = gridDim.x * blockDim.x / GKC_WAVE;
    const uint32_t total = plan->n_segs * MAX_PASS;
    for (uint32_t w = wave; w < total; w += nw) {
        if ((w % MAX_PASS) >= plan->n_pass) continue;
        uint32_t* hp = hist + (size_t)w * RADIX;
        uint32_t v[RADIX / GKC_WAVE];
        uint32_t sum = 0;
#pragma unroll
        for (int i = 0; i < RADIX / GKC_WAVE; i++) { v[i] = hp[lane * (RADIX / GKC_WAVE) + i]; sum += v[i]; }
        uint32_t inc = sum;
#pragma unroll
        for (int off = 1; off < GKC_WAVE; off <<= 1) {
            uint32_t y = __shfl_up(inc, off);
            if ((int)lane >= off) inc += y;
        }
        uint32_t run = inc - sum;
#pragma unroll
        for (int i = 0; i < RADIX / GKC_WAVE; i++) { hp[lane * (RADIX / GKC_WAVE) + i] = run; run += v[i]; }
    }
}

// register target of the onesweep pass: waves per SIMD (64-bit keys, no
// prefetch; 128-bit keys, next tile prefetched in registers).  A 512-thread
// workgroup with a tile of up to 6144 keys (the LDS of a 256-thread one, three
// workgroups per CU) is held to 80 VGPRs (6 waves per SIMD); with 8192 or
// more keys (runs of 32 or more keys per digit on the way out, two
// workgroups per CU) to 128 VGPRs (4 waves per SIMD).
template <typename K, int NT, int ITEMS> constexpr int onesweep_waves() {
    return NT == 512 ? (NT * ITEMS > 6144 ? 4 : 6) : sizeof(K) == 16 ? 3 : 4;
}

// Lanes of the wave whose item has this lane's digit dd (digits >= RD mark
// no item and match nobody): the valid lanes that agree with dd on each of
// its DB bits, one ballot per bit.  Returned as 32-bit halves.
template <int DB>
__device__ __forceinline__ void wave_peers(uint32_t dd, uint32_t RD, uint32_t& plo, uint32_t& phi) {
    const uint64_t v = __ballot(dd < RD);
    uint32_t alo = (uint32_t)v, ahi = (uint32_t)(v >> 32);
#pragma unroll
    for (int b = 0; b < DB; b++) {
        int32_t bm = __builtin_amdgcn_sbfe((int32_t)dd, b, 1);     // 0 or ~0: this lane's bit b
        asm("" : "+v"(bm));      // the ballot tests bm itself: one compare per bit
        const uint64_t bal = __ballot(bm != 0);
        alo &= ~((uint32_t)bal ^ (uint32_t)bm);
        ahi &= ~((uint32_t)(bal >> 32) ^ (uint32_t)bm);
    }
    plo = alo; phi = ahi;
}

// One onesweep pass over all segments, DB-bit digits (RD = 2^DB <= NT
// digits, one per thread of the first RD).  Tile = NT * ITEMS keys of one
// segment; tiles are claimed in order from the work lists (one atomic, then
// the tile's descriptor), so every tile's predecessors are claimed by running
// workgroups (forward progress of the look-back).  Per tile (claimed, then
// its keys loaded at once: a tile claimed ahead of time would publish its
// aggregate later and hold up its successors' look-back, which costs more
// than the load latency the other workgroups on the CU hide):
//   1. digits and the tile's digit histogram (LDS atomics, no returns)
//   2. publish the aggregate at once, so successors can look back early
//   4. stable rank within each wave in (item, lane) order: lanes with equal
//      digits are matched with DB ballots, the running per-wave digit
//      counters live in LDS
//   5. decoupled look-back per digit across the segment's earlier tiles
//   6. reorder the tile in LDS by digit and write each digit's run out
// The per-wave counters become each wave's tile positions per digit (digit
// start + earlier waves' counts), and a digit's output base is kept as a
// 32-bit segment offset less its tile start, so a key at tile position j
// goes to segment index base + j.
template <typename K, int ITEMS, int NT, int DB>
__global__ __launch_bounds__(NT, (onesweep_waves<K, NT, ITEMS>())) void k_onesweep(Plan* __restrict__ plan, SegArrays sa, WorkLists wl,
                                                      const TileDesc* __restrict__ tdesc, const K* __restrict__ in,
                                                      K* __restrict__ out, uint32_t* __restrict__ lookback, int pass) {
    constexpr int RD = 1 << DB;                    // digits of this pass
    static_assert(DB >= 6 && DB <= RADIX_BITS && NT % RD == 0, "one digit per thread of whole waves");
    constexpr int TILE = NT * ITEMS;
    static_assert(TILE < 65536, "keys of a tile fit the descriptor's 16 bits");
    constexpr int NWAVE = NT / GKC_WAVE;           // ranking waves
    constexpr int DWAVE = RD / GKC_WAVE;           // waves holding one digit per thread
    // per-wave digit counters, then tile positions (< TILE): 16 bits in
    // 512-thread workgroups, whose tiles take all but a few KB of their LDS
    using whist_t = typename std::conditional<(NT >= 512), uint16_t, uint32_t>::type;
    __shared__ K s_keys[TILE];
    __shared__ whist_t s_whist[NWAVE][RD];
    __shared__ uint32_t s_hist[RD];
    __shared__ uint32_t s_wsum[DWAVE];
    __shared__ uint32_t s_gbase[RD];
    __shared__ uint32_t s_g, s_ok;
    const int sh = (int)plan->sh0 + pass * DB;
    const uint32_t lane = __lane_id(), wv = threadIdx.x / GKC_WAVE;
    const uint32_t d = threadIdx.x;     // threads below RD: one digit each
    const bool dg = NT == RD || d < (uint32_t)RD;
    uint32_t* ctr = &plan->ctr[ctr_pass(pass)];
    uint32_t list = blockIdx.x % wl.n_lists;
    const uint32_t wbase = wv * GKC_WAVE * ITEMS;   // wave wv owns tile keys [wbase, wbase + 64*ITEMS)

    // steps 1-6 for the tile `tile` whose keys are in key[]; `mid` runs once
    // the tile's aggregate is out (the 128-bit prefetch of the next tile)
    auto sort_tile = [&](const K (&key)[ITEMS], uint32_t tile, uint32_t seg, uint32_t ti, uint32_t tn, uint64_t kbase,
                         auto&& mid) {
        // the segment's digit starts, loaded while the ranking runs
        const uint32_t hsd = dg ? sa.hist[((size_t)seg * MAX_PASS + pass) * RADIX + d] : 0u;
        // 1. digits + tile histogram
        uint32_t dig[ITEMS];          // digit in bits [8:0] (RD = no key), rank in [31:9]
#pragma unroll
        for (int i = 0; i < ITEMS; i++) {
            const uint32_t j = wbase + i * GKC_WAVE + lane;
            dig[i] = (j < tn) ? KeyTraits<K>::digit_n(key[i], sh, DB) : (uint32_t)RD;
            if (dig[i] < (uint32_t)RD) atomicAdd(&s_hist[dig[i]], 1u);
        }
        __syncthreads();
        // 2. aggregate out
        const uint32_t total = dg ? s_hist[d] : 0u;
        uint32_t* lb = lookback + (size_t)tile * RADIX;
        if (dg) st_agent(&lb[d], (ti == 0 ? LB_INC : LB_AGG) | total);
        mid();
        // 4. stable rank within the wave
#pragma unroll
        for (int i = 0; i < ITEMS; i++) {
            const uint32_t dd = dig[i];
            uint32_t plo, phi;
            wave_peers<DB>(dd, (uint32_t)RD, plo, phi);
            if (dd < (uint32_t)RD) {
                const uint32_t below = __builtin_amdgcn_mbcnt_hi(phi, __builtin_amdgcn_mbcnt_lo(plo, 0u));
                const uint32_t before = s_whist[wv][dd];
                dig[i] = dd | ((before + below) << 9);
                if (below == 0) s_whist[wv][dd] = (whist_t)(before + (uint32_t)__popc(plo) + (uint32_t)__popc(phi));
            }
        }
        // tile-local digit starts: exclusive scan over the digits, one
        // wave-level shuffle scan per 64 digits plus the wave sums
        uint32_t inc = total;
#pragma unroll
        for (int off = 1; off < GKC_WAVE; off <<= 1) {
            const uint32_t y = __shfl_up(inc, off);
            if ((int)lane >= off) inc += y;
        }
        if (dg && lane == GKC_WAVE - 1) s_wsum[wv] = inc;
        __syncthreads();                                         // ranking done, wave sums visible
        uint32_t wsum = 0;
#pragma unroll
        for (int w = 0; w < DWAVE; w++) wsum += (w < (int)wv) ? s_wsum[w] : 0u;
        uint32_t dstart = 0;
        if (dg) {   // per digit and wave: its first tile position
            dstart = wsum + inc - total;
            uint32_t acc = dstart;
#pragma unroll
            for (int w = 0; w < NWAVE; w++) { const uint32_t c = s_whist[w][d]; s_whist[w][d] = (whist_t)acc; acc += c; }
        }
        __syncthreads();
        // 5. decoupled look-back, one digit per thread
        uint32_t excl = 0;
        if (dg && ti > 0) {
            int64_t j = (int64_t)tile - 1;
            for (;;) {
                uint32_t v = ld_agent(&lookback[(size_t)j * RADIX + d]);
                while ((v & ~LB_VAL) == 0) { __builtin_amdgcn_s_sleep(1); v = ld_agent(&lookback[(size_t)j * RADIX + d]); }
                excl += v & LB_VAL;
                if (v & LB_INC) break;
                j--;
            }
            st_agent(&lb[d], LB_INC | (excl + total));
        }
        if (dg) s_gbase[d] = hsd + excl - dstart;    // mod 2^32; base + j is the key's segment index
        // 6. reorder in LDS by digit, write each digit's run out
#pragma unroll
        for (int i = 0; i < ITEMS; i++) {
            const uint32_t dd = dig[i] & 511u;
            if (dd < (uint32_t)RD) s_keys[s_whist[wv][dd] + (dig[i] >> 9)] = key[i];
        }
        __syncthreads();
        K* const sout = out + kbase;
        for (uint32_t j = threadIdx.x; j < tn; j += NT) {
            const K x = s_keys[j];
            sout[(uint32_t)(s_gbase[KeyTraits<K>::digit_n(x, sh, DB)] + j)] = x;
        }
    };
    auto claim_next = [&]() {
        if (threadIdx.x == 0) {
            uint32_t g = 0;
            s_ok = claim_tile(wl, ctr, list, g) ? 1u : 0u;
            s_g = g;
        }
        if (dg) s_hist[d] = 0;
        for (int i = threadIdx.x; i < NWAVE * RD; i += NT) (&s_whist[0][0])[i] = 0;
    };

    if constexpr (sizeof(K) == 16) {
        // 128-bit keys: the next tile's keys are prefetched into registers
        // during this one (their tiles are large enough that it pays)
        if (threadIdx.x == 0) {
            uint32_t g = 0;
            s_ok = claim_tile(wl, ctr, list, g) ? 1u : 0u;
            s_g = g;
        }
        __syncthreads();
        if (!s_ok) return;
        uint32_t tile = __builtin_amdgcn_readfirstlane(s_g);
        TileDesc td = tdesc[tile];
        K key[ITEMS];
#pragma unroll
        for (int i = 0; i < ITEMS; i++) {
            const uint32_t j = wbase + i * GKC_WAVE + lane;
            if (j < (td.seg_tn & 0xFFFFu)) key[i] = in[td.kbase + (uint64_t)td.ti * TILE + j];
        }
        for (;;) {
            __syncthreads();
            claim_next();
            __syncthreads();
            const bool have_next = s_ok;
            const uint32_t nxt = __builtin_amdgcn_readfirstlane(s_g);
            K nkey[ITEMS];
            TileDesc ntd{};
            sort_tile(key, tile, td.seg_tn >> 16, td.ti, td.seg_tn & 0xFFFFu, td.kbase, [&] {
                if (have_next) {     // 3. the next tile's keys load during this one
                    ntd = tdesc[nxt];
#pragma unroll
                    for (int i = 0; i < ITEMS; i++) {
                        const uint32_t j = wbase + i * GKC_WAVE + lane;
                        if (j < (ntd.seg_tn & 0xFFFFu)) nkey[i] = in[ntd.kbase + (uint64_t)ntd.ti * TILE + j];
                    }
                }
            });
            if (!have_next) break;
            tile = nxt; td = ntd;
#pragma unroll
            for (int i = 0; i < ITEMS; i++) key[i] = nkey[i];
        }
    } else {
        K key[ITEMS];
        for (;;) {
            __syncthreads();
            claim_next();
            __syncthreads();
            if (!s_ok) break;
            const uint32_t tile = __builtin_amdgcn_readfirstlane(s_g);     // look-back slot
            const TileDesc td = tdesc[tile];
            const uint32_t tn = td.seg_tn & 0xFFFFu;
            const K* tin = in + td.kbase + (uint64_t)td.ti * TILE + wbase + lane;
            if (tn == (uint32_t)TILE) {
#pragma unroll
                for (int i = 0; i < ITEMS; i++) key[i] = tin[i * GKC_WAVE];
            } else {
#pragma unroll
                for (int i = 0; i < ITEMS; i++)
                    if (wbase + i * GKC_WAVE + lane < tn) key[i] = tin[i * GKC_WAVE];
            }
            sort_tile(key, tile, td.seg_tn >> 16, td.ti, tn, td.kbase, [] {});
        }
    }
}

// Run-length count + clamp + threshold on sorted segments, compacted with a
// decoupled look-back (one value per tile).  Thread t of a tile holds keys
// [t0 + 16t, t0 + 16t + 16); the tile's last run is followed past the tile.
template <typename K>
__global__ __launch_bounds__(SORT_NT) void k_runlength(Plan* __restrict__ plan, SegArrays sa, WorkLists wl, const K* __restrict__ keys,
                                                       TableOut to, uint32_t* __restrict__ lookback) {
    __shared__ uint32_t s_first[SORT_NT];       // first head index (tile-local) of each thread, or NO_HEAD
    __shared__ uint32_t s_scan[SORT_NT];
    __shared__ uint32_t s_end, s_nd, s_seg, s_idx, s_ok;
    __shared__ uint64_t s_excl;
    uint32_t list = blockIdx.x % wl.n_lists;
    for (;;) {
        if (threadIdx.x == 0) {
            uint32_t sg = 0, ix = 0;
            s_ok = claim<1>(wl, &plan->ctr[ctr_rl()], list, sg, ix);
            s_seg = sg; s_idx = ix;
        }
        __syncthreads();
        if (!s_ok) break;
        const uint32_t seg = s_seg;
        const uint32_t ti = s_idx;
        const uint32_t tile = sa.rl_tile_start[seg] + ti;        // look-back slot
        const K* kp = keys + sa.key_start[seg];
        const uint32_t nseg = sa.seg_keys[seg];
        const uint32_t t0 = ti * RL_TILE;
        const uint32_t tn = min((uint32_t)RL_TILE, nseg - t0);
        const uint32_t my0 = threadIdx.x * RL_ITEMS;
        K x[RL_ITEMS];
        uint32_t headmask = 0;
#pragma unroll
        for (int i = 0; i < RL_ITEMS; i++) {
            const uint32_t j = my0 + i;
            if (j < tn) x[i] = kp[t0 + j];
        }
        K prevk;
        const bool has_prev = (t0 + my0) > 0 && my0 < tn;
        if (has_prev) prevk = kp[t0 + my0 - 1];
#pragma unroll
        for (int i = 0; i < RL_ITEMS; i++) {
            const uint32_t j = my0 + i;
            if (j < tn) {
                const bool head = (i == 0) ? (!has_prev || !(x[0] == prevk)) : !(x[i] == x[i - 1]);
                if (head) headmask |= 1u << i;
            }
        }
        // first head of each thread -> next-head lookup across threads
        // NO_HEAD must exceed every run end, including ends past the tile
        constexpr uint32_t NO_HEAD = 0xFFFFFFFFu;
        s_first[threadIdx.x] = headmask ? my0 + __builtin_ctz(headmask) : NO_HEAD;
        if (threadIdx.x == 0) { s_end = tn; s_nd = 0; }
        __syncthreads();
        // suffix minimum over threads of first-head positions
        for (int off = 1; off < SORT_NT; off <<= 1) {
            const uint32_t y = (threadIdx.x + off < SORT_NT) ? s_first[threadIdx.x + off] : NO_HEAD;
            __syncthreads();
            s_first[threadIdx.x] = min(s_first[threadIdx.x], y);
            __syncthreads();
        }
        // end of the tile's last run: wave 0 follows it past the tile, 64 keys
        // per step, until the key changes or the segment ends
        if (threadIdx.x < GKC_WAVE) {
            const K lk = kp[t0 + tn - 1];
            uint32_t e = t0 + tn;
            for (;;) {
                const uint32_t j = e + __lane_id();
                const bool diff = j >= nseg || !(kp[j] == lk);
                const uint64_t m = __ballot(diff);
                if (m) { e += (uint32_t)__builtin_ctzll(m); break; }
                e += GKC_WAVE;
            }
            if (threadIdx.x == 0) s_end = e - t0;     // tile-local end of the last run
        }
        __syncthreads();
        // counts and filter
        uint32_t keep = 0, nheads = 0;
        uint32_t cnt[RL_ITEMS];
        {
            const uint32_t nxt_thread = (threadIdx.x + 1 < SORT_NT) ? s_first[threadIdx.x + 1] : NO_HEAD;
            uint32_t next = min(nxt_thread, s_end);
#pragma unroll
            for (int i = RL_ITEMS - 1; i >= 0; i--) {
                cnt[i] = 0;
                if (headmask & (1u << i)) {
                    const uint32_t pos = my0 + i;
                    const uint32_t run = next - pos;
                    cnt[i] = run;
                    next = pos;
                    nheads++;
                    if (run >= to.min_cnt && (to.max_cnt == 0 || run <= to.max_cnt)) keep |= 1u << i;
                }
            }
        }
        const uint32_t nk = __popc(keep);
        s_scan[threadIdx.x] = nk;
        atomicAdd(&s_nd, nheads);
        __syncthreads();
        for (int off = 1; off < SORT_NT; off <<= 1) {
            const uint32_t y = threadIdx.x >= (uint32_t)off ? s_scan[threadIdx.x - off] : 0;
            __syncthreads();
            s_scan[threadIdx.x] += y;
            __syncthreads();
        }
        const uint32_t tile_total = s_scan[SORT_NT - 1];
        // decoupled look-back for the tile's output offset within the segment
        if (threadIdx.x == 0) {
            uint32_t* lb = lookback + tile;
            uint32_t excl = 0;
            if (ti == 0) st_agent(lb, LB_INC | tile_total);
            else {
                st_agent(lb, LB_AGG | tile_total);
                int64_t j = (int64_t)tile - 1;
                for (;;) {
                    uint32_t v = ld_agent(&lookback[j]);
                    while ((v & ~LB_VAL) == 0) { __builtin_amdgcn_s_sleep(1); v = ld_agent(&lookback[j]); }
                    excl += v & LB_VAL;
                    if (v & LB_INC) break;
                    j--;
                }
                st_agent(lb, LB_INC | (excl + tile_total));
            }
            s_excl = excl;
            atomicAdd(&to.n_distinct[seg], s_nd);
            atomicAdd(&to.n_entries[seg], tile_total);
        }
        __syncthreads();
        bool over = false;                      // an entry past the capacity: one flag per wave
        if (to.emit && keep) {
            const uint32_t cap = to.table_cap[seg];
            uint32_t idx = (uint32_t)s_excl + s_scan[threadIdx.x] - nk;
            K* tk = reinterpret_cast<K*>(to.keys) + to.table_start[seg];
            uint32_t* tc = to.counts + to.table_start[seg];
#pragma unroll
            for (int i = 0; i < RL_ITEMS; i++) {
                if (keep & (1u << i)) {
                    if (idx < cap) { tk[idx] = x[i]; tc[idx] = cnt[i]; }
                    else over = true;
                    idx++;
                }
            }
        }
        const uint64_t ob = __ballot(over);
        if (ob && __lane_id() == (uint32_t)__builtin_ctzll(ob)) atomicOr(to.overflow, 1u);
        __syncthreads();
    }
}

// ---------------------------------------------------------------- binned count
// Keys of a segment grouped by bin (their top 24 bits, i.e. sorted by them)
// are counted window by window.  Window t of a segment owns the bins whose
// first key lies in [t*WC_TILE, (t+1)*WC_TILE): its range [a, a+n) starts at
// the first bin start at or after t*WC_TILE and ends at the first bin start
// at or after (t+1)*WC_TILE (k_binranges, before any window's keys are
// reused as staging).  k_wavecount counts ranges of up to WC_CAP keys, one
// wave each; k_bincount the larger ones beforehand (a bin of more than
// WC_CAP - WC_TILE + 1 keys: skewed composition or a high-copy k-mer), one
// workgroup each: sorted on the bits that vary inside the range (LSD radix,
// ranking by wave64 ballots as in k_onesweep) in LDS, or in CAP-key chunks
// through global memory ping-ponging with the spare key buffer, then
// run-length counted, thresholded and staged.
// Key bits [bsh, bsh + nb), nb <= 32 (bits above the key width read as 0).
__device__ __forceinline__ uint32_t bits_of(uint64_t x, int bsh, int nb) {
    return (uint32_t)(x >> bsh) & (uint32_t)((1ull << nb) - 1);
}
__device__ __forceinline__ uint32_t bits_of(const u128k& x, int bsh, int nb) {
    uint64_t v;
    if (bsh >= 64) v = x.hi >> (bsh - 64);
    else v = (x.lo >> bsh) | (bsh ? x.hi << (64 - bsh) : 0ull);
    return (uint32_t)v & (uint32_t)((1ull << nb) - 1);
}
__device__ __forceinline__ void or_diff(uint64_t& lo, uint64_t&, uint64_t a, uint64_t b) { lo |= a ^ b; }
__device__ __forceinline__ void or_diff(uint64_t& lo, uint64_t& hi, const u128k& a, const u128k& b) {
    lo |= a.lo ^ b.lo; hi |= a.hi ^ b.hi;
}

// First index in [j, end) whose bin (key bits [bsh, bsh + bin_bits)) is not
// pb, or end; kp[j - 1] is in bin pb.  The keys are sorted by bin, so bin pb
// is one run: a short linear scan (a 3-pass bin ends within a few keys), then an
// exponential and a binary search.
template <typename K>
__device__ __forceinline__ uint32_t bin_run_end(const K* kp, uint32_t j, uint32_t end, uint32_t pb, int bsh, int bin_bits) {
    constexpr uint32_t SCAN = 4;
    const uint32_t lin = min(end, j + SCAN);
    for (; j < lin; j++)
        if (bits_of(kp[j], bsh, bin_bits) != pb) return j;
    if (j >= end) return end;
    uint32_t lo = j - 1, step = SCAN;        // kp[lo] is in bin pb
    uint32_t hi = end;                       // first index known past the run (or end)
    for (;;) {
        const uint32_t x = lo + step;
        if (x >= end) break;
        if (bits_of(kp[x], bsh, bin_bits) != pb) { hi = x; break; }
        lo = x;
        step *= 2;
    }
    while (hi - lo > 1) {                    // kp[lo] in pb; hi past the run or end
        const uint32_t mid = lo + (hi - lo) / 2;
        if (bits_of(kp[mid], bsh, bin_bits) != pb) hi = mid; else lo = mid;
    }
    return hi;
}

// Window ranges: wdesc[t] = {segment, n, first key of the range (absolute
// index) low, high}, the range starting at the window's first bin start, n = 0
// when no bin starts in the window; bins are key bits [bsh, bsh + bin_bits).  One lane per window:
// a = the first bin start in [lo, hi), the range ends at the first bin start
// at or after hi.  Windows whose range holds more than heavy_cap keys are
// appended to k_bincount's list (when one is given), except 64-bit ranges of
// up to WC_TRY_MAX keys, which k_wavecount tries first.
template <typename K>
__global__ __launch_bounds__(256) void k_binranges(Plan* __restrict__ plan, SegArrays sa, const K* __restrict__ keys,
                                                   uint4* __restrict__ wdesc, int tile_keys, int bsh, int bin_bits,
                                                   uint32_t heavy_cap, HeavyList hl) {
    const uint32_t TILE = (uint32_t)tile_keys;
    const uint32_t total = plan->tot_rl_tiles, S = plan->n_segs;
    const uint32_t nthreads = gridDim.x * blockDim.x;
    const uint32_t lane = __lane_id();
    // the waves' lanes take consecutive windows, so the range end of window t
    // (the first bin start at or after its end) is usually the next lane's a
    for (uint32_t t0 = blockIdx.x * blockDim.x + threadIdx.x - lane; t0 < total; t0 += nthreads) {
        const uint32_t t = t0 + lane;
        const bool act = t < total;
        uint32_t seg = 0, lo = 0, hi = 0, nseg = 0, a = 0;
        const K* kp = keys;
        if (act) {
            seg = find_seg(sa.rl_tile_start, S, t);
            kp = keys + sa.key_start[seg];
            nseg = sa.seg_keys[seg];
            lo = (t - sa.rl_tile_start[seg]) * TILE;
            hi = min(lo + TILE, nseg);
            a = lo == 0 ? 0u : bin_run_end(kp, lo, hi, bits_of(kp[lo - 1], bsh, bin_bits), bsh, bin_bits);
        }
        // the next window's a, if it is this segment's and holds a bin start
        const uint32_t a_next = __shfl_down(a, 1), seg_next = __shfl_down(seg, 1), hi_next = __shfl_down(hi, 1);
        const bool next_ok = lane + 1 < GKC_WAVE && t + 1 < total && seg_next == seg && a_next < hi_next;
        if (!act) continue;
        uint32_t n = 0;
        if (a < hi) {
            const uint32_t b = hi >= nseg ? nseg
                             : next_ok ? a_next
                                       : bin_run_end(kp, hi, nseg, bits_of(kp[hi - 1], bsh, bin_bits), bsh, bin_bits);
            n = b - a;
        }
        const uint64_t abs = sa.key_start[seg] + a;
        wdesc[t] = make_uint4(seg, n, (uint32_t)abs, (uint32_t)(abs >> 32));
        if (hl.tiles && n > heavy_cap && !(hl.mid[0] && n <= WC_TRY_MAX)) hl.tiles[atomicAdd(hl.count, 1u)] = t;
    }
}

// Stable rank of the wave's valid keys (index jb + 64 i < nv) by digit `sh`,
// continuing the wave's running digit counters s_whist[wv][*].  wb is the
// wave's first index (jb - lane): items the whole wave has past nv are skipped.
template <typename K, int ITEMS>
__device__ __forceinline__ void rank_in_wave(const K (&key)[ITEMS], uint32_t jb, uint32_t wb, uint32_t nv, int sh,
                                             uint32_t* __restrict__ whist, uint32_t (&dig)[ITEMS]) {
#pragma unroll
    for (int i = 0; i < ITEMS; i++) {
        dig[i] = RADIX;
        if (wb + i * GKC_WAVE >= nv) continue;
        const uint32_t j = jb + i * GKC_WAVE;
        const uint32_t dd = (j < nv) ? KeyTraits<K>::digit(key[i], sh) : RADIX;
        uint64_t peers = __ballot(dd < RADIX);
#pragma unroll
        for (int bb = 0; bb < RADIX_BITS; bb++) {
            const uint64_t bal = __ballot((dd >> bb) & 1u);
            peers &= ((dd >> bb) & 1u) ? bal : ~bal;
        }
        dig[i] = dd;
        if (dd < RADIX) {
            const uint32_t before = whist[dd];
            const uint64_t lower = peers & lanemask_lt();
            dig[i] = dd | ((before + (uint32_t)__popcll(lower)) << 9);
            if (lower == 0) whist[dd] = before + (uint32_t)__popcll(peers);
        }
    }
}

// Exclusive scan over the 256 digits of `v` (threads 0..255, four full
// waves); s_red holds the four wave sums.  Returns the exclusive prefix;
// *total gets the sum (valid in every thread after the call).
__device__ __forceinline__ uint32_t digit_scan(uint32_t v, uint32_t tid, uint32_t lane, uint32_t wv,
                                               uint32_t* s_red, uint32_t* total) {
    uint32_t inc = v;
    if (tid < RADIX) {
#pragma unroll
        for (int off = 1; off < GKC_WAVE; off <<= 1) {
            const uint32_t y = __shfl_up(inc, off);
            if ((int)lane >= off) inc += y;
        }
        if (lane == GKC_WAVE - 1) s_red[wv] = inc;
    }
    __syncthreads();
    uint32_t off = 0;
    if (tid < RADIX) for (uint32_t w = 0; w < wv; w++) off += s_red[w];
    *total = s_red[0] + s_red[1] + s_red[2] + s_red[3];
    return off + inc - v;
}

// Counts the windows of the heavy list, one workgroup per window at a time.
// The kept entries of a range at a go to staging, keys at alt[a + i] and
// counts as u32 over the range's own keys (they trail the reads: entry i
// never overtakes key 2i), and {kept, distinct} to hl.res[tile] (for
// k_wavecount).  late: the windows k_wavecount handed on (after it), whose
// kept count goes to hl.kept[tile] and both counts to the segment's totals.

template <typename K>
__global__ __launch_bounds__(BC_NT) void k_bincount(Plan* __restrict__ plan, SegArrays sa, K* __restrict__ keys,
                                                    K* __restrict__ alt, const uint4* __restrict__ wdesc,
                                                    TableOut to, HeavyList hl, int late) {
    constexpr int ITEMS = BinTraits<K>::ITEMS;
    constexpr uint32_t CAP = bc_cap<K>();
    constexpr int NWAVE = BC_NT / GKC_WAVE;
    constexpr int NCH = CAP / GKC_WAVE;          // 64-key chunks of a range, in key order
    constexpr int NCW = NCH / GKC_WAVE;          // waves that scan the chunk values
    constexpr uint32_t INF = 0xFFFFFFFFu;
    static_assert(NCH % GKC_WAVE == 0 && NCW <= NWAVE && RADIX == 4 * GKC_WAVE, "tile shape");
    __shared__ K s_keys[CAP];
    __shared__ uint32_t s_cnt[sizeof(K) == 8 ? CAP : 1];   // hash-table counts, then entry counts
    __shared__ uint32_t s_whist[NWAVE][RADIX];
    __shared__ uint32_t s_dstart[RADIX];
    __shared__ uint64_t s_hb[NCH];               // run-head ballot of each chunk
    __shared__ uint32_t s_ch[NCH];               // first head of each chunk, then the next head after it
    __shared__ uint32_t s_ck[NCH];               // kept entries of each chunk, then their exclusive prefix
    __shared__ uint32_t s_red[NWAVE];
    __shared__ uint32_t s_orv[NWAVE][4];
    __shared__ uint4 s_wd;                       // the window: segment, keys, first key (absolute) low, high
    __shared__ uint32_t s_tile, s_ok, s_nd, s_out, s_nh, s_occ, s_hfail;
    const uint32_t lane = __lane_id(), wv = threadIdx.x / GKC_WAVE, tid = threadIdx.x;
    const uint32_t wbase = wv * GKC_WAVE * ITEMS;     // wave wv holds chunk keys [wbase, wbase + 64*ITEMS)
    constexpr uint64_t EMPTY = ~0ull;  // never canonical: its reverse complement (0) is smaller
    constexpr int LOG_CAP = __builtin_ctz(CAP);
    constexpr uint32_t HASH_LIMIT = CAP / 4 * 3;   // distinct keys a hashed range may hold
    for (;;) {
        if (tid == 0) {
            uint32_t t = 0;
            uint4 wd = make_uint4(0u, 0u, 0u, 0u);
            const uint32_t q = atomicAdd(late ? hl.late_next : hl.next, 1u);
            const bool ok = q < *(late ? hl.late_count : hl.count);
            if (ok) {
                t = (late ? hl.late : hl.tiles)[q];
                wd = wdesc[t];
            }
            s_ok = ok; s_tile = t; s_wd = wd; s_nd = 0; s_out = 0; s_nh = 0; s_occ = 0; s_hfail = 0;
        }
        // 64-bit keys: equal keys are counted in an LDS hash table (while the
        // range is read, CAP keys at a time), so only the distinct keys are
        // sorted, with their counts; ranges with more than HASH_LIMIT distinct
        // keys are sorted whole instead
        if constexpr (sizeof(K) == 8)
            for (uint32_t i = tid; i < CAP; i += BC_NT) { s_keys[i] = EMPTY; s_cnt[i] = 0; }
        __syncthreads();
        if (!s_ok) break;
        const uint4 wd = s_wd;
        const uint32_t seg = wd.x, tile = s_tile;
        const uint64_t abs = ((uint64_t)wd.w << 32) | wd.z;
        // per-tile copy of the lane's base index, opaque to the compiler: the
        // per-item addresses are recomputed in the loop instead of being
        // hoisted out of it (which would exceed the 128 VGPRs of 16 waves/CU)
        uint32_t jb = wbase + lane;
        asm volatile("" : "+v"(jb));
        const uint32_t n = wd.y;
        const bool heavy = n > CAP;                             // larger than LDS: chunks through global memory
        K* src = keys + abs;
        K* dst = alt + abs;
        // 1. bits that vary inside the range (and, for 64-bit keys, the hash count)
        K key[ITEMS];
        uint64_t olo = 0, ohi = 0;
        {
            K ref{};
            if (n) ref = src[0];
            for (uint32_t c0 = 0; c0 < n; c0 += CAP) {
#pragma unroll(sizeof(K) == 8 ? 2 : ITEMS)
                for (int i = 0; i < ITEMS; i++) {
                    const uint32_t j = c0 + jb + i * GKC_WAVE;
                    if (j < n) {
                        if constexpr (sizeof(K) != 8) {
                            key[i] = src[j];
                            or_diff(olo, ohi, key[i], ref);
                        } else {
                            const K xk = src[j];          // not kept: the table holds the keys
                            or_diff(olo, ohi, xk, ref);
                            const unsigned long long x = xk;
                            uint32_t h = (uint32_t)((x * 0x9E3779B97F4A7C15ull) >> (64 - LOG_CAP));
                            // bounded probe: a full table (possible only past
                            // HASH_LIMIT) fails the range instead of looping
                            // (a failed range stops inserting: its probes would only lengthen)
                            bool done = *(volatile uint32_t*)&s_hfail != 0;
                            for (uint32_t pr = 0; pr < CAP && !done; pr++) {
                                const unsigned long long old =
                                    atomicCAS(reinterpret_cast<unsigned long long*>(&s_keys[h]), (unsigned long long)EMPTY, x);
                                if (old == EMPTY || old == x) {
                                    atomicAdd(&s_cnt[h], 1u);
                                    if (old == EMPTY && atomicAdd(&s_occ, 1u) + 1 > HASH_LIMIT) s_hfail = 1;
                                    done = true;
                                }
                                h = (h + 1) & (CAP - 1);
                            }
                            if (!done) s_hfail = 1;
                        }
                    }
                }
            }
        }
        uint32_t o0 = (uint32_t)olo, o1 = (uint32_t)(olo >> 32), o2 = (uint32_t)ohi, o3 = (uint32_t)(ohi >> 32);
#pragma unroll
        for (int off = 1; off < GKC_WAVE; off <<= 1) {
            o0 |= __shfl_xor(o0, off); o1 |= __shfl_xor(o1, off);
            o2 |= __shfl_xor(o2, off); o3 |= __shfl_xor(o3, off);
        }
        if (lane == 0) { s_orv[wv][0] = o0; s_orv[wv][1] = o1; s_orv[wv][2] = o2; s_orv[wv][3] = o3; }
        __syncthreads();
        o0 = o1 = o2 = o3 = 0;
#pragma unroll
        for (int w = 0; w < NWAVE; w++) { o0 |= s_orv[w][0]; o1 |= s_orv[w][1]; o2 |= s_orv[w][2]; o3 |= s_orv[w][3]; }
        const uint64_t vlo = ((uint64_t)o1 << 32) | o0, vhi = ((uint64_t)o3 << 32) | o2;
        const int nbits = vhi ? 128 - __builtin_clzll(vhi) : (vlo ? 64 - __builtin_clzll(vlo) : 0);
        const int npass = (nbits + RADIX_BITS - 1) / RADIX_BITS;
        uint32_t dig[ITEMS];
        uint32_t cnt[ITEMS];
        // 64-bit keys in LDS: equal keys are counted first in an LDS hash table,
        // so only the distinct keys are sorted (with their counts)
        // (the table and s_hfail are complete: the range's reads and inserts
        // precede the barrier above)
        const bool hashed = sizeof(K) == 8 && s_hfail == 0;
        if (sizeof(K) == 8 && !hashed && !heavy) {     // sorted whole in LDS: its keys in registers
#pragma unroll
            for (int i = 0; i < ITEMS; i++) {
                const uint32_t j = jb + i * GKC_WAVE;
                if (j < n) key[i] = src[j];
            }
        }
        bool ranked = false;                   // staged already by rank of comparison
        uint32_t nsort = n;                    // entries sorted in LDS
        if constexpr (sizeof(K) == 8) {
            if (hashed) {
                // compact the occupied slots (slot i*BC_NT + tid of thread tid)
                uint32_t occ = 0;
#pragma unroll
                for (int i = 0; i < ITEMS; i++) {
                    key[i] = s_keys[i * BC_NT + tid];
                    cnt[i] = s_cnt[i * BC_NT + tid];
                    occ += key[i] != EMPTY ? 1u : 0u;
                }
                uint32_t inc = occ;
#pragma unroll
                for (int off = 1; off < GKC_WAVE; off <<= 1) {
                    const uint32_t y = __shfl_up(inc, off);
                    if ((int)lane >= off) inc += y;
                }
                if (lane == GKC_WAVE - 1) s_red[wv] = inc;
                __syncthreads();
                uint32_t pos = inc - occ, dtot = 0;
#pragma unroll
                for (int w = 0; w < NWAVE; w++) { const uint32_t v = s_red[w]; pos += (uint32_t)w < wv ? v : 0u; dtot += v; }
#pragma unroll
                for (int i = 0; i < ITEMS; i++)
                    if (key[i] != EMPTY) { s_keys[pos] = key[i]; s_cnt[pos] = cnt[i]; pos++; }
                nsort = dtot;
                __syncthreads();
                {
                    // rank each kept distinct key among the kept ones (its place
                    // in key order) and stage it, without a full sort: bucket the
                    // entries by their 8 highest varying bits (one counting
                    // scatter in LDS), then compare within the bucket only
                    const int bsh = nbits > RADIX_BITS ? nbits - RADIX_BITS : 0;
                    uint32_t* b_n = s_whist[0];      // entries per bucket
                    uint32_t* b_k = s_whist[1];      // kept entries per bucket
                    uint32_t* b_s = s_whist[2];      // bucket start (entries)
                    uint32_t* b_ks = s_whist[3];     // bucket start (kept entries)
                    if (tid < RADIX) { b_n[tid] = 0; b_k[tid] = 0; }
                    __syncthreads();
                    uint32_t bp[ITEMS];              // bucket << 16 | place in bucket
#pragma unroll
                    for (int i = 0; i < ITEMS; i++) {
                        const uint32_t e = jb + i * GKC_WAVE;
                        bp[i] = 0xFFFFFFFFu;
                        if (e < nsort) {
                            key[i] = s_keys[e]; cnt[i] = s_cnt[e];
                            const uint32_t b = KeyTraits<K>::digit(key[i], bsh);
                            bp[i] = (b << 16) | atomicAdd(&b_n[b], 1u);
                            if (cnt[i] >= to.min_cnt && (to.max_cnt == 0 || cnt[i] <= to.max_cnt)) atomicAdd(&b_k[b], 1u);
                        }
                    }
                    __syncthreads();
                    uint32_t all_n, all_k;
                    const uint32_t ex_n = digit_scan(tid < RADIX ? b_n[tid] : 0u, tid, lane, wv, s_red, &all_n);
                    __syncthreads();
                    const uint32_t ex_k = digit_scan(tid < RADIX ? b_k[tid] : 0u, tid, lane, wv, s_red, &all_k);
                    if (tid < RADIX) { b_s[tid] = ex_n; b_ks[tid] = ex_k; }
                    __syncthreads();
#pragma unroll
                    for (int i = 0; i < ITEMS; i++) {
                        if (bp[i] != 0xFFFFFFFFu) {
                            const uint32_t p2 = b_s[bp[i] >> 16] + (bp[i] & 0xFFFFu);
                            s_keys[p2] = key[i]; s_cnt[p2] = cnt[i];
                        }
                    }
                    __syncthreads();
                    K* tk = alt + abs;
                    uint32_t* tc = reinterpret_cast<uint32_t*>(keys + abs);
#pragma unroll
                    for (int i = 0; i < ITEMS; i++) {
                        if (bp[i] == 0xFFFFFFFFu) continue;
                        const uint64_t x = key[i];
                        const uint32_t c = cnt[i];
                        if (c >= to.min_cnt && (to.max_cnt == 0 || c <= to.max_cnt)) {
                            const uint32_t b = bp[i] >> 16;
                            const uint32_t q0 = b_s[b], q1 = q0 + b_n[b];
                            uint32_t r = b_ks[b];
                            for (uint32_t q = q0; q < q1; q++) {
                                const uint32_t cq = s_cnt[q];
                                r += (s_keys[q] < x && cq >= to.min_cnt && (to.max_cnt == 0 || cq <= to.max_cnt)) ? 1u : 0u;
                            }
                            tk[r] = x; tc[r] = c;
                        }
                    }
                    if (tid == 0) { s_nd = nsort; s_out = all_k; }
                    ranked = true;
                    __syncthreads();
                }
            }
        }
        if (!ranked) {
        if (!heavy || hashed) {
            // 2a. LSD radix sort of the range (or of its distinct keys) in LDS
            //     (entries in registers, j = jb + 64 i)
            for (int ps = 0; ps < npass; ps++) {
                const int sh = ps * RADIX_BITS;
                for (int i = tid; i < NWAVE * RADIX; i += BC_NT) (&s_whist[0][0])[i] = 0;
                __syncthreads();
                rank_in_wave<K, ITEMS>(key, jb, wbase, nsort, sh, s_whist[wv], dig);
                __syncthreads();
                uint32_t tot = 0;
                if (tid < RADIX) {
#pragma unroll
                    for (int w = 0; w < NWAVE; w++) { const uint32_t c = s_whist[w][tid]; s_whist[w][tid] = tot; tot += c; }
                }
                uint32_t all;
                const uint32_t ex = digit_scan(tot, tid, lane, wv, s_red, &all);
                if (tid < RADIX) s_dstart[tid] = ex;
                __syncthreads();
#pragma unroll
                for (int i = 0; i < ITEMS; i++) {
                    const uint32_t dd = dig[i] & 511u;
                    if (dd < RADIX) {
                        const uint32_t p = s_dstart[dd] + s_whist[wv][dd] + (dig[i] >> 9);
                        s_keys[p] = key[i];
                        if (hashed) s_cnt[p] = cnt[i];
                    }
                }
                __syncthreads();
#pragma unroll
                for (int i = 0; i < ITEMS; i++) {
                    const uint32_t j = jb + i * GKC_WAVE;
                    if (wbase + i * GKC_WAVE < nsort && j < nsort) {
                        key[i] = s_keys[j];
                        if (hashed) cnt[i] = s_cnt[j];
                    }
                }
            }
            if (npass == 0 && !hashed) {   // 0 or 1 key, or all keys equal: already in order
#pragma unroll
                for (int i = 0; i < ITEMS; i++) {
                    const uint32_t j = jb + i * GKC_WAVE;
                    if (j < n) s_keys[j] = key[i];
                }
                __syncthreads();
            }
        } else {
            // 2b. LSD radix sort of the range through global memory, CAP keys
            //     at a time, stable across chunks (running digit positions)
            for (int ps = 0; ps < npass; ps++) {
                const int sh = ps * RADIX_BITS;
                if (tid < RADIX) s_dstart[tid] = 0;
                __syncthreads();
                for (uint32_t c0 = 0; c0 < n; c0 += CAP) {
#pragma unroll
                    for (int i = 0; i < ITEMS; i++) {
                        const uint32_t j = c0 + jb + i * GKC_WAVE;
                        if (j < n) atomicAdd(&s_dstart[KeyTraits<K>::digit(src[j], sh)], 1u);
                    }
                }
                __syncthreads();
                uint32_t all;
                const uint32_t h = tid < RADIX ? s_dstart[tid] : 0;
                const uint32_t ex = digit_scan(h, tid, lane, wv, s_red, &all);
                if (tid < RADIX) s_dstart[tid] = ex;
                for (uint32_t c0 = 0; c0 < n; c0 += CAP) {
                    const uint32_t nv = min(CAP, n - c0);
                    for (int i = tid; i < NWAVE * RADIX; i += BC_NT) (&s_whist[0][0])[i] = 0;
                    __syncthreads();
#pragma unroll
                    for (int i = 0; i < ITEMS; i++) {
                        const uint32_t j = jb + i * GKC_WAVE;
                        if (j < nv) key[i] = src[c0 + j];
                    }
                    rank_in_wave<K, ITEMS>(key, jb, wbase, nv, sh, s_whist[wv], dig);
                    __syncthreads();
                    uint32_t tot = 0;
                    if (tid < RADIX) {
#pragma unroll
                        for (int w = 0; w < NWAVE; w++) { const uint32_t c = s_whist[w][tid]; s_whist[w][tid] = tot; tot += c; }
                    }
                    __syncthreads();
#pragma unroll
                    for (int i = 0; i < ITEMS; i++) {
                        const uint32_t dd = dig[i] & 511u;
                        if (dd < RADIX) dst[s_dstart[dd] + s_whist[wv][dd] + (dig[i] >> 9)] = key[i];
                    }
                    __syncthreads();
                    if (tid < RADIX) s_dstart[tid] += tot;
                }
                K* t = src; src = dst; dst = t;
                __syncthreads();
            }
            if (npass & 1) {
                // the sorted range ended in the spare buffer: copy it back, so
                // the key buffer stays grouped by bin for a later pass over it
                for (uint32_t j = tid; j < n; j += BC_NT) dst[j] = src[j];
                K* t = src; src = dst; dst = t;
                __syncthreads();
            }
        }
        // 3. run-length count of the sorted range, CAP keys at a time (one
        //    chunk when the range was sorted in LDS), emitted in key order
        for (uint32_t c0 = 0; c0 == 0 || (!hashed && c0 < n); c0 += CAP) {
            const uint32_t nv = hashed ? nsort : (n > c0 ? min(CAP, n - c0) : 0);
            uint32_t keepmask = 0;
            if (hashed) {
                // entries are distinct and counted: threshold them
#pragma unroll
                for (int i = 0; i < ITEMS; i++) {
                    if (wbase + i * GKC_WAVE >= nv) {
                        if (lane == 0) s_ck[wv * ITEMS + i] = 0;
                        continue;
                    }
                    const uint32_t j = jb + i * GKC_WAVE;
                    const bool keep = j < nv && cnt[i] >= to.min_cnt && (to.max_cnt == 0 || cnt[i] <= to.max_cnt);
                    keepmask |= keep ? 1u << i : 0u;
                    const uint64_t kb = __ballot(keep);
                    if (lane == 0) s_ck[wv * ITEMS + i] = (uint32_t)__popcll(kb);
                }
                if (tid == 0) s_nd = nv;
            } else {
            if (heavy) {
#pragma unroll
                for (int i = 0; i < ITEMS; i++) {
                    const uint32_t j = jb + i * GKC_WAVE;
                    if (j < nv) key[i] = src[c0 + j];
                }
                // next head at or after the chunk's end (wave 0, kept across chunks)
                if (wv == 0 && s_nh < c0 + nv) {
                    uint32_t e = c0 + nv;
                    if (e < n) {
                        for (;; e += GKC_WAVE) {
                            const uint32_t j = e + lane;
                            const bool stop = j >= n || !(src[j] == src[j - 1]);
                            const uint64_t m = __ballot(stop);
                            if (m) { e += (uint32_t)__builtin_ctzll(m); break; }
                        }
                    }
                    if (lane == 0) s_nh = e;
                }
            }
            uint32_t headmask = 0;
#pragma unroll
            for (int i = 0; i < ITEMS; i++) {
                if (wbase + i * GKC_WAVE >= nv) {          // the whole wave is past the range
                    if (lane == 0) { s_hb[wv * ITEMS + i] = 0; s_ch[wv * ITEMS + i] = INF; }
                    continue;
                }
                const uint32_t j = jb + i * GKC_WAVE;
                bool head = false;
                if (j < nv) {
                    const uint32_t g = c0 + j;
                    head = g == 0 || !(key[i] == (heavy ? src[g - 1] : s_keys[j - 1]));
                }
                const uint64_t hb = __ballot(head);
                headmask |= head ? 1u << i : 0u;
                if (lane == 0) {
                    s_hb[wv * ITEMS + i] = hb;
                    s_ch[wv * ITEMS + i] = hb ? wbase + i * GKC_WAVE + (uint32_t)__builtin_ctzll(hb) : INF;
                    if (hb) atomicAdd(&s_nd, (uint32_t)__popcll(hb));
                }
            }
            __syncthreads();
            // next head after each 64-key chunk: exclusive suffix minimum
            uint32_t sfx = INF;
            if (tid < (uint32_t)NCH) {
                sfx = s_ch[tid];
#pragma unroll
                for (int off = 1; off < GKC_WAVE; off <<= 1) {
                    const uint32_t y = __shfl_down(sfx, off);
                    if (lane + off < GKC_WAVE) sfx = min(sfx, y);
                }
                if (lane == 0) s_red[wv] = sfx;
            }
            __syncthreads();
            if (tid < (uint32_t)NCH) {
                uint32_t e = __shfl_down(sfx, 1);
                if (lane == GKC_WAVE - 1) e = INF;
                for (int w = wv + 1; w < NCW; w++) e = min(e, s_red[w]);
                s_ch[tid] = e;
            }
            __syncthreads();
            // counts (chunk-local positions; the run after the chunk's last
            // head ends at the chunk end in LDS mode, at s_nh otherwise)
            const uint32_t tail = heavy ? s_nh - c0 : nv;
#pragma unroll
            for (int i = 0; i < ITEMS; i++) {
                const uint32_t q0 = wbase + i * GKC_WAVE;
                cnt[i] = 0;
                if (q0 >= nv) {
                    if (lane == 0) s_ck[wv * ITEMS + i] = 0;
                    continue;
                }
                bool keep = false;
                if ((headmask >> i) & 1u) {
                    const uint64_t hb = s_hb[wv * ITEMS + i];
                    const uint64_t above = lane == GKC_WAVE - 1 ? 0ull : hb & (~0ull << (lane + 1));
                    uint32_t nh = above ? q0 + (uint32_t)__builtin_ctzll(above) : s_ch[wv * ITEMS + i];
                    if (nh == INF) nh = tail;
                    cnt[i] = nh - (q0 + lane);
                    keep = cnt[i] >= to.min_cnt && (to.max_cnt == 0 || cnt[i] <= to.max_cnt);
                }
                keepmask |= keep ? 1u << i : 0u;
                const uint64_t kb = __ballot(keep);
                if (lane == 0) s_ck[wv * ITEMS + i] = (uint32_t)__popcll(kb);
            }
            }
            __syncthreads();
            // exclusive prefix of the kept entries over the 64-key chunks
            if (tid < (uint32_t)NCH) {
                const uint32_t v = s_ck[tid];
                uint32_t inc = v;
#pragma unroll
                for (int off = 1; off < GKC_WAVE; off <<= 1) {
                    const uint32_t y = __shfl_up(inc, off);
                    if ((int)lane >= off) inc += y;
                }
                s_ck[tid] = inc - v;
                if (lane == GKC_WAVE - 1) s_red[wv] = inc;
            }
            __syncthreads();
            if (tid < (uint32_t)NCH) {
                uint32_t off = 0;
                for (uint32_t w = 0; w < wv; w++) off += s_red[w];
                s_ck[tid] += off;
            }
            uint32_t ctot = 0;
            for (int w = 0; w < NCW; w++) ctot += s_red[w];
            __syncthreads();
            // 4. stage the kept (k-mer, count) entries
            if (ctot) {
                K* tk = alt + abs;
                uint32_t* tc = reinterpret_cast<uint32_t*>(keys + abs);
                const uint32_t base = s_out;
#pragma unroll
                for (int i = 0; i < ITEMS; i++) {
                    if (wbase + i * GKC_WAVE >= nv) continue;
                    const bool keep = (keepmask >> i) & 1u;
                    const uint64_t kb = __ballot(keep);
                    if (keep) {
                        const uint32_t idx = base + s_ck[wv * ITEMS + i] + (uint32_t)__popcll(kb & lanemask_lt());
                        tk[idx] = key[i]; tc[idx] = cnt[i];
                    }
                }
            }
            __syncthreads();
            if (tid == 0) s_out += ctot;
        }
        }   // !ranked
        if (tid == 0) {
            if (!late) hl.res[tile] = make_uint2(s_out, s_nd);
            else {
                hl.kept[tile] = s_out;
                atomicAdd(&to.n_distinct[seg], s_nd);
                if (s_out) atomicAdd(&to.n_entries[seg], s_out);
            }
        }
        __syncthreads();
    }
}

// Number of the first m keys of a 16-byte aligned LDS array below x (two
// 64-bit keys per 16-byte read).
__device__ __forceinline__ uint32_t count_below(const uint64_t* a, uint32_t m, uint64_t x) {
    const ulonglong2* a2 = reinterpret_cast<const ulonglong2*>(a);
    uint32_t r = 0, q = 0;
    for (; q + 2 <= m; q += 2) {
        const ulonglong2 v = a2[q >> 1];
        r += (v.x < x ? 1u : 0u) + (v.y < x ? 1u : 0u);
    }
    if (q < m) r += a[q] < x ? 1u : 0u;
    return r;
}
__device__ __forceinline__ uint32_t count_below(const u128k* a, uint32_t m, const u128k& x) {
    uint32_t r = 0;
    for (uint32_t q = 0; q < m; q++) r += less128(a[q], x) ? 1u : 0u;
    return r;
}

// Medium heavy windows, 64-bit keys, that k_wavecount handed on (more than
// WC_FILL distinct keys; list L: up to mid_max(L) keys, more than WC_CAP or
// mid_max(0)): one wave per window, no workgroup barriers.  The wave's LDS
// hash table has mid_slots(L) slots, at most 75 % full, so every insert finds
// its slot; the kept distinct keys are compacted to the front, ranked among
// themselves by comparison (their place in key order) and staged as
// k_bincount stages them (keys at alt[a + r], counts as u32 over the range's
// own keys, read already); the kept count to hl.kept[tile], both counts added
// to the segment's totals.
template <int L>
__global__ __launch_bounds__(GKC_WAVE) void k_midcount(Plan* __restrict__ plan, SegArrays sa, uint64_t* __restrict__ keys,
                                                       uint64_t* __restrict__ alt, const uint4* __restrict__ wdesc,
                                                       TableOut to, HeavyList hl) {
    constexpr uint32_t MID_SLOTS = mid_slots(L);
    __shared__ __align__(16) uint64_t s_key[MID_SLOTS];
    __shared__ uint32_t s_cnt[MID_SLOTS];
    constexpr uint64_t EMPTY = ~0ull;  // never canonical: its reverse complement (0) is smaller
    constexpr int LOG = __builtin_ctz(MID_SLOTS);
    static_assert((MID_SLOTS & (MID_SLOTS - 1)) == 0 && mid_max(L) < MID_SLOTS, "table size");
    const uint32_t lane = __lane_id();
    const uint64_t lt = lanemask_lt();
    const uint32_t nmid = *hl.mid_count[L];
    for (;;) {
        uint32_t q = 0;
        if (lane == 0) q = atomicAdd(hl.mid_next[L], 1u);
        q = __shfl(q, 0);
        if (q >= nmid) break;
        const uint32_t tile = hl.mid[L][q];
        const uint4 wd = wdesc[tile];
        const uint32_t n = wd.y;                 // n <= mid_max(L)
        const uint32_t seg = wd.x;
        const uint64_t abs = ((uint64_t)wd.w << 32) | wd.z;
        uint64_t* kp = keys + abs;
        for (uint32_t i = lane; i < MID_SLOTS; i += GKC_WAVE) { s_key[i] = EMPTY; s_cnt[i] = 0; }
        wave_lds_fence();    // other lanes' slots cleared
        for (uint32_t j = lane; j < n; j += GKC_WAVE) {
            const unsigned long long x = kp[j];
            uint32_t h = (uint32_t)((x * 0x9E3779B97F4A7C15ull) >> (64 - LOG));
            for (;;) {     // n <= mid_max(L) < MID_SLOTS keys: a free or matching slot exists
                const unsigned long long old = atomicCAS(reinterpret_cast<unsigned long long*>(&s_key[h]), EMPTY, x);
                if (old == EMPTY || old == x) { atomicAdd(&s_cnt[h], 1u); break; }
                h = (h + 1) & (MID_SLOTS - 1);
            }
        }
        wave_lds_fence();    // every lane's inserts
        // compact the kept distinct keys to the front (slot order)
        uint32_t m = 0, d = 0;
        for (uint32_t i0 = 0; i0 < MID_SLOTS; i0 += GKC_WAVE) {
            const uint32_t slot = i0 + lane;
            const uint64_t key = s_key[slot];
            const uint32_t c = s_cnt[slot];
            const bool occ = key != EMPTY;
            const bool keep = occ && c >= to.min_cnt && (to.max_cnt == 0 || c <= to.max_cnt);
            d += (uint32_t)__popcll(__ballot(occ));
            const uint64_t kb = __ballot(keep);
            if (keep) {
                const uint32_t p = m + (uint32_t)__popcll(kb & lt);
                s_key[p] = key; s_cnt[p] = c;      // p <= this slot's index: later slots are not disturbed
            }
            m += (uint32_t)__popcll(kb);
        }
        wave_lds_fence();    // compacted entries of every lane
        uint64_t* tk = alt + abs;
        uint32_t* tc = reinterpret_cast<uint32_t*>(kp);
        for (uint32_t e0 = 0; e0 < m; e0 += GKC_WAVE) {
            const uint32_t e = e0 + lane;
            uint64_t key = 0;
            if (e < m) key = s_key[e];
            const uint32_t r = count_below(s_key, m, key);
            if (e < m) { tk[r] = key; tc[r] = s_cnt[e]; }
        }
        if (lane == 0) {
            hl.kept[tile] = m;
            atomicAdd(&to.n_distinct[seg], d);
            if (m) atomicAdd(&to.n_entries[seg], m);
        }
        wave_lds_fence();    // this window's reads before the next clear
    }
}
template __global__ void k_midcount<0>(Plan*, SegArrays, uint64_t*, uint64_t*, const uint4*, TableOut, HeavyList);
template __global__ void k_midcount<1>(Plan*, SegArrays, uint64_t*, uint64_t*, const uint4*, TableOut, HeavyList);

// One wave's LDS hash table of a window (k_wavecount).  64-bit keys: the key
// word is the slot's state (EMPTY = all ones, never a canonical k-mer: its
// reverse complement, 0, is smaller) and is claimed by a 64-bit CAS.
// 128-bit keys: a 32-bit tag (0 = free, else a nonzero fingerprint of the
// key) is claimed by CAS and the winner writes the key; every lane then
// compares the full key.  All lanes of a wave run the loop in lockstep, so a
// key written in an iteration is visible to the compares of that iteration.
template <typename K> struct WaveTable;
template <> struct WaveTable<uint64_t> {
    uint64_t key[WC_CAP];
    uint32_t cnt[WC_CAP];
    static constexpr uint64_t EMPTY = ~0ull;
    __device__ void clear(uint32_t i) { key[i] = EMPTY; cnt[i] = 0; }
    __device__ bool used(uint32_t i) const { return key[i] != EMPTY; }
    // true when x took a free slot (a key not seen before); the table is its
    // first 2^lg slots
    __device__ bool insert(uint64_t x, int lg = __builtin_ctz(WC_CAP)) {
        const uint32_t mask = (1u << lg) - 1u;
        uint32_t h = (uint32_t)(((unsigned long long)x * 0x9E3779B97F4A7C15ull) >> (64 - lg));
        for (;;) {     // fewer distinct keys than slots: a free or matching slot exists
            const unsigned long long old = atomicCAS(reinterpret_cast<unsigned long long*>(&key[h]),
                                                     (unsigned long long)EMPTY, (unsigned long long)x);
            if (old == EMPTY || old == x) { atomicAdd(&cnt[h], 1u); return old == EMPTY; }
            h = (h + 1) & mask;
        }
    }
};
template <> struct WaveTable<u128k> {
    u128k key[WC_CAP];
    uint32_t cnt[WC_CAP];
    uint32_t tag[WC_CAP];
    __device__ void clear(uint32_t i) { tag[i] = 0; cnt[i] = 0; }
    __device__ bool used(uint32_t i) const { return tag[i] != 0; }
    __device__ bool insert(const u128k& x, int lg = __builtin_ctz(WC_CAP)) {
        const uint32_t mask = (1u << lg) - 1u;
        const unsigned long long hv = ((unsigned long long)x.lo ^ ((unsigned long long)x.hi * 0xC2B2AE3D27D4EB4Full)) *
                                      0x9E3779B97F4A7C15ull;
        uint32_t h = (uint32_t)(hv >> (64 - lg));
        const uint32_t fp = (uint32_t)hv | 1u;
        for (;;) {
            const uint32_t old = atomicCAS(&tag[h], 0u, fp);
            if (old == 0u) key[h] = x;
            // the winners' key stores before any lane's compare: keeps the
            // compiler from reading key[h] ahead of a store it does not make
            wave_lds_fence();
            if ((old == 0u || old == fp) && key[h] == x) { atomicAdd(&cnt[h], 1u); return old == 0u; }
            h = (h + 1) & mask;
        }
    }
};

// Wave-level binned count (see WC_TILE): every wave claims batches of
// windows and, per window range [a, a + n):
//   n <= WC_CAP: counts equal keys in its LDS hash table and ranks the kept
//                distinct keys by comparison (their place in key order), then
//                stages them like k_bincount (keys at alt[a + r],
//                counts as u32 over the range's own keys, read already);
//   n >  WC_CAP: takes {kept, distinct} k_bincount staged;
// and records the window's kept count in hl.kept[tile].  The per-segment
// distinct and kept totals are accumulated.  No window waits for another:
// k_tilescan turns the kept counts into table offsets, k_wavecopy writes.
// waves per SIMD the LDS tables allow (64-bit keys: 6 KB per wave; 128-bit: 12 KB)
template <typename K> constexpr int wavecount_waves() { return sizeof(K) == 16 ? 3 : 6; }
template <typename K>
__global__ __launch_bounds__(WC_NT, wavecount_waves<K>()) void k_wavecount(Plan* __restrict__ plan, SegArrays sa,
                                                     K* __restrict__ keys, K* __restrict__ alt,
                                                     const uint4* __restrict__ wdesc, HeavyList hl, TableOut to) {
    constexpr int NW = WC_NT / GKC_WAVE;
    constexpr int IT = WC_CAP / GKC_WAVE;
    static_assert((WC_CAP & (WC_CAP - 1)) == 0, "table size");
    __shared__ __align__(16) WaveTable<K> s_tab[NW];
    const uint32_t lane = __lane_id(), wv = threadIdx.x / GKC_WAVE;
    WaveTable<K>& T = s_tab[wv];
    uint32_t acc_seg = 0xFFFFFFFFu, acc_d = 0, acc_m = 0;     // per-segment totals, flushed on change
    constexpr bool MIDS = sizeof(K) == 8;                     // 64-bit keys: medium windows tried here
    uint32_t n_mid = 0;                                       // such windows taken (lane 0's count is used)
    uint32_t no_small = 0;                                    // windows left to count in the full table at once
    // windows in batches of WC_BATCH consecutive ones (segment order), the
    // following batch claimed when one starts; a window's descriptor is loaded
    // while the previous window is counted, and so are its keys for 64-bit
    // keys (128-bit ones would need more registers than their 3 waves a SIMD
    // leave)
    constexpr bool AHEAD = sizeof(K) == 8;
    const uint32_t TW = plan->tot_rl_tiles;
    uint32_t* ctr = &plan->ctr[ctr_rl()];
    uint32_t bt = 0, bt1 = 0, nb = 0;
    if (lane == 0) nb = atomicAdd(ctr, 1u);
    auto next_window = [&]() -> uint32_t {
        if (bt >= bt1) {
            const uint32_t b = __shfl(nb, 0);
            bt = min(b, TW / WC_BATCH + 1) * WC_BATCH;
            bt1 = min(TW, bt + WC_BATCH);
            if (bt >= TW) { bt1 = bt; return TW; }
            if (lane == 0) nb = atomicAdd(ctr, 1u);
        }
        return bt++;
    };
    // keys [0, min(n, WC_CAP)) of a window into x: the whole window when the
    // table takes it, a medium window's first chunk
    auto load_keys = [&](const uint4& w, K (&x)[IT]) {
        const uint32_t n = w.y;
        const K* kp = keys + (((uint64_t)w.w << 32) | w.z);
        if (n && (n <= (uint32_t)WC_CAP || (MIDS && hl.mid[0] && n <= WC_TRY_MAX))) {
#pragma unroll
            for (int i = 0; i < IT; i++) {
                const uint32_t j = i * GKC_WAVE + lane;
                if (i * GKC_WAVE < (int)n && j < n) x[i] = kp[j];
            }
        }
    };
    auto flush_acc = [&]() {
        if (lane == 0 && acc_seg != 0xFFFFFFFFu) {
            if (acc_d) atomicAdd(&to.n_distinct[acc_seg], acc_d);
            if (acc_m) atomicAdd(&to.n_entries[acc_seg], acc_m);
        }
    };
    uint32_t t = next_window();
    uint4 wd = t < TW ? wdesc[t] : make_uint4(0u, 0u, 0u, 0u);
    K x[IT];
    if constexpr (AHEAD) load_keys(wd, x);
    while (t < TW) {
        uint32_t tn = TW;
        uint4 wn = make_uint4(0u, 0u, 0u, 0u);
        if constexpr (AHEAD) {
            tn = next_window();
            if (tn < TW) wn = wdesc[tn];
        } else {
            load_keys(wd, x);
        }
        const uint32_t seg = wd.x, tile = t, n = wd.y;
        if (seg != acc_seg) {
            flush_acc();
            acc_seg = seg; acc_d = 0; acc_m = 0;
        }
        K* kp = keys + (((uint64_t)wd.w << 32) | wd.z);
        K* sk = alt + (((uint64_t)wd.w << 32) | wd.z);
        uint32_t* sc = reinterpret_cast<uint32_t*>(kp);
        uint32_t d = 0, m = 0;
        uint32_t slots = WC_CAP;                 // table slots in use
        bool handed = false;                     // more than WC_FILL distinct keys: k_midcount counts it
        const bool mid = MIDS && hl.mid[0] && n > (uint32_t)WC_CAP && n <= WC_TRY_MAX;
        if (n > (uint32_t)WC_CAP && !mid) {
            const uint2 hr = hl.res[tile];
            m = hr.x; d = hr.y;
        } else if (n) {
            // hash-count the range (its keys in x, loaded ahead) in the wave's table
            const bool small = !mid && no_small == 0;
#pragma unroll
            for (int i = 0; i < IT; i++) {
                const uint32_t j = i * GKC_WAVE + lane;
                if (!small || i * GKC_WAVE < (int)WC_SMALL) T.clear(j);
            }
            wave_lds_fence();    // other lanes' slots cleared
            if (small) {
                uint32_t taken = 0;
                bool over = false;
#pragma unroll
                for (int i = 0; i < IT; i++) {
                    if (i * GKC_WAVE >= (int)n) break;
                    if (taken > WC_SMALL_FILL) { over = true; break; }
                    bool fresh = false;
                    if (i * GKC_WAVE + lane < n) fresh = T.insert(x[i], __builtin_ctz(WC_SMALL));
                    taken += (uint32_t)__popcll(__ballot(fresh));
                }
                if (over) {                     // count again in the full table
                    no_small = 16;
#pragma unroll
                    for (int i = 0; i < IT; i++) T.clear(i * GKC_WAVE + lane);
                    wave_lds_fence();
#pragma unroll
                    for (int i = 0; i < IT; i++) {
                        if (i * GKC_WAVE < (int)n && i * GKC_WAVE + lane < n) T.insert(x[i]);
                    }
                } else {
                    slots = WC_SMALL;
                }
            } else if (!mid) {
                if (no_small) no_small--;
#pragma unroll
                for (int i = 0; i < IT; i++) {
                    if (i * GKC_WAVE < (int)n && i * GKC_WAVE + lane < n) T.insert(x[i]);
                }
            } else if constexpr (MIDS) {
                // WC_CAP-key chunks, 64 keys at a time while at most WC_FILL slots are taken
                n_mid++;
                uint32_t taken = 0;
                for (uint32_t c0 = 0; c0 < n && !handed; c0 += WC_CAP) {
                    if (c0) {
#pragma unroll
                        for (int i = 0; i < IT; i++) {
                            const uint32_t j = c0 + i * GKC_WAVE + lane;
                            if (c0 + i * GKC_WAVE < n && j < n) x[i] = kp[j];
                        }
                    }
#pragma unroll
                    for (int i = 0; i < IT; i++) {
                        if (c0 + i * GKC_WAVE >= n) break;
                        if (taken > WC_FILL) { handed = true; break; }
                        bool fresh = false;
                        if (c0 + i * GKC_WAVE + lane < n) fresh = T.insert(x[i]);
                        taken += (uint32_t)__popcll(__ballot(fresh));
                    }
                }
                if (handed && lane == 0) {
                    if (n <= MID_MAX) {
                        const int l = n <= mid_max(0) ? 0 : 1;
                        hl.mid[l][atomicAdd(hl.mid_count[l], 1u)] = tile;
                    } else {
                        hl.late[atomicAdd(hl.late_count, 1u)] = tile;
                    }
                }
            }
        }
        if constexpr (AHEAD) load_keys(wn, x);   // the next window's keys, in flight while this one is ranked
        if (n && !handed && !(n > (uint32_t)WC_CAP && !mid)) {
            // compact the kept distinct keys to the front of the table
#pragma unroll
            for (int i = 0; i < IT; i++) {
                if (i * GKC_WAVE >= (int)slots) break;
                const uint32_t slot = i * GKC_WAVE + lane;
                const bool occ = T.used(slot);
                K key{};
                uint32_t c = 0;
                if (occ) { key = T.key[slot]; c = T.cnt[slot]; }
                const bool keep = occ && c >= to.min_cnt && (to.max_cnt == 0 || c <= to.max_cnt);
                d += (uint32_t)__popcll(__ballot(occ));
                const uint64_t kb = __ballot(keep);
                if (keep) {
                    const uint32_t p = m + (uint32_t)__popcll(kb & lanemask_lt());
                    T.key[p] = key; T.cnt[p] = c;     // p <= this slot's index: reads of later slots are not disturbed
                }
                m += (uint32_t)__popcll(kb);
            }
            wave_lds_fence();    // compacted entries of every lane
            // rank of each kept key among the kept keys = its place in key order; stage
#pragma unroll
            for (int i = 0; i < IT; i++) {
                if (i * GKC_WAVE >= (int)m) continue;
                const uint32_t e = i * GKC_WAVE + lane;
                K key{};
                if (e < m) key = T.key[e];
                const uint32_t r = count_below(T.key, m, key);
                if (e < m) { sk[r] = key; sc[r] = T.cnt[e]; }
            }
        }
        if (lane == 0 && !handed) hl.kept[tile] = m;
        acc_d += d; acc_m += m;
        if constexpr (!AHEAD) {
            tn = next_window();
            if (tn < TW) wn = wdesc[tn];
        }
        t = tn;
        wd = wn;
    }
    flush_acc();
    if (MIDS && lane == 0 && n_mid && hl.mid_seen) atomicAdd(hl.mid_seen, n_mid);
}

// In-place exclusive scan of the kept counts of each segment's windows (one
// workgroup per segment): window offsets within the segment's table.
__global__ __launch_bounds__(1024) void k_tilescan(Plan* __restrict__ plan, SegArrays sa, uint32_t* __restrict__ kept) {
    __shared__ uint32_t s_w[1024 / GKC_WAVE];
    const uint32_t lane = __lane_id(), wv = threadIdx.x / GKC_WAVE;
    for (uint32_t seg = blockIdx.x; seg < plan->n_segs; seg += gridDim.x) {
        const uint32_t b = sa.rl_tile_start[seg], e = sa.rl_tile_start[seg + 1];
        uint32_t carry = 0;
        for (uint32_t base = b; base < e; base += 1024) {
            const uint32_t t = base + threadIdx.x;
            const uint32_t v = t < e ? kept[t] : 0u;
            uint32_t inc = v;
#pragma unroll
            for (int off = 1; off < GKC_WAVE; off <<= 1) {
                const uint32_t y = __shfl_up(inc, off);
                if ((int)lane >= off) inc += y;
            }
            if (lane == GKC_WAVE - 1) s_w[wv] = inc;
            __syncthreads();
            uint32_t off = 0, tot = 0;
#pragma unroll
            for (int w = 0; w < 1024 / GKC_WAVE; w++) { const uint32_t x = s_w[w]; off += (uint32_t)w < wv ? x : 0u; tot += x; }
            if (t < e) kept[t] = carry + off + inc - v;
            carry += tot;
            __syncthreads();
        }
    }
}

// Write each window's staged entries at its table offset (one wave per
// window): the binned count's emission; repeatable with other capacities.
template <typename K>
__global__ __launch_bounds__(256) void k_wavecopy(Plan* __restrict__ plan, SegArrays sa, const K* __restrict__ keys,
                                                  const K* __restrict__ alt, const uint4* __restrict__ wdesc,
                                                  HeavyList hl, TableOut to) {
    // a wave takes 64 consecutive windows: lane l reads window t0 + l's
    // offsets and its segment's bases (no chain of dependent loads per
    // window), then the wave copies the windows' entries U windows at a time,
    // the loads of all U in flight before their stores
    constexpr int U = 4;
    const uint32_t lane = __lane_id();
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) / GKC_WAVE;
    const uint32_t nw = gridDim.x * blockDim.x / GKC_WAVE;
    const uint32_t total = plan->tot_rl_tiles, S = plan->n_segs;
    for (uint32_t t0 = wave * GKC_WAVE; t0 < total; t0 += nw * GKC_WAVE) {
        const uint32_t t = t0 + lane;
        uint32_t off = 0, m = 0, a = 0, cap = 0;
        uint64_t kst = 0, tst = 0;
        if (t < total) {
            const uint4 wd = wdesc[t];
            const uint32_t seg = wd.x;
            const uint32_t last = sa.rl_tile_start[seg + 1] - 1;
            off = hl.kept[t];
            m = (t < last ? hl.kept[t + 1] : to.n_entries[seg]) - off;
            if (m) { kst = ((uint64_t)wd.w << 32) | wd.z; tst = to.table_start[seg]; cap = to.table_cap[seg]; }
        }
        const uint64_t over = __ballot(m && off + m > cap);
        if (over && lane == (uint32_t)__builtin_ctzll(over)) atomicOr(to.overflow, 1u);   // one flag per wave
        uint64_t todo = __ballot(m != 0);
        while (todo) {
            uint32_t mb[U], o[U], cp[U];
            uint64_t src[U], dst[U];
            K kv[U];
            uint32_t cv[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const uint32_t b = todo ? (uint32_t)__builtin_ctzll(todo) : 0u;
                const bool valid = todo != 0;
                todo &= todo - 1;
                mb[u] = valid ? __shfl(m, b) : 0u;
                o[u] = __shfl(off, b);
                cp[u] = __shfl(cap, b);
                src[u] = __shfl(kst, b) + __shfl(a, b);
                dst[u] = __shfl(tst, b);
                if (lane < mb[u] && o[u] + lane < cp[u]) {
                    kv[u] = alt[src[u] + lane];
                    cv[u] = reinterpret_cast<const uint32_t*>(keys + src[u])[lane];
                }
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
                if (lane < mb[u] && o[u] + lane < cp[u]) {
                    reinterpret_cast<K*>(to.keys)[dst[u] + o[u] + lane] = kv[u];
                    to.counts[dst[u] + o[u] + lane] = cv[u];
                }
            }
#pragma unroll
            for (int u = 0; u < U; u++) {             // windows of more than 64 entries: the rest
                for (uint32_t e = lane + GKC_WAVE; e < mb[u]; e += GKC_WAVE) {
                    const uint32_t idx = o[u] + e;
                    if (idx < cp[u]) {
                        reinterpret_cast<K*>(to.keys)[dst[u] + idx] = alt[src[u] + e];
                        to.counts[dst[u] + idx] = reinterpret_cast<const uint32_t*>(keys + src[u])[e];
                    }
                }
            }
        }
    }
}

template __global__ void k_binranges<uint64_t>(Plan*, SegArrays, const uint64_t*, uint4*, int, int, int, uint32_t, HeavyList);
template __global__ void k_binranges<u128k>(Plan*, SegArrays, const u128k*, uint4*, int, int, int, uint32_t, HeavyList);
template __global__ void k_bincount<uint64_t>(Plan*, SegArrays, uint64_t*, uint64_t*, const uint4*, TableOut, HeavyList, int);
template __global__ void k_bincount<u128k>(Plan*, SegArrays, u128k*, u128k*, const uint4*, TableOut, HeavyList, int);
template __global__ void k_wavecount<uint64_t>(Plan*, SegArrays, uint64_t*, uint64_t*, const uint4*, HeavyList, TableOut);
template __global__ void k_wavecount<u128k>(Plan*, SegArrays, u128k*, u128k*, const uint4*, HeavyList, TableOut);
template __global__ void k_wavecopy<uint64_t>(Plan*, SegArrays, const uint64_t*, const uint64_t*, const uint4*, HeavyList, TableOut);
template __global__ void k_wavecopy<u128k>(Plan*, SegArrays, const u128k*, const u128k*, const uint4*, HeavyList, TableOut);

template __global__ void k_extract<uint64_t>(Plan*, SegArrays, const uint32_t*, const SkmHdr*, const uint64_t*, const uint32_t*, const uint64_t*, uint64_t*, int);
template __global__ void k_extract<u128k>(Plan*, SegArrays, const uint32_t*, const SkmHdr*, const uint64_t*, const uint32_t*, const uint64_t*, u128k*, int);
#define GKC_OS(K, I, N, D) template __global__ void k_onesweep<K, I, N, D>(Plan*, SegArrays, WorkLists, const TileDesc*, const K*, K*, uint32_t*, int);
GKC_OS(uint64_t, 8, 256, 8) GKC_OS(uint64_t, 12, 256, 8) GKC_OS(uint64_t, 16, 256, 8) GKC_OS(uint64_t, 20, 256, 8)
GKC_OS(uint64_t, 24, 256, 8) GKC_OS(uint64_t, 8, 512, 8) GKC_OS(uint64_t, 10, 512, 8) GKC_OS(uint64_t, 12, 512, 8)
GKC_OS(uint64_t, 16, 512, 8) GKC_OS(uint64_t, 18, 512, 8)
GKC_OS(u128k, 4, 256, 8) GKC_OS(u128k, 8, 256, 8) GKC_OS(u128k, 12, 256, 8)
GKC_OS(uint64_t, 16, 256, 7) GKC_OS(uint64_t, 20, 256, 7) GKC_OS(uint64_t, 16, 256, 6) GKC_OS(uint64_t, 20, 256, 6)
GKC_OS(u128k, 8, 256, 7) GKC_OS(u128k, 12, 256, 7) GKC_OS(u128k, 8, 256, 6) GKC_OS(u128k, 12, 256, 6)
#undef GKC_OS
template __global__ void k_runlength<uint64_t>(Plan*, SegArrays, WorkLists, const uint64_t*, TableOut, uint32_t*);
template __global__ void k_runlength<u128k>(Plan*, SegArrays, WorkLists, const u128k*, TableOut, uint32_t*);

}  // namespace gkc
