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
__device__ __forceinline__ uint32_t find_read(const GKC_GLOBAL uint64_t* offs, uint32_t n, uint64_t pos) {
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

constexpr int P1_NCW = (P1_T + MAXK + 32) / 16 + 2;     // packed code words of a staged tile
constexpr int P1_MT = 8;                                  // tiles held per flush
constexpr int P1_RCAP = 3072;                             // superkmer records held per flush
// A flush follows any tile that leaves more than P1_RCAP - P1_T records, so
// the next tile's (at most P1_T) always fit.  Flush-local counts stay below
// 2^16: records <= P1_RCAP, k-mers <= P1_MT * P1_T, words <= k-mers / 16 +
// 6 per record.
static_assert(P1_RCAP % P1_NT == 0 && P1_RCAP >= P1_T, "record capacity");
static_assert(P1_MT * P1_T < 65536 && P1_MT * P1_T / 16 + 6 * P1_RCAP < 65536, "16-bit flush-local counts");
static_assert(P1_T <= 2048 && P1_MT <= 32, "record positions are slot << 11 | start in 16 bits");

// base_code of 4 bytes (byte 0 first) packed into 8 bits, byte 0 in the top two.
__device__ __forceinline__ uint32_t pack4(uint32_t v) {
    const uint32_t e = ((v >> 2) ^ (v >> 1)) & 0x03030303u;   // code of byte i at bits [8i, 8i + 2)
    return (e * 0x40100401u) >> 24;                            // byte i's code to bits [30 - 2i, 32 - 2i)
}
__device__ __forceinline__ uint32_t pack16(const uint4 v) {
    return (pack4(v.x) << 24) | (pack4(v.y) << 16) | (pack4(v.z) << 8) | pack4(v.w);
}
// Bit i set when byte i of v is not one of A C G T a c g t: the byte is
// compared with the lowercase letter of its own 2-bit code (v_perm picks
// "acgt"[code] per byte), so only those eight bytes give zero.
__device__ __forceinline__ uint32_t bad4(uint32_t v) {
    const uint32_t e = ((v >> 2) ^ (v >> 1)) & 0x03030303u;
    const uint32_t want = __builtin_amdgcn_perm(0u, 0x74676361u, e);    // 't' 'g' 'c' 'a'
    const uint32_t x = (v | 0x20202020u) ^ want;                          // zero byte <=> ACGT
    const uint32_t nz = (((x & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | x) & 0x80808080u;
    return ((nz >> 7) * 0x01020408u) >> 24;                               // byte i's flag to bit i
}
__device__ __forceinline__ uint32_t bad16(const uint4 v) {
    return bad4(v.x) | (bad4(v.y) << 4) | (bad4(v.z) << 8) | (bad4(v.w) << 12);
}
// Sketch sample of a superkmer: its first min(k, 16) bases (2-bit codes,
// first base highest) hashed by a multiplicative (Fibonacci) hash, kept when
// the top three bits are zero (one superkmer in eight, decided by its bases
// alone, so it does not depend on the launch shape).
__device__ __forceinline__ bool sketch_sampled(uint32_t lead) { return ((lead * 0x9E3779B1u) >> 29) == 0u; }
// Codes of tile bases [j, j + n), 1 <= n <= 16, in the low 2n bits, base j highest.
__device__ __forceinline__ uint32_t tile_codes(const uint32_t* sc, uint32_t j, uint32_t n) {
    const uint64_t x = ((uint64_t)sc[j >> 4] << 32) | sc[(j >> 4) + 1];
    return (uint32_t)((x << (2 * (j & 15))) >> (64 - 2 * n));
}
// The p-mer at tile base j (p <= 16 in 32 bits, p <= 32 in 64 bits).
template <typename MM>
__device__ __forceinline__ MM pmer_at(const uint32_t* sc, uint32_t j, int p);
template <>
__device__ __forceinline__ uint32_t pmer_at<uint32_t>(const uint32_t* sc, uint32_t j, int p) {
    return tile_codes(sc, j, (uint32_t)p);
}
template <>
__device__ __forceinline__ uint64_t pmer_at<uint64_t>(const uint32_t* sc, uint32_t j, int p) {
    const uint32_t q = j >> 4, sh = 2 * (j & 15);
    uint64_t x = ((uint64_t)sc[q] << 32) | sc[q + 1];
    if (sh) x = (x << sh) | (sc[q + 2] >> (32 - sh));
    return x >> (64 - 2 * p);
}
// Reverse complement of a p-mer held in the low 2p bits.
__device__ __forceinline__ uint32_t rc_pmer(uint32_t f, int p) {
    uint32_t y = __builtin_bitreverse32(~f);                      // bases reversed, each code's bits swapped
    y = ((y >> 1) & 0x55555555u) | ((y & 0x55555555u) << 1);
    return y >> (32 - 2 * p);
}
__device__ __forceinline__ uint64_t rc_pmer(uint64_t f, int p) { return revcomp64(f, p); }
// new_filter2 (gkc_cuda.cu:214-216): a p-mer whose first three bases are
// AAA, ACA or AGA (codes 0, 4, 8) is not a candidate; mode 0 accepts all.
template <typename MM>
__device__ __forceinline__ bool pmer_ok(MM v, int p, int mode) {
    const uint32_t top = (uint32_t)(v >> (2 * (p - 3)));
    return mode == 0 || ((0xFFFFFFFFFFFFFEEEull >> top) & 1ull);
}
template <typename MM>
__device__ __forceinline__ uint32_t part_mm(MM mm, uint32_t P, int mode) {
    if constexpr (sizeof(MM) == 4) {
        if (mode != 1) return mm % P;
    }
    return part_of((uint64_t)mm, P, mode);
}

// One tile = P1_T window starts, P1_V per thread.  The tile's bases are
// staged in LDS as packed 2-bit codes; every thread takes its first p-mer
// from them directly and rolls the rest.  Tiles go to workgroups round-robin
// and each workgroup loads the next tile's bases and read offsets into
// registers while it works on the current one.
// A workgroup keeps the superkmers of up to P1_MT tiles (their codes and one
// record each) in LDS and reserves bin space for all of them at once: one
// global atomic per (flush, partition, cursor) instead of per tile, since
// device-scope atomics, not the arithmetic, bound this kernel otherwise.
// With WIN, every candidate carries its position: the window minimum runs
// over composites (value << 14 | no-candidate << 13 | 8191 - start), so ties
// go to the later p-mer start (the reference's update order, gkc_cuda.cu:281-284)
// and the chosen start and strand are recovered per window.
// Dynamic LDS: P x (8 + 4) bytes (per-partition counters).
template <typename MM, bool WIN>
__global__ __launch_bounds__(P1_NT, (sizeof(MM) == 8 ? 3 : 4)) void k_superkmers(const ChunkDesc* __restrict__ cdp, P1Params prm, BinSet bins) {
    constexpr int CSH = WIN ? 14 : 0;
    constexpr int NQ = P1_RCAP / P1_NT;             // records per thread in a flush
    __shared__ int8_t s_str[WIN ? P1_T + MAXK : 1];
    __shared__ uint32_t s_code[P1_MT][P1_NCW];      // staged tiles, one slot each
    __shared__ uint16_t s_bad[P1_NCW];              // the current tile's non-ACGT bytes, a bit per base
    __shared__ MM s_v[P1_T + MAXK];
    __shared__ MM s_last[P1_NT];
    // superkmer records: slot << 11 | tile-local start; k-mers (window end
    // until the tile's records are complete); partition
    __shared__ uint16_t s_rpos[P1_RCAP], s_rn[P1_RCAP], s_rpart[P1_RCAP];
    __shared__ uint32_t s_scan[P1_NT / GKC_WAVE];
    __shared__ uint32_t s_ro[P1_RO];              // offsets (chunk positions) of reads r0 .. r0 + n_ro - 1
    __shared__ uint32_t s_ns, s_ovf;
    __shared__ uint32_t s_cr[2];                    // reads holding the current tile's window starts: first, last
    __shared__ uint32_t s_slot_tile[P1_MT];         // the tile held in each slot (superkmer positions)
    // per partition: flush-local [47:32] headers, [31:16] k-mers, [15:0] words;
    // then the flush's global bases [63:32] words, [31:0] headers
    extern __shared__ unsigned long long s_loc[];
    const uint32_t P = prm.P;
    uint32_t* s_gk = reinterpret_cast<uint32_t*>(s_loc + P);

    const ChunkDesc cd = *cdp;
    const int tid = threadIdx.x;
    const int k = prm.k, p = prm.p, w = k - p + 1;
    const MM M = (MM)((p >= (int)(sizeof(MM) * 4)) ? ~(MM)0 : (((MM)1 << (2 * p)) - 1));
    const MM MC = WIN ? (MM)((M << CSH) | (MM)0x3FFF) : M;     // largest composite
    const uint32_t G = gridDim.x;
    uint32_t tile = blockIdx.x;
    if (tile >= cd.n_tiles) return;

    const GKC_GLOBAL uint8_t* g_reads = as_global(cd.reads);
    const GKC_GLOBAL uint64_t* g_offs = as_global(cd.offs);
    const GKC_GLOBAL uint32_t* g_offs32 = as_global(reinterpret_cast<const uint32_t*>(cd.offs));   // low words at 2r
    const GKC_GLOBAL int32_t* g_lens = as_global(cd.lens);
    const GKC_GLOBAL uint32_t* g_tr0 = as_global(cd.tile_r0);
    const uint32_t base32 = (uint32_t)cd.base;      // chunk positions: low words of offsets minus this

    // The next tile's inputs load while this one is processed.  No loaded
    // value is computed with at load time (a use would wait for the load at
    // once): read offsets stay raw until staged, and the read holding the
    // tile after next is looked up at a clamped index, its end case applied
    // when it is used.
    // reads holding tile t's window starts: a .. b (b holds the next tile's first position)
    auto tile_reads_raw = [&](uint32_t t, uint32_t& a, uint32_t& braw) {
        uint32_t tv = min(t, cd.n_tiles - 1);
        asm volatile("" : "+v"(tv));                // per-lane loads, off the scalar cache path
        a = g_tr0[tv];
        braw = g_tr0[min(tv + 1, cd.n_tiles - 1)];
    };
    auto tile_b = [&](uint32_t t, uint32_t braw) { return t + 1 < cd.n_tiles ? braw : cd.n_reads - 1; };
    uint4 pf = make_uint4(0u, 0u, 0u, 0u);          // this thread's 16 bases of the tile
    uint32_t pro0 = 0, pro1 = 0;                    // low words of read offsets tid, tid + P1_NT of the tile
    // (offsets first: the range they need is worked out before any load of this call is in flight)
    auto fetch = [&](uint32_t t, uint32_t q0, uint32_t q1) {
        const uint32_t tt0 = t * P1_T;
        const uint32_t bl = min((uint32_t)(P1_T + MAXK + 32), cd.n_bases - tt0);
        const uint32_t n = q1 - q0 + 2;
        if (n <= P1_RO) {
            if ((uint32_t)tid < n) pro0 = g_offs32[2 * (q0 + tid)];
            if ((uint32_t)tid + P1_NT < n) pro1 = g_offs32[2 * (q0 + tid + P1_NT)];
        }
        if (tid < P1_NCW && (uint32_t)tid * 16 + 16 <= bl) {
            typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
            const u32x4 v = reinterpret_cast<const GKC_GLOBAL u32x4*>(g_reads + tt0)[tid];
            pf = make_uint4(v.x, v.y, v.z, v.w);
        }
    };
    // the current tile's read range passes from one iteration to the next in
    // LDS (a register copy of the value just loaded for the tile after next
    // would wait for that load)
    uint32_t n_r0 = 0, n_r1 = 0;                    // the tile after this one; n_r1 raw (tile_b applies the end case)
    {
        uint32_t a, braw;
        tile_reads_raw(tile, a, braw);
        const uint32_t b = tile_b(tile, braw);
        fetch(tile, a, b);
        if (tid == 0) { s_cr[0] = a; s_cr[1] = b; }
    }
    tile_reads_raw(tile + G, n_r0, n_r1);
    for (uint32_t i = tid; i < P; i += P1_NT) s_loc[i] = 0;
    uint32_t acc = 0, slot = 0;                     // records and tiles held
    __syncthreads();

    for (;;) {
        const uint32_t t0 = tile * P1_T;
        const uint32_t tlen = min((uint32_t)P1_T, cd.n_bases - t0);      // window starts owned
        const uint32_t blen = min((uint32_t)(P1_T + MAXK + 32), cd.n_bases - t0);  // bases staged
        const uint32_t r0 = s_cr[0], r1 = s_cr[1];
        const uint32_t n_ro = r1 - r0 + 2;                                 // offsets r0 .. r1 + 1
        const bool ro_lds = n_ro <= P1_RO;
        uint32_t* const sc = s_code[slot];
        if (tid == 0) s_slot_tile[slot] = tile;

        // 1. stage: packed codes of bases [t0, t0 + blen) (code 0 past blen), read offsets
        uint32_t badw = 0;                          // non-ACGT bytes of this thread's 16 (0 past blen)
        if (tid < P1_NCW) {
            const uint32_t b = tid * 16;
            uint32_t wd = 0;
            if (b + 16 <= blen) {
                wd = pack16(pf);
                badw = bad16(pf);
            } else {
                for (uint32_t t = 0; t < 16; t++)
                    if (b + t < blen) {
                        const uint32_t ch = g_reads[t0 + b + t];
                        wd |= base_code(ch) << (30 - 2 * t);
                        badw |= (bad4(ch | 0x41414100u) & 1u) << t;     // byte 0 tested, others 'A'
                    }
            }
            sc[tid] = wd;
            s_bad[tid] = (uint16_t)badw;
        }
        if (ro_lds) {
            if ((uint32_t)tid < n_ro) s_ro[tid] = pro0 - base32;
            if ((uint32_t)tid + P1_NT < n_ro) s_ro[tid + P1_NT] = pro1 - base32;
        }
        const bool tile_bad = __syncthreads_or(badw != 0);    // any non-ACGT byte staged: reads split there
        const uint32_t nt = tile + G;
        const bool more = nt < cd.n_tiles;
        if (more) {                                 // the next tile's inputs load during this one
            const uint32_t nb = tile_b(nt, n_r1);
            fetch(nt, n_r0, nb);
            if (tid == 0) { s_cr[0] = n_r0; s_cr[1] = nb; }     // read after this tile's later barriers
            tile_reads_raw(nt + G, n_r0, n_r1);
        }

        // 2. candidate value per p-mer start j: min over the passing forward and
        //    reverse-complement p-mer (M if neither passes), rolled over V starts.
        auto fill_v = [&](int j0, int cnt) {
            MM f = pmer_at<MM>(sc, j0, p);
            MM r = rc_pmer(f, p);
            uint32_t rw = tile_codes(sc, j0 + p, 16);           // bases rolled in next
            const int rs = 2 * (p - 1);
            for (int s = 0; s < cnt; s++) {
                if (s > 0) {
                    const uint32_t c = rw >> 30;
                    rw <<= 2;
                    f = ((f << 2) | c) & M;
                    r = (r >> 2) | ((MM)(3u - c) << rs);
                }
                const bool fp = pmer_ok(f, p, prm.fmode), rp = pmer_ok(r, p, prm.fmode);
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

        // 3b. a tile holding a non-ACGT byte: for each window start, the first
        //     such base at or after it (in the candidate array's LDS, free now):
        //     per staged word its first flagged base, the suffix minimum over
        //     the words (one wave), then each position from its own word
        uint16_t* const s_nb = reinterpret_cast<uint16_t*>(s_v);
        static_assert(P1_T * 2 + P1_NCW * 4 <= (P1_T + MAXK) * sizeof(uint32_t), "first-bad arrays fit in s_v");
        if (tile_bad) {
            uint32_t* const s_wf = reinterpret_cast<uint32_t*>(s_nb + P1_T);
            if (tid < P1_NCW) { const uint32_t m = s_bad[tid]; s_wf[tid] = m ? tid * 16 + (uint32_t)__builtin_ctz(m) : 0xFFFFu; }
            __syncthreads();
            if (tid < GKC_WAVE) {
                uint32_t carry = 0xFFFFu;
                for (int q0 = ((P1_NCW - 1) / GKC_WAVE) * GKC_WAVE; q0 >= 0; q0 -= GKC_WAVE) {
                    const int q = q0 + tid;
                    uint32_t v = q < P1_NCW ? s_wf[q] : 0xFFFFu;
#pragma unroll
                    for (int off = 1; off < GKC_WAVE; off <<= 1) v = min(v, (uint32_t)__shfl_down(v, off));
                    v = min(v, carry);
                    if (q < P1_NCW) s_wf[q] = v;
                    carry = __shfl(v, 0);
                }
            }
            __syncthreads();
            for (int s = 0; s < P1_V; s++) {
                const uint32_t g = tid * P1_V + s;      // < P1_T, so word g / 16 + 1 is staged
                const uint32_t m = (uint32_t)s_bad[g >> 4] >> (g & 15);
                s_nb[g] = (uint16_t)(m ? g + (uint32_t)__builtin_ctz(m) : s_wf[(g >> 4) + 1]);
            }
            __syncthreads();
        }

        // 4. validity and boundaries (chunk positions fit 32 bits)
        uint32_t bmask = 0;
        uint32_t wend[P1_V] = {};
        {
            const uint32_t g0 = t0 + tid * P1_V;
            uint32_t r = 0, rbeg = 0, rnext = 0, rend = 0;
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
                    r = find_read(g_offs, cd.n_reads, cd.base + gs);
                }
                rbeg = ro_lds ? s_ro[r - r0] : (uint32_t)(g_offs[r] - cd.base);
                rnext = ro_lds ? s_ro[r - r0 + 1] : (uint32_t)(g_offs[r + 1] - cd.base);
                rend = cd.lens ? rbeg + (uint32_t)g_lens[r] : rnext;
            }
            MM prev = tid > 0 ? s_last[tid - 1] : (MM)0;
            MM cmp[P1_V];
#pragma unroll
            for (int s = 0; s < P1_V; s++) cmp[s] = mm[s];
#pragma unroll
            for (int s = 0; s < P1_V; s++) mm[s] = cmp[s] >> CSH;      // minimizer values
            // Positions in at most two reads (read r, then r + 1 from rnext on),
            // no per-position outputs: the valid and read-start positions as bit
            // ranges per read.  A position is valid when it starts a k-mer inside
            // its read (rbeg <= g, g + k <= rend).
            const uint32_t nv = tid * P1_V < (int)tlen ? min((uint32_t)P1_V, tlen - tid * P1_V) : 0u;
            bool slow = nv != 0;
            if (!WIN && !prm.mm_out && ro_lds && nv && !tile_bad) {
                const uint32_t gend = g0 + nv;
                auto vbits = [&](uint32_t rb, uint32_t rn, uint32_t re) -> uint32_t {
                    const uint32_t lo = max(g0, rb);
                    if (re + 1 < (uint32_t)k + lo) return 0u;
                    const uint32_t hi = min(min(gend, rn), re + 1 - (uint32_t)k);
                    return hi <= lo ? 0u : ((1u << (hi - g0)) - 1u) & ~((1u << (lo - g0)) - 1u);
                };
                uint32_t split = P1_V, rnext2 = 0, rend2 = 0;
                bool two = true;
                if (rnext < gend) {             // r + 1 <= the tile's last read: its end is staged
                    split = rnext - g0;
                    rnext2 = s_ro[r - r0 + 2];
                    two = rnext2 >= gend;
                    rend2 = cd.lens ? rnext + (uint32_t)g_lens[r + 1] : rnext2;
                }
                if (two) {
                    slow = false;
                    const uint32_t va = vbits(rbeg, rnext, rend);
                    const uint32_t vb = split < P1_V ? vbits(rnext, rnext2, rend2) : 0u;
                    uint32_t chg = 0;
#pragma unroll
                    for (int s = 0; s < P1_V; s++) { chg |= (mm[s] != prev ? 1u : 0u) << s; prev = mm[s]; }
                    uint32_t fm = g0 == t0 ? 1u : 0u;
                    if (rbeg >= g0 && rbeg < gend) fm |= 1u << (rbeg - g0);
                    if (split < P1_V) fm |= 1u << split;
                    bmask = (va | vb) & (fm | chg);
                    const uint32_t wa = va ? min(rend + 1 - (uint32_t)k - t0, tlen) : 0u;
                    const uint32_t wb = vb ? min(rend2 + 1 - (uint32_t)k - t0, tlen) : 0u;
#pragma unroll
                    for (int s = 0; s < P1_V; s++) wend[s] = (uint32_t)s < split ? wa : wb;
                }
            }
#pragma unroll
            for (int s = 0; s < P1_V && slow; s++) {
                const uint32_t g = g0 + s;
                wend[s] = 0;
                if (tid * P1_V + s < (int)tlen) {
                    while (g >= rnext) {            // advance (skips empty reads)
                        r++;
                        rbeg = rnext;
                        rnext = ro_lds ? s_ro[r - r0 + 1] : (uint32_t)(g_offs[r + 1] - cd.base);
                        rend = cd.lens ? rbeg + (uint32_t)g_lens[r] : rnext;
                    }
                    const uint32_t gl = g - t0;     // tile-local
                    const bool valid = g >= rbeg && g < rend && rend - g >= (uint32_t)k &&
                                       !(tile_bad && s_nb[gl] < gl + (uint32_t)k);
                    if (valid) {
                        const bool after_bad = tile_bad && g > rbeg && gl > 0 && s_nb[gl - 1] == gl - 1;
                        const bool first = (g == rbeg) || (g == t0) || after_bad;
                        if (first || mm[s] != prev) bmask |= 1u << s;
                        // tile-local end of windows: the read's last window, or the window
                        // before the next non-ACGT base
                        uint32_t we = min(rend + 1 - (uint32_t)k - t0, tlen);
                        if (tile_bad) we = min(we, (uint32_t)s_nb[gl] + 1 - (uint32_t)k);
                        wend[s] = we;
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
                    }
                }
                prev = mm[s];
            }
        }

        // 5. append the tile's superkmer starts to the records, in tile order
        //    (wave scans + wave sums)
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
            uint32_t o = acc + inc - c, tot = 0;
#pragma unroll
            for (int w2 = 0; w2 < P1_NT / GKC_WAVE; w2++) { const uint32_t x = s_scan[w2]; o += (uint32_t)w2 < wv ? x : 0u; tot += x; }
            if (tid == 0) s_ns = tot;
            auto emit = [&](auto part) {
#pragma unroll
                for (int s = 0; s < P1_V; s++) {
                    if (bmask & (1u << s)) {
                        s_rpos[o] = (uint16_t)((slot << 11) | (tid * P1_V + s));
                        s_rn[o] = (uint16_t)wend[s];
                        s_rpart[o] = (uint16_t)part(mm[s]);
                        o++;
                    }
                }
            };
            if ((P & (P - 1)) == 0 && prm.pmode != 1) emit([&](MM v) { return (uint32_t)v & (P - 1); });
            else emit([&](MM v) { return part_mm<MM>(v, P, prm.pmode); });
        }
        __syncthreads();
        // 6. k-mers of each new record: up to the next superkmer's start or its read's last window
        const uint32_t ns = s_ns;
        for (uint32_t i = acc + tid; i < acc + ns; i += P1_NT) {
            const uint32_t a = s_rpos[i] & 2047u;
            uint32_t e = (i + 1 < acc + ns) ? (uint32_t)(s_rpos[i + 1] & 2047u) : tlen;
            e = min(e, (uint32_t)s_rn[i]);
            s_rn[i] = (uint16_t)(e - a);
        }
        acc += ns;
        slot++;

        // 7. flush: reserve bin space for the held records, then write them
        if (!more || slot == P1_MT || acc > P1_RCAP - P1_T) {
            __syncthreads();
            // flush-local offsets: one 64-bit LDS atomic per record reserves its
            // header, k-mer and word slots together, so the three run in the
            // same order inside a partition
            unsigned long long loc[NQ];
#pragma unroll
            for (int q = 0; q < NQ; q++) {
                const uint32_t i = tid + q * P1_NT;
                loc[q] = 0;
                if (i < acc) {
                    const unsigned long long n = s_rn[i];
                    const unsigned long long nw = (n + k - 1 + 15) >> 4;
                    loc[q] = atomicAdd(&s_loc[s_rpart[i]], (1ull << 32) | (n << 16) | nw);
                }
            }
            if (tid == 0) s_ovf = 0;
            __syncthreads();
            // one global reservation per (flush, partition)
            for (uint32_t q = tid; q < P; q += P1_NT) {
                const unsigned long long v = s_loc[q];
                const uint32_t h = (uint32_t)(v >> 32), kk = (uint32_t)(v >> 16) & 0xFFFFu, ww = (uint32_t)v & 0xFFFFu;
                if (h) {
                    // headers and k-mers in one atomic: k-mer offsets then rise with
                    // the header index inside a partition, across flushes too
                    const unsigned long long ghk = atomicAdd(&bins.hk_cur[q], ((unsigned long long)kk << 32) | h);
                    const uint32_t gh = (uint32_t)ghk, gk = (uint32_t)(ghk >> 32);
                    const uint32_t gw = atomicAdd(&bins.word_cur[q], ww);
                    if (prm.scatter && (gh + h > bins.hdr_cap[q] || gw + ww > bins.word_cap[q])) {
                        s_ovf = 1;
                        atomicOr(bins.overflow, 1u);
                    }
                    s_loc[q] = ((unsigned long long)gw << 32) | gh;
                    s_gk[q] = gk;
                }
            }
            __syncthreads();
            // headers and 2-bit packed bases (funnel shifts of the staged codes)
            if (prm.scatter && !s_ovf) {
#pragma unroll
                for (int q = 0; q < NQ; q++) {
                    const uint32_t i = tid + q * P1_NT;
                    if (i < acc) {
                        const uint32_t rp = s_rpos[i];
                        const uint32_t* csl = s_code[rp >> 11];
                        const uint32_t a = rp & 2047u;
                        const uint32_t n = s_rn[i];
                        const uint32_t nb = n + k - 1;
                        const uint32_t nw = (nb + 15) >> 4;
                        const uint32_t part = s_rpart[i];
                        const unsigned long long gb = s_loc[part];
                        const uint32_t wo = (uint32_t)(gb >> 32) + (uint32_t)(loc[q] & 0xFFFFu);
                        SkmHdr h;
                        h.word_off = wo;
                        h.kmer_off = s_gk[part] + (uint32_t)((loc[q] >> 16) & 0xFFFFu);
                        h.n_kmers = n;
                        h.reserved = 0;
                        const uint64_t hix = bins.hdr_base[part] + (uint32_t)gb + (uint32_t)(loc[q] >> 32);
                        bins.hdrs[hix] = h;
                        if (bins.pos) {     // the superkmer's read and in-read offset (search of the chunk's offsets)
                            const uint32_t cp = s_slot_tile[rp >> 11] * P1_T + a;
                            const uint32_t r = find_read(g_offs, cd.n_reads, cd.base + cp);
                            bins.pos[hix] = make_uint2(cd.read0 + r, cp - (uint32_t)(g_offs[r] - cd.base));
                        }
                        if (prm.sketch && k >= 4) {          // sampled superkmers' leading 4-mers
                            const uint32_t lead = tile_codes(csl, a, (uint32_t)min(k, 16));
                            if (sketch_sampled(lead))
                                atomicAdd(&prm.sketch[(size_t)part * 256 + (lead >> (2 * (min(k, 16) - 4)))], 1u);
                        }
                        // word wi = bases [a + 16 wi, a + 16 wi + 16): every word of a record
                        // has the same bit offset in the staged codes, so each is a funnel
                        // shift of two consecutive staged words, one of them the previous one's
                        uint32_t* dst = bins.words + bins.word_base[part] + wo;
                        const uint32_t q0 = a >> 4, sh = 2 * (a & 15);
                        uint32_t cur = csl[q0];
                        for (uint32_t wi = 0; wi < nw; wi++) {
                            const uint32_t nxt = csl[q0 + wi + 1];
                            uint32_t x = sh ? __builtin_amdgcn_alignbit(cur, nxt, 32 - sh) : cur;
                            const uint32_t rem = nb - wi * 16;
                            if (rem < 16) x &= ~0u << (32 - 2 * rem);     // zero past the last base
                            dst[wi] = x;
                            cur = nxt;
                        }
                    }
                }
            }
            __syncthreads();
            for (uint32_t i = tid; i < P; i += P1_NT) s_loc[i] = 0;
            acc = 0;
            slot = 0;
        }
        __syncthreads();
        if (!more) break;
        tile = nt;
    }
}

// Homopolymer compression (gkc_cuda.cu:178-190), one wave per read: keep
// byte j when it differs from byte j-1; out[beg + h] = byte of run h,
// lens[r] = runs, and (when orig_pos is not null) orig_pos[beg + h] = the
// in-read index of run h's first byte.  Each round a wave takes 1024 bytes,
// 16 per lane, read as the one or two aligned 16-byte blocks holding them
// (each holds a byte of the read, so no load leaves its buffer); a lane
// finds its kept bytes by comparing each byte with the one before (the
// previous lane's last byte, or the previous round's), places them in the
// wave's LDS buffer at their rank, and the wave stores the round's output
// from there, consecutive bytes across the lanes.
constexpr int HPC_NT = 256;
__device__ __forceinline__ uint32_t sel4(uint32_t q, uint32_t a, uint32_t b, uint32_t c, uint32_t d) {
    return q == 0 ? a : q == 1 ? b : q == 2 ? c : d;
}
__global__ __launch_bounds__(HPC_NT) void k_hpc(const uint8_t* __restrict__ reads, const uint64_t* __restrict__ offs,
                                                uint32_t n_reads, uint8_t* __restrict__ out,
                                                int32_t* __restrict__ lens, int32_t* __restrict__ orig_pos) {
    __shared__ uint8_t s_out[HPC_NT / GKC_WAVE][16 * GKC_WAVE];
    uint8_t* so = s_out[threadIdx.x / GKC_WAVE];
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) / GKC_WAVE;
    const uint32_t lane = __lane_id();
    const uint32_t nwaves = gridDim.x * blockDim.x / GKC_WAVE;
    const uint64_t base = offs[0];
    for (uint32_t r = wave; r < n_reads; r += nwaves) {
        const uint64_t beg = offs[r] - base, L = offs[r + 1] - offs[r];
        const uint8_t* rd = reads + beg;
        uint32_t cnt = 0;                       // runs so far
        uint32_t carry = 0x100u;                // the previous round's last byte (none: equal to no byte)
        for (uint64_t j0 = 0; j0 < L; j0 += 16 * GKC_WAVE) {
            const uint64_t j = j0 + 16u * lane;
            const uint32_t nb = j < L ? (uint32_t)(L - j < 16 ? L - j : 16) : 0u;    // this lane's bytes
            uint32_t w0 = 0, w1 = 0, w2 = 0, w3 = 0;    // bytes j .. j+15, byte j lowest
            if (nb) {
                const uintptr_t a0 = reinterpret_cast<uintptr_t>(rd + j);
                const uintptr_t al = a0 & ~(uintptr_t)15;
                const uint4 b0 = *reinterpret_cast<const uint4*>(al);
                uint4 b1 = make_uint4(0u, 0u, 0u, 0u);
                if (((a0 + nb - 1) & ~(uintptr_t)15) != al) b1 = *reinterpret_cast<const uint4*>(al + 16);
                const uint32_t sh = (uint32_t)(a0 & 15), q = sh >> 2, sb = sh & 3;
                // dwords q .. q+4 of the 32 bytes b0:b1, then shifted by sb bytes
                const uint32_t d0 = sel4(q, b0.x, b0.y, b0.z, b0.w), d1 = sel4(q, b0.y, b0.z, b0.w, b1.x);
                const uint32_t d2 = sel4(q, b0.z, b0.w, b1.x, b1.y), d3 = sel4(q, b0.w, b1.x, b1.y, b1.z);
                const uint32_t d4 = sel4(q, b1.x, b1.y, b1.z, b1.w);
                w0 = __builtin_amdgcn_alignbyte(d1, d0, sb);
                w1 = __builtin_amdgcn_alignbyte(d2, d1, sb);
                w2 = __builtin_amdgcn_alignbyte(d3, d2, sb);
                w3 = __builtin_amdgcn_alignbyte(d4, d3, sb);
            }
            // the byte before each of the 16: the previous lane's last, or the carry
            const uint32_t up = __shfl_up(w3 >> 24, 1);
            const uint32_t pb = lane == 0 ? carry : up;
            const uint32_t p0 = pb > 0xFFu ? (w0 ^ 1u) & 0xFFu : pb;         // no byte before: keep byte 0
            const uint32_t y0 = (w0 << 8) | p0, y1 = __builtin_amdgcn_alignbyte(w1, w0, 3),
                           y2 = __builtin_amdgcn_alignbyte(w2, w1, 3), y3 = __builtin_amdgcn_alignbyte(w3, w2, 3);
            auto nz4 = [](uint32_t d) {        // bit i: byte i of d nonzero
                const uint32_t t = (((d & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | d) & 0x80808080u;
                return ((t >> 7) * 0x01020408u) >> 24;
            };
            uint32_t keep = nz4(w0 ^ y0) | (nz4(w1 ^ y1) << 4) | (nz4(w2 ^ y2) << 8) | (nz4(w3 ^ y3) << 12);
            keep &= nb >= 16 ? 0xFFFFu : ((1u << nb) - 1u);
            const uint32_t c = (uint32_t)__popc(keep);
            uint32_t inc = c;
#pragma unroll
            for (int off = 1; off < GKC_WAVE; off <<= 1) {
                const uint32_t y = __shfl_up(inc, off);
                if ((int)lane >= off) inc += y;
            }
            const uint32_t tot = __shfl(inc, GKC_WAVE - 1);
            uint32_t o = inc - c;
            for (uint32_t m = keep; m; m &= m - 1) {
                const uint32_t t = (uint32_t)__builtin_ctz(m);
                const uint32_t wd = sel4(t >> 2, w0, w1, w2, w3);
                so[o] = (uint8_t)(wd >> (8 * (t & 3)));
                if (orig_pos) orig_pos[beg + cnt + o] = (int32_t)(j + t);
                o++;
            }
            wave_lds_fence();       // every lane's bytes placed
            for (uint32_t t = lane; t < tot; t += GKC_WAVE) out[beg + cnt + t] = so[t];
            wave_lds_fence();       // read before the next round's placements
            cnt += tot;
            carry = __shfl(w3 >> 24, GKC_WAVE - 1);
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

template __global__ void k_superkmers<uint32_t, false>(const ChunkDesc*, P1Params, BinSet);
template __global__ void k_superkmers<uint64_t, false>(const ChunkDesc*, P1Params, BinSet);
template __global__ void k_superkmers<uint64_t, true>(const ChunkDesc*, P1Params, BinSet);

}  // namespace gkc
