// Types and kernel declarations shared by the translation units of the
// gfx950 k-mer counting library.
#pragma once
#include "common.hpp"

namespace gkc {

// ---- phase 1 (partition.hip) ----
constexpr int P1_NT = 256;              // threads per block (4 waves)
constexpr int P1_V = 8;                 // window starts per thread
constexpr int P1_T = P1_NT * P1_V;      // window starts per tile
constexpr int MAXK = 64;
constexpr uint32_t MAX_PARTS = 1024;

struct BinSet {
    uint32_t* words;            // 2-bit packed superkmers, 16 bases / word, MSB first
    SkmHdr* hdrs;
    const uint64_t* word_base;  // [P] start of each partition's word region
    const uint64_t* hdr_base;   // [P]
    const uint32_t* word_cap;   // [P]
    const uint32_t* hdr_cap;    // [P]
    uint32_t* word_cur;         // [P] cursors (persist across chunks)
    unsigned long long* hk_cur; // [P] headers (low 32 bits) and k-mers (high 32 bits) appended so far
    uint32_t* overflow;         // [1] set when any capacity is exceeded
    uint2* pos;                 // beside each header: {read index since the last reset, in-read
                                // offset of its first base}, or null (GKC_WITH_POS)
};

struct P1Params {
    int k, p, fmode, pmode;
    uint32_t P;
    int scatter;                // 0: count only, 1: write bins
    uint32_t* sketch;           // [P][256] sampled 4-mer counts, or null
    uint32_t* mm_out;           // per-window minimizer (compat), or null
    uint32_t* win_mm;           // per-window (minimizer, position, strand) output, or null
    uint8_t* win_pos;           //   offset in the window of the chosen p-mer (255: none passes)
    int8_t* win_strand;         //   0 forward, 1 reverse complement, -1 tie / none
};

template <typename MM, bool WIN>
__global__ void k_superkmers(const ChunkDesc* __restrict__ cdp, P1Params prm, BinSet bins);
__global__ void k_tile_reads(const uint64_t* __restrict__ offs, uint32_t n_reads, uint64_t base, uint32_t n_tiles,
                             uint32_t* __restrict__ tile_r0);
__global__ void k_hpc(const uint8_t* __restrict__ reads, const uint64_t* __restrict__ offs, uint32_t n_reads,
                      uint8_t* __restrict__ out, int32_t* __restrict__ lens, int32_t* __restrict__ orig_pos);
__global__ void k_skm_lists(const uint64_t* __restrict__ offs, const int32_t* __restrict__ lens, uint32_t n_reads, int k,
                            const uint32_t* __restrict__ mm, int32_t* __restrict__ skm);

// ---- phase 2 (sort.hip) ----
constexpr int RADIX_BITS = 8;
constexpr int RADIX = 1 << RADIX_BITS;
constexpr int MAX_PASS = 16;
constexpr int SORT_NT = 256;
constexpr int RL_ITEMS = 16;
constexpr int RL_TILE = SORT_NT * RL_ITEMS;
constexpr int EX_HDRS = 256;
constexpr int NLIST = 8;
constexpr int NCTR = NLIST * (MAX_PASS + 2);
__host__ __device__ constexpr int ctr_extract() { return 0; }
__host__ __device__ constexpr int ctr_pass(int p) { return NLIST * (1 + p); }
__host__ __device__ constexpr int ctr_rl() { return NLIST * (1 + MAX_PASS); }

// Device-resident plan of one counting launch (all sizes live here so the
// launch sequence can be captured once in a hipGraph and replayed).
struct Plan {
    uint32_t n_segs;
    uint32_t n_pass;
    uint32_t tot_items;         // extraction work items (EX_HDRS headers each)
    uint32_t tot_sort_tiles;
    uint32_t tot_rl_tiles;
    uint32_t sh0;               // bit position of pass 0's digit (0: full LSD sort)
    uint32_t dbits;             // digit width of the passes (8; 6 or 7 in the binned form)
    uint64_t tot_keys;
    uint32_t ctr[NCTR];         // work counters, NLIST per stage: see ctr_extract / ctr_pass / ctr_rl
};

// Work lists.  Segments are grouped into lists (one per XCD in the batched
// form, a single list in the plain form); the tiles of a list are claimed in
// order from the list's own counter, so consecutive tiles of a segment run on
// one XCD (the partial cache lines at digit-run boundaries meet in one L2, and
// look-back neighbours sit on the same die).  A workgroup starts on list
// blockIdx % n_lists and moves to the next list when its list is empty, so
// every list drains even if an XCD gets no workgroup.
// Packed layout of w (n = n_segs):
//   [0, 8)        sort tiles per list        [8, 16)   run-length tiles per list
//   [16, 24)      extraction items per list  [24, 33)  first order index of each list
//   [33, 33+n)    segment order              then per order index: sort tile offset in
//   its list [n], run-length tile offset [n], extraction item offset [n]
//   [33+4n, 33+4n+9)  first sort tile of each list in the lists' concatenation
struct WorkLists {
    const uint32_t* w;
    uint32_t n_lists, n_segs;
};

// A sort tile of the lists' concatenation (k_tile_desc): its segment's first
// key, its index in the segment, segment << 16 | keys in the tile.  Tile g's
// predecessor in its segment is tile g - 1, so g is also its look-back slot.
struct TileDesc {
    uint64_t kbase;
    uint32_t ti;
    uint32_t seg_tn;
};

struct SegArrays {
    const uint32_t* seg_keys;      // [S] keys per segment
    const uint32_t* seg_hdrs;      // [S] superkmer headers per segment
    uint64_t* key_start;           // [S+1]
    uint32_t* item_start;          // [S+1]
    uint32_t* sort_tile_start;     // [S+1]
    uint32_t* rl_tile_start;       // [S+1]
    uint32_t* hist;                // [S][MAX_PASS][RADIX]
};

struct TableOut {
    void* keys;                    // K[], entry i of segment s at table_start[s] + i
    uint32_t* counts;
    const uint64_t* table_start;   // [S]
    const uint32_t* table_cap;     // [S]
    uint32_t* n_entries;           // [S] entries kept (may exceed cap => overflow)
    uint32_t* n_distinct;          // [S] distinct keys before the count filter
    uint32_t* overflow;            // [1]
    uint32_t min_cnt, max_cnt;     // max 0: unbounded
    int emit;                      // 0: count only
};

template <typename K> struct KeyTraits;
template <> struct KeyTraits<uint64_t> {
    static constexpr int ITEMS = 16;               // keys per thread in a sort tile
    __device__ static inline uint32_t digit(uint64_t x, int sh) { return (uint32_t)(x >> sh) & (RADIX - 1); }
    // bits [sh, sh + nb), nb <= 8
    __device__ static inline uint32_t digit_n(uint64_t x, int sh, int nb) { return (uint32_t)(x >> sh) & ((1u << nb) - 1u); }
};
template <> struct KeyTraits<u128k> {
    static constexpr int ITEMS = 8;
    __device__ static inline uint32_t digit(const u128k& x, int sh) {   // bits [sh, sh+8), may straddle bit 64
        if (sh >= 64) return (uint32_t)(x.hi >> (sh - 64)) & (RADIX - 1);
        uint64_t v = x.lo >> sh;
        if (sh > 64 - RADIX_BITS) v |= x.hi << (64 - sh);
        return (uint32_t)v & (RADIX - 1);
    }
    __device__ static inline uint32_t digit_n(const u128k& x, int sh, int nb) {   // bits [sh, sh+nb), nb <= 8
        if (sh >= 64) return (uint32_t)(x.hi >> (sh - 64)) & ((1u << nb) - 1u);
        uint64_t v = x.lo >> sh;
        if (sh > 64 - nb) v |= x.hi << (64 - sh);
        return (uint32_t)v & ((1u << nb) - 1u);
    }
};
template <typename K> constexpr int sort_tile() { return SORT_NT * KeyTraits<K>::ITEMS; }

// Heavy windows of the binned count (ranges of more than WC_CAP keys): one
// k_bincount workgroup counts a range in LDS when it holds at most bc_cap
// keys (64-bit keys: hash count, then an LSD sort of the distinct keys;
// 128-bit keys: LSD sort, run-length count), else sorts it in bc_cap-key
// chunks through global memory and counts it chunk by chunk.
constexpr int BC_NT = 1024;
template <typename K> struct BinTraits;
template <> struct BinTraits<uint64_t> { static constexpr int ITEMS = 8; };   // 128 VGPRs, no spills
template <> struct BinTraits<u128k> { static constexpr int ITEMS = 4; };
template <typename K> constexpr int bc_cap() { return BC_NT * BinTraits<K>::ITEMS; }

__global__ void k_plan(Plan* plan, SegArrays sa, int sort_tile_keys, int rl_tile_keys);
__global__ void k_zero32(uint32_t* p, uint64_t n);
__global__ void k_set_u64(uint64_t* p, uint64_t v);
__global__ void k_set_u32(uint32_t* p, uint32_t v);
__global__ void k_plan_init(Plan* plan, uint32_t n_segs, uint32_t n_pass, uint32_t sh0, uint32_t dbits);
__global__ void k_set_chunk(ChunkDesc* dst, ChunkDesc v);
template <typename K>
__global__ void k_extract(Plan* __restrict__ plan, SegArrays sa, const uint32_t* __restrict__ iseg, const SkmHdr* __restrict__ hdrs,
                          const uint64_t* __restrict__ hdr_base, const uint32_t* __restrict__ words,
                          const uint64_t* __restrict__ word_base, K* __restrict__ keys, int k);
__global__ void k_hist_scan(Plan* __restrict__ plan, uint32_t* __restrict__ hist);
__global__ void k_tile_desc(SegArrays sa, WorkLists wl, int tile_keys, TileDesc* __restrict__ desc, uint32_t* __restrict__ iseg);
template <typename K, int ITEMS, int NT, int DB>
__global__ void k_onesweep(Plan* __restrict__ plan, SegArrays sa, WorkLists wl, const TileDesc* __restrict__ tdesc,
                           const K* __restrict__ in, K* __restrict__ out, uint32_t* __restrict__ lookback, int pass);
template <typename K>
__global__ void k_runlength(Plan* __restrict__ plan, SegArrays sa, WorkLists wl, const K* __restrict__ keys, TableOut to,
                            uint32_t* __restrict__ lookback);
// Wave-level count: after three onesweep passes over key bits
// [sh0, sh0 + 24) each wave counts WC_TILE-key windows of whole 24-bit bins
// on its own (LDS hash table, rank of the distinct keys by comparison) into
// staging, with no workgroup barriers and no wait on other windows; a scan of
// the windows' kept counts gives their table offsets and a copy writes the
// table.  Ranges of more than WC_CAP keys are counted beforehand by
// k_bincount into the same staging; 64-bit ranges of up to WC_TRY_MAX keys
// are tried in the wave's table (64 keys at a time while at most WC_FILL
// slots are taken, so a free slot always remains) and handed on when they
// hold more distinct keys: to k_midcount up to MID_MAX keys, else to a second
// k_bincount pass.
constexpr int WC_NT = 256;
constexpr int WC_TILE = 256;
constexpr int WC_CAP = 512;
constexpr uint32_t WC_FILL = 384;
// Windows of up to WC_CAP keys are first counted in the first WC_SMALL
// slots of the table (cleared and compacted in half the time), and again in
// all WC_CAP slots once more than WC_SMALL_FILL of them are taken.
constexpr uint32_t WC_SMALL = 256;
constexpr uint32_t WC_SMALL_FILL = 191;
static_assert(WC_SMALL_FILL + 64 < WC_SMALL, "a free slot remains after any 64 inserts");
constexpr uint32_t WC_TRY_MAX = 8192;
constexpr uint32_t WC_BATCH = 16;   // windows claimed per atomic
static_assert(WC_FILL + 64 < (uint32_t)WC_CAP, "a free slot remains after any 64 inserts");

// Heavy-range list of a binned count: tiles whose range exceeds the
// capacity of the counting kernel, counted ahead by k_bincount.
struct HeavyList {
    uint32_t* tiles;          // [*count] tile ids of the windows k_bincount counts
    uint32_t* count;          // appended by k_binranges
    uint32_t* next;           // claim counter of k_bincount
    uint2* res;               // [tile] {kept entries, distinct keys} of each heavy tile
    uint32_t* kept;           // [tile] kept entries of each window, then their offsets (k_tilescan)
    uint32_t* mid[2];         // [*mid_count[i]] tile ids of the windows k_midcount<., i> counts, or null
    uint32_t* mid_count[2];   // appended by k_wavecount
    uint32_t* mid_next[2];    // claim counters of k_midcount
    uint32_t* mid_seen;       // windows of WC_CAP < n <= WC_TRY_MAX keys k_wavecount took (statistics)
    uint32_t* late;           // [*late_count] windows of more than MID_MAX keys k_wavecount handed on, or null
    uint32_t* late_count;     // appended by k_wavecount
    uint32_t* late_next;      // claim counter of the second k_bincount pass
};
// Medium heavy windows (64-bit keys, WC_CAP < n <= MID_MAX) with more than
// WC_FILL distinct keys: one wave each with an LDS hash table at most 75 %
// full: list 0 up to mid_max(0) keys in mid_slots(0) slots, list 1 the larger
// ones in mid_slots(1).
__host__ __device__ constexpr uint32_t mid_slots(int list) { return list == 0 ? 1024u : 2048u; }
__host__ __device__ constexpr uint32_t mid_max(int list) { return mid_slots(list) / 4 * 3; }
constexpr uint32_t MID_MAX = mid_max(1);

template <typename K>
__global__ void k_binranges(Plan* __restrict__ plan, SegArrays sa, const K* __restrict__ keys, uint4* __restrict__ wdesc,
                            int tile_keys, int bsh, int bin_bits, uint32_t heavy_cap, HeavyList hl);
template <typename K>
__global__ void k_bincount(Plan* __restrict__ plan, SegArrays sa, K* __restrict__ keys, K* __restrict__ alt,
                           const uint4* __restrict__ wdesc, TableOut to, HeavyList hl, int late);
template <int L>
__global__ void k_midcount(Plan* __restrict__ plan, SegArrays sa, uint64_t* __restrict__ keys, uint64_t* __restrict__ alt,
                           const uint4* __restrict__ wdesc, TableOut to, HeavyList hl);
template <typename K>
__global__ void k_wavecount(Plan* __restrict__ plan, SegArrays sa, K* __restrict__ keys,
                            K* __restrict__ alt, const uint4* __restrict__ wdesc, HeavyList hl, TableOut to);
__global__ void k_tilescan(Plan* __restrict__ plan, SegArrays sa, uint32_t* __restrict__ kept);
template <typename K>
__global__ void k_wavecopy(Plan* __restrict__ plan, SegArrays sa, const K* __restrict__ keys,
                           const K* __restrict__ alt, const uint4* __restrict__ wdesc, HeavyList hl, TableOut to);

// ---- positions (positions.hip) ----
__global__ void k_pos_scatter(const SkmHdr* __restrict__ hdrs, const uint2* __restrict__ pos, uint32_t n_hdrs,
                              const uint32_t* __restrict__ words, int k, uint32_t mod, const uint64_t* __restrict__ tkeys,
                              const uint32_t* __restrict__ tcnts, uint64_t n_tab, const uint64_t* __restrict__ entry_off, uint32_t* __restrict__ cursor,
                              uint64_t* __restrict__ out);

}  // namespace gkc
