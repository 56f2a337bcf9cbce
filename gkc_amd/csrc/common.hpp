// Device-side helpers shared by the gfx950 kernels of the k-mer counting path.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define GKC_WAVE 64

namespace gkc {

// Chunk of reads resident in device memory.  Position 0 of `reads` is the
// first base of read 0; offs are absolute (offs[0] is subtracted on use),
// the same CSR the reference's PinnedCSR holds (gkc_cuda.hpp:33-46).
struct ChunkDesc {
    const uint8_t* reads;
    const uint64_t* offs;     // n_reads + 1
    const int32_t* lens;      // optional per-read valid length (HPC), else null
    const uint32_t* tile_r0;  // [n_tiles] read holding each tile's first position (0 in the lead)
    uint64_t base;            // offs[0] - lead
    uint32_t n_reads;
    uint32_t n_bases;         // offs[n] - offs[0]
    uint32_t n_tiles;         // ceil(n_bases / tile)
    uint32_t lead;            // positions [0, lead) precede read 0 (16-byte alignment of `reads`)
    uint32_t read0;           // index of read 0 among all reads partitioned since the last reset
};

// Global-memory view of a pointer that the compiler cannot place (one read
// from memory rather than passed as a kernel argument): its loads become
// global loads, which complete in order, instead of flat loads, which make
// every wait cover all outstanding memory operations.
#define GKC_GLOBAL __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ const GKC_GLOBAL T* as_global(const T* p) { return (const GKC_GLOBAL T*)p; }

// Superkmer header in a partition bin (16 bytes).
struct SkmHdr {
    uint32_t word_off;   // partition-local offset of its first 2-bit word
    uint32_t kmer_off;   // partition-local index of its first k-mer
    uint32_t n_kmers;    // number of k-mers (windows) it carries
    uint32_t reserved;   // 0
};

struct u128k {           // 128-bit key: hi holds the most significant bits
    uint64_t lo, hi;
};

__device__ __forceinline__ bool operator==(const u128k& a, const u128k& b) { return a.lo == b.lo && a.hi == b.hi; }
__device__ __forceinline__ bool operator!=(const u128k& a, const u128k& b) { return !(a == b); }

// ASCII -> 2-bit code for A/C/G/T (either case): ((c>>2) ^ (c>>1)) & 3 maps
// A,a->0 C,c->1 G,g->2 T,t->3, the values of d_basemap (gkc_cuda.cu:25-58).
__device__ __forceinline__ uint32_t base_code(uint32_t c) { return ((c >> 2) ^ (c >> 1)) & 3u; }

// new_filter2 (gkc_cuda.cu:214-216) or accept-all (mm_filter, :196-209).
__device__ __forceinline__ bool pmer_pass(uint64_t v, int p, int mode) {
    if (mode == 0) return true;
    uint64_t top = v >> (2 * (p - 3));
    return (top != 0) & (top != 4) & (top != 8);
}

// murmur3's 64-bit finaliser: the hashed partition mode and the k-mer
// sampling of position lists.
__host__ __device__ __forceinline__ uint64_t mix64(uint64_t x) {
    x ^= x >> 33; x *= 0xff51afd7ed558ccdULL;
    x ^= x >> 33; x *= 0xc4ceb9fe1a85ec53ULL;
    x ^= x >> 33;
    return x;
}

__device__ __forceinline__ uint32_t part_of(uint64_t mm, uint32_t P, int mode) {
    if (mode == 1) mm = mix64(mm);
    return (uint32_t)(mm % P);
}

// Reverse complement of a k-mer held in the low 2k bits (k <= 32).
// (bit-reversing the word reverses the order of the 2-bit codes and swaps
// the two bits of each; the pair swap restores them)
__device__ __forceinline__ uint64_t rev2_64(uint64_t y) {   // reverse the order of the 2-bit groups
    y = __builtin_bitreverse64(y);
    return ((y >> 1) & 0x5555555555555555ULL) | ((y & 0x5555555555555555ULL) << 1);
}
__device__ __forceinline__ uint64_t revcomp64(uint64_t x, int k) {
    return rev2_64(~x) >> (64 - 2 * k);
}
__device__ __forceinline__ uint64_t canonical64(uint64_t x, int k) {
    uint64_t r = revcomp64(x, k);
    return r < x ? r : x;
}
// 128-bit: value = hi:lo, k <= 64, low 2k bits used.
__device__ __forceinline__ u128k revcomp128(u128k x, int k) {
    // reverse the 128-bit word's 2-bit groups: new_hi = rev(lo), new_lo = rev(hi)
    uint64_t nhi = rev2_64(~x.lo), nlo = rev2_64(~x.hi);
    int s = 128 - 2 * k;        // shift right by s
    u128k r;
    if (s == 0) { r.lo = nlo; r.hi = nhi; }
    else if (s < 64) { r.lo = (nlo >> s) | (nhi << (64 - s)); r.hi = nhi >> s; }
    else { r.lo = nhi >> (s - 64); r.hi = 0; }
    return r;
}
__device__ __forceinline__ bool less128(const u128k& a, const u128k& b) {
    return a.hi < b.hi || (a.hi == b.hi && a.lo < b.lo);
}

__device__ __forceinline__ uint64_t lanemask_lt() {
    uint32_t lane = __lane_id();
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// Orders the LDS accesses of one wave's lanes around this point (LDS work
// of a wave runs in issue order, so only the compiler must not move accesses
// across it).  A workgroup-scope fence would also wait for the wave's
// outstanding global loads and stores.
__device__ __forceinline__ void wave_lds_fence() { __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront"); }

// Device-scope (all XCDs) relaxed atomics used by decoupled look-back.
__device__ __forceinline__ uint32_t ld_agent(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_agent(uint32_t* p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Look-back word: [31:30] status (0 none, 1 aggregate, 2 inclusive), [29:0] value.
constexpr uint32_t LB_AGG = 1u << 30;
constexpr uint32_t LB_INC = 2u << 30;
constexpr uint32_t LB_VAL = (1u << 30) - 1;

}  // namespace gkc
