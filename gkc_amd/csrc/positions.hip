// Position-carrying counts (the reference's CountTask::SKMPartWithPos,
// gkc_cuda.hpp:50, declared for "count a portion of k-mers and store their
// positions", README.md:15-18, and never implemented): with GKC_WITH_POS the
// bins keep, beside each superkmer header, its read and in-read offset; this
// kernel then lists every occurrence of a partition's sampled k-mers (one in
// `mod` by a hash of the canonical k-mer) under its table entry.
#include "kernels.hpp"

namespace gkc {

// One thread per superkmer of the partition: roll its k-mers out of the
// 2-bit words (16 bases per word, first base highest), take each canonical
// form and strand (1 when the reverse complement is the smaller), and for a
// sampled k-mer find its table entry (binary search of the partition's
// sorted keys) and claim the next slot of that entry's range (its count
// long):
// out = read << 32 | (offset of the k-mer in its read) << 1 | strand.
__global__ __launch_bounds__(256) void k_pos_scatter(const SkmHdr* __restrict__ hdrs, const uint2* __restrict__ pos,
                                                     uint32_t n_hdrs, const uint32_t* __restrict__ words, int k,
                                                     uint32_t mod, const uint64_t* __restrict__ tkeys,
                                                     const uint32_t* __restrict__ tcnts, uint64_t n_tab,
                                                     const uint64_t* __restrict__ entry_off, uint32_t* __restrict__ cursor,
                                                     uint64_t* __restrict__ out) {
    const uint64_t mask = k == 32 ? ~0ull : ((1ull << (2 * k)) - 1);
    for (uint32_t h = blockIdx.x * blockDim.x + threadIdx.x; h < n_hdrs; h += gridDim.x * blockDim.x) {
        const SkmHdr hd = hdrs[h];
        const uint2 ps = pos[h];
        const uint32_t* w = words + hd.word_off;
        uint64_t x = 0;
        for (int j = 0; j < k - 1; j++) x = (x << 2) | ((w[j >> 4] >> (30 - 2 * (j & 15))) & 3u);
        for (uint32_t i = 0; i < hd.n_kmers; i++) {
            const uint32_t j = i + (uint32_t)k - 1;
            x = ((x << 2) | ((w[j >> 4] >> (30 - 2 * (j & 15))) & 3u)) & mask;
            const uint64_t r = revcomp64(x, k);
            const uint64_t cn = r < x ? r : x;
            if (mix64(cn) % mod != 0) continue;
            uint64_t lo = 0, hi = n_tab;                  // first key >= cn
            while (lo < hi) {
                const uint64_t mid = (lo + hi) >> 1;
                if (tkeys[mid] < cn) lo = mid + 1; else hi = mid;
            }
            if (lo >= n_tab || tkeys[lo] != cn || entry_off[lo] == ~0ull) continue;   // filtered out of the table
            const uint32_t slot = atomicAdd(&cursor[lo], 1u);
            if (slot >= tcnts[lo]) continue;              // never with a table counted from these bins
            out[entry_off[lo] + slot] = (uint64_t)ps.x << 32 | (uint64_t)(ps.y + i) << 1 | (r < x ? 1u : 0u);
        }
    }
}

}  // namespace gkc
