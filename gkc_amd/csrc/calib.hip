// Streaming-copy kernel used as the practical HBM ceiling that the
// measurement code reports kernel bandwidths against (read + write of the
// same bytes, the traffic shape of one radix pass), and as the reference
// kernel of known traffic for HBM byte counters.  A second kernel copies with
// a radix pass's write pattern and none of its work, the ceiling of that
// pattern (gkc_scatter_copy below).  Each workgroup of the first copies one
// contiguous 64 KiB range with 16-byte accesses, four in flight per thread:
// on the 9.6 GB buffers of the 50 Mbp workload's sort passes this ran
// faster than grid-stride copies (plain or nontemporal, 4-32 workgroups per
// CU) and than the runtime's device-to-device copy.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gkc_hip.h"

namespace gkc {

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
constexpr int CP_NT = 256, CP_U = 4;
constexpr uint64_t CP_PER = 4096;           // 16-byte words per workgroup (64 KiB)

__global__ __launch_bounds__(CP_NT) void k_stream_copy(u32x4* __restrict__ dst, const u32x4* __restrict__ src,
                                                       uint64_t n) {
    const uint64_t a = (uint64_t)blockIdx.x * CP_PER, b = a + CP_PER < n ? a + CP_PER : n;
    for (uint64_t i = a + threadIdx.x; i < b; i += CP_NT * CP_U) {
        u32x4 v[CP_U];
#pragma unroll
        for (int u = 0; u < CP_U; u++)
            if (i + u * CP_NT < b) v[u] = src[i + u * CP_NT];
#pragma unroll
        for (int u = 0; u < CP_U; u++)
            if (i + u * CP_NT < b) dst[i + u * CP_NT] = v[u];
    }
}

// The write pattern of one onesweep pass with uniform digits and no ranking,
// histogram or look-back: 64-bit keys in segments of seg_keys keys (the
// partitions), tiles of SC_TILE keys read in order, each tile written as
// `runs` runs of SC_TILE / runs keys into the segment's `runs` bucket regions
// (run r of tile t at bucket r's offset t * SC_TILE / runs), staged through
// LDS like the pass.  Segments go to 8 lists (segment s to list s % 8) and
// workgroup b takes list b % 8's tiles in order, one die per list as in the
// pass's work lists.  No workgroup waits for another.
constexpr int SC_NT = 256, SC_TILE = 5120;
__global__ __launch_bounds__(SC_NT) void k_scatter_copy(uint64_t* __restrict__ dst, const uint64_t* __restrict__ src,
                                                        uint64_t n_segs, uint64_t seg_tiles, uint32_t runs) {
    __shared__ uint64_t s[SC_TILE];
    const uint32_t L = blockIdx.x % 8, per_list = gridDim.x / 8;
    const uint32_t R = SC_TILE / runs;
    const uint64_t segs_l = n_segs / 8 + (L < n_segs % 8 ? 1 : 0);     // segments of list L
    const uint64_t tiles_l = segs_l * seg_tiles;
    for (uint64_t i = blockIdx.x / 8; i < tiles_l; i += per_list) {
        const uint64_t seg = (i / seg_tiles) * 8 + L, t = i % seg_tiles;
        const uint64_t base = seg * seg_tiles * SC_TILE;
        uint64_t v[SC_TILE / SC_NT];
#pragma unroll
        for (int u = 0; u < SC_TILE / SC_NT; u++) v[u] = src[base + t * SC_TILE + u * SC_NT + threadIdx.x];
#pragma unroll
        for (int u = 0; u < SC_TILE / SC_NT; u++) s[u * SC_NT + threadIdx.x] = v[u];
        __syncthreads();
        for (uint32_t j = threadIdx.x; j < SC_TILE; j += SC_NT) {
            const uint32_t r = j / R, o = j - r * R;
            dst[base + (uint64_t)r * seg_tiles * R + t * R + o] = s[j];
        }
        __syncthreads();
    }
}

}  // namespace gkc

extern "C" int gkc_scatter_copy(void* dst, const void* src, uint64_t n_keys, uint64_t seg_keys, uint32_t runs,
                                uint32_t workgroups, void* stream) {
    using namespace gkc;
    if (!dst || !src || seg_keys == 0 || seg_keys % SC_TILE || n_keys % seg_keys || runs == 0 || SC_TILE % runs ||
        workgroups < 8 || workgroups % 8 || ((uintptr_t)dst & 7) || ((uintptr_t)src & 7))
        return GKC_EINVAL;
    if (n_keys == 0) return 0;
    hipLaunchKernelGGL(k_scatter_copy, dim3(workgroups), dim3(SC_NT), 0, static_cast<hipStream_t>(stream),
                       static_cast<uint64_t*>(dst), static_cast<const uint64_t*>(src), n_keys / seg_keys,
                       seg_keys / SC_TILE, runs);
    return hipGetLastError() == hipSuccess ? 0 : GKC_EHIP;
}

extern "C" int gkc_stream_copy(void* dst, const void* src, uint64_t bytes, void* stream) {
    if ((bytes & 15) || ((uintptr_t)dst & 15) || ((uintptr_t)src & 15)) return GKC_EINVAL;
    const uint64_t n = bytes / 16;
    if (n == 0) return 0;
    const uint64_t grid = (n + gkc::CP_PER - 1) / gkc::CP_PER;
    if (grid > 0x7FFFFFFFull) return GKC_EINVAL;
    hipLaunchKernelGGL(gkc::k_stream_copy, dim3((uint32_t)grid), dim3(gkc::CP_NT), 0, static_cast<hipStream_t>(stream),
                       static_cast<gkc::u32x4*>(dst), static_cast<const gkc::u32x4*>(src), n);
    return hipGetLastError() == hipSuccess ? 0 : GKC_EHIP;
}
