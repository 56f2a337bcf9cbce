// Streaming-copy kernel used as the practical HBM ceiling that the
// measurement code reports kernel bandwidths against (read + write of the
// same bytes, the traffic shape of one radix pass), and as the reference
// kernel of known traffic for HBM byte counters.  Each workgroup copies one
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

}  // namespace gkc

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
