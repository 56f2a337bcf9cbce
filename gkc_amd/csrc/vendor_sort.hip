// The vendor library's radix sort (rocPRIM radix_sort_keys) over 64-bit keys,
// exported only so that measurement code can time it on the same keys and
// digit range as this library's onesweep passes: a baseline, not part of the
// counting path.
#include <cstring>
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>
#include <stdint.h>

#include "gkc_hip.h"

extern "C" int gkc_vendor_sort_u64(const void* keys_in, void* keys_out, uint64_t n, int begin_bit, int end_bit,
                                   void* stream, float* ms) {
    if (!keys_in || !keys_out || begin_bit < 0 || end_bit > 64 || begin_bit >= end_bit) return GKC_EINVAL;
    hipStream_t st = static_cast<hipStream_t>(stream);
    const uint64_t* in = static_cast<const uint64_t*>(keys_in);
    uint64_t* out = static_cast<uint64_t*>(keys_out);
    size_t tb = 0;
    if (rocprim::radix_sort_keys(nullptr, tb, in, out, (size_t)n, (unsigned)begin_bit, (unsigned)end_bit, st) !=
        hipSuccess)
        return GKC_EHIP;
    void* tmp = nullptr;
    if (hipMalloc(&tmp, tb ? tb : 16) != hipSuccess) return GKC_EHIP;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = 0;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = GKC_EHIP;
    if (!rc && hipEventRecord(e0, st) != hipSuccess) rc = GKC_EHIP;
    if (!rc && rocprim::radix_sort_keys(tmp, tb, in, out, (size_t)n, (unsigned)begin_bit, (unsigned)end_bit, st) !=
                   hipSuccess)
        rc = GKC_EHIP;
    if (!rc && hipEventRecord(e1, st) != hipSuccess) rc = GKC_EHIP;
    if (!rc && hipEventSynchronize(e1) != hipSuccess) rc = GKC_EHIP;
    if (!rc && ms && hipEventElapsedTime(ms, e0, e1) != hipSuccess) rc = GKC_EHIP;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(tmp);
    return rc;
}
