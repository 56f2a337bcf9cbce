"""Shared helpers of the test suite (comparison of GPU and oracle tables)."""
from __future__ import annotations

import numpy as np


def assert_tables_equal(gpu_table, oracle_table, n_parts: int, k: int) -> None:
    """Partition-by-partition, entry-by-entry equality (keys ascending)."""
    lo, hi, cnt, po, pd, pk = gpu_table
    o = oracle_table
    np.testing.assert_array_equal(pk, o.part_kmers, err_msg="k-mers per partition")
    np.testing.assert_array_equal(pd, o.part_distinct, err_msg="distinct per partition")
    np.testing.assert_array_equal(po, o.part_off, err_msg="table offsets per partition")
    np.testing.assert_array_equal(lo, o.key_lo, err_msg="keys (low 64 bits)")
    if k > 32:
        np.testing.assert_array_equal(hi, o.key_hi, err_msg="keys (high 64 bits)")
    np.testing.assert_array_equal(cnt, o.cnt, err_msg="counts")


def device_to_host(ptr: int, nbytes: int, dtype=np.uint8) -> np.ndarray:
    """Copy nbytes of device memory at ptr to a host array (hipMemcpy of the
    HIP runtime the library already loaded)."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    out = np.empty(nbytes, np.uint8)
    if nbytes:
        rc = hip.hipMemcpy(ctypes.c_void_p(out.ctypes.data), ctypes.c_void_p(ptr), ctypes.c_size_t(nbytes), 2)
        assert rc == 0, f"hipMemcpy failed: {rc}"
    return out.view(dtype)
