"""Learned partition pre-sizing: what the trained network gains over the
method without it, with the table still bit-exact.

Without it the batched counter sizes each partition's table by a counting
pass of the run-length kernel followed by a host sync (count-then-allocate);
the only other way to skip that pass is to size every table for its worst
case, the partition's k-mer total.  With it the table is sized from the
predicted distinct count: one host sync, no counting pass, and a table a
fraction of the worst-case size; under-predictions are caught on the device.
"""
from __future__ import annotations

import os

import numpy as np
import pytest

from gkc_amd import capi, learned
from gkc_amd import workloads as W
from oracle import oracle as O
from tests.helpers import assert_tables_equal

MODEL = learned.DEFAULT_MODEL


def test_model_file_is_plain_data():
    if not os.path.exists(MODEL):
        pytest.skip("no trained model committed")
    with np.load(MODEL, allow_pickle=False) as z:
        assert z["params"].shape == (256 * 128 + 128 + 128 * 128 + 128 + 128 * 2 + 2,)
        assert z["params"].dtype == np.float32
        assert float(z["margin"]) >= 1.0


@pytest.mark.gpu
@pytest.mark.parametrize("cid,scale,g", [(2, 0.01, 101), (1, 0.2, 102), (5, 0.004, 103)])
def test_learned_presizing_gain(gpu, cid, scale, g):
    pred = learned.Predictor.load_default()
    if pred is None:
        pytest.skip("no trained model committed")
    w = W.workload(cid, scale=scale)
    rs = W.generate(w, g=g)          # held-out genome seed
    ot = O.count(rs.reads, rs.offs, w.k, w.p, w.n_parts)
    print("[gpu-case] learned", w.name, flush=True)
    with capi.Counter(w.k, w.p, w.n_parts) as c:
        c.partition(rs.reads, rs.offs, flags=capi.BATCHED)
        c.count(capi.BATCHED)                      # count-then-allocate
        base = c.stats()
        assert_tables_equal(c.table(), ot, w.n_parts, w.k)
        caps = pred.caps(c)
        c.set_table_caps(caps)
        c.count(capi.BATCHED | capi.PRESIZED)      # learned presizing
        got = c.stats()
        assert_tables_equal(c.table(), ot, w.n_parts, w.k)
    exact = np.diff(ot.part_off).astype(np.float64)
    total = ot.part_kmers.astype(np.float64)
    # fewer host round trips and no counting pass unless the prediction fell short
    assert base["host_syncs"] == 2
    assert got["host_syncs"] == 1 + got["fallbacks"]
    # a table far below the worst-case bound, and not far above the exact size
    assert caps.sum() < 0.6 * total.sum()
    assert caps.sum() >= exact.sum() * 0.9
    under = int(np.sum(caps < exact))
    print(f"[learned] caps/exact {caps.sum() / exact.sum():.3f} caps/bound {caps.sum() / total.sum():.3f} "
          f"under-predicted {under}/{len(exact)} fallbacks {got['fallbacks']}", flush=True)


@pytest.mark.gpu
@pytest.mark.parametrize("cid,scale,mode", [(2, 0.01, 0), (5, 0.004, 0), (1, 0.2, 1)])
def test_learned_presizing_on_device(gpu, cid, scale, mode):
    """The predictor attached to the counter runs on the device inside the
    count (BF16 MFMA inference from the device sketch): its capacities match
    the host-side predictor's, the table is bit-exact, and the binned count
    needs only its final host sync (the full-LSD count also reads the
    capacities back)."""
    pred = learned.Predictor.load_default()
    if pred is None:
        pytest.skip("no trained model committed")
    w = W.workload(cid, scale=scale)
    rs = W.generate(w, g=104)
    ot = O.count(rs.reads, rs.offs, w.k, w.p, w.n_parts)
    print("[gpu-case] learned on device", w.name, mode, flush=True)
    with capi.Counter(w.k, w.p, w.n_parts) as c:
        c.set_tuning(count_mode=mode)
        c.partition(rs.reads, rs.offs, flags=capi.BATCHED)
        host_caps = pred.caps(c).astype(np.int64)
        pred.attach(c)
        for rep in range(2):                       # graph replays
            c.count(capi.BATCHED | capi.PRESIZED | capi.LEARNED)
            st = c.stats()
            assert_tables_equal(c.table(), ot, w.n_parts, w.k)
        dev_caps = c.learned_caps().astype(np.int64)
    # the same arithmetic on the device (log1pf / expf, f32) as on the host (numpy)
    assert np.all(np.abs(dev_caps - host_caps) <= host_caps // 200 + 2)
    assert abs(int(dev_caps.sum()) - int(host_caps.sum())) <= int(host_caps.sum()) // 1000 + 16
    assert st["fallbacks"] == 0
    assert st["host_syncs"] == (1 if mode == 0 else 2)
