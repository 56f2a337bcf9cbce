"""Learned partition pre-sizing.

A small MLP (256 -> 128 -> 128 -> 2, ReLU) reads the per-partition 4-mer
sketch the partition kernel takes during its scatter (log1p of the counts of
the leading 4-mers of one superkmer in eight, sampled by a hash of its first
bases) and predicts the partition's log total k-mer count and its log
distinct/total ratio.  The exact total is known
once the superkmers are scattered (the bins hold it), so the distinct count
is the exact total times the predicted ratio; those counts size the
(k-mer, count) table of the batched counter before counting, so the count
writes the table directly (no separate count pass, one host sync).  The
binned count lays the table out densely from the exact counts it measures on
the device, so only the sum of the predictions has to cover the table; a
shortfall is caught on the device and falls back to count-then-allocate (see
gkc_count(GKC_BATCHED | GKC_PRESIZED)).

Training, forward, backward and inference all run in the project's own HIP
kernels (gkc_amd/csrc/mlp.hip): FP32 master weights trained with Adam, BF16
matrix-core inference with FP32 accumulation.
"""
from __future__ import annotations

import ctypes as C
import json
import os

import numpy as np

from . import capi

MODEL_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "models")
DEFAULT_MODEL = os.path.join(MODEL_DIR, "presize_mlp.npz")
NI, NH, NO = 256, 128, 2

_f32p = C.POINTER(C.c_float)


def _p(a):
    return a.ctypes.data_as(_f32p)


class MLP:
    """The HIP MLP (FP32 master weights on the device)."""

    def __init__(self, seed: int = 0, device: int = 0):
        L = capi.lib()
        self.h = L.gkc_mlp_create(device, seed)
        if not self.h:
            raise capi.GkcError(L.gkc_mlp_last_error().decode())
        self._destroy = L.gkc_mlp_destroy      # held here: module globals may be gone when __del__ runs at exit
        self.n_params = int(L.gkc_mlp_n_params())

    def _chk(self, rc):
        if rc < 0:
            raise capi.GkcError(capi.lib().gkc_mlp_last_error().decode())

    def close(self):
        if getattr(self, "h", None):
            self._destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def get_params(self) -> np.ndarray:
        out = np.zeros(self.n_params, np.float32)
        self._chk(capi.lib().gkc_mlp_get_params(self.h, _p(out)))
        return out

    def set_params(self, w: np.ndarray) -> None:
        w = np.ascontiguousarray(w, np.float32)
        assert w.size == self.n_params
        self._chk(capi.lib().gkc_mlp_set_params(self.h, _p(w)))

    def forward(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32)
        y = np.zeros((len(x), NO), np.float32)
        self._chk(capi.lib().gkc_mlp_forward(self.h, _p(x), len(x), _p(y)))
        return y

    def forward_bf16(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32)
        y = np.zeros((len(x), NO), np.float32)
        self._chk(capi.lib().gkc_mlp_forward_bf16(self.h, _p(x), len(x), _p(y)))
        return y

    def loss_grad(self, x: np.ndarray, t: np.ndarray):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        loss = np.zeros(1, np.float32)
        g = np.zeros(self.n_params, np.float32)
        self._chk(capi.lib().gkc_mlp_loss_grad(self.h, _p(x), _p(t), len(x), _p(loss), _p(g)))
        return float(loss[0]), g

    def train(self, x, t, epochs=200, batch=256, lr=1e-3, weight_decay=0.0, seed=0) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        losses = np.zeros(max(epochs, 1), np.float32)
        self._chk(capi.lib().gkc_mlp_train(self.h, _p(x), _p(t), len(x), epochs, batch, lr, weight_decay, seed,
                                           _p(losses)))
        return losses[:epochs]


def split_params(w: np.ndarray):
    """Flat parameters -> (W1, b1, W2, b2, W3, b3) in the library's layout."""
    o = 0
    out = []
    for shape in [(NI, NH), (NH,), (NH, NH), (NH,), (NH, NO), (NO,)]:
        n = int(np.prod(shape))
        out.append(w[o:o + n].reshape(shape))
        o += n
    return out


class Predictor:
    """MLP + normalisation + safety margin -> per-partition table capacities."""

    def __init__(self, params, x_mean, x_std, t_mean, t_std, margin: float, info: dict | None = None):
        self.mlp = MLP()
        self.mlp.set_params(params)
        self.x_mean, self.x_std = np.asarray(x_mean, np.float32), np.asarray(x_std, np.float32)
        self.t_mean, self.t_std = np.asarray(t_mean, np.float32), np.asarray(t_std, np.float32)
        self.margin = float(margin)
        self.info = info or {}

    @staticmethod
    def load(path: str) -> "Predictor":
        with np.load(path, allow_pickle=False) as z:
            info = json.loads(str(z["info"])) if "info" in z.files else {}
            return Predictor(z["params"], z["x_mean"], z["x_std"], z["t_mean"], z["t_std"], float(z["margin"]), info)

    @staticmethod
    def load_default(n_parts: int | None = None):
        return Predictor.load(DEFAULT_MODEL) if os.path.exists(DEFAULT_MODEL) else None

    def save(self, path: str) -> None:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        np.savez(path, params=self.mlp.get_params(), x_mean=self.x_mean, x_std=self.x_std, t_mean=self.t_mean,
                 t_std=self.t_std, margin=np.float32(self.margin), info=json.dumps(self.info))

    def normalise(self, sketch: np.ndarray) -> np.ndarray:
        return ((sketch - self.x_mean) / self.x_std).astype(np.float32)

    def predict(self, sketch: np.ndarray, total: np.ndarray | None = None,
                bf16: bool = True) -> tuple[np.ndarray, np.ndarray]:
        """-> (predicted total, predicted distinct) k-mers per partition.  The
        distinct count is the predicted ratio times `total` (the exact totals)
        when given, else times the predicted total."""
        x = self.normalise(sketch)
        y = self.mlp.forward_bf16(x) if bf16 else self.mlp.forward(x)
        logs = y * self.t_std + self.t_mean
        tot = np.expm1(logs[:, 0]).clip(0)
        base = tot if total is None else np.asarray(total, np.float64)
        return tot, base * np.exp(np.minimum(logs[:, 1], 0.0))

    def attach(self, counter) -> None:
        """Run this predictor on the device inside counter's counts:
        count(BATCHED | PRESIZED | LEARNED) then needs no host round trip
        for its capacities (see gkc_set_predictor)."""
        counter.set_predictor(self.mlp.h, self.x_mean, self.x_std, self.t_mean, self.t_std, self.margin, keep=self)

    def caps(self, counter) -> np.ndarray:
        """Table capacities for counter's current bins: the exact k-mer total
        times the predicted ratio times the margin, never above the total."""
        total = counter.part_kmers().astype(np.float64)
        _, dist = self.predict(counter.sketch(), total)
        caps = np.minimum(np.ceil(dist * self.margin) + 16, total)
        return caps.astype(np.uint32)


def mape(pred: np.ndarray, true: np.ndarray) -> float:
    m = true > 0
    return float(np.mean(np.abs(pred[m] - true[m]) / true[m]) * 100.0)


def collect_pairs(cids, genome_indices, scale=1.0, device_flags=0, log=print):
    """(sketch, [log1p total, log(distinct / total)]) pairs per non-empty
    partition, from the plain counter on seeded genomes."""
    from . import workloads as W
    xs, ts = [], []
    for cid in cids:
        for g in genome_indices:
            w = W.workload(cid, scale=scale)
            rs = W.generate(w, g=g)
            with capi.Counter(w.k, w.p, w.n_parts, hpc=w.hpc, filter_mode=w.filter_mode, part_mode=w.part_mode,
                              min_count=1) as c:
                for a, b in W.chunk_bounds(rs, seed=0, lo_bytes=1 << 30, hi_bytes=1 << 30, fasta_overhead=False):
                    c.partition(rs.reads[int(rs.offs[a]):int(rs.offs[b])], rs.offs[a:b + 1])
                x = c.sketch()
                c.count(device_flags)
                po, pd, pk = c.table_info()
            keep = pk > 0
            xs.append(x[keep])
            tk, td = pk[keep].astype(np.float64), pd[keep].astype(np.float64)
            ts.append(np.stack([np.log1p(tk), np.log(td / tk)], 1))
            log(f"pairs C{cid} g={g}: {int(keep.sum())} partitions, {int(pk.sum())} k-mers")
    return np.concatenate(xs).astype(np.float32), np.concatenate(ts)
