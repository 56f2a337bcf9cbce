"""CPU oracle of the pre-sizing MLP (TEST INFRASTRUCTURE ONLY).

float64 numpy statement of the network 256 -> 128 -> 128 -> 2 with ReLU and
the mean-squared-error loss, with its analytic gradient; the tests pin this
gradient by central finite differences of the float64 loss and then compare
the HIP kernels' FP32 gradient against it.  The network has no counterpart in
the reference (its counting path sizes nothing ahead, kmer_counting.cu:95-191).
"""
from __future__ import annotations

import numpy as np

NI, NH, NO = 256, 128, 2


def split(w):
    o, out = 0, []
    for shape in [(NI, NH), (NH,), (NH, NH), (NH,), (NH, NO), (NO,)]:
        n = int(np.prod(shape))
        out.append(np.asarray(w[o:o + n], np.float64).reshape(shape))
        o += n
    return out


def forward(w, x):
    W1, b1, W2, b2, W3, b3 = split(w)
    x = np.asarray(x, np.float64)
    z1 = x @ W1 + b1
    a1 = np.maximum(z1, 0)
    z2 = a1 @ W2 + b2
    a2 = np.maximum(z2, 0)
    y = a2 @ W3 + b3
    return y, (x, z1, a1, z2, a2)


def loss(w, x, t):
    y, _ = forward(w, x)
    return float(np.mean((y - np.asarray(t, np.float64)) ** 2))


def loss_grad(w, x, t):
    W1, b1, W2, b2, W3, b3 = split(w)
    y, (x, z1, a1, z2, a2) = forward(w, x)
    d = y - np.asarray(t, np.float64)
    L = float(np.mean(d ** 2))
    dy = 2.0 * d / d.size
    gW3 = a2.T @ dy
    gb3 = dy.sum(0)
    d2 = (dy @ W3.T) * (z2 > 0)
    gW2 = a1.T @ d2
    gb2 = d2.sum(0)
    d1 = (d2 @ W2.T) * (z1 > 0)
    gW1 = x.T @ d1
    gb1 = d1.sum(0)
    return L, np.concatenate([g.ravel() for g in (gW1, gb1, gW2, gb2, gW3, gb3)])
