"""CPU restatement of the learned-warm-start network (test infrastructure).

Only tests, smoke() and the benchmark's checks may import this module.  The
network has no counterpart in the reference (its GMRES starts from zeros,
ErosionPackage/src/DensityStress.jl:45); this file restates what
erosion_amd/csrc/mlp.hip computes so that the GPU forward pass and gradients
can be checked:

  h_0 = x, h_{l+1} = tanh(W_l h_l + b_l) (l < L), y = W_L h_L + b_L,
  loss = sum (y - t)^2 / (B nout),

with every matrix product taken on BF16-rounded operands (round to nearest
even) and exact products summed in float64 (the GPU sums them in FP32).
``bf16=False`` gives the plain float64 network, whose backward pass the
tests pin against central finite differences.
"""
from __future__ import annotations

from typing import List, Tuple

import numpy as np


def bf16_round(a: np.ndarray) -> np.ndarray:
    """Round float32 values to the nearest BF16 (ties to even), as float64."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def layout(nin: int, nhid: int, nlayers: int, nout: int) -> List[Tuple[int, int, int, int]]:
    """(nin, nout, offset of W, offset of b) per layer of the flat parameter array."""
    out, off, prev = [], 0, nin
    for l in range(nlayers + 1):
        o = nhid if l < nlayers else nout
        out.append((prev, o, off, off + o * prev))
        off += o * prev + o
        prev = o
    return out


def param_count(nin: int, nhid: int, nlayers: int, nout: int) -> int:
    L = layout(nin, nhid, nlayers, nout)
    return L[-1][3] + L[-1][1]


def _mm(a, b, bf16):
    if bf16:
        return bf16_round(a) @ bf16_round(b)
    return np.asarray(a, np.float64) @ np.asarray(b, np.float64)


def forward(shape, params, x, bf16=True):
    """Activations [h_0, ..., h_L, y] (float64 arrays)."""
    nin, nhid, nlayers, nout = shape
    p = np.asarray(params, dtype=np.float64)
    hs = [np.asarray(x, dtype=np.float64)]
    for l, (ni, no, ow, ob) in enumerate(layout(*shape)):
        W = p[ow:ow + no * ni].reshape(no, ni)
        z = _mm(hs[-1], W.T, bf16) + p[ob:ob + no]
        hs.append(np.tanh(z) if l < nlayers else z)
    return hs


def loss_and_gradients(shape, params, x, t, bf16=True):
    """(loss, flat gradient) of the mean squared error."""
    nin, nhid, nlayers, nout = shape
    p = np.asarray(params, dtype=np.float64)
    hs = forward(shape, p, x, bf16)
    y = hs[-1]
    d = y - np.asarray(t, dtype=np.float64)
    n = d.size
    loss = float(np.sum(d * d) / n)
    g = np.zeros_like(p)
    gcur = 2.0 * d / n
    L = layout(*shape)
    for l in range(nlayers, -1, -1):
        ni, no, ow, ob = L[l]
        g[ow:ow + no * ni] = _mm(gcur.T, hs[l], bf16).ravel()
        g[ob:ob + no] = gcur.sum(axis=0)
        if l == 0:
            break
        W = p[ow:ow + no * ni].reshape(no, ni)
        gcur = _mm(gcur, W, bf16) * (1.0 - hs[l] ** 2)
    return loss, g
