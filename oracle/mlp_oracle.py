"""CPU restatement of the learned closure's network and feature vector.

TEST INFRASTRUCTURE ONLY: the checker for the GPU network kernels (forward, backward) and the
closure feature kernel; nothing in the library calls it.

Network (include/watercube.h, "Learned closure"): 208 padded inputs -> 128 -> 128 -> 128 -> 3
(padded to 32), SiLU, parameters as one fp32 vector in the library's layout.  ``bf16=True``
rounds to bfloat16 (nearest even) exactly where the GPU path stores bf16 values: weights,
hidden activations and back-propagated deltas; products accumulate in float64 here (fp32 on the
GPU), so results agree to the accumulated rounding of fp32 sums.  ``bf16=False`` is the plain
float64 network, whose analytic gradient is pinned by finite differences.
"""
from __future__ import annotations

import numpy as np

IN_PAD, HID, OUT, OUT_PAD = 208, 128, 3, 32


def layout():
    sizes = [("w1", IN_PAD * HID), ("b1", HID), ("w2", HID * HID), ("b2", HID), ("w3", HID * HID), ("b3", HID),
             ("w4", HID * OUT_PAD), ("b4", OUT_PAD)]
    out, o = {}, 0
    for k, n in sizes:
        out[k] = (o, n)
        o += n
    out["total"] = o
    return out


def unpack(params: np.ndarray):
    L = layout()
    g = lambda k: params[L[k][0]:L[k][0] + L[k][1]].astype(np.float64)
    return (g("w1").reshape(IN_PAD, HID), g("b1"), g("w2").reshape(HID, HID), g("b2"), g("w3").reshape(HID, HID),
            g("b3"), g("w4").reshape(HID, OUT_PAD), g("b4"))


def pack(w1, b1, w2, b2, w3, b3, w4, b4):
    return np.concatenate([np.ravel(a) for a in (w1, b1, w2, b2, w3, b3, w4, b4)]).astype(np.float64)


def bf16_round(a):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    u = np.asarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def bf16_bits_to_f64(h: np.ndarray) -> np.ndarray:
    return (h.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward_backward(params: np.ndarray, X: np.ndarray, T: np.ndarray | None, bf16: bool = True):
    """X: (M, 208) float64 inputs (already bf16 values); T: (M, 3) targets or None.
    Returns (y, loss, grad) with grad in the flat layout (None without targets)."""
    r = bf16_round if bf16 else (lambda a: np.asarray(a, np.float64))
    w1, b1, w2, b2, w3, b3, w4, b4 = unpack(params)
    W = [r(w1), r(w2), r(w3)]
    W4 = r(w4)
    H, Z = [X], []
    for l in range(3):
        z = H[-1] @ W[l] + (b1, b2, b3)[l]
        Z.append(z)
        H.append(r(silu(z)))
    y = (H[3] @ W4 + b4)[:, :OUT]
    if T is None:
        return y, None, None
    M = len(X)
    d = y - T
    loss = float(np.sum(d * d) / (3.0 * M))
    D = np.zeros((M, OUT_PAD))
    D[:, :OUT] = r(d * (2.0 / (3.0 * M)))
    gw4 = H[3].T @ D
    gb4 = D.sum(0)
    grads_w, grads_b = [None] * 3, [None] * 3
    G = D @ W4.T
    for l in (2, 1, 0):
        D = r(G * dsilu(Z[l]))
        grads_w[l] = H[l].T @ D
        grads_b[l] = D.sum(0)
        if l:
            G = D @ W[l].T
    return y, loss, pack(grads_w[0], grads_b[0], grads_w[1], grads_b[1], grads_w[2], grads_b[2], gw4, gb4)
