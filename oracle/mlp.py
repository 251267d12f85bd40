"""CPU oracle of the learned-binning network -- TEST INFRASTRUCTURE ONLY.

The 5 -> 64 -> 64 -> 8 level predictor has no counterpart in the reference
(the reference runs every particle's level search from level 0,
doAdaptiveFullLeap, nsphere.c:1648-1735); this module states the network's
mathematics in numpy so that the GPU kernels can be checked against it:

* features(): the five raw inputs (log r, v_r, log |l|, log M(<r), log dt) of
  a state in rank order, with the reference's rank-based enclosed mass
  ((i / npts) M, nsphere.c:4706-4717);
* forward() / gradient(): ReLU MLP, class-weighted mean cross entropy and its
  backward pass.  With bf16=True every matrix-product operand is rounded to
  bfloat16 (round to nearest even) at exactly the points where the matrix
  cores consume it -- weights, normalised inputs, activations, upstream
  gradients -- while biases, ReLU masks and softmax stay in float32; sums are
  taken in float64 (the GPU accumulates in float32, so GPU and oracle differ
  by accumulation order only);
* train_labels(): the (features, level) pairs of one step as the integrator
  logged them.

Parity pins (tests/test_oracle_mlp.py): the exact-arithmetic gradient against
central finite differences of the loss in float64 (every weight block and
bias); the bf16 gradient against the exact one within bfloat16's resolution;
the bf16 rounding against its bit-level definition.
"""
from __future__ import annotations

import numpy as np

N_IN, N_IN_PAD, N_HID, N_OUT, N_OUT_PAD = 5, 16, 64, 8, 16
W1 = 0
B1 = W1 + N_HID * N_IN_PAD
W2 = B1 + N_HID
B2 = W2 + N_HID * N_HID
W3 = B2 + N_HID
B3 = W3 + N_OUT_PAD * N_HID
N_PARAMS = B3 + N_OUT_PAD  # 6288

LVL_MASK, LVL_NOCONV, LVL_LC = 0x3F, 0x40, 0x80


def bf16(x) -> np.ndarray:
    """float32 -> nearest bfloat16 (ties to even), returned as float32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return (r & 0xFFFFFFFF).astype(np.uint32).view(np.float32)


def unpack(P):
    """Weight blocks (float32) of the flat parameter vector."""
    P = np.asarray(P, dtype=np.float32)
    w1 = P[W1:B1].reshape(N_HID, N_IN_PAD)[:, :N_IN]
    b1 = P[B1:W2]
    w2 = P[W2:B2].reshape(N_HID, N_HID)
    b2 = P[B2:W3]
    w3 = P[W3:B3].reshape(N_OUT_PAD, N_HID)[:N_OUT]
    b3 = P[B3:B3 + N_OUT]
    return w1, b1, w2, b2, w3, b3


def pack(w1, b1, w2, b2, w3, b3) -> np.ndarray:
    P = np.zeros(N_PARAMS, np.float32)
    P[W1:B1].reshape(N_HID, N_IN_PAD)[:, :N_IN] = w1
    P[B1:W2] = b1
    P[W2:B2] = np.asarray(w2, np.float32).ravel()
    P[B2:W3] = b2
    P[W3:B3].reshape(N_OUT_PAD, N_HID)[:N_OUT] = w3
    P[B3:B3 + N_OUT] = b3
    return P


def _q(x, on: bool):
    return bf16(x).astype(np.float64) if on else np.asarray(x, np.float64)


def normalise(X, norm) -> np.ndarray:
    """(x - mean) * (1 / std) per feature, in float32 as the kernels do."""
    norm = np.asarray(norm, np.float32)
    return (np.asarray(X, np.float32) - norm[:N_IN]) * norm[N_IN:2 * N_IN]


def forward(P, norm, X, bf16_ops: bool = True):
    """Activations (h1, h2) and logits z, samples along axis 0.

    bf16_ops=False gives the same network in exact float64 arithmetic."""
    w1, b1, w2, b2, w3, b3 = unpack(P)
    f = np.float32 if bf16_ops else np.float64
    x = normalise(X, norm) if bf16_ops else (np.asarray(X, np.float64) - np.float64(norm[:N_IN])) * \
        np.asarray(norm[N_IN:2 * N_IN], np.float64)
    a1 = (_q(x, bf16_ops) @ _q(w1, bf16_ops).T).astype(f) + b1.astype(f)
    h1 = np.maximum(a1, 0)
    a2 = (_q(h1, bf16_ops) @ _q(w2, bf16_ops).T).astype(f) + b2.astype(f)
    h2 = np.maximum(a2, 0)
    z = (_q(h2, bf16_ops) @ _q(w3, bf16_ops).T).astype(f) + b3.astype(f)
    return x, h1, h2, z


def predict(P, norm, X) -> np.ndarray:
    return np.argmax(forward(P, norm, X)[3], axis=1).astype(np.uint8)


def loss(P, norm, cw, X, y, bf16_ops: bool = False) -> float:
    """Class-weighted mean cross entropy: sum_s w_ys (-log p_ys) / sum_s w_ys."""
    z = forward(P, norm, X, bf16_ops)[3].astype(np.float64)
    z = z - z.max(axis=1, keepdims=True)
    logp = z - np.log(np.exp(z).sum(axis=1, keepdims=True))
    w = np.asarray(cw, np.float64)[y]
    return float(-(w * logp[np.arange(len(y)), y]).sum() / w.sum())


def gradient(P, norm, cw, X, y, bf16_ops: bool = True):
    """d loss / d P (flat, padding zero) and the loss."""
    w1, b1, w2, b2, w3, b3 = unpack(P)
    y = np.asarray(y, np.int64)
    x, h1, h2, z = forward(P, norm, X, bf16_ops)
    zz = z.astype(np.float64)
    zz = zz - zz.max(axis=1, keepdims=True)
    e = np.exp(zz)
    p = e / e.sum(axis=1, keepdims=True)
    w = np.asarray(cw, np.float64)[y]
    scale = 1.0 / w.sum()
    onehot = np.zeros_like(p)
    onehot[np.arange(len(y)), y] = 1.0
    dz = (w * scale)[:, None] * (p - onehot)
    if bf16_ops:
        dz = dz.astype(np.float32)
    q = lambda a: _q(a, bf16_ops)  # noqa: E731
    g_w3 = q(dz).T @ q(h2)
    g_b3 = dz.sum(axis=0)
    dh2 = (q(dz) @ q(w3)) * (h2 > 0)
    if bf16_ops:
        dh2 = dh2.astype(np.float32)
    g_w2 = q(dh2).T @ q(h1)
    g_b2 = dh2.sum(axis=0)
    dh1 = (q(dh2) @ q(w2)) * (h1 > 0)
    if bf16_ops:
        dh1 = dh1.astype(np.float32)
    g_w1 = q(dh1).T @ q(x)
    g_b1 = dh1.sum(axis=0)
    lossv = float(-(w * np.log(np.maximum(p[np.arange(len(y)), y], 1e-30))).sum() * scale)
    return pack(g_w1, g_b1, g_w2, g_b2, g_w3, g_b3), lossv


def gradient_f64(P, norm, cw, X, y) -> np.ndarray:
    """The exact-arithmetic gradient in float64 (no float32 repack)."""
    w1, b1, w2, b2, w3, b3 = (a.astype(np.float64) for a in unpack(P))
    y = np.asarray(y, np.int64)
    x, h1, h2, z = forward(P, norm, X, False)
    zz = z - z.max(axis=1, keepdims=True)
    p = np.exp(zz)
    p /= p.sum(axis=1, keepdims=True)
    w = np.asarray(cw, np.float64)[y]
    onehot = np.zeros_like(p)
    onehot[np.arange(len(y)), y] = 1.0
    dz = (w / w.sum())[:, None] * (p - onehot)
    dh2 = (dz @ w3) * (h2 > 0)
    dh1 = (dh2 @ w2) * (h1 > 0)
    g = np.zeros(N_PARAMS, np.float64)
    g[W1:B1].reshape(N_HID, N_IN_PAD)[:, :N_IN] = dh1.T @ x
    g[B1:W2] = dh1.sum(axis=0)
    g[W2:B2] = (dh2.T @ h1).ravel()
    g[B2:W3] = dh2.sum(axis=0)
    g[W3:B3].reshape(N_OUT_PAD, N_HID)[:N_OUT] = dz.T @ h2
    g[B3:B3 + N_OUT] = dz.sum(axis=0)
    return g


def loss_f64_of(Pvec64, norm, cw, X, y) -> float:
    """Loss of a float64 parameter vector in exact arithmetic (finite differences)."""
    P = np.asarray(Pvec64, np.float64)
    w1 = P[W1:B1].reshape(N_HID, N_IN_PAD)[:, :N_IN]
    b1 = P[B1:W2]
    w2 = P[W2:B2].reshape(N_HID, N_HID)
    b2 = P[B2:W3]
    w3 = P[W3:B3].reshape(N_OUT_PAD, N_HID)[:N_OUT]
    b3 = P[B3:B3 + N_OUT]
    x = (np.asarray(X, np.float64) - np.float64(norm[:N_IN])) * np.asarray(norm[N_IN:2 * N_IN], np.float64)
    h1 = np.maximum(x @ w1.T + b1, 0)
    h2 = np.maximum(h1 @ w2.T + b2, 0)
    z = h2 @ w3.T + b3
    z = z - z.max(axis=1, keepdims=True)
    logp = z - np.log(np.exp(z).sum(axis=1, keepdims=True))
    w = np.asarray(cw, np.float64)[np.asarray(y, np.int64)]
    return float(-(w * logp[np.arange(len(y)), y]).sum() / w.sum())


def features(r, v, ell, dt, halo_mass=1.0) -> np.ndarray:
    """Raw network inputs of a state in rank order: the natural logarithm of
    each value rounded to float32 (taken in float64 here and rounded to
    float32; the kernels use a single-precision log, within 2 ulp of this)."""
    n = len(r)
    menc = (np.arange(n, dtype=np.float64) / np.float64(n)) * np.float64(halo_mass)
    f32 = lambda x: np.asarray(x, np.float64).astype(np.float32).astype(np.float64)  # noqa: E731
    tiny = np.float64(np.float32(1e-30))
    out = np.empty((n, N_IN), np.float32)
    out[:, 0] = np.log(np.maximum(f32(r), tiny))
    out[:, 1] = np.asarray(v, np.float64)
    out[:, 2] = np.log(np.maximum(np.abs(f32(ell)), tiny))
    out[:, 3] = np.log(np.maximum(f32(menc), tiny))
    out[:, 4] = np.log(f32(dt))
    return out


def train_labels(levels) -> tuple[np.ndarray, np.ndarray]:
    """(mask, label) of one step's level records: leapfrog particles that
    converged, level clipped to 7."""
    lv = np.asarray(levels, np.uint8)
    keep = (lv & (LVL_LC | LVL_NOCONV)) == 0
    return keep, np.minimum(lv & LVL_MASK, 7).astype(np.uint8)


def class_weights(y) -> np.ndarray:
    """(n / count_c)^(1/2), normalised to mean 1 over the classes present."""
    cnt = np.bincount(np.asarray(y, np.int64), minlength=N_OUT).astype(np.float64)
    cw = np.where(cnt > 0, np.sqrt(len(y) / np.maximum(cnt, 1)), 0.0)
    return (cw * (cnt > 0).sum() / cw.sum()).astype(np.float32)
