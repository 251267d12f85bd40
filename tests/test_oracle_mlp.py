"""Pins of the level-predictor oracle (oracle/mlp.py) on the CPU.

The network has no reference counterpart, so its oracle is pinned by what
the mathematics fixes: the analytic gradient against central finite
differences of the loss (every block: a transposed or dropped term fails),
the float32-packed gradient against the float64 one, the bfloat16-operand
gradient against the exact one within bfloat16's resolution, and the
bfloat16 rounding against brute force over its two neighbours.
"""
import numpy as np
import pytest

from oracle import mlp as om


def he_params(seed):
    rng = np.random.default_rng(seed)
    w1 = rng.uniform(-1, 1, (om.N_HID, om.N_IN)) * np.sqrt(6 / om.N_IN)
    w2 = rng.uniform(-1, 1, (om.N_HID, om.N_HID)) * np.sqrt(6 / om.N_HID)
    w3 = rng.uniform(-1, 1, (om.N_OUT, om.N_HID)) * np.sqrt(6 / om.N_HID)
    b1, b2, b3 = (rng.normal(0, 0.1, k) for k in (om.N_HID, om.N_HID, om.N_OUT))
    return om.pack(w1, b1, w2, b2, w3, b3)


def problem(seed, n=96):
    rng = np.random.default_rng(seed)
    X = rng.normal(0, 2, (n, om.N_IN)).astype(np.float32) + np.float32(1.5)
    y = rng.integers(0, om.N_OUT, n).astype(np.uint8)
    norm = np.concatenate([rng.normal(1.5, 0.2, om.N_IN), rng.uniform(0.3, 0.8, om.N_IN)]).astype(np.float32)
    cw = rng.uniform(0.5, 2.0, om.N_OUT).astype(np.float32)
    return he_params(seed), norm, cw, X, y


def test_bf16_rounding_brute_force():
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.normal(0, 1, 20000), rng.normal(0, 1e-3, 1000) * 10.0 ** rng.integers(-20, 20, 1000),
                        [1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -1.0 - 2 ** -8, 0.0, -0.0]]).astype(np.float32)
    got = om.bf16(x)
    bits = x.view(np.uint32)
    lo = (bits & 0xFFFF0000).view(np.float32)                       # truncation toward zero
    hi = ((bits & 0xFFFF0000) + 0x10000).astype(np.uint32).view(np.float32)  # next magnitude up
    dlo = np.abs(x.astype(np.float64) - lo)
    dhi = np.abs(x.astype(np.float64) - hi)
    lo_even = ((bits >> 16) & 1) == 0
    want = np.where(dlo < dhi, lo, np.where(dhi < dlo, hi, np.where(lo_even, lo, hi)))
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))
    assert om.bf16(np.float32(1.0 + 2 ** -8)) == 1.0                  # tie to even (down)
    assert om.bf16(np.float32(1.0 + 3 * 2 ** -8)) == np.float32(1.015625)  # tie to even (up)


def test_pack_unpack_roundtrip_and_padding():
    P = he_params(3)
    assert P.size == om.N_PARAMS == 6288
    blocks = om.unpack(P)
    np.testing.assert_array_equal(om.pack(*blocks), P)
    pad = np.ones(om.N_PARAMS, bool)
    pad[om.W1:om.B1].reshape(om.N_HID, om.N_IN_PAD)[:, :om.N_IN] = False
    pad[om.B1:om.W3] = False
    pad[om.W3:om.B3].reshape(om.N_OUT_PAD, om.N_HID)[:om.N_OUT] = False
    pad[om.B3:om.B3 + om.N_OUT] = False
    assert not P[pad].any()
    assert pad.sum() == om.N_HID * (om.N_IN_PAD - om.N_IN) + (om.N_OUT_PAD - om.N_OUT) * (om.N_HID + 1)


@pytest.mark.parametrize("seed", [11, 12])
def test_gradient_matches_finite_differences(seed):
    P, norm, cw, X, y = problem(seed)
    g = om.gradient_f64(P, norm, cw, X, y)
    rng = np.random.default_rng(seed)
    P64 = P.astype(np.float64)
    h = 1e-6
    picks = []
    for lo, hi in [(om.W1, om.B1), (om.B1, om.W2), (om.W2, om.B2), (om.B2, om.W3), (om.W3, om.B3),
                   (om.B3, om.B3 + om.N_OUT)]:
        live = [k for k in range(lo, hi) if not (lo == om.W1 and (k - lo) % om.N_IN_PAD >= om.N_IN)
                and not (lo == om.W3 and (k - lo) // om.N_HID >= om.N_OUT)]
        picks += list(rng.choice(live, min(12, len(live)), replace=False))
    for k in picks:
        e = np.zeros_like(P64)
        e[k] = h
        fd = (om.loss_f64_of(P64 + e, norm, cw, X, y) - om.loss_f64_of(P64 - e, norm, cw, X, y)) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(g[k])) + 1e-8, f"param {k}: fd {fd} vs {g[k]}"
    # loss of the exact network equals loss_f64_of (same mathematics, two codings)
    assert abs(om.loss(P, norm, cw, X, y) - om.loss_f64_of(P64, norm, cw, X, y)) < 1e-12


def test_float32_gradient_equals_float64():
    P, norm, cw, X, y = problem(21, n=500)
    g64 = om.gradient_f64(P, norm, cw, X, y)
    g32, loss = om.gradient(P, norm, cw, X, y, bf16_ops=False)
    np.testing.assert_allclose(g32, g64, rtol=1e-5, atol=1e-9)
    assert abs(loss - om.loss_f64_of(P.astype(np.float64), norm, cw, X, y)) < 1e-6


def test_bf16_gradient_close_to_exact():
    P, norm, cw, X, y = problem(22, n=2000)
    g64 = om.gradient_f64(P, norm, cw, X, y)
    gb, _ = om.gradient(P, norm, cw, X, y, bf16_ops=True)
    for lo, hi in [(om.W1, om.B1), (om.B1, om.W2), (om.W2, om.B2), (om.B2, om.W3), (om.W3, om.B3)]:
        err = np.linalg.norm(gb[lo:hi] - g64[lo:hi]) / np.linalg.norm(g64[lo:hi])
        assert err < 2e-2, (lo, err)  # bfloat16 keeps 8 significant bits
    z64 = om.forward(P, norm, X, False)[3]
    zb = om.forward(P, norm, X, True)[3]
    assert np.max(np.abs(zb - z64)) < 0.05 * np.max(np.abs(z64))


def test_features_closed_form():
    r = np.array([0.5, 1.0, 2.0, 4.0])
    v = np.array([0.1, -0.2, 0.3, -0.4])
    ell = np.array([0.01, -0.02, 0.0, 1.0])
    f = om.features(r, v, ell, 1e-3, halo_mass=2.0)
    np.testing.assert_array_equal(f[:, 0], np.log(r).astype(np.float32))
    np.testing.assert_array_equal(f[:, 1], v.astype(np.float32))
    assert f[2, 2] == np.float32(np.log(1e-30)) and f[1, 2] == np.float32(np.log(0.02))
    np.testing.assert_array_equal(f[:, 3], np.float32([np.log(1e-30), np.log(0.5), np.log(1.0), np.log(1.5)]))
    assert np.all(f[:, 4] == np.float32(np.log(1e-3)))


def test_labels_and_class_weights():
    lv = np.array([0, 3, 9, om.LVL_LC | 2, om.LVL_NOCONV | 5, 7, 1, 1], np.uint8)
    keep, y = om.train_labels(lv)
    assert keep.tolist() == [True, True, True, False, False, True, True, True]
    assert y[keep].tolist() == [0, 3, 7, 7, 1, 1]
    cw = om.class_weights(y[keep])
    present = cw > 0
    assert present.tolist() == [True, True, False, True, False, False, False, True]
    assert abs(cw[present].mean() - 1.0) < 1e-6
    assert cw[1] < cw[0]  # the commoner class weighs less
