"""Learned refinement binning: the level predictor's kernels and its use.

* features: the kernel's raw inputs equal the oracle's (log r, v_r, log |l|,
  log M(<r), log dt) of the same rank-ordered state;
* gradient: the MFMA backward pass (BF16 operands, FP32 accumulation)
  equals the oracle's gradient with the same operand rounding to within
  accumulation order, and the exact gradient to within bfloat16 resolution;
* predict: the MFMA forward pass gives the oracle's argmax except at
  near-ties of the two largest logits;
* training: the network trained by the library's own kernels on levels
  logged by the adaptive integrator beats the majority-class predictor;
* binning never changes results: with any network (trained, random, or one
  that puts every particle in the lowest or the highest bin) every step is
  bit-identical to the step without it, in both precisions, both binned
  forms, an ensemble and graph replay -- the adaptive convergence test runs
  on every particle from level 0 whatever the prediction was.
"""
import numpy as np
import pytest

import oracle
from oracle import mlp as om

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ns():
    import nsphere_amd
    return nsphere_amd


def assert_state(a, b, what):
    for k, name in enumerate(("r", "v", "ell", "id")):
        x, y = np.asarray(a[k]), np.asarray(b[k])
        bad = np.flatnonzero(x.view(np.uint64) != y.view(np.uint64)) if x.dtype == np.float64 else \
            np.flatnonzero(x != y)
        assert bad.size == 0, f"{what} {name}: {bad.size} differ"


def sorted_ic(ns, kind, n, seed, lowl=0.10):
    r, v, ell, ids = ns.generate_ic(kind, n, seed, 100.0, lowl, 1e-3)
    q = oracle.sort_perm(r)
    return r[q].copy(), v[q].copy(), ell[q].copy(), ids[q].copy()


def he_params(seed):
    rng = np.random.default_rng(seed)
    w1 = rng.uniform(-1, 1, (om.N_HID, om.N_IN)) * np.sqrt(6 / om.N_IN)
    w2 = rng.uniform(-1, 1, (om.N_HID, om.N_HID)) * np.sqrt(6 / om.N_HID)
    w3 = rng.uniform(-1, 1, (om.N_OUT, om.N_HID)) * np.sqrt(6 / om.N_HID)
    b1, b2, b3 = (rng.normal(0, 0.1, k) for k in (om.N_HID, om.N_HID, om.N_OUT))
    return om.pack(w1, b1, w2, b2, w3, b3)


def random_problem(seed, n):
    rng = np.random.default_rng(seed)
    X = rng.normal(0, 2, (n, om.N_IN)).astype(np.float32) + np.float32(1.5)
    y = rng.integers(0, om.N_OUT, n).astype(np.uint8)
    norm = np.concatenate([rng.normal(1.5, 0.2, om.N_IN), rng.uniform(0.3, 0.8, om.N_IN)]).astype(np.float32)
    cw = rng.uniform(0.5, 2.0, om.N_OUT).astype(np.float32)
    return he_params(seed), norm, cw, X, y


@pytest.mark.parametrize("precision", [0, 1])
def test_features_match_oracle(ns, precision):
    r, v, ell, ids = sorted_ic(ns, 2, 50_000, 3)
    p = ns.unit_params(halo_mass=2.0, dt=5e-4)
    if precision:
        r, v, ell = (x.astype(np.float32).astype(np.float64) for x in (r, v, ell))
    with ns.GpuContext(r.size) as ctx:
        ctx.set_option(ns.NS_OPT_PRECISION, precision)
        ctx.upload(r, v, ell, ids)
        got = ctx.features(p)
    want = om.features(r, v, ell, p.dt, p.halo_mass)
    np.testing.assert_array_equal(got[:, 1], want[:, 1])
    for k in (0, 2, 3, 4):  # single-precision logs: within 2 float32 ulp of the rounded float64 log
        ulp = np.spacing(np.abs(want[:, k]))
        assert np.all(np.abs(got[:, k] - want[:, k]) <= 2 * ulp), k


@pytest.mark.parametrize("n", [16, 4999, 100_000])
def test_gradient_matches_oracle(ns, n):
    P, norm, cw, X, y = random_problem(100 + n, n)
    m = ns.Mlp(seed=1)
    m.set_params(P, norm, cw)
    g, loss = m.gradient(X, y)
    gb, loss_b = om.gradient(P, norm, cw, X, y, bf16_ops=True)
    g64 = om.gradient_f64(P, norm, cw, X, y)
    for name, lo, hi in [("W1", om.W1, om.B1), ("b1", om.B1, om.W2), ("W2", om.W2, om.B2), ("b2", om.B2, om.W3),
                         ("W3", om.W3, om.B3), ("b3", om.B3, om.N_PARAMS)]:
        scale = np.linalg.norm(gb[lo:hi])
        err_b = np.linalg.norm(g[lo:hi] - gb[lo:hi]) / scale
        err_x = np.linalg.norm(g[lo:hi] - g64[lo:hi]) / np.linalg.norm(g64[lo:hi])
        assert err_b < 3e-3, f"{name}: {err_b} vs the bf16-operand oracle"
        assert err_x < 3e-2, f"{name}: {err_x} vs the exact gradient"
    live = om.pack(*(np.ones_like(b) for b in om.unpack(np.ones(om.N_PARAMS, np.float32)))) != 0
    assert not np.any(g[~live]), "gradient on padding"
    assert abs(loss - loss_b) <= 1e-4 * abs(loss_b)


def test_predict_matches_oracle(ns):
    P, norm, cw, X, _ = random_problem(7, 60_000)
    m = ns.Mlp(seed=1)
    m.set_params(P, norm, cw)
    got = m.predict(X)
    z = om.forward(P, norm, X, True)[3]
    want = np.argmax(z, axis=1)
    bad = np.flatnonzero(got != want)
    zs = np.sort(z, axis=1)
    margin = (zs[:, -1] - zs[:, -2]) / np.maximum(np.abs(zs[:, -1]), 1.0)
    assert bad.size <= X.shape[0] // 1000, bad.size
    assert np.all(margin[bad] < 1e-3), margin[bad].max()
    # the chosen logit is the largest up to the same near-tie
    assert np.all(z[bad, got[bad]] >= zs[bad, -1] - 1e-3 * np.maximum(np.abs(zs[bad, -1]), 1.0))


def test_train_sets_normalisation_and_class_weights(ns):
    rng = np.random.default_rng(4)
    X = (rng.normal(0, 1, (20_000, 5)) * [1, 2, 3, 4, 5] + [5, 4, 3, 2, 1]).astype(np.float32)
    y = rng.choice(8, 20_000, p=[0.5, 0.3, 0.1, 0.05, 0.05, 0, 0, 0]).astype(np.uint8)
    m = ns.Mlp(seed=2)
    m.train(X, y, epochs=1, batch=4096, lr=1e-3)
    _, norm, cw = m.params()
    X64 = X.astype(np.float64)
    np.testing.assert_allclose(norm[:5], X64.mean(axis=0), rtol=1e-6)
    np.testing.assert_allclose(norm[5:], 1.0 / X64.std(axis=0), rtol=1e-5)
    np.testing.assert_allclose(cw, om.class_weights(y), rtol=1e-6)
    with pytest.raises(ns.NsError):
        m.train(X, np.full(20_000, 8, np.uint8))  # level outside 0..7


def level_data(ns, n, seed, steps, precision=1):
    state = sorted_ic(ns, 2, n, seed)
    p = ns.unit_params()
    with ns.GpuContext(n) as ctx:
        ctx.set_option(ns.NS_OPT_PRECISION, precision)
        ctx.set_option(ns.NS_OPT_INTEGRATOR, 1)
        ctx.upload(*state)
        return ns.level_samples(ctx, p, steps)


def test_training_learns_levels(ns):
    X, y = level_data(ns, 400_000, 11, 4)
    Xt, yt = level_data(ns, 200_000, 12, 2)
    m = ns.Mlp(seed=3)
    m.train(X, y, epochs=1, batch=8192, lr=1e-3)
    P, norm, cw = m.params()
    first = om.loss(P, norm, cw, Xt, yt)
    for _ in range(4):
        m.train(X, y, epochs=1, batch=8192, lr=3e-3)
    P, norm, cw = m.params()
    last = om.loss(P, norm, cw, Xt, yt)
    pred = m.predict(Xt)
    conf = np.zeros((8, 8), np.int64)
    np.add.at(conf, (yt, pred), 1)
    # what binning needs: particles above level 0 found, level-0 particles left in bin 0
    deep, shallow = yt > 0, yt == 0
    recall = float(np.mean(pred[deep] > 0))
    false_pos = float(np.mean(pred[shallow] > 0))
    common = [c for c in range(8) if (yt == c).sum() >= 50]
    ba = float(np.mean([np.mean(pred[yt == c] == c) for c in common]))
    print(f"held-out loss {first:.4f} -> {last:.4f}; level>0 recall {recall:.3f}, level-0 false positives "
          f"{false_pos:.4f}; balanced accuracy over classes {common} {ba:.3f}\nconfusion (rows true, cols "
          f"predicted):\n{conf}")
    assert last < first
    assert recall - false_pos > 0.5          # majority predictor: 0
    assert ba > 1.6 / len(common)            # majority predictor: 1 / len(common)


def steps_with(ns, state, setup, mlp, nsteps=4, p=None):
    p = p or ns.unit_params()
    with ns.GpuContext(state[0].size) as ctx:
        setup(ctx)
        ctx.upload(*state)
        if mlp is not None:
            ctx.set_mlp(mlp)
        ctx.run(p, nsteps)
        out = ctx.download(), ctx.levels(), ctx.energy(p), ctx.cap_hits()
        ctx.set_mlp(None)
    return out


def networks(ns):
    """Random network; every particle in bin 0; every particle in bin 7."""
    nets = []
    for bias in (None, 0, 7):
        m = ns.Mlp(seed=5)
        if bias is not None:
            P, norm, cw = m.params()
            P[om.B3 + bias] = 1e4
            m.set_params(P)
        nets.append(m)
    return nets


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("form", [1, 2])
def test_binning_never_changes_results(ns, form, precision):
    state = sorted_ic(ns, 2, 120_000, 21)

    def setup(ctx):
        ctx.set_option(ns.NS_OPT_PRECISION, precision)
        ctx.set_option(ns.NS_OPT_INTEGRATOR, form)
        ctx.set_option(ns.NS_OPT_BUDGET, 1 << 16)

    base = steps_with(ns, state, setup, None)
    for k, m in enumerate(networks(ns)):
        got = steps_with(ns, state, setup, m)
        assert_state(got[0], base[0], f"network {k}")
        np.testing.assert_array_equal(got[1], base[1])
        assert got[2] == base[2] and got[3] == base[3]


def test_binning_ensemble_and_graph(ns):
    sizes = [20_000, 3_000, 45_000, 7_777]
    parts = [sorted_ic(ns, h % 2, s, 40 + h) for h, s in enumerate(sizes)]
    base = np.cumsum([0] + sizes[:-1])
    state = [np.concatenate([pt[k] + (b if k == 3 else 0) for pt, b in zip(parts, base)]) for k in range(4)]
    state[3] = state[3].astype(np.uint32)
    nets = networks(ns)
    for graph in (0, 1):
        def setup(ctx):
            ctx.set_option(ns.NS_OPT_INTEGRATOR, 2)
            ctx.set_option(ns.NS_OPT_PRECISION, 1)
            ctx.set_option(ns.NS_OPT_GRAPH, graph)
            ctx.set_halos(sizes, [1.0, 2.0, 0.5, 1.0])

        ref = steps_with(ns, state, setup, None, nsteps=5)
        for k, m in enumerate(nets):
            got = steps_with(ns, state, setup, m, nsteps=5)
            assert_state(got[0], ref[0], f"graph {graph} network {k}")
            np.testing.assert_array_equal(got[1], ref[1])


def test_binning_with_masses(ns):
    state = sorted_ic(ns, 0, 60_000, 33)
    mass = np.random.default_rng(5).uniform(0.5, 1.5, state[0].size) / state[0].size

    def setup(ctx):
        ctx.set_option(ns.NS_OPT_INTEGRATOR, 2)
        ctx.set_option(ns.NS_OPT_PRECISION, 1)
        ctx.upload(*state)
        ctx.set_masses(mass)

    ref = steps_with(ns, state, setup, None)
    for k, m in enumerate(networks(ns)):
        got = steps_with(ns, state, setup, m)
        assert_state(got[0], ref[0], f"masses network {k}")


def test_set_mlp_rejections(ns):
    from nsphere_amd import _lib
    m = ns.Mlp(seed=1)
    with ns.GpuContext(1000) as ctx:
        ctx.set_option(_lib.NS_OPT_SORT_ALG, 1)
        with pytest.raises(ns.NsError):
            ctx.set_mlp(m)
        ctx.set_option(_lib.NS_OPT_SORT_ALG, 0)
        ctx.set_mlp(m)
        with pytest.raises(ns.NsError):
            ctx.set_option(_lib.NS_OPT_SORT_ALG, 1)
        ctx.set_mlp(None)
        ctx.set_option(_lib.NS_OPT_SORT_ALG, 1)
    with pytest.raises(ns.NsError):
        m.predict(np.zeros((0, 5), np.float32))
    with pytest.raises(ValueError):
        m.predict(np.zeros((10, 4), np.float32))
