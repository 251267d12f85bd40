"""Pins the oracle's integrators with mathematics, not by retyping them.

* closed forms: the force law, straight-line motion under identity gravity,
  a circular orbit as a fixed point;
* an independent high-accuracy ODE solution (scipy DOP853, rtol 1e-13) of
  dr/dt = v, dv/dt = -G M_i / r^2 + ell^2 / r^3 at frozen rank;
* the error structure of the adaptive output: a second-order method gives
  coarse/fine errors in ratio 4, and the reference's "Richardson" output
  4F - 3C (nsphere.c:1703-1708) then carries -2x the coarse error -- which
  the textbook (4F - C)/3 would not;
* the Levi-Civita branch (nsphere.c:2231-2334) and the leapfrog branch
  approximate the same ODE and must agree with it and with each other;
* the rank-based potential equals the brute-force pairwise shell sum.
"""
import math

import numpy as np
import pytest
from scipy.integrate import solve_ivp

import oracle


def exact(r0, v0, ell, gm, t):
    def f(_, y):
        return [y[1], -gm / y[0] ** 2 + ell * ell / y[0] ** 3]
    s = solve_ivp(f, (0.0, t), [r0, v0], method="DOP853", rtol=1e-13, atol=1e-16)
    return s.y[0, -1], s.y[1, -1]


def gm_rank(i, npts, p):
    return p.vel_conv_sq * p.G * (i / npts) * p.halo_mass


def test_force_closed_form():
    p = oracle.reference_units()
    L = oracle.lib()
    for (r, i, n) in [(1.5, 10, 100), (0.01, 99, 100), (30.0, 0, 100)]:
        want = -(p.vel_conv_sq * p.G) * (i / n) * p.halo_mass / (r * r)
        assert L.orc_gravitational_force(r, i, n, p) == pytest.approx(want, rel=1e-15)
    assert math.copysign(1, L.orc_gravitational_force(2.0, 0, 10, p)) == -1  # rank 0: -0.0
    assert L.orc_effective_angular_force(2.0, 3.0) == pytest.approx(9.0 / 8.0, rel=1e-15)
    q = oracle.params(identity_gravity=1)
    assert L.orc_gravitational_force(1.0, 5, 10, q) == 0.0


def test_identity_gravity_radial_line_is_exact():
    p = oracle.params(identity_gravity=1)
    r, v, lv = oracle.adaptive_leap(3, 10, 2.0, -0.7, 0.0, 0.05, p)
    assert lv == 0
    # exact up to the rounding of the drift and of 4F - 3C (a few ulp)
    assert r == pytest.approx(2.0 - 0.7 * 0.05, abs=8 * np.spacing(2.0))
    assert v == pytest.approx(-0.7, abs=8 * np.spacing(0.7))


@pytest.mark.parametrize("r0,v0,ell,h", [(1.0, 0.3, 0.8, 0.05), (0.5, -0.2, 0.3, 0.02), (2.0, 0.0, 1.0, 0.1)])
def test_identity_gravity_matches_straight_line(r0, v0, ell, h):
    p = oracle.params(identity_gravity=1, dt=h)
    r, v, lv = oracle.adaptive_leap(5, 10, r0, v0, ell, h, p)
    b = v0 * v0 + ell * ell / (r0 * r0)
    rx = math.sqrt(r0 * r0 + 2 * r0 * v0 * h + b * h * h)
    vx = (r0 * v0 + b * h) / rx
    # converged means |F - C| < tol |F|; with second-order errors C - X = (4/3)(C - F)
    # and the 4F - 3C output carries -2 (C - X): |R - X| < (8/3) tol |F|
    tol = 1e-7
    assert abs(r - rx) < 8.0 / 3.0 * tol * abs(rx) * 1.01
    assert abs(v - vx) < 8.0 / 3.0 * tol * max(abs(vx), 1e-30) * 1.01 + 1e-15


@pytest.mark.parametrize("r0,v0,ell,i,h", [(1.0, 0.3, 0.8, 400, 0.05), (0.6, -0.4, 0.35, 200, 0.03)])
def test_output_error_structure_pins_4f_minus_3c(r0, v0, ell, i, h):
    """Converge at level 0 (tolerance 1): C = one KDK step of h, F = two of h/2."""
    npts = 1000
    outs = {}
    for ot in (0, 1, 2):
        p = oracle.params(out_type=ot, radius_tol=1.0, velocity_tol=1.0, dt=h)
        outs[ot] = oracle.adaptive_leap(i, npts, r0, v0, ell, h, p)
        assert outs[ot][2] == 0
    rx, vx = exact(r0, v0, ell, gm_rank(i, npts, oracle.params()), h)
    eC = outs[0][0] - rx
    eF = outs[1][0] - rx
    eR = outs[2][0] - rx
    assert eC / eF == pytest.approx(4.0, rel=0.05), "coarse/fine must be h and h/2 second-order passes"
    assert eR == pytest.approx(-2.0 * eC, rel=0.05), "output must be 4F - 3C"
    assert abs(eR) > 10 * abs((4 * outs[1][0] - outs[0][0]) / 3 - rx)


@pytest.mark.parametrize("r0,v0,ell,i", [(1.0, 0.3, 0.8, 400), (0.3, 0.05, 0.1, 30), (5.0, -0.2, 1.5, 900)])
def test_adaptive_leap_matches_ode_solution(r0, v0, ell, i):
    npts, h = 1000, 1e-3
    p = oracle.params(dt=h)
    r, v, lv = oracle.adaptive_leap(i, npts, r0, v0, ell, h, p)
    rx, vx = exact(r0, v0, ell, gm_rank(i, npts, p), h)
    assert abs(r - rx) <= 1e-9 * rx
    assert abs(v - vx) <= 1e-9 * max(abs(vx), 1.0)


def test_level_is_first_converging_level():
    """The returned level is the first k at which the coarse (2^k) and fine
    (2 * 2^k) passes agree to the tolerance (nsphere.c:1670-1716)."""
    npts, i, h = 1000, 500, 0.2
    r0, v0, ell = 0.8, 0.25, 0.5
    p = oracle.params(dt=h, radius_tol=1e-9, velocity_tol=1e-9)
    _, _, lv = oracle.adaptive_leap(i, npts, r0, v0, ell, h, p)
    first = None
    for k in range(20):
        N = 1 << k
        rc, vc = oracle.micro_leapfrog(i, npts, r0, v0, ell, h, N, 2 * N + 1, p)
        rf, vf = oracle.micro_leapfrog(i, npts, r0, v0, ell, h, N, 4 * N + 1, p)
        if abs(rf - rc) / (abs(rf) + 1e-30) < 1e-9 and abs(vf - vc) / (abs(vf) + 1e-30) < 1e-9:
            first = k
            break
    assert first is not None and first >= 2
    assert lv == first


def test_nonconvergence_returns_last_level_flagged():
    p = oracle.params(dt=0.2, radius_tol=1e-15, velocity_tol=1e-15, max_subdiv=8)
    _, _, lv = oracle.adaptive_leap(500, 1000, 0.8, 0.25, 0.5, 0.2, p)
    assert lv == (3 | oracle.LVL_NOCONV)


def test_circular_orbit_is_fixed_point():
    npts, i = 1000, 600
    gm = 0.6
    ell = 0.9
    rc = ell * ell / gm
    p = oracle.params(max_subdiv=16)
    r, v, _ = oracle.adaptive_leap(i, npts, rc, 0.0, ell, 1e-3, p)
    assert r == pytest.approx(rc, rel=1e-13)
    assert abs(v) < 1e-12


@pytest.mark.parametrize("r0,v0,ell,i", [(0.5, 0.1, 0.3, 500), (0.05, -0.2, 0.02, 100), (0.2, 0.3, 0.05, 300)])
def test_levi_civita_matches_ode_solution(r0, v0, ell, i):
    """The regularised branch (rho = sqrt r, dt = rho^2 dtau) integrates the
    same ODE: its result approaches the exact solution as the tau step
    shrinks (the scheme is first order in dtau -- its drift uses rho at the
    start of the micro-step -- and the final overshoot is interpolated
    linearly), and agrees with the leapfrog branch."""
    npts, dt = 1000, 1e-3
    rx, vx = exact(r0, v0, ell, gm_rank(i, npts, oracle.params()), dt)
    errs = []
    for nt in (10, 40, 640):
        p = oracle.params(dt=dt, N_taumin=nt, max_subdiv=1 << 16)
        r, v, lv, ntau = oracle.adaptive_lc(i, npts, r0, v0, ell, dt, p)
        assert lv & oracle.LVL_LC
        assert 2 * nt <= ntau <= 2 * nt + 8      # dtau = dt / (2 r N_taumin), t ~ r tau
        errs.append(abs(r - rx) / rx)
        assert abs(v - vx) <= 1e-7 * max(abs(vx), 1.0)
    assert errs[0] < 1e-7
    assert errs[2] < errs[0] / 4
    p = oracle.params(dt=dt)
    rl, _, _ = oracle.adaptive_leap(i, npts, r0, v0, ell, dt, p)
    r10, _, _, _ = oracle.adaptive_lc(i, npts, r0, v0, ell, dt, p)
    assert r10 == pytest.approx(rl, rel=1e-7)


def test_levi_civita_level_search_is_first_order():
    """A strong-field plunge: coarse/fine differences halve per level, so the
    tau step needs deep levels; with max_subdiv 64 it stops unconverged at
    level 6, flagged, after about 2 N_taumin tau steps."""
    p = oracle.params(max_subdiv=64)
    r, v, lv, ntau = oracle.adaptive_lc(20, 1000, 1e-3, -0.5, 1e-4, 1e-3, p)
    assert lv == (oracle.LVL_LC | oracle.LVL_NOCONV | 7)
    assert 20 <= ntau <= 60


def test_levi_civita_zero_radius_passthrough():
    p = oracle.params()
    r, v, lv, ntau = oracle.adaptive_lc(3, 10, 0.0, 0.25, 0.1, 1e-3, p)
    assert (r, v, ntau) == (0.0, 0.25, 0)


def test_energy_equals_pairwise_shell_sum():
    rng = np.random.default_rng(1)
    n = 400
    r = np.sort(rng.uniform(0.1, 10, n))
    v = rng.normal(0, 0.3, n)
    ell = rng.uniform(0, 1, n)
    p = oracle.params(G=1.3, halo_mass=2.0)
    K, W = oracle.energy(r, v, ell, p)
    m = p.halo_mass / n
    Kx = np.sum(0.5 * m * (v * v + ell * ell / (r * r)))
    rr = np.maximum.outer(r, r)
    iu = np.triu_indices(n, 1)
    Wx = -p.G * m * m * np.sum(1.0 / rr[iu])
    assert K == pytest.approx(Kx, rel=1e-13)
    assert W == pytest.approx(Wx, rel=1e-12)


def test_step_inverse_map_and_order():
    from nsphere_amd import generate_ic
    r, v, ell, ids = generate_ic(0, 5000, 3, 100.0, 0.05, 1e-3)
    r0 = r.copy()
    lv, inv = oracle.step(r, v, ell, ids, oracle.params())
    # ids are the initial positions, now in radius order of the step's start
    np.testing.assert_array_equal(r0[ids], np.sort(r0))
    np.testing.assert_array_equal(inv[ids], np.arange(5000))


def test_strict_and_fast_builds_agree():
    from nsphere_amd import generate_ic
    r, v, ell, ids = generate_ic(0, 20000, 9, 100.0, 0.05, 1e-3)
    q = oracle.sort_perm(r)
    a = [x[q].copy() for x in (r, v, ell, ids)]
    b = [x.copy() for x in a]
    p = oracle.params()
    oracle.step(*a, p)
    oracle.step(*b, p, fast=True)
    assert np.array_equal(a[3], b[3])
    assert np.max(np.abs(a[0] - b[0]) / a[0]) < 1e-6


# ---- Levi-Civita branch: its quirks pinned by mathematics ----------------
#
# In tau (dt = rho^2 dtau, rho = sqrt r) the motion is
#   drho/dtau = rho v / 2,  dv/dtau = -G M_i / rho^2 + ell^2 / rho^4,  dt/dtau = rho^2
# (nsphere.c:1757-1787).  The reference's micro-step drifts rho with rho at
# the start of the step (nsphere.c:2017), so a pass of M substeps has a
# first-order error in rho, e(M) ~ K / M.  With no force (identity gravity,
# ell = 0) a pass has a closed form: v stays v0, rho_k = rho0 q^k with
# q = 1 + a / 2, a = dtau v0, and t = sum dtau rho_k^2 (1 + a / 4)^2.


def lc_pass_closed_form(rho0, v0, h, m):
    """(rho, t) of a force-free pass of m substeps over h."""
    dtau = h / m
    a = dtau * v0
    q = 1.0 + 0.5 * a
    rho = rho0 * q ** m
    t = dtau * rho0 * rho0 * (1.0 + 0.25 * a) ** 2 * (q ** (2 * m) - 1.0) / (q * q - 1.0)
    return rho, t


def exact_tau(rho0, v0, ell, gm, h):
    def f(_, y):
        rho, v, _t = y
        return [0.5 * rho * v, -gm / rho ** 2 + ell * ell / rho ** 4, rho * rho]
    s = solve_ivp(f, (0.0, h), [rho0, v0, 0.0], method="DOP853", rtol=1e-13, atol=1e-18)
    return s.y[:, -1]


def test_lc_micro_pass_closed_form():
    rho0, v0, h = 0.3, -0.8, 0.6
    p = oracle.params(identity_gravity=1)
    for m in (1, 3, 5, 33, 1000):
        rho, v, t = oracle.micro_lc(3, 10, rho0, v0, 0.0, m, h, 0.0, p)
        rx, tx = lc_pass_closed_form(rho0, v0, h, m)
        assert rho == pytest.approx(rx, rel=1e-13)
        assert t == pytest.approx(tx, rel=1e-13)
        assert v == v0


def test_lc_single_tau_outputs_closed_form():
    """Level 0 (tolerance 1): the coarse pass has 3 substeps, the fine 5, and
    out_type 2 returns 4F - 3C on rho and t (nsphere.c:2120-2157)."""
    rho0, v0, h = 0.3, -0.8, 0.6
    C = lc_pass_closed_form(rho0, v0, h, 3)
    F = lc_pass_closed_form(rho0, v0, h, 5)
    R = (4.0 * F[0] - 3.0 * C[0], 4.0 * F[1] - 3.0 * C[1])
    for ot, want in ((0, C), (1, F), (2, R)):
        p = oracle.params(identity_gravity=1, out_type=ot, radius_tol=1.0, velocity_tol=1.0)
        rho, v, t, lv = oracle.single_tau_lc(3, 10, rho0, v0, 0.0, h, 0.0, p)
        assert lv == 0
        assert rho == pytest.approx(want[0], rel=1e-13)
        assert t == pytest.approx(want[1], rel=1e-13)
        assert v == pytest.approx(v0, abs=4 * np.spacing(abs(v0)))
    # the three candidates are far apart at this step size
    assert abs(R[0] - F[0]) > 1e-3 * R[0] and abs((4 * F[0] - C[0]) / 3 - R[0]) > 1e-3 * R[0]


@pytest.mark.parametrize("r0,v0,ell,i,h", [(0.04, -0.3, 0.02, 100, 0.02), (0.1, -0.5, 0.05, 300, 0.01)])
def test_lc_single_tau_error_structure_pins_4f_minus_3c(r0, v0, ell, i, h):
    """Against an independent DOP853 solution of the tau ODE over one tau
    step: the first-order passes of 3 and 5 substeps err in ratio
    e_F / e_C = 3/5, and 4F - 3C then errs by -3/5 e_C.  F alone (+3/5), C
    (1) or the textbook (4F - C)/3 (+7/15) would fail."""
    npts = 1000
    rho0 = math.sqrt(r0)
    outs = {}
    for ot in (0, 1, 2):
        p = oracle.params(out_type=ot, radius_tol=1.0, velocity_tol=1.0)
        outs[ot] = oracle.single_tau_lc(i, npts, rho0, v0, 0.0, h, ell, p)
        assert outs[ot][3] == 0
    rx = exact_tau(rho0, v0, ell, gm_rank(i, npts, oracle.params()), h)[0]
    eC, eF, eR = (outs[k][0] - rx for k in (0, 1, 2))
    assert abs(eC) > 1e-9 * rho0
    assert 0.55 < eF / eC < 0.65
    assert -0.7 < eR / eC < -0.5


def test_lc_nonconvergence_falls_back_to_fine_pass_at_doubled_n():
    """No level converges up to max_subdiv = 4 (N = 1, 2, 4): the result is
    the fine pass at the next N, 4 * 8 + 1 = 33 substeps (nsphere.c:2168-2185),
    flagged unconverged at level 3."""
    rho0, v0, h = 0.3, -0.8, 0.6
    p = oracle.params(identity_gravity=1, radius_tol=1e-30, velocity_tol=1.0, max_subdiv=4)
    rho, v, t, lv = oracle.single_tau_lc(3, 10, rho0, v0, 0.0, h, 0.0, p)
    assert lv == (3 | oracle.LVL_NOCONV)
    rx, tx = lc_pass_closed_form(rho0, v0, h, 33)
    assert rho == pytest.approx(rx, rel=1e-13)
    assert t == pytest.approx(tx, rel=1e-13)
    assert v == v0
    r17, _ = lc_pass_closed_form(rho0, v0, h, 17)  # the last tried fine pass is distinguishable
    assert abs(r17 - rx) > 1e-4 * rx


@pytest.mark.parametrize("r0,v0,nt", [(0.01, -0.5, 10), (0.02, 0.7, 10), (0.005, -0.3, 3)])
def test_lc_overshoot_interpolation_on_a_radial_line(r0, v0, nt):
    """With no force a radial orbit is the straight line r0 + v0 t.  The tau
    loop stops at the step that passes t = dt and interpolates rho and v
    linearly in t (nsphere.c:2294-2308): the result lies on the line to the
    scheme's tolerance, while the states on either side of dt miss it by
    about 1e-3."""
    dt = 1e-3 * (2.0 if v0 > 0 else 1.0)
    p = oracle.params(identity_gravity=1, dt=dt, N_taumin=nt)
    r, v, lv, ntau = oracle.adaptive_lc(5, 100, r0, v0, 0.0, dt, p)
    assert lv & oracle.LVL_LC and ntau >= 2
    line = r0 + v0 * dt
    assert abs(r - line) < 1e-5 * line
    assert v == pytest.approx(v0, abs=4 * np.spacing(abs(v0)))
    # the tau step that crossed dt, from the state before it
    rho, vv, t = math.sqrt(r0), v0, 0.0
    dtau = dt / (2.0 * r0 * nt)
    while True:
        rn, vn, tn, _ = oracle.single_tau_lc(5, 100, rho, vv, t, dtau, 0.0, p)
        if tn > dt:
            break
        rho, vv, t = rn, vn, tn
    assert abs(rho * rho - line) > 1e-4 * line and abs(rn * rn - line) > 1e-4 * line
