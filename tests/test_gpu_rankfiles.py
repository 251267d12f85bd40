"""Rank_Mass_Rad_VRad files of the current state against the oracle's
restatement of the reference's post-processing (oracle/rankfiles.py) fed the
same snapshot records.

Rank, mass, r, v_r and ell are exact; PsiA, E and density are float32 of
values whose float64 forms differ from the oracle's only by summation order
(the suffix sum of 1/r, the Gaussian's normalisation and the FFT), so they
agree to a few float32 ulp -- and E, a difference of nearly equal terms for
bound particles, to a few ulp of its largest term.
"""
import numpy as np
import pytest

import oracle
from oracle import rankfiles as orf

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ns():
    import nsphere_amd
    return nsphere_amd


def close_f32(got, want, scale, ulps, what):
    tol = ulps * np.spacing(np.maximum(np.abs(want), scale).astype(np.float32)).astype(np.float64)
    bad = np.flatnonzero(np.abs(got.astype(np.float64) - want.astype(np.float64)) > tol)
    assert bad.size == 0, f"{what}: {bad.size} rows, first {bad[0]}: {got[bad[0]]} vs {want[bad[0]]}"


@pytest.mark.parametrize("kind,n,precision", [(0, 50_000, 0), (1, 200_000, 1)])
def test_rank_files_match_oracle(ns, kind, n, precision, tmp_path):
    r, v, ell, ids = ns.generate_ic(kind, n, 500 + n, 100.0, 0.05, 1e-3)
    p = ns.default_params()
    with ns.GpuContext(n) as ctx:
        ctx.set_option(ns.NS_OPT_PRECISION, precision)
        ctx.set_option(ns.NS_OPT_INTEGRATOR, 2)
        ctx.upload(r, v, ell, ids)
        ctx.run(ns.unit_params(), 2)  # positions no longer in id or radius order
        rec = ctx.snapshot_records()
        ctx.rank_files(p, tmp_path / "unsorted.dat", tmp_path / "sorted.dat")
    un = np.fromfile(tmp_path / "unsorted.dat", orf.UNSORTED)
    so = np.fromfile(tmp_path / "sorted.dat", orf.SORTED)
    wun, wso = orf.rank_files(rec, p.G, p.vel_conv_sq, p.halo_mass)
    assert un.size == n and so.size == n
    for got, want in ((un, wun), (so, wso)):
        for col in ("rank", "mass", "r", "v", "ell"):
            np.testing.assert_array_equal(got[col], want[col], err_msg=col)
        close_f32(got["psi"], want["psi"], 0.0, 4, "psi")
        kin = np.maximum(np.abs(want["psi"]), 0.5 * want["v"].astype(np.float64) ** 2)
        close_f32(got["E"], want["E"], kin, 4, "E")
    close_f32(so["density"], wso["density"], 1e-4 * np.max(wso["density"]), 8, "density")


def test_rank_files_rejections(ns, tmp_path):
    r, v, ell, ids = ns.generate_ic(0, 1000, 1, 100.0, 0.0, 1e-3)
    with ns.GpuContext(1000) as ctx:
        with pytest.raises(ns.NsError):
            ctx.rank_files(ns.default_params(), tmp_path / "u.dat")  # no state
        ctx.upload(r, v, ell, ids)
        with pytest.raises(ns.NsError):
            ctx.rank_files(ns.default_params(), tmp_path / "missing_dir" / "u.dat")
        ctx.set_halos([500, 500])
        with pytest.raises(ns.NsError):
            ctx.rank_files(ns.default_params(), tmp_path / "u.dat")
