"""CPU oracle for the NSphere per-step loop -- TEST INFRASTRUCTURE ONLY.

ctypes wrapper of libns_oracle.so (strict IEEE restatement of the reference,
see ns_oracle.h for the file:line of every function) and, when it has been
built from the reference tree, of _ref/libref_quadsort.so (the reference's
own quadsort).  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may import this package; the product never does.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LVL_MASK, LVL_NOCONV, LVL_LC = 0x3F, 0x40, 0x80


class OrcParams(ctypes.Structure):
    _fields_ = [
        ("dt", ctypes.c_double),
        ("G", ctypes.c_double),
        ("halo_mass", ctypes.c_double),
        ("vel_conv_sq", ctypes.c_double),
        ("radius_tol", ctypes.c_double),
        ("velocity_tol", ctypes.c_double),
        ("alpha", ctypes.c_double),
        ("max_subdiv", ctypes.c_int32),
        ("out_type", ctypes.c_int32),
        ("N_taumin", ctypes.c_int32),
        ("identity_gravity", ctypes.c_int32),
        ("budget", ctypes.c_int64),
        ("menc", ctypes.c_void_p),   # NULL: the reference's (rank/npts) M
    ]


def params(dt=1e-3, G=1.0, halo_mass=1.0, vel_conv_sq=1.0, radius_tol=1e-7, velocity_tol=1e-7,
           alpha=0.05, max_subdiv=4096 * 4096 * 16, out_type=2, N_taumin=10,
           identity_gravity=0, budget=0) -> OrcParams:
    """Method-5 constants (nsphere.c:5777-5782); G = M = 1 units by default.
    budget: force evaluations per particle-step (0 = unlimited, as the reference)."""
    return OrcParams(dt, G, halo_mass, vel_conv_sq, radius_tol, velocity_tol, alpha, max_subdiv,
                     out_type, N_taumin, identity_gravity, budget)


def reference_units(**kw) -> OrcParams:
    """The reference's kpc/Myr/Msun constants, nsphere.c:364-373."""
    d = dict(G=4.3e-6, halo_mass=1.0e12, vel_conv_sq=1.02271e-3 * 1.02271e-3)
    d.update(kw)
    return params(**d)


def from_step_params(sp, budget: int = 0) -> OrcParams:
    """Copy a nsphere_amd.StepParams into OrcParams (budget: the GPU context's
    NS_OPT_BUDGET when it is to be mirrored, else 0 = unlimited)."""
    vals = [getattr(sp, f) for f, _ in OrcParams._fields_ if f not in ("budget", "menc")]
    return OrcParams(*vals, budget)


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)


def build_ref() -> bool:
    """Build _ref/ from the reference's quadsort sources when they exist here."""
    if not os.path.isdir("/root/reference/src/c"):
        return False
    subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)
    return True


def _load(name: str):
    path = os.path.join(HERE, name)
    if not os.path.exists(path):
        raise OSError(f"{path} not built (run `make -C oracle`)")
    return ctypes.CDLL(path)


_libs: dict = {}


def lib(fast: bool = False):
    key = "fast" if fast else "strict"
    if key not in _libs:
        L = _load("libns_oracle_fast.so" if fast else "libns_oracle.so")
        vp, i64, i32, dbl = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_double
        P = ctypes.POINTER(OrcParams)
        pd = ctypes.POINTER(ctypes.c_double)
        L.orc_gravitational_force.restype = dbl
        L.orc_gravitational_force.argtypes = [dbl, i32, i32, P]
        L.orc_effective_angular_force.restype = dbl
        L.orc_effective_angular_force.argtypes = [dbl, dbl]
        L.orc_micro_leapfrog.restype = None
        L.orc_micro_leapfrog.argtypes = [i32, i32, dbl, dbl, dbl, dbl, i32, i32, P, pd, pd]
        L.orc_adaptive_full_leap.restype = i32
        L.orc_adaptive_full_leap.argtypes = [i32, i32, dbl, dbl, dbl, dbl, P, pd, pd]
        L.orc_micro_lc.restype = None
        L.orc_micro_lc.argtypes = [i32, i32, dbl, dbl, dbl, i64, dbl, P, dbl, pd, pd, pd]
        L.orc_single_tau_step_lc.restype = i32
        L.orc_single_tau_step_lc.argtypes = [i32, i32, dbl, dbl, dbl, dbl, P, dbl, pd, pd, pd]
        L.orc_adaptive_full_lc.restype = i32
        L.orc_adaptive_full_lc.argtypes = [i32, i32, dbl, dbl, dbl, dbl, P, pd, pd,
                                           ctypes.POINTER(ctypes.c_int64)]
        L.orc_radius_key.restype = ctypes.c_uint64
        L.orc_radius_key.argtypes = [dbl]
        L.orc_sort_perm.restype = None
        L.orc_sort_perm.argtypes = [vp, vp, i64, i32, vp]
        L.orc_sort_perm_chunked.restype = None
        L.orc_sort_perm_chunked.argtypes = [vp, i64, vp]
        L.orc_step.restype = None
        L.orc_step.argtypes = [vp, vp, vp, vp, i64, P, i32, i32, vp, vp]
        L.orc_integrate.restype = None
        L.orc_integrate.argtypes = [vp, vp, vp, i64, P, vp]
        L.orc_energy.restype = None
        L.orc_energy.argtypes = [vp, vp, vp, i64, P, pd, pd]
        L.orc_energy_masses.restype = None
        L.orc_energy_masses.argtypes = [vp, vp, vp, vp, i64, P, pd, pd]
        L.orc_mass_prefix.restype = None
        L.orc_mass_prefix.argtypes = [vp, i64, vp]
        L.orc_num_threads.restype = i32
        L.orc_set_num_threads.argtypes = [i32]
        _libs[key] = L
    return _libs[key]


def ref_lib():
    """The reference's own quadsort (None when _ref/ was not built)."""
    if "ref" not in _libs:
        path = os.path.join(HERE, "_ref", "libref_quadsort.so")
        if not os.path.exists(path):
            return None
        L = ctypes.CDLL(path)
        L.ref_quadsort_perm.restype = ctypes.c_int
        L.ref_quadsort_perm.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p]
        _libs["ref"] = L
    return _libs["ref"]


def _p(a):
    return None if a is None else a.ctypes.data


# -- thin numpy-facing helpers ------------------------------------------

def sort_perm(r, ids=None, tiebreak_id=False) -> np.ndarray:
    r = np.ascontiguousarray(r, np.float64)
    ids_c = None if ids is None else np.ascontiguousarray(ids, np.uint32)
    perm = np.empty(r.size, np.uint32)
    lib().orc_sort_perm(_p(r), _p(ids_c), r.size, int(tiebreak_id), _p(perm))
    return perm


def sort_perm_chunked(r) -> np.ndarray:
    r = np.ascontiguousarray(r, np.float64)
    perm = np.empty(r.size, np.uint32)
    lib().orc_sort_perm_chunked(_p(r), r.size, _p(perm))
    return perm


def ref_quadsort_perm(r) -> np.ndarray:
    L = ref_lib()
    if L is None:
        raise OSError("oracle/_ref not built (needs the reference tree)")
    r = np.ascontiguousarray(r, np.float64)
    perm = np.empty(r.size, np.uint32)
    if L.ref_quadsort_perm(_p(r), r.size, _p(perm)) != 0:
        raise MemoryError("ref_quadsort_perm")
    return perm


def radius_key(x: float) -> int:
    return lib().orc_radius_key(float(x))


def step(r, v, ell, ids, p: OrcParams, chunked=False, tiebreak_id=False, fast=False):
    """One method-5 step in place on float64/uint32 arrays; returns (levels, inverse_map)."""
    n = r.size
    lv = np.empty(n, np.uint8)
    inv = np.empty(n, np.int32)
    lib(fast).orc_step(_p(r), _p(v), _p(ell), _p(ids), n, ctypes.byref(p), int(chunked),
                       int(tiebreak_id), _p(inv), _p(lv))
    return lv, inv


def integrate(r, v, ell, p: OrcParams, fast=False) -> np.ndarray:
    n = r.size
    lv = np.empty(n, np.uint8)
    lib(fast).orc_integrate(_p(r), _p(v), _p(ell), n, ctypes.byref(p), _p(lv))
    return lv


def mass_prefix(mass) -> np.ndarray:
    """Enclosed mass of rank-ordered masses: [i] below rank i (long double sum,
    one rounding), [n] total, [n + 1] the rank-0 mass."""
    m = np.ascontiguousarray(mass, np.float64)
    out = np.empty(m.size + 2, np.float64)
    lib().orc_mass_prefix(_p(m), m.size, _p(out))
    return out


def with_menc(p: OrcParams, menc) -> OrcParams:
    """A copy of p whose force uses the enclosed-mass array menc (layout of
    mass_prefix); the array is kept alive on the returned object."""
    q = OrcParams()
    ctypes.pointer(q)[0] = p
    arr = np.ascontiguousarray(menc, np.float64)
    q.menc = arr.ctypes.data
    q._menc_ref = arr
    return q


def energy_masses(r, v, ell, mass, p: OrcParams):
    k = ctypes.c_double()
    w = ctypes.c_double()
    lib().orc_energy_masses(_p(np.ascontiguousarray(r, np.float64)), _p(np.ascontiguousarray(v, np.float64)),
                            _p(np.ascontiguousarray(ell, np.float64)),
                            _p(np.ascontiguousarray(mass, np.float64)), r.size, ctypes.byref(p),
                            ctypes.byref(k), ctypes.byref(w))
    return k.value, w.value


def energy(r, v, ell, p: OrcParams):
    k = ctypes.c_double()
    w = ctypes.c_double()
    lib().orc_energy(_p(np.ascontiguousarray(r, np.float64)), _p(np.ascontiguousarray(v, np.float64)),
                     _p(np.ascontiguousarray(ell, np.float64)), r.size, ctypes.byref(p),
                     ctypes.byref(k), ctypes.byref(w))
    return k.value, w.value


def adaptive_leap(i, npts, r, v, ell, h, p: OrcParams):
    ro, vo = ctypes.c_double(), ctypes.c_double()
    lv = lib().orc_adaptive_full_leap(i, npts, r, v, ell, h, ctypes.byref(p), ctypes.byref(ro),
                                      ctypes.byref(vo))
    return ro.value, vo.value, lv


def adaptive_lc(i, npts, r, v, ell, dt, p: OrcParams):
    ro, vo = ctypes.c_double(), ctypes.c_double()
    nt = ctypes.c_int64()
    lv = lib().orc_adaptive_full_lc(i, npts, r, v, ell, dt, ctypes.byref(p), ctypes.byref(ro),
                                    ctypes.byref(vo), ctypes.byref(nt))
    return ro.value, vo.value, lv, nt.value


def micro_lc(i, npts, rho, v, t, substeps, h_tau, ell, p: OrcParams):
    """doMicroLeviCivita (nsphere.c:1990-2036): (rho, v, t) after substeps KDK steps in tau."""
    ro, vo, to = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    lib().orc_micro_lc(i, npts, rho, v, t, substeps, h_tau, ctypes.byref(p), ell, ctypes.byref(ro),
                       ctypes.byref(vo), ctypes.byref(to))
    return ro.value, vo.value, to.value


def single_tau_lc(i, npts, rho, v, t, h_tau, ell, p: OrcParams):
    """doSingleTauStepAdaptiveLeviCivita (nsphere.c:2077-2186): (rho, v, t, level record)."""
    ro, vo, to = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    lv = lib().orc_single_tau_step_lc(i, npts, rho, v, t, h_tau, ctypes.byref(p), ell, ctypes.byref(ro),
                                      ctypes.byref(vo), ctypes.byref(to))
    return ro.value, vo.value, to.value, lv


def micro_leapfrog(i, npts, r, v, ell, h, N, sub, p: OrcParams):
    ro, vo = ctypes.c_double(), ctypes.c_double()
    lib().orc_micro_leapfrog(i, npts, r, v, ell, h, N, sub, ctypes.byref(p), ctypes.byref(ro),
                             ctypes.byref(vo))
    return ro.value, vo.value


def set_threads(n: int) -> None:
    lib().orc_set_num_threads(n)
    lib(True).orc_set_num_threads(n)
