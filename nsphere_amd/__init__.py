"""nsphere_amd -- MI355X-native per-step evolution loop of NSphere.

Host-side mirror of the reference's hot-path interface over the C-ABI of
libnsphere_gpu.so (include/nsphere_gpu.h):

* ``sort_particles(particles, npts)`` -- nsphere.c:8662-8665 (in place, all
  columns move together, stable on radius);
* ``method5_step(particles, npts, dt, inverse_map)`` -- the method-5 loop body,
  nsphere.c:5775-5838 (sort, inverse map, adaptive integration at frozen rank);
* ``GpuContext`` -- device-resident state for running many steps without
  host round trips.

Particles use the reference's layout: ``particles`` is a (5, npts) float64
array holding r, v_r, ell, id (as a double) and mu (not used by the step, but
moved with the others by every reordering, as the reference moves it).
"""
from __future__ import annotations

import ctypes

import numpy as np

from ._lib import (FORM_BINNED, FORM_BINNED_FAST, FORM_PLAIN, NS_LVL_LC, NS_LVL_MASK, NS_LVL_NOCONV,
                   NS_OPT_BUDGET, NS_OPT_INTEGRATOR, NS_OPT_PRECISION, NS_OPT_PROFILE, NS_OPT_PROFILE_KERNELS,
                   NS_OPT_SORT_KEYS, NS_OPT_GRAPH_STEPS,
                   NS_OPT_GRAPH, NS_OPT_OVERLAP, NS_OPT_SORT_ALG, NS_OPT_SORT_TILE, NS_OPT_TIEBREAK,
                   SORT_ONESWEEP, SORT_REDUCE_SCAN, NsError, StepParams, check, lib)

__all__ = [
    "GpuContext", "StepParams", "NsError", "default_params", "unit_params", "generate_ic",
    "sort_particles", "method5_step", "philox4x32_10",
    "NS_LVL_MASK", "NS_LVL_NOCONV", "NS_LVL_LC",
    "NS_OPT_TIEBREAK", "NS_OPT_PROFILE", "NS_OPT_INTEGRATOR", "NS_OPT_BUDGET", "NS_OPT_PRECISION",
    "NS_OPT_SORT_TILE", "NS_OPT_SORT_ALG", "NS_OPT_GRAPH", "NS_OPT_OVERLAP", "NS_OPT_PROFILE_KERNELS",
    "NS_OPT_SORT_KEYS", "NS_OPT_GRAPH_STEPS",
    "Mlp", "level_samples", "ApdWriter", "APD_RECORD",
    "read_all_particle_data", "write_ic_file", "read_ic_file", "SORT_REDUCE_SCAN", "SORT_ONESWEEP",
]

IC_PLUMMER, IC_HERNQUIST, IC_MIX = 0, 1, 2


def default_params(**kw) -> StepParams:
    """Method-5 constants in the reference's kpc / Myr / Msun units."""
    p = StepParams()
    lib.ns_step_params_default(ctypes.byref(p))
    return p.replace(**kw) if kw else p


def unit_params(**kw) -> StepParams:
    """Method-5 constants with G = M = 1 and no velocity conversion."""
    p = StepParams()
    lib.ns_step_params_unit(ctypes.byref(p))
    return p.replace(**kw) if kw else p


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def _c_f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def generate_ic(kind: int, n: int, seed: int, rmax: float = 100.0, lowl_frac: float = 0.0,
                lowl_scale: float = 1e-3):
    """Seeded Plummer (0) / Hernquist (1) / mixed (2) sphere in G = M = a = 1 units.

    Returns (r, v_r, ell, id) as float64, float64, float64, uint32 arrays.
    """
    r = np.empty(n, np.float64)
    v = np.empty(n, np.float64)
    ell = np.empty(n, np.float64)
    ids = np.empty(n, np.uint32)
    check(lib.ns_ic_generate(int(kind), int(n), ctypes.c_uint64(seed), float(rmax), float(lowl_frac),
                             float(lowl_scale), _ptr(r), _ptr(v), _ptr(ell), _ptr(ids)))
    return r, v, ell, ids


def chain_floor(device: int = 0, micro_steps: int = 1 << 22) -> float:
    """ns per micro-step of one wave running the late kernel's fast
    Levi-Civita micro-step alone on the device (the floor of its chains)."""
    out = ctypes.c_double()
    check(lib.ns_gpu_chain_floor(int(device), int(micro_steps), ctypes.byref(out)))
    return out.value


def philox4x32_10(ctr, key) -> tuple:
    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    lib.ns_philox4x32_10(c, k, o)
    return tuple(o)


class GpuContext:
    """One GPU's copy of the particle state plus its HIP stream."""

    def __init__(self, npts: int, device: int = 0):
        h = ctypes.c_void_p()
        check(lib.ns_gpu_ctx_create(int(npts), int(device), ctypes.byref(h)))
        self._h = h
        self.npts = int(npts)

    # -- lifetime -------------------------------------------------------
    def close(self) -> None:
        if getattr(self, "_h", None):
            lib.ns_gpu_ctx_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- options --------------------------------------------------------
    def set_option(self, option: int, value: int) -> None:
        check(lib.ns_gpu_set_option(self._h, int(option), int(value)))

    # -- data movement --------------------------------------------------
    def upload(self, r, v, ell, ids) -> None:
        r, v, ell = _c_f64(r), _c_f64(v), _c_f64(ell)
        ids = np.ascontiguousarray(ids, dtype=np.uint32)
        for a in (r, v, ell, ids):
            if a.shape != (self.npts,):
                raise ValueError(f"expected arrays of shape ({self.npts},), got {a.shape}")
        check(lib.ns_gpu_upload_soa(self._h, _ptr(r), _ptr(v), _ptr(ell), _ptr(ids)))

    def upload_particles(self, particles: np.ndarray) -> None:
        """Upload the reference's particles[5][npts] layout."""
        cols = [np.ascontiguousarray(particles[k], dtype=np.float64) for k in range(5)]
        arr = (ctypes.POINTER(ctypes.c_double) * 5)(
            *[c.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) for c in cols])
        check(lib.ns_gpu_upload(self._h, arr))

    def download(self):
        r = np.empty(self.npts, np.float64)
        v = np.empty(self.npts, np.float64)
        ell = np.empty(self.npts, np.float64)
        ids = np.empty(self.npts, np.uint32)
        check(lib.ns_gpu_download_soa(self._h, _ptr(r), _ptr(v), _ptr(ell), _ptr(ids)))
        return r, v, ell, ids

    def download_particles(self, particles: np.ndarray) -> None:
        """Write r, v_r, ell, id back into a (5, npts) float64 array in place."""
        if particles.dtype != np.float64 or not particles.flags.c_contiguous:
            raise ValueError("particles must be a C-contiguous float64 (5, npts) array")
        arr = (ctypes.POINTER(ctypes.c_double) * 5)(
            *[particles[k].ctypes.data_as(ctypes.POINTER(ctypes.c_double)) for k in range(5)])
        check(lib.ns_gpu_download(self._h, arr))

    # -- the loop -------------------------------------------------------
    def sort_particles(self) -> None:
        check(lib.ns_gpu_sort_particles(self._h))

    def permutation(self) -> np.ndarray:
        """The last reordering: position k now holds what sat at position perm[k]."""
        out = np.empty(self.npts, np.uint32)
        check(lib.ns_gpu_permutation(self._h, _ptr(out)))
        return out

    def permute_column(self, col) -> np.ndarray:
        """A caller-held float64 column (e.g. mu) reordered by the last reordering, on the device."""
        a = _c_f64(col)
        if a.shape != (self.npts,):
            raise ValueError(f"expected shape ({self.npts},)")
        out = np.empty(self.npts, np.float64)
        check(lib.ns_gpu_permute_column(self._h, _ptr(a), _ptr(out)))
        return out

    def inverse_map(self) -> np.ndarray:
        out = np.empty(self.npts, np.int32)
        check(lib.ns_gpu_inverse_map(self._h, _ptr(out)))
        return out

    def integrate(self, params: StepParams) -> None:
        check(lib.ns_gpu_integrate(self._h, ctypes.byref(params)))

    def step(self, params: StepParams) -> None:
        check(lib.ns_gpu_step(self._h, ctypes.byref(params)))

    def run(self, params: StepParams, nsteps: int) -> None:
        check(lib.ns_gpu_run(self._h, ctypes.byref(params), int(nsteps)))

    def graph_prepare(self, params: StepParams, nsteps: int) -> None:
        """Capture ahead the graph that run(params, nsteps) replays (NS_OPT_GRAPH)."""
        check(lib.ns_gpu_graph_prepare(self._h, ctypes.byref(params), int(nsteps)))

    def sync(self) -> None:
        check(lib.ns_gpu_sync(self._h))

    def cap_hits(self, reset: bool = False) -> int:
        """Particle-steps stopped by the work cap (NS_OPT_BUDGET)."""
        n = ctypes.c_uint64()
        check(lib.ns_gpu_cap_hits(self._h, ctypes.byref(n), int(reset)))
        return n.value

    def levels(self) -> np.ndarray:
        out = np.empty(self.npts, np.uint8)
        check(lib.ns_gpu_levels(self._h, _ptr(out)))
        return out

    def energy(self, params: StepParams):
        k = ctypes.c_double()
        w = ctypes.c_double()
        check(lib.ns_gpu_energy(self._h, ctypes.byref(params), ctypes.byref(k), ctypes.byref(w)))
        return k.value, w.value

    def kernel_stats(self) -> dict:
        n = ctypes.c_int()
        from ._lib import NS_MAX_KERNEL_STATS as cap
        names = (ctypes.c_char_p * cap)()
        launches = (ctypes.c_int64 * cap)()
        ms = (ctypes.c_double * cap)()
        check(lib.ns_gpu_kernel_stats(self._h, ctypes.byref(n), names, launches, ms))
        return {names[i].decode(): (launches[i], ms[i]) for i in range(n.value) if launches[i]}

    def reset_kernel_stats(self) -> None:
        check(lib.ns_gpu_kernel_stats_reset(self._h))

    def set_masses(self, mass) -> None:
        """Per-particle masses in the current state order (None: equal masses)."""
        if mass is None:
            check(lib.ns_gpu_set_masses(self._h, None))
            return
        m = np.ascontiguousarray(mass, dtype=np.float64)
        if m.shape != (self.npts,):
            raise ValueError(f"mass must have shape ({self.npts},)")
        check(lib.ns_gpu_set_masses(self._h, m.ctypes.data))

    def set_halos(self, counts, halo_mass=None) -> None:
        """Ensemble: consecutive runs of counts[h] particles are independent
        halos of mass halo_mass[h] (default 1 each); None ends the ensemble."""
        if counts is None:
            check(lib.ns_gpu_set_halos(self._h, 0, None, None))
            return
        c = np.ascontiguousarray(counts, dtype=np.int64)
        m = np.ones(c.size) if halo_mass is None else np.ascontiguousarray(halo_mass, dtype=np.float64)
        if m.shape != c.shape:
            raise ValueError("halo_mass must match counts")
        check(lib.ns_gpu_set_halos(self._h, int(c.size), c.ctypes.data, m.ctypes.data))

    def autotune(self, p: StepParams, steps: int = 5) -> dict:
        """Time each launch configuration on the resident state, keep the
        fastest (state restored); returns the best and every trial."""
        from ._lib import TuneResult
        res = TuneResult()
        check(lib.ns_gpu_autotune(self._h, ctypes.byref(p), int(steps), ctypes.byref(res)))
        return {"best": res.best.as_dict(), "tried": [res.tried[i].as_dict() for i in range(res.ntried)]}

    def generate_ic(self, kind: int, seed: int, rmax: float = 100.0, lowl_frac: float = 0.0,
                    lowl_scale: float = 1e-3) -> None:
        """Draw the seeded initial conditions of generate_ic on the device, as the state."""
        check(lib.ns_gpu_generate_ic(self._h, int(kind), ctypes.c_uint64(seed), float(rmax), float(lowl_frac),
                                     float(lowl_scale)))

    def upload_ic(self, records, conv: float = 1.0) -> None:
        """Upload initial-conditions records (npts x 5: r, |v|, ell, id, mu) with
        the reference's conversion v_r = |v| mu conv, ell = ell conv."""
        rec = np.ascontiguousarray(records, np.float64)
        if rec.shape != (self.npts, 5):
            raise ValueError(f"records must be ({self.npts}, 5)")
        check(lib.ns_gpu_upload_ic(self._h, rec.ctypes.data, float(conv)))

    def potential(self, params: StepParams):
        """Dynamic potential, method A (nsphere.c:7698-7738), over the current
        radii in the current order taken as rank order: (psi per rank, psi(0))."""
        psi = np.empty(self.npts, np.float64)
        psi0 = ctypes.c_double()
        check(lib.ns_gpu_potential(self._h, ctypes.byref(params), psi.ctypes.data, ctypes.byref(psi0)))
        return psi, psi0.value

    def density(self, params: StepParams):
        """The reference's snapshot density estimate (nsphere.c:7236-7686) at
        the current radii in the current order taken as rank order: (density
        per rank, info dict)."""
        from ._lib import DensityInfo
        out = np.empty(self.npts, np.float64)
        info = DensityInfo()
        check(lib.ns_gpu_density(self._h, ctypes.byref(params), out.ctypes.data, ctypes.byref(info)))
        return out, {k: getattr(info, k) for k, _ in DensityInfo._fields_}

    def rank_files(self, params: StepParams, unsorted_path=None, sorted_path=None) -> None:
        """The reference's Rank_Mass_Rad_VRad files of the current state."""
        enc = (lambda p: None if p is None else str(p).encode())  # noqa: E731
        check(lib.ns_gpu_rank_files(self._h, ctypes.byref(params), enc(unsorted_path), enc(sorted_path)))

    def snapshot_records(self) -> np.ndarray:
        """all_particle_data records of the current state in id order
        (rank, float32 r, v_r, ell), as nsphere.c:6456-6471 fills them."""
        out = np.empty(self.npts, APD_RECORD)
        check(lib.ns_gpu_snapshot_records(self._h, out.ctypes.data))
        return out

    def gather_ids(self, ids):
        """(r, v_r, ell, rank) of the particles with these original ids."""
        ids = np.ascontiguousarray(ids, np.uint32)
        k = ids.size
        r, v, ell = (np.empty(k) for _ in range(3))
        rank = np.empty(k, np.int32)
        check(lib.ns_gpu_gather_ids(self._h, ids.ctypes.data, k, r.ctypes.data, v.ctypes.data, ell.ctypes.data,
                                    rank.ctypes.data))
        return r, v, ell, rank

    def set_mlp(self, mlp) -> None:
        """Learned binning with a trained Mlp (None: off)."""
        check(lib.ns_gpu_set_mlp(self._h, None if mlp is None else mlp._h))
        self._mlp = mlp  # keep the network alive while the context uses it

    def features(self, p: StepParams) -> np.ndarray:
        """The network's raw inputs for the current state (current order = rank order)."""
        out = np.empty((self.npts, 5), dtype=np.float32)
        check(lib.ns_gpu_features(self._h, ctypes.byref(p), out.ctypes.data))
        return out

    def enclosed_mass(self) -> np.ndarray:
        """M_enc of the last scan: [i] below rank i, [n] total, [n + 1] rank-0 mass."""
        out = np.empty(self.npts + 2, dtype=np.float64)
        check(lib.ns_gpu_enclosed_mass(self._h, out.ctypes.data))
        return out

    def force_evals(self) -> dict:
        """Force evaluations spent while profiling, per integrate kernel."""
        out = (ctypes.c_uint64 * 3)()
        check(lib.ns_gpu_force_evals(self._h, out))
        return {"bulk_or_plain": out[0], "deep": out[1], "late": out[2]}

    def chain_steps(self) -> int:
        """Sum over the late kernel's launches while profiling of each
        launch's longest Levi-Civita chain (sequential micro-steps)."""
        out = ctypes.c_uint64()
        check(lib.ns_gpu_chain_steps(self._h, ctypes.byref(out)))
        return out.value

    def kernel_families(self) -> list:
        """Kernel family names in id order (bit k of NS_OPT_PROFILE_KERNELS)."""
        n = ctypes.c_int()
        from ._lib import NS_MAX_KERNEL_STATS as cap
        names = (ctypes.c_char_p * cap)()
        check(lib.ns_gpu_kernel_stats(self._h, ctypes.byref(n), names, None, None))
        return [names[i].decode() for i in range(n.value)]

    def profile_kernels(self, names) -> None:
        """Time only these kernel families while profiling (None: all)."""
        if names is None:
            mask = 0xFFFFFFFF
        else:
            fam = self.kernel_families()
            unknown = set(names) - set(fam)
            if unknown:
                raise ValueError(f"unknown kernel families {sorted(unknown)}")
            mask = sum(1 << fam.index(x) for x in names)
        self.set_option(NS_OPT_PROFILE_KERNELS, mask)


APD_RECORD = np.dtype([("rank", "<i4"), ("r", "<f4"), ("v", "<f4"), ("ell", "<f4")])


class ApdWriter:
    """Appends snapshots of a context's state to all_particle_data<suffix>.dat
    in the reference's format (16-byte records, snapshot-major, whole blocks
    of block_size snapshots kept at close)."""

    def __init__(self, ctx: "GpuContext", path: str, block_size: int = 100):
        h = ctypes.c_void_p()
        check(lib.ns_apd_writer_open(ctx._h, str(path).encode(), int(block_size), ctypes.byref(h)))
        self._h, self._ctx = h, ctx

    def write(self) -> None:
        check(lib.ns_apd_writer_write(self._h))

    def close(self) -> int:
        """Finish the file; returns the number of snapshots kept."""
        if not self._h:
            return 0
        kept = ctypes.c_int64()
        h, self._h = self._h, None
        check(lib.ns_apd_writer_close(h, ctypes.byref(kept)))
        return kept.value

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def write_ic_file(path: str, records) -> None:
    """The reference's initial-conditions file (nsphere.c:2386-2416)."""
    rec = np.ascontiguousarray(records, np.float64)
    if rec.ndim != 2 or rec.shape[1] != 5:
        raise ValueError("records must be (npts, 5)")
    check(lib.ns_ic_file_write(str(path).encode(), rec.shape[0], rec.ctypes.data))


def read_ic_file(path: str) -> np.ndarray:
    """Records (npts x 5) of an initial-conditions file (nsphere.c:2445-2487)."""
    n = ctypes.c_int32()
    check(lib.ns_ic_file_read(str(path).encode(), ctypes.byref(n), None))
    rec = np.empty((n.value, 5), np.float64)
    check(lib.ns_ic_file_read(str(path).encode(), ctypes.byref(n), rec.ctypes.data))
    return rec


def read_all_particle_data(path: str, npts: int, snap: int) -> np.ndarray:
    """Snapshot `snap` of an all_particle_data file (offset snap * npts * 16,
    as the reference's reader computes it, nsphere.c:1380-1385)."""
    return np.fromfile(path, APD_RECORD, count=npts, offset=int(snap) * int(npts) * APD_RECORD.itemsize)


class Mlp:
    """The 5 -> 64 -> 64 -> 8 level predictor (BF16 MFMA, FP32 accumulation)."""

    NPARAMS = 6288

    def __init__(self, device: int = 0, seed: int = 1):
        h = ctypes.c_void_p()
        check(lib.ns_mlp_create(int(device), int(seed), ctypes.byref(h)))
        self._h = h

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib.ns_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _xy(X, y=None):
        X = np.ascontiguousarray(X, dtype=np.float32)
        if X.ndim != 2 or X.shape[1] != 5:
            raise ValueError("X must be (n, 5)")
        if y is None:
            return X, None
        y = np.ascontiguousarray(y, dtype=np.uint8)
        if y.shape != (X.shape[0],):
            raise ValueError("y must be (n,)")
        return X, y

    def train(self, X, y, epochs: int = 2, batch: int = 16384, lr: float = 3e-3) -> float:
        X, y = self._xy(X, y)
        loss = ctypes.c_float()
        check(lib.ns_mlp_train(self._h, X.ctypes.data, y.ctypes.data, X.shape[0], int(epochs), int(batch),
                               float(lr), ctypes.byref(loss)))
        return loss.value

    def gradient(self, X, y):
        X, y = self._xy(X, y)
        g = np.empty(self.NPARAMS, dtype=np.float32)
        loss = ctypes.c_float()
        check(lib.ns_mlp_gradient(self._h, X.ctypes.data, y.ctypes.data, X.shape[0], g.ctypes.data,
                                  ctypes.byref(loss)))
        return g, loss.value

    def predict(self, X) -> np.ndarray:
        X, _ = self._xy(X)
        out = np.empty(X.shape[0], dtype=np.uint8)
        check(lib.ns_mlp_predict(self._h, X.ctypes.data, X.shape[0], out.ctypes.data))
        return out

    def params(self):
        P = np.empty(self.NPARAMS, np.float32)
        norm = np.empty(10, np.float32)
        cw = np.empty(8, np.float32)
        check(lib.ns_mlp_get_params(self._h, P.ctypes.data, norm.ctypes.data, cw.ctypes.data))
        return P, norm, cw

    def set_params(self, P=None, norm=None, cw=None) -> None:
        arrs = [None if a is None else np.ascontiguousarray(a, np.float32) for a in (P, norm, cw)]
        check(lib.ns_mlp_set_params(self._h, *[None if a is None else a.ctypes.data for a in arrs]))


def level_samples(ctx: GpuContext, params: StepParams, steps: int, every: int = 1):
    """Training pairs for the level predictor, logged by the adaptive
    integrator itself: before every `every`-th of `steps` steps the state is
    put in rank order and its features read; the step's level records then
    label them.  Levi-Civita particles and particles that did not converge
    are left out; levels above 7 count as 7.  Returns (X float32 (m, 5),
    y uint8 (m,))."""
    xs, ys = [], []
    for s in range(int(steps)):
        if s % every == 0:
            ctx.sort_particles()
            X = ctx.features(params)
            ctx.step(params)
            lv = ctx.levels()
            keep = (lv & (NS_LVL_LC | NS_LVL_NOCONV)) == 0
            xs.append(X[keep])
            ys.append(np.minimum(lv[keep] & NS_LVL_MASK, 7).astype(np.uint8))
        else:
            ctx.step(params)
    return np.concatenate(xs), np.concatenate(ys)


# -- reference-interface mirror -------------------------------------------

_ctx_cache: dict = {}


def _ctx_for(npts: int) -> GpuContext:
    ctx = _ctx_cache.get(npts)
    if ctx is None:
        _ctx_cache.clear()
        ctx = GpuContext(npts)
        _ctx_cache[npts] = ctx
    return ctx


def sort_particles(particles: np.ndarray, npts: int) -> None:
    """sort_particles(particles, npts), nsphere.c:8662-8665: stable radial sort
    of all columns of a (5, npts) float64 array, in place, on the GPU."""
    ctx = _ctx_for(npts)
    ctx.upload_particles(particles)
    ctx.sort_particles()
    ctx.download_particles(particles)
    particles[4] = ctx.permute_column(particles[4])


def method5_step(particles: np.ndarray, npts: int, params: StepParams,
                 inverse_map: np.ndarray | None = None) -> np.ndarray:
    """One method-5 step (nsphere.c:5775-5838) on a (5, npts) array in place.

    Returns the level records (rank order).  When ``inverse_map`` is given it
    receives inverse_map[id] = rank as in nsphere.c:5789-5794.
    """
    ctx = _ctx_for(npts)
    ctx.upload_particles(particles)
    ctx.step(params)
    ctx.sync()
    if inverse_map is not None:
        inverse_map[:] = ctx.inverse_map()
    ctx.download_particles(particles)
    particles[4] = ctx.permute_column(particles[4])
    return ctx.levels()
