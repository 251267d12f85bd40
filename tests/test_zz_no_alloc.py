"""ns_gpu_sort_particles_noalloc: the allocation-free radial sort.

It has no counterpart in the reference; its contract is that of
ns_gpu_sort_particles (the stable radial sort of sort_particles,
nsphere.c:8605-8665) on caller-owned device arrays, with its workspace in a
static device array.  The GPU test checks it bit for bit against the context
sort on two calls in a row, the second smaller, so that the second runs on
the workspace the first left behind.  The caller's device arrays come from
the HIP runtime the library itself loaded.
"""
import ctypes

import numpy as np
import pytest

H2D, D2H = 1, 2


def _hip():
    import nsphere_amd  # noqa: F401  (loads libnsphere_gpu.so and with it libamdhip64.so.7)
    hip = ctypes.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return hip


def _ok(rc, what):
    assert rc == 0, f"{what} returned {rc}"


def test_noalloc_rejects_sizes_beyond_its_workspace():
    """Refused on the host before any device work (runs without a GPU)."""
    from nsphere_amd._lib import NS_ERR_ARG, lib
    cap = 1 << 20  # NS_SORT_NOALLOC_MAX
    for n in (cap + 1, 1 << 40, -1):
        rc = lib.ns_gpu_sort_particles_noalloc(None, None, None, None, None, None, None, None, n, None)
        assert rc == NS_ERR_ARG
        assert b"NS_SORT_NOALLOC_MAX" in lib.ns_gpu_last_error()
    assert lib.ns_gpu_sort_particles_noalloc(None, None, None, None, None, None, None, None, 0, None) == 0


def _case(n, seed):
    import nsphere_amd as ns
    r, v, ell, ids = ns.generate_ic(0, n, seed, 100.0, 0.05, 1e-3)
    r[::9] = r[5]          # exact ties, kept in input order
    r[1] = 0.0
    r[2] = -0.0            # ties with +0
    r[3] = np.inf
    return r, v, ell, ids


@pytest.mark.gpu
def test_noalloc_sort_equals_context_sort():
    import nsphere_amd as ns
    from nsphere_amd._lib import check, lib
    hip = _hip()
    for n, seed in ((400_000, 11), (70_001, 12)):  # the second call smaller
        print(f"[noalloc] n={n} seed={seed}", flush=True)
        host = _case(n, seed)
        with ns.GpuContext(n) as ctx:
            ctx.upload(*host)
            ctx.sort_particles()
            want = ctx.download()
        sizes = [a.nbytes for a in host]
        bufs = []
        try:
            for nb in sizes * 2:
                p = ctypes.c_void_p()
                _ok(hip.hipMalloc(ctypes.byref(p), nb), "hipMalloc")
                bufs.append(p)
            for p, a in zip(bufs[:4], host):
                a = np.ascontiguousarray(a)
                _ok(hip.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D), "hipMemcpy")
            check(lib.ns_gpu_sort_particles_noalloc(*bufs[:4], *bufs[4:], n, None))
            _ok(hip.hipDeviceSynchronize(), "hipDeviceSynchronize")
            got = [np.empty_like(a) for a in host]
            for p, a in zip(bufs[4:], got):
                _ok(hip.hipMemcpy(a.ctypes.data, p, a.nbytes, D2H), "hipMemcpy")
        finally:
            for p in bufs:
                hip.hipFree(p)
        for g, w, what in zip(got, want, ("r", "v", "ell", "id")):
            gb = g.view(np.uint64) if g.dtype == np.float64 else g
            wb = w.view(np.uint64) if w.dtype == np.float64 else w
            bad = np.flatnonzero(gb != wb)
            assert bad.size == 0, f"n={n}: {what} differs at {bad.size} positions, first {bad[:5]}"
