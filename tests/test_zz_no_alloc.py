"""GPU test of drape_prolong_state_noalloc, the allocation-free variant of drape_prolong_state, through the
C ABI: on two calls in a row (the second with a smaller grid and other colliders) its fine positions are
bit-identical to drape_prolong_state's and its moved / infeasible counts equal; host pointers and too many
colliders are refused before any launch. The caller's device arrays are allocated here through the HIP
runtime (the entry point itself allocates nothing)."""
import ctypes as C

import numpy as np
import pytest

import drape_amd as D

pytestmark = pytest.mark.gpu

f32p = C.POINTER(C.c_float)
f64p = C.POINTER(C.c_double)
i32p = C.POINTER(C.c_int32)


def hip():
    h = C.CDLL("libamdhip64.so")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


H2D, D2H = 1, 2


def entry():
    L = D.lib()
    L.drape_prolong_state_noalloc.argtypes = [C.c_int32, C.c_void_p, f64p, C.c_int32, C.c_void_p, i32p, i32p,
                                              C.c_void_p]
    return L


def coarse_state(n, seed):
    """A coarse sheet state partly pressed into a sphere of radius 0.15 m at (0.25, 0.25, 0.15): some fine
    vertices land within dhat (moved), a few inside (infeasible)."""
    rng = np.random.default_rng(seed)
    g = np.linspace(0.0, 0.5, n)
    x, y = np.meshgrid(g, g)
    r2 = (x - 0.25) ** 2 + (y - 0.25) ** 2
    z = np.where(r2 < 0.15 ** 2, 0.15 + np.sqrt(np.maximum(0.15 ** 2 - r2, 0.0)), 0.15)
    z = z + rng.uniform(-2e-3, 2e-3, z.shape)
    return np.stack([x, y, z], axis=-1).reshape(-1).astype(np.float32)


def colliders(kinds):
    rows = [[2.0, 0.25, 0.25, 0.15, 0.15, 0, 0, 0] if k == "sphere" else [1.0, 0.0, 0.0, 1.0, 0.01, 0, 0, 0]
            for k in kinds]
    return np.array([len(rows), 1e-3, 1.0, 0.0] + [v for r in rows for v in r], np.float64)


def reference(n, coarse, cols):
    nf = 2 * n - 1
    fine = np.zeros(3 * nf * nf, np.float32)
    moved, bad = C.c_int32(), C.c_int32()
    D._check(D.lib().drape_prolong_state(n, coarse.ctypes.data_as(f32p), cols.ctypes.data_as(f64p), cols.size,
                                          fine.ctypes.data_as(f32p), C.byref(moved), C.byref(bad)))
    return fine, moved.value, bad.value


def no_alloc(h, L, n, coarse, cols):
    nf = 2 * n - 1
    dc, df = C.c_void_p(), C.c_void_p()
    assert h.hipMalloc(C.byref(dc), coarse.nbytes) == 0
    assert h.hipMalloc(C.byref(df), 3 * nf * nf * 4) == 0
    try:
        assert h.hipMemcpy(dc, coarse.ctypes.data, coarse.nbytes, H2D) == 0
        moved, bad = C.c_int32(-1), C.c_int32(-1)
        D._check(L.drape_prolong_state_noalloc(n, dc, cols.ctypes.data_as(f64p), cols.size, df, C.byref(moved),
                                               C.byref(bad), None))
        fine = np.zeros(3 * nf * nf, np.float32)
        assert h.hipMemcpy(fine.ctypes.data, df, fine.nbytes, D2H) == 0
        return fine, moved.value, bad.value
    finally:
        h.hipFree(dc)
        h.hipFree(df)


def test_no_alloc_prolong_matches_the_allocating_entry_point():
    h, L = hip(), entry()
    for n, kinds, seed in ((257, ("sphere",), 1), (33, ("plane", "sphere"), 2)):
        coarse = coarse_state(n, seed)
        cols = colliders(kinds)
        print(f"[gpu] no-alloc prolong: coarse {n}x{n}, colliders {kinds}", flush=True)
        want, wm, wb = reference(n, coarse, cols)
        got, gm, gb = no_alloc(h, L, n, coarse, cols)
        print(f"[gpu]   moved {gm} (allocating: {wm}), infeasible {gb} ({wb})", flush=True)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
        assert (gm, gb) == (wm, wb)
        assert gm > 0


def test_no_alloc_prolong_refuses_host_arrays_and_too_many_colliders():
    h, L = hip(), entry()
    n = 9
    coarse = coarse_state(n, 3)
    nf = 2 * n - 1
    fine = np.zeros(3 * nf * nf, np.float32)
    cols = colliders(("sphere",))
    rc = L.drape_prolong_state_noalloc(n, coarse.ctypes.data, cols.ctypes.data_as(f64p), cols.size,
                                       fine.ctypes.data, None, None, None)
    assert rc == -1 and b"device memory" in L.drape_last_error()
    dc, df = C.c_void_p(), C.c_void_p()
    assert h.hipMalloc(C.byref(dc), coarse.nbytes) == 0
    assert h.hipMalloc(C.byref(df), fine.nbytes) == 0
    try:
        many = colliders(("sphere",) * 17)
        rc = L.drape_prolong_state_noalloc(n, dc, many.ctypes.data_as(f64p), many.size, df, None, None, None)
        assert rc == -1 and b"at most 16" in L.drape_last_error()
        rc = L.drape_prolong_state_noalloc(8194, dc, cols.ctypes.data_as(f64p), cols.size, df, None, None, None)
        assert rc == -1
    finally:
        h.hipFree(dc)
        h.hipFree(df)
