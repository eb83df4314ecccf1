"""ust_k_apply_K_f64_noalloc, the K u entry point that allocates nothing (its workspace is a static device
array), against ust_k_apply_K_f64 with a caller workspace: the same kernels in the same order, so bitwise
equal, on two calls in a row, the second on a smaller grid."""
import ctypes as C

import numpy as np
import pytest

from tests.test_gpu_operators import _Dev, _grid

pytestmark = pytest.mark.gpu


def _compare(lib, nel, seed):
    g = _grid(len(nel), nel)
    dim = len(nel)
    n = int(np.prod([k + 1 for k in nel])) * dim
    wl = lib.ust_k_weights_len(C.byref(g))
    rng = np.random.default_rng(seed)
    W = _Dev(lib, np.zeros(wl))
    assert lib.ust_k_fill_hash_f64(W.p, wl, seed, None) == 0
    dm = _Dev(lib, (rng.random(n) < 0.1).astype(np.uint8))
    du = _Dev(lib, rng.normal(size=n))
    work = _Dev(lib, np.zeros(lib.ust_k_apply_work_len(C.byref(g))))
    ya, yb = _Dev(lib, np.zeros(n)), _Dev(lib, np.full(n, np.nan))
    assert lib.ust_k_apply_K_f64(C.byref(g), W.p, dm.p, du.p, ya.p, work.p, None) == 0
    assert lib.ust_k_apply_K_f64_noalloc(C.byref(g), W.p, dm.p, du.p, yb.p, None) == 0
    a, b = ya.get(), yb.get()
    for x in (W, dm, du, work, ya, yb):
        x.free()
    return a, b


def test_noalloc_apply_matches_workspace_apply():
    from earthbyte_underworld_amd import _lib as L
    lib = L.load()
    assert lib.ust_k_apply_noalloc_max_work() == 1 << 22
    # the column kernel on the 128x128x64 grid, then the tile kernel on a smaller 3D grid, then a 2D grid
    for nel, seed in (((128, 128, 64), 3), ((33, 20, 7), 4), ((40, 24), 5)):
        a, b = _compare(lib, nel, seed)
        assert np.array_equal(a, b), nel
        assert np.isfinite(b).all()


def test_noalloc_apply_rejects_grid_beyond_its_workspace():
    from earthbyte_underworld_amd import _lib as L
    lib = L.load()
    g = _grid(3, (512, 512, 256))
    assert lib.ust_k_apply_work_len(C.byref(g)) > lib.ust_k_apply_noalloc_max_work()
    dummy = C.c_void_p(16)  # never dereferenced: the size check comes first
    assert lib.ust_k_apply_K_f64_noalloc(C.byref(g), dummy, dummy, dummy, dummy, None) == -1
    assert b"static workspace" in lib.ust_dev_error_string()
