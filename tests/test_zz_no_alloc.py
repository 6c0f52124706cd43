"""The allocation-free kernel evaluation against the allocating one, through
the C ABI: identical W and dW/dr on two calls in a row (the second smaller,
so stale values left in the shared device scratch by the first call would
show), for every kernel type in 2D and 3D, and the capacity limit."""
import ctypes as C

import numpy as np
import pytest

from fluidized_bed_master_amd import abi
from fluidized_bed_master_amd.abi import ptr

pytestmark = pytest.mark.gpu

KERNELS = [abi.KNL_GAUSSIAN, abi.KNL_QUINTIC, abi.KNL_QUARTIC, abi.KNL_WENDLAND2, abi.KNL_WENDLAND4,
           abi.KNL_WENDLAND6]


def _eval(fn, kt, dim, h, r):
    r = np.ascontiguousarray(r, np.float64)
    w = np.full_like(r, np.nan)
    dw = np.full_like(r, np.nan)
    rc = fn(kt, dim, h, ptr(r), len(r), ptr(w), ptr(dw))
    return rc, w, dw


def test_noalloc_matches_allocating_entry_point():
    lib = abi.load_library()
    h = 1.3e-3
    rng = np.random.default_rng(7)
    for n in (50000, 1237):  # second call smaller than the first
        r = rng.uniform(0.0, 2.2 * h, n)
        for kt in KERNELS:
            for dim in (2, 3):
                rc0, w0, dw0 = _eval(lib.sophia_knl_eval, kt, dim, h, r)
                rc1, w1, dw1 = _eval(lib.sophia_knl_eval_noalloc, kt, dim, h, r)
                assert rc0 == 0 and rc1 == 0, (rc0, rc1, abi.last_error())
                assert np.array_equal(w0, w1), (n, kt, dim)
                assert np.array_equal(dw0, dw1), (n, kt, dim)


def test_noalloc_rejects_more_than_its_scratch():
    lib = abi.load_library()
    r = np.zeros(abi.KNL_NOALLOC_MAX + 1)
    w = np.zeros_like(r)
    dw = np.zeros_like(r)
    rc = lib.sophia_knl_eval_noalloc(abi.KNL_WENDLAND2, 3, 1e-3, ptr(r), len(r), ptr(w), ptr(dw))
    assert rc == -6, rc  # SOPHIA_E_CAPACITY, before any device work
    assert "scratch capacity" in abi.last_error()
    # and the largest size it accepts still works
    r = np.linspace(0.0, 2e-3, abi.KNL_NOALLOC_MAX)
    rc0, w0, dw0 = _eval(lib.sophia_knl_eval, abi.KNL_WENDLAND2, 3, 1e-3, r)
    rc1, w1, dw1 = _eval(lib.sophia_knl_eval_noalloc, abi.KNL_WENDLAND2, 3, 1e-3, r)
    assert rc0 == 0 and rc1 == 0
    assert np.array_equal(w0, w1) and np.array_equal(dw0, dw1)
