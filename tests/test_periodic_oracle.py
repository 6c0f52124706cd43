"""CPU pins of the oracle's periodic ISPH box (C3, Taylor-Green).

No build of the reference is periodic (its KERNEL_periodic_3D,
functions_BC.cuh:1522, copies velocities and is never called), so the
periodic wrap is pinned by what a periodic lattice must satisfy:
- translation invariance: on a regular lattice every particle has the same
  neighbour count (equal to an O(N^2) minimum-image count), the same
  Shepard sum and the same KGC matrix;
- the Taylor-Green field is divergence-free: the first PPE right-hand
  side vanishes to rounding on the lattice, and the pressure the projection
  builds tends to the analytic Taylor-Green pressure
  rho/16 (cos 2x + cos 2y)(cos 2z + 2) as the lattice refines;
- the zero-mean solve leaves a zero-mean pressure that meets rtol;
- a uniform flow translates the lattice rigidly across the boundary.
"""
import math

import numpy as np
import pytest

from fluidized_bed_master_amd import abi, configs
from oracle.oracle import Oracle, knl_eval

L = 2.0 * math.pi


def _opts(rtol=1e-10):
    return abi.ppe_opts(method=abi.PPE_BICGSTAB, rtol=rtol, zero_mean=1, max_iter=4000)


@pytest.mark.parametrize("jitter", [0.0, 0.2])
def test_periodic_neighbour_counts(jitter):
    p, parts = configs.taylor_green(12, jitter=jitter)
    o = Oracle(p, parts)
    o.nnps()
    d = o.domain
    assert (d.NI, d.NJ, d.NK) == (4, 4, 4)
    c = o.get_nnps(abi.SPH)["nbr_count"]
    b = o.brute_neighbor_counts()
    assert np.array_equal(c, b)
    if jitter == 0:
        assert np.all(c == c[0]) and c[0] == 81


def test_periodic_lattice_is_translation_invariant():
    p, parts = configs.taylor_green(12, noise=0.0)
    p.kgc_solve = 1
    o = Oracle(p, parts)
    o.nnps()
    o.prep()
    p1, _, p3 = o.get_particles(abi.SPH)
    dx = L / 12
    h = float(p1["h"][0])
    a, b, c = np.meshgrid(np.arange(-3, 4), np.arange(-3, 4), np.arange(-3, 4))
    rr = np.sqrt(a * a + b * b + c * c).ravel() * dx
    w, _ = knl_eval(p.kernel_type, 3, h, rr[rr < p.kappa * h])
    lattice = dx ** 3 * w.sum()
    assert np.allclose(p1["flt_s"], lattice, rtol=1e-12)
    assert abs(lattice - 1.0) < 0.02
    C = p3["Cm"]
    assert np.allclose(C, C[0][None], rtol=0, atol=1e-12 * np.abs(C[0]).max())


def test_taylor_green_rhs_vanishes_and_pressure_tends_to_analytic():
    # the first PPE of a divergence-free field on the lattice: b ~ rounding
    p, parts = configs.taylor_green(16, noise=0.0)
    o = Oracle(p, parts)
    st = o.step(_opts())
    _, rhs, rows = o.get_ppe_system()
    dx = L / 16
    scale = 1000.0 / p.dt * 1.0 / dx
    assert np.max(np.abs(rhs[rows.astype(bool)])) < 1e-10 * scale
    err = {}
    for n in (16, 24):
        p, parts = configs.taylor_green(n, noise=0.0)
        o = Oracle(p, parts)
        for _ in range(8):
            st = o.step(_opts())
            assert st["converged"]
        p1 = o.get_particles(abi.SPH)[0]
        pa = 1000.0 / 16.0 * (np.cos(2 * p1["x"]) + np.cos(2 * p1["y"])) * (np.cos(2 * p1["z"]) + 2.0)
        assert np.corrcoef(p1["pres"], pa)[0, 1] > 0.995
        err[n] = np.linalg.norm(p1["pres"] - pa) / np.linalg.norm(pa)
    assert err[24] < 0.75 * err[16] and err[24] < 0.2, err


def test_zero_mean_solution():
    p, parts = configs.taylor_green(12, noise=0.02, jitter=0.1)
    o = Oracle(p, parts)
    for _ in range(3):
        st = o.step(_opts(1e-8))
    assert st["converged"] and st["r_rel"] <= 1e-8
    pr = o.get_particles(abi.SPH)[0]["pres"]
    assert np.max(np.abs(pr)) > 1.0
    assert abs(pr.mean()) < 1e-12 * np.max(np.abs(pr))


def test_uniform_flow_translates_across_the_boundary():
    p, parts = configs.taylor_green(10, noise=0.0)
    p.fv_solve = 0
    U = np.array([0.9, -0.6, 0.4])
    parts["ux"], parts["uy"], parts["uz"] = U
    p.dt = 0.05  # the lattice moves a quarter of a spacing per step
    x0 = np.stack([parts["x"], parts["y"], parts["z"]], 1)
    o = Oracle(p, parts)
    n = 12
    for _ in range(n):
        o.step(_opts())
    p1 = o.get_particles(abi.SPH)[0]
    got = np.stack([p1["x"], p1["y"], p1["z"]], 1)
    assert np.all(got >= 0) and np.all(got < L)
    assert np.max(np.abs(p1["pres"])) < 1e-8
    want = np.mod(x0 + U[None] * p.dt * n, L)
    # match each moved particle to the translated lattice point
    key = lambda a: np.round(a / (L / 10) * 4).astype(np.int64) % 40
    ka, kb = key(got), key(want)
    ia = np.lexsort(ka.T)
    ib = np.lexsort(kb.T)
    assert np.array_equal(ka[ia], kb[ib])
    assert np.allclose(got[ia], want[ib], rtol=0, atol=1e-12)
    c = o.get_nnps(abi.SPH)["nbr_count"]
    assert np.all(c == c[0])
