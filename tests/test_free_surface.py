"""Free-surface rows of the PPE on the oracle (CPU).

KERNEL_clc_surface_detect3D (functions_PST.cuh:263-320) labels a particle as
surface when the mean unit vector to its neighbours is longer than 0.3 or it
has fewer than 10 neighbours; the reference's PPE would then set p = 0 there
(functions_PPE2.cuh:533-534).  Both calls are commented out in the reference
(ISPH_Calc_seperate.cuh:289-296), which leaves the pressure level of a
converged PPE undetermined up to the wall average; the pins below show the
labels' geometry on a lattice and that a column with the Dirichlet surface
stays at rest under the converged solve while the same column without it
drifts.
"""
import numpy as np
import pytest

from fluidized_bed_master_amd import abi, configs
from oracle import oracle
from oracle.oracle import Oracle


def _labels(fs, nx=8, nz=10):
    p, parts = configs.isph_dem_column(nx, nx, nz, 1e-3, seed=5, dt=2e-6)
    p.free_surface = fs
    o = Oracle(p, parts)
    o.nnps(); o.prep(); o.interaction(); o.dem_interaction(); o.sph_dem_coupling(); o.predict()
    o.ppe(abi.ppe_opts(method=abi.PPE_BICGSTAB, rtol=1e-8))
    a, _, a3 = o.get_particles(abi.SPH)
    f = a["p_type"] == abi.PT_FLUID
    ijk = np.round(np.stack([a[c][f] for c in ("x", "y", "z")]) / 1e-3 - 0.5).astype(int)
    return a[f], a3["lbl_surf"][f], ijk


def test_surface_labels_on_a_lattice():
    """Walls counted (free_surface = 1): on a uniform lattice only the top
    layer can be surface, and its particles at least kappa*h from the side
    walls all are (a half-space: the mean unit vector points down with
    length well above 0.3); nothing below the top layer is labelled.
    Same-type counting (free_surface = 2, the reference's rule) also labels
    the bottom layer and every fluid particle next to a side wall, because
    walls do not count as neighbours there.  Labelled particles take
    p = p_ref = 0."""
    nx, nz = 8, 10
    a, lab, (i, j, k) = _labels(1, nx, nz)
    assert lab[k < nz - 1].sum() == 0
    far = (k == nz - 1) & (i >= 3) & (i <= nx - 4) & (j >= 3) & (j <= nx - 4)
    assert far.sum() > 0 and np.all(lab[far] == 1)
    assert np.all(a["pres"][lab == 1] == 0.0)
    a2, lab2, (i2, j2, k2) = _labels(2, nx, nz)
    assert np.all(lab2[k2 == 0] == 1) and np.all(lab2[k2 == nz - 1] == 1)
    side = (i2 == 0) | (i2 == nx - 1) | (j2 == 0) | (j2 == nx - 1)
    assert np.all(lab2[side] == 1)
    mid = (k2 > 2) & (k2 < nz - 3) & (i2 >= 3) & (i2 <= nx - 4) & (j2 >= 3) & (j2 <= nx - 4)
    assert mid.sum() > 0 and np.all(lab2[mid] == 0)


def _column_run(steps, fs, pgrad_form, every):
    oracle.set_threads(0)
    p, parts = configs.isph_dem_column(8, 8, 10, 1e-3, seed=5, jitter=0.02, dt=2e-6)
    p.free_surface = fs
    p.pgrad_form = pgrad_form
    o = Oracle(p, parts)
    opts = abi.ppe_opts(method=abi.PPE_BICGSTAB, precision=abi.PPE_FP64, guess=abi.GUESS_PREVIOUS, rtol=1e-8)
    hist = []
    for s in range(steps):
        st = o.step(opts)
        assert st["converged"], (s, st)
        if (s + 1) % every == 0:
            a, _, _ = o.get_particles(abi.SPH)
            f = a["p_type"] == abi.PT_FLUID
            umax = float(np.sqrt(a["ux"] ** 2 + a["uy"] ** 2 + a["uz"] ** 2)[f].max())
            hist.append((s + 1, umax, float(a["pres"][f].max()), float(a["pres"][f].min())))
    return hist, p


def test_converged_ppe_with_free_surface_stays_at_rest():
    """A jittered 8x8x10 column at rest, 1000 steps, BiCGStab to rtol 1e-8
    every step, with the reference's symmetric pressure gradient: max |u|
    stays below 1e-3 m/s (the first steps' adjustment to the jitter is
    ~5e-4) and the pressure stays hydrostatic, between -2 % and +10 % of
    rho g H around it."""
    hist, p = _column_run(1000, 1, 0, 100)
    rho_g_h = 1000.0 * 9.8 * 10e-3
    for s, umax, pmax, pmin in hist:
        assert umax < 1e-3, (s, umax)
        assert 0.9 * rho_g_h < pmax < 1.1 * rho_g_h, (s, pmax)
        assert pmin > -0.02 * rho_g_h, (s, pmin)
    # and it has settled: the last 500 steps change max |u| by under 2 %
    assert abs(hist[-1][1] - hist[4][1]) < 0.02 * hist[-1][1]


def test_converged_ppe_without_free_surface_drifts():
    """The same column without surface rows: the wall pressure is the
    Neumann average of the fluid's, so a uniform pressure shift is
    undetermined and a converged solve lets it grow; within 100 steps the
    velocity is two orders of magnitude above the free-surface run."""
    hist, _ = _column_run(100, 0, 0, 100)
    assert hist[-1][1] > 0.1, hist
