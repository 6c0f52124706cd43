"""Heat transfer (con_solve = 1) on the GPU against the oracle, stage by
stage, on a seeded ISPH-DEM column with a temperature gradient in the fluid,
hotter spheres and a wall heat flux through the bottom wall layer:

  interaction   conduction and Eulerian heat advection + flux-wall heat in
                the SPH temperature rate (functions_PPE2.cuh:1550-1591)
  DEM contact   radiation and contact conduction between spheres
                (functions_DEM_INTERACTION2.cuh:1080-1194, 1300-1315)
  coupling      Ranz-Marshall convective exchange: sphere Q_sd and rate,
                fluid reaction (functions_SPH_DEM_COUPLING2.cuh:996-1106, 1951)
  corrector     T = T0 + dT dt on both species

each with the reference's quirks switched on (reference_compat) and off;
then three full steps, and DEM spin integrated from the contact torque.
"""
import numpy as np
import pytest

from fluidized_bed_master_amd import abi, configs
from tests.helpers import compare_fields

pytestmark = pytest.mark.gpu

TOL = {abi.PREC_FP64: 1e-9, abi.PREC_MIXED: 1e-5}


def _heated(seed=5, compat=0):
    p, parts = configs.small_isph_dem(seed=seed)
    p.con_solve = 1
    p.reference_compat = compat
    parts = parts.copy()
    H = parts["z"][parts["p_type"] == abi.PT_FLUID].max()
    sph = parts["p_type"] < 1000
    parts["temp"][sph] = 293.0 + 40.0 * (1.0 - parts["z"][sph] / H)
    dem = parts["p_type"] > 1000
    parts["temp"][dem] = 330.0 + 20.0 * np.sin(np.arange(dem.sum()))
    p.heat_flux = 5000.0
    n = 10
    for k, v in enumerate([0.0, 0.0, -1.0e-3, n * 1e-3, n * 1e-3, 0.0]):
        p.heat_flux_box[k] = v
    return p, parts


def _pair(prec, compat=0, seed=5):
    from fluidized_bed_master_amd.sophia import Sophia
    from oracle.oracle import Oracle
    p, parts = _heated(seed, compat)
    return Sophia(p, parts, precision=prec), Oracle(p, parts)


def _cmp(g, o, sp, fields, p3fields, tol, where):
    a1, _, a3 = g.get_particles(sp)
    b1, _, b3 = o.get_particles(sp)
    if fields:
        compare_fields(a1, b1, fields, tol, where)
    if p3fields:
        compare_fields(a3, b3, p3fields, tol, where + "/p3")
    return a1, b1, a3, b3


@pytest.mark.parametrize("compat", [0, abi.COMPAT_ALL], ids=["corrected", "reference_compat"])
@pytest.mark.parametrize("prec", [abi.PREC_FP64, abi.PREC_MIXED], ids=["fp64", "mixed"])
def test_heat_stages_match_oracle(prec, compat):
    g, o = _pair(prec, compat)
    tol = TOL[prec]
    for s in (g, o):
        s.nnps(); s.prep(); s.interaction()
    a1, b1, a3, b3 = _cmp(g, o, abi.SPH, None, ["dtemp"], tol, "interaction")
    assert np.max(np.abs(b3["dtemp"])) > 0
    for s in (g, o):
        s.dem_interaction()
    _, d1, _, d3 = _cmp(g, o, abi.DEM, ["test1", "test2"], ["dtemp"], tol, "dem_interaction")
    assert np.max(np.abs(d1["test1"])) > 0 and np.max(np.abs(d1["test2"])) > 0
    for s in (g, o):
        s.sph_dem_coupling()
    _, d1, _, _ = _cmp(g, o, abi.DEM, ["Q_sd"], ["dtemp"], tol, "coupling")
    assert np.max(np.abs(d1["Q_sd"])) > 0
    _cmp(g, o, abi.SPH, None, ["dtemp"], tol, "coupling")
    for s in (g, o):
        s.predict()
        s.ppe(abi.ppe_opts(method=abi.PPE_JACOBI))
        s.correct()
    _cmp(g, o, abi.SPH, ["temp"], None, tol, "correct")
    _cmp(g, o, abi.DEM, ["temp"], None, tol, "correct")


@pytest.mark.parametrize("prec", [abi.PREC_FP64, abi.PREC_MIXED], ids=["fp64", "mixed"])
def test_heat_full_steps_and_spin(prec):
    """Three full steps with heat on: temperatures of both species, and the
    DEM spin, which the corrected path integrates from the contact torque
    (the reference's ISPH corrector leaves it alone, reproduced with
    COMPAT_NO_SPIN)."""
    g, o = _pair(prec, 0, seed=8)
    opts = abi.ppe_opts(method=abi.PPE_JACOBI)
    for _ in range(3):
        g.step(opts)
        o.step(opts)
    tol = TOL[prec] * 10
    _cmp(g, o, abi.SPH, ["temp", "ux", "uz", "pres"], None, tol, "3 steps/sph")
    _, d1, _, _ = _cmp(g, o, abi.DEM, ["temp", "wx", "wy", "wz", "ux", "uz"], None, tol, "3 steps/dem")
    d0 = _heated(8)[1]
    d0 = d0[d0["p_type"] > 1000]
    # spin changed from the start (contacts exert torque)
    assert np.max(np.abs(np.sort(d1["wx"]) - np.sort(d0["wx"]))) > 0


def _linear_table():
    from fluidized_bed_master_amd.sophia import Table
    t = Table()
    T = [250.0, 300.0, 350.0, 400.0]
    off = 0
    for pt, (k0, mu0) in ((0, (0.9, 1.2e-3)), (1, (0.6, 1.0e-3))):
        for i, x in enumerate(T):
            t.T[off + i] = x
            t.k[off + i] = k0 + 2e-3 * (x - 300.0) + 1e-4 * i * i
            t.mu[off + i] = mu0 - 1e-5 * (x - 300.0)
            t.h[off + i] = 4000.0 * x
        t.index[pt], t.size[pt] = off, 4
        off += 4
    t.n_tables, t.n_values = 2, off
    return t


@pytest.mark.parametrize("prec", [abi.PREC_FP64, abi.PREC_MIXED], ids=["fp64", "mixed"])
def test_property_tables_match_oracle(prec):
    """property_table YES: viscosity and conductivity of walls (p_type 0) and
    fluid (1) interpolated in temperature from data.txt blocks, on the ISPH
    advection and on the WCSPH interaction."""
    from fluidized_bed_master_amd.sophia import Sophia
    from oracle.oracle import Oracle
    tab = _linear_table()
    p, parts = _heated(seed=6)
    p.prop_table = 1
    g, o = Sophia(p, parts, precision=prec), Oracle(p, parts)
    g.set_table(tab)
    o.set_table(tab)
    for s in (g, o):
        s.nnps(); s.prep(); s.interaction()
    _cmp(g, o, abi.SPH, None, ["dtemp", "ftotalx", "ftotaly", "ftotalz"], TOL[prec], "advection/table")
    # WCSPH: the interaction stage of two whole steps
    p, parts = configs.wcsph_column(8, 8, 12, con=1, seed=2)
    p.prop_table = 1
    g, o = Sophia(p, parts, precision=prec), Oracle(p, parts)
    g.set_table(tab)
    o.set_table(tab)
    for _ in range(2):
        for k in range(9):
            g.wcsph_stage(k)
            o.wcsph_stage(k)
    _cmp(g, o, abi.SPH, ["temp", "ux", "uz"], None, TOL[prec] * 10, "wcsph/table")
