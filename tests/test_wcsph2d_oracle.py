"""CPU pins of the oracle's 2D WCSPH step (SOPHIA_multi's 2D kernels,
WCSPH_Calc.cuh:368-563).

Each check is fixed by mathematics, not by the kernels' text:
- KGC_2D: sum_j V_j (x_j - x_i) (x) C_i grad W_ij is the identity for every
  particle whose moment matrix is invertible;
- prep2D: rho_ref = m / (h / h_over_dx)^2, the Shepard sum of a regular
  lattice is 1 in the interior, and the density gradient of a linear field
  is exact with KGC: the reference stores minus the gradient;
- continuity_norm2D: -rho^2 / rho_ref div u for a linear velocity field;
  the Molteni term is delta h c0 times the Laplacian of a quadratic density
  in the interior, and the reference's Antuono term (with its stored minus
  gradient) doubles it there instead of cancelling it;
- boundary2D: walls take minus the fluid velocity for a uniform flow;
- EOS: B ((rho / rho_ref)^gamma - 1) with B = eos_b_factor c0^2 / gamma;
- interaction2D: -G / rho in a linear pressure field, (mu / rho) lap u
  (Morris) and k lap T / (rho c_p m_i) (the reference's division by m_i)
  for quadratic fields, gravity alone gives (0, -g), artificial viscosity
  is off under expansion and dissipates energy;
- time update and XSPH: an isolated particle falls with symplectic Euler
  kinematics; a uniform flow is translated exactly with XSPH on.
"""
import numpy as np
import pytest

from fluidized_bed_master_amd import abi, configs
from oracle.oracle import Oracle, knl_eval


def _run(o, stages):
    for k in stages:
        o.wcsph_stage(k)


def _pairs(p1, kappa):
    """All SPH pairs (i, j) with 0 <= r_ij < kappa h_i, brute force."""
    x, y, h = p1["x"], p1["y"], p1["h"]
    dx = x[:, None] - x[None, :]
    dy = y[:, None] - y[None, :]
    r = np.sqrt(dx * dx + dy * dy)
    i, j = np.nonzero(r < kappa * h[:, None])
    return i, j, dx[i, j], dy[i, j], r[i, j]


def _interior(p1, lo, hi, margin):
    return (p1["p_type"] == 1) & (p1["x"] > lo[0] + margin) & (p1["x"] < hi[0] - margin) & \
        (p1["y"] > lo[1] + margin) & (p1["y"] < hi[1] - margin)


def _fluid_only(**kw):
    """A fluid block with no walls (regular lattice unless jitter > 0)."""
    p, parts = configs.dam_break_2d(**kw)
    return p, parts[parts["p_type"] == 1].copy()


def test_kgc2d_reproduces_the_identity_on_linear_fields():
    p, parts = configs.dam_break_2d(nx=10, ny=14, tank=(20, 18), jitter=0.1, u_noise=0.1)
    o = Oracle(p, parts)
    _run(o, range(9))  # one step: non-uniform densities
    _run(o, range(3))
    p1, _, p3 = o.get_particles(abi.SPH)
    i, j, dx, dy, r = _pairs(p1, p.kappa)
    keep = r > 0
    i, j, dx, dy, r = i[keep], j[keep], dx[keep], dy[keep], r[keep]
    h = float(p1["h"][0])
    _, dw = knl_eval(p.kernel_type, 2, h, r)
    gx, gy = dw * dx / r, dw * dy / r
    C = p3["Cm"][i]
    cx = C[:, 0, 0] * gx + C[:, 0, 1] * gy
    cy = C[:, 1, 0] * gx + C[:, 1, 1] * gy
    V = p1["m"][j] / p1["rho"][j]
    n = len(p1)
    S = np.zeros((n, 2, 2))
    for a, ra in enumerate((-dx, -dy)):
        for b, cb in enumerate((cx, cy)):
            S[:, a, b] = np.bincount(i, V * ra * cb, minlength=n)
    err = np.abs(S - np.eye(2)[None]).max()
    assert err < 1e-11, err
    assert np.all(p3["Cm"][:, 2, :] == 0) and np.all(p3["Cm"][:, :, 2] == 0)


def test_prep2d_reference_density_shepard_and_gradient_sign():
    p, parts = _fluid_only(nx=14, ny=14, tank=(20, 20), jitter=0.0)
    a, b = 2.0, -3.0  # relative density gradient per metre
    parts["rho"] = 1000.0 * (1.0 + a * parts["x"] + b * parts["y"])
    parts["m"] = parts["rho"] * 1e-4  # m / rho stays dx^2
    o = Oracle(p, parts)
    _run(o, range(3))
    p1, p2, _ = o.get_particles(abi.SPH)
    hv = p1["h"] / p.h_over_dx
    assert np.allclose(p2["rho_ref"], p1["m"] / (hv * hv), rtol=1e-15)
    dx = 0.01
    inner = _interior(p1, (0, 0), (14 * dx, 14 * dx), 2.6 * dx)
    assert inner.sum() >= 16
    # interior Shepard sum = the lattice sum of dx^2 W over the offsets
    # within kappa h (close to, not exactly, 1)
    h = float(p1["h"][0])
    a_, b_ = np.meshgrid(np.arange(-3, 4), np.arange(-3, 4))
    rr = np.sqrt(a_ * a_ + b_ * b_).ravel() * dx
    w, _ = knl_eval(p.kernel_type, 2, h, rr[rr < p.kappa * h])
    lattice = dx * dx * w.sum()
    assert abs(lattice - 1.0) < 0.02
    assert np.allclose(p1["flt_s"][inner], lattice, rtol=1e-12), (p1["flt_s"][inner].min(), lattice)
    # the KGC gradient of a linear field is exact at every particle; the
    # reference accumulates -(rho_j - rho_i) m_j / rho_j grad~W: minus it
    gx, gy = -1000.0 * a, -1000.0 * b
    assert np.allclose(p1["grad_rhox"], gx, rtol=0, atol=1e-9 * 1000 * abs(b))
    assert np.allclose(p1["grad_rhoy"], gy, rtol=0, atol=1e-9 * 1000 * abs(b))


@pytest.mark.parametrize("jitter", [0.0, 0.1])
def test_continuity2d_linear_velocity(jitter):
    p, parts = _fluid_only(nx=12, ny=12, tank=(20, 20), jitter=jitter, delsph=0, kgc=1)
    A = np.array([[1.5, 0.7], [-0.4, -0.5]])
    c = np.array([0.06, 0.06])
    parts["ux"] = A[0, 0] * (parts["x"] - c[0]) + A[0, 1] * (parts["y"] - c[1])
    parts["uy"] = A[1, 0] * (parts["x"] - c[0]) + A[1, 1] * (parts["y"] - c[1])
    o = Oracle(p, parts)
    _run(o, range(4))
    p1, p2, p3 = o.get_particles(abi.SPH)
    # the predictor moved every particle by u dt / 2: in the new positions
    # the field is linear with matrix A (I + A dt / 2)^-1
    Ap = A @ np.linalg.inv(np.eye(2) + A * p.dt / 2)
    want = -p1["rho"] ** 2 / p2["rho_ref"] * np.trace(Ap)
    err = np.max(np.abs(p3["drho"] - want)) / np.max(np.abs(want))
    assert err < 1e-10, err


def test_continuity2d_delta_sph_laplacian_and_antuono_sign():
    out = {}
    for model in (1, 2):
        p, parts = _fluid_only(nx=20, ny=20, tank=(24, 24), jitter=0.0, delsph=model, kgc=1)
        cq = 40.0  # rho = rho0 (1 + cq |x - xc|^2)
        r2 = (parts["x"] - 0.1) ** 2 + (parts["y"] - 0.1) ** 2
        parts["rho"] = 1000.0 * (1.0 + cq * r2)
        parts["m"] = parts["rho"] * 1e-4
        o = Oracle(p, parts)
        _run(o, range(4))
        p1, _, p3 = o.get_particles(abi.SPH)
        dx = 0.01
        # neighbours of these particles are interior too: their corrected
        # gradients of the quadratic field are exact by symmetry
        inner = _interior(p1, (0, 0), (20 * dx, 20 * dx), 5.3 * dx)
        assert inner.sum() >= 36
        # rho_ref = rho0 and u = 0: drho is the delta-SPH term alone
        out[model] = p3["drho"][inner]
        h = float(p1["h"][0])
        lap = 4.0 * cq * 1000.0
        if model == 1:
            want = p.delta_sph * h * p.soundspeed * lap
            err = np.max(np.abs(out[1] - want)) / want
            assert err < 0.03, err
    ratio = out[2] / out[1]
    assert np.allclose(ratio, 2.0, rtol=1e-9), (ratio.min(), ratio.max())


def test_boundary2d_walls_take_minus_the_fluid_velocity():
    U = np.array([0.3, -0.2])
    for noslip in (1, 0):
        p, parts = configs.dam_break_2d(nx=10, ny=12, tank=(16, 14), jitter=0.05, noslip=noslip)
        fl = parts["p_type"] == 1
        parts["ux"][fl], parts["uy"][fl] = U
        o = Oracle(p, parts)
        _run(o, range(3))
        p1, _, _ = o.get_particles(abi.SPH)
        w = p1["p_type"] == 0
        f = ~w
        assert np.all(p1["ux"][f] == U[0]) and np.all(p1["uy"][f] == U[1])
        if noslip == 0:
            assert np.all(p1["ux"][w] == 0) and np.all(p1["uy"][w] == 0)
            continue
        i, j, _, _, r = _pairs(p1, p.kappa)
        has = np.zeros(len(p1), bool)
        has[i[(p1["p_type"][j] == 1) & (r + 1e-20 < p.kappa * p1["h"][i])]] = True
        near, far = w & has, w & ~has
        assert near.sum() > 20 and far.sum() > 20
        assert np.allclose(p1["ux"][near], -U[0], rtol=1e-12)
        assert np.allclose(p1["uy"][near], -U[1], rtol=1e-12)
        assert np.all(p1["ux"][far] == 0) and np.all(p1["uy"][far] == 0)


def test_eos_uses_eos_b_factor():
    p, parts = configs.dam_break_2d(nx=8, ny=8, tank=(12, 12), jitter=0.0)
    parts["rho"] = 1001.0  # m stays rho0 dx^2: rho_ref = 1000
    o = Oracle(p, parts)
    _run(o, range(5))
    p1, p2, _ = o.get_particles(abi.SPH)
    assert np.all(p2["rho_ref"] == 1000.0)
    B = p.eos_b_factor * p.soundspeed ** 2 / p.gamma
    assert p.eos_b_factor == 1000.0
    assert np.allclose(p1["pres"], B * (1.001 ** 7 - 1.0), rtol=1e-12)


def _forces(p, parts, pres=None):
    o = Oracle(p, parts)
    _run(o, range(4))
    if pres is not None:
        p1 = o.get_particles(abi.SPH)[0]
        o.set_pressure(pres(p1))
    o.wcsph_stage(6)
    return o.get_particles(abi.SPH)


def _quiet(p):
    p.fp_solve = p.fv_solve = p.fva_solve = p.fg_solve = 0
    p.con_solve = 0
    return p


def test_interaction2d_linear_pressure_gives_minus_gradient_over_rho():
    p, parts = _fluid_only(nx=14, ny=14, tank=(20, 20), jitter=0.0, kgc=1)
    _quiet(p).fp_solve = 1
    G = np.array([300.0, -800.0])
    p1, _, p3 = _forces(p, parts, pres=lambda q: 50.0 + G[0] * q["x"] + G[1] * q["y"])
    dx = 0.01
    inner = _interior(p1, (0, 0), (14 * dx, 14 * dx), 2.6 * dx)
    assert inner.sum() >= 16
    assert np.allclose(p3["ftotalx"][inner], -G[0] / 1000.0, rtol=1e-9)
    assert np.allclose(p3["ftotaly"][inner], -G[1] / 1000.0, rtol=1e-9)
    assert np.allclose(p1["pgf_x"][inner], -G[0], rtol=1e-9)
    assert np.allclose(p1["pgf_y"][inner], -G[1], rtol=1e-9)


def test_interaction2d_morris_viscosity_and_conduction_laplacians():
    p, parts = _fluid_only(nx=16, ny=16, tank=(20, 20), jitter=0.0, kgc=0, con=1)
    _quiet(p)
    p.fv_solve = 1
    p.con_solve = 1
    p.mu_fluid = 0.05
    c = 7.0
    parts["ux"] = c * (parts["y"] - 0.08) ** 2
    parts["uy"] = 0.0
    parts["temp"] = 300.0 + c * ((parts["x"] - 0.08) ** 2 + (parts["y"] - 0.08) ** 2)
    # stages 0-3 with ux != 0 would move nothing (predictor: no force yet)
    p1, _, p3 = _forces(p, parts)
    dx = 0.01
    inner = _interior(p1, (0, 0), (16 * dx, 16 * dx), 2.6 * dx)
    assert inner.sum() >= 36
    rho = 1000.0
    want_u = p.mu_fluid / rho * 2.0 * c
    err = np.max(np.abs(p3["ftotalx"][inner] - want_u)) / want_u
    assert err < 0.03, err
    assert np.max(np.abs(p3["ftotaly"][inner])) < 1e-9 * want_u
    m = rho * dx * dx
    want_t = p.k_fluid * 4.0 * c / (rho * p.cp_fluid * m)
    err = np.max(np.abs(p3["dtemp"][inner] - want_t)) / want_t
    assert err < 0.03, err


def test_interaction2d_gravity_and_artificial_viscosity():
    p, parts = configs.dam_break_2d(nx=10, ny=12, tank=(16, 14), jitter=0.05)
    _quiet(p).fg_solve = 1
    p1, _, p3 = _forces(p, parts)
    assert np.all(p3["ftotalx"] == 0) and np.all(p3["ftotaly"] == -p.gravity)
    # artificial viscosity: off under a uniform expansion, dissipative otherwise
    p, parts = _fluid_only(nx=12, ny=12, tank=(20, 20), jitter=0.05)
    _quiet(p).fva_solve = 1
    parts["ux"] = 2.0 * (parts["x"] - 0.06)
    parts["uy"] = 2.0 * (parts["y"] - 0.06)
    p1, _, p3 = _forces(p, parts)
    assert np.all(p3["ftotalx"] == 0) and np.all(p3["ftotaly"] == 0)
    rng = np.random.default_rng(3)
    parts["ux"] = rng.uniform(-0.5, 0.5, len(parts))
    parts["uy"] = rng.uniform(-0.5, 0.5, len(parts))
    p1, _, p3 = _forces(p, parts)
    power = np.sum(p1["m"] * (p1["ux"] * p3["ftotalx"] + p1["uy"] * p3["ftotaly"]))
    assert power < 0 and np.max(np.abs(p3["ftotalx"])) > 0


@pytest.mark.parametrize("xsph", [0, 1])
def test_isolated_particle_falls_with_symplectic_euler(xsph):
    from fluidized_bed_master_amd.configs import _records
    p, parts = configs.dam_break_2d(nx=6, ny=6, tank=(12, 30), jitter=0.0, xsph=xsph)
    s = _records(1)
    s["x"], s["y"], s["uy"] = 0.09, 0.25, 0.1
    s["h"], s["m"], s["rho"], s["temp"] = parts["h"][0], parts["m"][0], 1000.0, 293.0
    s["p_type"] = abi.PT_FLUID
    o = Oracle(p, np.concatenate([parts, s]))
    n = 10
    for _ in range(n):
        o.step()
    p1 = o.get_particles(abi.SPH)[0]
    q = p1[np.argmax(np.where(p1["p_type"] == 1, p1["y"], -1.0))]
    g, dt = p.gravity, p.dt
    u = [0.1 - k * g * dt for k in range(1, n + 1)]
    assert abs(q["uy"] - u[-1]) < 1e-15
    assert abs(q["y"] - (0.25 + dt * sum(u))) < 1e-15
    assert q["x"] == 0.09 and q["ux"] == 0.0
    w = p1["p_type"] == 0
    assert np.all(p1["ux"][w] == 0) and np.all(p1["uy"][w] == 0)


def test_xsph2d_translates_a_uniform_flow_exactly():
    p, parts = _fluid_only(nx=10, ny=10, tank=(20, 20), jitter=0.05)
    _quiet(p)
    p.xsph_solve = 1
    p.c_xsph = 0.5
    parts["ux"], parts["uy"] = 0.4, -0.25
    o = Oracle(p, parts)
    o.step()
    p1, p2, _ = o.get_particles(abi.SPH)
    assert np.allclose(p1["x"], p2["x0"] + 0.4 * p.dt, rtol=0, atol=1e-17)
    assert np.allclose(p1["y"], p2["y0"] - 0.25 * p.dt, rtol=0, atol=1e-17)


def test_2d_spheres_and_mass_summation_are_refused():
    p, parts = configs.dam_break_2d(nx=6, ny=6, tank=(10, 10))
    s = parts[:1].copy()
    s["p_type"] = abi.PT_DEM
    o = Oracle(p, np.concatenate([parts, s]))
    with pytest.raises(RuntimeError):
        o.wcsph_stage(0)
    p.rho_type = 0
    o = Oracle(p, parts)
    with pytest.raises(RuntimeError):
        o.wcsph_stage(0)
