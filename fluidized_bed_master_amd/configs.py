"""Seeded synthetic inputs for the SPH-DEM step (no datasets are read).

Random numbers come from splitmix64 (Steele, Lea & Flood 2014) so the same
seed gives the same particles on every machine.  Particles are emitted in
the reference's input order: walls, then fluid, then DEM spheres
(input/p_type.txt of the reference lists 0s, then 1s, then 1001s).

Workloads (sizes in metres):
  isph_dem_column(...)   fluid lattice in a walled column (3 wall layers on
                         the bottom and four sides) with DEM spheres on a
                         seeded jittered lattice in the bottom region
  c5_1m()                0.1 x 0.1 column filled to 0.1 at dx = 1 mm:
                         1,000,000 fluid + 30,000 spheres d = 2 mm
  c5()                   0.1 x 0.1 x 0.2 fill: 2,000,000 fluid + 60,000 spheres
  c2_isph()              0.05 x 0.05 x 0.15 fill at dx = 1 mm: 375,000 fluid +
                         8,000 spheres d = 3 mm (ISPH PPE on C2's particle set)
  heated(p, parts)       C5's heating on any column: conduction on, a wall heat
                         flux through the innermost side-wall layer over the
                         fill height, free-surface PPE rows
  small_isph_dem(...)    parity-test sizes
"""
from __future__ import annotations

import math

import numpy as np

from .abi import KNL_WENDLAND2, PART1, PT_BOUNDARY, PT_DEM, PT_FLUID, Params, default_params

_MASK = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(seed: int, n: int) -> np.ndarray:
    """n consecutive outputs of splitmix64 started at `seed`."""
    with np.errstate(over="ignore"):
        z = (np.uint64(seed) + np.uint64(0x9E3779B97F4A7C15) * np.arange(1, n + 1, dtype=np.uint64)) & _MASK
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed: int, n: int, lo: float = 0.0, hi: float = 1.0) -> np.ndarray:
    u = (splitmix64(seed, n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return lo + (hi - lo) * u


def config_params(dx: float, dt: float, rho_f: float = 1000.0, mu_f: float = 1.0e-3, h_ratio: float = 1.3,
                  E: float = 1.0e7, nu: float = 0.3, mu_fric: float = 0.3, con: int = 0) -> Params:
    """Parameters of the synthetic configs (model constants set to the
    config values; everything else at the reference defaults)."""
    p = default_params()
    p.solver_type = 1
    p.dim = 3
    p.kernel_type = KNL_WENDLAND2
    p.kappa = 2.0
    p.dt = dt
    p.rho_fluid = rho_f
    p.rho0_eos = rho_f
    p.mu_fluid = mu_f
    p.h_over_dx = h_ratio
    p.dem_shear_modulus = E / (2.0 * (1.0 + nu))
    p.dem_poisson = nu
    p.dem_friction = mu_fric
    p.con_solve = con
    p.open_boundary = 0
    p.u_limit = 40.0
    p.sph_max_neighbors = 96
    return p


def _records(n: int) -> np.ndarray:
    a = np.zeros(n, PART1)
    a["i_type"] = 1
    return a


def isph_dem_column(nx: int, ny: int, nz: int, dx: float, *, wall_layers: int = 3, tank_extra: int = 2,
                    d_dem: float = 0.0, rho_s: float = 2500.0, dem_height: float = 0.0, phi: float = 0.35,
                    max_dem: int | None = None, h_ratio: float = 1.3, rho_f: float = 1000.0, g: float = 9.8,
                    jitter: float = 0.0, seed: int = 5, dt: float = 2.0e-6, temp: float = 293.0,
                    u_noise: float = 0.0, dem_u_noise: float = 0.0, dem_w_noise: float = 0.0):
    """Walled column of fluid with DEM spheres; returns (params, part1 records)."""
    h = h_ratio * dx
    H = nz * dx
    Lx, Ly = nx * dx, ny * dx
    # walls: 3 layers below and around, up to the fill height plus tank_extra layers
    nzw = nz + tank_extra
    wi = np.arange(-wall_layers, nx + wall_layers)
    wj = np.arange(-wall_layers, ny + wall_layers)
    wk = np.arange(-wall_layers, nzw)
    I, J, K = np.meshgrid(wi, wj, wk, indexing="ij")
    inside = (I >= 0) & (I < nx) & (J >= 0) & (J < ny) & (K >= 0)
    I, J, K = I[~inside], J[~inside], K[~inside]
    walls = _records(I.size)
    walls["x"] = (I + 0.5) * dx
    walls["y"] = (J + 0.5) * dx
    walls["z"] = (K + 0.5) * dx
    walls["p_type"] = PT_BOUNDARY
    # fluid lattice
    fi, fj, fk = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    fluid = _records(fi.size)
    fluid["x"] = (fi.ravel() + 0.5) * dx
    fluid["y"] = (fj.ravel() + 0.5) * dx
    fluid["z"] = (fk.ravel() + 0.5) * dx
    if jitter > 0:
        nf = fi.size
        fluid["x"] += uniform(seed * 3 + 0, nf, -jitter, jitter) * dx
        fluid["y"] += uniform(seed * 3 + 1, nf, -jitter, jitter) * dx
        fluid["z"] += uniform(seed * 3 + 2, nf, -jitter, jitter) * dx
    fluid["p_type"] = PT_FLUID
    if u_noise > 0:
        nf = fi.size
        fluid["ux"] = uniform(seed * 11 + 1, nf, -u_noise, u_noise)
        fluid["uy"] = uniform(seed * 11 + 2, nf, -u_noise, u_noise)
        fluid["uz"] = uniform(seed * 11 + 3, nf, -u_noise, u_noise) + u_noise * fluid["x"] / Lx
    for a in (walls, fluid):
        a["h"] = h
        a["m"] = rho_f * dx ** 3
        a["rho"] = rho_f
        a["temp"] = temp
        a["pres"] = rho_f * g * np.maximum(H - a["z"], 0.0)
        a["DEMpor"] = 1.0
    # DEM spheres on a jittered cubic lattice with solid fraction phi
    dem = _records(0)
    if d_dem > 0 and dem_height > 0:
        s = d_dem * (math.pi / (6.0 * phi)) ** (1.0 / 3.0)
        gx = int(Lx // s)
        gy = int(Ly // s)
        gz = int(dem_height // s)
        di, dj, dk = np.meshgrid(np.arange(gx), np.arange(gy), np.arange(gz), indexing="ij")
        nd = di.size if max_dem is None else min(di.size, max_dem)
        dem = _records(nd)
        amp = 0.5 * (s - d_dem) * 0.98
        ox = (Lx - gx * s) / 2
        oy = (Ly - gy * s) / 2
        dem["x"] = ox + (di.ravel()[:nd] + 0.5) * s + uniform(seed * 7 + 1, nd, -amp, amp)
        dem["y"] = oy + (dj.ravel()[:nd] + 0.5) * s + uniform(seed * 7 + 2, nd, -amp, amp)
        dem["z"] = (dk.ravel()[:nd] + 0.5) * s + uniform(seed * 7 + 3, nd, -amp, amp)
        r = d_dem / 2
        dem["rad"] = r
        dem["m"] = rho_s * 4.0 / 3.0 * math.pi * r ** 3
        dem["ri"] = 0.4 * dem["m"] * r * r
        dem["rho"] = rho_s
        dem["h"] = h
        dem["temp"] = temp
        dem["p_type"] = PT_DEM
        dem["dem_idx"] = np.arange(1, nd + 1)
        if dem_u_noise > 0:
            dem["ux"] = uniform(seed * 13 + 1, nd, -dem_u_noise, dem_u_noise)
            dem["uy"] = uniform(seed * 13 + 2, nd, -dem_u_noise, dem_u_noise)
            dem["uz"] = uniform(seed * 13 + 3, nd, -dem_u_noise, dem_u_noise)
        if dem_w_noise > 0:
            dem["wx"] = uniform(seed * 17 + 1, nd, -dem_w_noise, dem_w_noise)
            dem["wy"] = uniform(seed * 17 + 2, nd, -dem_w_noise, dem_w_noise)
            dem["wz"] = uniform(seed * 17 + 3, nd, -dem_w_noise, dem_w_noise)
    parts = np.concatenate([walls, fluid, dem])
    p = config_params(dx, dt, rho_f=rho_f, h_ratio=h_ratio)
    return p, parts


def small_isph_dem(n: int = 10, nz: int = 12, dx: float = 1.0e-3, d_dem: float = 2.0e-3, seed: int = 5,
                   jitter: float = 0.02, dem_height: float = 6.0e-3, phi: float = 0.6, u_noise: float = 0.05,
                   dem_u_noise: float = 0.05, dem_w_noise: float = 20.0):
    """A parity-test column: n x n x nz fluid + walls + overlapping spheres
    (phi > pi/6 puts lattice neighbours in contact), with seeded random
    fluid and sphere velocities so that every force term is exercised."""
    return isph_dem_column(n, n, nz, dx, d_dem=d_dem, dem_height=dem_height, seed=seed, jitter=jitter,
                           phi=phi, u_noise=u_noise, dem_u_noise=dem_u_noise, dem_w_noise=dem_w_noise)


def c5_1m(seed: int = 5):
    """ISPH-DEM 1M: 0.1 x 0.1 x 0.1 fill at dx = 1 mm (1,000,000 fluid),
    30,000 spheres d = 2 mm (jittered lattice, phi = 0.35), dt = 2e-6."""
    dx = 1.0e-3
    s = 2.0e-3 * (math.pi / (6.0 * 0.35)) ** (1.0 / 3.0)
    layers = math.ceil(30000 / (int(0.1 // s) ** 2))
    return isph_dem_column(100, 100, 100, dx, d_dem=2.0e-3, dem_height=layers * s + 1e-9, max_dem=30000,
                           seed=seed, dt=2.0e-6)


def c5(seed: int = 5):
    """Full C5 geometry: 0.1 x 0.1 column filled to 0.2 (2,000,000 fluid),
    60,000 spheres d = 2 mm, dt = 2e-6."""
    dx = 1.0e-3
    s = 2.0e-3 * (math.pi / (6.0 * 0.35)) ** (1.0 / 3.0)
    layers = math.ceil(60000 / (int(0.1 // s) ** 2))
    return isph_dem_column(100, 100, 200, dx, d_dem=2.0e-3, dem_height=layers * s + 1e-9, max_dem=60000,
                           seed=seed, dt=2.0e-6)


def wcsph_column(nx: int = 10, ny: int = 10, nz: int = 20, dx: float = 1.0e-3, *, seed: int = 1,
                 jitter: float = 0.02, u_noise: float = 0.0, dt: float | None = None, con: int = 0,
                 delsph: int = 2, kgc: int = 1):
    """Weakly compressible fluid column in a walled tank (no spheres): Tait
    EOS with gamma = 7, continuity density, delta-SPH, predictor-corrector,
    Wendland C2 h = 1.3 dx.  The reference EOS uses B = 1.186 c0^2 / gamma
    with no rho0 factor (functions_PROP.cuh:21-26), so c0 is set to
    c * sqrt(rho0 / 1.186) with c = 10 sqrt(2 g H): B = rho0 c^2 / gamma."""
    p, parts = isph_dem_column(nx, ny, nz, dx, seed=seed, jitter=jitter, u_noise=u_noise, dt=2.0e-6)
    H = nz * dx
    p.solver_type = 0
    p.time_type = 1
    p.rho_type = 1
    p.delSPH_solve = delsph
    p.kgc_solve = kgc
    p.soundspeed = 10.0 * math.sqrt(2.0 * 9.8 * H) * math.sqrt(p.rho_fluid / 1.186)
    p.gamma = 7.0
    p.con_solve = con
    p.dt = dt if dt is not None else 0.2 * 1.3 * dx / p.soundspeed
    if con:
        parts = parts.copy()
        parts["temp"] = 293.0 + 50.0 * (parts["z"] < 0.3 * H)
    return p, parts


def wcsph_dem_column(nx: int = 10, ny: int = 10, nz: int = 16, dx: float = 1.0e-3, *, d_dem: float = 2.0e-3,
                     dem_height: float = 6.0e-3, phi: float = 0.6, seed: int = 2, jitter: float = 0.02,
                     u_noise: float = 0.02, dem_u_noise: float = 0.02, dem_w_noise: float = 5.0, con: int = 0,
                     dt: float | None = None):
    """The WCSPH column of wcsph_column with DEM spheres on a seeded jittered
    lattice in its bottom region (phi > pi/6 puts lattice neighbours in
    contact), run by the WCSPH step with Hertz-Mindlin contact and the WCSPH
    drag coupling."""
    p, parts = wcsph_column(nx, ny, nz, dx, seed=seed, jitter=jitter, u_noise=u_noise, dt=dt, con=con)
    _, with_dem = isph_dem_column(nx, ny, nz, dx, d_dem=d_dem, dem_height=dem_height, phi=phi, seed=seed + 100,
                                  dem_u_noise=dem_u_noise, dem_w_noise=dem_w_noise)
    dem = with_dem[with_dem["p_type"] == PT_DEM].copy()
    if con:
        dem["temp"] = 330.0
    return p, np.concatenate([parts, dem])


def dam_break_2d(nx: int = 40, ny: int = 80, dx: float = 0.01, *, tank: tuple = (160, 100), wall_layers: int = 3,
                 seed: int = 1, jitter: float = 0.01, h_ratio: float = 1.3, rho0: float = 1000.0,
                 mu: float = 1.0e-3, gamma: float = 7.0, dt: float = 1.0e-5, delsph: int = 2, kgc: int = 1,
                 noslip: int = 1, xsph: int = 1, fva: int = 1, con: int = 0, u_noise: float = 0.0):
    """C1: 2D WCSPH dam break.  A water column of nx x ny particles at
    ((i + 1/2) dx, (j + 1/2) dx) +- U(-jitter dx, jitter dx) in the left
    corner of a tank of tank = (cells wide, cells high) with `wall_layers`
    wall layers below and on both sides (no lid): 40 x 80 = 3,200 fluid and
    1,098 wall particles at the defaults (a 1.6 x 1.0 m tank, 0.4 x 0.8 m
    column).  Wendland C2 h = 1.3 dx, m = rho0 dx^2, Tait gamma = 7 with
    c0 = 10 sqrt(2 g H) and B = rho0 c0^2 / gamma (eos_b_factor = rho0),
    continuity density, delta-SPH delta = 0.1, predictor-corrector at
    dt = 1e-5.  h_over_dx = h_ratio makes the 2D reference density
    m / (h / h_over_dx)^2 equal rho0.  con = 1 adds conduction with a warm
    lower half; u_noise adds seeded velocity noise to the fluid."""
    NX, NY = tank
    h = h_ratio * dx
    wi = np.arange(-wall_layers, NX + wall_layers)
    wj = np.arange(-wall_layers, NY)
    I, J = np.meshgrid(wi, wj, indexing="ij")
    inside = (I >= 0) & (I < NX) & (J >= 0)
    I, J = I[~inside], J[~inside]
    walls = _records(I.size)
    walls["x"] = (I + 0.5) * dx
    walls["y"] = (J + 0.5) * dx
    walls["p_type"] = PT_BOUNDARY
    fi, fj = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    nf = fi.size
    fluid = _records(nf)
    fluid["x"] = (fi.ravel() + 0.5) * dx
    fluid["y"] = (fj.ravel() + 0.5) * dx
    if jitter > 0:
        fluid["x"] += uniform(seed * 3 + 0, nf, -jitter, jitter) * dx
        fluid["y"] += uniform(seed * 3 + 1, nf, -jitter, jitter) * dx
    fluid["p_type"] = PT_FLUID
    if u_noise > 0:
        fluid["ux"] = uniform(seed * 11 + 1, nf, -u_noise, u_noise)
        fluid["uy"] = uniform(seed * 11 + 2, nf, -u_noise, u_noise)
    parts = np.concatenate([walls, fluid])
    parts["h"] = h
    parts["m"] = rho0 * dx * dx
    parts["rho"] = rho0
    parts["temp"] = 293.0
    if con:
        parts["temp"] = 293.0 + 50.0 * (parts["y"] < 0.5 * ny * dx)
    H = ny * dx
    p = config_params(dx, dt, rho_f=rho0, mu_f=mu, h_ratio=h_ratio, con=con)
    p.solver_type = 0
    p.dim = 2
    p.time_type = 1
    p.rho_type = 1
    p.delSPH_solve = delsph
    p.kgc_solve = kgc
    p.delta_sph = 0.1
    p.soundspeed = 10.0 * math.sqrt(2.0 * 9.81 * H)
    p.gamma = gamma
    p.eos_b_factor = rho0
    p.noslip_bc = noslip
    p.xsph_solve = xsph
    p.fva_solve = fva
    p.sph_max_neighbors = 48
    return p, parts


def taylor_green(n: int = 100, *, seed: int = 3, noise: float = 0.01, jitter: float = 0.0, rho0: float = 1000.0,
                 mu: float = 0.1, h_ratio: float = 1.3, dt: float = 1.0e-3, kgc: int = 0,
                 drift: tuple = (0.0, 0.0, 0.0)):
    """C3: triply periodic Taylor-Green box [0, 2 pi)^3 of n^3 fluid
    particles at ((i + 1/2) dx, ...) +- U(-jitter dx, jitter dx), dx = 2 pi / n,
    u = (sin x cos y cos z, -cos x sin y cos z, 0) (1 + noise U(-1, 1)) per
    component, rho0 = 1000, Wendland C2 h = 1.3 dx; Lagrangian ISPH with a
    zero-mean PPE (the operator of a periodic box without walls is singular
    with the constant null space; without KGC it is symmetric, so a zero-mean
    right-hand side is consistent).  drift adds a uniform velocity (a
    Galilean shift that carries particles through the faces; the Taylor-Green
    field itself has no flux through them).  n = 100: 1,000,000 particles."""
    dx = 2.0 * math.pi / n
    i, j, k = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    parts = _records(i.size)
    parts["x"] = (i.ravel() + 0.5) * dx
    parts["y"] = (j.ravel() + 0.5) * dx
    parts["z"] = (k.ravel() + 0.5) * dx
    N = i.size
    if jitter > 0:
        parts["x"] += uniform(seed * 3 + 0, N, -jitter, jitter) * dx
        parts["y"] += uniform(seed * 3 + 1, N, -jitter, jitter) * dx
        parts["z"] += uniform(seed * 3 + 2, N, -jitter, jitter) * dx
    x, y, z = parts["x"], parts["y"], parts["z"]
    f = [1.0 + noise * uniform(seed * 7 + c, N, -1.0, 1.0) for c in range(3)] if noise > 0 else [1.0, 1.0, 1.0]
    parts["ux"] = np.sin(x) * np.cos(y) * np.cos(z) * f[0] + drift[0]
    parts["uy"] = -np.cos(x) * np.sin(y) * np.cos(z) * f[1] + drift[1]
    parts["uz"] = drift[2]
    parts["p_type"] = PT_FLUID
    parts["h"] = h_ratio * dx
    parts["m"] = rho0 * dx ** 3
    parts["rho"] = rho0
    parts["temp"] = 293.0
    p = config_params(dx, dt, rho_f=rho0, mu_f=mu, h_ratio=h_ratio)
    p.solver_type = 1
    p.eulerian_sph = 0
    p.kgc_solve = kgc
    p.delSPH_solve = 0
    p.fg_solve = 0
    p.fva_solve = 0
    p.xsph_solve = 0
    p.con_solve = 0
    p.domain_fixed = 1
    p.periodic_mask = 7
    for a in range(3):
        p.domain_min[a] = 0.0
        p.domain_max[a] = 2.0 * math.pi
    p.u_limit = 100.0
    return p, parts


def c3(seed: int = 3):
    """C3 at its full size: 100^3 = 1,000,000 particles."""
    return taylor_green(100, seed=seed)


def c1(seed: int = 1):
    """C1 at its full size: 3,200 fluid + 1,098 wall particles."""
    return dam_break_2d(seed=seed)


def c2(seed: int = 2):
    """C2: WCSPH-DEM fluidised bed, 0.05 x 0.05 column filled to 0.15 at
    dx = 1 mm (375,000 fluid), 8000 glass spheres d = 3 mm rho = 2500 on a
    seeded jittered lattice in the bottom 8.25 cm (phi = 0.56, the packing
    8000 such spheres need there), Hertz-Mindlin
    E = 1e7, nu = 0.3, mu = 0.3, Di Felice/Rong drag, a 0.3 m/s inlet
    through the bottom 3 fluid layers and an outlet through the top 3,
    dt = 5e-7."""
    dx = 1.0e-3
    p, parts = wcsph_column(50, 50, 150, dx, seed=seed, dt=5.0e-7)
    _, with_dem = isph_dem_column(50, 50, 150, dx, d_dem=3.0e-3, dem_height=0.0825, phi=0.56, max_dem=8000,
                                  seed=seed)
    dem = with_dem[with_dem["p_type"] == PT_DEM].copy()
    p.open_boundary = 1
    for k, v in enumerate([-1.0, 3.0e-3, 147.0e-3, 1.0]):
        p.ob_band[k] = v
    p.ob_inlet_u = 0.3
    p.ob_inlet_temp = 293.0
    return p, np.concatenate([parts, dem])


def c4_train(seed: int = 4):
    """Warm-start training column: the C5 sphere size and packing (d = 2 mm,
    phi = 0.35) in a 0.06 x 0.06 box filled to 0.06 (216,000 fluid) at
    dx = 1 mm, dt = 2e-6.  (C2's 3 mm spheres are wider than the kernel
    support at h = 1.3 mm and drive the ISPH porosity estimate below zero.)"""
    dx = 1.0e-3
    s = 2.0e-3 * (math.pi / (6.0 * 0.35)) ** (1.0 / 3.0)
    layers = math.ceil(0.3 * 0.06 / s)
    return isph_dem_column(60, 60, 60, dx, d_dem=2.0e-3, dem_height=layers * s + 1e-9, seed=seed, dt=2.0e-6)


def c2_isph(seed: int = 2):
    """C2's particle set (50 x 50 x 150 fluid at dx = 1 mm, 8000 spheres
    d = 3 mm, rho = 2500) run through the ISPH step, dt = 5e-7."""
    dx = 1.0e-3
    s = 3.0e-3 * (math.pi / (6.0 * 0.35)) ** (1.0 / 3.0)
    layers = math.ceil(8000 / (int(0.05 // s) ** 2))
    return isph_dem_column(50, 50, 150, dx, d_dem=3.0e-3, dem_height=layers * s + 1e-9, max_dem=8000,
                           seed=seed, dt=5.0e-7)


def heated(p: Params, parts: np.ndarray, flux: float = 5000.0) -> Params:
    """C5's heat transfer on a column from isph_dem_column: particle and
    sphere conduction, radiation and convective exchange (con_solve = 1), a
    wall heat flux (W/m^2) through the innermost layer of the four side
    walls over the fill height, and Dirichlet free-surface rows in the PPE
    (walls counted as neighbours), which keep a converged solve at rest."""
    fl = parts["p_type"] == PT_FLUID
    wall = parts["p_type"] == PT_BOUNDARY
    dx = float(parts["h"][0]) / p.h_over_dx
    lo = [float(parts[c][fl].min()) - 1.25 * dx for c in ("x", "y")]
    hi = [float(parts[c][fl].max()) + 1.25 * dx for c in ("x", "y")]
    zt = float(parts["z"][fl].max()) + 0.25 * dx
    p.con_solve = 1
    p.heat_flux = flux
    box = [lo[0], lo[1], 0.0, hi[0], hi[1], zt]
    for k, v in enumerate(box):
        p.heat_flux_box[k] = v
    p.free_surface = 1
    assert np.any(wall & (parts["z"] > 0) & (parts["x"] >= lo[0]) & (parts["x"] <= hi[0]))
    return p
