"""Throughput of the SOPHIA ISPH-DEM step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c5_1m]

A step is one full ISPH-DEM time step (NNPS, prep, interaction, DEM
contact, SPH-DEM coupling, predictor, PPE solve to rtol 1e-8, pressure force,
corrector, XSPH) over the whole particle set, inputs resident in HBM.
The headline path is the mixed-precision one (FP32 fields, FP32 PPE operator
inside FP64 iterative refinement) with the learned warm start when trained
weights are available, else the previous-step pressure; it is compared with
the plain FP64 path (FP64 fields and FP64 BiCGStab) on the same input.

Prints one JSON line (rank 0, N = 1).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
# compulsory bytes per particle-step of the whole step (FP32 fields, FP64
# positions): 981 B outside the PPE, 232 B per mixed BiCGStab iteration
# (two matrix-free matvecs of 72 B and the Krylov vectors) or 344 B in FP64,
# 1500 B per sphere (its 16-slot contact history read and written)
STEP_BYTES_BASE, STEP_BYTES_ITER_MIXED, STEP_BYTES_ITER_FP64, STEP_BYTES_SPHERE = 981.0, 232.0, 344.0, 1500.0
MF_MATVEC_BYTES = 72.0


def _matvec_bytes(layout: dict, n_sph: int, fp32: bool) -> float:
    """Compulsory HBM bytes of one inner PPE matvec y = A x over the compact
    operator: every stored entry's coefficient and column index once (on the
    mixed path one 4-byte word holds both), and per particle the entry count
    (4), run lengths (8), cell key (4), row flag (4), diagonal (4 FP32 inner
    diagonal, 8 FP64), own x, y and the fused dot operand (3 vector words);
    the neighbour gathers of x hit the cache and count once through the
    own-x read."""
    vec = 4 if fp32 else 8
    coef = layout["coef_bytes"] if fp32 else 8
    return layout["nnz"] * (coef + layout["index_bytes"]) + n_sph * ((24.0 if fp32 else 28.0) + 3 * vec)


def run_gpu(params, parts, precision, ppe_opts, steps, warmup, timers=False, mlp_weights=None, keep_pressure=False,
            probe=None):
    """warmup untimed steps, then `steps` timed ones between two device
    synchronisations.  With keep_pressure the final pressure of the PPE rows
    (fluid particles, in the particles' original numbering) is returned in
    the last slot, for comparisons between legs run from the same input."""
    from fluidized_bed_master_amd.sophia import Sophia
    from fluidized_bed_master_amd import abi
    g = Sophia(params, parts, precision=precision)
    if mlp_weights is not None:
        g.mlp_set_params(mlp_weights)
    stats = []
    for _ in range(warmup):
        stats.append(g.step(ppe_opts))
    g.sync()
    if timers:
        g.timers(True)
    t0 = time.perf_counter()
    st_timed = []
    for _ in range(steps):
        st_timed.append(g.step(ppe_opts))
    g.sync()
    dt = time.perf_counter() - t0
    rep = g.timers_report() if timers else {}
    layout = g.ppe_layout()
    n_sph, n_dem = g.count(abi.SPH), g.count(abi.DEM)
    if probe is not None:
        # the inner FP32 matvec on this step's operator: the solver's
        # stored-operator kernel against the matrix-free one
        _, _, rows = g.get_ppe_system()
        x = np.random.default_rng(1).normal(0.0, 1.0, len(rows)) * (rows == 1)
        _, us_stored = g.matvec_probe(0, x, reps=30)
        _, us_mf = g.matvec_probe(1, x, reps=30)
        probe.update({"stored_us": us_stored, "matrix_free_us": us_mf, "reps": 30})
    pres = None
    if keep_pressure:
        a, _, a3 = g.get_particles(abi.SPH)
        # the SPH records are in cell order; key them by position to compare legs
        fl = (a["p_type"] >= 1) & (a["p_type"] < 1000) & (a3["lbl_surf"] == 0)
        order = np.lexsort((a["x"][fl], a["y"][fl], a["z"][fl]))
        pres = a["pres"][fl][order]
    g.close()
    return dt, st_timed, rep, layout, n_sph, n_dem, pres


PMC_FILE = ROOT / "profiles" / "ppe_matvec_fp32_c5_1m_pmc.txt"


def _pmc_traffic():
    """HBM bytes per launch of the inner matvec on c5_1m from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (KiB per dispatch, each scaled
    by the factor the same counter reads on a copy kernel of known size);
    None when the summary is absent."""
    if not PMC_FILE.exists():
        return None
    kv = {}
    for line in PMC_FILE.read_text().splitlines():
        if "=" in line and not line.lstrip().startswith("#"):
            k, v = line.split("=", 1)
            kv[k.strip()] = float(v)
    try:
        return round(kv["fetch_kib_per_dispatch"] * 1024 * kv["fetch_factor"]
                     + kv["write_kib_per_dispatch"] * 1024 * kv["write_factor"])
    except KeyError:
        return None


def _step_roofline(n_sph: int, n_dem: int, iters: float, ms_step: float, per_iter: float) -> dict:
    """The whole step against HBM: compulsory bytes per step (the per-stage
    figures above times the particles and the measured iterations) over the
    measured step time."""
    b = n_sph * (STEP_BYTES_BASE + iters * per_iter) + n_dem * STEP_BYTES_SPHERE
    ach = b / (ms_step * 1e-3) / 1e9
    return {"bound": "hbm", "bytes_per_step": round(b), "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4)}


def cpu_baseline(params, parts, target_seconds: float = 15.0):
    """The FP64 C++/OpenMP oracle (same formulas, same BiCGStab with the
    Jacobi preconditioner, previous-step guess, rtol 1e-8) on the benchmark's
    own input: one untimed step, then whole steps until target_seconds have
    passed (at least one), with every host thread the process may use;
    beside it a single-thread run of the same step on a 24^3-fluid column of
    the same kind, bounded the same way."""
    import copy
    from fluidized_bed_master_amd import abi, configs
    from oracle import oracle
    opts = abi.ppe_opts(method=abi.PPE_BICGSTAB, precision=abi.PPE_FP64, guess=abi.GUESS_PREVIOUS, rtol=1e-8)

    def timed(p, pa, secs):
        o = oracle.Oracle(p, pa)
        o.step(opts)  # first step (count 0 extras) untimed
        steps, t0 = 0, time.perf_counter()
        while True:
            o.step(opts)
            steps += 1
            el = time.perf_counter() - t0
            if el >= secs or steps >= 200:
                break
        o.close()
        return len(pa) * steps / el, steps, el

    oracle.set_threads(0)
    threads = oracle.get_threads()
    v, steps, el = timed(params, parts, target_seconds)
    ps, pas = configs.isph_dem_column(24, 24, 24, 1.0e-3, d_dem=2.0e-3, dem_height=0.008, seed=5, dt=params.dt)
    ps = copy.copy(ps)
    for f in ("con_solve", "heat_flux", "free_surface", "pgrad_form"):
        setattr(ps, f, getattr(params, f))
    if params.heat_flux != 0.0:
        configs.heated(ps, pas, params.heat_flux)
    oracle.set_threads(1)
    v1, steps1, el1 = timed(ps, pas, max(5.0, target_seconds / 2))
    oracle.set_threads(0)
    return {
        "value": v,
        "unit": "particle-steps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"the benchmark input itself ({len(parts)} particles), {steps} step(s) after one untimed, "
                  f"{el:.1f} s on {threads} threads; BiCGStab FP64, Jacobi preconditioner, rtol 1e-8",
        "single_thread": {"value": v1, "unit": "particle-steps/s", "cores": 1,
                          "sample": f"24x24x24-fluid ISPH-DEM column ({len(pas)} particles), same settings, "
                                    f"{steps1} step(s), {el1:.1f} s"},
    }


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="c5_1m", choices=["c5_1m", "c5", "c2_isph", "small"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64-path comparison leg")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-heat", action="store_true", help="heat transfer off (free-surface rows stay on)")
    ap.add_argument("--pgrad-form", type=int, default=0, choices=[0, 1],
                    help="pressure gradient: 0 the reference's symmetric sum V_j (p_i + p_j) grad W, "
                         "1 the difference form sum V_j (p_j - p_i) grad W")
    ap.add_argument("--weights", default=str(ROOT / "fluidized_bed_master_amd" / "weights" / "ppe_mlp.npy"))
    ap.add_argument("--precond", default="two_level", choices=["two_level", "jacobi"],
                    help="preconditioner of the mixed-precision solve")
    ap.add_argument("--inner-rtol", type=float, default=0.001,
                    help="mixed solve: fold into the FP64 solution each time the FP32 residual drops by this factor")
    ap.add_argument("--legs", default="fp64,fp64_pc,prev,extrap,bc,c3",
                    help="comparison legs to run: fp64 (plain FP64 path), fp64_pc (FP64, same preconditioner), "
                         "prev (mixed, previous-step guess), extrap (mixed, extrapolated guess without the "
                         "learned correction), bc (open boundaries and a DEM wall on the same column: "
                         "cost of the two boundary kernels), c3 (the periodic Taylor-Green box at 100^3: "
                         "zero-mean PPE, FP64 path vs mixed)")
    a = ap.parse_args(argv)

    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {a.gpus}); "
              "multi-GPU slab decomposition is a design, not code", file=sys.stderr)
        return 2
    if int(os.environ.get("RANK", "0")) != 0:
        return 0

    from fluidized_bed_master_amd import abi, configs
    gen = {"c5_1m": configs.c5_1m, "c5": configs.c5, "c2_isph": configs.c2_isph,
           "small": lambda: configs.small_isph_dem()}[a.config]
    params, parts = gen()
    # C5's heat transfer (conduction, radiation, convective exchange, a
    # 5 kW/m^2 side-wall flux) and free-surface PPE rows, under which the
    # converged solve with the reference's symmetric pressure gradient stays
    # at rest
    if not a.no_heat:
        configs.heated(params, parts)
    else:
        params.free_surface = 1
    params.pgrad_form = a.pgrad_form
    n_total = len(parts)

    weights = None
    if Path(a.weights).exists():
        weights = np.load(a.weights).astype(np.float32)
    guess = abi.GUESS_LEARNED if weights is not None else abi.GUESS_EXTRAPOLATE
    precond = abi.PRECOND_TWO_LEVEL if a.precond == "two_level" else abi.PRECOND_JACOBI
    mixed_opts = abi.ppe_opts(method=abi.PPE_BICGSTAB, precision=abi.PPE_MIXED_IR, guess=guess, rtol=1e-8,
                              precond=precond, inner_rtol=a.inner_rtol)
    legs = set(a.legs.split(",")) if not a.no_fp64 else set()
    # the plain FP64 path: the reference's Jacobi-preconditioned system, FP64
    # operator and vectors, previous-step guess
    fp64_opts = abi.ppe_opts(method=abi.PPE_BICGSTAB, precision=abi.PPE_FP64, guess=abi.GUESS_PREVIOUS, rtol=1e-8)

    probe = {}
    dt, st, rep, layout, n_sph, n_dem, pres_head = run_gpu(params, parts, abi.PREC_MIXED, mixed_opts, a.steps, a.warmup,
                                                     timers=True, mlp_weights=weights, keep_pressure=True,
                                                     probe=probe)
    value = n_total * a.steps / dt
    ms_step = dt / a.steps * 1e3
    ppe_ms = float(np.mean([s["ms"] for s in st]))
    iters = float(np.mean([s["iterations"] for s in st]))

    # dominant kernel group and its roofline: the inner PPE matvec.  Launches
    # past convergence inside the last host-checked chunk return at once, so
    # the average is taken over the 2 matvecs of each counted iteration with
    # the time of every launch included.
    roofline = None
    name = "ppe_matvec_fp32"
    if name in rep:
        # HIP events bracket one launch in four (kMvSample); every launch is
        # counted under name + ".launches", so the timed launches are a
        # uniform sample of all of them, early exits included
        cnt, tot = rep[name]
        allc = rep.pop(name + ".launches", (cnt, 0.0))[0]
        tot_all = tot * allc / cnt if cnt else 0.0
        rep[name] = (allc, tot_all)
        real = 2 * int(sum(s["iterations"] for s in st))
        if real > 0 and cnt > 0:
            ab = _matvec_bytes(layout, n_sph, fp32=True)
            avg_s = tot_all / real / 1e3
            ach = ab / avg_s / 1e9
            roofline = {"bound": "hbm", "kernel": "k_bi_mv_staged (inner FP32 matvec)", "achieved": round(ach, 1),
                        "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                        "traffic": _pmc_traffic() if a.config == "c5_1m" else None,
                        "traffic_source": str(PMC_FILE.relative_to(ROOT)) if a.config == "c5_1m" else None,
                        "launches": real, "launches_incl_early_exit": allc, "launches_timed": cnt,
                        "avg_us": round(avg_s * 1e6, 2), "bytes_per_launch": ab, "nnz": layout["nnz"],
                        "index_bytes": layout["index_bytes"]}
            # the same launches against the compulsory traffic of a
            # matrix-free matvec (72 B per particle: no stored operator)
            mf = MF_MATVEC_BYTES * n_total
            roofline["matrix_free_bytes_per_launch"] = mf
            roofline["frac_vs_matrix_free_bytes"] = round(mf / avg_s / 1e9 / HBM_PEAK_GBS, 4)

    fp64 = None
    if "fp64" in legs:
        dt64, st64, _, _, _, _, pres64 = run_gpu(params, parts, abi.PREC_FP64, fp64_opts, a.steps, a.warmup,
                                                 keep_pressure=True)
        fp64 = {"value": n_total * a.steps / dt64, "ms_per_step": dt64 / a.steps * 1e3,
                "step_roofline": _step_roofline(n_sph, n_dem, float(np.mean([s["iterations"] for s in st64])),
                                                dt64 / a.steps * 1e3, STEP_BYTES_ITER_FP64),
                "ppe_ms": float(np.mean([s["ms"] for s in st64])),
                "ppe_iters": float(np.mean([s["iterations"] for s in st64])),
                "ppe_r0_rel": float(np.mean([s["r0_rel"] for s in st64])),
                "ppe_converged_all": bool(all(s["converged"] for s in st64)),
                "ppe_r_rel_max": float(np.max([s["r_rel"] for s in st64]))}
        # the headline's pressure against the FP64 path's after the same
        # warm-up + timed steps from the same input (fluid rows): the
        # FP64 path matches the FP64 oracle to 1e-9 in the parity tests
        if pres64 is not None and pres_head is not None and len(pres64) == len(pres_head):
            fp64["headline_p_rel_err"] = float(np.max(np.abs(pres_head - pres64)) / np.max(np.abs(pres64)))

    # the FP64 path with the same preconditioner: what precision alone buys
    fp64_pc = None
    if "fp64_pc" in legs and precond != abi.PRECOND_JACOBI:
        pc_opts = abi.ppe_opts(method=abi.PPE_BICGSTAB, precision=abi.PPE_FP64, guess=abi.GUESS_PREVIOUS, rtol=1e-8,
                               precond=precond)
        dtq, stq, _, _, _, _, _ = run_gpu(params, parts, abi.PREC_FP64, pc_opts, a.steps, a.warmup)
        fp64_pc = {"value": n_total * a.steps / dtq, "ms_per_step": dtq / a.steps * 1e3,
                   "ppe_ms": float(np.mean([s["ms"] for s in stq])),
                   "ppe_iters": float(np.mean([s["iterations"] for s in stq])),
                   "precond": a.precond, "guess": "previous"}

    # the same mixed path with the plain previous-step guess: what the warm start buys
    prev = None
    if "prev" in legs:
        prev_opts = abi.ppe_opts(method=abi.PPE_BICGSTAB, precision=abi.PPE_MIXED_IR, guess=abi.GUESS_PREVIOUS,
                                 rtol=1e-8, precond=precond, inner_rtol=a.inner_rtol)
        dtp, stp, _, _, _, _, _ = run_gpu(params, parts, abi.PREC_MIXED, prev_opts, a.steps, a.warmup)
        prev = {"ms_per_step": dtp / a.steps * 1e3, "ppe_ms": float(np.mean([s["ms"] for s in stp])),
                "ppe_iters": float(np.mean([s["iterations"] for s in stp])),
                "ppe_r0_rel": float(np.mean([s["r0_rel"] for s in stp]))}

    # the same mixed path without the learned correction: what the network adds
    extrap = None
    if "extrap" in legs and weights is not None:
        ex_opts = abi.ppe_opts(method=abi.PPE_BICGSTAB, precision=abi.PPE_MIXED_IR, guess=abi.GUESS_EXTRAPOLATE,
                               rtol=1e-8, precond=precond, inner_rtol=a.inner_rtol)
        dte, ste, _, _, _, _, _ = run_gpu(params, parts, abi.PREC_MIXED, ex_opts, a.steps, a.warmup)
        extrap = {"value": n_total * a.steps / dte, "ms_per_step": dte / a.steps * 1e3,
                  "ppe_ms": float(np.mean([s["ms"] for s in ste])),
                  "ppe_iters": float(np.mean([s["iterations"] for s in ste])),
                  "ppe_r0_rel": float(np.mean([s["r0_rel"] for s in ste]))}

    boundaries = None
    if "bc" in legs:
        # the same step with open boundaries (bottom and top 2 mm of the fluid
        # as inlet and outlet buffers) and a DEM box wall around the column;
        # reports what the two boundary kernels cost per step
        import copy
        pb = copy.copy(params)
        fl = parts["p_type"] == 1
        z0, z1 = float(parts["z"][fl].min()), float(parts["z"][fl].max())
        pb.open_boundary = 1
        for k, v in enumerate([-1.0, z0 + 2e-3, z1 - 2e-3, 1.0]):
            pb.ob_band[k] = v
        pb.ob_inlet_u = 0.0
        pb.ob_inlet_temp = 293.0
        dm = parts["p_type"] > 1000
        lo = [float(parts[c][dm].min()) for c in ("x", "y", "z")]
        hi = [float(parts[c][dm].max()) for c in ("x", "y", "z")]
        rad = float(parts["rad"][dm].max())
        pb.dem_walls = 2
        for k in range(3):
            pb.wall_box[k] = 0.0 if k < 2 else 0.5 * (lo[k] + hi[k])
            pb.wall_box[3 + k] = (2 * max(abs(lo[k]), abs(hi[k])) if k < 2 else hi[k] - lo[k]) + 4 * rad
        steps_bc = max(1, min(a.steps, 3))
        dtb, stb, repb, _, _, _, _ = run_gpu(pb, parts, abi.PREC_MIXED, mixed_opts, steps_bc, 1, timers=True,
                                          mlp_weights=weights)
        boundaries = {"ms_per_step": dtb / steps_bc * 1e3,
                      "open_boundary_ms": repb.get("open_boundary", (0, 0.0))[1] / steps_bc,
                      "dem_walls_ms": repb.get("dem_walls", (0, 0.0))[1] / steps_bc,
                      "buffer_particles": int(((parts["z"] < z0 + 2e-3) | (parts["z"] >= z1 - 2e-3))[fl].sum()),
                      "steps": steps_bc}

    # C3: the periodic Taylor-Green box (1,000,000 fluid particles, no walls
    # or spheres), Lagrangian ISPH with a zero-mean PPE, Jacobi-preconditioned
    # BiCGStab to rtol 1e-8 from the previous pressure: FP64 path vs mixed
    c3 = None
    if "c3" in legs:
        p3, parts3 = configs.c3()
        c3 = {"workload": "c3: periodic Taylor-Green box [0, 2 pi)^3, 100^3 = 1,000,000 fluid, h = 1.3 dx, "
                          f"Wendland C2, dt = {p3.dt:g}, zero-mean PPE, Jacobi preconditioner, previous-step guess",
              "particles": len(parts3)}
        steps3 = max(1, min(a.steps, 5))
        for name, prec, pp in (("fp64", abi.PREC_FP64, abi.PPE_FP64), ("mixed", abi.PREC_MIXED, abi.PPE_MIXED_IR)):
            o3 = abi.ppe_opts(method=abi.PPE_BICGSTAB, precision=pp, guess=abi.GUESS_PREVIOUS, rtol=1e-8,
                              zero_mean=1, inner_rtol=a.inner_rtol)
            dt3, st3, _, _, _, _, _ = run_gpu(p3, parts3, prec, o3, steps3, 1)
            c3[name] = {"value": len(parts3) * steps3 / dt3, "ms_per_step": dt3 / steps3 * 1e3,
                        "ppe_ms": float(np.mean([s["ms"] for s in st3])),
                        "ppe_iters": float(np.mean([s["iterations"] for s in st3])),
                        "ppe_converged_all": bool(all(s["converged"] for s in st3)),
                        "ppe_r_rel_max": float(np.max([s["r_rel"] for s in st3]))}
        c3["steps"] = steps3
        c3["ppe_speedup_mixed_vs_fp64"] = c3["fp64"]["ppe_ms"] / c3["mixed"]["ppe_ms"]

    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(params, parts, a.cpu_seconds)

    out = {
        "metric": "particle-steps/s ISPH-DEM 1M; PPE ms learned+FP32 vs FP64 @rtol 1e-8",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+fp64",
        "data": "synthetic (seeded splitmix64 column, jittered-lattice spheres)",
        "config": {
            "workload": f"{a.config}: ISPH-DEM column, {n_sph - int(np.sum(parts['p_type'] == 0))} fluid + "
                        f"{int(np.sum(parts['p_type'] == 0))} wall + {n_dem} DEM, dx=1mm, h=1.3dx, "
                        f"Wendland C2, dt={params.dt:g}, "
                        + (f"heat on (conduction, sphere radiation and contact conduction, Ranz-Marshall "
                           f"exchange, {params.heat_flux:g} W/m^2 side-wall flux)" if params.con_solve else "heat off")
                        + ", free-surface PPE rows",
            "particles": n_total,
            "ppe": "BiCGStab rtol 1e-8, FP32 iterations with FP64 reliable updates, "
                   + ("two-level (Jacobi + 3x3x3-cell aggregate coarse correction) preconditioner, "
                      if precond == abi.PRECOND_TWO_LEVEL else "Jacobi preconditioner, ")
                   + ("learned correction of the extrapolated guess" if weights is not None
                      else "quadratic extrapolation of the last three pressures"),
            "pressure_gradient": ("difference form sum_j V_j (p_j - p_i) grad~W_ij" if params.pgrad_form else
                                  "the reference's symmetric form sum_j V_j (p_i + p_j) grad~W_ij"),
            "parallelism": "single GPU",
        },
        "ppe_ms_per_step": ppe_ms,
        "ppe_iters_per_step": iters,
        "ppe_folds_per_step": float(np.mean([s["outer"] for s in st])),
        "ppe_inner_rtol": a.inner_rtol,
        "ppe_r0_rel": float(np.mean([s["r0_rel"] for s in st])),
        "ppe_converged_all": bool(all(s["converged"] for s in st)),
        "ppe_r_rel_max": float(np.max([s["r_rel"] for s in st])),
        "ppe_guess_learned_frac": float(np.mean([s["guess_used"] == abi.GUESS_LEARNED for s in st])),
        "ppe_precond_used": sorted({int(s["precond_used"]) for s in st}),
        "ppe_n_coarse": int(st[-1]["n_coarse"]),
        "ppe_coarse_res": float(np.max([s["coarse_res"] for s in st])),
        "fp64_path": fp64,
        "fp64_same_precond": fp64_pc,
        "mixed_previous_guess": prev,
        "mixed_extrapolated_guess": extrap,
        "boundaries": boundaries,
        "c3_periodic": c3,
        "speedup_vs_fp64_path": (value / fp64["value"]) if fp64 else None,
        "speedup_vs_fp64_same_precond": (value / fp64_pc["value"]) if fp64_pc else None,
        "kernel_ms": {k: {"launches": v[0], "total_ms": round(v[1], 3)} for k, v in sorted(rep.items())},
        "roofline": roofline,
        "step_roofline": _step_roofline(n_sph, n_dem, iters, ms_step, STEP_BYTES_ITER_MIXED),
        "matvec_probe": probe,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
