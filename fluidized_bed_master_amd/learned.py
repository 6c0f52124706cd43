"""Learned PPE warm start: initialisation, on-GPU training from the solver's
own pressure fields, and the A/B evaluation against the plain guesses.

The network and its kernels live in csrc/learned.hip (BF16 weights on MFMA,
FP32 accumulation).  Training data are the converged pressures the solver
produces: each training step is one time step whose PPE solve is followed by
one Adam step on (inputs before the solve, converged pressure), all on the
device (sophia_ppe_opts.train).
"""
from __future__ import annotations

import math
import time
from pathlib import Path

import numpy as np

from . import abi
from .configs import splitmix64

H, EIN = 64, 6
OFF_W1 = 0
OFF_B1 = OFF_W1 + H * EIN
OFF_W2 = OFF_B1 + H
OFF_B2 = OFF_W2 + H * H
OFF_W3 = OFF_B2 + H
OFF_B3 = OFF_W3 + H * H
OFF_W4 = OFF_B3 + H
OFF_B4 = OFF_W4 + H
NPARAM = OFF_B4 + 1

WEIGHTS_DIR = Path(__file__).resolve().parent / "weights"


def normal(seed: int, n: int) -> np.ndarray:
    """Standard normals from splitmix64 (Box-Muller)."""
    m = (n + 1) // 2
    u = (splitmix64(seed, 2 * m) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    u1 = np.maximum(u[:m], 1e-300)
    u2 = u[m:]
    r = np.sqrt(-2.0 * np.log(u1))
    z = np.concatenate([r * np.cos(2 * math.pi * u2), r * np.sin(2 * math.pi * u2)])
    return z[:n]


def init_params(seed: int = 4, out_scale: float = 0.1) -> np.ndarray:
    """He-normal weights for the ReLU layers, zero biases, small output layer."""
    w = np.zeros(NPARAM, np.float64)
    w[OFF_W1:OFF_B1] = normal(seed * 101 + 1, H * EIN) * math.sqrt(2.0 / EIN)
    w[OFF_W2:OFF_B2] = normal(seed * 101 + 2, H * H) * math.sqrt(2.0 / H)
    w[OFF_W3:OFF_B3] = normal(seed * 101 + 3, H * H) * math.sqrt(2.0 / H)
    w[OFF_W4:OFF_B4] = normal(seed * 101 + 4, H) * math.sqrt(1.0 / H) * out_scale
    return w.astype(np.float32)


def train(sim, snapshots: int = 400, lr: float = 1e-3, precision: int = abi.PPE_MIXED_IR, rtol: float = 1e-8,
          params: np.ndarray | None = None, log_every: int = 0, precond: int = abi.PRECOND_TWO_LEVEL):
    """Advance `sim` (a Sophia context) `snapshots` steps; each PPE solve is
    followed by one Adam step of the warm-start network on that solve."""
    if params is None:
        params = init_params()
    sim.mlp_set_params(params)
    opts = abi.ppe_opts(method=abi.PPE_BICGSTAB, precision=precision, guess=abi.GUESS_LEARNED, rtol=rtol,
                        train=1, lr=lr, precond=precond)
    hist = []
    t0 = time.perf_counter()
    for k in range(snapshots):
        st = sim.step(opts)
        hist.append({"step": k, "loss": st["loss"], "iterations": st["iterations"], "guess_used": st["guess_used"],
                     "r0_rel": st["r0_rel"], "r0_rel_learned": st["r0_rel_learned"]})
        if log_every and (k % log_every == 0 or k == snapshots - 1):
            print(f"  snapshot {k}: loss {st['loss']:.4e} iters {st['iterations']} "
                  f"r0 learned {st['r0_rel_learned']:.3e} prev {st['r0_rel']:.3e}", flush=True)
    return sim.mlp_get_params(), hist, time.perf_counter() - t0


def evaluate(make_sim, params: np.ndarray, steps: int = 20, warm: int = 0,
             precision: int = abi.PPE_MIXED_IR, rtol: float = 1e-8, precond: int = abi.PRECOND_TWO_LEVEL):
    """Run the same state with zero, previous-step, extrapolated and learned guesses in
    lockstep; returns per-guess lists of PPE iterations and initial
    residuals, plus the learned path's fallback count."""
    res = {}
    for name, guess in (("zero", abi.GUESS_ZERO), ("previous", abi.GUESS_PREVIOUS),
                        ("extrapolate", abi.GUESS_EXTRAPOLATE), ("learned", abi.GUESS_LEARNED)):
        sim = make_sim()
        if guess == abi.GUESS_LEARNED:
            sim.mlp_set_params(params)
        opts = abi.ppe_opts(method=abi.PPE_BICGSTAB, precision=precision, guess=guess, rtol=rtol, precond=precond)
        warm_opts = abi.ppe_opts(method=abi.PPE_BICGSTAB, precision=precision, guess=abi.GUESS_PREVIOUS, rtol=rtol,
                                 precond=precond)
        for _ in range(warm):
            sim.step(warm_opts)
        its, r0, fb, ms, r0l, r0b = [], [], 0, [], [], []
        for _ in range(steps):
            st = sim.step(opts)
            its.append(st["iterations"])
            ms.append(st["ms"])
            r0.append(st["r0_rel"])
            if guess == abi.GUESS_LEARNED:
                # the network's own guess and the extrapolation it corrects,
                # on the same state, whichever the solve started from
                r0l.append(st["r0_rel_learned"])
                r0b.append(st["r0_rel_base"])
                if st["guess_used"] != abi.GUESS_LEARNED:
                    fb += 1
        sim.close()
        res[name] = {"iterations": its, "r0_rel": r0, "ppe_ms": ms, "fallbacks": fb,
                     "mean_iterations": float(np.mean(its)), "mean_r0_rel": float(np.mean(r0)),
                     "mean_ppe_ms": float(np.mean(ms))}
        if guess == abi.GUESS_LEARNED:
            ratio = np.array(r0l) / np.maximum(np.array(r0b), 1e-300)
            res[name].update({"r0_rel_learned": r0l, "r0_rel_base": r0b, "kept_frac": 1.0 - fb / steps,
                              "mean_r0_ratio_learned_to_base": float(np.mean(ratio))})
    return res
