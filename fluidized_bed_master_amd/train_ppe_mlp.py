"""Train the learned PPE warm start and report the iteration counts of every
guess.

    python -m fluidized_bed_master_amd.train_ppe_mlp [--snapshots 400] [--out weights/ppe_mlp.npy]

Workload: a 60 x 60 x 60 column (216,000 fluid at dx = 1 mm) with spheres of
the benchmark's size and packing (d = 2 mm, phi = 0.35), through the ISPH
step with a converged PPE (difference-form pressure gradient).  Every training snapshot is one time step whose mixed
BiCGStab solve to rtol 1e-8 is followed by one Adam step of the network on
(features before the solve, converged pressure), all on the device.  The
evaluation then runs the same initial state with the zero, previous-step,
extrapolated and learned guesses and prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import sys
from pathlib import Path

import numpy as np

from . import abi, configs, learned


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--snapshots", type=int, default=400)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--eval-steps", type=int, default=20)
    ap.add_argument("--eval-warm", type=int, default=20)
    ap.add_argument("--seed", type=int, default=4)
    ap.add_argument("--out", default=str(learned.WEIGHTS_DIR / "ppe_mlp.npy"))
    a = ap.parse_args(argv)

    from .sophia import Sophia
    params, parts = configs.c4_train()
    params.pgrad_form = 1

    def make():
        return Sophia(params, parts, precision=abi.PREC_MIXED)

    sim = make()
    w0 = learned.init_params(seed=a.seed)
    w, hist, secs = learned.train(sim, snapshots=a.snapshots, lr=a.lr, params=w0, log_every=50)
    sim.close()
    out = Path(a.out)
    out.parent.mkdir(parents=True, exist_ok=True)
    np.save(out, w.astype(np.float32))
    ev = learned.evaluate(make, w, steps=a.eval_steps, warm=a.eval_warm)
    losses = np.array([h["loss"] for h in hist])
    rep = {
        "workload": f"60x60x60 column ({len(parts)} particles, 2 mm spheres phi 0.35), ISPH, converged mixed PPE rtol 1e-8",
        "snapshots": a.snapshots,
        "train_seconds": secs,
        "loss_first20": float(losses[:20].mean()),
        "loss_last20": float(losses[-20:].mean()),
        "weights": str(out),
        "eval": {k: {kk: v[kk] for kk in ("mean_iterations", "mean_r0_rel", "mean_ppe_ms", "fallbacks")}
                 for k, v in ev.items()},
    }
    print(json.dumps(rep))
    return 0


if __name__ == "__main__":
    sys.exit(main())
