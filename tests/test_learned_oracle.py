"""CPU pins of the numpy warm-start oracle: its backward is the exact
gradient (finite differences in FP64), and BF16 rounding matches the
round-to-nearest-even definition."""
import numpy as np

from fluidized_bed_master_amd import learned
from oracle import mlp


def test_bf16_rounding_rne():
    x = np.array([1.0, 1.00390625, 1.005859375, -3.1415926, 65504.0, 1e-30], np.float32)
    r = mlp.bf16(x)
    # 1 + 2^-8 is a tie -> even (1.0); 1 + 1.5*2^-8 rounds up to 1 + 2^-7
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0078125
    assert abs(r[3] + 3.140625) < 1e-12
    bits = np.asarray(r, np.float32).view(np.uint32)
    assert np.all((bits & 0xFFFF) == 0)


def test_param_layout_matches_package():
    assert mlp.NPARAM == learned.NPARAM == 8833
    assert (mlp.OFF_W2, mlp.OFF_W3, mlp.OFF_B4) == (learned.OFF_W2, learned.OFF_W3, learned.OFF_B4)


def _toy(seed=3, n=60):
    rng = np.random.default_rng(seed)
    X = rng.uniform(0, 4e-3, (n, 3))
    h = np.full(n, 1.3e-3)
    nf = np.column_stack([rng.uniform(0.5, 1.0, n), rng.normal(0, 1, n)])
    rows = (rng.uniform(0, 1, n) < 0.8).astype(np.int32)
    src, dst = mlp.edges(X, h, 2.0)
    t = rng.normal(0, 0.5, n)
    return X, h, nf, rows, src, dst, t


def test_backward_matches_finite_differences():
    X, h, nf, rows, src, dst, t = _toy()
    w = learned.init_params(seed=7, out_scale=1.0).astype(np.float64)
    w[mlp.OFF_B1:mlp.OFF_W2] = 0.05
    w[mlp.OFF_B2:mlp.OFF_W3] = 0.05
    w[mlp.OFF_B3:mlp.OFF_W4] = 0.05
    loss, g, _ = mlp.loss_grad(w, X, h, nf, rows, src, dst, t, bf=False)
    rng = np.random.default_rng(0)
    picks = np.concatenate([rng.choice(np.arange(a, b), 6, replace=False)
                            for a, b in [(mlp.OFF_W1, mlp.OFF_B1), (mlp.OFF_B1, mlp.OFF_W2), (mlp.OFF_W2, mlp.OFF_B2),
                                         (mlp.OFF_B2, mlp.OFF_W3), (mlp.OFF_W3, mlp.OFF_B3), (mlp.OFF_B3, mlp.OFF_W4),
                                         (mlp.OFF_W4, mlp.OFF_B4)]] + [[mlp.OFF_B4]])
    eps = 1e-6
    bad = []
    for p in picks:
        wp = w.copy()
        wp[p] += eps
        wm = w.copy()
        wm[p] -= eps
        lp = mlp.loss_grad(wp, X, h, nf, rows, src, dst, t, bf=False)[0]
        lm = mlp.loss_grad(wm, X, h, nf, rows, src, dst, t, bf=False)[0]
        fd = (lp - lm) / (2 * eps)
        if abs(fd - g[p]) > 1e-5 * max(1.0, abs(fd)) + 1e-7:
            bad.append((p, fd, g[p]))
    # ReLU kinks may sit inside +-eps for an isolated parameter
    assert len(bad) <= 1, bad
