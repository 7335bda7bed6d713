"""The numpy reference of the surrogate network (tests/mlp_ref.py), pinned by
central finite differences of its own loss in float64 (CPU)."""
import numpy as np

import mlp_ref as R


def test_parameter_count():
    assert R.NPARAM == 34050


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = R.glorot(1)
    x = rng.normal(size=(37, R.IN))
    t = rng.normal(size=(37, R.OUT))
    g, _ = R.grad(p, x, t, emulate=False)
    idx = np.concatenate([rng.choice(R.NPARAM, 60, replace=False), np.arange(R.NPARAM - 2, R.NPARAM)])
    h = 1e-6
    for i in idx:
        q = p.copy()
        q[i] += h
        lp = np.mean((R.forward(q, x, False)[0] - t) ** 2)
        q[i] -= 2 * h
        lm = np.mean((R.forward(q, x, False)[0] - t) ** 2)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_bf16_rounding():
    v = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -9, -2.5e-3, 3.0e38], dtype=np.float32)
    r = R.bf16(v)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1 + 2 ** -7   # ties to even
    assert abs(r[3] / -2.5e-3 - 1) < 2 ** -8
