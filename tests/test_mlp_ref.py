"""CPU pins of the warm-start network's reference (oracle/mlp_ref.py): the exact
(float64) gradient against central finite differences, the bf16 rounding rule against
known bit patterns, and the bf16-rounded gradient against the exact one."""
import numpy as np
import pytest

import mlp_ref as R


def _problem(M=12, seed=0, scale=0.3):
    rng = np.random.default_rng(seed)
    theta = rng.normal(scale=scale, size=R.PARAMS) / np.sqrt(np.r_[np.full(R.OFF_B1, R.IN), np.ones(R.H),
                                                                   np.full(R.H * R.H, R.H), np.ones(R.H),
                                                                   np.full(R.H * R.OUT, R.H), np.ones(R.OUT)])
    X = rng.uniform(-1, 1, size=(M, R.IN))
    X[:, R.GATE + 6:] = 0.0
    t = rng.normal(size=M)
    return theta, X, t


def test_bf16_rounding_rule():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 1.0 + 2 ** -9, 3.0e-3], dtype=np.float32)
    bits = R.bf16_bits(vals)
    # ties to even: 1+2^-8 -> 1.0 (0x3F80); 1+3*2^-8 -> 1+2^-6 (0x3F82); below half -> down
    assert bits.tolist()[:5] == [0x3F80, 0x3F80, 0x3F82, 0xC020, 0x3F80]
    assert abs(R.bf16_to_f32(bits[5:]) - vals[5]) / vals[5] < 2 ** -8


def test_exact_gradient_matches_finite_differences():
    theta, X, t = _problem()
    g, _ = R.gradient(theta, X, t, bf16_points=False)
    rng = np.random.default_rng(1)
    blocks = [(R.OFF_W1, R.OFF_B1), (R.OFF_B1, R.OFF_W2), (R.OFF_W2, R.OFF_B2), (R.OFF_B2, R.OFF_W3),
              (R.OFF_W3, R.OFF_B3), (R.OFF_B3, R.PARAMS)]
    h = 1e-6
    for a, b in blocks:
        for p in rng.integers(a, b, size=min(6, b - a)):
            if a == R.OFF_W1 and (p - a) // R.H >= R.TRUNK:
                assert g[p] == 0.0  # gate columns do not feed the trunk
                continue
            tp = theta.copy(); tp[p] += h
            tm = theta.copy(); tm[p] -= h
            fd = (R.loss(tp, X, t) - R.loss(tm, X, t)) / (2 * h)
            assert abs(fd - g[p]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (p, fd, g[p])


def test_bf16_gradient_close_to_exact():
    theta, X, t = _problem(M=64, seed=3)
    Xb = R.bf16(X).astype(np.float64)
    g64, y64 = R.gradient(theta, Xb, t, bf16_points=False)
    g16, y16 = R.gradient(theta, Xb, t, bf16_points=True)
    assert np.linalg.norm(y16 - y64) / np.linalg.norm(y64) < 2e-2
    assert np.linalg.norm(g16 - g64) / np.linalg.norm(g64) < 5e-2
