"""Pins the float64 MLP restatement (oracle/mlp_ref.py): its analytic
gradient equals central finite differences of its own loss, and a planted
error in the gradient (a dropped ReLU mask) is caught by the same check."""
import numpy as np

from oracle import mlp_ref as R


def _problem(seed=0, n=12):
    rng = np.random.default_rng(seed)
    prm = np.zeros(R.TOTAL)
    prm[:R.TRAINED] = rng.normal(0, 0.4, R.TRAINED)
    prm[R.FMU:R.FSIG] = rng.normal(0, 1, 16)
    prm[R.FSIG:R.YMU] = rng.uniform(0.5, 2, 16)
    prm[R.YMU:R.YSIG] = rng.normal(0, 1, 2)
    prm[R.YSIG:R.TOTAL] = rng.uniform(0.5, 2, 2)
    x = rng.normal(0, 1.5, (n, 16))
    y = rng.normal(0, 2, (n, 2))
    return prm, x, y


def _fd(prm, x, y, idx, h=1e-6):
    out = []
    for i in idx:
        a, b = prm.copy(), prm.copy()
        a[i] += h
        b[i] -= h
        out.append((R.loss_and_grad(a, x, y, half=False)[0] -
                    R.loss_and_grad(b, x, y, half=False)[0]) / (2 * h))
    return np.array(out)


def test_gradient_matches_finite_differences():
    prm, x, y = _problem()
    _, g = R.loss_and_grad(prm, x, y, half=False)
    rng = np.random.default_rng(1)
    idx = np.concatenate([rng.choice(np.arange(lo, hi), 6, replace=False)
                          for lo, hi in ((R.W1, R.B1), (R.B1, R.W2), (R.W2, R.B2), (R.B2, R.W3),
                                         (R.W3, R.B3))] + [np.arange(R.B3, R.TRAINED)])
    fd = _fd(prm, x, y, idx)
    assert np.allclose(g[idx], fd, rtol=1e-4, atol=1e-7)


def test_planted_mistake_is_caught():
    """Dropping the ReLU mask of the second layer changes the gradient of the
    first layer's weights beyond the finite-difference tolerance."""
    prm, x, y = _problem(seed=3)
    p = R.unpack(prm)
    z3, (_, xn, z1, h1, z2, h2) = R.forward(prm, x, half=False)
    n = len(x)
    e = z3 - (y - p["ymu"]) / p["ysig"]
    d3 = 2 * e / n
    d2 = d3 @ p["w3"].T  # mask dropped
    d1 = (d2 @ p["w2"].T) * (z1 > 0)
    wrong = (xn.T @ d1).ravel()
    idx = np.arange(R.W1, R.B1)[:64]
    fd = _fd(prm, x, y, idx)
    assert not np.allclose(wrong[idx - R.W1], fd, rtol=1e-4, atol=1e-7)
