"""CPU pins of the numpy reference of the learned preconditioner network
(oracle/lcc_ref.py): its analytic gradient against central finite differences,
and its structural properties (positive homogeneity, Wout = 0 gives zero)."""
import numpy as np

import lcc_ref as R


def _graph(n):
    # ring + chords, symmetric, no self loops
    rows = [sorted({(i - 1) % n, (i + 1) % n, (i + 3) % n, (i - 3) % n}) for i in range(n)]
    cia = np.zeros(n + 1, dtype=np.int64)
    for i, r in enumerate(rows):
        cia[i + 1] = cia[i] + len(r)
    cja = np.concatenate([np.array(r) for r in rows]).astype(np.int64)
    return cia, cja


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    bs, F, L, nc, ns = 3, 5, 3, 11, 2
    cia, cja = _graph(nc)
    theta = 0.4 * rng.standard_normal(R.nparams(bs, F, L))
    X = rng.standard_normal((ns, nc, bs))
    T = rng.standard_normal((ns, nc, bs))
    w = rng.uniform(0.5, 2.0, nc)
    f, g = R.loss_and_grad(theta, X, T, w, cia, cja, bs, F, L)
    h = 1e-6
    for k in rng.choice(len(theta), 40, replace=False):
        tp = theta.copy(); tp[k] += h
        tm = theta.copy(); tm[k] -= h
        fd = (R.loss_and_grad(tp, X, T, w, cia, cja, bs, F, L)[0] -
              R.loss_and_grad(tm, X, T, w, cia, cja, bs, F, L)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, g[k])


def test_homogeneity_and_zero_output():
    rng = np.random.default_rng(1)
    bs, F, L, nc = 1, 4, 3, 9
    cia, cja = _graph(nc)
    theta = rng.standard_normal(R.nparams(bs, F, L))
    X = rng.standard_normal((1, nc, bs))
    Z1, _ = R.forward(theta, X, cia, cja, bs, F, L)
    Z2, _ = R.forward(theta, 3.5 * X, cia, cja, bs, F, L)
    assert np.allclose(Z2, 3.5 * Z1)
    theta[-F * bs:] = 0.0
    Z0, _ = R.forward(theta, X, cia, cja, bs, F, L)
    assert np.all(Z0 == 0.0)
