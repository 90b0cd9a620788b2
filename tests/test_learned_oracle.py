"""The numpy statement of the learned element smoother (oracle/les_ref.py)
pinned by finite differences: every parameter block's gradient against
central differences of the loss, on a small random element graph with
random pairs and a partial dof mask."""
import numpy as np

import les_ref as R


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    ne, nel, nr = 6, 7, 3
    o = R.offsets(ne)
    theta = rng.standard_normal(o["n"]) * 0.3
    G = rng.standard_normal((ne, ne))
    theta[: ne * ne] = (0.5 * (G + G.T)).ravel()
    # ring graph with a chord, one isolated element
    nb = {0: [1, 5], 1: [0, 2], 2: [1, 3, 5], 3: [2, 4], 4: [3], 5: [0, 2], 6: []}
    cia = np.zeros(nel + 1, np.int64)
    cja = []
    for e in range(nel):
        cja += nb[e]
        cia[e + 1] = len(cja)
    cja = np.array(cja, np.int64)
    X = rng.standard_normal((nel, R.NF))
    Rr = rng.standard_normal((nel, nr, ne))
    E = rng.standard_normal((nel, nr, ne))
    Mk = (rng.random((nel, ne)) > 0.2).astype(float)
    loss, grad = R.loss_grad(theta, ne, X, cia, cja, Rr, E, Mk)
    h = 1e-6
    idx = list(range(ne * ne)) + list(rng.choice(np.arange(ne * ne, o["n"]), 60, replace=False))
    for i in idx:
        tp, tm = theta.copy(), theta.copy()
        if i < ne * ne:  # G is symmetric: perturb both (i, j) and (j, i)
            a, b = divmod(i, ne)
            for t, sg in ((tp, h), (tm, -h)):
                t[a * ne + b] += sg
                if a != b:
                    t[b * ne + a] += sg
            fd = (R.loss_grad(tp, ne, X, cia, cja, Rr, E, Mk)[0] - R.loss_grad(tm, ne, X, cia, cja, Rr, E, Mk)[0]) / (2 * h)
            an = grad[a * ne + b] * (2 if a != b else 1)
        else:
            tp[i] += h
            tm[i] -= h
            fd = (R.loss_grad(tp, ne, X, cia, cja, Rr, E, Mk)[0] - R.loss_grad(tm, ne, X, cia, cja, Rr, E, Mk)[0]) / (2 * h)
            an = grad[i]
        assert abs(fd - an) <= 1e-6 * max(1.0, abs(fd)), (i, fd, an)


def test_zero_output_layer_gives_unit_scales():
    rng = np.random.default_rng(1)
    ne, nel = 4, 5
    o = R.offsets(ne)
    theta = rng.standard_normal(o["n"])
    theta[o["w"]:o["w"] + R.NH] = 0.0
    cia = np.array([0, 1, 2, 3, 4, 4])
    cja = np.array([1, 0, 3, 2])
    z, s, _ = R.forward(theta, ne, rng.standard_normal((nel, R.NF)), cia, cja)
    assert np.all(z == 0) and np.all(s == 1)


def test_smoother_is_symmetric_and_local_inverse_exact():
    """With G the inverse of a single element's matrix and no overlap, the
    smoother inverts it; for any symmetric G the operator is symmetric."""
    rng = np.random.default_rng(2)
    ne = 5
    A = rng.standard_normal((ne, ne))
    A = A @ A.T + ne * np.eye(ne)
    dests = np.arange(ne)[None, :]
    pen = np.zeros(ne)
    r = rng.standard_normal(ne)
    z = R.smoother_apply(np.linalg.inv(A), np.ones(1), dests, pen, r)
    assert np.allclose(A @ z, r)
    # two overlapping elements, symmetric G: <z(a), b> = <a, z(b)>
    G = rng.standard_normal((3, 3))
    G = G + G.T
    dests = np.array([[0, 1, 2], [2, 3, 4]])
    pen = np.array([0, 0, 0, 0, 1.0])
    a, b = rng.standard_normal(5), rng.standard_normal(5)
    s = np.array([0.7, 1.3])
    za, zb = R.smoother_apply(G, s, dests, pen, a), R.smoother_apply(G, s, dests, pen, b)
    keep = pen == 0
    assert abs((za * b)[keep].sum() - (a * zb)[keep].sum()) < 1e-12
