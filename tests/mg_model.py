"""Float64 model of the multigrid V-cycle of the mixed solver (test
infrastructure, built from the oracle's assembled matrix with scipy).

Level 0: S = D^-1/2 K D^-1/2.  Level 1: the vertex-vertex part K_vv of K
(the trilinear subspace of the hierarchical H1 space), in lattice order
(x fastest); level 0 -> 1 restriction r_1 = D^1/2 r at the vertex
equations, prolongation x_0 += D^1/2 x_1 there.  Levels 2..: trilinear
P (kron of the 1D prolongations), A_c = P^T A P, while the level has more
than 128 unknowns and every element count is even.  Smoother: Chebyshev
iteration (Saad, Alg. 12.1) on M^-1 A with M the nstate x nstate diagonal
blocks, interval [1.1 lmax / range, 1.1 lmax]; m pre- and m post-steps
(m = degree on level 0, coarse_degree = 1 on the lattice levels);
dense solve on the coarsest level.  The largest eigenvalues are inputs
(the device estimates them by power iteration; tests bound them apart).
"""
import numpy as np
import scipy.sparse as sp

DENSE = 128
MAXDENSE = 1536  # beyond: the coarsest level gets a degree-8 Chebyshev iteration (range 100)


def full_from_upper(ia, ja, a, n):
    r = np.repeat(np.arange(n), np.diff(ia))
    U = sp.csr_matrix((a, (r, ja)), shape=(n, n))
    return (U + sp.triu(U, 1).T).tocsr()


def vertex_lattice(md, bs):
    """Equation of each lattice node's first component (x fastest) and the
    node counts per direction."""
    vn = np.asarray(md["vol_nodes"]).reshape(len(md["vol_nodes"]), -1)
    nc = vn.shape[1]  # corners (sides 0 .. nc-1)
    vc = np.asarray(md["vol_conn"]).reshape(len(md["vol_conn"]), -1)[:, :nc]
    xyz = np.asarray(md["xyz"]).reshape(-1, 3)
    nodes, first = np.unique(vn.ravel(), return_index=True)
    eq = np.asarray(md["blk_pos"])[vc.ravel()[first]]
    ids, dims = [], []
    for d in range(3):
        u, inv = np.unique(np.round(xyz[nodes, d], 9), return_inverse=True)
        ids.append(inv)
        dims.append(len(u))
    lat = ids[0] + dims[0] * (ids[1] + dims[1] * ids[2])
    assert len(np.unique(lat)) == len(lat) == np.prod(dims)
    veq = np.empty(len(lat), dtype=np.int64)
    veq[lat] = eq
    return veq, dims


def prolong_1d(nf):
    nc = (nf - 1) // 2 + 1
    P = sp.lil_matrix((nf, nc))
    for i in range(nf):
        if i % 2 == 0:
            P[i, i // 2] = 1.0
        else:
            P[i, i // 2] = 0.5
            P[i, i // 2 + 1] = 0.5
    return P.tocsr()


def block_diag_inv(A, bs):
    n = A.shape[0]
    blocks = np.zeros((n // bs, bs, bs))
    A = A.tocsr()
    for k in range(bs):
        for l in range(bs):
            blocks[:, k, l] = np.asarray(A[np.arange(k, n, bs), np.arange(l, n, bs)]).ravel()
    inv = np.linalg.inv(blocks)
    rows = np.repeat(np.arange(n), bs)
    cols = (np.arange(n)[:, None] // bs * bs + np.arange(bs)[None, :]).ravel()
    return sp.csr_matrix((inv.reshape(n // bs, bs, bs).reshape(-1), (rows, cols)), shape=(n, n))


class Model:
    def __init__(self, K, md, bs):
        self.bs = bs
        n = K.shape[0]
        d = K.diagonal()
        self.dsq = np.sqrt(d)
        Dis = sp.diags(1.0 / self.dsq)
        self.S = (Dis @ K @ Dis).tocsr()
        veq, dims = vertex_lattice(md, bs)
        self.veq = (veq[:, None] + np.arange(bs)[None, :]).ravel()
        A1 = K[self.veq][:, self.veq].tocsr()
        self.levels = [(A1, list(dims))]
        while True:
            A, dm = self.levels[-1]
            if A.shape[0] <= DENSE or any(x != 1 and ((x - 1) % 2 or x < 3) for x in dm):
                break
            Px, Py, Pz = (prolong_1d(x) for x in dm)
            P = sp.kron(sp.kron(Pz, Py), Px)
            P = sp.kron(P, sp.identity(bs)).tocsr()
            self.levels.append(((P.T @ A @ P).tocsr(), [1 if x == 1 else (x - 1) // 2 + 1 for x in dm]))
            self.levels[-2] = self.levels[-2] + (P,)
        self.M0 = block_diag_inv(self.S, bs)
        self.Ml = [block_diag_inv(A, bs) for A, *_ in self.levels]
        Ac = self.levels[-1][0]
        self.coarse_inv = np.linalg.inv(Ac.toarray()) if Ac.shape[0] <= MAXDENSE else None
        self.n = n

    def dims(self):
        return [lv[1] for lv in self.levels]

    @staticmethod
    def cheb(lmax, rng, m):
        hi = 1.1 * lmax
        lo = hi / rng
        theta, delta = 0.5 * (hi + lo), 0.5 * (hi - lo)
        sigma = theta / delta
        rho = 1.0 / sigma
        c1, c2 = [], []
        for _ in range(1, m):
            rn = 1.0 / (2 * sigma - rho)
            c1.append(rn * rho)
            c2.append(2 * rn / delta)
            rho = rn
        return 1.0 / theta, c1, c2

    @staticmethod
    def pre(A, M, b, co):
        c0, c1, c2 = co
        r = b.copy()
        d = c0 * (M @ b)
        x = d.copy()
        for a1, a2 in zip(c1, c2):
            r = r - A @ d
            d = a1 * d + a2 * (M @ r)
            x = x + d
        r = r - A @ d
        return x, r

    @staticmethod
    def post(A, M, b, x, co):
        c0, c1, c2 = co
        r = b - A @ x
        d = c0 * (M @ r)
        x = x + d
        for a1, a2 in zip(c1, c2):
            r = r - A @ d
            d = a1 * d + a2 * (M @ r)
            x = x + d
        return x

    def apply(self, r, lmax, degree=2, rng=30.0, coarse_degree=1):
        co0 = self.cheb(lmax[0], rng, degree)
        # smoothing coefficients of the levels above the coarsest (the coarsest
        # is solved directly or by its own polynomial below)
        cos = [self.cheb(lmax[1 + k], rng, coarse_degree) for k in range(len(self.levels) - 1)]
        x0, r0 = self.pre(self.S, self.M0, r, co0)
        b = [None] * len(self.levels)
        x = [None] * len(self.levels)
        b[0] = self.dsq[self.veq] * r0[self.veq]
        L = len(self.levels)
        for k in range(L - 1):
            A, _, P = self.levels[k]
            x[k], rk = self.pre(A, self.Ml[k], b[k], cos[k])
            b[k + 1] = P.T @ rk
        if self.coarse_inv is not None:
            x[L - 1] = self.coarse_inv @ b[L - 1]
        else:
            x[L - 1] = self.pre(self.levels[-1][0], self.Ml[-1], b[L - 1],
                                self.cheb(lmax[L], 100.0, 8))[0]
        for k in range(L - 2, -1, -1):
            A, _, P = self.levels[k]
            x[k] = x[k] + P @ x[k + 1]
            x[k] = self.post(A, self.Ml[k], b[k], x[k], cos[k])
        x0 = x0.copy()
        x0[self.veq] += self.dsq[self.veq] * x[0]
        return self.post(self.S, self.M0, r, x0, co0)

    def lmax_exact(self):
        """Largest eigenvalue of M^-1 A on level 0 and every lattice level."""
        import scipy.sparse.linalg as spla
        out = []
        for A, M in [(self.S, self.M0)] + [(lv[0], Ml) for lv, Ml in zip(self.levels, self.Ml)]:
            op = spla.LinearOperator(A.shape, matvec=lambda v, A=A, M=M: M @ (A @ v))
            out.append(float(np.max(np.real(spla.eigs(op, k=1, which="LM", tol=1e-8,
                                                      maxiter=20000)[0]))))
        return out
