"""Pins of the CPU oracle (oracle/pz_oracle.c) -- CPU only.

Each oracle function is checked against something other than itself:

=====================  =====================================================
function               pinned by
=====================  =====================================================
orc_gauss_legendre     Line.txt known integrals (reference test data), point
                       count rule, symmetry, weight sum
orc_shape              closed forms (corner Kronecker property, edge/face/
                       volume bubbles 4x0x1, 16.., 64..), finite differences
orc_contribute         CubeStiffMatrix.txt (elasticity), TestDarcyFlow matrix,
                       row sums / rigid-body null space
orc_calcstiff          Jacobian: element matrices of a scaled/shifted element
                       against the closed-form h-scaling; linear patch tests
orc_mesh_build         hand-derived first-touch numbering of a 2x1 mesh, the
                       closed-form equation count
orc_setup_pattern      brute-force pattern from element dof sets, closed-form
                       nnz, the 1-D known matrix of StructMatrixUnitTest
orc_addkel             dense sum of element matrices; the 1e-12 drop rule
orc_multadd            dense multiply (TestMatrix.cpp:1072-1106 analogue)
orc_pcg                scipy direct solve; exact quadratic reproduction
=====================  =====================================================
"""
import json
import os

import numpy as np
import pytest

import oracle as O
from neopz_amd import problems as P

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _load(name):
    with open(os.path.join(GOLD, name)) as f:
        return json.load(f)


# --------------------------------------------------------------------------- quadrature
def test_gauss_line_known_integrals():
    vals = _load("line_integrals.json")["values"]
    for order in range(0, 31):
        x, w = O.gauss_legendre(order)
        f = sum((i + 1) * (x + 0.1) ** i for i in range(order + 1))
        got = float(np.dot(w, f))
        ref = float(vals[order])
        tol = 6e-13
        if ref > 10:
            tol *= 10
        if ref > 100:
            tol *= 10
        if ref > 1000:
            tol *= 10
        assert abs(got - ref) < tol, (order, got, ref)


def test_triangle_rules_known_integrals():
    """The simplex rule of each order p integrates the reference test's
    polynomial of total degree p over the triangle to the values of
    UnitTest_PZ/TestIntegNum/Triangle.txt, with that test's tolerance."""
    vals = _load("triangle_integrals.json")["values"]
    for p in range(0, 32):
        pts, w = O.simplex_rule(O.TRI, p)
        x, y = pts[:, 0], pts[:, 1]
        f = sum((i + 1) * (x + 0.1) ** i * (j + 1) * (y + 0.1) ** j
                for i in range(p + 1) for j in range(p + 1 - i))
        got, ref = float(np.dot(w, f)), float(vals[p])
        tol = 6e-13 * (10 if ref > 10 else 1) * (10 if ref > 100 else 1) * (10 if ref > 1000 else 1)
        assert abs(got - ref) < tol, (p, got, ref)


def test_gauss_rule_structure():
    for order in range(0, 20):
        x, w = O.gauss_legendre(order)
        n = max(1, int(0.51 * (order + 2)))
        assert len(x) == n
        assert abs(w.sum() - 2.0) < 1e-14
        # (-z, +z) pairs, |z| descending, centre point last for odd n
        for i in range(n // 2):
            assert x[2 * i] == -x[2 * i + 1] and x[2 * i + 1] > 0
            assert w[2 * i] == w[2 * i + 1]
        assert np.all(np.diff(np.abs(x[::2])) <= 0)
        # exact for degree 2n-1 against the closed form of int x^k
        for k in range(2 * n):
            exact = 0.0 if k % 2 else 2.0 / (k + 1)
            assert abs(np.dot(w, x ** k) - exact) < 1e-13


# --------------------------------------------------------------------------- shapes
def _corners(etype):
    if etype == O.LINE:
        return np.array([[-1.0], [1.0]])
    if etype == O.QUAD:
        return np.array([[-1, -1], [1, -1], [1, 1], [-1, 1]], float)
    return np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                     [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], float)


@pytest.mark.parametrize("etype", [O.LINE, O.QUAD, O.HEX])
def test_shape_vertex_kronecker_and_bubbles(etype):
    cs = _corners(etype)
    nc = len(cs)
    for p in (1, 2):
        for i, c in enumerate(cs):
            phi, _ = O.shape(etype, p, c)
            e = np.zeros(len(phi))
            e[i] = 1.0
            assert np.allclose(phi, e, atol=1e-15)  # bubbles vanish at vertices
    # p=2 closed forms: 4 x0 x1 (times linear in other dirs), 16.., 64..
    rng = np.random.default_rng(1)
    for _ in range(5):
        xi = rng.uniform(-1, 1, size=cs.shape[1])
        phi, _ = O.shape(etype, 2, xi)
        l0 = (1 - xi) / 2
        l1 = (1 + xi) / 2
        if etype == O.LINE:
            assert np.isclose(phi[2], 4 * l0[0] * l1[0])
        elif etype == O.QUAD:
            # edge 4 = nodes (0,1): y = -1 side
            assert np.isclose(phi[4], 4 * l0[0] * l1[0] * l0[1])
            assert np.isclose(phi[8], 16 * l0[0] * l1[0] * l0[1] * l1[1])
        else:
            assert np.isclose(phi[8], 4 * l0[0] * l1[0] * l0[1] * l0[2])       # edge 0-1
            assert np.isclose(phi[8 + 4], 4 * l0[0] * l0[1] * l0[2] * l1[2])   # edge 0-4
            assert np.isclose(phi[20], 16 * l0[0] * l1[0] * l0[1] * l1[1] * l0[2])  # face z=-1
            assert np.isclose(phi[26], 64 * np.prod(l0 * l1))
    # unitary bubbles at edge midpoints
    if etype == O.QUAD:
        phi, _ = O.shape(etype, 2, np.array([0.0, -1.0]))
        assert np.isclose(phi[4], 1.0)
        phi, _ = O.shape(etype, 2, np.array([0.0, 0.0]))
        assert np.isclose(phi[8], 1.0)
    assert nc == len(O.shape(etype, 1, cs[0])[0])


@pytest.mark.parametrize("etype", [O.LINE, O.QUAD, O.HEX])
def test_shape_derivatives_finite_difference(etype):
    rng = np.random.default_rng(2)
    dim = {O.LINE: 1, O.QUAD: 2, O.HEX: 3}[etype]
    for _ in range(4):
        xi = rng.uniform(-0.9, 0.9, size=dim)
        _, dphi = O.shape(etype, 2, xi)
        h = 1e-6
        for d in range(dim):
            e = np.zeros(dim)
            e[d] = h
            fd = (O.shape(etype, 2, xi + e)[0] - O.shape(etype, 2, xi - e)[0]) / (2 * h)
            assert np.allclose(dphi[d], fd, atol=1e-8)


# --------------------------------------------------------------------------- materials
def test_elasticity_cube_stiffness_golden():
    g = _load("cube_stiff_matrix.json")
    ek, _ = O.contribute(O.ELAST3D, 3, g["weight"], np.zeros(3), np.eye(3), E=g["E"], nu=g["nu"])
    assert np.allclose(ek, ek.T, atol=1e-8)
    assert np.max(np.abs(ek - np.array(g["ek"]))) < g["margin"]


def test_darcy_q1_golden():
    g = _load("darcy_q1_matrix.json")
    x, w = O.gauss_legendre(2)  # TPZShapeQuad::IntruleType integrule(2)
    ek = np.zeros((4, 4))
    ef = np.zeros(4)
    for ik in range(len(x)):
        for ie in range(len(x)):
            phi, dphi = O.shape(O.QUAD, 1, np.array([x[ik], x[ie]]))
            O.contribute(O.DARCY, 2, w[ik] * w[ie], phi, dphi, f=(0, 0, 0), kappa=1.0, ek=ek, ef=ef)
    assert np.max(np.abs(ek - np.array(g["ek"]))) < g["margin"]


def _rigid_modes(xyz):
    """6 rigid-body modes evaluated at points (n,3) -> (6, 3n) shape-major."""
    modes = []
    for k in range(3):
        u = np.zeros_like(xyz)
        u[:, k] = 1.0
        modes.append(u.ravel())
    for w in np.eye(3):
        modes.append(np.cross(w, xyz).ravel())
    return np.array(modes)


def test_element_invariants():
    # Poisson: the constant function (vertex coefficients 1, bubbles 0) is in the
    # null space; elasticity: rigid modes are
    pb = P.Problem(material=P.POISSON, dim=2, p=2, nx=3, ny=2)
    o = O.Oracle(pb)
    for e in range(o.m.nvol):
        ek, _, _ = o.calcstiff(e)
        assert np.allclose(ek[:, :4].sum(axis=1), 0.0, atol=1e-13)
        assert np.allclose(ek, ek.T, atol=1e-14)
    pb = P.Problem(material=P.ELAST3D, dim=3, p=2, nx=2, ny=2, nz=1, E=3.0, nu=0.3,
                   el_coef=np.array([1.0, 2.0, 0.5, 4.0]))
    o = O.Oracle(pb)
    for e in range(o.m.nvol):
        ek, _, dest = o.calcstiff(e)
        # coefficient vector of a linear field: vertex values, zero bubbles
        xyz = np.array([o.m.xyz[3 * o.m.enodes[8 * e + n] + d] for n in range(8) for d in range(3)]).reshape(8, 3)
        for mode in _rigid_modes(xyz):
            u = np.zeros(81)
            u[:24] = mode
            assert np.max(np.abs(ek @ u)) < 1e-12 * np.max(np.abs(ek))


def test_jacobian_scaling():
    # 3D elasticity element matrices scale with h (E h); 2D Poisson is h-invariant
    for n in (1, 3):
        pb = P.Problem(material=P.ELAST3D, dim=3, p=2, nx=n, ny=n, nz=n, E=1.0, nu=0.25)
        ek, _, _ = O.Oracle(pb).calcstiff(0)
        if n == 1:
            ek1 = ek
        else:
            assert np.allclose(ek * 3.0, ek1, rtol=1e-13, atol=1e-15)
    for n in (1, 4):
        pb = P.Problem(material=P.POISSON, dim=2, p=2, nx=n, ny=n)
        ek, _, _ = O.Oracle(pb).calcstiff(0)
        if n == 1:
            ek1 = ek
        else:
            assert np.allclose(ek, ek1, rtol=1e-13, atol=1e-15)


# --------------------------------------------------------------------------- numbering / pattern
def test_first_touch_numbering_2x1_quads():
    pb = P.Problem(material=P.POISSON, dim=2, p=2, nx=2, ny=1)
    o = O.Oracle(pb)
    econn, pos, size = o.connects()
    assert list(econn[0, :9]) == list(range(9))
    # element 1 = nodes (n1, n2, n5, n4): shares corner n1 (c1), corner n4 (c2), edge (n1,n4) (c5)
    assert list(econn[1, :9]) == [1, 9, 10, 2, 11, 12, 13, 5, 14]
    assert o.neq == 15 == pb.n_equations()
    assert list(pos) == list(range(15)) and list(size) == [1] * 15


@pytest.mark.parametrize("dim,p,nstate,n", [(2, 1, 1, (3, 2, 1)), (2, 2, 1, (4, 3, 1)),
                                            (3, 1, 3, (2, 2, 2)), (3, 2, 3, (3, 2, 2)),
                                            (3, 2, 1, (2, 3, 2))])
def test_pattern_brute_force_and_closed_form(dim, p, nstate, n):
    mat = P.ELAST3D if nstate == 3 else P.POISSON
    pb = P.Problem(material=mat, dim=dim, p=p, nx=n[0], ny=n[1], nz=n[2],
                   bc_type=[0, 1, -1, -1, -1, -1])
    o = O.Oracle(pb)
    ia, ja = o.pattern()
    assert o.neq == pb.n_equations()
    assert len(ja) == O.nnz_closed_form(dim, p, nstate, *n)
    # brute force: union over elements of dof pairs, upper part, sorted rows
    rows = [set() for _ in range(o.neq)]
    for e in range(o.m.nel):
        _, _, dest = o.calcstiff(e)
        for i in dest:
            for j in dest:
                if j >= i:
                    rows[i].add(j)
    exp_ia = np.zeros(o.neq + 1, dtype=np.int64)
    exp_ja = []
    for r in range(o.neq):
        cols = sorted(rows[r])
        exp_ja += cols
        exp_ia[r + 1] = exp_ia[r] + len(cols)
    assert np.array_equal(ia, exp_ia)
    assert np.array_equal(ja, np.array(exp_ja))


def test_known_matrix_1d_mock():
    """UnitTest_PZ/TestStruct/StructMatrixUnitTest.cpp:183-213."""
    pb = P.mock1d(20)
    o = O.Oracle(pb)
    ia, ja = o.pattern()
    a, rhs = o.assemble()
    K = O.upper_to_dense(ia, ja, a)
    n = 21
    assert o.neq == n
    assert K[0, 0] == 2 and K[n - 1, n - 1] == 2 and K[0, 1] == 1 and K[n - 1, n - 2] == 1
    for i in range(1, n - 1):
        assert K[i, i - 1] == 1 and K[i, i] == 4 and K[i, i + 1] == 1


# --------------------------------------------------------------------------- assembly
def test_addkel_drop_rule_and_dense_sum():
    ia = np.array([0, 2, 3], dtype=np.int64)
    ja = np.array([0, 1, 1], dtype=np.int64)
    a = np.zeros(3)
    ek = np.array([[1.0, 5e-13], [5e-13, 2e-12]])
    O.addkel(ia, ja, a, ek, np.array([0, 1]))
    assert a[0] == 1.0 and a[1] == 0.0 and a[2] == 2e-12
    # a mesh: CSR assembly equals the dense sum of element matrices
    pb = P.poisson2d(n=3, p=2, seed=3)
    o = O.Oracle(pb)
    ia, ja = o.pattern()
    a, rhs = o.assemble()
    Kd = np.zeros((o.neq, o.neq))
    fd = np.zeros(o.neq)
    for e in range(o.m.nel):
        ek, ef, dest = o.calcstiff(e)
        ek = np.where(np.abs(ek) < 1e-12, 0.0, ek)
        Kd[np.ix_(dest, dest)] += ek
        fd[dest] += ef
    K = O.upper_to_dense(ia, ja, a)
    assert np.allclose(K, Kd, rtol=1e-14, atol=1e-14 * np.abs(Kd).max())
    assert np.allclose(rhs, fd, rtol=1e-14, atol=1e-300)
    # Dirichlet penalty: fBigNumber = pow(10, 17) * 2 / 3 times the boundary mass
    assert K.max() > 1e15


def _coeffs_of_linear(o, fun, nstate):
    """Coefficient vector of a linear field: vertex dofs = values, bubbles = 0."""
    u = np.zeros(o.neq)
    econn, pos, size = o.connects()
    for e in range(o.m.nvol):
        for n in range(8 if o.m.dim == 3 else 4):
            node = o.m.enodes[8 * e + n]
            x = np.array([o.m.xyz[3 * node + d] for d in range(3)])
            c = econn[e, n]
            u[pos[c]:pos[c] + size[c]] = fun(x)[:nstate]
    return u


def test_linear_patch_tests():
    # Poisson: interior residual of a linear field vanishes (no BC elements)
    pb = P.Problem(material=P.POISSON, dim=2, p=2, nx=4, ny=3)
    o = O.Oracle(pb)
    ia, ja = o.pattern()
    a, _ = o.assemble()
    u = _coeffs_of_linear(o, lambda x: np.array([1.0 + 2 * x[0] - 3 * x[1]]), 1)
    r = O.multadd(ia, ja, a, u)
    econn, pos, size = o.connects()
    interior = [e for e in range(o.m.nvol)
                if 0 < e % 4 < 3 and 0 < e // 4 < 2]
    for e in interior:
        c = econn[e, 0]
        assert abs(r[pos[c]]) < 1e-13
    # elasticity, constant E: every rigid mode is in the null space of the global K
    pb = P.Problem(material=P.ELAST3D, dim=3, p=2, nx=2, ny=2, nz=2, E=2.0, nu=0.3)
    o = O.Oracle(pb)
    ia, ja = o.pattern()
    a, _ = o.assemble()
    for k in range(6):
        def f(x, k=k):
            return _rigid_modes(x[None, :])[k]
        u = _coeffs_of_linear(o, f, 3)
        r = O.multadd(ia, ja, a, u)
        assert np.max(np.abs(r)) < 1e-12 * np.abs(a).max()


def test_multadd_matches_dense():
    pb = P.elasticity3d(n=2, p=2, seed=4)
    o = O.Oracle(pb)
    ia, ja = o.pattern()
    a, _ = o.assemble()
    K = O.upper_to_dense(ia, ja, a)
    rng = np.random.default_rng(0)
    x = rng.standard_normal(o.neq)
    y = rng.standard_normal(o.neq)
    z = O.multadd(ia, ja, a, x, y, alpha=0.7, beta=-1.3)
    assert np.allclose(z, 0.7 * K @ x - 1.3 * y, rtol=1e-13, atol=1e-13 * np.abs(K @ x).max())


def test_pcg_matches_direct_and_quadratic_exactness():
    import scipy.sparse.linalg as spla
    import scipy.sparse as sp
    # -u'' = 2 on [0,1] in x, u(0)=u(1)=0, natural BC in y: u = x(1-x) is in Q2
    n = 4
    pb = P.Problem(material=P.POISSON, dim=2, p=2, nx=n, ny=2, force_kind=P.FORCE_CONST,
                   force_const=(2.0, 0, 0), force_porder=0, bc_type=[0, 0, -1, -1, -1, -1])
    o = O.Oracle(pb)
    ia, ja = o.pattern()
    a, rhs = o.assemble()
    res = O.pcg(ia, ja, a, rhs, tol=1e-14)
    K = O.upper_to_dense(ia, ja, a)
    xd = np.linalg.solve(K, rhs)
    assert np.allclose(res["x"], xd, rtol=1e-8, atol=1e-10)
    econn, pos, size = o.connects()
    h = 1.0 / n
    for e in range(o.m.nvol):
        for s in range(9):
            c = econn[e, s]
            v = xd[pos[c]]
            if s < 4:
                node = o.m.enodes[8 * e + s]
                x = o.m.xyz[3 * node]
                assert abs(v - x * (1 - x)) < 1e-12
            elif s in (4, 6):   # x-directed edges: bubble coefficient h^2/4
                assert abs(v - h * h / 4) < 1e-12
            else:               # y-directed edges and the face bubble: zero
                assert abs(v) < 1e-12
    # larger random problem vs sparse direct
    pb = P.poisson2d(n=8, p=2, seed=5)
    o = O.Oracle(pb)
    ia, ja = o.pattern()
    a, rhs = o.assemble()
    res = O.pcg(ia, ja, a, rhs, tol=1e-13)
    assert res["converged"]
    rows = np.repeat(np.arange(o.neq), np.diff(ia))
    U = sp.csr_matrix((a, (rows, ja)), shape=(o.neq, o.neq))
    Kf = U + sp.triu(U, 1).T
    xd = spla.spsolve(Kf.tocsc(), rhs)
    assert np.linalg.norm(res["x"] - xd) <= 1e-8 * np.linalg.norm(xd)


def test_pcg_block_jacobi_special_cases():
    """orc_pcg_block (EBlockJacobi with overlap) pinned by what the algebra
    fixes: blocks of one equation are point Jacobi (the same iterates as
    orc_pcg, bit for bit); one block holding every equation is the exact
    inverse (one iteration reaches the direct solution); on a matrix that is
    block diagonal in blocks of mixed sizes each block is solved exactly (one
    iteration), which fails for a wrong block offset, a transposed factor or a
    dropped substitution term."""
    pb = P.elasticity3d(n=3, p=2, seed=4, field_kind="grf", load="body")
    o = O.Oracle(pb)
    ia, ja = o.pattern()
    a, rhs = o.assemble()
    n = o.neq
    r1 = O.pcg(ia, ja, a, rhs, maxit=5000, tol=1e-12, wasted_spmv=False, history=True)
    r2 = O.pcg_block(ia, ja, a, rhs, np.ones(n, np.int32), maxit=5000, tol=1e-12, history=True)
    assert r1["iters"] == r2["iters"] and np.array_equal(r1["x"], r2["x"])
    assert np.array_equal(r1["history"], r2["history"])
    # the connect blocks (nstate equations) converge, to the same solution
    r3 = O.pcg_block(ia, ja, a, rhs, np.full(n // 3, 3, np.int32), maxit=5000, tol=1e-12)
    assert r3["converged"] and np.linalg.norm(r3["x"] - r1["x"]) <= 1e-9 * np.linalg.norm(r1["x"])
    # one block: the exact solve
    pb = P.poisson2d(n=4, p=2, seed=2)
    o = O.Oracle(pb)
    ia, ja = o.pattern()
    a, rhs = o.assemble()
    K = O.upper_to_dense(ia, ja, a)
    xd = np.linalg.solve(K, rhs)
    r = O.pcg_block(ia, ja, a, rhs, [o.neq], maxit=10, tol=1e-12)
    assert r["iters"] == 1 and np.allclose(r["x"], xd, rtol=1e-10, atol=1e-12 * np.abs(xd).max())
    # block diagonal SPD matrix, blocks of sizes 1..4
    rng = np.random.default_rng(3)
    sizes = rng.integers(1, 5, 40).astype(np.int32)
    nn = int(sizes.sum())
    D = np.zeros((nn, nn))
    e = 0
    for m in sizes:
        B = rng.standard_normal((m, m))
        D[e:e + m, e:e + m] = B @ B.T + m * np.eye(m)
        e += m
    iu, ju = np.triu_indices(nn)
    keep = D[iu, ju] != 0
    iu, ju, av = iu[keep], ju[keep], D[iu, ju][keep]
    ia_d = np.concatenate([[0], np.cumsum(np.bincount(iu, minlength=nn))]).astype(np.int64)
    b = rng.standard_normal(nn)
    r = O.pcg_block(ia_d, ju.astype(np.int64), av, b, sizes, maxit=10, tol=1e-13)
    assert r["iters"] == 1 and np.allclose(r["x"], np.linalg.solve(D, b), rtol=1e-11, atol=1e-13)
    with pytest.raises(ValueError):
        O.pcg_block(ia_d, ju.astype(np.int64), av, b, sizes[:-1])


# --------------------------------------------------------------------------- hp (min rule)
def _hp_problem(kind, seed):
    rng = np.random.default_rng(seed)
    if kind == "quad":
        pb = P.Problem(material=P.POISSON, dim=2, p=2, nx=5, ny=4, bc_type=[0, 1, 1, 1, -1, -1],
                       bc_val=[[0, 0, 0], [1.0, 0, 0], [0, 0, 0], [0, 0, 0], [0, 0, 0], [0, 0, 0]])
    else:
        pb = P.Problem(material=P.POISSON, dim=3, p=2, nx=3, ny=3, nz=2, bc_type=[0, 1, 1, 1, 1, 1],
                       bc_val=[[0, 0, 0], [1.0, 0, 0], [0, 0, 0], [0, 0, 0], [0, 0, 0], [0, 0, 0]])
    pb.el_porder = rng.integers(1, 3, pb.nvol).astype(np.int32)
    return pb


@pytest.mark.parametrize("kind", ["quad", "hex"])
def test_hp_min_rule_pattern_and_sizes(kind):
    """hp orders: a side's connect has shapes iff every element sharing it
    prefers order 2 (pzintel.cpp:1370-1390); element dof sets follow; the CSR
    pattern equals the brute-force union of element dof pairs."""
    pb = _hp_problem(kind, 3)
    o = O.Oracle(pb)
    econn, pos, size = o.connects()
    nvol = pb.nvol
    ns = 9 if pb.dim == 2 else 27
    nc = 4 if pb.dim == 2 else 8
    # expected connect sizes from the rule, recomputed here from element lists
    mins = {}
    for e in range(nvol):
        for s in range(nc, ns):
            c = econn[e, s]
            mins[c] = min(mins.get(c, 2), int(pb.el_porder[e]))
    for c, pmin in mins.items():
        assert size[c] == (1 if pmin >= 2 else 0)
    assert o.neq == int(size.sum())
    # independent of the formula: an element of order 1 spans exactly the
    # (bi/tri)linear functions, and an order-2 element whose neighbours are all
    # of order 2 keeps its full quadratic set
    nx, ny, nz = pb.nx, pb.ny, (pb.nz if pb.dim == 3 else 1)
    po = np.asarray(pb.el_porder).reshape(nz, ny, nx)
    for e in range(nvol):
        nd = sum(size[c] for c in econn[e, :ns])
        iz, iy, ix = e // (nx * ny), (e // nx) % ny, e % nx
        nb = po[max(iz - 1, 0):iz + 2, max(iy - 1, 0):iy + 2, max(ix - 1, 0):ix + 2]
        if pb.el_porder[e] == 1:
            assert nd == nc
        elif nb.min() == 2:
            assert nd == ns
        else:
            assert nc < nd <= ns  # the interior shape is always there
    ia, ja = o.pattern()
    rows = [set() for _ in range(o.neq)]
    for e in range(o.m.nel):
        _, _, dest = o.calcstiff(e)
        for i in dest:
            for j in dest:
                if j >= i:
                    rows[i].add(j)
    exp = [sorted(r) for r in rows]
    assert np.array_equal(np.diff(ia), [len(r) for r in exp])
    assert np.array_equal(ja, np.concatenate([np.array(r, dtype=np.int64) for r in exp]))


@pytest.mark.parametrize("kind", ["quad", "hex"])
def test_hp_patch_test_linear_solution(kind):
    """u = x solves -lap u = 0 with u = 0 on x=0, du/dn = 1 on x=1 and 0 on the
    other faces.  A conforming hp space reproduces it exactly (corner values =
    x, hierarchical coefficients = 0); a wrong side order (a shape kept on one
    side of an interface only) breaks conformity and the patch test."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    pb = _hp_problem(kind, 5)
    o = O.Oracle(pb)
    ia, ja = o.pattern()
    a, rhs = o.assemble()
    rows = np.repeat(np.arange(o.neq), np.diff(ia))
    U = sp.csr_matrix((a, (rows, ja)), shape=(o.neq, o.neq))
    K = (U + sp.triu(U, 1).T).tocsc()
    u = spla.spsolve(K, rhs)
    econn, pos, size = o.connects()
    enodes = np.ctypeslib.as_array(o.m.enodes, shape=(8 * o.m.nel,)).reshape(o.m.nel, 8)
    xyz = np.ctypeslib.as_array(o.m.xyz, shape=(3 * o.m.nnodes,)).reshape(-1, 3)
    nc = 4 if pb.dim == 2 else 8
    ns = 9 if pb.dim == 2 else 27
    for e in range(pb.nvol):
        for k in range(nc):
            c = econn[e, k]
            assert abs(u[pos[c]] - xyz[enodes[e, k], 0]) < 1e-9
        for s in range(nc, ns):
            c = econn[e, s]
            if size[c]:
                assert abs(u[pos[c]]) < 1e-9


def test_permuted_numbering_is_a_similarity():
    """Renumbering the connects (what TPZCompMesh::Permute does after a
    bandwidth optimisation) permutes K and f symmetrically: K' = P K P^T."""
    pb = P.Problem(material=P.POISSON, dim=2, p=2, nx=4, ny=3, bc_type=[0, 1, 0, 1, -1, -1],
                   force_kind=P.FORCE_CONST, force_const=(1.0, 0, 0))
    o = O.Oracle(pb)
    ia, ja = o.pattern()
    a, rhs = o.assemble()
    K = O.upper_to_dense(ia, ja, a)
    econn, pos, size = o.connects()
    perm = np.random.default_rng(0).permutation(o.m.nconn)
    o2 = O.Oracle(pb)
    o2.permute(perm)
    ia2, ja2 = o2.pattern()
    a2, rhs2 = o2.assemble()
    K2 = O.upper_to_dense(ia2, ja2, a2)
    _, pos2, size2 = o2.connects()
    # equation map old -> new (one equation per connect here)
    emap = np.empty(o.neq, dtype=np.int64)
    for s_old in range(o.m.nconn):
        for k in range(size[s_old]):
            emap[pos[s_old] + k] = pos2[perm[s_old]] + k
    assert np.array_equal(size2[perm], size)
    assert np.allclose(K2[np.ix_(emap, emap)], K, rtol=0, atol=1e-9 * np.abs(K).max())
    assert np.allclose(rhs2[emap], rhs, rtol=0, atol=1e-12 * np.abs(rhs).max())


# --------------------------------------------------------------------------- simplices
@pytest.mark.parametrize("etype,order,deg", [(O.TRI, 2, 2), (O.TRI, 4, 4), (O.TET, 2, 2), (O.TET, 4, 5)])
def test_simplex_rules_integrate_monomials(etype, order, deg):
    """Exact moments: int x^a y^b (z^c) over the unit simplex = a! b! c! / (a+b+c+dim)!"""
    from math import factorial
    pts, w = O.simplex_rule(etype, order)
    assert (w > 0).all()
    dim = 2 if etype == O.TRI else 3
    for a in range(deg + 1):
        for b in range(deg + 1 - a):
            for c in range(deg + 1 - a - b if dim == 3 else 1):
                v = np.sum(w * pts[:, 0] ** a * pts[:, 1] ** b * pts[:, 2] ** c)
                ex = factorial(a) * factorial(b) * factorial(c) / factorial(a + b + c + dim)
                assert abs(v - ex) < 1e-15


@pytest.mark.parametrize("etype", [O.TRI, O.TET])
def test_simplex_shapes(etype):
    """Corners are barycentric coordinates; edge functions are 4*la*lb (value 1 at
    the edge midpoint, 0 at the corners and on the other edges); derivatives
    agree with finite differences."""
    dim = 2 if etype == O.TRI else 3
    nc = dim + 1
    edges = [(0, 1), (1, 2), (2, 0)] if dim == 2 else [(0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3)]
    corner = np.vstack([np.zeros(dim), np.eye(dim)])
    for k in range(nc):
        phi, _ = O.shape(etype, 2, corner[k])
        assert np.allclose(phi[:nc], np.eye(nc)[k]) and np.allclose(phi[nc:], 0)
    for e, (a, b) in enumerate(edges):
        phi, _ = O.shape(etype, 2, 0.5 * (corner[a] + corner[b]))
        assert abs(phi[nc + e] - 1.0) < 1e-15
        assert np.allclose(np.delete(phi[nc:], e), 0)
    rng = np.random.default_rng(0)
    x = rng.random(dim) / dim
    phi, dphi = O.shape(etype, 2, x)
    assert abs(phi[:nc].sum() - 1) < 1e-15
    h = 1e-6
    for d in range(dim):
        xp = x.copy(); xp[d] += h
        xm = x.copy(); xm[d] -= h
        fd = (O.shape(etype, 2, xp)[0] - O.shape(etype, 2, xm)[0]) / (2 * h)
        assert np.allclose(dphi[d], fd, atol=1e-8)


def _tet_problem(p, bc, bcval, force=0.0, n=(2, 2, 2)):
    return P.Problem(material=P.POISSON, dim=3, p=p, nx=n[0], ny=n[1], nz=n[2], simplex=1,
                     force_kind=P.FORCE_CONST if force else P.FORCE_NONE,
                     force_const=(force, 0, 0), bc_type=bc,
                     bc_val=[[v, 0, 0] for v in bcval])


def _solve_oracle(o):
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    ia, ja = o.pattern()
    a, rhs = o.assemble()
    rows = np.repeat(np.arange(o.neq), np.diff(ia))
    U = sp.csr_matrix((a, (rows, ja)), shape=(o.neq, o.neq))
    return spla.spsolve((U + sp.triu(U, 1).T).tocsc(), rhs)


@pytest.mark.parametrize("p", [1, 2])
def test_tet_patch_tests(p):
    """u = x (Dirichlet 0 at x=0, flux 1 at x=1) for p = 1, 2, and for p = 2
    u = x(1-x)/2 (-lap u = 1, u = 0 at x = 0 and x = 1): corner dofs equal u
    at the nodes; edge dofs equal u(mid) - (u_a + u_b)/2, NeoPZ's edge
    function being 1 at the midpoint."""
    cases = [(lambda X: X, [0, 1, 1, 1, 1, 1], [0, 1, 0, 0, 0, 0], 0.0)]
    if p == 2:
        cases.append((lambda X: X * (1 - X) / 2, [0, 0, 1, 1, 1, 1], [0] * 6, 1.0))
    for u_ex, bc, bcv, f in cases:
        o = O.Oracle(_tet_problem(p, bc, bcv, f))
        u = _solve_oracle(o)
        econn, pos, size = o.connects()
        enodes = np.ctypeslib.as_array(o.m.enodes, shape=(8 * o.m.nel,)).reshape(o.m.nel, 8)
        xyz = np.ctypeslib.as_array(o.m.xyz, shape=(3 * o.m.nnodes,)).reshape(-1, 3)
        edges = [(0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3)]
        for e in range(o.m.nvol):
            for k in range(4):
                assert abs(u[pos[econn[e, k]]] - u_ex(xyz[enodes[e, k], 0])) < 1e-9
            for j, (a, b) in enumerate(edges):
                c = econn[e, 4 + j]
                if size[c]:
                    xa, xb = xyz[enodes[e, a], 0], xyz[enodes[e, b], 0]
                    exp = u_ex(0.5 * (xa + xb)) - 0.5 * (u_ex(xa) + u_ex(xb))
                    assert abs(u[pos[c]] - exp) < 1e-9


def test_tet_pattern_brute_force_and_count():
    """P2 on the Kuhn mesh of n^3 cubes has (2n+1)^3 equations (vertices + edge
    midpoints); the pattern is the union of element dof pairs."""
    n = (2, 3, 2)
    o = O.Oracle(_tet_problem(2, [0] * 6, [0] * 6, 1.0, n))
    assert o.neq == (2 * n[0] + 1) * (2 * n[1] + 1) * (2 * n[2] + 1)
    ia, ja = o.pattern()
    rows = [set() for _ in range(o.neq)]
    for e in range(o.m.nel):
        _, _, dest = o.calcstiff(e)
        for i in dest:
            for j in dest:
                if j >= i:
                    rows[i].add(j)
    assert np.array_equal(np.diff(ia), [len(r) for r in rows])
    assert np.array_equal(ja, np.concatenate([np.array(sorted(r), dtype=np.int64) for r in rows]))


# --------------------------------------------------------------------------- simplices, p > 2
def test_collapsed_rules_exact_and_jacobi_points():
    """Collapsed rules of n^dim points integrate every monomial of degree
    <= 2n-1 over the unit simplex; their 1-D factors are the Gauss-Jacobi
    points of weights 1 and (1-b) (scipy's roots_jacobi on [-1, 1])."""
    from math import factorial
    from scipy.special import roots_jacobi
    for n in range(1, 8):
        for etype, dim in ((O.TRI, 2), (O.TET, 3)):
            pts, w = O.collapsed_rule(etype, n)
            assert len(w) == n ** dim and (w > 0).all()
            deg = 2 * n - 1
            for a in range(deg + 1):
                for b in range(deg + 1 - a):
                    for c in range(deg + 1 - a - b if dim == 3 else 1):
                        v = np.sum(w * pts[:, 0] ** a * pts[:, 1] ** b * pts[:, 2] ** c)
                        ex = factorial(a) * factorial(b) * factorial(c) / factorial(a + b + c + dim)
                        assert abs(v - ex) < 1e-14 * max(ex, 1e-3)
        pts, w = O.collapsed_rule(O.TRI, n)
        b = np.unique(np.round(pts[:, 1], 14))
        tb, _ = roots_jacobi(n, 1.0, 0.0)
        assert np.allclose(np.sort(b), np.sort((1 + tb) / 2), atol=1e-13)
        ta, _ = roots_jacobi(n, 0.0, 0.0)
        a = pts[:, 0] / (1 - pts[:, 1])
        assert np.allclose(np.sort(np.unique(np.round(a, 13))), np.sort((1 + ta) / 2), atol=1e-12)
    for etype, top in ((O.TRI, 12), (O.TET, 12)):
        for order in range(1, top + 1):
            pts, w = O.simplex_rule(etype, order)
            dim = 2 if etype == O.TRI else 3
            for a in range(order + 1):
                v = np.sum(w * pts[:, 0] ** a)
                ex = factorial(a) / factorial(a + dim)
                assert abs(v - ex) < 1e-14


def _h1(etype, ids, orders, x):
    return O.shape_h1(etype, ids, orders, x)


def _monomials(x, p):
    dim = x.shape[1]
    cols = []
    for a in range(p + 1):
        for b in range(p + 1 - a):
            if dim == 2:
                cols.append(x[:, 0] ** a * x[:, 1] ** b)
            else:
                for c in range(p + 1 - a - b):
                    cols.append(x[:, 0] ** a * x[:, 1] ** b * x[:, 2] ** c)
    return np.array(cols).T


@pytest.mark.parametrize("etype", [O.TRI, O.TET])
@pytest.mark.parametrize("p", [1, 2, 3, 4, 5, 6])
def test_h1_simplex_span_count_and_gradients(etype, p):
    """With every side at order p the hierarchical set has dim P_p functions
    and spans P_p exactly (a dropped or duplicated function, or a wrong blend
    or Chebyshev degree, breaks the span); gradients match finite
    differences for random corner ids."""
    dim = 2 if etype == O.TRI else 3
    nc, ns = dim + 1, (7 if dim == 2 else 15)
    rng = np.random.default_rng(p)
    ids = rng.permutation(50)[:nc]
    orders = np.full(ns - nc, p)
    nP = (p + 1) * (p + 2) // 2 if dim == 2 else (p + 1) * (p + 2) * (p + 3) // 6
    x = rng.dirichlet(np.ones(nc), 3 * nP)[:, 1:]
    Phi = np.array([_h1(etype, ids, orders, xi)[0] for xi in x])
    assert Phi.shape[1] == nP
    assert np.linalg.matrix_rank(Phi, tol=1e-10) == nP
    M = _monomials(x, p)
    coef, *_ = np.linalg.lstsq(Phi, M, rcond=None)
    assert np.abs(Phi @ coef - M).max() < 1e-10
    xi = x[3]
    phi, dphi = _h1(etype, ids, orders, xi)
    h = 1e-6
    for d in range(dim):
        xp = xi.copy(); xp[d] += h
        xm = xi.copy(); xm[d] -= h
        fd = (_h1(etype, ids, orders, xp)[0] - _h1(etype, ids, orders, xm)[0]) / (2 * h)
        assert np.allclose(dphi[d], fd, atol=1e-7 * max(1, np.abs(fd).max()))


_TET_SIDES = [(0,), (1,), (2,), (3,), (0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3),
              (0, 1, 2), (0, 1, 3), (1, 2, 3), (0, 2, 3), (0, 1, 2, 3)]
_TRI_SIDES = [(0,), (1,), (2,), (0, 1), (1, 2), (2, 0), (0, 1, 2)]


def _side_blocks(sides, nc, orders, etype):
    """offset and count of each side's functions in the element's list"""
    def nshape(s, o):
        k = len(sides[s])
        if k == 1:
            return 1
        if k == 2:
            return max(o - 1, 0)
        if k == 3:
            return max((o - 1) * (o - 2) // 2, 0)
        return max((o - 1) * (o - 2) * (o - 3) // 6, 0)
    off, out = 0, []
    for s in range(len(sides)):
        n = nshape(s, 1 if s < nc else orders[s - nc])
        out.append((off, n))
        off += n
    return out


def _ref_coords(X, P):
    """reference coordinates of physical point(s) P in the simplex with corners X"""
    J = (X[1:] - X[0]).T
    return np.linalg.lstsq(J, (P - X[0]).T, rcond=None)[0].T


@pytest.mark.parametrize("p", [3, 4, 5])
def test_h1_conformity_across_faces_random_ids(p):
    """Two tetrahedra sharing a face, and the boundary triangle on it, in
    random local orders with random global ids: the functions of every shared
    connect (corners, edges, face) agree on the face.  This pins the id-based
    side transforms: a wrong edge flip or face permutation makes the traces
    of the two elements differ."""
    rng = np.random.default_rng(10 + p)
    for trial in range(12):
        G = rng.random((5, 3))
        gid = rng.permutation(1000)[:5]
        A = rng.permutation([0, 1, 2, 3])             # global node 4 is B's apex
        B = rng.permutation([1, 2, 3, 4])
        T = rng.permutation([1, 2, 3])
        oA = np.full(11, p); oB = np.full(11, p); oT = np.full(4, p)
        blkA = _side_blocks(_TET_SIDES, 4, oA, O.TET)
        blkB = _side_blocks(_TET_SIDES, 4, oB, O.TET)
        blkT = _side_blocks(_TRI_SIDES, 3, oT, O.TRI)
        w = rng.dirichlet(np.ones(3), 6)
        P = w @ G[[1, 2, 3]]
        for Pk in P:
            xa = _ref_coords(G[A], Pk[None])[0]
            xb = _ref_coords(G[B], Pk[None])[0]
            xt = _ref_coords(G[T], Pk[None])[0]
            pa = _h1(O.TET, gid[A], oA, xa)[0]
            pb = _h1(O.TET, gid[B], oB, xb)[0]
            pt = _h1(O.TRI, gid[T], oT, xt)[0]
            for gs in [(1,), (2,), (3,), (1, 2), (2, 3), (1, 3), (1, 2, 3)]:
                def loc(order, sides):
                    key = sorted(int(np.where(order == g)[0][0]) for g in gs)
                    return next(s for s, c in enumerate(sides) if sorted(c) == key)
                sa, sb, st = loc(A, _TET_SIDES), loc(B, _TET_SIDES), loc(T, _TRI_SIDES)
                (a0, n), (b0, nb_), (t0, nt) = blkA[sa], blkB[sb], blkT[st]
                assert n == nb_ == nt and n > 0
                assert np.allclose(pa[a0:a0 + n], pb[b0:b0 + n], atol=1e-12)
                assert np.allclose(pa[a0:a0 + n], pt[t0:t0 + n], atol=1e-12)


def _tet_hp_problem(seed, pmax=5, n=(2, 2, 2), bc=None, bcv=None, force=0.0, material=None):
    rng = np.random.default_rng(seed)
    pb = P.Problem(material=P.DARCY if material is None else material, dim=3, p=pmax,
                   nx=n[0], ny=n[1], nz=n[2], simplex=1,
                   force_kind=P.FORCE_CONST if force else P.FORCE_NONE,
                   force_const=(force, 0, 0),
                   bc_type=bc if bc is not None else [0] * 6,
                   bc_val=[[v, 0, 0] for v in (bcv if bcv is not None else [0] * 6)])
    pb.el_porder = rng.integers(1, pmax + 1, pb.nvol).astype(np.int32)
    return pb


def test_tet_hp_sizes_and_pattern():
    """p_e ~ U{1..5} on a Kuhn mesh: each connect carries NConnectShapeF of the
    minimum preferred order of the tetrahedra sharing its side (boundary
    triangles keep the default order), recomputed here from the element lists;
    the CSR pattern is the union of element dof pairs."""
    pb = _tet_hp_problem(3)
    o = O.Oracle(pb)
    econn, pos, size = o.connects()
    mins = {}
    for e in range(pb.nvol):
        for s in range(4, 15):
            c = econn[e, s]
            mins[c] = min(mins.get(c, pb.p), int(pb.el_porder[e]))
    kinds = {}
    for e in range(pb.nvol):
        for s in range(4, 15):
            kinds[econn[e, s]] = len(_TET_SIDES[s])
    for c, pmin in mins.items():
        k = kinds[c]
        exp = {2: pmin - 1, 3: (pmin - 1) * (pmin - 2) // 2,
               4: (pmin - 1) * (pmin - 2) * (pmin - 3) // 6}[k]
        assert size[c] == max(exp, 0), (c, k, pmin)
    assert o.neq == int(size.sum())
    ia, ja = o.pattern()
    rows = [set() for _ in range(o.neq)]
    for e in range(o.m.nel):
        _, _, dest = o.calcstiff(e)
        for i in dest:
            for j in dest:
                if j >= i:
                    rows[i].add(j)
    assert np.array_equal(np.diff(ia), [len(r) for r in rows])
    assert np.array_equal(ja, np.concatenate([np.array(sorted(r), dtype=np.int64) for r in rows]))


def test_tet_high_order_element_energy():
    """For tetrahedra of order 2..5 in an hp mesh: the coefficients c of a
    quadratic q in the element's own basis (interpolated by least squares)
    give c^T K_e c = kappa_e * int_e |grad q|^2, the integral taken with an
    independent collapsed rule; a wrong Jacobian, rule or gradient
    transformation changes the energy."""
    pb = _tet_hp_problem(5)
    pb.el_porder = np.random.default_rng(5).integers(2, 6, pb.nvol).astype(np.int32)
    pb.el_coef = np.exp(np.random.default_rng(1).standard_normal(pb.nvol))
    o = O.Oracle(pb)
    rng = np.random.default_rng(2)
    q = lambda X: X[:, 0] ** 2 + 0.5 * X[:, 1] * X[:, 2] - 0.3 * X[:, 2]
    gq = lambda X: np.stack([2 * X[:, 0], 0.5 * X[:, 2], 0.5 * X[:, 1] - 0.3], 1)
    checked = 0
    for e in range(pb.nvol):
        el = o.element(e)
        ek, _, _ = o.calcstiff(e)
        assert el["orders"].min() >= 2 and ek.shape[0] >= 10
        X = el["xyz"]
        xi = rng.dirichlet(np.ones(4), 200)[:, 1:]
        Phi = np.array([_h1(O.TET, el["ids"], el["orders"], x)[0] for x in xi])
        phys = X[0] + xi @ (X[1:] - X[0])
        c, *_ = np.linalg.lstsq(Phi, q(phys), rcond=None)
        assert np.abs(Phi @ c - q(phys)).max() < 1e-10
        pts, w = O.collapsed_rule(O.TET, 4)
        J = (X[1:] - X[0]).T
        P_ = X[0] + pts @ (X[1:] - X[0])
        ex = pb.el_coef[e] * abs(np.linalg.det(J)) * np.sum(w * (gq(P_) ** 2).sum(1))
        assert abs(c @ ek @ c - ex) < 1e-11 * ex
        checked += 1
    assert checked >= 5


@pytest.mark.parametrize("pmax", [3, 5])
def test_tet_hp_patch_tests(pmax):
    """Random p_e in 1..pmax: u = x (Dirichlet 0 at x = 0, flux 1 at x = 1) is
    reproduced exactly (corner dofs = x, all others 0); with every p_e >= 2,
    u = x(1-x)/2 (-lap u = 1, u = 0 at x = 0, 1) has corner dofs u(node),
    first edge dofs u(mid) - (u_a + u_b)/2 and every other dof 0."""
    cases = [(lambda X: X, [0, 1, 1, 1, 1, 1], [0, 1, 0, 0, 0, 0], 0.0, 1),
             (lambda X: X * (1 - X) / 2, [0, 0, 1, 1, 1, 1], [0] * 6, 1.0, 2)]
    for u_ex, bc, bcv, f, pmin in cases:
        pb = _tet_hp_problem(7, pmax, bc=bc, bcv=bcv, force=f, material=P.POISSON)
        rng = np.random.default_rng(7)
        pb.el_porder = rng.integers(pmin, pmax + 1, pb.nvol).astype(np.int32)
        o = O.Oracle(pb)
        u = _solve_oracle(o)
        econn, pos, size = o.connects()
        enodes = np.ctypeslib.as_array(o.m.enodes, shape=(8 * o.m.nel,)).reshape(o.m.nel, 8)
        xyz = np.ctypeslib.as_array(o.m.xyz, shape=(3 * o.m.nnodes,)).reshape(-1, 3)
        for e in range(pb.nvol):
            for s in range(15):
                c = econn[e, s]
                if not size[c]:
                    continue
                blk = u[pos[c]:pos[c] + size[c]]
                if s < 4:
                    assert abs(blk[0] - u_ex(xyz[enodes[e, s], 0])) < 1e-9
                elif s < 10:
                    a, b = _TET_SIDES[s]
                    xa, xb = xyz[enodes[e, a], 0], xyz[enodes[e, b], 0]
                    exp = u_ex(0.5 * (xa + xb)) - 0.5 * (u_ex(xa) + u_ex(xb))
                    assert abs(blk[0] - exp) < 1e-9
                    assert np.abs(blk[1:]).max(initial=0) < 1e-9
                else:
                    assert np.abs(blk).max() < 1e-9


# --------------------------------------------------------------------------- non-symmetric format
@pytest.mark.parametrize("kind", ["poisson2d", "elast3d", "tet_hp5"])
def test_full_pattern_and_values(kind):
    """TPZSpStructMatrix / TPZFYsmpMatrix: every row lists all coupled
    equations (both triangles) sorted; the pattern equals the brute-force
    union of element dof pairs and the symmetric closure of the upper pattern;
    TPZFYsmpMatrix::AddKel values agree with the symmetric matrix (lower
    entries with their transposes up to the rounding of ek's asymmetry)."""
    if kind == "poisson2d":
        pb = P.poisson2d(n=5, p=2, seed=1)
    elif kind == "elast3d":
        pb = P.elasticity3d(n=2, p=2, seed=2, load="traction")
    else:
        pb = _tet_hp_problem(9, force=1.0)
    o = O.Oracle(pb)
    fia, fja, fa, frhs = o.assemble_full()
    rows = [set() for _ in range(o.neq)]
    for e in range(o.m.nel):
        _, _, dest = o.calcstiff(e)
        for i in dest:
            rows[i].update(dest)
    assert np.array_equal(np.diff(fia), [len(r) for r in rows])
    assert np.array_equal(fja, np.concatenate([np.array(sorted(r), dtype=np.int64) for r in rows]))
    ia, ja = o.pattern()
    a, rhs = o.assemble()
    K = O.upper_to_dense(ia, ja, a)
    F = np.zeros_like(K)
    for r in range(o.neq):
        F[r, fja[fia[r]:fia[r + 1]]] = fa[fia[r]:fia[r + 1]]
    scale = np.abs(K).max(axis=1, keepdims=True)
    assert np.all(np.abs(F - K) <= 1e-13 * scale)
    assert np.array_equal(frhs, rhs)


def test_assemble_rows_equals_whole_assembly():
    """oracle.assemble_rows (the elements touching the given rows only) gives
    those rows' entries and loads bit for bit as the whole assembly does."""
    for pb in (P.elasticity3d(n=4, p=2, seed=1, field_kind="grf", load="body"),
               P.darcy3d_tet(n=2, pmax=4, seed=3), P.poisson2d(n=6, p=2, seed=2)):
        o = O.Oracle(pb)
        ia, ja = o.pattern()
        a, rhs = o.assemble()
        rows = np.sort(np.random.default_rng(0).choice(o.neq, min(150, o.neq // 2), replace=False))
        lens = ia[rows + 1] - ia[rows]
        ia_s = np.concatenate([[0], np.cumsum(lens)])
        idx = np.repeat(ia[rows] - ia_s[:-1], lens) + np.arange(ia_s[-1])
        a_s, rhs_s, nel = O.assemble_rows(o, rows, ia_s, ja[idx])
        assert 0 < nel <= o.m.nel
        assert np.array_equal(a_s, a[idx]) and np.array_equal(rhs_s, rhs[rows])
        # the rows of one element touch only its neighbours
        r1 = np.unique(o.calcstiff(0)[2])
        l1 = ia[r1 + 1] - ia[r1]
        i1 = np.concatenate([[0], np.cumsum(l1)])
        x1 = np.repeat(ia[r1] - i1[:-1], l1) + np.arange(i1[-1])
        a1, f1, n1 = O.assemble_rows(o, r1, i1, ja[x1])
        assert n1 < o.m.nel and np.array_equal(a1, a[x1]) and np.array_equal(f1, rhs[r1])
        # a column outside the given row pattern is reported
        with pytest.raises(ValueError):
            O.assemble_rows(o, rows, ia_s, ja[idx] + 10 ** 9)
