"""GPU parity of the HIP path against the CPU oracle (through the C ABI).

Bars (written here, checked below):
  * equation numbering and CSR pattern (IA/JA): bit-exact;
  * assembled FP64 entries: |a - a_ref| <= 1e-12 |a_ref| + 1e-15 max|row| (the
    reference's |v| < 1e-12 drop rule is reproduced, structural zeros match up
    to cancellation residues below 1e-15 max|row|);
  * rhs: |f - f_ref| <= 1e-12 max|f_ref|;
  * PCG64 solution: ||u - u_ref|| <= 1e-8 ||u_ref|| when both are converged
    to a recursive residual of 1e-13;
  * mixed FP32-FCG + FP64 refinement: true FP64 residual <= 1e-8 (the
    benchmark target) and the solution within 1e-5 of a sparse direct solve
    of the oracle's matrix (homogeneous Dirichlet data: with an inhomogeneous
    penalty value ||f|| ~ 1e16 and a relative residual says little about
    the solution).
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle as O
from neopz_amd import _lib as L
from neopz_amd import mesh as M
from neopz_amd import problems as P
from neopz_amd.analysis import Context, TPZLinearAnalysis, TPZSSpStructMatrixHIP, TPZStepSolver, \
    TPZSYsmpMatrixHIP, TPZSpStructMatrixHIP

pytestmark = pytest.mark.gpu

CASES = {
    "poisson2d_p2": lambda: P.poisson2d(n=12, p=2, seed=1),
    "poisson2d_p1": lambda: P.poisson2d(n=10, p=1, seed=2),
    "darcy2d_p2": lambda: P.darcy2d(n=16, p=2, seed=3),
    "darcy2d_p1_neumann": lambda: P.Problem(material=P.DARCY, dim=2, p=1, nx=9, ny=7,
                                            el_coef=np.exp(np.random.default_rng(4).standard_normal(63)),
                                            force_kind=P.FORCE_CONST, force_const=(1.0, 0, 0),
                                            bc_type=[0, 1, -1, 0, -1, -1],
                                            bc_val=[[0, 0, 0], [2.0, 0, 0], [0, 0, 0], [0.5, 0, 0],
                                                    [0, 0, 0], [0, 0, 0]]),
    "darcy2d_p1_neumann_h": lambda: P.Problem(material=P.DARCY, dim=2, p=1, nx=9, ny=7,
                                              el_coef=np.exp(np.random.default_rng(4).standard_normal(63)),
                                              force_kind=P.FORCE_CONST, force_const=(1.0, 0, 0),
                                              bc_type=[0, 1, -1, 0, -1, -1],
                                              bc_val=[[0, 0, 0], [2.0, 0, 0], [0, 0, 0], [0, 0, 0],
                                                      [0, 0, 0], [0, 0, 0]]),
    "elast3d_p2_traction": lambda: P.elasticity3d(n=5, p=2, seed=5, load="traction"),
    "elast3d_p2_body": lambda: P.elasticity3d(n=4, p=2, seed=6, field_kind="grf", load="body"),
    "elast3d_p1": lambda: P.elasticity3d(n=6, p=1, seed=7, load="traction"),
    "poisson3d_p2": lambda: P.Problem(material=P.POISSON, dim=3, p=2, nx=4, ny=3, nz=5,
                                      force_kind=P.FORCE_CONST, force_const=(1.0, 0, 0),
                                      bc_type=[0, 0, 1, 1, 0, 0]),
    # tetrahedra (Kuhn split) with boundary triangles
    "darcy_tet_p2": lambda: P.darcy3d_tet(n=3, pmax=2, seed=12, hp=False),
    "darcy_tet_hp": lambda: P.darcy3d_tet(n=4, pmax=2, seed=13, hp=True),
    # high-order tetrahedra: uniform p = 3, Poisson p = 4 with flux data, hp p_e in 1..5
    "darcy_tet_p3": lambda: P.darcy3d_tet(n=2, pmax=3, seed=14, hp=False),
    "poisson_tet_p4_flux": lambda: P.Problem(material=P.POISSON, dim=3, p=4, nx=2, ny=1, nz=2,
                                             simplex=1, scale=1.5, force_kind=P.FORCE_CONST,
                                             force_const=(1.0, 0, 0), bc_type=[0, 1, -1, 1, 0, -1],
                                             bc_val=[[0, 0, 0], [0.7, 0, 0], [0, 0, 0], [-0.2, 0, 0],
                                                     [0, 0, 0], [0, 0, 0]]),
    "darcy_tet_hp5": lambda: P.darcy3d_tet(n=3, pmax=5, seed=15, hp=True),
    "elast_tet_p2": lambda: P.Problem(material=P.ELAST3D, dim=3, p=2, nx=2, ny=3, nz=2, simplex=1,
                                      E=2.0, force_kind=P.FORCE_CONST, force_const=(0.0, 0.0, -1.0),
                                      bc_type=[0, 1, -1, -1, 0, -1],
                                      bc_val=[[0, 0, 0], [0.5, 0, 0], [0, 0, 0], [0, 0, 0],
                                              [0, 0, 0], [0, 0, 0]]),
    # TPZElasticity3D on high-order tetrahedra: uniform p = 3 with a body load
    # and a traction, hp p_e in 1..4 with heterogeneous E and prestress
    "elast_tet_p3": lambda: P.Problem(material=P.ELAST3D, dim=3, p=3, nx=2, ny=1, nz=2, simplex=1,
                                      E=2.0, force_kind=P.FORCE_CONST, force_const=(0.3, 0.0, -1.0),
                                      bc_type=[0, 1, -1, 1, 0, -1],
                                      bc_val=[[0, 0, 0], [0.5, 0, 0.1], [0, 0, 0], [0, -0.2, 0],
                                              [0, 0, 0], [0, 0, 0]]),
    "elast_tet_hp4": lambda: _elast_tet_hp(),
    # hp: per-element orders 1/2, shared sides at the minimum order
    "poisson2d_hp": lambda: _hp(P.poisson2d(n=12, p=2, seed=8), 8),
    "darcy2d_hp": lambda: _hp(P.darcy2d(n=14, p=2, seed=9), 9),
    "elast3d_hp": lambda: _hp(P.elasticity3d(n=4, p=2, seed=10, field_kind="grf", load="body"), 10),
    "poisson3d_hp": lambda: _hp(P.Problem(material=P.POISSON, dim=3, p=2, nx=4, ny=3, nz=5,
                                          force_kind=P.FORCE_CONST, force_const=(1.0, 0, 0),
                                          bc_type=[0, 0, 1, 1, 0, 0]), 11),
}

def _elast_tet_hp():
    rng = np.random.default_rng(16)
    pb = P.Problem(material=P.ELAST3D, dim=3, p=4, nx=2, ny=2, nz=2, simplex=1,
                   el_coef=np.exp(0.5 * rng.standard_normal(48)),
                   el_porder=rng.integers(1, 5, 48).astype(np.int32),
                   force_kind=P.FORCE_CONST, force_const=(0.0, 1.0, -1.0),
                   bc_type=[0, -1, -1, -1, 1, -1],
                   bc_val=[[0, 0, 0], [0, 0, 0], [0, 0, 0], [0, 0, 0], [0.2, 0, 0.4], [0, 0, 0]])
    pb.prestress = (0.1, -0.2, 0.05)
    pb.name = "elast_tet_hp4"
    return pb


def _hp(pb, seed):
    pb.el_porder = np.random.default_rng(seed).integers(1, 3, pb.nvol).astype(np.int32)
    pb.name = (pb.name or "") + "_hp"
    return pb


_CTX = None


def ctx():
    global _CTX
    if _CTX is None:
        _CTX = Context(0)
    return _CTX


def _assemble(pb):
    print(f"[gpu] assemble {pb.name or pb.material} n=({pb.nx},{pb.ny},{pb.nz}) p={pb.p}", flush=True)
    c = ctx()
    sm = TPZSSpStructMatrixHIP(c, M.build(pb))
    K = sm.Create()
    sm.Assemble(K)
    rhs = np.zeros(K.neq)
    L.check(c.lib.pzg_get_system(c.h, None, L.ptr(rhs, L.dp)))
    return K, K.IA().copy(), K.JA().copy(), K.A().copy(), rhs


def _oracle(pb):
    o = O.Oracle(pb)
    ia, ja = o.pattern()
    a, rhs = o.assemble()
    return o, ia, ja, a, rhs


def _check_values(ia, a, a_ref):
    rowmax = np.maximum.reduceat(np.abs(a_ref), ia[:-1])
    rows = np.repeat(np.arange(len(ia) - 1), np.diff(ia))
    tol = 1e-12 * np.abs(a_ref) + 1e-15 * rowmax[rows]
    bad = np.abs(a - a_ref) > tol
    assert not bad.any(), (np.flatnonzero(bad)[:10], a[bad][:5], a_ref[bad][:5])
    # structural zeros (the drop rule) agree; a sum of element contributions
    # that cancels may be exactly 0 in one summation order and a rounding
    # residue in the other, both below the floor 1e-15 max|row|
    zmis = (a == 0) != (a_ref == 0)
    floor = 1e-15 * rowmax[rows]
    assert not (zmis & ((np.abs(a) > floor) | (np.abs(a_ref) > floor))).any()


@pytest.mark.parametrize("name", list(CASES))
def test_pattern_and_values(name):
    pb = CASES[name]()
    K, ia, ja, a, rhs = _assemble(pb)
    o, ia_r, ja_r, a_r, rhs_r = _oracle(pb)
    assert np.array_equal(ia, ia_r)
    assert np.array_equal(ja, ja_r)
    _check_values(ia, a, a_r)
    assert np.max(np.abs(rhs - rhs_r)) <= 1e-12 * max(np.max(np.abs(rhs_r)), 1e-300)


@pytest.mark.parametrize("name", ["poisson2d_p2", "elast3d_p2_traction", "poisson3d_p2", "darcy_tet_hp5"])
def test_full_csr_matches_oracle(name):
    """TPZSpStructMatrix format: pattern bit-exact with the oracle's
    TPZFYsmpMatrix, values within 1e-12, and the operator of the non-symmetric
    matrix object solves like the symmetric one."""
    pb = CASES[name]()
    c = ctx()
    sm = TPZSpStructMatrixHIP(c, M.build(pb))
    K = sm.Create()
    sm.Assemble(K)
    o = O.Oracle(pb)
    fia, fja, fa, frhs = o.assemble_full()
    assert np.array_equal(K.IA(), fia) and np.array_equal(K.JA(), fja)
    _check_values(fia, K.A(), fa)
    r0 = int(fia[1] - 1)
    assert K.GetVal(0, int(fja[r0])) == pytest.approx(fa[r0], rel=1e-12)
    x = np.random.default_rng(3).standard_normal(o.neq)
    rows = np.repeat(np.arange(o.neq), np.diff(fia))
    Kf = sp.csr_matrix((fa, (rows, fja)), shape=(o.neq, o.neq))
    y = K.Multiply(x)
    assert np.linalg.norm(y - Kf @ x) <= 1e-12 * np.linalg.norm(Kf @ x)


def test_hp_assembly_autotuner():
    """The high-order assembly picks one kernel variant per rule bucket by
    timing (the fastest of the five), the tuning launches leave no trace in the
    system, and later assemblies reproduce the first bit for bit."""
    pb = CASES["darcy_tet_hp5"]()
    K, ia, ja, a1, r1 = _assemble(pb)
    tun = L.hp_tuning(ctx())
    assert len(tun) == len(set(M.build(pb)["vol_rule"]))
    for t in tun:
        ms = t["ms_per_variant"]
        assert 1 <= t["order"] <= 5 and 0 <= t["variant"] < 5
        assert all(m > 0 for m in ms) and ms[t["variant"]] == min(ms)
    o, ia_r, ja_r, a_r, rhs_r = _oracle(pb)
    _check_values(ia, a1, a_r)
    L.check(ctx().lib.pzg_assemble(ctx().h))
    a2 = np.zeros(len(ja))
    r2 = np.zeros(len(ia) - 1)
    L.check(ctx().lib.pzg_get_system(ctx().h, L.ptr(a2, L.dp), L.ptr(r2, L.dp)))
    assert np.array_equal(a1, a2) and np.array_equal(r1, r2)


def test_assembly_is_deterministic():
    pb = CASES["elast3d_p2_body"]()
    _, _, _, a1, r1 = _assemble(pb)
    _, _, _, a2, r2 = _assemble(pb)
    assert np.array_equal(a1, a2) and np.array_equal(r1, r2)


@pytest.mark.parametrize("name", ["poisson2d_p2", "darcy2d_p2", "elast3d_p2_traction", "elast3d_hp",
                                  "darcy_tet_hp", "darcy_tet_hp5", "elast_tet_hp4"])
def test_pcg64_matches_oracle(name):
    pb = CASES[name]()
    an = TPZLinearAnalysis(pb, False, ctx=ctx())
    st = TPZStepSolver()
    st.SetCG(20000, TPZStepSolver.jacobi(), 1e-13, 0)
    an.SetSolver(st)
    an.Assemble()
    u = an.Solve()
    o, ia, ja, a, rhs = _oracle(pb)
    ref = O.pcg(ia, ja, a, rhs, maxit=20000, tol=1e-13, wasted_spmv=False)
    assert ref["converged"] and st.info["converged"]
    assert np.linalg.norm(u - ref["x"]) <= 1e-8 * np.linalg.norm(ref["x"])
    assert abs(st.fNumIterations - ref["iters"]) <= max(3, 0.05 * ref["iters"])
    assert st.fTol <= 1e-13


@pytest.mark.parametrize("name", ["poisson2d_p2", "elast3d_p2_traction", "elast3d_p1", "elast3d_hp",
                                  "elast_tet_p2", "darcy_tet_hp"])
def test_pcg64_block_jacobi_matches_oracle(name):
    """CG with BuildPreconditioner(EBlockJacobi, true) against the oracle's
    CG with the LU-solved connect blocks (nstate equations each): solution
    within 1e-8, iterations within max(3, 5%)."""
    pb = CASES[name]()
    print(f"[gpu] PCG64 block Jacobi {name}", flush=True)
    an = TPZLinearAnalysis(pb, False, ctx=ctx())
    st = TPZStepSolver()
    an.SetSolver(st)
    an.Assemble()
    assert an.BuildPreconditioner(an.EJacobi, True) is None
    pre = an.BuildPreconditioner(an.EBlockJacobi, True)
    st.SetCG(20000, pre, 1e-13, 0)
    u = an.Solve()
    o, ia, ja, a, rhs = _oracle(pb)
    ref = O.pcg_block(ia, ja, a, rhs, np.full(o.neq // pb.nstate, pb.nstate, np.int32),
                      maxit=20000, tol=1e-13)
    pj = O.pcg(ia, ja, a, rhs, maxit=20000, tol=1e-13, wasted_spmv=False)
    print(f"[gpu]   iterations: device {st.fNumIterations}, oracle {ref['iters']}, "
          f"point Jacobi {pj['iters']}", flush=True)
    assert ref["converged"] and st.info["converged"]
    assert np.linalg.norm(u - ref["x"]) <= 1e-8 * np.linalg.norm(ref["x"])
    assert abs(st.fNumIterations - ref["iters"]) <= max(3, 0.05 * ref["iters"])
    assert st.fTol <= 1e-13
    if pb.nstate == 1:  # one equation per connect: point Jacobi
        assert abs(st.fNumIterations - pj["iters"]) <= max(3, 0.05 * pj["iters"])


@pytest.mark.parametrize("name", ["darcy_tet_hp5", "elast_tet_hp4", "elast_tet_p3"])
def test_pcg64_block_jacobi_high_order_matches_oracle(name):
    """High-order connects carry several shapes: BuildPreconditioner(
    EBlockJacobi, true) inverts each connect's whole block of consecutive
    equations (up to 18 here).  Against the oracle's CG with the LU-solved
    blocks of the same sizes (the mesh's Block() sizes in equation order):
    solution within 1e-8, iterations within max(3, 5%), and fewer iterations
    than point Jacobi."""
    pb = CASES[name]()
    md = M.build(pb)
    order = np.argsort(md["blk_pos"], kind="stable")
    sizes = np.asarray(md["blk_size"])[order]
    sizes = sizes[sizes > 0].astype(np.int32)
    print(f"[gpu] PCG64 high-order block Jacobi {name}: {len(sizes)} blocks, largest {sizes.max()}", flush=True)
    an = TPZLinearAnalysis(pb, False, ctx=ctx())
    st = TPZStepSolver()
    an.SetSolver(st)
    an.Assemble()
    st.SetCG(20000, an.BuildPreconditioner(an.EBlockJacobi, True), 1e-13, 0)
    u = an.Solve()
    o, ia, ja, a, rhs = _oracle(pb)
    assert sizes.sum() == o.neq
    ref = O.pcg_block(ia, ja, a, rhs, sizes, maxit=20000, tol=1e-13)
    pj = O.pcg(ia, ja, a, rhs, maxit=20000, tol=1e-13, wasted_spmv=False)
    print(f"[gpu]   iterations: device {st.fNumIterations}, oracle {ref['iters']}, point Jacobi {pj['iters']}",
          flush=True)
    assert ref["converged"] and st.info["converged"]
    assert np.linalg.norm(u - ref["x"]) <= 1e-8 * np.linalg.norm(ref["x"])
    assert abs(st.fNumIterations - ref["iters"]) <= max(3, 0.05 * ref["iters"])
    assert ref["iters"] < pj["iters"]


def test_pcg64_block_jacobi_refusals():
    """PCG64 with a multigrid preconditioner: PZG_EINVAL; the other
    BuildPreconditioner forms: NotImplementedError."""
    an = TPZLinearAnalysis(CASES["darcy_tet_hp5"](), False, ctx=ctx())
    st = TPZStepSolver()
    an.SetSolver(st)
    an.Assemble()
    with pytest.raises(NotImplementedError):
        an.BuildPreconditioner(an.EElement, True)
    with pytest.raises(NotImplementedError):
        an.BuildPreconditioner(an.EBlockJacobi, False)
    st.SetCG(100, TPZStepSolver.jacobi(), 1e-10, 0)
    st.precond = L.PRECOND_MULTIGRID
    with pytest.raises(L.PzgError) as e:
        an.Solve()
    assert e.value.code == L.PZG_EINVAL


@pytest.mark.parametrize("inner_op", [L.INNER_MATFREE, L.INNER_BSR])
@pytest.mark.parametrize("name", ["poisson2d_p2", "darcy2d_p1_neumann_h", "elast3d_p2_body",
                                  "elast3d_p1"])
def test_mixed_refinement(name, inner_op):
    pb = CASES[name]()
    an = TPZLinearAnalysis(pb, False, ctx=ctx())
    st = TPZStepSolver()
    st.SetMixedFCG(50000, 1e-8, inner_tol=1e-4, inner_op=inner_op)
    an.SetSolver(st)
    an.Assemble()
    u = an.Solve()
    print("mixed info", st.info, flush=True)
    assert st.info["converged"] and st.info["true_residual"] <= 1e-8
    o, ia, ja, a, rhs = _oracle(pb)
    rows = np.repeat(np.arange(o.neq), np.diff(ia))
    U = sp.csr_matrix((a, (rows, ja)), shape=(o.neq, o.neq))
    Kf = (U + sp.triu(U, 1).T).tocsc()
    ud = spla.spsolve(Kf, rhs)
    assert np.linalg.norm(u - ud) <= 1e-5 * np.linalg.norm(ud)
    rel = np.linalg.norm(rhs - Kf @ u) / np.linalg.norm(rhs)
    assert rel <= 2e-8


def _scaled_apply(c, op, x):
    y = np.zeros_like(x)
    L.check(c.lib.pzg_scaled_apply(c.h, op, L.ptr(x, L.dp), L.ptr(y, L.dp)))
    return y


@pytest.mark.parametrize("name", ["poisson2d_p2", "poisson2d_p1", "darcy2d_p2", "darcy2d_p1_neumann",
                                  "elast3d_p2_traction", "elast3d_p2_body", "elast3d_p1",
                                  "poisson3d_p2"])
def test_inner_operators_match_scaled_matrix(name):
    """Both FP32 inner operators apply S = D^-1/2 K D^-1/2 of the oracle's K
    (FP64) to within FP32 accuracy: ||y - S x|| <= 1e-5 ||S x||, and the
    matrix-free result is bitwise repeatable."""
    pb = CASES[name]()
    K, ia, ja, a, rhs = _assemble(pb)
    o, ia_r, ja_r, a_r, _ = _oracle(pb)
    rows = np.repeat(np.arange(o.neq), np.diff(ia_r))
    U = sp.csr_matrix((a_r, (rows, ja_r)), shape=(o.neq, o.neq))
    Kf = (U + sp.triu(U, 1).T).tocsr()
    dis = 1.0 / np.sqrt(np.abs(Kf.diagonal()))
    x = np.random.default_rng(11).standard_normal(o.neq).astype(np.float32).astype(np.float64)
    ref = dis * (Kf @ (dis * x))
    c = ctx()
    for op in (L.INNER_BSR, L.INNER_MATFREE):
        y = _scaled_apply(c, op, x)
        err = np.linalg.norm(y - ref) / np.linalg.norm(ref)
        print(f"{name} op {op}: rel err {err:.2e}", flush=True)
        assert err <= 1e-5
    assert np.array_equal(_scaled_apply(c, L.INNER_MATFREE, x), _scaled_apply(c, L.INNER_MATFREE, x))


@pytest.mark.parametrize("name", ["elast3d_p2_body", "poisson3d_p2", "elast3d_p1"])
def test_block_assembly_matches_element_scatter(name):
    """Identical affine hexahedra are assembled by blocks (every stored block
    summed over the elements sharing it, from the coefficient-1 element
    matrix); PZG_NO_BLKASM=1 keeps the coloured element scatter.  Both meet
    the oracle's bar, and agree with each other entry by entry within 1e-13
    of the row's largest entry."""
    import os
    pb = CASES[name]()
    res = {}
    for flag in ("0", "1"):
        os.environ["PZG_NO_BLKASM"] = flag
        try:
            K, ia, ja, a, rhs = _assemble(pb)
        finally:
            del os.environ["PZG_NO_BLKASM"]
        res[flag] = (ia, a, rhs)
    o, ia_r, ja_r, a_r, rhs_r = _oracle(pb)
    for flag, (ia, a, rhs) in res.items():
        assert np.array_equal(ia, ia_r)
        _check_values(ia, a, a_r)
        assert np.max(np.abs(rhs - rhs_r)) <= 1e-12 * max(np.max(np.abs(rhs_r)), 1e-300)
    ia, a0, r0 = res["0"]
    a1, r1 = res["1"][1], res["1"][2]
    rowmax = np.maximum.reduceat(np.abs(a1), ia[:-1])
    rows = np.repeat(np.arange(len(ia) - 1), np.diff(ia))
    assert np.all(np.abs(a0 - a1) <= 1e-13 * rowmax[rows])
    assert np.array_equal(r0, r1)


def test_matfree_needs_mesh_built_system():
    pb = CASES["poisson3d_p2"]()
    o, ia, ja, a, rhs = _oracle(pb)
    K = TPZSYsmpMatrixHIP(ctx(), o.neq)
    K.SetData(ia, ja, a, rhs)
    x = np.ones(o.neq)
    with pytest.raises(L.PzgError):
        _scaled_apply(ctx(), L.INNER_MATFREE, x)
    st = TPZStepSolver(K)
    st.SetMixedFCG(10000, 1e-8, inner_op=L.INNER_AUTO)   # falls back to the stored operator
    st.Solve(rhs)
    assert st.info["converged"]


@pytest.mark.parametrize("name", ["poisson2d_hp", "elast3d_hp", "darcy_tet_hp", "darcy_tet_hp5", "elast_tet_hp4"])
def test_mixed_refinement_hp(name):
    """hp meshes (several integration rules): the mixed solver uses the stored
    FP32 operator; the matrix-free one is refused."""
    pb = CASES[name]()
    an = TPZLinearAnalysis(pb, False, ctx=ctx())
    st = TPZStepSolver()
    st.SetMixedFCG(50000, 1e-8, inner_tol=1e-4, inner_op=L.INNER_AUTO)
    an.SetSolver(st)
    an.Assemble()
    u = an.Solve()
    assert st.info["converged"] and st.info["true_residual"] <= 1e-8
    o, ia, ja, a, rhs = _oracle(pb)
    rows = np.repeat(np.arange(o.neq), np.diff(ia))
    U = sp.csr_matrix((a, (rows, ja)), shape=(o.neq, o.neq))
    Kf = (U + sp.triu(U, 1).T).tocsc()
    ud = spla.spsolve(Kf, rhs)
    assert np.linalg.norm(u - ud) <= 1e-5 * np.linalg.norm(ud)
    with pytest.raises(L.PzgError):
        _scaled_apply(ctx(), L.INNER_MATFREE, np.ones(o.neq))


def test_known_matrix_1d_addkel():
    """StructMatrixUnitTest "Assemble known matrix" through SetData + batched AddKel."""
    o = O.Oracle(P.mock1d(20))
    ia, ja = o.pattern()
    K = TPZSYsmpMatrixHIP(ctx(), o.neq)
    K.SetData(ia, ja, np.zeros(len(ja)))
    ek = np.tile(np.array([[2.0, 1.0], [1.0, 2.0]]), (20, 1, 1))
    dest = np.stack([np.arange(20), np.arange(1, 21)], 1)
    K.AddKel(ek, dest)
    n = 21
    assert K.GetVal(0, 0) == 2 and K.GetVal(n - 1, n - 1) == 2
    assert K.GetVal(0, 1) == 1 and K.GetVal(n - 1, n - 2) == 1
    for i in range(1, n - 1):
        assert K.GetVal(i, i - 1) == 1 and K.GetVal(i, i) == 4 and K.GetVal(i, i + 1) == 1
    # the drop rule: contributions below 1e-12 are skipped
    K.AddKel(np.array([[[5e-13, 5e-13], [5e-13, 3e-12]]]), np.array([[0, 1]]))
    assert K.GetVal(0, 0) == 2 and K.GetVal(0, 1) == 1 and K.GetVal(1, 1) == 4 + 3e-12


def test_multadd_matches_oracle():
    pb = CASES["elast3d_p2_traction"]()
    K, ia, ja, a, rhs = _assemble(pb)
    rng = np.random.default_rng(0)
    x = rng.standard_normal(K.neq)
    y = rng.standard_normal(K.neq)
    z = K.MultAdd(x, y, alpha=0.5, beta=2.0)
    zr = O.multadd(ia, ja, a, x, y, alpha=0.5, beta=2.0)
    assert np.max(np.abs(z - zr)) <= 1e-12 * np.max(np.abs(zr))


def test_csr_set_then_solve_matches_direct():
    """A caller-assembled matrix (TPZSYsmpMatrix::SetData) solved on the GPU."""
    pb = CASES["poisson3d_p2"]()
    o, ia, ja, a, rhs = _oracle(pb)
    K = TPZSYsmpMatrixHIP(ctx(), o.neq)
    K.SetData(ia, ja, a, rhs)
    st = TPZStepSolver(K)
    st.SetCG(10000, TPZStepSolver.jacobi(), 1e-12, 0)
    u = st.Solve(rhs)
    ref = O.pcg(ia, ja, a, rhs, maxit=10000, tol=1e-12, wasted_spmv=False)
    assert np.linalg.norm(u - ref["x"]) <= 1e-8 * np.linalg.norm(ref["x"])


def test_errors_are_reported():
    c = Context(0)
    with pytest.raises(L.PzgError):
        L.check(c.lib.pzg_assemble(c.h))  # before mesh/struct: PZG_ESTATE
    K = TPZSYsmpMatrixHIP(c, 3)
    with pytest.raises(L.PzgError):
        K.SetData(np.array([0, 1, 2, 3]), np.array([0, 2, 1]), np.ones(3))  # row 1 lacks diagonal
    c.close()


@pytest.mark.parametrize("name", ["elast3d_p2_traction", "poisson2d_hp", "darcy_tet_hp5"])
def test_renumbered_connects(name):
    """Any connect numbering (e.g. after TPZAnalysis::OptimizeBandwidth and
    TPZCompMesh::Permute) goes through the boundary unchanged: pattern and
    entries equal the oracle's under the same renumbering; the mixed solve
    converges."""
    pb = CASES[name]()
    md = M.build(pb)
    perm = np.random.default_rng(21).permutation(md["nconn"])
    md = M.permute(md, perm)
    print(f"[gpu] renumbered {name}", flush=True)
    c = ctx()
    sm = TPZSSpStructMatrixHIP(c, md)
    K = sm.Create()
    sm.Assemble(K)
    rhs = np.zeros(K.neq)
    L.check(c.lib.pzg_get_system(c.h, None, L.ptr(rhs, L.dp)))
    o = O.Oracle(pb)
    o.permute(perm)
    ia_r, ja_r = o.pattern()
    a_r, rhs_r = o.assemble()
    assert np.array_equal(K.IA(), ia_r) and np.array_equal(K.JA(), ja_r)
    _check_values(ia_r, K.A(), a_r)
    assert np.max(np.abs(rhs - rhs_r)) <= 1e-12 * np.max(np.abs(rhs_r))
    st = TPZStepSolver(K)
    st.SetMixedFCG(50000, 1e-8)
    st.Solve(None, want_solution=False)
    assert st.info["converged"] and st.info["true_residual"] <= 1e-8


@pytest.mark.parametrize("name", ["elast3d_p2_body", "poisson2d_hp", "darcy_tet_hp5"])
def test_sloan_optimized_analysis(name):
    """TPZLinearAnalysis(cmesh, mustOptimizeBandwidth=True): the connects are
    renumbered by the library's TPZSloanRenumbering (the oracle's permutation),
    then assembly and the FP64 PCG solution equal the oracle's on the
    renumbered mesh."""
    pb = CASES[name]()
    print(f"[gpu] Sloan-renumbered analysis {name}", flush=True)
    an = TPZLinearAnalysis(pb, True, ctx=ctx())
    o = O.Oracle(pb)
    perm = o.sloan()
    assert np.array_equal(an.fPermutation, perm)
    o.permute(perm)
    st = TPZStepSolver()
    st.SetCG(20000, TPZStepSolver.jacobi(), 1e-13, 0)
    an.SetSolver(st)
    an.Assemble()
    ia_r, ja_r = o.pattern()
    a_r, rhs_r = o.assemble()
    assert np.array_equal(an.K.IA(), ia_r) and np.array_equal(an.K.JA(), ja_r)
    _check_values(ia_r, an.K.A(), a_r)
    assert np.max(np.abs(an.Rhs() - rhs_r)) <= 1e-12 * np.max(np.abs(rhs_r))
    u = an.Solve()
    ref = O.pcg(ia_r, ja_r, a_r, rhs_r, maxit=20000, tol=1e-13, wasted_spmv=False)
    assert np.linalg.norm(u - ref["x"]) <= 1e-8 * np.linalg.norm(ref["x"])
