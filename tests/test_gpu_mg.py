"""GPU checks of the multigrid preconditioner of the mixed solver
(PZG_PRECOND_MULTIGRID) against the float64 model of tests/mg_model.py
built from the oracle's matrix.

Bars (written here, checked below):
  * levels: the lattice node counts equal the model's;
  * eigenvalue estimates (20 CG-Lanczos steps, from below): 0.9 lmax <=
    est <= (1 + 1e-3) lmax, lmax from scipy on the model's operators;
  * one V-cycle z = M r (with the device's estimates): |z - z_model| <=
    2e-4 |z_model| (FP32 operators and vectors through every level);
  * the mixed solve with the V-cycle: true residual <= 1e-8, the oracle's
    residual of the solution <= 2e-8 and the solution within 1e-5 of a
    direct solve of the oracle's system (the bars of the block-Jacobi mixed
    solves), in fewer iterations than block Jacobi;
  * refusals: tetrahedra.  Lattices whose halving stops above the dense
    limit (odd element counts) get a Chebyshev coarsest solve, checked the
    same way.
Cases: hexahedra (p = 1, 2; Poisson, Darcy, elasticity) and quadrilaterals
(a plane lattice: one node in z).
"""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse.linalg as spla

import mg_model as G
import oracle as O
from neopz_amd import _lib as L
from neopz_amd import problems as P
from neopz_amd.analysis import Context, TPZLinearAnalysis, TPZStepSolver

pytestmark = pytest.mark.gpu

_CTX = None


def ctx():
    global _CTX
    if _CTX is None:
        _CTX = Context(0)
    return _CTX


CASES = {
    "elast8_p2": lambda: P.elasticity3d(n=8, p=2, seed=3, field_kind="grf", load="body"),
    "poisson16_p2": lambda: P.Problem(material=P.POISSON, dim=3, p=2, nx=16, ny=16, nz=16,
                                      force_kind=P.FORCE_CONST, force_const=(1.0, 0, 0),
                                      bc_type=[0, 1, 0, -1, -1, -1]),
    "darcy8_p1": lambda: P.Problem(material=P.DARCY, dim=3, p=1, nx=8, ny=8, nz=8,
                                   el_coef=np.exp(np.random.default_rng(5).standard_normal(512)),
                                   force_kind=P.FORCE_CONST, force_const=(1.0, 0, 0), bc_type=[0] * 6),
    "elast4_p2": lambda: P.elasticity3d(n=4, p=2, seed=1, field_kind="iid", load="traction"),
    "poisson2d16_p2": lambda: P.poisson2d(n=16, p=2, seed=1),
    # odd element counts stop the halving above the dense limit: the coarsest
    # level is a degree-8 Chebyshev iteration
    "poisson11_p2": lambda: P.Problem(material=P.POISSON, dim=3, p=2, nx=11, ny=11, nz=11,
                                      force_kind=P.FORCE_CONST, force_const=(1.0, 0, 0), bc_type=[0] * 6),
    "darcy30_p1": lambda: P.Problem(material=P.DARCY, dim=3, p=1, nx=30, ny=30, nz=30,
                                    el_coef=np.exp(np.random.default_rng(6).standard_normal(27000)),
                                    force_kind=P.FORCE_CONST, force_const=(1.0, 0, 0), bc_type=[0] * 6),
    "darcy2d32_p1": lambda: P.darcy2d(n=32, p=1, seed=2),
}

_MODELS = {}


def _model(name):
    if name not in _MODELS:
        pb = CASES[name]()
        o = O.Oracle(pb)
        ia, ja = o.pattern()
        a, rhs = o.assemble()
        K = G.full_from_upper(ia, ja, a, o.neq)
        an = TPZLinearAnalysis(pb, False, ctx=ctx())
        m = G.Model(K, an.mesh, pb.nstate)
        _MODELS[name] = (pb, o, K, rhs, m, m.lmax_exact())
    return _MODELS[name]


def _mg_apply(op, r, degree=0, rng=0.0):
    z = np.zeros_like(r)
    nlev = C.c_int32(0)
    dims = np.zeros(48, dtype=np.int32)
    lmax = np.zeros(16)
    L.check(ctx().lib.pzg_mg_apply(ctx().h, op, degree, rng, L.ptr(r, L.dp), L.ptr(z, L.dp),
                                   C.byref(nlev), L.ptr(dims, L.i32p), L.ptr(lmax, L.dp)))
    nl = nlev.value
    return z, [list(dims[3 * k:3 * k + 3]) for k in range(nl)], lmax[:nl + 1]


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("op", [L.INNER_BSR, L.INNER_MATFREE])
def test_vcycle_matches_model(name, op):
    pb, o, K, rhs, m, lex = _model(name)
    print(f"[gpu] multigrid V-cycle {name} op {op}", flush=True)
    an = TPZLinearAnalysis(pb, False, ctx=ctx())
    an.Assemble()
    r = np.random.default_rng(7).standard_normal(o.neq)
    for degree, rng in ((0, 0.0), (3, 25.0)):
        z, dims, lmax = _mg_apply(op, r, degree, rng)
        assert dims == m.dims()
        lex_ = np.asarray(lex)
        # every estimate, the coarsest level's too unless it is solved densely
        # (then the device reports 0 there)
        k = len(lmax) if m.coarse_inv is None else len(lmax) - 1
        assert np.all(lmax[:k] >= 0.9 * lex_[:k]) and np.all(lmax[:k] <= (1 + 1e-3) * lex_[:k]), \
            (lmax, lex_)
        if m.coarse_inv is not None:
            assert lmax[-1] == 0.0
        zm = m.apply(r, lmax, degree or 2, rng or 30.0)
        err = np.linalg.norm(z - zm) / np.linalg.norm(zm)
        print(f"[gpu]   degree {degree} range {rng}: levels {dims}, lmax {np.round(lmax, 4)}, "
              f"rel diff {err:.2e}", flush=True)
        assert err <= 2e-4


@pytest.mark.parametrize("name", list(CASES))
def test_mixed_solve_with_multigrid(name):
    pb, o, K, rhs, m, lex = _model(name)
    print(f"[gpu] multigrid mixed solve {name}", flush=True)
    x = spla.spsolve(K.tocsc(), rhs)
    its = {}
    for pre in (L.PRECOND_BLOCK_JACOBI, L.PRECOND_MULTIGRID):
        an = TPZLinearAnalysis(pb, False, ctx=ctx())
        st = TPZStepSolver()
        st.SetMixedFCG(200000, 1e-8, precond=pre)
        an.SetSolver(st)
        an.Assemble()
        u = an.Solve()
        info = st.info
        assert info["converged"] and info["true_residual"] <= 1e-8
        # the residual bar of the mixed solves on unconstrained meshes: the
        # oracle's residual of u, and u against the direct solve as far as
        # the condition number lets a 1e-8 residual fix it
        assert np.linalg.norm(rhs - K @ u) <= 2e-8 * np.linalg.norm(rhs)
        assert np.linalg.norm(u - x) <= 1e-5 * np.linalg.norm(x)
        its[pre] = info["iterations"]
    print(f"[gpu]   iterations: block Jacobi {its[L.PRECOND_BLOCK_JACOBI]}, "
          f"multigrid {its[L.PRECOND_MULTIGRID]}", flush=True)
    assert its[L.PRECOND_MULTIGRID] < its[L.PRECOND_BLOCK_JACOBI]


def test_multigrid_refusals():
    for pb in (P.darcy3d_tet(n=2, pmax=2, seed=1, hp=False),):
        an = TPZLinearAnalysis(pb, False, ctx=ctx())
        an.Assemble()
        r = np.ones(an.mesh["neq"])
        z = np.zeros_like(r)
        rc = ctx().lib.pzg_mg_apply(ctx().h, L.INNER_BSR, 0, 0.0, L.ptr(r, L.dp), L.ptr(z, L.dp),
                                    None, None, None)
        assert rc == L.PZG_EUNSUP


@pytest.mark.parametrize("pre", [L.PRECOND_BLOCK_JACOBI, L.PRECOND_MULTIGRID])
def test_coefficient_change_rebuilds_the_preconditioner(pre):
    """pzg_mesh_set_coef + a new assembly on the same mesh: the next mixed
    solve (matrix-free operator) uses the new D^-1/2, block-Jacobi inverses
    and multigrid levels; its solution meets the bars against the oracle's
    system of the new coefficients."""
    import dataclasses
    pb = CASES["elast8_p2"]()
    print(f"[gpu] coefficient change, precond {pre}", flush=True)
    an = TPZLinearAnalysis(pb, False, ctx=ctx())
    st = TPZStepSolver()
    st.SetMixedFCG(200000, 1e-8, precond=pre)
    an.SetSolver(st)
    an.Assemble()
    an.Solve()
    coef = np.asarray(pb.el_coef) * np.exp(np.random.default_rng(4).standard_normal(pb.nvol))
    L.check(ctx().lib.pzg_mesh_set_coef(ctx().h, L.ptr(coef, L.dp)))
    an.Assemble()
    u = an.Solve()
    assert st.info["converged"] and st.info["true_residual"] <= 1e-8
    pb2 = dataclasses.replace(pb, el_coef=coef)
    o = O.Oracle(pb2)
    ia, ja = o.pattern()
    a, rhs = o.assemble()
    K = G.full_from_upper(ia, ja, a, o.neq)
    x = spla.spsolve(K.tocsc(), rhs)
    assert np.linalg.norm(rhs - K @ u) <= 2e-8 * np.linalg.norm(rhs)
    assert np.linalg.norm(u - x) <= 1e-5 * np.linalg.norm(x)


def test_hp_hexahedra_stored_operator():
    """p_e in {1, 2} on hexahedra (several integration rules: the stored FP32
    operator, no matrix-free one): the vertex functions are still the
    trilinear subspace; the V-cycle matches the model and the mixed solve
    meets the bars."""
    pb = P.elasticity3d(n=4, p=2, seed=8, field_kind="grf", load="body")
    pb.el_porder = np.random.default_rng(9).integers(1, 3, pb.nvol).astype(np.int32)
    print("[gpu] multigrid hp hexahedra", flush=True)
    o = O.Oracle(pb)
    ia, ja = o.pattern()
    a, rhs = o.assemble()
    K = G.full_from_upper(ia, ja, a, o.neq)
    an = TPZLinearAnalysis(pb, False, ctx=ctx())
    m = G.Model(K, an.mesh, pb.nstate)
    an.Assemble()
    r = np.random.default_rng(7).standard_normal(o.neq)
    z, dims, lmax = _mg_apply(L.INNER_BSR, r)
    assert dims == m.dims()
    zm = m.apply(r, lmax)
    assert np.linalg.norm(z - zm) <= 2e-4 * np.linalg.norm(zm)
    st = TPZStepSolver()
    st.SetMixedFCG(200000, 1e-8, precond=L.PRECOND_MULTIGRID)
    an.SetSolver(st)
    u = an.Solve()
    assert st.info["converged"] and st.info["true_residual"] <= 1e-8
    x = spla.spsolve(K.tocsc(), rhs)
    assert np.linalg.norm(rhs - K @ u) <= 2e-8 * np.linalg.norm(rhs)
    assert np.linalg.norm(u - x) <= 1e-5 * np.linalg.norm(x)
