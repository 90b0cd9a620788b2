"""GPU parity of the boundary-condition types beyond Dirichlet/Neumann
(TPZMatPoisson type 2, TPZDarcyFlow type 2, TPZElasticity3D types 2-8) and of
TPZElasticity3D's prestress load against the oracle, with the bars of test_gpu_parity (pattern bit-exact,
entries 1e-12 relative, rhs 1e-12 of max|rhs|), and the refusal of types
outside each material's ContributeBC switch.  The mixed solver's FP32
operator falls back from matrix-free to the stored block CSR when a
boundary matrix is not the Dirichlet penalty; the solve still meets 1e-8."""
import numpy as np
import pytest

import oracle as O
from neopz_amd import _lib as L
from neopz_amd import mesh as M
from neopz_amd import problems as P
from neopz_amd.analysis import Context, TPZLinearAnalysis, TPZSSpStructMatrixHIP, TPZStepSolver

from test_gpu_parity import _assemble, _check_values, _oracle, ctx

pytestmark = pytest.mark.gpu

S = [[1.0, 0.3, -0.2, 0.3, -0.5, 0.7, -0.2, 0.7, 2.0]]
A = [[3.0, 0.5, -0.25, 0.5, 2.0, 0.125, -0.25, 0.125, 1.0]]


def _elast(types, vals=None, val1=None, n=3, p=2, simplex=0, seed=5):
    pb = P.elasticity3d(n=n, p=p, seed=seed, field_kind="grf", load="body")
    pb.simplex = simplex
    if simplex:
        pb.el_coef = np.exp(np.random.default_rng(seed).standard_normal(pb.nvol))
    pb.bc_type = list(types)
    pb.bc_val = [list(v) for v in (vals or [[0, 0, 0]] * 6)]
    pb.bc_val1 = [list(v) for v in (val1 or [[0.0] * 9] * 6)]
    pb.name = f"elast_bc_{types}"
    return pb


def _scalar(material, types, vals, val1, p=2, n=6, simplex=0):
    nv = n * n if not simplex else 6 * n ** 3
    pb = P.Problem(material=material, dim=3 if simplex else 2, p=p, nx=n, ny=n, nz=n if simplex else 1,
                   simplex=simplex, scale=1.25,
                   el_coef=np.exp(np.random.default_rng(7).standard_normal(nv)) if material == P.DARCY else None,
                   force_kind=P.FORCE_CONST, force_const=(1.0, 0, 0), bc_type=list(types),
                   bc_val=[list(v) for v in vals], bc_val1=[list(v) for v in val1])
    pb.name = f"scalar_bc_{material}_{types}"
    return pb


def _prestress(pb):
    pb.prestress = (0.7, -0.3, 0.2)  # TPZElasticity3D::fPreStress
    pb.name += "_prestress"
    return pb


Z3 = [0, 0, 0]
Z9 = [0.0] * 9
CASES = {
    # every Elasticity3D type on a face of its own
    "elast_types_0_1_2_3_4_5": lambda: _elast(
        [0, 1, 2, 3, 4, 5],
        [Z3, [0.2, -0.1, 0.3], [0.1, 0.0, -0.4], [1.0, 0.0, 1.0], Z3, [0.25, 0, 0]],
        [Z9, Z9, A[0], Z9, S[0], Z9]),
    "elast_types_6_7_8": lambda: _elast(
        [6, 7, 8, -1, 0, 4],
        [[0, 0.5, 0], [0, 0, -0.5], [0.3, 0, 0.2], Z3, Z3, Z3],
        [Z9, Z9, Z9, Z9, Z9, S[0]]),
    "elast_types_p1": lambda: _elast([0, 2, 4, 7, -1, 3], [Z3, [0, 1, 0], Z3, [0, 0, 2.0], Z3, [0, 1, 1]],
                                     [Z9, A[0], S[0], Z9, Z9, Z9], n=4, p=1),
    "elast_types_tet": lambda: _elast([0, 2, 4, 5, -1, 8], [Z3, [0, 1, 0], Z3, [1.0, 0, 0], Z3, [0.5, 0, 0.5]],
                                      [Z9, A[0], S[0], Z9, Z9, Z9], n=2, p=2, simplex=1),
    # order-3 tetrahedra (hp boundary kernel, 3 states): the boundary mass
    # types and the stress field
    "elast_types_tet_p3": lambda: _elast([0, 2, 4, 5, 3, 8],
                                         [Z3, [0, 1, 0], Z3, [1.0, 0, 0], [1.0, 0, 1.0], [0.5, 0, 0.5]],
                                         [Z9, A[0], S[0], Z9, Z9, Z9], n=2, p=3, simplex=1),
    "elast_prestress_hex": lambda: _prestress(_elast([0, 1, -1, 2, -1, 4], [Z3, [0.1, 0.2, 0.3], Z3, Z3, Z3, Z3],
                                                     [Z9, Z9, Z9, A[0], Z9, S[0]], n=3, p=2)),
    "elast_prestress_tet": lambda: _prestress(_elast([0, 1, -1, -1, -1, -1], n=2, p=2, simplex=1)),
    "darcy_robin_q2": lambda: _scalar(P.DARCY, [0, 2, 1, 2, -1, -1],
                                      [Z3, [0.5, 0, 0], [0.25, 0, 0], [-1.0, 0, 0], Z3, Z3],
                                      [Z9, [2.0] + [0.0] * 8, Z9, [0.5] + [0.0] * 8, Z9, Z9]),
    "darcy_robin_q1": lambda: _scalar(P.DARCY, [2, 2, 0, 2, -1, -1],
                                      [[1.0, 0, 0], [0.5, 0, 0], Z3, Z3, Z3, Z3],
                                      [[3.0] + [0.0] * 8, [2.0] + [0.0] * 8, Z9, [0.5] + [0.0] * 8, Z9, Z9],
                                      p=1, n=9),
    "darcy_robin_tet_p3": lambda: _scalar(P.DARCY, [0, 2, 2, 0, 0, 2],
                                          [Z3, [0.5, 0, 0], Z3, Z3, Z3, [1.0, 0, 0]],
                                          [Z9, [2.0] + [0.0] * 8, [1.0] + [0.0] * 8, Z9, Z9, [0.25] + [0.0] * 8],
                                          p=3, n=2, simplex=1),
    "poisson_robin_q2": lambda: _scalar(P.POISSON, [0, 2, 1, 0, -1, -1],
                                        [Z3, [0.5, 0, 0], [0.25, 0, 0], Z3, Z3, Z3],
                                        [Z9, [2.0] + [0.0] * 8, Z9, Z9, Z9, Z9]),
    "poisson_robin_tet_p2": lambda: _scalar(P.POISSON, [0, 2, 0, 0, 0, 0],
                                            [Z3, [0.5, 0, 0], Z3, Z3, Z3, Z3],
                                            [Z9, [2.0] + [0.0] * 8, Z9, Z9, Z9, Z9], p=2, n=2, simplex=1),
}


@pytest.mark.parametrize("name", list(CASES))
def test_bc_types_match_oracle(name):
    pb = CASES[name]()
    K, ia, ja, a, rhs = _assemble(pb)
    o, ia_r, ja_r, a_r, rhs_r = _oracle(pb)
    assert np.array_equal(ia, ia_r) and np.array_equal(ja, ja_r)
    _check_values(ia, a, a_r)
    assert np.max(np.abs(rhs - rhs_r)) <= 1e-12 * np.max(np.abs(rhs_r))


@pytest.mark.parametrize("name", ["elast_types_0_1_2_3_4_5", "darcy_robin_q2"])
def test_bc_types_mixed_solve(name):
    """Mixed FP32 FCG + FP64 refinement on a system with general boundary
    matrices (stored FP32 operator): true residual <= 1e-8."""
    pb = CASES[name]()
    c = ctx()
    an = TPZLinearAnalysis(pb, False, ctx=c)
    st = TPZStepSolver()
    st.SetMixedFCG(20000, 1e-8, precond=L.PRECOND_BLOCK_JACOBI)
    an.SetSolver(st)
    an.Assemble()
    an.Solve()
    assert st.info["converged"] == 1 and st.info["true_residual"] <= 1e-8, st.info


@pytest.mark.parametrize("material,t", [(P.DARCY, 3), (P.POISSON, 4), (P.ELAST3D, 9), (P.DARCY, -7)])
def test_unknown_bc_type_is_refused(material, t):
    if material == P.ELAST3D:
        pb = _elast([0, 1, -1, -1, -1, -1], n=2)
    else:
        pb = _scalar(material, [0, 1, -1, -1, -1, -1], [Z3] * 6, [Z9] * 6, p=1, n=3)
    md = M.build(pb)
    md["bc_type"] = np.array(md["bc_type"], dtype=np.int32)
    md["bc_type"][1] = t
    c = ctx()
    d, keep = L.make_mesh_desc(md)
    rc = c.lib.pzg_mesh_set(c.h, L.C.byref(d))
    assert rc == L.PZG_EUNSUP, rc
    assert b"ContributeBC" in c.lib.pzg_last_error()
