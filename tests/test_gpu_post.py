"""GPU parity of the post-processing entry points (pzg_load_solution,
pzg_error_points, pzg_evaluate_error) against the oracle's restatement of
TPZCompMesh::EvaluateError.

Bars (written here, checked below):
  * error-rule points: |x - x_ref| <= 1e-14 max|x_ref| (same rule, same map);
  * per-element errors and mesh totals: |e - e_ref| <= 1e-12 |e_ref| + 1e-13
    max|e_ref| (sums of squares of FP64 values over the same points; the
    floor covers elements whose error is a cancellation residue);
  * at the full sizes of the 64^3 p = 2 elasticity and 32^3 hp tetrahedra
    workloads, size-independent properties: the interpolant of a linear
    field has zero error, and the zero solution's errors equal closed forms.
"""
import time

import numpy as np
import pytest

import oracle as O
from neopz_amd import _lib as L
from neopz_amd import mesh as M
from neopz_amd import problems as P
from neopz_amd.analysis import Context, TLaplaceExample1, TPZLinearAnalysis, TPZStepSolver

pytestmark = pytest.mark.gpu


def _hp(pb, seed):
    pb.el_porder = np.random.default_rng(seed).integers(1, 3, pb.nvol).astype(np.int32)
    return pb


CASES = {
    "poisson2d_p2": lambda: P.poisson2d(n=10, p=2, seed=1),
    "poisson2d_hp": lambda: _hp(P.poisson2d(n=10, p=2, seed=2), 3),
    "darcy2d_p1": lambda: P.darcy2d(n=12, p=1, seed=2),
    "elast3d_p2": lambda: P.elasticity3d(n=4, p=2, seed=6, field_kind="grf", load="body"),
    "elast3d_hp": lambda: _hp(P.elasticity3d(n=3, p=2, seed=7, field_kind="grf", load="body"), 4),
    "darcy_tet_p2": lambda: P.darcy3d_tet(n=3, pmax=2, seed=12, hp=False),
    "darcy_tet_hp5": lambda: P.darcy3d_tet(n=2, pmax=5, seed=5, hp=True),
    "elast_tet_p2": lambda: P.Problem(material=P.ELAST3D, dim=3, p=2, nx=2, ny=2, nz=2, simplex=1,
                                      E=2.0, nu=0.27, force_kind=P.FORCE_CONST,
                                      force_const=(0.0, 0.0, -1.0), bc_type=[0] * 6),
    "elast_tet_hp4": lambda: P.Problem(material=P.ELAST3D, dim=3, p=4, nx=2, ny=1, nz=2, simplex=1,
                                       E=2.0, nu=0.27, force_kind=P.FORCE_CONST,
                                       force_const=(0.0, 0.0, -1.0), bc_type=[0] * 6,
                                       el_porder=np.random.default_rng(9).integers(1, 5, 24).astype(np.int32)),
}

_CTX = None


def ctx():
    global _CTX
    if _CTX is None:
        _CTX = Context(0)
    return _CTX


def _exact_np(dim, ns, x):
    """Host exact solution: scalar prod sin(pi x_i); elasticity
    u_j = (j+1) sin(pi x_j) + x0 x1 x2.  du[q, i, j] = d u_j / d x_i."""
    x = np.asarray(x)
    if ns == 1:
        s, c = np.sin(np.pi * x[:, :dim]), np.pi * np.cos(np.pi * x[:, :dim])
        u = np.prod(s, 1)[:, None]
        du = np.stack([c[:, i] * np.prod(np.delete(s, i, 1), 1) for i in range(dim)], 1)[:, :, None]
        return u, du
    u = np.stack([(j + 1) * np.sin(np.pi * x[:, j]) + x[:, 0] * x[:, 1] * x[:, 2] for j in range(3)], 1)
    prod = np.stack([x[:, 1] * x[:, 2], x[:, 0] * x[:, 2], x[:, 0] * x[:, 1]], 1)
    du = np.repeat(prod[:, :, None], 3, axis=2)
    for j in range(3):
        du[:, j, j] += (j + 1) * np.pi * np.cos(np.pi * x[:, j])
    return u, du


def _close(e, e_ref):
    floor = 1e-13 * max(np.max(np.abs(e_ref)), 1e-300)
    bad = np.abs(e - e_ref) > 1e-12 * np.abs(e_ref) + floor
    assert not bad.any(), (np.argwhere(bad)[:5], e[bad][:5], e_ref[bad][:5])


@pytest.mark.parametrize("name", list(CASES))
def test_error_points_match_oracle(name):
    pb = CASES[name]()
    print(f"[gpu] error points {name}", flush=True)
    an = TPZLinearAnalysis(pb, False, ctx=ctx())
    for porder in (0, 5, 40):
        an.SetExact(lambda x: None, porder)
        xq = an.ErrorPoints()
        xr = O.Oracle(pb).error_points(porder)
        assert xq.shape == xr.shape
        assert np.max(np.abs(xq - xr)) <= 1e-14 * np.max(np.abs(xr))


@pytest.mark.parametrize("name", list(CASES))
def test_errors_of_solution_match_oracle(name):
    """Assemble, solve (FP64 PCG, the loaded solution), PostProcessError with a
    host exact solution and, for scalar problems, a device TLaplaceExample1;
    then a random LoadSolution."""
    pb = CASES[name]()
    print(f"[gpu] evaluate error {name}", flush=True)
    an = TPZLinearAnalysis(pb, False, ctx=ctx())
    st = TPZStepSolver()
    st.SetCG(20000, TPZStepSolver.jacobi(), 1e-11, 0)
    an.SetSolver(st)
    an.Assemble()
    u = an.Solve()
    o = O.Oracle(pb)
    ns, dim = pb.nstate, pb.dim
    porder = 6
    an.SetExact(lambda x: _exact_np(dim, ns, x), porder)
    tot = an.PostProcessError(store_error=True)
    el = an.ElementSolution()
    xr = o.error_points(porder)
    uq, duq = _exact_np(dim, ns, xr)
    el_r, tot_r = o.evaluate_error(u, kind=O.EXACT_POINTS, porder=porder, uq=uq, duq=duq)
    _close(el, el_r)
    _close(tot, tot_r)
    if ns == 1:
        ex = TLaplaceExample1(TLaplaceExample1.ECosCos, center=(0.1, -0.2, 0.3))
        an.SetExact(ex.ExactSolution(), 3)
        tot = an.PostProcessError(store_error=True)
        el_r, tot_r = o.evaluate_error(u, kind=O.EXACT_COSCOS, porder=3, center=(0.1, -0.2, 0.3))
        _close(an.ElementSolution(), el_r)
        _close(tot, tot_r)
    # a caller-loaded solution (TPZCompMesh::LoadSolution)
    v = np.random.default_rng(3).standard_normal(len(u))
    an.LoadSolution(v)
    back = np.zeros(len(v))
    L.check(ctx().lib.pzg_get_solution(ctx().h, L.ptr(back, L.dp)))
    assert np.array_equal(back, v)
    an.SetExact(lambda x: _exact_np(dim, ns, x), 0)
    tot = an.PostProcessError(store_error=True)
    xr = o.error_points(0)
    uq, duq = _exact_np(dim, ns, xr)
    el_r, tot_r = o.evaluate_error(v, kind=O.EXACT_POINTS, porder=0, uq=uq, duq=duq)
    _close(an.ElementSolution(), el_r)
    _close(tot, tot_r)


def test_errors_are_reported():
    pb = CASES["elast3d_p2"]()
    an = TPZLinearAnalysis(pb, False, ctx=ctx())
    an.SetExact(TLaplaceExample1(TLaplaceExample1.ESinSin), 2)
    with pytest.raises(L.PzgError) as e:  # no solution loaded yet
        an.PostProcessError()
    assert e.value.code == L.PZG_ESTATE
    an.LoadSolution(np.zeros(an.mesh["neq"]))
    with pytest.raises(L.PzgError) as e:  # device kinds are scalar
        an.PostProcessError()
    assert e.value.code == L.PZG_EINVAL
    an.SetExact(lambda x: (np.zeros((len(x) - 1, 3)), np.zeros((len(x) - 1, 3, 3))), 2)
    with pytest.raises((L.PzgError, ValueError)):  # wrong number of points
        an.PostProcessError()


def test_c5_elasticity_full_size_properties():
    """64^3 p = 2 heterogeneous elasticity (6.4 M equations): the interpolant
    of u = G^T x + b has zero error; the zero solution's energy, L2 and H1
    semi-norms equal Lame's closed forms with the element-wise E."""
    pb = P.config("C5")
    print("[gpu] C5 error norms", flush=True)
    an = TPZLinearAnalysis(pb, False, ctx=ctx())
    m = an.mesh
    rng = np.random.default_rng(11)
    G = rng.standard_normal((3, 3))
    b = rng.standard_normal(3)
    u = np.zeros(m["neq"])
    vn = np.asarray(m["vol_nodes"]).reshape(pb.nvol, -1)
    vc = np.asarray(m["vol_conn"]).reshape(pb.nvol, -1)
    xyz = np.asarray(m["xyz"]).reshape(-1, 3)
    pos = np.asarray(m["blk_pos"])[vc[:, :8]].ravel()
    vals = xyz[vn[:, :8].ravel()] @ G + b
    for j in range(3):
        u[pos + j] = vals[:, j]
    an.LoadSolution(u)
    an.SetExact(lambda x: (x @ G + b, np.broadcast_to(G, (len(x), 3, 3))), 2)
    t0 = time.perf_counter()
    tot = an.PostProcessError()
    print(f"[gpu] C5 PostProcessError (host exact values at the points) "
          f"{time.perf_counter() - t0:.3f} s", flush=True)
    ref = np.sqrt(np.sum(G * G))
    assert np.all(np.abs(tot) <= 1e-10 * ref), tot
    an.LoadSolution(np.zeros(m["neq"]))
    tot = an.PostProcessError()
    nu = pb.nu
    eps = 0.5 * (G + G.T)
    vol = 1.0 / pb.nvol
    lam = pb.el_coef * nu / ((1 + nu) * (1 - 2 * nu))
    mu = pb.el_coef / (2 * (1 + nu))
    energy = np.sum(vol * (lam * np.trace(eps) ** 2 + 2 * mu * np.sum(eps * eps)))
    Mm = np.full((3, 3), 0.25) + np.eye(3) * (1 / 3 - 0.25)
    l2 = sum(G[:, j] @ Mm @ G[:, j] + b[j] * G[:, j].sum() + b[j] ** 2 for j in range(3))
    assert abs(tot[0] ** 2 - energy) <= 1e-10 * energy
    assert abs(tot[1] ** 2 - l2) <= 1e-10 * l2
    assert abs(tot[2] ** 2 - np.sum(G * G)) <= 1e-10 * np.sum(G * G)


def test_c3_tet_hp_full_size_closed_form():
    """32^3 x 6 tetrahedra with p_e in 1..5: the zero solution against
    u = prod sin(pi x_i) evaluated on the device, rule raised to order 14:
    ||u||^2 = 1/8 and each flux term pi^2/8 (TPZDarcyFlow's flux terms with
    u_h = 0 are the squared exact derivatives)."""
    pb = P.config("C3")
    print("[gpu] C3 error norms", flush=True)
    an = TPZLinearAnalysis(pb, False, ctx=ctx())
    an.LoadSolution(np.zeros(an.mesh["neq"]))
    an.SetExact(TLaplaceExample1(TLaplaceExample1.ESinSin), 14)
    t0 = time.perf_counter()
    tot = an.PostProcessError()
    print(f"[gpu] C3 PostProcessError (device exact solution) {time.perf_counter() - t0:.3f} s",
          flush=True)
    assert abs(tot[1] ** 2 - 0.125) <= 1e-9
    for i in range(3):
        assert abs(tot[3 + i] ** 2 - np.pi ** 2 / 8) <= 1e-9 * np.pi ** 2
