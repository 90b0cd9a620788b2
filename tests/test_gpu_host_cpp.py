"""The C++ host mirror (neopz_amd/host/TPZHIP.h) driven end to end.

``pzhip_run`` runs the reference's analysis sequence (TPZLinearAnalysis with a
TPZSSpStructMatrix and a Jacobi-preconditioned TPZStepSolver CG) in its own
process over the C ABI.  Its pattern must equal the oracle's bit for bit, its
entries and rhs must meet the 1e-12 relative bar, and its CG solution must be
within 1e-8 of the oracle's CG solution (both converged to 1e-13).
"""
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle as O
from neopz_amd import mesh as M
from neopz_amd import problems as P

def _hp(pb, seed):
    pb.el_porder = np.random.default_rng(seed).integers(1, 3, pb.nvol).astype(np.int32)
    return pb


RUN = Path(__file__).resolve().parents[1] / "neopz_amd" / "host" / "pzhip_run"


def _read_result(path, extra=()):
    buf = path.read_bytes()
    off = 0

    def take(dt, n):
        nonlocal off
        a = np.frombuffer(buf, dt, n, off)
        off += a.nbytes
        return a

    neq, nnz = take(np.int64, 2)
    out = {"neq": int(neq), "nnz": int(nnz)}
    for name, dt in (("ia", np.int64), ("ja", np.int64), ("a", np.float64), ("rhs", np.float64),
                     ("u", np.float64)):
        n = int(take(np.int64, 1)[0])
        out[name] = take(dt, n)
    out["iterations"] = int(take(np.int64, 1)[0])
    out["achieved"] = float(take(np.float64, 1)[0])
    for name in extra:  # errors: ervec, el_errors; sloan: perm
        n = int(take(np.int64, 1)[0])
        out[name] = take(np.int64 if name == "perm" else np.float64, n)
    assert off == len(buf)
    return out


def test_runner_rejects_malformed_mesh(tmp_path):
    assert RUN.exists(), "build the library first (neopz_amd/csrc make)"
    bad = tmp_path / "bad.bin"
    bad.write_bytes(b"PZGMESH5" + b"\0" * 6)
    r = subprocess.run([str(RUN), str(bad), str(tmp_path / "o.bin")], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 2 and "malformed" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("name,pb", [
    ("poisson2d_p2", lambda: P.poisson2d(n=10, p=2, seed=1)),
    ("elast3d_p2_body", lambda: P.elasticity3d(n=4, p=2, seed=6, field_kind="grf", load="body")),
    ("poisson2d_hp", lambda: _hp(P.poisson2d(n=10, p=2, seed=2), 3)),
    ("darcy_tet_hp", lambda: P.darcy3d_tet(n=3, pmax=2, seed=4, hp=True)),
    ("darcy_tet_hp5", lambda: P.darcy3d_tet(n=2, pmax=5, seed=5, hp=True)),
])
def test_cpp_analysis_matches_oracle(tmp_path, name, pb):
    pb = pb()
    mfile, ofile = tmp_path / "mesh.bin", tmp_path / "out.bin"
    M.save(M.build(pb), mfile)
    print(f"[gpu] pzhip_run {name}", flush=True)
    r = subprocess.run([str(RUN), str(mfile), str(ofile), "cg", "1e-13", "20000"],
                       capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    res = _read_result(ofile)
    o = O.Oracle(pb)
    ia, ja = o.pattern()
    a, rhs = o.assemble()
    assert np.array_equal(res["ia"], ia) and np.array_equal(res["ja"], ja)
    assert np.max(np.abs(res["a"] - a) - 1e-12 * np.abs(a)) <= 1e-15 * np.max(np.abs(a))
    assert np.max(np.abs(res["rhs"] - rhs)) <= 1e-12 * np.max(np.abs(rhs))
    ref = O.pcg(ia, ja, a, rhs, maxit=20000, tol=1e-13, wasted_spmv=False)
    assert ref["converged"] and res["achieved"] <= 1e-13
    assert np.linalg.norm(res["u"] - ref["x"]) <= 1e-8 * np.linalg.norm(ref["x"])


@pytest.mark.gpu
@pytest.mark.parametrize("name,pb", [
    ("elast3d_p2_body", lambda: P.elasticity3d(n=4, p=2, seed=6, field_kind="grf", load="body")),
    ("poisson2d_p2", lambda: P.poisson2d(n=10, p=2, seed=1)),
])
def test_cpp_block_jacobi_cg_matches_oracle(tmp_path, name, pb):
    """SetCG with BuildPreconditioner(EBlockJacobi, true) in the C++ mirror:
    within 1e-8 of the oracle's CG with the LU-solved connect blocks."""
    pb = pb()
    mfile, ofile = tmp_path / "mesh.bin", tmp_path / "out.bin"
    M.save(M.build(pb), mfile)
    print(f"[gpu] pzhip_run cg-bj {name}", flush=True)
    r = subprocess.run([str(RUN), str(mfile), str(ofile), "cg-bj", "1e-13", "20000"],
                       capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    res = _read_result(ofile)
    o = O.Oracle(pb)
    ia, ja = o.pattern()
    a, rhs = o.assemble()
    ref = O.pcg_block(ia, ja, a, rhs, np.full(o.neq // pb.nstate, pb.nstate, np.int32), maxit=20000, tol=1e-13)
    assert ref["converged"] and res["achieved"] <= 1e-13
    assert abs(res["iterations"] - ref["iters"]) <= max(3, 0.05 * ref["iters"])
    assert np.linalg.norm(res["u"] - ref["x"]) <= 1e-8 * np.linalg.norm(ref["x"])


@pytest.mark.gpu
@pytest.mark.parametrize("name,pb", [
    ("elast3d_p2_body", lambda: P.elasticity3d(n=3, p=2, seed=6, field_kind="grf", load="body")),
    ("darcy_tet_hp5", lambda: P.darcy3d_tet(n=2, pmax=5, seed=5, hp=True)),
])
def test_cpp_full_format_matches_oracle(tmp_path, name, pb):
    """The C++ mirror with TPZSpStructMatrixHIP: IA/JA/A of the non-symmetric
    TPZFYsmpMatrix equal the oracle's, and the CG solution its solve."""
    pb = pb()
    mfile, ofile = tmp_path / "mesh.bin", tmp_path / "out.bin"
    M.save(M.build(pb), mfile)
    print(f"[gpu] pzhip_run full {name}", flush=True)
    r = subprocess.run([str(RUN), str(mfile), str(ofile), "cg", "1e-13", "20000", "full"],
                       capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    res = _read_result(ofile)
    o = O.Oracle(pb)
    fia, fja, fa, frhs = o.assemble_full()
    assert np.array_equal(res["ia"], fia) and np.array_equal(res["ja"], fja)
    assert np.max(np.abs(res["a"] - fa) - 1e-12 * np.abs(fa)) <= 1e-15 * np.max(np.abs(fa))
    assert np.max(np.abs(res["rhs"] - frhs)) <= 1e-12 * np.max(np.abs(frhs))
    ia, ja = o.pattern()
    a, rhs = o.assemble()
    ref = O.pcg(ia, ja, a, rhs, maxit=20000, tol=1e-13, wasted_spmv=False)
    assert np.linalg.norm(res["u"] - ref["x"]) <= 1e-8 * np.linalg.norm(ref["x"])


def _exact_np(dim, ns, x):
    """numpy statement of pzhip_run's exact_solution."""
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


@pytest.mark.gpu
@pytest.mark.parametrize("name,pb", [
    ("poisson2d_hp", lambda: _hp(P.poisson2d(n=10, p=2, seed=2), 3)),
    ("elast3d_p2_body", lambda: P.elasticity3d(n=3, p=2, seed=6, field_kind="grf", load="body")),
    ("darcy_tet_hp5", lambda: P.darcy3d_tet(n=2, pmax=5, seed=5, hp=True)),
])
def test_cpp_post_process_error_matches_oracle(tmp_path, name, pb):
    """TPZLinearAnalysis::SetExact(std::function, 6) + PostProcessError(ervec,
    true) in the C++ mirror: totals and element errors of its CG solution
    within 1e-12 of the oracle's EvaluateError on the same solution."""
    pb = pb()
    mfile, ofile = tmp_path / "mesh.bin", tmp_path / "out.bin"
    M.save(M.build(pb), mfile)
    print(f"[gpu] pzhip_run errors {name}", flush=True)
    r = subprocess.run([str(RUN), str(mfile), str(ofile), "cg", "1e-12", "20000", "sym", "errors"],
                       capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    res = _read_result(ofile, ("ervec", "el_errors"))
    o = O.Oracle(pb)
    xq = o.error_points(6)
    uq, duq = _exact_np(pb.dim, pb.nstate, xq)
    el, tot = o.evaluate_error(res["u"], kind=O.EXACT_POINTS, porder=6, uq=uq, duq=duq)
    assert np.allclose(res["ervec"], tot, rtol=1e-12, atol=0)
    el_g = res["el_errors"].reshape(el.shape)
    assert np.max(np.abs(el_g - el) - 1e-12 * np.abs(el)) <= 1e-13 * np.max(np.abs(el))


@pytest.mark.gpu
@pytest.mark.parametrize("name,pb", [
    ("poisson2d_hp", lambda: _hp(P.poisson2d(n=10, p=2, seed=2), 3)),
    ("elast3d_p2_body", lambda: P.elasticity3d(n=3, p=2, seed=6, field_kind="grf", load="body")),
    ("darcy_tet_hp5", lambda: P.darcy3d_tet(n=2, pmax=5, seed=5, hp=True)),
])
def test_cpp_sloan_renumbered_analysis_matches_oracle(tmp_path, name, pb):
    """TPZLinearAnalysis(cmesh, mustOptimizeBandwidth = true) in the C++ mirror:
    the permutation equals the oracle's TPZSloanRenumbering, and IA/JA/A/rhs of
    the renumbered system and the CG solution equal the oracle's on the
    renumbered mesh."""
    pb = pb()
    mfile, ofile = tmp_path / "mesh.bin", tmp_path / "out.bin"
    M.save(M.build(pb), mfile)
    print(f"[gpu] pzhip_run sloan {name}", flush=True)
    r = subprocess.run([str(RUN), str(mfile), str(ofile), "cg", "1e-13", "20000", "sloan"],
                       capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    res = _read_result(ofile, ("perm",))
    o = O.Oracle(pb)
    perm = o.sloan()
    assert np.array_equal(res["perm"], perm)
    o.permute(perm)
    ia, ja = o.pattern()
    a, rhs = o.assemble()
    assert np.array_equal(res["ia"], ia) and np.array_equal(res["ja"], ja)
    assert np.max(np.abs(res["a"] - a) - 1e-12 * np.abs(a)) <= 1e-15 * np.max(np.abs(a))
    assert np.max(np.abs(res["rhs"] - rhs)) <= 1e-12 * np.max(np.abs(rhs))
    ref = O.pcg(ia, ja, a, rhs, maxit=20000, tol=1e-13, wasted_spmv=False)
    assert np.linalg.norm(res["u"] - ref["x"]) <= 1e-8 * np.linalg.norm(ref["x"])


@pytest.mark.gpu
def test_cpp_constrained_mesh_matches_oracle(tmp_path):
    """A quadrilateral mesh with hanging sides through the C++ mirror: the
    dependencies travel in the mesh file; pattern, entries, rhs and the CG
    solution equal the oracle's constrained assembly."""
    pb = P.Problem(material=P.DARCY, dim=2, p=2, nx=6, ny=6,
                   el_coef=np.exp(np.random.default_rng(7).standard_normal(36)),
                   force_kind=P.FORCE_CONST, force_const=(1.0, 0, 0), bc_type=[0, 1, 0, -1, -1, -1],
                   bc_val=[[0, 0, 0], [0.4, 0, 0], [0, 0, 0], [0, 0, 0], [0, 0, 0], [0, 0, 0]])
    md = M.build_refined2d(pb, np.random.default_rng(8).random(36) < 0.4)
    mfile, ofile = tmp_path / "mesh.bin", tmp_path / "out.bin"
    M.save(md, mfile)
    print("[gpu] pzhip_run constrained", flush=True)
    r = subprocess.run([str(RUN), str(mfile), str(ofile), "cg", "1e-13", "20000"],
                       capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    res = _read_result(ofile)
    o = O.Oracle(pb, mesh=md)
    ia, ja = o.pattern()
    a, rhs = o.assemble()
    assert np.array_equal(res["ia"], ia) and np.array_equal(res["ja"], ja)
    assert np.max(np.abs(res["a"] - a) - 1e-12 * np.abs(a)) <= 1e-15 * np.max(np.abs(a))
    assert np.max(np.abs(res["rhs"] - rhs)) <= 1e-12 * np.max(np.abs(rhs))
    ref = O.pcg(ia, ja, a, rhs, maxit=20000, tol=1e-13, wasted_spmv=False)
    assert np.linalg.norm(res["u"] - ref["x"]) <= 1e-8 * np.linalg.norm(ref["x"])


@pytest.mark.gpu
def test_cpp_multigrid_mixed_solve(tmp_path):
    """TPZStepSolver::SetMixedFCG(n, tol, PZG_PRECOND_MULTIGRID) through the C++
    mirror: the solution meets the mixed solver's bars against the oracle's
    system (residual 2e-8, direct solution 1e-5)."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    pb = P.elasticity3d(n=8, p=2, seed=3, field_kind="grf", load="body")
    mfile, ofile = tmp_path / "mesh.bin", tmp_path / "out.bin"
    M.save(M.build(pb), mfile)
    print("[gpu] pzhip_run mixed-mg", flush=True)
    r = subprocess.run([str(RUN), str(mfile), str(ofile), "mixed-mg", "1e-8", "20000"],
                       capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    res = _read_result(ofile)
    o = O.Oracle(pb)
    ia, ja = o.pattern()
    a, rhs = o.assemble()
    rows = np.repeat(np.arange(o.neq), np.diff(ia))
    U = sp.csr_matrix((a, (rows, ja)), shape=(o.neq, o.neq))
    K = (U + sp.triu(U, 1).T).tocsc()
    x = spla.spsolve(K, rhs)
    assert np.linalg.norm(rhs - K @ res["u"]) <= 2e-8 * np.linalg.norm(rhs)
    assert np.linalg.norm(res["u"] - x) <= 1e-5 * np.linalg.norm(x)
    assert res["iterations"] < 200  # block Jacobi needs about 520 here


@pytest.mark.gpu
@pytest.mark.parametrize("method", ["cg-mixed", "cg-mixed-mg"])
def test_cpp_mixed_solver_through_setcg(tmp_path, method):
    """TPZStepSolver::SetCG(maxit, jac, tol, 0) with the struct matrix set to
    the mixed-precision FCG: TPZStepSolver::Solve -> SolveCG runs it, writes
    back the iterations and the achieved (true FP64) residual; the solution
    agrees with the oracle's CG solution within 1e-6."""
    pb = P.elasticity3d(n=4, p=2, seed=6, field_kind="grf", load="body")
    mfile, ofile = tmp_path / "mesh.bin", tmp_path / "out.bin"
    M.save(M.build(pb), mfile)
    print(f"[gpu] pzhip_run {method}", flush=True)
    r = subprocess.run([str(RUN), str(mfile), str(ofile), method, "1e-8", "20000"],
                       capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    res = _read_result(ofile)
    assert 0 < res["iterations"] < 20000 and res["achieved"] <= 1e-8
    o = O.Oracle(pb)
    ia, ja = o.pattern()
    a, rhs = o.assemble()
    ref = O.pcg(ia, ja, a, rhs, maxit=20000, tol=1e-12, wasted_spmv=False)
    assert np.linalg.norm(res["u"] - ref["x"]) <= 1e-6 * np.linalg.norm(ref["x"])


@pytest.mark.gpu
def test_cpp_equation_filter(tmp_path):
    """The struct matrix's TPZEquationFilter::SetMinMaxEq through the C++
    mirror: condensed IA/JA/A equal the oracle's filtered SetupMatrixData and
    assembly, the rhs comes back in the mesh numbering and the CG solution
    of the condensed system matches the oracle's."""
    pb = P.darcy2d(n=8, p=2, seed=3)
    mfile, ofile = tmp_path / "mesh.bin", tmp_path / "out.bin"
    M.save(M.build(pb), mfile)
    o = O.Oracle(pb)
    lo, hi = 7, o.neq - 5
    print("[gpu] pzhip_run minmax", flush=True)
    r = subprocess.run([str(RUN), str(mfile), str(ofile), "cg", "1e-13", "20000", f"minmax:{lo}:{hi}"],
                       capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    res = _read_result(ofile)
    act = np.arange(lo, hi)
    fia, fja, fa, frhs = o.filtered(act)
    assert np.array_equal(res["ia"], fia) and np.array_equal(res["ja"], fja)
    assert np.max(np.abs(res["a"] - fa) - 1e-12 * np.abs(fa)) <= 1e-15 * np.max(np.abs(fa))
    assert len(res["rhs"]) == o.neq
    assert np.max(np.abs(res["rhs"] - frhs)) <= 1e-12 * np.max(np.abs(frhs))
    ref = O.pcg(fia, fja, fa, frhs[act], maxit=20000, tol=1e-13, wasted_spmv=False)
    u = res["u"]
    assert len(u) == o.neq and np.all(u[:lo] == 0) and np.all(u[hi:] == 0)
    assert np.linalg.norm(u[act] - ref["x"]) <= 1e-8 * np.linalg.norm(ref["x"])
