"""Parity at the full sizes of the benchmark configurations (C1 64^2 Poisson
p=2, C2 48^3 elasticity p=2, C3 32^3x6 hp tetrahedra p 1..5, C4 256^2 Darcy
p=1 and p=2, C5 64^3 heterogeneous elasticity p=2), where the oracle cannot
assemble the whole system inside a test:

* the number of stored entries equals the closed form of the structured
  quad/hex pattern (orc_nnz_closed_form), and for C1 and C4 the whole IA/JA
  pattern equals the oracle's (TPZSSpStructMatrix::SetupMatrixData,
  StrMatrix/TPZSSpStructMatrix.cpp:49-193);
* C1 and C4: the whole system against the oracle's assembly; C2, C3, C5: at
  least 10^4 seeded rows (every equation of randomly drawn elements, the
  first and the last element among them): their stored columns, values and
  loads against the oracle's AddKel/AddFel over the elements touching them
  (oracle.assemble_rows: the same terms in the same element order as its whole
  assembly, Matrix/pzsysmp.cpp:369-411).  Entries agree within 1e-12
  relative, or within 1e-14 of the row's largest entry for the entries that
  are cancellations of much larger element terms (the summation order of the
  coloured assembly differs from the element order; measured at C5: 6.6e-15);
* the mixed solve (FP32 FCG + FP64 refinement) reaches the FP64 true residual
  1e-8 and lands within 1e-6 of the FP64 Jacobi-PCG solution (CG<>,
  Solvers/LinearSolvers/cg.h:29-132, run to a recursive residual of 1e-10;
  on C4p2 its true residual stalls near 2e-8, the limit of its recursive
  residual update there); on C4p2 also with the block-Jacobi inner
  preconditioner, whose FP32 iterations stall there and hand over to FP64
  Jacobi-PCG."""
import numpy as np
import pytest

import oracle as O
from neopz_amd import _lib as L
from neopz_amd import mesh as M
from neopz_amd import problems as P
from neopz_amd.analysis import Context, TPZLinearAnalysis, TPZSSpStructMatrixHIP, TPZStepSolver

pytestmark = pytest.mark.gpu

NAMES = ["C1", "C2", "C3", "C4p1", "C4p2", "C5"]
WHOLE = {"C1", "C4p1", "C4p2"}
STRUCTURED = {"C1", "C2", "C4p1", "C4p2", "C5"}
NSAMPLE_EL = {"C2": 200, "C3": 1200, "C5": 200}


def _sample_rows(o, nel_vol, name):
    """Every equation of seeded volume elements (the first and the last
    included): at least 10^4 rows."""
    rng = np.random.default_rng(12345)
    els = np.unique(np.concatenate([[0, nel_vol - 1],
                                    rng.choice(nel_vol, NSAMPLE_EL[name], replace=False)]))
    rows = np.unique(np.concatenate([o.calcstiff(int(e))[2] for e in els]))
    return rows


def _row_slices(ia, rows):
    """Index array of the stored entries of the given rows, and their sub-IA."""
    lens = ia[rows + 1] - ia[rows]
    ia_s = np.concatenate([[0], np.cumsum(lens)])
    idx = np.repeat(ia[rows] - ia_s[:-1], lens) + np.arange(ia_s[-1])
    return idx, ia_s


@pytest.mark.parametrize("name", NAMES)
def test_full_size_pattern_and_sampled_rows(name):
    pb = P.config(name, seed=0)
    print(f"[gpu] full-size parity {name}", flush=True)
    md = M.build(pb)
    c = Context(0)
    sm = TPZSSpStructMatrixHIP(c, md)
    K = sm.Create()
    sm.Assemble(K)
    sizes = c.sizes()
    o = O.Oracle(pb)
    assert o.neq == sizes["neq"]
    ia = K.IA().copy()
    assert ia[-1] == sizes["nnz_upper"]
    if name in STRUCTURED:
        nnz_cf = O.nnz_closed_form(pb.dim, pb.p, pb.nstate, pb.nx, pb.ny, pb.nz)
        assert sizes["nnz_upper"] == nnz_cf, (sizes["nnz_upper"], nnz_cf)
    ja = K.JA()
    rhs = np.zeros(sizes["neq"])
    L.check(c.lib.pzg_get_system(c.h, None, L.ptr(rhs, L.dp)))
    if name in WHOLE:
        ia_o, ja_o = o.pattern()
        assert np.array_equal(ia, ia_o) and np.array_equal(ja, ja_o)
        rows = np.arange(o.neq)
        ia_s, ja_s, a_s = ia, ja, K.A().copy()
        a_o, rhs_o = o.assemble()
        nel = o.m.nel
    else:
        rows = _sample_rows(o, o.m.nvol, name)
        idx, ia_s = _row_slices(ia, rows)
        ja_s = np.ascontiguousarray(ja[idx])
        del ja
        a_s = np.ascontiguousarray(K.A()[idx])
        a_o, rhs_o, nel = O.assemble_rows(o, rows, ia_s, ja_s)
    assert len(rows) >= 10_000, len(rows)
    print(f"[gpu]   {len(rows)} rows, {len(ja_s)} entries, {nel} elements integrated by the oracle",
          flush=True)
    # entries: within 1e-12 relative (floor 1e-15 of the row's largest entry)
    rowmax = np.maximum.reduceat(np.abs(a_o), ia_s[:-1])
    rr = np.repeat(np.arange(len(rows)), np.diff(ia_s))
    d = np.abs(a_s - a_o)
    print(f"[gpu]   max |diff|/|entry| {np.max(d / np.maximum(np.abs(a_o), 1e-300)):.2e}, "
          f"max |diff|/rowmax {np.max(d / rowmax[rr]):.2e}", flush=True)
    bad = d > 1e-12 * np.abs(a_o) + 1e-14 * rowmax[rr]
    if bad.any():
        k = np.flatnonzero(bad)
        print("[gpu]   worst:", [(float(a_s[i]), float(a_o[i]), float(rowmax[rr[i]])) for i in k[np.argsort(-d[k])][:5]],
              flush=True)
    assert not bad.any(), (np.flatnonzero(bad)[:5], a_s[bad][:3], a_o[bad][:3])
    assert np.max(np.abs(rhs[rows] - rhs_o)) <= 1e-12 * max(np.max(np.abs(rhs)), 1e-300)
    c.close()


SOLVE_CASES = [("C1", L.PRECOND_MULTIGRID), ("C2", L.PRECOND_MULTIGRID), ("C3", L.PRECOND_BLOCK_JACOBI),
               ("C4p1", L.PRECOND_MULTIGRID), ("C4p2", L.PRECOND_MULTIGRID),
               ("C4p2", L.PRECOND_BLOCK_JACOBI), ("C5", L.PRECOND_MULTIGRID)]


@pytest.mark.parametrize("name,pre", SOLVE_CASES)
def test_full_size_mixed_solve_meets_target_and_matches_pcg64(name, pre):
    pb = P.config(name, seed=0)
    print(f"[gpu] full-size mixed solve {name} precond {pre}", flush=True)
    c = Context(0)
    an = TPZLinearAnalysis(pb, False, ctx=c)
    ref = TPZStepSolver()
    ref.SetCG(200000, TPZStepSolver.jacobi(), 1e-10, 0)
    ref.check_every = 100
    an.SetSolver(ref)
    an.Assemble()
    u_ref = an.Solve()
    assert ref.info["converged"] == 1 and ref.info["true_residual"] <= 5e-8, ref.info
    st = TPZStepSolver()
    st.SetMixedFCG(200000, 1e-8, precond=pre)
    an.SetSolver(st)
    u = an.Solve()
    print(f"[gpu]   PCG64 {ref.info['iterations']} it, mixed {st.info['iterations']} it, "
          f"{st.info['outer']} refinements, fallback {st.info['fallback']}, "
          f"true residual {st.info['true_residual']:.2e}", flush=True)
    assert st.info["converged"] == 1 and st.info["true_residual"] <= 1e-8, st.info
    assert np.linalg.norm(u - u_ref) <= 1e-6 * np.linalg.norm(u_ref)
    c.close()
