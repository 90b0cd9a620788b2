"""The mixed solver's Jacobi-PCG fallback: when the FP32 iterations stall the
solve continues with plain FP64 Jacobi-PCG (CG<> with the EJacobi step,
Solvers/LinearSolvers/cg.h:29-132, Matrix/pzmatrix.cpp:642-667) from the
current solution and is judged by the true FP64 residual.  The stall is
forced from the host with a refinement budget of one step whose inner
reduction cannot reach the target; the result must meet the target, report
the fallback, and agree with the plain FP64 PCG solution within 1e-6."""
import numpy as np
import pytest

from neopz_amd import _lib as L
from neopz_amd import problems as P
from neopz_amd.analysis import TPZLinearAnalysis, TPZStepSolver

from test_gpu_parity import ctx

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["elast", "darcy"])
def test_stalled_mixed_solve_falls_back_to_jacobi_pcg(name):
    pb = (P.elasticity3d(n=6, p=2, seed=3, field_kind="grf", load="body") if name == "elast"
          else P.darcy2d(n=32, p=2, seed=4))
    c = ctx()
    an = TPZLinearAnalysis(pb, False, ctx=c)
    ref = TPZStepSolver()
    ref.SetCG(100000, TPZStepSolver.jacobi(), 1e-12, 0)
    an.SetSolver(ref)
    an.Assemble()
    u_ref = an.Solve()
    st = TPZStepSolver()
    st.SetMixedFCG(100000, 1e-8, precond=L.PRECOND_BLOCK_JACOBI, inner_tol=1e-2, max_outer=1)
    an.SetSolver(st)
    u = an.Solve()
    info = st.info
    print(f"[gpu] fallback {name}: {info}", flush=True)
    assert info["fallback"] == 1
    assert info["converged"] == 1 and info["true_residual"] <= 1e-8
    assert np.linalg.norm(u - u_ref) <= 1e-6 * np.linalg.norm(u_ref)
    # without the forced stall no fallback happens
    st2 = TPZStepSolver()
    st2.SetMixedFCG(100000, 1e-8, precond=L.PRECOND_BLOCK_JACOBI)
    an.SetSolver(st2)
    an.Solve()
    assert st2.info["fallback"] == 0 and st2.info["true_residual"] <= 1e-8
