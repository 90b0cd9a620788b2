"""Host-side mirror of NeoPZ's analysis interface over the HIP C ABI.

Names, argument meaning and error behaviour follow the reference classes on
the path, so that a test reads like NeoPZ's own:

* :class:`TPZSSpStructMatrixHIP` -- ``TPZSSpStructMatrix``
  (``StrMatrix/TPZSSpStructMatrix.h:17-38``): ``Create()`` builds the upper
  CSR pattern, ``Assemble(K, rhs)`` integrates and scatters on the GPU.
* :class:`TPZSYsmpMatrixHIP` -- ``TPZSYsmpMatrix`` (``Matrix/pzsysmp.h:20-230``):
  ``IA/JA/A``, ``GetVal``, ``MultAdd``, ``SolveCG``.
* :class:`TPZStepSolver` -- ``SetCG``/``SetJacobi``/``Solve``
  (``Solvers/pzstepsolver.cpp:65-122,190-234``), plus ``SetMixedFCG`` for the
  mixed-precision path.
* :class:`TPZLinearAnalysis` -- ``Assemble()``/``Solve()``
  (``Analysis/TPZLinearAnalysis.cpp:42-227``).

The reference stops with ``DebugStop()`` on errors; here every non-zero
status of the C ABI raises :class:`neopz_amd._lib.PzgError`.
"""
import ctypes as C

import numpy as np

from . import _lib as L
from . import mesh as M


class Context:
    """One GPU context (device buffers + stream) of libpzg."""

    def __init__(self, device=0):
        self.lib = L.load()
        self.h = C.c_void_p()
        L.check(self.lib.pzg_create(device, C.byref(self.h)))

    def close(self):
        if self.h:
            self.lib.pzg_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sizes(self):
        s = L.Sizes()
        L.check(self.lib.pzg_get_sizes(self.h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in L.Sizes._fields_}


class TPZSYsmpMatrixHIP:
    """Symmetric CSR (upper triangle + diagonal) living on the GPU."""

    def __init__(self, ctx, neq):
        self.ctx = ctx
        self.neq = neq
        self._ia = self._ja = None

    def Rows(self):
        return self.neq

    def SetData(self, ia, ja, a, rhs=None):
        """TPZSYsmpMatrix::SetData (Matrix/pzsysmp.h:232-240)."""
        ia = np.ascontiguousarray(ia, np.int64)
        ja = np.ascontiguousarray(ja, np.int64)
        a = np.ascontiguousarray(a, np.float64)
        r = None if rhs is None else np.ascontiguousarray(rhs, np.float64)
        L.check(self.ctx.lib.pzg_csr_set(self.ctx.h, len(ia) - 1, L.ptr(ia, L.i64p),
                                         L.ptr(ja, L.i64p), L.ptr(a, L.dp), L.ptr(r, L.dp)))
        self.neq = len(ia) - 1
        self._ia, self._ja = ia, ja

    def IA(self):
        if self._ia is None:
            self._pattern()
        return self._ia

    def JA(self):
        if self._ja is None:
            self._pattern()
        return self._ja

    def _pattern(self):
        s = self.ctx.sizes()
        ia = np.zeros(self.neq + 1, np.int64)
        ja = np.zeros(max(s["nnz_upper"], 1), np.int64)
        L.check(self.ctx.lib.pzg_struct_get_pattern(self.ctx.h, L.ptr(ia, L.i64p), L.ptr(ja, L.i64p)))
        # with an equation filter the condensed pattern is shorter than the
        # device's mesh-numbered one
        self._ia, self._ja = ia, ja[: ia[-1]]

    def A(self):
        a = np.zeros(max(len(self.JA()), 1))
        L.check(self.ctx.lib.pzg_get_system(self.ctx.h, L.ptr(a, L.dp), None))
        return a[: len(self.JA())]

    def GetVal(self, r, c):
        """TPZSYsmpMatrix::GetVal (Matrix/pzsysmp.cpp:59-77): lower entries read the transpose."""
        if r > c:
            r, c = c, r
        ia, ja = self.IA(), self.JA()
        row = ja[ia[r]:ia[r + 1]]
        k = np.searchsorted(row, c)
        if k < len(row) and row[k] == c:
            return float(self.A()[ia[r] + k])
        return 0.0

    def AddKel(self, ek_batch, dest_batch):
        """Batched TPZSYsmpMatrix::AddKel (Matrix/pzsysmp.cpp:369-411)."""
        ek = np.ascontiguousarray(ek_batch, np.float64)
        dest = np.ascontiguousarray(dest_batch, np.int64)
        nel, ne = dest.shape
        L.check(self.ctx.lib.pzg_add_kel_batch(self.ctx.h, nel, ne, L.ptr(dest, L.i64p),
                                               L.ptr(ek, L.dp)))

    def MultAdd(self, x, y, alpha=1.0, beta=0.0):
        """z = beta*y + alpha*K*x (TPZSYsmpMatrix::MultAdd, Matrix/pzsysmp.cpp:189-232)."""
        x = np.ascontiguousarray(x, np.float64)
        yy = None if y is None else np.ascontiguousarray(y, np.float64)
        z = np.zeros(self.neq)
        L.check(self.ctx.lib.pzg_multadd(self.ctx.h, L.ptr(x, L.dp), L.ptr(yy, L.dp),
                                         L.ptr(z, L.dp), alpha, beta))
        return z

    def Multiply(self, x):
        return self.MultAdd(x, None, 1.0, 0.0)

    def SolveCG(self, numiterations, F, tol, precond="jacobi"):
        """TPZMatrix::SolveCG (Matrix/pzmatrix.cpp:750-765) with EJacobi."""
        st = TPZStepSolver(self)
        st.SetCG(numiterations, TPZStepSolver.jacobi(self), tol, 0)
        return st.Solve(F)


class TPZFYsmpMatrixHIP(TPZSYsmpMatrixHIP):
    """Non-symmetric sparse rows (both triangles, sorted columns) of the same
    device system: TPZFYsmpMatrix (Matrix/pzysmp.h), as TPZSpStructMatrix
    creates it.  MultAdd and the solvers are those of the symmetric matrix."""

    def SetData(self, ia, ja, a, rhs=None):
        raise L.PzgError(L.PZG_EUNSUP, "TPZFYsmpMatrixHIP: caller matrices go through TPZSYsmpMatrixHIP")

    def AddKel(self, ek_batch, dest_batch):
        raise L.PzgError(L.PZG_EUNSUP, "TPZFYsmpMatrixHIP: AddKel goes through TPZSYsmpMatrixHIP")

    def _pattern(self):
        s = self.ctx.sizes()
        ia = np.zeros(self.neq + 1, np.int64)
        ja = np.zeros(max(s["nnz_full"], 1), np.int64)
        L.check(self.ctx.lib.pzg_full_csr(self.ctx.h, L.ptr(ia, L.i64p), L.ptr(ja, L.i64p), None))
        self._ia, self._ja = ia, ja[: s["nnz_full"]]

    def A(self):
        a = np.zeros(max(len(self.JA()), 1))
        L.check(self.ctx.lib.pzg_full_csr(self.ctx.h, None, None, L.ptr(a, L.dp)))
        return a[: len(self.JA())]

    def GetVal(self, r, c):
        """TPZFYsmpMatrix::GetVal: the row holds both triangles."""
        ia, ja = self.IA(), self.JA()
        row = ja[ia[r]:ia[r + 1]]
        k = np.searchsorted(row, c)
        if k < len(row) and row[k] == c:
            return float(self.A()[ia[r] + k])
        return 0.0


class TPZStepSolver:
    """TPZStepSolver subset used on the path."""

    ECG, EJacobi, EMixedFCG, EBlockJacobi = 0, 1, 2, 3

    def __init__(self, matrix=None):
        self.matrix = matrix
        self.kind = None
        self.maxit = 0
        self.tol = 0.0
        self.pre = None
        self.precond = L.PRECOND_BLOCK_JACOBI
        self.inner_tol = 1e-4
        self.refine = 0
        self.inner_op = L.INNER_AUTO
        self.max_outer = 200
        self.check_every = 50
        self.use_graph = True
        self.info = None
        self.fNumIterations = 0
        self.fTol = 0.0

    @staticmethod
    def jacobi(matrix=None):
        s = TPZStepSolver(matrix)
        s.SetJacobi(1, 0.0, 0)
        return s

    def SetMatrix(self, matrix):
        self.matrix = matrix

    def SetJacobi(self, numiterations, tol, FromCurrent):
        """EJacobi with one iteration is the only form the GPU path accepts."""
        if numiterations != 1 or FromCurrent != 0:
            raise ValueError("EJacobi preconditioner: only SetJacobi(1, tol, 0) is supported")
        self.kind = self.EJacobi

    def SetCG(self, numiterations, pre, tol, FromCurrent):
        """CG preconditioned by pre: an EJacobi step solver (point Jacobi) or
        the connect-block Jacobi that TPZLinearAnalysis.BuildPreconditioner
        (EBlockJacobi, overlap=True) returns."""
        if FromCurrent != 0:
            raise ValueError("SetCG: FromCurrent=1 is not supported (x starts at zero)")
        if pre is None or pre.kind not in (self.EJacobi, self.EBlockJacobi):
            raise ValueError("SetCG: the GPU path needs an EJacobi step solver or the "
                             "BuildPreconditioner(EBlockJacobi, True) preconditioner")
        self.kind = self.ECG
        self.maxit = int(numiterations)
        self.tol = float(tol)
        self.pre = pre
        self.precond = L.PRECOND_BLOCK_JACOBI if pre.kind == self.EBlockJacobi else L.PRECOND_JACOBI

    def SetMixedFCG(self, numiterations, tol, precond=L.PRECOND_BLOCK_JACOBI, inner_tol=1e-4,
                    max_outer=200, refine=0, inner_op=L.INNER_AUTO, mg_degree=0, mg_range=0.0):
        """FP32 flexible CG inside FP64 iterative refinement; stops on the true
        residual.  inner_tol is the reduction of the FP32 recursive residual
        between two true-residual evaluations; refine 0 restarts FCG from each
        true residual, 1 keeps the search direction (reliable update);
        inner_op selects the FP32 operator (INNER_BSR: stored block CSR,
        INNER_MATFREE: element-by-element, INNER_AUTO: matrix-free for
        mesh-built systems).  precond PRECOND_MULTIGRID preconditions each FCG
        step with a V-cycle (hexahedral lattice meshes; mg_degree and mg_range:
        Chebyshev degree and interval ratio of the smoothing, 0 = 2 and 30, 60
        with PRECOND_LEARNED)."""
        self.kind = self.EMixedFCG
        self.mg_degree = int(mg_degree)
        self.mg_range = float(mg_range)
        self.refine = refine
        self.inner_op = inner_op
        self.maxit = int(numiterations)
        self.tol = float(tol)
        self.precond = precond
        self.inner_tol = inner_tol
        self.max_outer = max_outer

    def params(self):
        p = L.SolveParams()
        p.solver = L.SOLVER_PCG64 if self.kind == self.ECG else L.SOLVER_MIXED
        p.precond = self.precond
        p.max_iter = self.maxit
        p.tol = self.tol
        p.inner_tol = self.inner_tol
        p.max_outer = self.max_outer
        p.check_every = self.check_every
        p.use_graph = 1 if self.use_graph else 0
        p.refine = self.refine
        p.inner_op = self.inner_op
        p.mg_degree = getattr(self, "mg_degree", 0)
        p.mg_range = getattr(self, "mg_range", 0.0)
        if self.precond == L.PRECOND_MULTIGRID and self.check_every == 50:
            p.check_every = 0  # the solver's short batches
        return p

    def Solve(self, F=None, want_solution=True):
        """Solves K u = F (F=None: the assembled rhs on the device).  Like
        TPZStepSolver::Solve (Solvers/pzstepsolver.cpp:107-112) it records the
        iteration count and overwrites the tolerance with the achieved one."""
        if self.kind not in (self.ECG, self.EMixedFCG):
            raise ValueError("TPZStepSolver::Solve: solver not set")
        ctx = self.matrix.ctx
        n = self.matrix.neq
        f = None if F is None else np.ascontiguousarray(F, np.float64)
        u = np.zeros(n) if want_solution else None
        info = L.SolveInfo()
        p = self.params()
        rc = ctx.lib.pzg_solve(ctx.h, C.byref(p), L.ptr(f, L.dp), L.ptr(u, L.dp), C.byref(info))
        L.check(rc, allow=(L.PZG_ENOCONV,))
        self.info = {k: getattr(info, k) for k, _ in L.SolveInfo._fields_}
        self.fNumIterations = info.iterations
        self.fTol = info.achieved
        return u


class TPZEquationFilter:
    """TPZEquationFilter (StrMatrix/TPZEquationFilter.h) as held by a struct
    matrix: SetActiveEquations / SetMinMaxEq / Reset, applied by the GPU
    struct matrix at Create()."""

    def __init__(self, neq=0):
        self.fNumEq = neq
        self.fActiveEqs = None

    def SetNumEq(self, neq):
        self.fNumEq = neq

    def SetActiveEquations(self, active):
        if self.fActiveEqs is not None:
            raise RuntimeError("TPZEquationFilter::SetActiveEquations: call Reset first")
        self.fActiveEqs = np.unique(np.asarray(active, dtype=np.int64))

    def SetMinMaxEq(self, mineq, maxeq):
        if not (0 <= mineq <= maxeq <= self.fNumEq):
            raise ValueError("TPZEquationFilter::SetMinMaxEq: bad range")
        self.SetActiveEquations(np.arange(mineq, maxeq))

    def Reset(self):
        self.fActiveEqs = None

    def IsActive(self):
        return self.fActiveEqs is not None

    def NActiveEquations(self):
        return len(self.fActiveEqs) if self.IsActive() else self.fNumEq

    def Scatter(self, small):
        if not self.IsActive():
            return np.array(small, dtype=np.float64)
        big = np.zeros(self.fNumEq)
        big[self.fActiveEqs] = small
        return big

    def Gather(self, big):
        return np.asarray(big)[self.fActiveEqs] if self.IsActive() else np.asarray(big)


class TPZSSpStructMatrixHIP:
    """GPU TPZStructMatrix: Create() -> pattern, Assemble() -> values + rhs."""

    def __init__(self, ctx, mesh_dict):
        self.ctx = ctx
        self.mesh = mesh_dict
        self._desc, self._keep = L.make_mesh_desc(mesh_dict)
        self._flattened = False
        self._filter = TPZEquationFilter(mesh_dict["neq"])

    def Clone(self):
        c = type(self)(self.ctx, self.mesh)
        if self._filter.IsActive():
            c._filter.SetActiveEquations(self._filter.fActiveEqs)
        return c

    def EquationFilter(self):
        """TPZStructMatrix::EquationFilter (StrMatrix/TPZStructMatrix.h)."""
        return self._filter

    def NReducedEquations(self):
        return self._filter.NActiveEquations()

    def _flatten(self):
        if not self._flattened:
            L.check(self.ctx.lib.pzg_mesh_set(self.ctx.h, C.byref(self._desc)))
            self._flattened = True

    def _apply_filter(self):
        f = self._filter
        if f.IsActive():
            act = np.ascontiguousarray(f.fActiveEqs, dtype=np.int64)
            L.check(self.ctx.lib.pzg_set_equation_filter(self.ctx.h, len(act), L.ptr(act, L.i64p)))
        else:
            L.check(self.ctx.lib.pzg_set_equation_filter(self.ctx.h, -1, None))

    def Create(self):
        """TPZSSpStructMatrix::Create (StrMatrix/TPZSSpStructMatrix.cpp:30-47)."""
        self._flatten()
        self._apply_filter()
        neq = C.c_int64()
        nnz = C.c_int64()
        L.check(self.ctx.lib.pzg_struct_create(self.ctx.h, C.byref(neq), C.byref(nnz)))
        return TPZSYsmpMatrixHIP(self.ctx, neq.value)

    def Assemble(self, K=None, rhs=None):
        """TPZStructMatrixOR::Assemble(K, rhs) on the GPU; values stay on the device."""
        L.check(self.ctx.lib.pzg_assemble(self.ctx.h))

    def CreateAssemble(self):
        K = self.Create()
        self.Assemble(K)
        return K


class TPZSpStructMatrixHIP(TPZSSpStructMatrixHIP):
    """GPU TPZSpStructMatrix (StrMatrix/TPZSpStructMatrix.cpp:33-190): the same
    assembly, the matrix handed out in the non-symmetric TPZFYsmpMatrix format."""

    def Create(self):
        K = super().Create()
        return TPZFYsmpMatrixHIP(self.ctx, K.neq)


class TLaplaceExample1:
    """Scalar exact solutions evaluated on the device: the EConst, EX, ESinSin
    and ECosCos cases of TLaplaceExample1::uxy (Pre/TPZAnalyticSolution.cpp:
    1254-1310) with fCenter = center (gC = constant for EConst)."""

    EConst, EX, ESinSin, ECosCos = L.EXACT_CONST, L.EXACT_X, L.EXACT_SINSIN, L.EXACT_COSCOS

    def __init__(self, kind, center=(0.0, 0.0, 0.0), constant=1.0):
        if kind not in (self.EConst, self.EX, self.ESinSin, self.ECosCos):
            raise ValueError("TLaplaceExample1: EConst, EX, ESinSin or ECosCos")
        self.fExact, self.fCenter, self.gC = kind, tuple(center), float(constant)

    def ExactSolution(self):
        return self


def element_graph(mesh):
    """TPZCompMesh::ComputeElGraph (Mesh/pzcmesh.cpp:1223-1266) of a flattened
    mesh: (elgraphindex, elgraph) over the volume then the boundary elements,
    each element's connect sequence numbers in side order."""
    vc = np.asarray(mesh["vol_conn"], dtype=np.int64)
    vc = vc.reshape(len(mesh["vol_nodes"]), -1)
    bc = np.asarray(mesh["bc_conn"], dtype=np.int64)
    nbc = len(mesh["bc_nodes"])
    bc = bc.reshape(nbc, -1) if nbc else np.zeros((0, 1), dtype=np.int64)
    eptr = np.concatenate([[0], np.cumsum([vc.shape[1]] * vc.shape[0] + [bc.shape[1]] * nbc)])
    return eptr.astype(np.int64), np.concatenate([vc.ravel(), bc.ravel()]).astype(np.int64)


def OptimizeBandwidth(mesh):
    """TPZAnalysis::OptimizeBandwidth (Analysis/TPZAnalysis.cpp:225-300) with
    the default TPZSloanRenumbering, computed by the library on the host:
    returns the renumbered mesh description (TPZCompMesh::Permute) and the
    permutation perm[old sequence number] = new one."""
    if mesh.get("dep_ptr") is not None:
        raise ValueError("OptimizeBandwidth: meshes with constrained connects are not renumbered here")
    eptr, elist = element_graph(mesh)
    n = int(mesh["nconn"])
    perm = np.zeros(n, dtype=np.int64)
    L.check(L.load().pzg_sloan_renumber(n, len(eptr) - 1, L.ptr(eptr, L.i64p), L.ptr(elist, L.i64p),
                                        L.ptr(perm, L.i64p)))
    return M.permute(mesh, perm), perm


class TPZLinearAnalysis:
    """TPZLinearAnalysis(cmesh, mustOptimizeBandwidth) over the GPU path; with
    mustOptimizeBandwidth the connects are renumbered by TPZSloanRenumbering
    first (TPZAnalysis::SetCompMeshInit, Analysis/TPZAnalysis.cpp:150-194)."""

    def __init__(self, problem_or_mesh, mustOptimizeBandwidth=False, ctx=None):
        self.ctx = ctx or Context(0)
        self.mesh = problem_or_mesh if isinstance(problem_or_mesh, dict) else M.build(problem_or_mesh)
        self.fPermutation = None
        if mustOptimizeBandwidth:
            self.mesh, self.fPermutation = OptimizeBandwidth(self.mesh)
        self.fStructMatrix = None
        self.fSolver = None
        self.K = None
        self.fSolution = None

    EJacobi, EBlockJacobi, EElement, ENodeCentered = 0, 1, 2, 3  # TPZAnalysis::EPrecond

    def BuildPreconditioner(self, preconditioner, overlap):
        """TPZAnalysis::BuildPreconditioner (Analysis/TPZAnalysis.cpp:995-1086).
        EJacobi returns None as there.  EBlockJacobi with overlap returns the
        block diagonal of the connects (TPZBlockDiagonalStructMatrix::
        AssembleBlockDiagonal, StrMatrix/pzbdstrmatrix.cpp:13-56: one block per
        connect, its nshape x nstate equations, the stiffness restricted to
        them) solved by LU; the solver inverts the blocks of the assembled
        matrix on the device when CG starts.  The other forms (overlapping
        element and vertex blocks, the coloured sequence solver of
        overlap=False) are not built."""
        if preconditioner == self.EJacobi:
            return None
        if preconditioner != self.EBlockJacobi or not overlap:
            raise NotImplementedError("BuildPreconditioner: only EJacobi and EBlockJacobi with overlap")
        if self.fSolver is None:
            raise ValueError("BuildPreconditioner called with uninitialized stiffness matrix")
        step = TPZStepSolver(self.fSolver.matrix)
        step.kind = TPZStepSolver.EBlockJacobi
        return step

    def SetStructuralMatrix(self, strmat):
        self.fStructMatrix = strmat.Clone()  # TPZAnalysis clones (Analysis/TPZAnalysis.cpp:83-104)

    def SetSolver(self, solver):
        self.fSolver = solver

    def Assemble(self):
        if self.fStructMatrix is None:
            self.fStructMatrix = TPZSSpStructMatrixHIP(self.ctx, self.mesh)
        if self.K is None:
            self.K = self.fStructMatrix.CreateAssemble()
        else:
            self.fStructMatrix.Assemble(self.K)
        if self.fSolver is not None:
            self.fSolver.SetMatrix(self.K)

    def Rhs(self):
        """fRhs: the load in the mesh numbering (with an equation filter the
        condensed load scattered back, StrMatrix/pzstrmatrixor.cpp:63)."""
        r = np.zeros(self.K.neq)
        L.check(self.ctx.lib.pzg_get_system(self.ctx.h, None, L.ptr(r, L.dp)))
        return self.fStructMatrix.EquationFilter().Scatter(r)

    def Solve(self, want_solution=True):
        """TPZLinearAnalysis::SolveT (Analysis/TPZLinearAnalysis.cpp:130-227):
        with an equation filter the condensed solution is scattered into the
        mesh numbering (inactive equations 0) before LoadSolution."""
        if self.fSolver is None:
            raise ValueError("TPZLinearAnalysis::Solve: no solver")
        self.fSolver.SetMatrix(self.K)
        u = self.fSolver.Solve(None, want_solution=want_solution)
        self.fSolution = None if u is None else self.fStructMatrix.EquationFilter().Scatter(u)
        return self.fSolution

    def Solution(self):
        return self.fSolution

    # --- post-processing ------------------------------------------------------
    def _mesh_on_device(self):
        # the mesh is flattened to the device by the struct matrix (Assemble);
        # post-processing a loaded solution needs it without an assembly
        if self.fStructMatrix is None:
            self.fStructMatrix = TPZSSpStructMatrixHIP(self.ctx, self.mesh)
        self.fStructMatrix._flatten()

    def LoadSolution(self, sol=None):
        """TPZLinearAnalysis::LoadSolution -> TPZCompMesh::LoadSolution
        (Mesh/pzcmesh.cpp:524-540): the solution the error norms use.  Solve
        loads its solution already; sol replaces it."""
        if sol is not None:
            self._mesh_on_device()
            sol = np.ascontiguousarray(sol, dtype=np.float64)
            if sol.shape != (self.mesh["neq"],):
                raise ValueError("LoadSolution: solution of the wrong length")
            L.check(self.ctx.lib.pzg_load_solution(self.ctx.h, L.ptr(sol, L.dp)))
            self.fSolution = sol

    def SetExact(self, f, p):
        """TPZAnalysis::SetExact (Analysis/TPZAnalysis.cpp:655-676).  f is a
        TLaplaceExample1 (evaluated on the device) or a vectorised exact
        solution f(x[npts, 3]) -> (u[npts, nstate], du[npts, dim, nstate]),
        du[q, i, j] = d u_j / d x_i, evaluated on the host at the error rule's
        points."""
        self.fExact, self.fExactOrder = f, int(p)

    def ErrorPoints(self):
        """Physical coordinates [npts, 3] of the error rule's points."""
        self._mesh_on_device()
        n = C.c_int64(0)
        L.check(self.ctx.lib.pzg_error_points(self.ctx.h, self.fExactOrder, C.byref(n), None))
        xq = np.zeros((n.value, 3))
        L.check(self.ctx.lib.pzg_error_points(self.ctx.h, self.fExactOrder, C.byref(n),
                                              L.ptr(xq, L.dp)))
        return xq

    def PostProcessError(self, store_error=False):
        """TPZAnalysis::PostProcessError (Analysis/TPZAnalysis.cpp:480-487,
        625-648): the mesh totals of the material's error norms; with
        store_error the per-element values go to ElementSolution()."""
        f = getattr(self, "fExact", None)
        if f is None:
            raise ValueError("PostProcessError: no exact solution (SetExact)")
        self._mesh_on_device()
        ex = L.Exact()
        ex.porder = self.fExactOrder
        keep = []
        if isinstance(f, TLaplaceExample1):
            ex.kind = f.fExact
            for d in range(3):
                ex.center[d] = f.fCenter[d]
            ex.constant = f.gC
        else:
            xq = self.ErrorPoints()
            u, du = f(xq)
            ns, dim = self.mesh["nstate"], self.mesh["dim"]
            u = np.ascontiguousarray(np.asarray(u, dtype=np.float64).reshape(len(xq), ns))
            du = np.ascontiguousarray(np.asarray(du, dtype=np.float64).reshape(len(xq), dim, ns))
            keep += [u, du]
            ex.kind, ex.npts = L.EXACT_POINTS, len(xq)
            ex.u, ex.du = L.ptr(u, L.dp), L.ptr(du, L.dp)
        nvol = len(self.mesh["vol_nodes"])
        el = np.zeros((nvol, 6))
        tot = np.zeros(6)
        nerr = C.c_int32(0)
        L.check(self.ctx.lib.pzg_evaluate_error(self.ctx.h, C.byref(ex), L.ptr(el, L.dp),
                                                L.ptr(tot, L.dp), C.byref(nerr)))
        k = nerr.value
        if store_error:
            self.fElementSolution = el.reshape(-1)[: nvol * k].reshape(nvol, k).copy()
        return tot[:k].copy()

    def ElementSolution(self):
        """TPZCompMesh::ElementSolution: per volume element the error norms of
        the last PostProcessError(store_error=True)."""
        return self.fElementSolution
