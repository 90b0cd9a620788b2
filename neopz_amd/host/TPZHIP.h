// TPZHIP.h -- C++17 host-side mirror of NeoPZ's interface for the linear FEM
// path, over the C ABI of include/pzg.h.
//
// The classes keep the reference's names, argument meaning and error
// behaviour (a failing call prints to stderr and stops, as PZError +
// DebugStop() do), so that code written against
//   TPZSSpStructMatrix (StrMatrix/TPZSSpStructMatrix.h:17-38),
//   TPZSpStructMatrix  (StrMatrix/TPZSpStructMatrix.h),
//   TPZSYsmpMatrix     (Matrix/pzsysmp.h:20-230),
//   TPZEquationFilter  (StrMatrix/TPZEquationFilter.h),
//   TPZFYsmpMatrix     (Matrix/pzysmp.h),
//   TPZStepSolver      (Solvers/pzstepsolver.h:34-89),
//   TPZLinearAnalysis  (Analysis/TPZLinearAnalysis.h:20-45)
// reads the same.  NeoPZ's own containers are replaced by std::vector; the
// mesh arrives flattened (pzg_mesh_desc) instead of as a TPZCompMesh.  The
// subclasses that plug into NeoPZ itself are given in INTEGRATION.md.
#pragma once

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <memory>
#include <vector>

#include "../../include/pzg.h"

namespace pzhip {

inline void DebugStop() { std::abort(); }

inline void Check(int rc, const char *where) {
  if (rc != PZG_OK) {
    std::fprintf(stderr, "%s: pzg error %d: %s\n", where, rc, pzg_last_error());
    DebugStop();
  }
}

using Ctx = std::shared_ptr<pzg_ctx>;

inline Ctx MakeContext(int device = 0) {
  pzg_ctx *c = nullptr;
  Check(pzg_create(device, &c), "pzg_create");
  return Ctx(c, pzg_destroy);
}

// Solver that TPZSYsmpMatrixHIP::SolveCG runs for a TPZStepSolver set up with
// SetCG(n, EJacobi step, tol, 0): the reference's FP64 Jacobi-PCG (default),
// or, when the struct matrix was asked for it (SetMixedPrecision), the FP32
// flexible CG inside FP64 refinement with the given preconditioner.  Either
// way the iteration count and the achieved tolerance are written back.
struct SolveMode {
  int solver = PZG_SOLVER_PCG64;
  int precond = PZG_PRECOND_JACOBI;
  double inner_tol = 1e-4;
};

// TPZSYsmpMatrix: symmetric CSR, upper triangle with the diagonal first.
class TPZSYsmpMatrixHIP {
 public:
  TPZSYsmpMatrixHIP(Ctx ctx, int64_t neq, SolveMode mode = SolveMode())
      : fCtx(std::move(ctx)), fNeq(neq), fMode(mode) {}
  virtual ~TPZSYsmpMatrixHIP() = default;
  int64_t Rows() const { return fNeq; }
  // TPZSYsmpMatrix::SetData (Matrix/pzsysmp.h:232-240)
  void SetData(std::vector<int64_t> IA, std::vector<int64_t> JA, std::vector<double> A) {
    fNeq = (int64_t)IA.size() - 1;
    Check(pzg_csr_set(fCtx.get(), fNeq, IA.data(), JA.data(), A.data(), nullptr), "SetData");
    fIA = std::move(IA);
    fJA = std::move(JA);
  }
  const std::vector<int64_t> &IA() {
    Pattern();
    return fIA;
  }
  const std::vector<int64_t> &JA() {
    Pattern();
    return fJA;
  }
  virtual std::vector<double> A() {
    std::vector<double> a(JA().size());
    Check(pzg_get_system(fCtx.get(), a.data(), nullptr), "A");
    return a;
  }
  // TPZSYsmpMatrix::GetVal (Matrix/pzsysmp.cpp:59-77): lower entries read the transpose
  virtual double GetVal(int64_t r, int64_t c) {
    if (r > c) std::swap(r, c);
    const auto &ia = IA();
    const auto &ja = JA();
    for (int64_t k = ia[r]; k < ia[r + 1]; k++)
      if (ja[k] == c) return A()[k];
    return 0.;
  }
  // batched TPZSYsmpMatrix::AddKel (Matrix/pzsysmp.cpp:369-411)
  void AddKel(int64_t nel, int32_t ne, const std::vector<int64_t> &dest,
              const std::vector<double> &ek) {
    Check(pzg_add_kel_batch(fCtx.get(), nel, ne, dest.data(), ek.data()), "AddKel");
  }
  // z = beta*y + alpha*K*x (TPZSYsmpMatrix::MultAdd, Matrix/pzsysmp.cpp:189-232)
  void MultAdd(const std::vector<double> &x, const std::vector<double> &y, std::vector<double> &z,
               double alpha, double beta) const {
    z.resize(fNeq);
    Check(pzg_multadd(fCtx.get(), x.data(), beta != 0. ? y.data() : nullptr, z.data(), alpha, beta),
          "MultAdd");
  }
  pzg_ctx *Context() const { return fCtx.get(); }
  const SolveMode &Mode() const { return fMode; }
  // TPZMatrix::SolveCG (Matrix/pzmatrix.cpp:750-765, Matrix/pzmatrix.h:460-463)
  // as TPZStepSolver::Solve calls it for ECG with an EJacobi preconditioner
  // (Solvers/pzstepsolver.cpp:107-112): F and result in the matrix numbering;
  // numiterations and tol receive the iterations and the achieved residual.
  // pcg_precond: the preconditioner step solver's kind for the plain CG
  // (PZG_PRECOND_JACOBI for EJacobi, PZG_PRECOND_BLOCK_JACOBI for the
  // EBlockJacobi block diagonal of BuildPreconditioner)
  int SolveCG(int64_t &numiterations, const std::vector<double> *F, std::vector<double> &result,
              double &tol, pzg_solve_info &info, int pcg_precond = PZG_PRECOND_JACOBI) {
    pzg_solve_params p{};
    p.solver = fMode.solver;
    p.precond = fMode.solver == PZG_SOLVER_PCG64 ? pcg_precond : fMode.precond;
    p.max_iter = numiterations;
    p.tol = tol;
    p.inner_tol = fMode.inner_tol;
    p.max_outer = 200;
    p.check_every = p.precond == PZG_PRECOND_MULTIGRID ? 0 : 50;  // 0: the solver's batch
    p.use_graph = 1;
    p.refine = 0;
    p.inner_op = PZG_INNER_AUTO;
    result.resize(fNeq);
    int rc = pzg_solve(fCtx.get(), &p, F ? F->data() : nullptr, result.data(), &info);
    if (rc != PZG_OK && rc != PZG_ENOCONV) Check(rc, "SolveCG");
    numiterations = info.iterations;
    tol = info.achieved;
    return rc;
  }

 protected:
  virtual void Pattern() {
    if (!fIA.empty()) return;
    pzg_sizes s;
    Check(pzg_get_sizes(fCtx.get(), &s), "sizes");
    fIA.resize(fNeq + 1);
    fJA.resize(s.nnz_upper);
    Check(pzg_struct_get_pattern(fCtx.get(), fIA.data(), fJA.data()), "IA/JA");
    fJA.resize(fIA[fNeq]);  // condensed (equation filter) patterns are shorter
  }
  Ctx fCtx;
  int64_t fNeq;
  std::vector<int64_t> fIA, fJA;
  SolveMode fMode;
};

// TPZFYsmpMatrix: non-symmetric sparse rows (both triangles, sorted columns)
// over the same device system; MultAdd and the solvers are the symmetric ones.
class TPZFYsmpMatrixHIP : public TPZSYsmpMatrixHIP {
 public:
  using TPZSYsmpMatrixHIP::TPZSYsmpMatrixHIP;
  std::vector<double> A() override {
    std::vector<double> a(JA().size());
    Check(pzg_full_csr(fCtx.get(), nullptr, nullptr, a.data()), "A");
    return a;
  }
  // TPZFYsmpMatrix::GetVal (Matrix/pzysmp.cpp): the row holds both triangles
  double GetVal(int64_t r, int64_t c) override {
    const auto &ia = IA();
    const auto &ja = JA();
    for (int64_t k = ia[r]; k < ia[r + 1]; k++)
      if (ja[k] == c) return A()[k];
    return 0.;
  }

 protected:
  void Pattern() override {
    if (!fIA.empty()) return;
    pzg_sizes s;
    Check(pzg_get_sizes(fCtx.get(), &s), "sizes");
    fIA.resize(fNeq + 1);
    fJA.resize(s.nnz_full);
    Check(pzg_full_csr(fCtx.get(), fIA.data(), fJA.data(), nullptr), "IA/JA");
  }
};

// TPZStepSolver: ECG with an EJacobi(1) preconditioner or with the
// EBlockJacobi block diagonal of TPZLinearAnalysis::BuildPreconditioner
// (EDirectBlocks here: a TPZBlockDiagonal solved by ELU), or the mixed solver.
class TPZStepSolver {
 public:
  enum MSolver { ENoSolver, EJacobi, ECG, EMixedFCG, EDirectBlocks };
  void SetJacobi(int64_t numiterations, double /*tol*/, int FromCurrent) {
    if (numiterations != 1 || FromCurrent) DebugStop();
    fSolver = EJacobi;
  }
  void SetCG(int64_t numiterations, const TPZStepSolver &pre, double tol, int FromCurrent) {
    if ((pre.fSolver != EJacobi && pre.fSolver != EDirectBlocks) || FromCurrent) DebugStop();
    fSolver = ECG;
    fMaxIterations = numiterations;
    fTol = tol;
    fPcgPrecond = pre.fSolver == EDirectBlocks ? PZG_PRECOND_BLOCK_JACOBI : PZG_PRECOND_JACOBI;
  }
  // the step solver BuildPreconditioner(EBlockJacobi, true) returns
  static TPZStepSolver ConnectBlocks() {
    TPZStepSolver s;
    s.fSolver = EDirectBlocks;
    return s;
  }
  void SetMixedFCG(int64_t numiterations, double tol, int precond = PZG_PRECOND_BLOCK_JACOBI) {
    fSolver = EMixedFCG;
    fMaxIterations = numiterations;
    fTol = tol;
    fPrecond = precond;
  }
  void SetMatrix(TPZSYsmpMatrixHIP *K) { fMatrix = K; }
  MSolver Solver() const { return fSolver; }
  // TPZStepSolver::Solve (Solvers/pzstepsolver.cpp:65-122): records the
  // iteration count and overwrites the tolerance with the achieved one.
  void Solve(const std::vector<double> *F, std::vector<double> &result) {
    if (!fMatrix || (fSolver != ECG && fSolver != EMixedFCG)) DebugStop();
    if (fSolver == ECG) {  // mat->SolveCG(...), the matrix picks the algorithm
      int64_t its = fMaxIterations;
      double tol = fTol;
      fMatrix->SolveCG(its, F, result, tol, fInfo, fPcgPrecond);
      fNumIterations = its;
      fTol = tol;
      return;
    }
    pzg_solve_params p{};
    p.solver = fSolver == ECG ? PZG_SOLVER_PCG64 : PZG_SOLVER_MIXED;
    p.precond = fPrecond;
    p.max_iter = fMaxIterations;
    p.tol = fTol;
    p.inner_tol = 1e-4;
    p.max_outer = 200;
    p.check_every = fPrecond == PZG_PRECOND_MULTIGRID ? 0 : 50;  // 0: the solver's batch
    p.use_graph = 1;
    p.refine = 0;
    p.inner_op = PZG_INNER_AUTO;
    result.resize(fMatrix->Rows());
    int rc = pzg_solve(fMatrix->Context(), &p, F ? F->data() : nullptr, result.data(), &fInfo);
    if (rc != PZG_OK && rc != PZG_ENOCONV) Check(rc, "Solve");
    fNumIterations = fInfo.iterations;
    fTol = fInfo.achieved;
  }
  int64_t NumIterations() const { return fNumIterations; }
  double GetTolerance() const { return fTol; }
  const pzg_solve_info &Info() const { return fInfo; }

 private:
  MSolver fSolver = ENoSolver;
  int64_t fMaxIterations = 0, fNumIterations = 0;
  double fTol = 0.;
  int fPrecond = PZG_PRECOND_BLOCK_JACOBI;
  int fPcgPrecond = PZG_PRECOND_JACOBI;
  TPZSYsmpMatrixHIP *fMatrix = nullptr;
  pzg_solve_info fInfo{};
};

// TPZEquationFilter (StrMatrix/TPZEquationFilter.h): the active equations of
// the mesh numbering, sorted without duplicates; Scatter / Gather between the
// condensed and the mesh numbering.
class TPZEquationFilter {
 public:
  explicit TPZEquationFilter(int64_t numeq = 0) : fNumEq(numeq) {}
  void SetNumEq(int64_t n) { fNumEq = n; }
  void SetActiveEquations(std::vector<int64_t> active) {
    if (fIsActive) DebugStop();  // call Reset first
    std::sort(active.begin(), active.end());
    active.erase(std::unique(active.begin(), active.end()), active.end());
    if (!active.empty() && (active.front() < 0 || active.back() >= fNumEq)) DebugStop();
    fActiveEqs = std::move(active);
    fIsActive = true;
  }
  void SetMinMaxEq(int64_t mineq, int64_t maxeq) {
    if (mineq < 0 || maxeq > fNumEq || mineq > maxeq) DebugStop();
    std::vector<int64_t> a(maxeq - mineq);
    for (int64_t i = 0; i < maxeq - mineq; i++) a[i] = mineq + i;
    SetActiveEquations(std::move(a));
  }
  void Reset() {
    fIsActive = false;
    fActiveEqs.clear();
  }
  bool IsActive() const { return fIsActive; }
  int64_t NActiveEquations() const { return fIsActive ? (int64_t)fActiveEqs.size() : fNumEq; }
  int64_t NEqExpand() const { return fNumEq; }
  const std::vector<int64_t> &ActiveEquations() const { return fActiveEqs; }
  void Scatter(const std::vector<double> &small, std::vector<double> &big) const {
    if (!fIsActive) { big = small; return; }
    big.assign(fNumEq, 0.);
    for (size_t i = 0; i < fActiveEqs.size(); i++) big[fActiveEqs[i]] = small[i];
  }
  void Gather(const std::vector<double> &big, std::vector<double> &small) const {
    if (!fIsActive) { small = big; return; }
    small.resize(fActiveEqs.size());
    for (size_t i = 0; i < fActiveEqs.size(); i++) small[i] = big[fActiveEqs[i]];
  }

 private:
  int64_t fNumEq = 0;
  bool fIsActive = false;
  std::vector<int64_t> fActiveEqs;
};

// TPZSSpStructMatrix over the GPU: Create() -> pattern, Assemble() -> values + rhs.
class TPZSSpStructMatrixHIP {
 public:
  TPZSSpStructMatrixHIP(Ctx ctx, const pzg_mesh_desc &mesh)
      : fCtx(std::move(ctx)), fMesh(mesh), fFilter(mesh.neq) {}
  virtual ~TPZSSpStructMatrixHIP() = default;
  // TPZStructMatrix::EquationFilter (StrMatrix/TPZStructMatrix.h)
  TPZEquationFilter &EquationFilter() { return fFilter; }
  int64_t NReducedEquations() const { return fFilter.NActiveEquations(); }
  // the matrices this struct matrix creates solve SetCG step solvers with the
  // mixed-precision FCG (FP32 iterations with the given preconditioner, FP64
  // refinement to the FP64 true residual) instead of FP64 Jacobi-PCG
  void SetMixedPrecision(int precond, double inner_tol = 1e-4) {
    fMode.solver = PZG_SOLVER_MIXED;
    fMode.precond = precond;
    fMode.inner_tol = inner_tol;
  }
  // TPZSSpStructMatrix::Create (StrMatrix/TPZSSpStructMatrix.cpp:30-47)
  virtual std::unique_ptr<TPZSYsmpMatrixHIP> Create() {
    return std::make_unique<TPZSYsmpMatrixHIP>(fCtx, Setup(), fMode);
  }
  // the analysis' mesh (after a renumbering) replaces the one given here
  void SetMesh(const pzg_mesh_desc &mesh) {
    fMesh = mesh;
    fFlattened = false;
  }
  // TPZStructMatrixOR::Assemble(K, rhs) (StrMatrix/pzstrmatrixor.cpp:41-71):
  // rhs in the mesh numbering (with a filter the condensed load scattered,
  // inactive equations 0); the values stay on the device
  void Assemble(TPZSYsmpMatrixHIP & /*K*/, std::vector<double> &rhs) {
    Check(pzg_assemble(fCtx.get()), "Assemble");
    std::vector<double> small(fFilter.NActiveEquations());
    Check(pzg_get_system(fCtx.get(), nullptr, small.data()), "rhs");
    fFilter.Scatter(small, rhs);
  }

 protected:
  int64_t Setup() {
    if (!fFlattened) {
      Check(pzg_mesh_set(fCtx.get(), &fMesh), "mesh");
      fFlattened = true;
    }
    const auto &act = fFilter.ActiveEquations();
    Check(pzg_set_equation_filter(fCtx.get(), fFilter.IsActive() ? (int64_t)act.size() : -1,
                                  fFilter.IsActive() ? act.data() : nullptr),
          "EquationFilter");
    int64_t neq = 0, nnz = 0;
    Check(pzg_struct_create(fCtx.get(), &neq, &nnz), "Create");
    return neq;
  }
  Ctx fCtx;
  pzg_mesh_desc fMesh;
  bool fFlattened = false;
  TPZEquationFilter fFilter;
  SolveMode fMode;
};

// TPZSpStructMatrix over the GPU: the same assembly, the matrix in the
// non-symmetric format (StrMatrix/TPZSpStructMatrix.cpp:33-190).
class TPZSpStructMatrixHIP : public TPZSSpStructMatrixHIP {
 public:
  using TPZSSpStructMatrixHIP::TPZSSpStructMatrixHIP;
  std::unique_ptr<TPZSYsmpMatrixHIP> Create() override {
    return std::make_unique<TPZFYsmpMatrixHIP>(fCtx, Setup(), fMode);
  }
};

// TPZLinearAnalysis(cmesh, mustOptimizeBandwidth=false): Assemble()/Solve().
// The exact solution of TPZAnalysis::SetExact: the reference's
// std::function(loc, result, deriv) with result[nstate] and deriv the
// dim x nstate matrix stored row-major (deriv[i * nstate + j] = d u_j / d x_i).
using ExactSolType = std::function<void(const std::vector<double> &loc, std::vector<double> &result,
                                        std::vector<double> &deriv)>;

class TPZLinearAnalysis {
 public:
  // TPZLinearAnalysis(cmesh, mustOptimizeBandwidth) -> TPZAnalysis::
  // SetCompMeshInit (Analysis/TPZAnalysis.cpp:150-194): with
  // mustOptimizeBandwidth the connects are renumbered (OptimizeBandwidth)
  TPZLinearAnalysis(Ctx ctx, const pzg_mesh_desc &mesh, bool mustOptimizeBandwidth = false)
      : fCtx(std::move(ctx)), fMesh(mesh), fDim(mesh.dim), fNState(mesh.nstate), fNVol(mesh.nvol) {
    if (mustOptimizeBandwidth) OptimizeBandwidth();
    fStr = std::make_unique<TPZSSpStructMatrixHIP>(fCtx, fMesh);
  }
  // TPZAnalysis::OptimizeBandwidth (Analysis/TPZAnalysis.cpp:225-300) with the
  // default TPZSloanRenumbering; the analysis keeps the renumbered connect
  // lists and blocks (TPZCompMesh::Permute)
  void OptimizeBandwidth() {
    if (fMesh.dep_ptr) DebugStop();  // constrained connects: not renumbered here
    const bool tet = fMesh.vol_etype == PZG_TET;
    const int nsv = fMesh.dim == 2 ? 9 : tet ? 15 : 27, nsb = fMesh.dim == 2 ? 3 : tet ? 7 : 9;
    std::vector<int64_t> eptr(1, 0), elist;
    for (int64_t e = 0; e < fMesh.nvol; e++) {
      elist.insert(elist.end(), fMesh.vol_conn + e * nsv, fMesh.vol_conn + (e + 1) * nsv);
      eptr.push_back(int64_t(elist.size()));
    }
    for (int64_t e = 0; e < fMesh.nbc; e++) {
      elist.insert(elist.end(), fMesh.bc_conn + e * nsb, fMesh.bc_conn + (e + 1) * nsb);
      eptr.push_back(int64_t(elist.size()));
    }
    const int64_t n = fMesh.nconn;
    fPerm.assign(n, 0);
    Check(pzg_sloan_renumber(n, int64_t(eptr.size()) - 1, eptr.data(), elist.data(), fPerm.data()),
          "OptimizeBandwidth");
    fVolConn.resize(fMesh.nvol * nsv);
    for (size_t k = 0; k < fVolConn.size(); k++) fVolConn[k] = fPerm[fMesh.vol_conn[k]];
    fBcConn.resize(fMesh.nbc * nsb);
    for (size_t k = 0; k < fBcConn.size(); k++) fBcConn[k] = fPerm[fMesh.bc_conn[k]];
    fBlkSize.assign(n, 0);
    for (int64_t c = 0; c < n; c++) fBlkSize[fPerm[c]] = fMesh.blk_size[c];
    fBlkPos.assign(n, 0);
    for (int64_t c = 1; c < n; c++) fBlkPos[c] = fBlkPos[c - 1] + fBlkSize[c - 1];
    fMesh.vol_conn = fVolConn.data();
    fMesh.bc_conn = fBcConn.data();
    fMesh.blk_size = fBlkSize.data();
    fMesh.blk_pos = fBlkPos.data();
    if (fStr) fStr->SetMesh(fMesh);
  }
  // perm[old sequence number] = new one (empty without OptimizeBandwidth)
  const std::vector<int64_t> &Permutation() const { return fPerm; }
  // TPZAnalysis::SetStructuralMatrix (Analysis/TPZAnalysis.cpp:83-104): the
  // struct matrix works on the analysis' mesh
  void SetStructuralMatrix(std::unique_ptr<TPZSSpStructMatrixHIP> str) {
    fStr = std::move(str);
    fStr->SetMesh(fMesh);
    fK.reset();
  }
  void SetSolver(const TPZStepSolver &s) { fSolver = s; }
  // TPZAnalysis::BuildPreconditioner (Analysis/TPZAnalysis.cpp:995-1086):
  // EBlockJacobi with overlap gives the block diagonal of the connects, the
  // blocks of the assembled matrix solved directly (inverted on the device
  // when CG starts); EJacobi has no preconditioner object there (an
  // EJacobi step solver stands for it here); the other forms are not built.
  enum EPrecond { EJacobi, EBlockJacobi, EElement, ENodeCentered };
  TPZStepSolver BuildPreconditioner(EPrecond pre, bool overlap) const {
    TPZStepSolver s;
    if (pre == EJacobi) {
      s.SetJacobi(1, 0., 0);
      return s;
    }
    if (pre != EBlockJacobi || !overlap) DebugStop();
    return TPZStepSolver::ConnectBlocks();
  }
  // Analysis/TPZLinearAnalysis.cpp:42-94
  void Assemble() {
    if (!fK) fK = fStr->Create();
    fStr->Assemble(*fK, fRhs);
    fSolver.SetMatrix(fK.get());
  }
  // Analysis/TPZLinearAnalysis.cpp:130-227: with an equation filter the rhs
  // is gathered to the condensed numbering and the solution scattered back
  void Solve() {
    const TPZEquationFilter &f = fStr->EquationFilter();
    if (!f.IsActive()) {
      fSolver.Solve(&fRhs, fSolution);
      return;
    }
    std::vector<double> r, delu;
    f.Gather(fRhs, r);
    fSolver.Solve(&r, delu);
    f.Scatter(delu, fSolution);
  }
  TPZSSpStructMatrixHIP &StructMatrix() { return *fStr; }
  const std::vector<double> &Solution() const { return fSolution; }
  const std::vector<double> &Rhs() const { return fRhs; }
  TPZSYsmpMatrixHIP &Matrix() { return *fK; }
  TPZStepSolver &Solver() { return fSolver; }

  // TPZLinearAnalysis::LoadSolution / TPZCompMesh::LoadSolution
  // (Mesh/pzcmesh.cpp:524-540); Solve loads its own solution already.
  void LoadSolution(const std::vector<double> &sol) {
    fSolution = sol;
    Check(pzg_load_solution(fCtx.get(), fSolution.data()), "LoadSolution");
  }
  // TPZAnalysis::SetExact (Analysis/TPZAnalysis.cpp:655-676): a host
  // function, evaluated at the error rule's points ...
  void SetExact(ExactSolType f, int p) {
    fExact = std::move(f);
    fExactDev = pzg_exact{};
    fExactDev.kind = PZG_EXACT_POINTS;
    fExactDev.porder = p;
  }
  // ... or a TLaplaceExample1 case evaluated on the device (PZG_EXACT_CONST,
  // _X, _SINSIN, _COSCOS; Pre/TPZAnalyticSolution.cpp:1254-1310)
  void SetExactLaplace(int kind, int p, const double center[3] = nullptr, double constant = 1.) {
    fExact = nullptr;
    fExactDev = pzg_exact{};
    fExactDev.kind = kind;
    fExactDev.porder = p;
    fExactDev.constant = constant;
    for (int d = 0; d < 3; d++) fExactDev.center[d] = center ? center[d] : 0.;
  }
  // TPZAnalysis::PostProcessError (Analysis/TPZAnalysis.cpp:480-487,
  // 625-648): ervec receives the mesh totals; with store_error the element
  // values go to ElementSolution() (nvol x nerr, row-major).
  void PostProcessError(std::vector<double> &ervec, bool store_error = false) {
    pzg_exact ex = fExactDev;
    std::vector<double> u, du;
    if (ex.kind == PZG_EXACT_POINTS) {
      if (!fExact) DebugStop();  // no exact solution set
      int64_t n = 0;
      Check(pzg_error_points(fCtx.get(), ex.porder, &n, nullptr), "PostProcessError");
      std::vector<double> xq(3 * n);
      Check(pzg_error_points(fCtx.get(), ex.porder, &n, xq.data()), "PostProcessError");
      u.assign(n * fNState, 0.);
      du.assign(n * fDim * fNState, 0.);
      std::vector<double> loc(3), res(fNState), der(fDim * fNState);
      for (int64_t q = 0; q < n; q++) {
        for (int d = 0; d < 3; d++) loc[d] = xq[3 * q + d];
        std::fill(res.begin(), res.end(), 0.);
        std::fill(der.begin(), der.end(), 0.);
        fExact(loc, res, der);
        for (int k = 0; k < fNState; k++) u[q * fNState + k] = res[k];
        for (int k = 0; k < fDim * fNState; k++) du[q * fDim * fNState + k] = der[k];
      }
      ex.npts = n;
      ex.u = u.data();
      ex.du = du.data();
    }
    std::vector<double> el(fNVol * 6);
    ervec.assign(6, 0.);
    int32_t nerr = 0;
    Check(pzg_evaluate_error(fCtx.get(), &ex, el.data(), ervec.data(), &nerr), "PostProcessError");
    ervec.resize(nerr);
    if (store_error) {
      el.resize(fNVol * nerr);
      fElementSolution = std::move(el);
    }
  }
  const std::vector<double> &ElementSolution() const { return fElementSolution; }

 private:
  Ctx fCtx;
  std::unique_ptr<TPZSSpStructMatrixHIP> fStr;
  std::unique_ptr<TPZSYsmpMatrixHIP> fK;
  TPZStepSolver fSolver;
  std::vector<double> fRhs, fSolution, fElementSolution;
  pzg_mesh_desc fMesh;
  std::vector<int64_t> fPerm, fVolConn, fBcConn, fBlkPos;
  std::vector<int32_t> fBlkSize;
  int fDim = 0, fNState = 1;
  int64_t fNVol = 0;
  ExactSolType fExact;
  pzg_exact fExactDev{};
};

}  // namespace pzhip
