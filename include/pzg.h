/*
 * pzg.h -- C ABI of the MI355X (gfx950) linear FEM hot path.
 *
 * The entry points are what a NeoPZ drop-in pair of subclasses binds:
 *   TPZSSpStructMatrixHIP : TPZSSpStructMatrix<STATE, TPZStructMatrixOR<STATE>>
 *   TPZSYsmpMatrixHIP     : TPZSYsmpMatrix<STATE>
 * so that TPZLinearAnalysis::Assemble()/Solve() stay unchanged.  Each function
 * names the reference interface it replaces (paths relative to the NeoPZ
 * source tree).
 *
 * Conventions
 *  - Plain C types only; host pointers are borrowed for the duration of the
 *    call; device memory is owned by the context.
 *  - Every function returns PZG_OK (0) or an error code; pzg_last_error()
 *    gives the message of the last failure on the calling thread.  The C++
 *    adapter maps a non-zero code to PZError + DebugStop(), which is the
 *    reference's error behaviour (e.g. Matrix/pzsysmp.cpp:99-101).
 *  - One context = one GPU, one HIP stream; a context is used from one host
 *    thread at a time (the reference calls the struct matrix from one thread).
 *  - Equation numbering and the upper-triangular CSR pattern are exactly the
 *    reference's (StrMatrix/TPZSSpStructMatrix.cpp:49-193): IA/JA are int64,
 *    values double, rows sorted, diagonal first.
 */
#ifndef PZG_H
#define PZG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define PZG_OK 0
#define PZG_EINVAL 1   /* bad argument / inconsistent mesh description      */
#define PZG_EHIP 2     /* HIP runtime failure                               */
#define PZG_ENOCONV 3  /* solver hit max_iter before reaching tol           */
#define PZG_ESTATE 4   /* call out of order (e.g. assemble before create)   */
#define PZG_ENOMEM 5   /* device allocation failed                          */
#define PZG_EUNSUP 6   /* element/material combination not implemented     */

/* ---- element types, materials, loads ----------------------------------- */
#define PZG_LINE 1
#define PZG_QUAD 2
#define PZG_HEX 3
#define PZG_TRI 4  /* TPZTriangle: boundary elements of tetrahedral meshes */
#define PZG_TET 5  /* TPZTetrahedron, H1 p <= 5 (hp: per-element orders)  */

#define PZG_MAT_POISSON 0  /* TPZMatPoisson   (Material/Poisson/TPZMatPoisson.cpp) */
#define PZG_MAT_DARCY 1    /* TPZDarcyFlow    (Material/DarcyFlow/TPZDarcyFlow.cpp) */
#define PZG_MAT_ELAST3D 2  /* TPZElasticity3D (Material/Elasticity/TPZElasticity3D.cpp) */

#define PZG_FORCE_NONE 0
#define PZG_FORCE_CONST 1
#define PZG_FORCE_BUMPS 2  /* f(x) = sum_k a_k exp(-|x-c_k|^2/(2 w^2)), rows (cx,cy,cz,a0,a1,a2) */

/* TPZBndCondT types (TPZBndCondT::Type), per material:
 *   TPZMatPoisson::ContributeBC   (Material/Poisson/TPZMatPoisson.cpp:45-121):
 *     0 Dirichlet (penalty fBigNumber), 1 Neumann (Val2 * fScale),
 *     2 Robin as written there: penalty * Val2 load, penalty * Val1(0,0) times
 *       the first axis derivative (dphix row 0) of the boundary element;
 *   TPZDarcyFlow::ContributeBC    (Material/DarcyFlow/TPZDarcyFlow.cpp:49-91):
 *     0 Dirichlet, 1 Neumann, 2 Robin (Val1(0,0) mass term + Val2 load);
 *   TPZElasticity3D::ContributeBC (Material/Elasticity/TPZElasticity3D.cpp:
 *     616-773, penalty 1e12): 0 Dirichlet, 1 Neumann, 2 mixed (Val1 3x3 mass
 *     coupling + Val2 load), 3 directional null Dirichlet (penalty * Val2[k]
 *     on component k), 4 stress field (load -Val1 . n, n the outward normal of
 *     TPZInterpolationSpace::ComputeNormal), 5/6/7 Dirichlet on x/y/z only,
 *     8 Dirichlet on x and z.
 * Any other type is refused by pzg_mesh_set with PZG_EUNSUP (the reference
 * prints an error there and adds nothing, or stops). */
#define PZG_BC_DIRICHLET 0
#define PZG_BC_NEUMANN 1
#define PZG_BC_MIXED 2
#define PZG_BC_DIR_NULL 3
#define PZG_BC_STRESS 4
#define PZG_BC_DIR_X 5
#define PZG_BC_DIR_Y 6
#define PZG_BC_DIR_Z 7
#define PZG_BC_DIR_XZ 8

typedef struct pzg_ctx pzg_ctx;

/*
 * Flattened computational mesh: what TPZSSpStructMatrixHIP reads once from
 * TPZCompMesh (connect sequence numbers, Block() position/size, element
 * connect lists from TPZCompEl::BuildConnectList, geometric node coordinates,
 * material parameters and boundary conditions).
 */
typedef struct {
  int32_t dim;        /* 2 or 3                                               */
  int32_t material;   /* PZG_MAT_*                                            */
  int32_t nstate;     /* 1 (Poisson/Darcy) or 3 (Elasticity3D)                */
  int32_t porder;     /* H1 order: 1 or 2; 1..5 on tetrahedral meshes (the
                         highest order when vol_rule carries hp orders)      */
  int64_t nnodes;
  const double *xyz;  /* 3 * nnodes                                           */
  /* volume elements: PZG_QUAD (dim 2), PZG_HEX or PZG_TET (dim 3)           */
  int64_t nvol;
  const int64_t *vol_nodes; /* nvol * ncorner (4 or 8), NeoPZ corner order  */
  const int64_t *vol_conn;  /* nvol * nsides (9 or 27): connect seq per side */
  const double *vol_coef;   /* nvol: E (elasticity) or permeability (Darcy); NULL = constant */
  /* boundary elements: PZG_LINE (dim 2) or PZG_QUAD (dim 3)                  */
  int64_t nbc;
  const int64_t *bc_nodes;  /* nbc * ncorner (2 or 4)                        */
  const int64_t *bc_conn;   /* nbc * nsides (3 or 9)                          */
  const int32_t *bc_id;     /* nbc: index into the bc tables below            */
  int32_t nbcid;
  const int32_t *bc_type;   /* nbcid: PZG_BC_*                                */
  const double *bc_val;     /* 3 * nbcid: TPZBndCondT::Val2                   */
  /* (TPZBndCondT::Val1: bc_val1 at the end of this struct)                   */
  /* connects by sequence number: TPZCompMesh::Block() (Matrix/pzblock.h)     */
  int64_t nconn;
  const int64_t *blk_pos;
  const int32_t *blk_size;
  int64_t neq;
  /* material constants and loads                                             */
  double scale;       /* TPZMatPoisson fScale / constant Darcy permeability   */
  double E, nu;       /* TPZElasticity3D (E used when vol_coef == NULL)       */
  int32_t rule_order_vol; /* TPZMatSingleSpace::IntegrationRuleOrder result   */
  int32_t rule_order_bc;
  int32_t force_kind;
  double force_const[3];
  int32_t nbumps;
  const double *bumps;    /* 6 * nbumps                                       */
  double bump_width;
  double big_number;      /* Dirichlet penalty (TPZMaterial::fBigNumber or 1e12) */
  /* hp meshes: integration rule order of every element (the material's
   * IntegrationRuleOrder of the element's highest connect order,
   * TPZInterpolationSpace::AdjustIntegrationRule, Mesh/pzinterpolationspace.
   * cpp:69-83); NULL = rule_order_vol / rule_order_bc for all.  Connects of
   * order 1 on higher sides carry no shapes: give them blk_size 0.          */
  const int32_t *vol_rule;  /* nvol or NULL */
  const int32_t *bc_rule;   /* nbc or NULL  */
  /* volume element type: 0 = by dim (PZG_QUAD / PZG_HEX), or PZG_TET (then
   * vol_nodes has 4 corners and vol_conn 15 sides per element, boundary
   * elements are PZG_TRI with 3 corners and 7 sides)                        */
  int32_t vol_etype;
  /* constrained connects of hanging sides (TPZConnect::TPZDepend,
   * Mesh/pzconnect.h; TPZInterpolatedElement::RestrainSide,
   * Mesh/pzintel.cpp:847-1010): connect c depends on the connects
   * dep_conn[dep_ptr[c] .. dep_ptr[c+1]) with the coefficients dep_coef
   * (TPZDepend::fDepMatrix, one shape per connect).  Constrained connects
   * carry no equations: their blocks come after all the others in sequence
   * order.  The connects they depend on are unconstrained (one level).
   * Quadrilateral and hexahedral meshes (p <= 2), volume and boundary
   * elements (tetrahedral meshes: PZG_EUNSUP).
   * NULL = no constraints. */
  const int64_t *dep_ptr;   /* nconn + 1, or NULL */
  const int64_t *dep_conn;
  const double *dep_coef;
  /* TPZBndCondT::Val1 per bc id, 9 * nbcid row-major (nstate x nstate used:
   * Val1(0,0) for Poisson/Darcy, the 3x3 matrix for Elasticity3D); read by
   * types 2 (all materials) and 4 (Elasticity3D).  NULL = zero.  Type 4 also
   * needs each such boundary element on exactly one volume element (the
   * neighbour that orients the normal; an interior face is PZG_EUNSUP). */
  const double *bc_val1;
  /* TPZElasticity3D::fPreStress (XX, YY, ZZ): the load term
   * ef(3j+k) -= weight * prestress[k] * dphix(k, j) of Contribute
   * (Material/Elasticity/TPZElasticity3D.cpp:271-278); zeros = none. */
  double prestress[3];
} pzg_mesh_desc;

/* ---- solver selection -------------------------------------------------- */
#define PZG_SOLVER_PCG64 0   /* FP64 CG + Jacobi, reference semantics:
                                CG<> (Solvers/LinearSolvers/cg.h:29-132),
                                stop on the recursive ||r||/||b|| <= tol      */
#define PZG_SOLVER_MIXED 1   /* FP32 flexible CG on D^-1/2 K D^-1/2 inside an
                                FP64 iterative refinement (the shape of IR<>,
                                Solvers/LinearSolvers/ir.h:23-66); stop on the
                                true FP64 ||f - K u|| / ||f|| <= tol          */

#define PZG_PRECOND_JACOBI 0       /* point Jacobi                              */
#define PZG_PRECOND_BLOCK_JACOBI 1 /* connect-block Jacobi: the nstate x nstate
                                      diagonal block of each connect (one
                                      shape per connect on these meshes)
                                      inverted; with PZG_SOLVER_PCG64 the
                                      preconditioner of TPZAnalysis::
                                      BuildPreconditioner(EBlockJacobi, true),
                                      Analysis/TPZAnalysis.cpp:1062-1070      */
#define PZG_PRECOND_LEARNED 2      /* learned message-passing preconditioner    */
#define PZG_PRECOND_MULTIGRID 3    /* V-cycle: p = 2 -> vertex functions ->
                                      halved lattices, Chebyshev / block-Jacobi
                                      smoothing (hexahedral or quadrilateral
                                      meshes on a lattice)                      */

typedef struct {
  int32_t solver;        /* PZG_SOLVER_*                                       */
  int32_t precond;       /* PZG_PRECOND_*; PZG_SOLVER_PCG64 takes JACOBI
                            (EJacobi) or BLOCK_JACOBI (EBlockJacobi: each
                            connect's equations, up to 24 on high-order
                            tetrahedra; PZG_EUNSUP beyond or when filtered)   */
  int64_t max_iter;      /* total inner iterations allowed                     */
  double tol;            /* relative residual target                           */
  double inner_tol;      /* mixed: relative reduction per refinement step      */
  int32_t max_outer;     /* mixed: refinement steps allowed                    */
  int32_t check_every;   /* iterations per device batch between host checks   */
  int32_t use_graph;     /* capture the iteration batch in a hipGraph          */
  int32_t refine;        /* mixed: 0 = restart FCG from each true residual,
                                   1 = reliable update (keep the direction)    */
  int32_t inner_op;      /* mixed: PZG_INNER_* operator of the FP32 iterations */
  int32_t mg_degree;     /* multigrid: Chebyshev degree of the smoothing on the
                            p level (0 = 2; the lattice levels take 1)         */
  double mg_range;       /* multigrid: lmax / lmin of the smoothed interval
                            (0 = 30; 60 with PZG_PRECOND_LEARNED)              */
} pzg_solve_params;

/* Operator of the mixed solver's FP32 iterations.  Both approximate the
 * scaled S = D^-1/2 K D^-1/2 in FP32; the FP64 true residual uses the
 * assembled CSR either way. */
#define PZG_INNER_AUTO 0     /* matrix-free when the system is mesh-built, else BSR */
#define PZG_INNER_BSR 1      /* stored FP32 3x3 (or 1x1) block CSR of S             */
#define PZG_INNER_MATFREE 2  /* element-by-element FP32 application from geometry   */

typedef struct {
  int64_t iterations;     /* CG iterations (inner total for the mixed solver)  */
  int32_t outer;          /* refinement steps                                   */
  int32_t converged;
  double achieved;        /* PCG64: recursive ||r||/||b|| (TPZStepSolver fTol)  */
  double true_residual;   /* ||f - K64 u|| / ||f|| in FP64                      */
  double solve_ms;        /* device time of the solve                           */
  int32_t fallback;       /* mixed: 1 if the FP32 iterations stalled (the true
                             residual did not halve over three refinement
                             steps, the refinement budget ran out, or with the
                             learned preconditioner one step gained < 10%)
                             and plain FP64 Jacobi-PCG (CG<> + EJacobi, from
                             the current solution) finished the solve        */
} pzg_solve_info;

/* ---- lifecycle ---------------------------------------------------------- */
const char *pzg_last_error(void);
int pzg_create(int device, pzg_ctx **out);
void pzg_destroy(pzg_ctx *ctx);

/* Flattens the computational mesh into device arrays (geometry, connect
 * blocks, coefficients, shape tables at the quadrature points).
 * Replaces the element loop set-up of TPZStructMatrixOR::Serial_Assemble
 * (StrMatrix/pzstrmatrixor.cpp:104-180). */
int pzg_mesh_set(pzg_ctx *ctx, const pzg_mesh_desc *mesh);

/* Symbolic phase: the reference's upper-triangular CSR pattern, computed from
 * the connect graph.  Replaces TPZSSpStructMatrix::Create / SetupMatrixData
 * (StrMatrix/TPZSSpStructMatrix.cpp:30-193) with
 * TPZRenumbering::ConvertGraph (External/TPZRenumbering.cpp:76-110). */
int pzg_struct_create(pzg_ctx *ctx, int64_t *neq, int64_t *nnz);

/* Equation filter of the struct matrix: TPZEquationFilter::SetActiveEquations
 * (StrMatrix/TPZEquationFilter.h; active sorted, duplicates removed) on
 * TPZStructMatrix::EquationFilter(), or Reset() when active is NULL or
 * nactive < 0.  Call after pzg_mesh_set (a new mesh resets the filter) and
 * before pzg_struct_create, which then builds the condensed pattern of
 * SetupMatrixData with the filter (StrMatrix/TPZSSpStructMatrix.cpp:49-193;
 * neq = the number of active equations) and pzg_assemble the condensed
 * system (TPZStructMatrixOR::Assemble filtering every element's destinations,
 * StrMatrix/pzstrmatrixor.cpp:41-71,226-233).  With a filter, every vector of
 * the matrix-level entry points (pzg_get_system's rhs, pzg_multadd,
 * pzg_solve, pzg_add_kel_batch destinations) is in the condensed numbering;
 * the post-processing entry points (pzg_load_solution, pzg_get_solution,
 * pzg_evaluate_error) keep the mesh numbering, inactive equations 0
 * (TPZEquationFilter::Scatter, as TPZLinearAnalysis::SolveT loads it).
 * The mixed solver then uses the stored FP32 operator, and pzg_full_csr
 * returns PZG_EUNSUP. */
int pzg_set_equation_filter(pzg_ctx *ctx, int64_t nactive, const int64_t *active);

/* Copies the pattern to host arrays ia[neq+1], ja[nnz] (TPZSYsmpMatrix::IA/JA,
 * Matrix/pzsysmp.h:104-127). */
int pzg_struct_get_pattern(pzg_ctx *ctx, int64_t *ia, int64_t *ja);

/* Numeric assembly into the pattern: element integration + deterministic
 * coloured scatter.  Replaces TPZStructMatrixOR::Assemble (StrMatrix/
 * pzstrmatrixor.cpp:41-71) -> CalcStiff (Mesh/pzinterpolationspace.cpp:403-474)
 * -> TPZSYsmpMatrix::AddKel (Matrix/pzsysmp.cpp:369-411) and
 * TPZFMatrix::AddFel (Matrix/pzfmatrix.cpp:284-298), then
 * EndCreateAssemble -> ComputeDiagonal (StrMatrix/TPZSSpStructMatrix.cpp:24-28). */
int pzg_assemble(pzg_ctx *ctx);

/* Copies the assembled values a[nnz] and rhs[neq] to the host (either may be
 * NULL).  TPZSYsmpMatrix::A (Matrix/pzsysmp.h:104-127) and the TPZFMatrix rhs. */
int pzg_get_system(pzg_ctx *ctx, double *a, double *rhs);

/* The assembled system in the non-symmetric sparse format of
 * TPZSpStructMatrix / TPZFYsmpMatrix: every equation row holds the equations
 * of its own and of every adjacent connect, sorted (both triangles;
 * StrMatrix/TPZSpStructMatrix.cpp:53-190), values as TPZFYsmpMatrix::AddKel
 * accumulates them (Matrix/pzysmp.cpp:135-176, |v| < 1e-12 dropped).
 * ia[neq+1], ja[nnz_full], a[nnz_full] (nnz_full from pzg_get_sizes); any of
 * them may be NULL.  Needs an assembled system. */
int pzg_full_csr(pzg_ctx *ctx, int64_t *ia, int64_t *ja, double *a);

/* Installs a caller-built matrix (and optional rhs): TPZSYsmpMatrix::SetData
 * (Matrix/pzsysmp.h:232-240). */
int pzg_csr_set(pzg_ctx *ctx, int64_t neq, const int64_t *ia, const int64_t *ja,
                const double *a, const double *rhs);

/* AddKel of a batch of caller-given element matrices (row-major ne x ne each,
 * with dest[nel*ne] global equations) into the current pattern, on the
 * device, including the |v| < 1e-12 drop rule (Matrix/pzsysmp.cpp:381,
 * Common/pzreal.h:736-741). */
int pzg_add_kel_batch(pzg_ctx *ctx, int64_t nel, int32_t ne, const int64_t *dest,
                      const double *ek);

/* z = beta*y + alpha*K*x on host vectors of length neq.
 * TPZSYsmpMatrix::MultAdd (Matrix/pzsysmp.cpp:189-232); y may be NULL when
 * beta == 0. */
int pzg_multadd(pzg_ctx *ctx, const double *x, const double *y, double *z, double alpha,
                double beta);

/* Solve K u = rhs.  rhs: host vector or NULL (use the assembled rhs on the
 * device).  u: host output or NULL (keep on the device).
 * PZG_SOLVER_PCG64 replaces TPZMatrix::SolveCG (Matrix/pzmatrix.cpp:750-765)
 * with the EJacobi TPZStepSolver preconditioner (Solvers/pzstepsolver.cpp:
 * 65-122, Matrix/pzmatrix.cpp:642-667), x starting at zero. */
int pzg_solve(pzg_ctx *ctx, const pzg_solve_params *params, const double *rhs, double *u,
              pzg_solve_info *info);

/* FP64 true relative residual ||rhs - K u|| / ||rhs|| of the device solution. */
int pzg_true_residual(pzg_ctx *ctx, double *rel);

/* pzg_true_residual without any allocation: no counterpart in the reference;
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc).  The FP64 block
 * CSR and the solution must already be on the device (after a solve,
 * pzg_multadd or pzg_true_residual of the current assembly), else
 * PZG_ESTATE.  Reduces in a fixed per-GPU static scratch of 32.8 KB (2 x 2048
 * partial sums, the two norms, a counter), set up by pzg_create and held
 * until the process ends; any system size fits (the residual grid is capped
 * at 2048 workgroups).  Every context on one GPU shares that scratch: calls
 * that overlap on one GPU (different streams or threads) race. */
int pzg_true_residual_noalloc(pzg_ctx *ctx, double *rel);

/* Average device duration (HIP events on the context stream) of nrep launches
 * of the SpMV kernel of the given format: 0 = FP64 full CSR (plain PCG path),
 * 1 = FP32 block CSR of the mixed path, 2 = element assembly (whole pass),
 * 3 = matrix-free FP32 operator (element, boundary and row-sum kernels),
 * 4 = its volume element kernel alone, for which *bytes receives the
 * algorithmic flops of one launch instead of bytes. */
int pzg_time_kernel(pzg_ctx *ctx, int32_t which, int32_t nrep, double *avg_ms, double *bytes);

/* No counterpart in the reference (measurement): the choices of the
 * assembly-kernel autotuner of high-order tetrahedral meshes.  On entry *n is
 * the capacity of the arrays; on return it is the number of rule buckets,
 * order[b] the bucket's highest element order, variant[b] the kernel variant
 * kept and ms[b*5 + v] the measured milliseconds per launch of variant v
 * (scalar materials 0: 16 points/pass, 4x4 tiles; 1: 8, 4x4; 2: 32, 4x4;
 * 3: 16, 2x2; 4: 32, 2x2; Elasticity3D 0: 16, 2x2; 1: 8, 2x2; 2: 32, 2x2;
 * 3: 16, 1x1; 4: 8, 1x1).  *n = 0 when the mesh has no high-order
 * tetrahedra or has not been assembled. */
int pzg_hp_tuning(pzg_ctx *ctx, int32_t *n, int32_t *order, int32_t *variant, float *ms);

/* y = S x with S = D^-1/2 K D^-1/2 applied by the FP32 inner operator op
 * (PZG_INNER_BSR or PZG_INNER_MATFREE) on host vectors of length neq; D is
 * the diagonal of the assembled K.  Exposes the operator of the mixed
 * solver's inner iterations, which has no reference counterpart. */
int pzg_scaled_apply(pzg_ctx *ctx, int32_t op, const double *x, double *y);

/* z = M r with M the V-cycle of PZG_PRECOND_MULTIGRID for the scaled system
 * S (level 0 applied by op: PZG_INNER_BSR or PZG_INNER_MATFREE; op |
 * PZG_MG_LEARNED smooths level 0 with the learned element smoother, the
 * cycle of PZG_PRECOND_LEARNED, and lmax[0] is then that of M_learned S), on host
 * vectors of length neq; degree and range as mg_degree and mg_range of
 * pzg_solve_params, 0 for the defaults.  *nlev receives the number of lattice levels (the first is the
 * trilinear subspace of the mesh's vertices, the last the coarsest: solved
 * densely when it has at most 1536 unknowns, else by a degree-8 Chebyshev
 * polynomial), dims[3l .. 3l+2] their node counts per direction, lmax[0] the
 * estimated largest eigenvalue of the block-Jacobi scaled S and lmax[1+l]
 * that of lattice level l (0 for a densely solved coarsest level; arrays of
 * 48 and 16 entries; any may be NULL).  The
 * levels are built on the first call after an assembly.  Exposes the mixed
 * solver's multigrid preconditioner, which has no reference counterpart. */
#define PZG_MG_LEARNED 4
int pzg_mg_apply(pzg_ctx *ctx, int32_t op, int32_t degree, double range, const double *r, double *z,
                 int32_t *nlev, int32_t *dims, double *lmax);

/* ---- bandwidth renumbering ------------------------------------------------ */

/* TPZAnalysis::OptimizeBandwidth (Analysis/TPZAnalysis.cpp:225-300) with the
 * default TPZSloanRenumbering (RENUMBER, :74-78): the permutation of
 * TPZSloanRenumbering::Resequence (External/TPZSloanRenumbering.cpp:23-161)
 * of the element graph of TPZCompMesh::ComputeElGraph (Mesh/pzcmesh.cpp:
 * 1223-1266): elgraph[elgraphindex[e] .. elgraphindex[e+1]) are the sequence
 * numbers of element e's independent connects, nnodes the number of
 * independent connects.  perm[seq] receives the new sequence number (the
 * permGather handed to TPZCompMesh::Permute); apply it to blk_pos/blk_size
 * and the element connect lists before pzg_mesh_set.  Host computation, no
 * context and no GPU.  PZG_EINVAL if the graph is not connected (the
 * reference stops there too). */
int pzg_sloan_renumber(int64_t nnodes, int64_t nel, const int64_t *elgraphindex,
                       const int64_t *elgraph, int64_t *perm);

/* ---- post-processing: solution and error norms -------------------------- */

/* Exact solutions for pzg_evaluate_error.  PZG_EXACT_POINTS: the caller
 * evaluates its own exact-solution function (the std::function of
 * TPZMatError::SetExactSol, Material/TPZMatError.h:35-48) at the points that
 * pzg_error_points returns and hands over the values.  The other kinds are
 * evaluated on the device: scalar solutions of TLaplaceExample1::uxy
 * (Pre/TPZAnalyticSolution.cpp:1254-1310) with fCenter = center over the
 * mesh dimension d, and their exact gradients:
 *   CONST  u = constant (TLaplaceExample1::gC)
 *   X      u = x_0 - c_0
 *   SINSIN u = prod_{i<d} sin(pi (x_i - c_i))
 *   COSCOS u = prod_{i<d} cos(pi (x_i - c_i) / 2) */
#define PZG_EXACT_POINTS 0
#define PZG_EXACT_CONST 1
#define PZG_EXACT_X 2
#define PZG_EXACT_SINSIN 3
#define PZG_EXACT_COSCOS 4

typedef struct {
  int32_t kind;       /* PZG_EXACT_*                                          */
  int32_t porder;     /* polynomial order given with the exact solution
                         (TPZMatError::fExactPOrder): raises the element rules */
  double center[3];   /* device kinds: TLaplaceExample1::fCenter              */
  double constant;    /* PZG_EXACT_CONST: the value                           */
  int64_t npts;       /* PZG_EXACT_POINTS: number of points (pzg_error_points) */
  const double *u;    /* PZG_EXACT_POINTS: [npts][nstate] exact values        */
  const double *du;   /* PZG_EXACT_POINTS: [npts][dim][nstate] derivatives,
                         du[q][i][j] = d u_j / d x_i (the layout of dsol and of
                         the exact function's deriv matrix)                  */
} pzg_exact;

/* Replaces the device solution used by post-processing with u[neq] (host).
 * TPZCompMesh::LoadSolution (Mesh/pzcmesh.cpp:524-540); pzg_solve leaves its
 * solution loaded already (TPZLinearAnalysis::Solve calls LoadSolution). */
int pzg_load_solution(pzg_ctx *ctx, const double *u);

/* Copies the loaded solution (the last solve's or pzg_load_solution's) to
 * u[neq]: TPZLinearAnalysis::Solution / TPZCompMesh::Solution
 * (Analysis/TPZAnalysis.h, Mesh/pzcmesh.h). */
int pzg_get_solution(pzg_ctx *ctx, double *u);

/* Integration points of the error norms of every volume element for an exact
 * solution of order porder: each element's rule is raised to
 * max(its rule order, porder), or to the rule family's highest order when
 * porder reaches it (TPZInterpolationSpace::EvaluateError,
 * Mesh/pzinterpolationspace.cpp:1140-1160; highest orders 36 line/quad/hex,
 * 21 triangle, 14 tetrahedron, Integral/pzquad.h, tpzintrulet.h:72,
 * tpzintrulet3d.h:72).  *npts receives the number of points; xq (NULL or
 * [npts][3]) their physical coordinates, element by element in element
 * order, points in rule order. */
int pzg_error_points(pzg_ctx *ctx, int32_t porder, int64_t *npts, double *xq);

/* Error norms of the loaded solution against an exact solution.
 * TPZAnalysis::PostProcessError -> TPZCompMesh::EvaluateError
 * (Analysis/TPZAnalysis.cpp:625-648, Mesh/pzcmesh.cpp:1497-1552) ->
 * TPZInterpolationSpace::EvaluateError (Mesh/pzinterpolationspace.cpp:
 * 1108-1207) with the material's Errors at every point:
 *   TPZMatPoisson::Errors (Material/Poisson/TPZMatPoisson.cpp:173-218),
 *     3 values: H1, L2, H1 semi-norm;
 *   TPZDarcyFlow::Errors (Material/DarcyFlow/TPZDarcyFlow.cpp:268-326),
 *     6 values: energy, L2, semi-norm, flux components;
 *   TPZElasticity3D::Errors (Material/Elasticity/TPZElasticity3D.cpp:
 *     1081-1133), 3 values: energy, L2, H1 semi-norm.
 * el_errors (NULL or [nvol][nerr]): per volume element the square roots of
 * the integrated values (TPZCompMesh::ElementSolution with store_error);
 * errors (NULL or [nerr]): the mesh totals sqrt(sum_e el^2); *nerr (may be
 * NULL) receives the number of values. */
int pzg_evaluate_error(pzg_ctx *ctx, const pzg_exact *exact, double *el_errors, double *errors,
                       int32_t *nerr);

/* Sizes of the device formats (for roofline accounting). */
typedef struct {
  int64_t neq, nnz_upper, nnz_full, nblocks, nnz_blocks, nvol, ncolors;
  int32_t bs;  /* block size of the mixed-path block CSR */
} pzg_sizes;
int pzg_get_sizes(pzg_ctx *ctx, pzg_sizes *s);

/* Synchronise the context stream. */
int pzg_sync(pzg_ctx *ctx);

/* Replaces the per-element coefficients of the current mesh (a new
 * TPZElasticity3D E / TPZDarcyFlow permeability field on the same mesh);
 * the next pzg_assemble uses them.  Same role as re-running the element loop
 * of TPZStructMatrixOR::Assemble (StrMatrix/pzstrmatrixor.cpp:104-365) after
 * changing material data. */
int pzg_mesh_set_coef(pzg_ctx *ctx, const double *vol_coef);

/* ---- learned preconditioner (no counterpart in the reference) -------------
 * PZG_PRECOND_LEARNED: the multigrid V-cycle of PZG_PRECOND_MULTIGRID whose
 * smoothing on the p level uses, in place of the block-Jacobi M^-1, the
 * learned element smoother
 *     z = sum_e s_e P_e^T G P_e r  (z_i = r_i on Dirichlet-penalty equations),
 * G one learned symmetric ne x ne matrix (ne = shapes x states of the volume
 * element) and s_e the output of a message-passing network (3 layers, 32
 * features) on the element graph (elements adjacent through a face) from
 * four features of the element's coefficient field and boundary.  Trained
 * with the project's kernels on (residual, error) pairs of the solver's local
 * problems: per element of a training system, random residuals r and the
 * errors e = S_e^-1 r of its block S_e of the assembled scaled matrix.
 * Quadrilateral / hexahedral lattice meshes with one rule, no constraints,
 * no equation filter.  NeoPZ has no learned component; these entry points
 * serve the mixed solver only. */
/* Parameters for the context's mesh: G = diag(1 / elements sharing the side
 * of each dof's shape), network weights from seed, output layer 0 (s = 1). */
int pzg_learned_setup(pzg_ctx *ctx, uint64_t seed);
/* Training pairs of the assembled system (replacing earlier ones): nrhs
 * (1..8) Gaussian residuals per element (zero on penalized dofs, seeded) and
 * their local errors. */
int pzg_learned_collect(pzg_ctx *ctx, int32_t nrhs, uint64_t seed);
/* Training: G = the least-squares fit B A^-1 (A = sum r r^T, B = sum e r^T
 * over the pairs of every collect since pzg_learned_setup, elements without
 * penalized dofs), symmetrised and projected onto the positive definite
 * matrices; then steps Adam steps (learning rate lr) of the network on the
 * stored pairs.  The loss sum |mask (s_e G r - e)|^2 / sum |e|^2 after the
 * fit and after the last step. */
int pzg_learned_train(pzg_ctx *ctx, int32_t steps, double lr, double *loss_first, double *loss_last);
/* Reads (get) and/or writes (set) the FP32 parameters; *n = their number.
 * Layout: G (ne x ne, row-major), W0 (4 x 32), A_l, B_l (32 x 32, l = 1..3,
 * interleaved A_1 B_1 A_2 ...), w (32); weights [in][out] row-major. */
int pzg_learned_weights(pzg_ctx *ctx, float *get, const float *set, int32_t *n);
/* The element graph (cia[nel+1], cja[cia[nel]]), the features of the
 * current assembly ([nel][4]) and the penalized-equation mask ([neq], 1 =
 * penalized); any pointer may be NULL. */
int pzg_learned_graph(pzg_ctx *ctx, int64_t *nel, int64_t *cia, int32_t *cja, float *features,
                      float *pen);
/* The stored pairs R, E ([element][pair][ne]) and the dof mask ([element][ne],
 * 0 on penalized or absent dofs); *ne and *nr receive their sizes. */
int pzg_learned_samples(pzg_ctx *ctx, int32_t *ne, int32_t *nr, float *R, float *E, float *mask);
/* Loss and gradient (nparams floats) of the current parameters on the stored
 * pairs. */
int pzg_learned_loss_grad(pzg_ctx *ctx, double *loss, float *grad);
/* z = M r of the learned smoother of the current assembly on host vectors
 * of length neq (scaled unknowns, FP32 on the device). */
int pzg_learned_apply(pzg_ctx *ctx, const double *r, double *z);
/* The scales s_e ([nel]) the network gives for the current assembly. */
int pzg_learned_scales(pzg_ctx *ctx, float *s);

#ifdef __cplusplus
}
#endif
#endif
