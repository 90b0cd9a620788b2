// pzg_internal.h -- shared declarations of the HIP linear FEM path.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/pzg.h"

namespace pzg {

void set_error(const std::string &msg);

#define PZG_HIP(call)                                                                 \
  do {                                                                                \
    hipError_t e_ = (call);                                                           \
    if (e_ != hipSuccess) {                                                           \
      ::pzg::set_error(std::string(#call) + ": " + hipGetErrorString(e_));            \
      return PZG_EHIP;                                                                \
    }                                                                                 \
  } while (0)

#define PZG_TRY(call)              \
  do {                             \
    int rc_ = (call);              \
    if (rc_ != PZG_OK) return rc_; \
  } while (0)

// Device buffer owned by the context.
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  template <class T> T *as() const { return static_cast<T *>(p); }
};

int dev_alloc(DevBuf &b, size_t bytes);
void dev_free(DevBuf &b);
template <class T> int dev_upload(DevBuf &b, const T *h, size_t n, hipStream_t s) {
  PZG_TRY(dev_alloc(b, n * sizeof(T)));
  if (n) PZG_HIP(hipMemcpyAsync(b.p, h, n * sizeof(T), hipMemcpyHostToDevice, s));
  return PZG_OK;
}

// Learned preconditioner network: hidden features, message-passing layers,
// leaky-ReLU slope.
constexpr int NNF = 32;
constexpr int NNL = 3;
constexpr float NN_SLOPE = 0.1f;

// Boundary condition of one bc id as coefficients of the boundary element's
// integrals (filled by mesh_set from the material's ContributeBC):
//   ek(I,J) += ck[ci][cj] M_ij + cd D_ij,   ef(I) += cf[ci] m_i
//            + stress ? sum_q w_q phi_i(q) (-sum_m val1[ci][m] n_m(q)) : 0
// with M_ij = sum_q w|J| phi_i phi_j, m_i = sum_q w|J| phi_i and
// D_ij = sum_q w|J| dphix0_i dphix0_j (derivative along the first axis).
constexpr int BCC = 24;
enum { BCC_CK = 0, BCC_CF = 9, BCC_CD = 12, BCC_STRESS = 13, BCC_VAL1 = 14 };

// Maximum sizes of the element tables (hexahedron p=2).
constexpr int MAXQ = 27;      // quadrature points per element
constexpr int MAXSH = 27;     // shape functions per element
constexpr int MAXSIDES = 27;  // sides per element

// Reference-element tables at the quadrature points (host-computed once).
struct ElemTables {
  int etype = 0, dim = 0, ncorner = 0, nsides = 0, nsh = 0, nq = 0;
  int shape_side[MAXSH];       // side that owns shape i (p <= 2: one shape per side)
  double w[MAXQ];              // quadrature weights
  double phi[MAXQ][MAXSH];     // shape values
  double dphi[MAXQ][3][MAXSH]; // reference derivatives
  double cphi[MAXQ][8];        // corner (geometric) shape values
  double cdphi[MAXQ][3][8];    // corner reference derivatives
};

// High-order triangles / tetrahedra (orders 3..5).  Every shape of the
// maximum-order set is a "virtual side" with one connect of nstate equations:
// a side of order k < P carries the first NConnectShapeF(side, k) functions of
// the order-P set (edge: Chebyshev degree; face: total degree; interior at
// its own order, prefix of the order-P set for P <= 5).  A shape's value
// depends on the id-based transform of its real side (edge: 2 variants,
// face: 6), so every variant has its own table column.
constexpr int HP_MAXP = 5;       // highest supported order
constexpr int HP_MAXQ = 343;     // collapsed rule, 7 points per direction
constexpr int HP_MAXSH = 56;     // tetrahedron of order 5
constexpr int HP_MAXCOL = 200;   // tetrahedron of order 5, all variants

struct HpTables {
  int etype = 0, dim = 0, ncorner = 0, order = 0, nsh = 0, nq = 0, ncol = 0;
  int shape_side[HP_MAXSH];      // real side of shape i
  int shape_loc[HP_MAXSH];       // index of shape i among its side's functions
  int col_off[HP_MAXSH];         // table column of variant 0 (variant v: col_off + v)
  double w[HP_MAXQ];             // quadrature weights
  double pt[HP_MAXQ][3];         // quadrature points
  double phi[HP_MAXQ][HP_MAXCOL];
  double dphi[HP_MAXQ][3][HP_MAXCOL];
};

// Tables of an error integration rule (post-processing), for kernels with
// lanes over quadrature points: every array has the point index fastest.
// d = w[nq] | phi[ncol][nq] | dphi[dim][ncol][nq] | cphi[nc][nq] |
// cdphi[dim][nc][nq] (cphi: corner functions of the geometric map).  Columns
// are the shapes of order p (p <= 2, column = side) or, for high-order
// simplices, one column per shape and side variant (col_off[i] + variant).
struct PostTables {
  int etype = 0, dim = 0, nc = 0, nq = 0, ncol = 0, nsh = 0;
  int col_off[HP_MAXSH] = {}, shape_side[HP_MAXSH] = {};
  std::vector<double> d;
};

// Transform id of a triangle from its three corner ids
// (TPZShapeTriang::GetTransformId2dT, Shape/pzshapetriang.cpp:410-431).
__host__ __device__ inline int tri_tid(int64_t i0, int64_t i1, int64_t i2) {
  const int a = i0 < i1 ? 0 : 1;
  const int64_t ia = a ? i1 : i0;
  const int mn = i2 < ia ? 2 : a;
  const int64_t n1 = mn == 0 ? i1 : mn == 1 ? i2 : i0;
  const int64_t n2 = mn == 0 ? i2 : mn == 1 ? i0 : i1;
  return 2 * mn + (n1 < n2 ? 0 : 1);
}

// Variant of side s of a tetrahedron / triangle with corner ids id[]
// (TPZTetrahedron / TPZTriangle::GetTransformId, Topology/tpztetrahedron.cpp:
// 1128-1168, Topology/tpztriangle.cpp:629-650).
__host__ __device__ inline int hp_variant(bool tet, int s, const int64_t *id) {
  if (tet) {
    if (s < 4 || s == 14) return 0;
    if (s < 10) {
      const int a = s < 7 ? s - 4 : s - 7, b = s < 7 ? (s - 3) % 3 : 3;
      return id[a] < id[b] ? 0 : 1;
    }
    const int f = s - 10;
    const int f0 = f == 2 ? 1 : 0, f1 = f == 0 || f == 1 ? 1 : 2, f2 = f == 0 ? 2 : 3;
    return tri_tid(id[f0], id[f1], id[f2]);
  }
  if (s < 3) return 0;
  if (s < 6) return id[s - 3] < id[(s - 2) % 3] ? 0 : 1;
  return tri_tid(id[0], id[1], id[2]);
}

// Full block CSR values (FP64 K, FP32 S): the blocks of a block row come in
// chunks of BSR_CHUNK, component-major inside a chunk, so that the BSR_CHUNK
// lanes that share a block row read consecutive addresses with every load.
// Position of value k (row-major inside the block) of the i-th block of a
// row whose blocks are [b0, b0 + n).
constexpr int BSR_CHUNK = 16;
constexpr int VBJ_MAX = 24;  // largest connect block of the high-order block Jacobi
template <int BB>
__host__ __device__ inline int64_t bsr_pos(int64_t b0, int64_t n, int64_t i, int k) {
  const int64_t j = i - i % BSR_CHUNK;
  const int64_t m = n - j < BSR_CHUNK ? n - j : BSR_CHUNK;
  return (b0 + j) * BB + (int64_t)k * m + (i - j);
}

// Device-side scalar state of an iterative solve.
struct SolveScalars {
  double rho, rho_old, pq, alpha, beta, rr, normb, resid, tol, pr, zq;
  double rlast;   // mixed: scaled residual norm at the last reliable update
  double delta;   // mixed: reduction of the recursive residual that triggers an update
  int64_t iter, maxit;
  // done: 0 running, 1 tol, 2 maxit, 3 divergence, 4 update due (set by the
  // iteration's last reduction; its completion kernels still run), 5 update
  // due (completion kernels already ran)
  int32_t done, pad;
};

// A lattice level of the multigrid preconditioner (pzg_mg.hip): n[0] x n[1] x
// n[2] nodes, a 27-point block stencil (entry (s, e) of node v at
// A[(s*bs*bs + e)*nn + v]) and the inverse diagonal blocks ([e*nn + v]).
struct MgLevel {
  int n[3] = {0, 0, 0};
  int64_t nn = 0;
  double lmax = 0.;  // largest eigenvalue of M^-1 A (CG-Lanczos estimate)
  DevBuf A, minv, x, b, r, d, t;
};

// Elements (volume or boundary) that hold a constrained connect with shapes.
struct CelSet {
  std::vector<int32_t> cel;          // element ids, grouped by rule bucket
  std::vector<char> is_cel;          // per element
  std::vector<int64_t> xptr, xconn;  // per element: the connects its constrained ones depend on
  std::vector<int64_t> bptr;         // bucket b: cel[bptr[b] .. bptr[b+1])
  std::vector<int64_t> cptr;         // colour k: d_cslot[cptr[k] .. cptr[k+1])
  DevBuf d_cel, d_cslot, d_T, d_blk, d_n, d_ek, d_ef;
};

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;

  // ---- mesh (host copies of the description) --------------------------
  bool has_mesh = false;
  int dim = 0, material = 0, nstate = 1, porder = 1, bs = 1;
  int64_t nnodes = 0, nvol = 0, nbc = 0, nconn = 0, neq = 0;
  std::vector<double> xyz, vol_coef, bc_val, bc_val1, bumps;
  // boundary conditions as coefficients of the boundary element's matrices
  // (bc_coef_of): per bc id BCC doubles; per boundary element the vector from
  // its volume neighbour's centre to its own (stress-field type only)
  std::vector<double> bc_coef, bc_vec;
  bool bc_mf_ok = true;  // every boundary matrix is the Dirichlet penalty or zero
  std::vector<int64_t> vol_nodes, vol_conn, bc_nodes, bc_conn, blk_pos;
  std::vector<int32_t> blk_size, bc_id, bc_type;
  double scale = 1, E = 1, nu = 0.3, big = 1e12, bump_width = 0.1, force_const[3] = {0, 0, 0};
  double prestress[3] = {0, 0, 0};  // TPZElasticity3D::fPreStress
  int force_kind = 0, nbumps = 0, rule_vol = 0, rule_bc = 0, nbcid = 0;
  ElemTables tvol, tbc;        // tables of the first rule bucket
  // hp: elements bucketed by integration rule (one table per bucket); the
  // colour lists are grouped by bucket: list k covers colour k / nbk, bucket
  // k % nbk
  std::vector<int> vol_rules, bc_rules;
  std::vector<int32_t> vol_bucket, bc_bucket;
  std::vector<ElemTables> tvol_all, tbc_all;
  int vol_ncorner = 0, vol_nsides = 0, bc_ncorner = 0, bc_nsides = 0;
  int vol_etype = 0, bc_etype = 0;  // PZG_QUAD/PZG_HEX/PZG_TET, PZG_LINE/PZG_QUAD/PZG_TRI
  // high-order simplices: connects expanded into one virtual connect per
  // shape (vol_conn/bc_conn/blk_* then hold the virtual ones); hp_order = P
  bool hp_simplex = false;
  int hp_order = 0;
  int64_t nconn_in = 0;              // connects as given by the caller
  std::vector<int> vol_side_real;    // virtual side -> real side (volume elements)
  std::vector<HpTables> thp_vol, thp_bc;
  // per rule bucket: highest order of its elements and the sub-list of the
  // virtual sides an element of that order can use (HP_MAXSH slots a bucket)
  std::vector<int> hp_vol_P, hp_bc_P;
  DevBuf d_hp_vsub, d_hp_bsub;
  // volume kernel variant per bucket, picked by timing at the first assembly
  // (hp_tune), and the measured milliseconds per launch of every variant
  bool hp_tuned = false;
  std::vector<int> hp_vol_var;
  std::vector<float> hp_vol_ms;

  // connect -> active block index (-1 for size-0 and constrained connects)
  std::vector<int64_t> aidx;
  // affine volume elements first in each colour (count per colour list) and
  // the reference Gram matrices of the rule (k_vol_aff)
  std::vector<int64_t> vcolor_aff;
  DevBuf d_vol_R;
  bool aff_on = true;
  // all volume elements one affine hexahedron: stiffness assembled by blocks
  // (k_vol_blk) from the coefficient-1 element matrix and the connects'
  // element lists
  bool blk_asm = false;
  int blk_grid = 0;
  DevBuf d_blk_iptr, d_blk_inc, d_blk_kref;
  int64_t nb = 0;  // active blocks

  // constrained connects (hanging sides): dependencies per connect, and the
  // volume elements that hold a constrained connect with shapes.  Such an
  // element is assembled apart: its element matrix (shapes of every side)
  // goes to a buffer, then K_c = T^T K T over its constrained list (the
  // unconstrained blocks of its connects and of those they depend on), and
  // K_c is added with AddKel's drop rule.  T[i][a]: shape i -> list entry a.
  bool has_dep = false;
  std::vector<int64_t> dep_ptr, dep_conn;
  std::vector<double> dep_coef;
  std::vector<char> is_dep;
  CelSet cv, cb;                          // constrained volume / boundary elements
  DevBuf d_aidx_ek;                       // connect -> block, -2 constrained, -1 no shapes
  DevBuf d_dep_ptr, d_dep_blk, d_dep_coef;

  // ---- equation filter (TPZEquationFilter held by the struct matrix) -------
  // filt_active: active equations of the mesh numbering, ascending;
  // filt_dest[eq]: condensed index or -1.  The device keeps the system in the
  // mesh numbering with every inactive equation decoupled (unit diagonal, zero
  // row, column and load), which has the condensed system's iterates on the
  // active equations; the condensed pattern and values (the reference's
  // TPZSYsmpMatrix) are gathered through filt_keep, the positions of the
  // condensed entries in the mesh-numbered upper CSR.
  bool filt = false;
  std::vector<int64_t> filt_active, filt_dest, filt_keep, filt_ia;
  DevBuf d_filt_dest;

  // ---- pattern (device) -------------------------------------------------
  bool has_struct = false, external_csr = false;
  int64_t nnz_u = 0, nnz_full = 0, nbu = 0, nbf = 0;
  DevBuf d_ia, d_ja;          // reference upper CSR: int64 IA[neq+1], JA[nnz_u]
  DevBuf d_a, d_rhs, d_diag;  // values (fp64), rhs, diagonal (ComputeDiagonal)
  DevBuf d_ubia, d_ubja;      // upper block graph: int64 [nb+1], int32 [nbu]
  DevBuf d_fbia, d_fbja;      // full block graph:  int64 [nb+1], int32 [nbf]
  DevBuf d_rank_fwd, d_rank_bwd; // per upper block entry: rank of (a,b) in full row a, (b,a) in row b
  DevBuf d_lowsrc;               // per lower block (a,b) of the full rows: the upper block (b,a)
  bool lowsrc_ready = false;
  DevBuf d_fa;                // full block CSR values fp64 (K, unscaled), block-major
  DevBuf d_bsr;               // full block CSR fp32 values, scaled D^-1/2 K D^-1/2
  DevBuf d_dinvsqrt;          // fp64 D^-1/2
  DevBuf d_bjinv;             // fp32 inverse diagonal blocks of the scaled matrix
  bool full_valid = false, bsr_valid = false;
  bool scal_valid = false;    // d_dinvsqrt and d_bjinv (also without the d_bsr values)

  // ---- assembly (device) -------------------------------------------------
  DevBuf d_xyz, d_vnodes, d_vconn, d_vcoef, d_bnodes, d_bconn, d_bcid, d_bctype, d_bcval;
  DevBuf d_bccoef, d_bcvec;
  DevBuf d_aidx, d_bumps, d_vmap, d_bmap, d_tvol, d_tbc, d_thp_vol, d_thp_bc;
  DevBuf d_vcolor_list, d_bcolor_list;
  std::vector<int64_t> vcolor_ptr, bcolor_ptr;  // color c: list[ptr[c] .. ptr[c+1])
  bool assembled = false;

  // ---- solver workspace -----------------------------------------------------
  DevBuf d_x, d_r, d_z, d_p, d_p2, d_q, d_u, d_f, d_w;        // fp64
  DevBuf d_xs, d_rs, d_zs, d_ps, d_ps2, d_qs, d_bs;          // fp32
  DevBuf d_scal, d_partials, d_counter;
  DevBuf d_res2;  // |r|^2, |f|^2 of the FP64 true residual
  int reduce_blocks = 1024;
  bool has_solution = false;
  bool f_valid = false;  // d_f holds the load of the current system (set by a solve)

  // matrix-free FP32 inner operator (pzg_matfree.hip)
  bool mf_ready = false, mf_vals_valid = false;
  bool kel_added = false;  // AddKel contributions beyond the mesh's elements
  int64_t mf_nbcd = 0;     // Dirichlet boundary elements
  DevBuf d_mf_blk, d_mf_xe, d_mf_bblk, d_mf_bm, d_mf_ptr, d_mf_pos, d_mf_bpos, d_mf_tab, d_mf_ye;
  int64_t mf_nout = 0;     // element contributions (entries of the row lists)
  int mf_resident = 0;     // workgroups of the element kernel resident at once
  DevBuf d_mf_coef, d_mf_dis, d_mf_mfma, d_mf_geo, d_mf_pt;

  // learned element smoother of the multigrid cycle (pzg_learned.hip):
  // parameters (G, network), Adam moments and gradient; element graph,
  // penalized-equation mask and features; training pairs of the current set
  bool les_ready = false, les_graph = false;
  int les_ne = 0, les_nfaces = 0, les_adam_t = 0;
  int64_t les_nr = 0;
  double les_den = 1.;
  std::vector<int64_t> les_cia;
  std::vector<int32_t> les_cja;
  std::vector<float> les_pen, les_penfrac;
  std::vector<double> les_A, les_B;  // least-squares sums of the G fit
  int64_t les_fit_rows = 0;
  bool les_fit_stale = false;
  DevBuf d_les_th, d_les_m, d_les_v, d_les_g;
  DevBuf d_les_cia, d_les_cja, d_les_pen, d_les_X, d_les_z, d_les_s, d_les_ye;
  DevBuf d_les_Se, d_les_M, d_les_R, d_les_E, d_les_D, d_les_le, d_les_ds, d_les_loss, d_les_fail;
  DevBuf d_nn_trH, d_nn_trM, d_nn_trP, d_nn_dH, d_nn_G, d_nn_tmp, d_nn_dZ, d_nn_part;

  // multigrid preconditioner (pzg_mg.hip): lattice levels 1.. (level 0 is the
  // FCG system), vertex block of each lattice node of level 1, D^1/2 at the
  // vertex equations, the dense inverse of the coarsest level
  bool mg_ready = false, mg_use_mf = false;
  bool mg_les = false;     // level 0 smoothed with the learned element smoother
  DevBuf d_bjinv64;        // FP64 inverses of the connect blocks (PCG64 with block Jacobi)
  // tiled learned smoother on hexahedral lattice meshes (pzg_learned.hip)
  bool les_tile_ready = false, les_tile_ok = false;
  int les_tP = 0, les_tn[3] = {0, 0, 0};
  int16_t les_toff[27] = {}, les_tinv[27] = {};
  int64_t les_nhalo = 0;
  DevBuf d_les_cmap, d_les_tmapb, d_les_tmap, d_les_part, d_les_hmapb, d_les_hmap, d_les_hptr, d_les_hsrc,
      d_les_tscale;
  DevBuf d_les_blkm;       // element blocks with penalized-component bits (learned apply)
  DevBuf d_les_eoff;       // row-ordered slot -> element-major output offset (learned apply)
  // penalty rows of the learned path, solved in FP64 by the outer refinement
  // (pzg_penalty.hip): K_pp CSR over the penalized equations, the Schwarz
  // sets of the penalized boundary elements with their inverses, the row ->
  // slot gather, CG work vectors and scalars
  bool pen_ready = false, pen_struct = false;  // values of this assembly; structure of this mesh
  // connect blocks of high-order meshes (consecutive equations of one real
  // connect, at most VBJ_MAX): starts, sizes, inverse offsets, equation -> block
  int64_t vbj_nblk = 0, vbj_ninv = 0;
  DevBuf d_vbj_pos, d_vbj_m, d_vbj_off, d_vbj_eqb, d_vbjinv64, d_vbjinv32;
  int64_t pen_n = 0, pen_nnz = 0, pen_nset = 0, pen_nslot = 0, pen_narow = 0;
  DevBuf d_pen_arow, d_pen_delta;
  DevBuf d_pen_idx, d_pen_pmap, d_pen_ia, d_pen_ja, d_pen_a, d_pen_sptr, d_pen_sidx, d_pen_sset, d_pen_soff,
      d_pen_sval, d_pen_sinv, d_pen_gptr, d_pen_gslot, d_pen_w, d_pen_scal;
  hipGraphExec_t pen_graph = nullptr;  // one batch of its CG iterations
  bool mg_struct = false;  // lattice, levels and buffers of the current mesh
  bool mg_dense = true;    // the coarsest level has a dense inverse (else a Chebyshev solve)
  int mg_degree = 2;       // Chebyshev degree of the pre- and post-smoothing
  double mg_range = 30.;   // ratio of the smoothed interval's ends
  double mg_lmax0 = 0.;    // largest eigenvalue of the block-Jacobi scaled S
  int64_t mg_ncoarse = 0;  // unknowns of the coarsest level
  std::vector<MgLevel> mg;
  DevBuf d_mg_vmap, d_mg_vinv, d_mg_dsq, d_mg_cinv, d_mg_r0, d_mg_d0, d_mg_t0, d_mg_w0, d_mg_lz, d_mg_y0;

  // post-processing (pzg_post.hip): error-rule tables per rule bucket for the
  // exact-solution order post_porder (-1: not built), the elements of every
  // bucket, and per element the offset of its first point
  int post_porder = -1;
  int64_t post_npts = 0;
  std::vector<PostTables> post_tab;
  std::vector<int64_t> post_tab_off, post_lptr;
  DevBuf d_post_tab, d_post_list, d_post_slot, d_post_off, d_post_out, d_post_uq, d_post_duq,
      d_post_xq;
};

// constrained elements: at most CEL_NC unconstrained connects in the list
constexpr int CEL_NC = 64;

// ---- element tables (pzg_shape.cpp) ----------------------------------------
int build_tables(int etype, int p, int rule_order, ElemTables &t);
// high-order simplex tables of order P at the rule of rule_order
int build_hp_tables(int etype, int P, int rule_order, HpTables &t);
// canonical shape list of a simplex of order P: real side and local index
int hp_shape_list(int etype, int P, int *side, int *loc);
// NConnectShapeF of side s of a triangle / tetrahedron at order p
int simplex_nshape(int etype, int s, int p);
int etype_ncorner(int t);
int etype_nsides(int t);
int gauss_legendre(int order, double *loc, double *w);
// error integration rule order and its tables (post-processing)
int post_rule_order(int etype, int rule, int porder);
int build_post_tables(int etype, int p, bool hp, int rule_order, PostTables &t);

// ---- kernels launchers (pzg_assemble.hip / pzg_solve.hip) --------------------
int launch_expand_upper_ja(Ctx &c);
int launch_elem_maps(Ctx &c);
int launch_assemble(Ctx &c);
int launch_diag(Ctx &c);
// decouples the inactive equations of the equation filter (mesh numbering)
int launch_filter_mask(Ctx &c);
int launch_expand_full(Ctx &c);
// equation-level full CSR (device arrays, any may be null; ia without ia[neq])
int launch_full_csr(Ctx &c, int64_t *d_ia, int64_t *d_ja, double *d_a);
// D^-1/2 and the block-Jacobi inverses of S; values: also the FP32 block CSR of S
int launch_build_bsr(Ctx &c, bool values = true);
int launch_spmv64(Ctx &c, const double *x, double *y);
int launch_spmv_bsr(Ctx &c, const float *x, float *y);
int launch_multadd_upper(Ctx &c, const double *x, const double *y, double *z, double alpha,
                         double beta);
int launch_addkel_batch(Ctx &c, int64_t nel, int ne, const int64_t *d_dest, const double *d_ek,
                        const int32_t *d_list, int64_t nlist);
int side_corner_list_pub(int etype, int s, int *loc);
// learned element smoother (pzg_learned.hip)
int les_nparams(int ne);
int les_setup(Ctx &c, uint64_t seed);
int les_prepare(Ctx &c);  // per-element scales of the current assembly, apply buffers
// Chebyshev step fused into the learned smoother's gather (d, x, pt, dis:
// the smoothing's direction, iterate and pre-scaled direction)
struct LesCheb {
  int mode;
  float c1, c2;
  float *d, *x, *pt;
  const float *dis;
};
// z = M r (learned smoother); with ch, the Chebyshev step with y = M r instead
int les_apply(Ctx &c, const float *r, float *z, const SolveScalars *sc, const LesCheb *ch = nullptr);
int les_collect(Ctx &c, int nrhs, uint64_t seed);
int les_loss_grad(Ctx &c, double *loss_out);
int les_adam_step(Ctx &c, double lr);
int les_fit_G(Ctx &c);  // G from the accumulated least-squares sums
// penalty rows of the learned path (pzg_penalty.hip)
int pen_setup(Ctx &c);
int pen_solve(Ctx &c, double *r64, double *u64, double stop_abs, bool use_graph, int64_t *its, double *rr0,
              double *rr1, double frac = 0., double tot2 = 0., double *rr_full = nullptr);
int pen_zero32(Ctx &c, float *r32);
void pen_free(Ctx &c);
int mf_setup(Ctx &c);
// multigrid preconditioner: levels, operators and eigenvalue estimates of the
// current system (use_mf: level 0 applied matrix-free); z = V-cycle(r), then
// p = z and the z.q partials of the FCG update (when p / q are given)
int mg_setup(Ctx &c, bool use_mf);
int mg_apply(Ctx &c, const float *r, float *z, float *p, const float *q, SolveScalars *sc,
             double *partials);
void mg_free(Ctx &c);
// reduction partials and device scalars of the solvers
int ensure_reduce_pub(Ctx &c);
int mf_prepare(Ctx &c);
// volume element kernel alone (for timing)
int mf_launch_vol(Ctx &c, const SolveScalars *sc);
// pt = D^-1/2 p (input of the element kernels)
int mf_scale(Ctx &c, const float *p, const SolveScalars *sc);
// q = S p; scaled_ready: pt already holds D^-1/2 p (fused into the p update)
int mf_apply(Ctx &c, const float *p, const float *r, float *q, const SolveScalars *sc,
             double *partials, int *gather_grid, bool scaled_ready = false);
// the matrix-free operator reproduces the assembled K (mesh-built, no AddKel extras)
inline bool mf_available(const Ctx &c) {
  return c.has_mesh && c.has_struct && !c.external_csr && !c.kel_added && c.assembled && !c.has_dep &&
         !c.filt &&
         c.vol_rules.size() == 1 && c.bc_rules.size() <= 1 && c.vol_etype != PZG_TET && c.bc_mf_ok;
}
int solve_pcg64(Ctx &c, const pzg_solve_params &prm, pzg_solve_info &info);
int solve_mixed(Ctx &c, const pzg_solve_params &prm, pzg_solve_info &info);
int true_residual(Ctx &c, const double *u, double *rel);
// allocation-free variant on the per-device static scratch (pzg_solve.hip)
int noalloc_scratch_setup(int device);
int true_residual_noalloc(Ctx &c, double *rel);
// post-processing of the device solution (pzg_post.hip)
int post_points(Ctx &c, int porder, int64_t *npts, double *xq);
int post_errors(Ctx &c, const pzg_exact &ex, double *el_errors, double *errors, int *nerr);
// Sloan renumbering of the connect graph (pzg_renumber.cpp)
int sloan_renumber(int64_t nnodes, int64_t nel, const int64_t *eptr, const int64_t *elist,
                   int64_t *perm);

}  // namespace pzg
