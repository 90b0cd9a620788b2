// pzg_host.cpp -- C ABI of the HIP linear FEM path: context, mesh flattening,
// symbolic phase (pattern + maps + colouring) and the entry points that
// dispatch to the device kernels in pzg_assemble.hip / pzg_solve.hip.
#include <algorithm>
#include <map>
#include <set>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <thread>

#include "pzg_internal.h"

namespace pzg {

static thread_local std::string g_err;
void set_error(const std::string &msg) { g_err = msg; }

int dev_alloc(DevBuf &b, size_t bytes) {
  if (b.p && b.bytes >= bytes && bytes > 0) return PZG_OK;
  dev_free(b);
  if (bytes == 0) return PZG_OK;
  hipError_t e = hipMalloc(&b.p, bytes);
  if (e != hipSuccess) {
    b.p = nullptr;
    set_error("hipMalloc(" + std::to_string(bytes) + " bytes): " + hipGetErrorString(e));
    return PZG_ENOMEM;
  }
  b.bytes = bytes;
  return PZG_OK;
}

void dev_free(DevBuf &b) {
  if (b.p) (void)hipFree(b.p);
  b.p = nullptr;
  b.bytes = 0;
}

static int nthreads() {
  unsigned n = std::thread::hardware_concurrency();
  if (n == 0) n = 4;
  return (int)std::min(16u, n);
}

template <class F> static void parallel_for(int64_t n, F f) {
  int T = nthreads();
  if (n < 4096) T = 1;
  std::vector<std::thread> th;
  for (int t = 0; t < T; t++)
    th.emplace_back([=]() {
      int64_t lo = n * t / T, hi = n * (t + 1) / T;
      for (int64_t i = lo; i < hi; i++) f(i);
    });
  for (auto &x : th) x.join();
}

static void free_all(Ctx &c) {
  DevBuf *bufs[] = {&c.d_ia, &c.d_ja, &c.d_a, &c.d_rhs, &c.d_diag, &c.d_ubia, &c.d_ubja,
                    &c.d_fbia, &c.d_fbja, &c.d_rank_fwd, &c.d_rank_bwd, &c.d_lowsrc,
                    &c.d_fa, &c.d_bsr, &c.d_dinvsqrt, &c.d_bjinv, &c.d_xyz, &c.d_vnodes,
                    &c.d_vconn, &c.d_vcoef, &c.d_bnodes, &c.d_bconn, &c.d_bcid, &c.d_bctype,
                    &c.d_bcval, &c.d_bccoef, &c.d_bcvec, &c.d_aidx, &c.d_bumps, &c.d_vmap, &c.d_bmap, &c.d_tvol,
                    &c.d_tbc, &c.d_vcolor_list, &c.d_vol_R, &c.d_blk_iptr, &c.d_blk_inc, &c.d_blk_kref, &c.d_vbj_pos, &c.d_vbj_m, &c.d_vbj_off, &c.d_vbj_eqb, &c.d_vbjinv64, &c.d_vbjinv32, &c.d_bcolor_list, &c.d_x, &c.d_r, &c.d_z,
                    &c.d_p, &c.d_p2, &c.d_q, &c.d_u, &c.d_f, &c.d_w, &c.d_xs, &c.d_rs,
                    &c.d_zs, &c.d_ps, &c.d_ps2, &c.d_qs, &c.d_bs, &c.d_scal, &c.d_partials,
                    &c.d_counter, &c.d_res2, &c.d_les_th, &c.d_les_m, &c.d_les_v, &c.d_les_g,
                    &c.d_les_cia, &c.d_les_cja, &c.d_les_pen, &c.d_les_X, &c.d_les_z, &c.d_les_s,
                    &c.d_les_ye, &c.d_les_blkm, &c.d_les_eoff, &c.d_bjinv64, &c.d_les_cmap, &c.d_les_tmapb, &c.d_les_tmap, &c.d_les_part, &c.d_les_hmapb, &c.d_les_hmap, &c.d_les_hptr, &c.d_les_hsrc, &c.d_les_tscale, &c.d_les_Se, &c.d_les_M, &c.d_les_R, &c.d_les_E, &c.d_les_D,
                    &c.d_les_le, &c.d_les_ds, &c.d_les_loss, &c.d_les_fail, &c.d_nn_trH, &c.d_nn_trM,
                    &c.d_nn_trP, &c.d_nn_dH, &c.d_nn_G, &c.d_nn_tmp, &c.d_nn_dZ, &c.d_nn_part,
                    &c.d_filt_dest, &c.d_post_tab, &c.d_post_list, &c.d_post_slot,
                    &c.d_post_off, &c.d_post_out, &c.d_post_uq, &c.d_post_duq, &c.d_post_xq,
                    &c.d_hp_vsub, &c.d_hp_bsub, &c.d_thp_vol, &c.d_thp_bc, &c.d_mf_blk,
                    &c.d_mf_xe, &c.d_mf_bblk, &c.d_mf_bm, &c.d_mf_ptr, &c.d_mf_pos, &c.d_mf_bpos,
                    &c.d_mf_tab, &c.d_mf_ye, &c.d_mf_coef, &c.d_mf_dis, &c.d_mf_mfma, &c.d_mf_geo,
                    &c.d_mf_pt, &c.d_aidx_ek, &c.d_dep_ptr, &c.d_dep_blk, &c.d_dep_coef,
                    &c.cv.d_cel, &c.cv.d_cslot, &c.cv.d_T, &c.cv.d_blk, &c.cv.d_n, &c.cv.d_ek,
                    &c.cv.d_ef, &c.cb.d_cel, &c.cb.d_cslot, &c.cb.d_T, &c.cb.d_blk, &c.cb.d_n,
                    &c.cb.d_ek, &c.cb.d_ef};
  for (DevBuf *b : bufs) dev_free(*b);
  mg_free(c);
}

// Greedy first-fit colouring over connects (the idea of
// TPZStructMatrix::ComputeElementColors, StrMatrix/TPZStructMatrix.cpp:104-161):
// elements of one colour share no active connect, so their scatter is race-free.
static int color_elements(int64_t nel, int nsides, const std::vector<int64_t> &conn,
                          const std::vector<int64_t> &aidx, int64_t nb,
                          std::vector<int64_t> &ptr, std::vector<int32_t> &list) {
  std::vector<uint64_t> used(nb, 0);
  std::vector<int> col(nel);
  int ncol = 0;
  for (int64_t e = 0; e < nel; e++) {
    uint64_t m = 0;
    for (int s = 0; s < nsides; s++) {
      int64_t a = aidx[conn[e * nsides + s]];
      if (a >= 0) m |= used[a];
    }
    if (m == ~0ULL) {
      set_error("element colouring needs more than 64 colours");
      return PZG_EUNSUP;
    }
    int k = __builtin_ctzll(~m);
    col[e] = k;
    ncol = std::max(ncol, k + 1);
    for (int s = 0; s < nsides; s++) {
      int64_t a = aidx[conn[e * nsides + s]];
      if (a >= 0) used[a] |= (1ULL << k);
    }
  }
  ptr.assign(ncol + 1, 0);
  for (int64_t e = 0; e < nel; e++) ptr[col[e] + 1]++;
  for (int k = 0; k < ncol; k++) ptr[k + 1] += ptr[k];
  list.resize(nel);
  std::vector<int64_t> fill(ptr.begin(), ptr.end() - 1);
  for (int64_t e = 0; e < nel; e++) list[fill[col[e]]++] = (int32_t)e;
  return PZG_OK;
}

// Splits every colour list into one list per rule bucket (stable): list k of
// the result is colour k / nbk, bucket k % nbk.
static void group_by_bucket(std::vector<int64_t> &ptr, std::vector<int32_t> &list,
                            const std::vector<int32_t> &bucket, int nbk) {
  if (nbk <= 1) return;
  const size_t ncol = ptr.size() - 1;
  std::vector<int64_t> nptr(ncol * nbk + 1, 0);
  std::vector<int32_t> nlist(list.size());
  int64_t o = 0;
  for (size_t k = 0; k < ncol; k++)
    for (int b = 0; b < nbk; b++) {
      nptr[k * nbk + b] = o;
      for (int64_t i = ptr[k]; i < ptr[k + 1]; i++)
        if (bucket[list[i]] == b) nlist[o++] = list[i];
    }
  nptr[ncol * nbk] = o;
  ptr.swap(nptr);
  list.swap(nlist);
}

static int etype_nc(int t) { return etype_ncorner(t); }
static int etype_ns(int t) { return etype_nsides(t); }

// High-order tetrahedra: replace every connect of m shapes by m virtual
// connects of nstate equations (positions pos + k*nstate, so the equation
// numbering is unchanged) and every element's side list by its list of
// shapes of the order-P set, the k-th shape of a side mapped to the side's
// k-th virtual connect, or to an empty connect where the side's order
// carries fewer shapes.  SetupMatrixData on the virtual connects gives the
// same IA/JA: the connects of one side are consecutive, coupled exactly when
// the side is, and listed in ascending order.
static int expand_hp(Ctx &c) {
  const int P = c.porder, ns = c.nstate;
  int vside[HP_MAXSH], vloc[HP_MAXSH], bside[HP_MAXSH], bloc[HP_MAXSH];
  const int nsv = hp_shape_list(PZG_TET, P, vside, vloc);
  const int nsb = hp_shape_list(PZG_TRI, P, bside, bloc);
  std::vector<int64_t> vbase(c.nconn + 1, 0);
  for (int64_t s = 0; s < c.nconn; s++) {
    if (c.blk_size[s] < 0 || c.blk_size[s] % ns) {
      set_error("pzg_mesh_set: block size not a multiple of nstate");
      return PZG_EINVAL;
    }
    vbase[s + 1] = vbase[s] + c.blk_size[s] / ns;
  }
  const int64_t nvirt = vbase[c.nconn], empty = nvirt;
  auto expand = [&](const std::vector<int64_t> &conn, int nsr, int etype, int nsx, const int *side,
                    const int *loc, int64_t nel, std::vector<int64_t> &out) -> int {
    for (int64_t e = 0; e < nel; e++)
      for (int r = 0; r < nsr; r++) {
        const int64_t cn = conn[e * nsr + r];
        if (vbase[cn + 1] - vbase[cn] > simplex_nshape(etype, r, P)) {
          set_error("pzg_mesh_set: a connect carries more shapes than its side allows at order p");
          return PZG_EINVAL;
        }
      }
    out.assign(nel * nsx, empty);
    for (int64_t e = 0; e < nel; e++)
      for (int i = 0; i < nsx; i++) {
        const int64_t cn = conn[e * nsr + side[i]];
        if (loc[i] < vbase[cn + 1] - vbase[cn]) out[e * nsx + i] = vbase[cn] + loc[i];
      }
    return PZG_OK;
  };
  std::vector<int64_t> vconn, bconn;
  PZG_TRY(expand(c.vol_conn, c.vol_nsides, PZG_TET, nsv, vside, vloc, c.nvol, vconn));
  if (c.nbc > 0) PZG_TRY(expand(c.bc_conn, c.bc_nsides, PZG_TRI, nsb, bside, bloc, c.nbc, bconn));
  // the real connects as the blocks of the connect-block Jacobi
  {
    std::vector<int64_t> bp, bo;
    std::vector<int32_t> bmv, eqb(c.neq, -1);
    int64_t off = 0;
    bool ok = true;
    for (int64_t s = 0; s < c.nconn; s++) {
      const int m = c.blk_size[s];
      if (m <= 0) continue;
      if (m > VBJ_MAX || c.blk_pos[s] + m > c.neq) { ok = false; break; }
      for (int k = 0; k < m; k++) eqb[c.blk_pos[s] + k] = (int32_t)bp.size();
      bp.push_back(c.blk_pos[s]);
      bmv.push_back(m);
      bo.push_back(off);
      off += (int64_t)m * m;
    }
    for (int64_t i = 0; ok && i < c.neq; i++) ok = eqb[i] >= 0;
    c.vbj_nblk = 0;
    if (ok && !bp.empty()) {
      hipStream_t st = c.stream;
      PZG_TRY(dev_upload(c.d_vbj_pos, bp.data(), bp.size(), st));
      PZG_TRY(dev_upload(c.d_vbj_m, bmv.data(), bmv.size(), st));
      PZG_TRY(dev_upload(c.d_vbj_off, bo.data(), bo.size(), st));
      PZG_TRY(dev_upload(c.d_vbj_eqb, eqb.data(), eqb.size(), st));
      c.vbj_nblk = (int64_t)bp.size();
      c.vbj_ninv = off;
    }
  }
  std::vector<int64_t> pos(nvirt + 1);
  std::vector<int32_t> size(nvirt + 1, ns);
  for (int64_t s = 0; s < c.nconn; s++)
    for (int64_t k = 0; k < vbase[s + 1] - vbase[s]; k++) pos[vbase[s] + k] = c.blk_pos[s] + k * ns;
  pos[empty] = c.neq;
  size[empty] = 0;
  c.vol_conn.swap(vconn);
  c.bc_conn.swap(bconn);
  c.blk_pos.swap(pos);
  c.blk_size.swap(size);
  c.nconn = nvirt + 1;
  c.vol_nsides = nsv;
  c.bc_nsides = nsb;
  c.vol_side_real.assign(vside, vside + nsv);
  c.hp_simplex = true;
  c.hp_order = P;
  return PZG_OK;
}

// Block assembly of identical affine hexahedra (k_vol_blk): when every
// volume element has the Jacobian of element 0 (within 1e-12), its element
// matrix is coef_e K_ref, K_ref formed here in FP64 exactly as k_vol_aff forms
// an element's (P^ab = |det J| Jinv^T R Jinv, then the material's entries
// with coefficient 1).  Also the element list of every block row (ascending
// element order, packed element << 5 | local shape).  PZG_NO_BLKASM=1 keeps
// the element scatter.
static int blk_setup(Ctx &c, const std::vector<double> &R) {
  const char *off = std::getenv("PZG_NO_BLKASM");
  if (off && off[0] == '1') return PZG_OK;
  const ElemTables &T = c.tvol;
  const int NC = c.vol_ncorner, NSD = c.vol_nsides, NSH = T.nsh, ns = c.nstate, NE = NSH * ns;
  if (c.dim != 3 || NC != 8 || NE > 81 || c.nvol <= 0 || c.nvol >= ((int64_t)1 << 26) || c.has_dep ||
      c.hp_simplex)
    return PZG_OK;
  // Jacobian of each element at the first point (the affine elements' constant one)
  auto jac = [&](int64_t e, double J[3][3]) {
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) {
        double v = 0.;
        for (int k = 0; k < NC; k++) v += c.xyz[3 * c.vol_nodes[e * NC + k] + i] * T.cdphi[0][j][k];
        J[i][j] = v;
      }
  };
  int64_t naff = 0;  // every element affine
  for (int64_t v : c.vcolor_aff) naff += v;
  if (naff != c.nvol) return PZG_OK;
  double J0[3][3], sc = 0.;
  jac(0, J0);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) sc = std::max(sc, std::fabs(J0[i][j]));
  for (int64_t e = 1; e < c.nvol; e++) {
    double J[3][3];
    jac(e, J);
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++)
        if (std::fabs(J[i][j] - J0[i][j]) > 1e-12 * sc) return PZG_OK;
  }
  const double(*j)[3] = J0;
  const double det = -j[0][2] * j[1][1] * j[2][0] + j[0][1] * j[1][2] * j[2][0] + j[0][2] * j[1][0] * j[2][1] -
                     j[0][0] * j[1][2] * j[2][1] - j[0][1] * j[1][0] * j[2][2] + j[0][0] * j[1][1] * j[2][2];
  if (det == 0.) return PZG_OK;
  double inv[3][3];
  inv[0][0] = (-j[1][2] * j[2][1] + j[1][1] * j[2][2]) / det;
  inv[0][1] = (j[0][2] * j[2][1] - j[0][1] * j[2][2]) / det;
  inv[0][2] = (-j[0][2] * j[1][1] + j[0][1] * j[1][2]) / det;
  inv[1][0] = (j[1][2] * j[2][0] - j[1][0] * j[2][2]) / det;
  inv[1][1] = (-j[0][2] * j[2][0] + j[0][0] * j[2][2]) / det;
  inv[1][2] = (j[0][2] * j[1][0] - j[0][0] * j[1][2]) / det;
  inv[2][0] = (-j[1][1] * j[2][0] + j[1][0] * j[2][1]) / det;
  inv[2][1] = (j[0][1] * j[2][0] - j[0][0] * j[2][1]) / det;
  inv[2][2] = (-j[0][1] * j[1][0] + j[0][0] * j[1][1]) / det;
  const double adet = std::fabs(det);
  auto Rv = [&](int m, int n, int i, int jj) {  // R_mn(i, j), stored for m <= n
    const int lo = std::min(m, n), hi = std::max(m, n), r = lo * 3 - lo * (lo - 1) / 2 + (hi - lo);
    return m <= n ? R[((size_t)r * NSH + i) * NSH + jj] : R[((size_t)r * NSH + jj) * NSH + i];
  };
  std::vector<double> K((size_t)NE * NE, 0.);
  const double nu = c.nu, den = -1. + nu + 2. * nu * nu;
  const double C1 = 1. / (2. + 2. * nu), C2 = nu / den, C3 = (nu - 1.) / den;
  for (int i = 0; i < NSH; i++)
    for (int jj = 0; jj < NSH; jj++) {
      double Q[3][3], P[3][3];
      for (int m = 0; m < 3; m++)
        for (int b = 0; b < 3; b++) {
          double v = 0.;
          for (int n = 0; n < 3; n++) v += Rv(m, n, i, jj) * inv[n][b];
          Q[m][b] = v;
        }
      for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
          double v = 0.;
          for (int m = 0; m < 3; m++) v += inv[m][a] * Q[m][b];
          P[a][b] = v * adet;
        }
      if (ns == 3) {
        for (int ci = 0; ci < 3; ci++)
          for (int cj = 0; cj < 3; cj++) {
            double v;
            if (ci == cj) {
              double tr = 0.;
              for (int cc = 0; cc < 3; cc++)
                if (cc != ci) tr += P[cc][cc];
              v = tr * C1 + P[ci][ci] * C3;
            } else {
              v = P[cj][ci] * C1 - P[ci][cj] * C2;
            }
            K[(size_t)(3 * i + ci) * NE + 3 * jj + cj] = v;
          }
      } else {
        K[(size_t)i * NE + jj] = P[0][0] + P[1][1] + P[2][2];
      }
    }
  // element lists of the block rows
  std::vector<int64_t> iptr(c.nb + 1, 0);
  for (int64_t e = 0; e < c.nvol; e++)
    for (int sh = 0; sh < NSH; sh++) {
      const int64_t b = c.aidx[c.vol_conn[e * NSD + T.shape_side[sh]]];
      if (b >= 0) iptr[b + 1]++;
    }
  for (int64_t b = 0; b < c.nb; b++) {
    if (iptr[b + 1] > 8) return PZG_OK;  // a connect of more than 8 elements
    iptr[b + 1] += iptr[b];
  }
  std::vector<int32_t> inc(std::max<int64_t>(iptr[c.nb], 1));
  std::vector<int64_t> fill(iptr.begin(), iptr.end() - 1);
  for (int64_t e = 0; e < c.nvol; e++)
    for (int sh = 0; sh < NSH; sh++) {
      const int64_t b = c.aidx[c.vol_conn[e * NSD + T.shape_side[sh]]];
      if (b >= 0) inc[fill[b]++] = (int32_t)((e << 5) | sh);
    }
  hipStream_t s = c.stream;
  PZG_TRY(dev_upload(c.d_blk_kref, K.data(), K.size(), s));
  PZG_TRY(dev_upload(c.d_blk_iptr, iptr.data(), iptr.size(), s));
  PZG_TRY(dev_upload(c.d_blk_inc, inc.data(), inc.size(), s));
  c.blk_asm = true;
  c.blk_grid = 0;
  return PZG_OK;
}

// Boundary conditions of every bc id as coefficients of the boundary
// element's integrals (BCC layout, pzg_internal.h), from the ContributeBC
// switches of TPZMatPoisson (Material/Poisson/TPZMatPoisson.cpp:45-121),
// TPZDarcyFlow (Material/DarcyFlow/TPZDarcyFlow.cpp:49-91) and
// TPZElasticity3D (Material/Elasticity/TPZElasticity3D.cpp:616-773); big is
// the material's penalty.  Types outside a switch are refused.  For the
// stress-field type the vector from the volume neighbour's centre to the
// boundary element's centre (TPZInterpolationSpace::ComputeNormal,
// Mesh/pzinterpolationspace.cpp:301-383) is kept per boundary element.
static int bc_setup(Ctx &c, double big) {
  c.bc_coef.assign((size_t)BCC * std::max(0, c.nbcid), 0.);
  c.bc_mf_ok = true;
  bool any_stress = false;
  std::vector<char> used(std::max(0, c.nbcid), 0);
  for (int64_t e = 0; e < c.nbc; e++) used[c.bc_id[e]] = 1;
  for (int id = 0; id < c.nbcid; id++) {
    const int type = c.bc_type[id];
    const double *v2 = &c.bc_val[3 * id], *v1 = &c.bc_val1[9 * id];
    double *cf = &c.bc_coef[(size_t)BCC * id];
    double *ck = cf + BCC_CK, *cl = cf + BCC_CF;
    bool ok = true, penalty_only = true;
    if (c.material == PZG_MAT_ELAST3D) {
      int pen[3] = {0, 0, 0};
      switch (type) {
        case PZG_BC_DIRICHLET: pen[0] = pen[1] = pen[2] = 1; break;
        case PZG_BC_DIR_X: pen[0] = 1; break;
        case PZG_BC_DIR_Y: pen[1] = 1; break;
        case PZG_BC_DIR_Z: pen[2] = 1; break;
        case PZG_BC_DIR_XZ: pen[0] = pen[2] = 1; break;
        case PZG_BC_NEUMANN: case PZG_BC_MIXED: case PZG_BC_DIR_NULL: case PZG_BC_STRESS: break;
        default: ok = false;
      }
      for (int k = 0; k < 3; k++) {
        if (pen[k]) { ck[4 * k] = big; cl[k] = big * v2[k]; }
        else if (type == PZG_BC_NEUMANN || type == PZG_BC_MIXED) cl[k] = v2[k];
        if (type == PZG_BC_DIR_NULL) ck[4 * k] = big * v2[k];
      }
      if (type == PZG_BC_MIXED)
        for (int k = 0; k < 9; k++) ck[k] = v1[k];
      if (type == PZG_BC_STRESS) {
        cf[BCC_STRESS] = 1.;
        for (int k = 0; k < 9; k++) cf[BCC_VAL1 + k] = v1[k];
        any_stress |= used[id] != 0;
      }
      penalty_only = type == PZG_BC_DIRICHLET || type == PZG_BC_NEUMANN || type == PZG_BC_STRESS;
    } else {
      switch (type) {
        case PZG_BC_DIRICHLET: ck[0] = big; cl[0] = big * v2[0]; break;
        case PZG_BC_NEUMANN:
          cl[0] = v2[0] * (c.material == PZG_MAT_POISSON ? c.scale : 1.0);
          break;
        case PZG_BC_MIXED:
          if (c.material == PZG_MAT_POISSON) { cf[BCC_CD] = big * v1[0]; cl[0] = big * v2[0]; }
          else { ck[0] = v1[0]; cl[0] = v2[0]; }
          break;
        default: ok = false;
      }
      penalty_only = type == PZG_BC_DIRICHLET || type == PZG_BC_NEUMANN;
    }
    if (!ok && used[id]) {  // ids without boundary elements are never read
      set_error("pzg_mesh_set: boundary condition type " + std::to_string(type) + " of bc id " +
                std::to_string(id) + " is not one of the material's ContributeBC types");
      return PZG_EUNSUP;
    }
    if (used[id] && !penalty_only) c.bc_mf_ok = false;
    if (used[id] && c.material == PZG_MAT_POISSON && type == PZG_BC_MIXED && c.hp_simplex) {
      set_error("pzg_mesh_set: TPZMatPoisson type 2 on high-order tetrahedral meshes is not "
                "implemented");
      return PZG_EUNSUP;
    }
  }
  c.bc_vec.clear();
  if (!any_stress) return PZG_OK;
  if (c.has_dep) {
    set_error("pzg_mesh_set: stress-field boundary conditions on constrained meshes are not "
              "implemented");
    return PZG_EUNSUP;
  }
  // volume neighbour of every stress-field boundary element: the volume
  // element with a side whose corners are the boundary element's corners
  // (the element type's own sides: on high-order meshes vol_nsides counts
  // the virtual sides)
  const int vt = c.vol_etype, ncb = c.bc_ncorner, ncv = c.vol_ncorner;
  std::vector<std::vector<int>> faces;
  for (int sd = 0; sd < etype_ns(vt); sd++) {
    int loc[8];
    const int n = side_corner_list_pub(vt, sd, loc);
    if (n == ncb) faces.emplace_back(loc, loc + n);
  }
  auto sorted = [](std::vector<int64_t> v) {
    std::sort(v.begin(), v.end());
    return v;
  };
  std::vector<std::vector<int64_t>> want(c.nbc);
  std::map<std::vector<int64_t>, std::pair<int64_t, int>> owner;  // corners -> (volume, count)
  for (int64_t e = 0; e < c.nbc; e++) {
    if (c.bc_coef[(size_t)BCC * c.bc_id[e] + BCC_STRESS] == 0.) continue;
    want[e] = sorted(std::vector<int64_t>(c.bc_nodes.begin() + e * ncb, c.bc_nodes.begin() + (e + 1) * ncb));
    owner[want[e]] = {-1, 0};
  }
  for (int64_t v = 0; v < c.nvol; v++)
    for (const auto &f : faces) {
      std::vector<int64_t> k(f.size());
      for (size_t i = 0; i < f.size(); i++) k[i] = c.vol_nodes[v * ncv + f[i]];
      auto it = owner.find(sorted(k));
      if (it == owner.end()) continue;
      if (it->second.second++ == 0) it->second.first = v;
    }
  c.bc_vec.assign(3 * c.nbc, 0.);
  for (int64_t e = 0; e < c.nbc; e++) {
    if (want[e].empty()) continue;
    const auto &o = owner[want[e]];
    if (o.second != 1) {
      set_error("pzg_mesh_set: a stress-field boundary element needs exactly one volume neighbour");
      return PZG_EUNSUP;
    }
    for (int d = 0; d < 3; d++) {
      double xb = 0., xv = 0.;
      for (int k = 0; k < ncb; k++) xb += c.xyz[3 * c.bc_nodes[e * ncb + k] + d] / ncb;
      for (int k = 0; k < ncv; k++) xv += c.xyz[3 * c.vol_nodes[o.first * ncv + k] + d] / ncv;
      c.bc_vec[3 * e + d] = -xv + xb;
    }
  }
  return PZG_OK;
}

static int mesh_set(Ctx &c, const pzg_mesh_desc *m) {
  const int pmax = (m && m->vol_etype == PZG_TET) ? HP_MAXP : 2;
  if (!m || (m->dim != 2 && m->dim != 3) || m->porder < 1 || m->porder > pmax ||
      (m->nstate != 1 && m->nstate != 3) || m->material < 0 || m->material > 2 ||
      (m->material == PZG_MAT_ELAST3D && (m->nstate != 3 || m->dim != 3)) ||
      (m->material != PZG_MAT_ELAST3D && m->nstate != 1) || m->nvol <= 0 || !m->xyz ||
      !m->vol_nodes || !m->vol_conn || !m->blk_pos || !m->blk_size || m->nconn <= 0) {
    set_error("pzg_mesh_set: invalid or unsupported mesh description");
    return PZG_EINVAL;
  }
  if (m->nbc > 0 && (!m->bc_nodes || !m->bc_conn || !m->bc_id || !m->bc_type || !m->bc_val)) {
    set_error("pzg_mesh_set: boundary element arrays missing");
    return PZG_EINVAL;
  }
  c.dim = m->dim;
  c.material = m->material;
  c.nstate = m->nstate;
  c.porder = m->porder;
  c.nnodes = m->nnodes;
  c.nvol = m->nvol;
  c.nbc = m->nbc;
  c.nconn = m->nconn;
  c.neq = m->neq;
  if (m->vol_etype != 0 && !(m->vol_etype == PZG_TET && m->dim == 3)) {
    set_error("pzg_mesh_set: unsupported volume element type");
    return PZG_EUNSUP;
  }
  int vt = m->dim == 2 ? PZG_QUAD : PZG_HEX, bt = m->dim == 2 ? PZG_LINE : PZG_QUAD;
  if (m->vol_etype == PZG_TET) {
    vt = PZG_TET;
    bt = PZG_TRI;
  }
  c.vol_etype = vt;
  c.bc_etype = bt;
  c.vol_ncorner = etype_nc(vt);
  c.vol_nsides = etype_ns(vt);
  c.bc_ncorner = etype_nc(bt);
  c.bc_nsides = etype_ns(bt);
  c.xyz.assign(m->xyz, m->xyz + 3 * m->nnodes);
  c.vol_nodes.assign(m->vol_nodes, m->vol_nodes + m->nvol * c.vol_ncorner);
  c.vol_conn.assign(m->vol_conn, m->vol_conn + m->nvol * c.vol_nsides);
  c.vol_coef.clear();
  if (m->vol_coef) c.vol_coef.assign(m->vol_coef, m->vol_coef + m->nvol);
  c.bc_nodes.clear();
  c.bc_conn.clear();
  c.bc_id.clear();
  if (m->nbc > 0) {
    c.bc_nodes.assign(m->bc_nodes, m->bc_nodes + m->nbc * c.bc_ncorner);
    c.bc_conn.assign(m->bc_conn, m->bc_conn + m->nbc * c.bc_nsides);
    c.bc_id.assign(m->bc_id, m->bc_id + m->nbc);
  }
  c.nbcid = m->nbcid;
  c.bc_type.assign(m->bc_type, m->bc_type + std::max(0, m->nbcid));
  c.bc_val.assign(m->bc_val, m->bc_val + 3 * std::max(0, m->nbcid));
  c.bc_val1.assign(9 * std::max(0, m->nbcid), 0.);
  if (m->bc_val1) c.bc_val1.assign(m->bc_val1, m->bc_val1 + 9 * std::max(0, m->nbcid));
  c.blk_pos.assign(m->blk_pos, m->blk_pos + m->nconn);
  c.blk_size.assign(m->blk_size, m->blk_size + m->nconn);
  c.scale = m->scale;
  c.E = m->E;
  c.nu = m->nu;
  for (int k = 0; k < 3; k++) c.prestress[k] = m->material == PZG_MAT_ELAST3D ? m->prestress[k] : 0.;
  c.big = m->big_number;
  c.rule_vol = m->rule_order_vol;
  c.rule_bc = m->rule_order_bc;
  c.force_kind = m->force_kind;
  for (int k = 0; k < 3; k++) c.force_const[k] = m->force_const[k];
  c.nbumps = m->force_kind == PZG_FORCE_BUMPS ? m->nbumps : 0;
  c.bumps.clear();
  if (c.nbumps > 0) c.bumps.assign(m->bumps, m->bumps + 6 * c.nbumps);
  c.bump_width = m->bump_width;

  for (int64_t v : c.vol_conn)
    if (v < 0 || v >= c.nconn) { set_error("pzg_mesh_set: connect index out of range"); return PZG_EINVAL; }
  for (int64_t v : c.bc_conn)
    if (v < 0 || v >= c.nconn) { set_error("pzg_mesh_set: connect index out of range"); return PZG_EINVAL; }
  c.hp_simplex = false;
  c.hp_tuned = false;
  c.hp_order = 0;
  c.nconn_in = c.nconn;
  c.vol_side_real.clear();
  c.thp_vol.clear();
  c.thp_bc.clear();
  {
    bool multi = false;
    for (int64_t s = 0; s < c.nconn; s++) multi |= c.blk_size[s] > c.nstate;
    if (multi || (vt == PZG_TET && c.porder > 2)) {
      if (vt != PZG_TET) {
        set_error("pzg_mesh_set: connects with several shapes are supported on tetrahedra only");
        return PZG_EUNSUP;
      }
      PZG_TRY(expand_hp(c));
    }
  }

  // constrained connects (hanging sides)
  c.has_dep = false;
  c.dep_ptr.clear();
  c.dep_conn.clear();
  c.dep_coef.clear();
  c.is_dep.assign(c.nconn, 0);
  if (m->dep_ptr) {
    if (!m->dep_conn || !m->dep_coef || m->dep_ptr[0] != 0) {
      set_error("pzg_mesh_set: incomplete constraint arrays");
      return PZG_EINVAL;
    }
    c.dep_ptr.assign(m->dep_ptr, m->dep_ptr + c.nconn + 1);
    for (int64_t s = 0; s < c.nconn; s++)
      if (c.dep_ptr[s + 1] < c.dep_ptr[s]) {
        set_error("pzg_mesh_set: dep_ptr is not monotone");
        return PZG_EINVAL;
      }
    c.dep_conn.assign(m->dep_conn, m->dep_conn + c.dep_ptr[c.nconn]);
    c.dep_coef.assign(m->dep_coef, m->dep_coef + c.dep_ptr[c.nconn]);
    for (int64_t s = 0; s < c.nconn; s++) c.is_dep[s] = c.dep_ptr[s + 1] > c.dep_ptr[s];
    for (int64_t s = 0; s < c.nconn; s++)
      for (int64_t k = c.dep_ptr[s]; k < c.dep_ptr[s + 1]; k++)
        if (c.dep_conn[k] < 0 || c.dep_conn[k] >= c.nconn || c.is_dep[c.dep_conn[k]]) {
          set_error("pzg_mesh_set: a connect depends on a missing or constrained connect");
          return PZG_EINVAL;
        }
    c.has_dep = c.dep_ptr[c.nconn] > 0;
    if (c.has_dep && (c.vol_etype == PZG_TET || c.hp_simplex)) {
      set_error("pzg_mesh_set: constrained connects on quadrilateral / hexahedral meshes only");
      return PZG_EUNSUP;
    }
  }

  // Uniform active block size: every (virtual) connect carries 0 or nstate
  // equations, block positions must be consecutive in seq order; the blocks
  // of constrained connects carry no equations and come last.
  c.bs = c.nstate;
  c.aidx.assign(c.nconn, -1);
  int64_t nb = 0, pos = 0;
  bool seen_dep = false;
  for (int64_t s = 0; s < c.nconn; s++) {
    int sz = c.blk_size[s];
    if (sz != 0 && sz != c.bs) {
      set_error("pzg_mesh_set: block sizes other than 0 or nstate are not supported");
      return PZG_EUNSUP;
    }
    if (c.is_dep[s]) {
      seen_dep |= sz > 0;
      continue;
    }
    if (sz && seen_dep) {
      set_error("pzg_mesh_set: constrained connects must come after the others in sequence order");
      return PZG_EINVAL;
    }
    if (c.blk_pos[s] != pos) {
      set_error("pzg_mesh_set: Block() positions are not cumulative in sequence order");
      return PZG_EINVAL;
    }
    if (sz) c.aidx[s] = nb++;
    pos += sz;
  }
  if (pos != c.neq) {
    set_error("pzg_mesh_set: neq does not match the block sizes");
    return PZG_EINVAL;
  }
  c.nb = nb;
  for (int64_t v : c.vol_conn)
    if (v < 0 || v >= c.nconn) { set_error("pzg_mesh_set: connect index out of range"); return PZG_EINVAL; }
  for (int64_t v : c.bc_conn)
    if (v < 0 || v >= c.nconn) { set_error("pzg_mesh_set: connect index out of range"); return PZG_EINVAL; }
  for (int64_t v : c.vol_nodes)
    if (v < 0 || v >= c.nnodes) { set_error("pzg_mesh_set: node index out of range"); return PZG_EINVAL; }
  for (int64_t v : c.bc_nodes)
    if (v < 0 || v >= c.nnodes) { set_error("pzg_mesh_set: node index out of range"); return PZG_EINVAL; }
  for (int32_t v : c.bc_id)
    if (v < 0 || v >= c.nbcid) { set_error("pzg_mesh_set: bc id out of range"); return PZG_EINVAL; }
  PZG_TRY(bc_setup(c, m->big_number));

  // integration-rule buckets (hp meshes) and their tables
  auto bucket = [&](const int32_t *rules, int64_t n, int dflt, int etype, std::vector<int> &uniq,
                    std::vector<int32_t> &bk, std::vector<ElemTables> &tabs) -> int {
    uniq.clear();
    bk.assign(n, 0);
    if (!rules) uniq.push_back(dflt);
    for (int64_t e = 0; rules && e < n; e++) {
      const int r = rules[e];
      if (r < 0 || r > 64) {
        set_error("pzg_mesh_set: integration rule order out of range");
        return PZG_EINVAL;
      }
      size_t k = 0;
      while (k < uniq.size() && uniq[k] != r) k++;
      if (k == uniq.size()) uniq.push_back(r);
      bk[e] = (int32_t)k;
    }
    if (uniq.empty()) uniq.push_back(dflt);
    tabs.assign(uniq.size(), ElemTables());
    if (c.hp_simplex) {
      std::vector<HpTables> &ht = etype == PZG_TET ? c.thp_vol : c.thp_bc;
      ht.resize(uniq.size());
      for (size_t k = 0; k < uniq.size(); k++)
        PZG_TRY(build_hp_tables(etype, c.hp_order, uniq[k], ht[k]));
      return PZG_OK;
    }
    for (size_t k = 0; k < uniq.size(); k++) PZG_TRY(build_tables(etype, c.porder, uniq[k], tabs[k]));
    return PZG_OK;
  };
  PZG_TRY(bucket(m->vol_rule, c.nvol, c.rule_vol, vt, c.vol_rules, c.vol_bucket, c.tvol_all));
  PZG_TRY(bucket(c.nbc > 0 ? m->bc_rule : nullptr, c.nbc, c.rule_bc, bt, c.bc_rules, c.bc_bucket,
                 c.tbc_all));
  c.tvol = c.tvol_all[0];
  c.tbc = c.tbc_all[0];
  // high-order buckets: the highest order any element of the bucket uses
  // (from the shapes its connects carry) and the matching sub-list
  std::vector<int32_t> vsub, bsub;
  if (c.hp_simplex) {
    auto subs = [&](int etype, int nsx, const std::vector<int64_t> &conn, int64_t nel,
                    const std::vector<int32_t> &bk, size_t nbk, std::vector<int> &Pb,
                    std::vector<int32_t> &flat) {
      int side[HP_MAXSH], loc[HP_MAXSH];
      hp_shape_list(etype, c.hp_order, side, loc);
      int need[HP_MAXSH];
      for (int i = 0; i < nsx; i++) {
        int o = 1;
        while (o < c.hp_order && simplex_nshape(etype, side[i], o) <= loc[i]) o++;
        need[i] = o;
      }
      Pb.assign(nbk, 1);
      for (int64_t e = 0; e < nel; e++)
        for (int i = 0; i < nsx; i++)
          if (c.blk_size[conn[e * nsx + i]] > 0) Pb[bk[e]] = std::max(Pb[bk[e]], need[i]);
      flat.assign(nbk * HP_MAXSH, 0);
      for (size_t b = 0; b < nbk; b++) {
        int n = 0;
        for (int i = 0; i < nsx; i++)
          if (loc[i] < simplex_nshape(etype, side[i], Pb[b])) flat[b * HP_MAXSH + n++] = i;
      }
    };
    subs(PZG_TET, c.vol_nsides, c.vol_conn, c.nvol, c.vol_bucket, c.vol_rules.size(), c.hp_vol_P, vsub);
    subs(PZG_TRI, c.bc_nsides, c.bc_conn, c.nbc, c.bc_bucket, c.bc_rules.size(), c.hp_bc_P, bsub);
  }

  // constrained elements (hanging sides): the volume and boundary elements
  // with a constrained connect that has shapes, grouped by rule bucket; per
  // element the unconstrained blocks of its connects and of the connects they
  // depend on (TPZConnect::BuildConnectList, sorted by connect) and T
  std::vector<int64_t> aidx_ek(c.nconn), dep_blk;
  for (int64_t s = 0; s < c.nconn; s++)
    aidx_ek[s] = c.aidx[s] >= 0 ? c.aidx[s] : (c.is_dep[s] && c.blk_size[s] > 0 ? -2 : -1);
  std::vector<double> cvT, cbT;
  std::vector<int64_t> cvB, cbB;
  std::vector<int32_t> cvN, cbN;
  auto build_cel = [&](CelSet &cs, int64_t nel, int nsx, const std::vector<int64_t> &conn,
                       const std::vector<int32_t> &bucket, size_t nbk, const ElemTables &tab,
                       std::vector<double> &T, std::vector<int64_t> &B, std::vector<int32_t> &N) -> int {
    cs.cel.clear();
    cs.is_cel.assign(nel, 0);
    cs.xptr.assign(nel + 1, 0);
    cs.xconn.clear();
    cs.bptr.assign(nbk + 1, 0);
    if (!c.has_dep) return PZG_OK;
    for (int64_t e = 0; e < nel; e++) {
      for (int k = 0; k < nsx; k++) {
        const int64_t cc = conn[e * nsx + k];
        if (!c.is_dep[cc] || c.blk_size[cc] == 0) continue;
        cs.is_cel[e] = 1;
        for (int64_t d = c.dep_ptr[cc]; d < c.dep_ptr[cc + 1]; d++) cs.xconn.push_back(c.dep_conn[d]);
      }
      cs.xptr[e + 1] = (int64_t)cs.xconn.size();
    }
    for (size_t b = 0; b < nbk; b++) {
      for (int64_t e = 0; e < nel; e++)
        if (cs.is_cel[e] && bucket[e] == (int32_t)b) cs.cel.push_back((int32_t)e);
      cs.bptr[b + 1] = (int64_t)cs.cel.size();
    }
    for (int32_t e : cs.cel) {
      std::vector<int64_t> L;
      for (int k = 0; k < nsx; k++) L.push_back(conn[(int64_t)e * nsx + k]);
      for (int64_t x = cs.xptr[e]; x < cs.xptr[e + 1]; x++) L.push_back(cs.xconn[x]);
      std::sort(L.begin(), L.end());
      L.erase(std::unique(L.begin(), L.end()), L.end());
      std::vector<int64_t> Bl, Lc;
      for (int64_t cc : L)
        if (c.aidx[cc] >= 0) {
          Bl.push_back(c.aidx[cc]);
          Lc.push_back(cc);
        }
      if ((int)Bl.size() > CEL_NC) {
        set_error("pzg_mesh_set: a constrained element couples too many connects");
        return PZG_EUNSUP;
      }
      auto slot = [&](int64_t cc) {
        return (int)(std::lower_bound(Lc.begin(), Lc.end(), cc) - Lc.begin());
      };
      const size_t t0 = T.size();
      T.resize(t0 + (size_t)nsx * CEL_NC, 0.);
      for (int i = 0; i < tab.nsh; i++) {  // p <= 2: shape i lives on side shape_side[i]
        const int64_t cc = conn[(int64_t)e * nsx + tab.shape_side[i]];
        if (c.blk_size[cc] == 0) continue;
        if (c.aidx[cc] >= 0) {
          T[t0 + (size_t)i * CEL_NC + slot(cc)] = 1.;
          continue;
        }
        for (int64_t d = c.dep_ptr[cc]; d < c.dep_ptr[cc + 1]; d++)
          if (c.aidx[c.dep_conn[d]] >= 0) T[t0 + (size_t)i * CEL_NC + slot(c.dep_conn[d])] += c.dep_coef[d];
      }
      for (int a = 0; a < CEL_NC; a++) B.push_back(a < (int)Bl.size() ? Bl[a] : -1);
      N.push_back((int32_t)Bl.size());
    }
    return PZG_OK;
  };
  PZG_TRY(build_cel(c.cv, c.nvol, c.vol_nsides, c.vol_conn, c.vol_bucket, c.vol_rules.size(),
                    c.tvol_all[0], cvT, cvB, cvN));
  PZG_TRY(build_cel(c.cb, c.nbc, c.bc_nsides, c.bc_conn, c.bc_bucket, c.bc_rules.size(),
                    c.tbc_all[0], cbT, cbB, cbN));
  if (c.has_dep) {
    dep_blk.resize(c.dep_conn.size());
    for (size_t k = 0; k < c.dep_conn.size(); k++) dep_blk[k] = c.aidx[c.dep_conn[k]];
  }

  hipStream_t s = c.stream;
  PZG_TRY(dev_upload(c.d_xyz, c.xyz.data(), c.xyz.size(), s));
  PZG_TRY(dev_upload(c.d_aidx_ek, aidx_ek.data(), aidx_ek.size(), s));
  if (c.has_dep) {
    PZG_TRY(dev_upload(c.cv.d_cel, c.cv.cel.data(), c.cv.cel.size(), s));
    PZG_TRY(dev_upload(c.cv.d_T, cvT.data(), cvT.size(), s));
    PZG_TRY(dev_upload(c.cv.d_blk, cvB.data(), cvB.size(), s));
    PZG_TRY(dev_upload(c.cv.d_n, cvN.data(), cvN.size(), s));
    PZG_TRY(dev_upload(c.cb.d_cel, c.cb.cel.data(), c.cb.cel.size(), s));
    PZG_TRY(dev_upload(c.cb.d_T, cbT.data(), cbT.size(), s));
    PZG_TRY(dev_upload(c.cb.d_blk, cbB.data(), cbB.size(), s));
    PZG_TRY(dev_upload(c.cb.d_n, cbN.data(), cbN.size(), s));
    PZG_TRY(dev_upload(c.d_dep_ptr, c.dep_ptr.data(), c.dep_ptr.size(), s));
    PZG_TRY(dev_upload(c.d_dep_blk, dep_blk.data(), dep_blk.size(), s));
    PZG_TRY(dev_upload(c.d_dep_coef, c.dep_coef.data(), c.dep_coef.size(), s));
  }
  PZG_TRY(dev_upload(c.d_vnodes, c.vol_nodes.data(), c.vol_nodes.size(), s));
  PZG_TRY(dev_upload(c.d_vconn, c.vol_conn.data(), c.vol_conn.size(), s));
  if (!c.vol_coef.empty()) PZG_TRY(dev_upload(c.d_vcoef, c.vol_coef.data(), c.vol_coef.size(), s));
  else dev_free(c.d_vcoef);
  if (c.nbc > 0) {
    PZG_TRY(dev_upload(c.d_bnodes, c.bc_nodes.data(), c.bc_nodes.size(), s));
    PZG_TRY(dev_upload(c.d_bconn, c.bc_conn.data(), c.bc_conn.size(), s));
    PZG_TRY(dev_upload(c.d_bcid, c.bc_id.data(), c.bc_id.size(), s));
  }
  if (c.nbcid > 0) {
    PZG_TRY(dev_upload(c.d_bctype, c.bc_type.data(), c.bc_type.size(), s));
    PZG_TRY(dev_upload(c.d_bcval, c.bc_val.data(), c.bc_val.size(), s));
    PZG_TRY(dev_upload(c.d_bccoef, c.bc_coef.data(), c.bc_coef.size(), s));
  }
  if (!c.bc_vec.empty()) PZG_TRY(dev_upload(c.d_bcvec, c.bc_vec.data(), c.bc_vec.size(), s));
  else dev_free(c.d_bcvec);
  if (c.nbumps > 0) PZG_TRY(dev_upload(c.d_bumps, c.bumps.data(), c.bumps.size(), s));
  PZG_TRY(dev_upload(c.d_aidx, c.aidx.data(), c.aidx.size(), s));
  PZG_TRY(dev_upload(c.d_tvol, reinterpret_cast<const char *>(c.tvol_all.data()),
                     sizeof(ElemTables) * c.tvol_all.size(), s));
  PZG_TRY(dev_upload(c.d_tbc, reinterpret_cast<const char *>(c.tbc_all.data()),
                     sizeof(ElemTables) * c.tbc_all.size(), s));
  if (c.hp_simplex) {
    PZG_TRY(dev_upload(c.d_hp_vsub, vsub.data(), vsub.size(), s));
    PZG_TRY(dev_upload(c.d_hp_bsub, bsub.data(), bsub.size(), s));
    PZG_TRY(dev_upload(c.d_thp_vol, reinterpret_cast<const char *>(c.thp_vol.data()),
                       sizeof(HpTables) * c.thp_vol.size(), s));
    if (!c.thp_bc.empty())
      PZG_TRY(dev_upload(c.d_thp_bc, reinterpret_cast<const char *>(c.thp_bc.data()),
                         sizeof(HpTables) * c.thp_bc.size(), s));
  }
  PZG_HIP(hipStreamSynchronize(s));
  c.les_graph = false;
  // a new mesh has its own numbering: no equation filter
  c.filt = false;
  c.filt_active.clear();
  c.filt_dest.clear();
  c.filt_keep.clear();
  c.filt_ia.clear();
  c.has_mesh = true;
  c.has_struct = false;
  c.mf_ready = c.mf_vals_valid = false;
  c.les_tile_ready = c.les_tile_ok = false;
  mg_free(c);  // the multigrid lattice belongs to the mesh
  c.assembled = false;
  c.external_csr = false;
  c.has_solution = c.f_valid = false;
  c.post_porder = -1;
  return PZG_OK;
}

// Node graph of the active blocks (TPZRenumbering::ConvertGraph semantics:
// union of the connects of all elements that touch a block, sorted, self
// excluded), split into the upper part (SetupMatrixData's column order) and
// the full symmetric part.
static int struct_create(Ctx &c) {
  if (!c.has_mesh) {
    set_error("pzg_struct_create: call pzg_mesh_set first");
    return PZG_ESTATE;
  }
  const int64_t nb = c.nb, bs = c.bs;
  const int nsv = c.vol_nsides, nsb = c.bc_nsides;
  const int64_t nel = c.nvol + c.nbc;
  // a constrained element also couples the connects its constrained ones
  // depend on (TPZCompMesh::ComputeElGraph over BuildConnectList)
  const bool dep = c.has_dep;
  auto elem_conn = [&](int64_t e, int s) -> int64_t {
    if (e >= c.nvol) {
      const int64_t b = e - c.nvol;
      if (s < nsb) return c.aidx[c.bc_conn[b * nsb + s]];
      return c.aidx[c.cb.xconn[c.cb.xptr[b] + s - nsb]];
    }
    if (s < nsv) return c.aidx[c.vol_conn[e * nsv + s]];
    return c.aidx[c.cv.xconn[c.cv.xptr[e] + s - nsv]];
  };
  auto elem_ns = [&](int64_t e) -> int {
    if (e >= c.nvol) {
      const int64_t b = e - c.nvol;
      return dep ? nsb + (int)(c.cb.xptr[b + 1] - c.cb.xptr[b]) : nsb;
    }
    return dep ? nsv + (int)(c.cv.xptr[e + 1] - c.cv.xptr[e]) : nsv;
  };
  // block -> element incidence
  std::vector<int64_t> bptr(nb + 1, 0);
  for (int64_t e = 0; e < nel; e++)
    for (int s = 0; s < elem_ns(e); s++) {
      int64_t a = elem_conn(e, s);
      if (a >= 0) bptr[a + 1]++;
    }
  for (int64_t a = 0; a < nb; a++) bptr[a + 1] += bptr[a];
  std::vector<int64_t> belem(bptr[nb]);
  {
    std::vector<int64_t> fill(bptr.begin(), bptr.end() - 1);
    for (int64_t e = 0; e < nel; e++)
      for (int s = 0; s < elem_ns(e); s++) {
        int64_t a = elem_conn(e, s);
        if (a >= 0) belem[fill[a]++] = e;
      }
  }
  auto neighbours = [&](int64_t a, std::vector<int64_t> &buf) {
    buf.clear();
    for (int64_t k = bptr[a]; k < bptr[a + 1]; k++) {
      int64_t e = belem[k];
      for (int s = 0; s < elem_ns(e); s++) {
        int64_t b = elem_conn(e, s);
        if (b >= 0 && b != a) buf.push_back(b);
      }
    }
    std::sort(buf.begin(), buf.end());
    buf.erase(std::unique(buf.begin(), buf.end()), buf.end());
  };
  std::vector<int32_t> nup(nb), nlow(nb);
  {
    int T = nthreads();
    std::vector<std::thread> th;
    for (int t = 0; t < T; t++)
      th.emplace_back([&, t]() {
        std::vector<int64_t> buf;
        for (int64_t a = nb * t / T; a < nb * (t + 1) / T; a++) {
          neighbours(a, buf);
          int64_t lo = std::lower_bound(buf.begin(), buf.end(), a) - buf.begin();
          nlow[a] = (int32_t)lo;
          nup[a] = (int32_t)(buf.size() - lo);
        }
      });
    for (auto &x : th) x.join();
  }
  std::vector<int64_t> ubia(nb + 1, 0), fbia(nb + 1, 0);
  for (int64_t a = 0; a < nb; a++) {
    ubia[a + 1] = ubia[a] + nup[a];
    fbia[a + 1] = fbia[a] + nlow[a] + 1 + nup[a];
  }
  c.nbu = ubia[nb];
  c.nbf = fbia[nb];
  std::vector<int32_t> ubja(std::max<int64_t>(c.nbu, 1)), fbja(std::max<int64_t>(c.nbf, 1));
  {
    int T = nthreads();
    std::vector<std::thread> th;
    for (int t = 0; t < T; t++)
      th.emplace_back([&, t]() {
        std::vector<int64_t> buf;
        for (int64_t a = nb * t / T; a < nb * (t + 1) / T; a++) {
          neighbours(a, buf);
          int64_t f = fbia[a], u = ubia[a];
          for (int64_t k = 0; k < nlow[a]; k++) fbja[f++] = (int32_t)buf[k];
          fbja[f++] = (int32_t)a;
          for (int64_t k = nlow[a]; k < (int64_t)buf.size(); k++) {
            fbja[f++] = (int32_t)buf[k];
            ubja[u++] = (int32_t)buf[k];
          }
        }
      });
    for (auto &x : th) x.join();
  }
  std::vector<int32_t> rank_fwd(std::max<int64_t>(c.nbu, 1)), rank_bwd(std::max<int64_t>(c.nbu, 1));
  parallel_for(nb, [&](int64_t a) {
    for (int64_t k = ubia[a]; k < ubia[a + 1]; k++) {
      int64_t b = ubja[k];
      rank_fwd[k] = (int32_t)(nlow[a] + 1 + (k - ubia[a]));
      const int32_t *row = fbja.data() + fbia[b];
      rank_bwd[k] = (int32_t)(std::lower_bound(row, row + nlow[b], (int32_t)a) - row);
    }
  });
  // equation-level row pointers: reference upper CSR and full CSR
  const int64_t N = c.neq;
  std::vector<int64_t> ia(N + 1);
  int64_t ub = 0, fb = 0;
  for (int64_t a = 0; a < nb; a++) {
    for (int64_t k = 0; k < bs; k++) {
      ia[a * bs + k] = ub;
      ub += (bs - k) + bs * (int64_t)nup[a];
      fb += bs * (int64_t)(nlow[a] + 1 + nup[a]);
    }
  }
  ia[N] = ub;
  c.nnz_u = ub;
  c.nnz_full = fb;
  hipStream_t s = c.stream;
  PZG_TRY(dev_upload(c.d_ubia, ubia.data(), ubia.size(), s));
  PZG_TRY(dev_upload(c.d_ubja, ubja.data(), ubja.size(), s));
  PZG_TRY(dev_upload(c.d_fbia, fbia.data(), fbia.size(), s));
  PZG_TRY(dev_upload(c.d_fbja, fbja.data(), fbja.size(), s));
  PZG_TRY(dev_upload(c.d_rank_fwd, rank_fwd.data(), rank_fwd.size(), s));
  PZG_TRY(dev_upload(c.d_rank_bwd, rank_bwd.data(), rank_bwd.size(), s));
  c.lowsrc_ready = false;
  PZG_TRY(dev_upload(c.d_ia, ia.data(), ia.size(), s));
  PZG_TRY(dev_alloc(c.d_ja, sizeof(int64_t) * std::max<int64_t>(c.nnz_u, 1)));
  PZG_TRY(dev_alloc(c.d_a, sizeof(double) * std::max<int64_t>(c.nnz_u, 1)));
  PZG_TRY(dev_alloc(c.d_rhs, sizeof(double) * std::max<int64_t>(N, 1)));
  PZG_TRY(dev_alloc(c.d_diag, sizeof(double) * std::max<int64_t>(N, 1)));
  PZG_TRY(launch_expand_upper_ja(c));
  if (c.filt) {
    // condensed pattern: the mesh-numbered upper CSR restricted to the active
    // rows and columns (what SetupMatrixData builds with the filter,
    // StrMatrix/TPZSSpStructMatrix.cpp:49-193)
    std::vector<int64_t> ja(std::max<int64_t>(c.nnz_u, 1));
    PZG_HIP(hipMemcpyAsync(ja.data(), c.d_ja.p, sizeof(int64_t) * c.nnz_u, hipMemcpyDeviceToHost, s));
    PZG_HIP(hipStreamSynchronize(s));
    const int64_t na = (int64_t)c.filt_active.size();
    c.filt_ia.assign(na + 1, 0);
    c.filt_keep.clear();
    for (int64_t i = 0; i < na; i++) {
      const int64_t r = c.filt_active[i];
      for (int64_t k = ia[r]; k < ia[r + 1]; k++)
        if (c.filt_dest[ja[k]] >= 0) c.filt_keep.push_back(k);
      c.filt_ia[i + 1] = (int64_t)c.filt_keep.size();
    }
    PZG_TRY(dev_upload(c.d_filt_dest, c.filt_dest.data(), c.filt_dest.size(), s));
  }
  // colourings and element -> upper-block maps
  std::vector<int64_t> vptr, bptr2;
  std::vector<int32_t> vlist, blist;
  PZG_TRY(color_elements(c.nvol, nsv, c.vol_conn, c.aidx, nb, vptr, vlist));
  group_by_bucket(vptr, vlist, c.vol_bucket, (int)c.vol_rules.size());
  // constrained elements leave their colours; they get colours of their own
  // over their unconstrained blocks (first fit)
  auto drop_cel = [&](std::vector<int64_t> &ptr, std::vector<int32_t> &list, const CelSet &cs) {
    std::vector<int64_t> p2(ptr.size(), 0);
    std::vector<int32_t> l2;
    for (size_t k = 0; k + 1 < ptr.size(); k++) {
      for (int64_t i = ptr[k]; i < ptr[k + 1]; i++)
        if (!cs.is_cel[list[i]]) l2.push_back(list[i]);
      p2[k + 1] = (int64_t)l2.size();
    }
    ptr.swap(p2);
    list.swap(l2);
  };
  auto color_cel = [&](CelSet &cs, int64_t eoff) -> int {
    const int64_t ncel = (int64_t)cs.cel.size();
    std::vector<uint64_t> used(nb, 0);
    std::vector<int> col(ncel);
    int ncol = 0;
    for (int64_t k = 0; k < ncel; k++) {
      const int64_t e = eoff + cs.cel[k];
      std::vector<int64_t> bl;
      for (int q = 0; q < elem_ns(e); q++)
        if (elem_conn(e, q) >= 0) bl.push_back(elem_conn(e, q));
      uint64_t msk = 0;
      for (int64_t b : bl) msk |= used[b];
      if (msk == ~0ULL) {
        set_error("constrained element colouring needs more than 64 colours");
        return PZG_EUNSUP;
      }
      col[k] = __builtin_ctzll(~msk);
      ncol = std::max(ncol, col[k] + 1);
      for (int64_t b : bl) used[b] |= 1ULL << col[k];
    }
    cs.cptr.assign(ncol + 1, 0);
    for (int64_t k = 0; k < ncel; k++) cs.cptr[col[k] + 1]++;
    for (int k = 0; k < ncol; k++) cs.cptr[k + 1] += cs.cptr[k];
    std::vector<int32_t> cslot(ncel);
    std::vector<int64_t> fill(cs.cptr.begin(), cs.cptr.end() - 1);
    for (int64_t k = 0; k < ncel; k++) cslot[fill[col[k]]++] = (int32_t)k;
    PZG_TRY(dev_upload(cs.d_cslot, cslot.data(), cslot.size(), s));
    return PZG_OK;
  };
  if (c.has_dep) {
    drop_cel(vptr, vlist, c.cv);
    PZG_TRY(color_cel(c.cv, 0));
    PZG_TRY(color_cel(c.cb, c.nvol));
  } else {
    c.cv.cptr.assign(1, 0);
    c.cb.cptr.assign(1, 0);
  }
  // affine volume elements (parallelepipeds / parallelograms) first in each
  // colour: the assembly integrates them from the rule's reference Gram
  // matrices (k_vol_aff); single-rule quad/hex meshes only
  c.vcolor_aff.clear();
  c.blk_asm = false;
  c.aff_on = !(std::getenv("PZG_NO_AFFINE") && std::getenv("PZG_NO_AFFINE")[0] == '1');
  if (c.vol_rules.size() == 1 && !c.hp_simplex && c.vol_etype != PZG_TET &&
      ((c.dim == 3 && c.vol_ncorner == 8) || (c.dim == 2 && c.vol_ncorner == 4))) {
    const int NC = c.vol_ncorner;
    auto affine = [&](int64_t e) {
      const double *x[8];
      for (int k = 0; k < NC; k++) x[k] = &c.xyz[3 * c.vol_nodes[e * NC + k]];
      double scale = 0., dev = 0.;
      for (int i = 0; i < 3; i++) {
        const double a1 = x[1][i] - x[0][i], a3 = x[3][i] - x[0][i];
        scale = std::max(scale, std::max(std::fabs(a1), std::fabs(a3)));
        dev = std::max(dev, std::fabs(x[2][i] - (x[1][i] + a3)));
        if (NC == 8) {
          const double a4 = x[4][i] - x[0][i];
          scale = std::max(scale, std::fabs(a4));
          dev = std::max(dev, std::fabs(x[5][i] - (x[1][i] + a4)));
          dev = std::max(dev, std::fabs(x[7][i] - (x[3][i] + a4)));
          dev = std::max(dev, std::fabs(x[6][i] - (x[1][i] + a3 + a4)));
        }
      }
      return scale > 0. && dev <= 1e-12 * scale;
    };
    const size_t ncol = vptr.size() - 1;
    c.vcolor_aff.assign(ncol, 0);
    for (size_t k = 0; k < ncol; k++) {
      auto b = vlist.begin() + vptr[k], en = vlist.begin() + vptr[k + 1];
      auto mid = std::stable_partition(b, en, [&](int32_t e) { return affine(e); });
      c.vcolor_aff[k] = mid - b;
    }
    // reference Gram matrices R_mn(i,j) = sum_q w_q dphi_m(i) dphi_n(j), m <= n
    const ElemTables &T = c.tvol;
    const int D = c.dim, ns = T.nsh;
    std::vector<double> R((size_t)(D * (D + 1) / 2) * ns * ns, 0.);
    int r = 0;
    for (int m = 0; m < D; m++)
      for (int n = m; n < D; n++, r++)
        for (int i = 0; i < ns; i++)
          for (int j = 0; j < ns; j++) {
            double v = 0.;
            for (int q = 0; q < T.nq; q++) v += T.w[q] * T.dphi[q][m][i] * T.dphi[q][n][j];
            R[((size_t)r * ns + i) * ns + j] = v;
          }
    PZG_TRY(dev_upload(c.d_vol_R, R.data(), R.size(), s));
    PZG_TRY(blk_setup(c, R));
  }
  c.vcolor_ptr = vptr;
  PZG_TRY(dev_upload(c.d_vcolor_list, vlist.data(), vlist.size(), s));
  if (c.nbc > 0) {
    PZG_TRY(color_elements(c.nbc, nsb, c.bc_conn, c.aidx, nb, bptr2, blist));
    group_by_bucket(bptr2, blist, c.bc_bucket, (int)c.bc_rules.size());
    if (c.has_dep) drop_cel(bptr2, blist, c.cb);
    c.bcolor_ptr = bptr2;
    PZG_TRY(dev_upload(c.d_bcolor_list, blist.data(), blist.size(), s));
  } else {
    c.bcolor_ptr.assign(1, 0);
  }
  PZG_TRY(launch_elem_maps(c));
  PZG_HIP(hipStreamSynchronize(s));
  c.has_struct = true;
  c.external_csr = false;
  c.assembled = false;
  c.kel_added = false;
  c.mf_ready = c.mf_vals_valid = false;
  c.les_tile_ready = c.les_tile_ok = false;
  c.full_valid = c.bsr_valid = c.scal_valid = false;
  c.has_solution = c.f_valid = false;
  return PZG_OK;
}

}  // namespace pzg

using namespace pzg;

struct pzg_ctx {
  Ctx c;
};

extern "C" {

const char *pzg_last_error(void) { return g_err.c_str(); }

int pzg_create(int device, pzg_ctx **out) {
  if (!out) {
    set_error("pzg_create: null output pointer");
    return PZG_EINVAL;
  }
  *out = nullptr;
  int n = 0;
  PZG_HIP(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) {
    set_error("pzg_create: no HIP device " + std::to_string(device));
    return PZG_EINVAL;
  }
  PZG_HIP(hipSetDevice(device));
  PZG_TRY(noalloc_scratch_setup(device));
  pzg_ctx *p = new pzg_ctx();
  p->c.device = device;
  hipError_t e = hipStreamCreateWithFlags(&p->c.stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&p->c.ev0);
  if (e == hipSuccess) e = hipEventCreate(&p->c.ev1);
  if (e != hipSuccess) {
    set_error(std::string("pzg_create: ") + hipGetErrorString(e));
    delete p;
    return PZG_EHIP;
  }
  *out = p;
  return PZG_OK;
}

void pzg_destroy(pzg_ctx *ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->c.device);
  (void)hipStreamSynchronize(ctx->c.stream);
  free_all(ctx->c);
  if (ctx->c.ev0) (void)hipEventDestroy(ctx->c.ev0);
  if (ctx->c.ev1) (void)hipEventDestroy(ctx->c.ev1);
  if (ctx->c.stream) (void)hipStreamDestroy(ctx->c.stream);
  delete ctx;
}

#define CTX_CHECK()                                      \
  do {                                                   \
    if (!ctx) {                                          \
      set_error("null context");                         \
      return PZG_EINVAL;                                 \
    }                                                    \
    PZG_HIP(hipSetDevice(ctx->c.device));                \
  } while (0)

int pzg_mesh_set(pzg_ctx *ctx, const pzg_mesh_desc *mesh) {
  CTX_CHECK();
  return mesh_set(ctx->c, mesh);
}

int pzg_struct_create(pzg_ctx *ctx, int64_t *neq, int64_t *nnz) {
  CTX_CHECK();
  PZG_TRY(struct_create(ctx->c));
  const Ctx &c = ctx->c;
  if (neq) *neq = c.filt ? (int64_t)c.filt_active.size() : c.neq;
  if (nnz) *nnz = c.filt ? (int64_t)c.filt_keep.size() : c.nnz_u;
  return PZG_OK;
}

int pzg_set_equation_filter(pzg_ctx *ctx, int64_t nactive, const int64_t *active) {
  CTX_CHECK();
  Ctx &c = ctx->c;
  if (!c.has_mesh) {
    set_error("pzg_set_equation_filter: call pzg_mesh_set first");
    return PZG_ESTATE;
  }
  if (nactive < 0 || !active) {
    if (!c.filt) return PZG_OK;  // no filter before either: nothing changes
    c.filt = false;
    c.filt_active.clear();
    c.filt_dest.clear();
  } else {
    std::vector<int64_t> v(active, active + nactive);
    for (int64_t e : v)
      if (e < 0 || e >= c.neq) {
        set_error("pzg_set_equation_filter: active equation out of range");
        return PZG_EINVAL;
      }
    // TPZEquationFilter::SetActiveEquations: sorted, duplicates removed
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
    c.filt_active.swap(v);
    c.filt_dest.assign(c.neq, -1);
    for (size_t i = 0; i < c.filt_active.size(); i++) c.filt_dest[c.filt_active[i]] = (int64_t)i;
    c.filt = true;
  }
  c.filt_keep.clear();
  c.filt_ia.clear();
  // the pattern belongs to the filter: pzg_struct_create builds it again
  c.has_struct = false;
  c.assembled = false;
  c.has_solution = c.f_valid = false;
  c.mf_ready = c.mf_vals_valid = false;
  c.les_tile_ready = c.les_tile_ok = false;
  c.full_valid = c.bsr_valid = c.scal_valid = false;
  mg_free(c);
  return PZG_OK;
}

// TPZEquationFilter::Scatter / Gather of host vectors between the condensed
// numbering (the matrix's rows) and the mesh numbering of the device system.
static void filt_scatter(const Ctx &c, const double *small, std::vector<double> &big) {
  big.assign(c.neq, 0.);
  for (size_t i = 0; i < c.filt_active.size(); i++) big[c.filt_active[i]] = small[i];
}
static void filt_gather(const Ctx &c, const double *big, double *small) {
  for (size_t i = 0; i < c.filt_active.size(); i++) small[i] = big[c.filt_active[i]];
}

int pzg_struct_get_pattern(pzg_ctx *ctx, int64_t *ia, int64_t *ja) {
  CTX_CHECK();
  Ctx &c = ctx->c;
  if (!c.has_struct) {
    set_error("pzg_struct_get_pattern: no pattern");
    return PZG_ESTATE;
  }
  if (c.filt) {
    if (ia) std::copy(c.filt_ia.begin(), c.filt_ia.end(), ia);
    if (ja) {
      std::vector<int64_t> full(std::max<int64_t>(c.nnz_u, 1));
      PZG_HIP(hipMemcpyAsync(full.data(), c.d_ja.p, sizeof(int64_t) * c.nnz_u, hipMemcpyDeviceToHost, c.stream));
      PZG_HIP(hipStreamSynchronize(c.stream));
      for (size_t k = 0; k < c.filt_keep.size(); k++) ja[k] = c.filt_dest[full[c.filt_keep[k]]];
    }
    return PZG_OK;
  }
  if (ia) PZG_HIP(hipMemcpyAsync(ia, c.d_ia.p, sizeof(int64_t) * (c.neq + 1), hipMemcpyDeviceToHost, c.stream));
  if (ja && c.nnz_u) PZG_HIP(hipMemcpyAsync(ja, c.d_ja.p, sizeof(int64_t) * c.nnz_u, hipMemcpyDeviceToHost, c.stream));
  PZG_HIP(hipStreamSynchronize(c.stream));
  return PZG_OK;
}

int pzg_assemble(pzg_ctx *ctx) {
  CTX_CHECK();
  Ctx &c = ctx->c;
  if (!c.has_struct || c.external_csr || !c.has_mesh) {
    set_error("pzg_assemble: call pzg_mesh_set and pzg_struct_create first");
    return PZG_ESTATE;
  }
  PZG_TRY(launch_assemble(c));
  PZG_TRY(launch_filter_mask(c));
  PZG_TRY(launch_diag(c));
  c.assembled = true;
  c.kel_added = false;
  c.full_valid = c.bsr_valid = c.scal_valid = false;
  c.has_solution = c.f_valid = false;
  return PZG_OK;
}

int pzg_get_system(pzg_ctx *ctx, double *a, double *rhs) {
  CTX_CHECK();
  Ctx &c = ctx->c;
  if (!c.assembled) {
    set_error("pzg_get_system: nothing assembled");
    return PZG_ESTATE;
  }
  if (c.filt) {  // the condensed values and load (matrix numbering)
    std::vector<double> full(std::max<int64_t>(std::max(c.nnz_u, c.neq), 1));
    if (a) {
      PZG_HIP(hipMemcpyAsync(full.data(), c.d_a.p, sizeof(double) * c.nnz_u, hipMemcpyDeviceToHost, c.stream));
      PZG_HIP(hipStreamSynchronize(c.stream));
      for (size_t k = 0; k < c.filt_keep.size(); k++) a[k] = full[c.filt_keep[k]];
    }
    if (rhs) {
      PZG_HIP(hipMemcpyAsync(full.data(), c.d_rhs.p, sizeof(double) * c.neq, hipMemcpyDeviceToHost, c.stream));
      PZG_HIP(hipStreamSynchronize(c.stream));
      filt_gather(c, full.data(), rhs);
    }
    return PZG_OK;
  }
  if (a && c.nnz_u) PZG_HIP(hipMemcpyAsync(a, c.d_a.p, sizeof(double) * c.nnz_u, hipMemcpyDeviceToHost, c.stream));
  if (rhs) PZG_HIP(hipMemcpyAsync(rhs, c.d_rhs.p, sizeof(double) * c.neq, hipMemcpyDeviceToHost, c.stream));
  PZG_HIP(hipStreamSynchronize(c.stream));
  return PZG_OK;
}

int pzg_hp_tuning(pzg_ctx *ctx, int32_t *n, int32_t *order, int32_t *variant, float *ms) {
  CTX_CHECK();
  Ctx &c = ctx->c;
  if (!n) return PZG_EINVAL;
  const int cap = *n;
  if (!c.hp_simplex || !c.hp_tuned) {
    *n = 0;
    return PZG_OK;
  }
  const int nbk = (int)c.hp_vol_var.size();
  *n = nbk;
  for (int b = 0; b < nbk && b < cap; b++) {
    if (order) order[b] = c.hp_vol_P[b];
    if (variant) variant[b] = c.hp_vol_var[b];
    if (ms)
      for (int v = 0; v < 5; v++) ms[b * 5 + v] = c.hp_vol_ms[b * 5 + v];
  }
  return PZG_OK;
}

int pzg_full_csr(pzg_ctx *ctx, int64_t *ia, int64_t *ja, double *a) {
  CTX_CHECK();
  Ctx &c = ctx->c;
  if (!c.assembled) {
    set_error("pzg_full_csr: needs an assembled system");
    return PZG_ESTATE;
  }
  if (c.filt) {
    set_error("pzg_full_csr: the non-symmetric format with an equation filter is not implemented");
    return PZG_EUNSUP;
  }
  const int64_t nnz = c.nnz_full;
  DevBuf dia, dja, da;
  if (ia) PZG_TRY(dev_alloc(dia, sizeof(int64_t) * (c.neq + 1)));
  if (ja) PZG_TRY(dev_alloc(dja, sizeof(int64_t) * std::max<int64_t>(nnz, 1)));
  if (a) PZG_TRY(dev_alloc(da, sizeof(double) * std::max<int64_t>(nnz, 1)));
  int rc = launch_full_csr(c, ia ? dia.as<int64_t>() : nullptr, ja ? dja.as<int64_t>() : nullptr,
                           a ? da.as<double>() : nullptr);
  if (rc == PZG_OK) {
    hipError_t e = hipSuccess;
    if (ia && e == hipSuccess) e = hipMemcpyAsync(ia, dia.p, sizeof(int64_t) * c.neq, hipMemcpyDeviceToHost, c.stream);
    if (ja && e == hipSuccess) e = hipMemcpyAsync(ja, dja.p, sizeof(int64_t) * nnz, hipMemcpyDeviceToHost, c.stream);
    if (a && e == hipSuccess) e = hipMemcpyAsync(a, da.p, sizeof(double) * nnz, hipMemcpyDeviceToHost, c.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c.stream);
    if (e != hipSuccess) {
      set_error(std::string("pzg_full_csr: ") + hipGetErrorString(e));
      rc = PZG_EHIP;
    }
  }
  if (ia && rc == PZG_OK) ia[c.neq] = nnz;
  dev_free(dia);
  dev_free(dja);
  dev_free(da);
  return rc;
}

int pzg_csr_set(pzg_ctx *ctx, int64_t neq, const int64_t *ia, const int64_t *ja, const double *a,
                const double *rhs) {
  CTX_CHECK();
  Ctx &c = ctx->c;
  if (neq <= 0 || !ia || !ja || !a || ia[0] != 0) {
    set_error("pzg_csr_set: invalid CSR");
    return PZG_EINVAL;
  }
  // validate: rows sorted, upper (col >= row), diagonal first
  for (int64_t r = 0; r < neq; r++) {
    if (ia[r + 1] <= ia[r] || ja[ia[r]] != r) {
      set_error("pzg_csr_set: row " + std::to_string(r) + " has no leading diagonal");
      return PZG_EINVAL;
    }
    for (int64_t k = ia[r] + 1; k < ia[r + 1]; k++)
      if (ja[k] <= ja[k - 1] || ja[k] >= neq) {
        set_error("pzg_csr_set: row " + std::to_string(r) + " not sorted/upper");
        return PZG_EINVAL;
      }
  }
  // A caller matrix is handled as a scalar problem (block size 1), in its
  // own numbering (no equation filter).
  c.has_mesh = false;
  c.filt = false;
  c.filt_active.clear();
  c.filt_dest.clear();
  c.neq = neq;
  c.bs = 1;
  c.nb = neq;
  c.nnz_u = ia[neq];
  std::vector<int32_t> nlow(neq, 0);
  std::vector<int64_t> rowof(c.nnz_u);
  for (int64_t r = 0; r < neq; r++)
    for (int64_t k = ia[r]; k < ia[r + 1]; k++) {
      rowof[k] = r;
      if (ja[k] != r) nlow[ja[k]]++;
    }
  std::vector<int64_t> ubia(neq + 1), fbia(neq + 1);
  ubia[0] = fbia[0] = 0;
  for (int64_t r = 0; r < neq; r++) {
    int64_t nup = ia[r + 1] - ia[r] - 1;
    ubia[r + 1] = ubia[r] + nup;
    fbia[r + 1] = fbia[r] + nlow[r] + 1 + nup;
  }
  c.nbu = ubia[neq];
  c.nbf = fbia[neq];
  c.nnz_full = c.nbf;
  std::vector<int32_t> ubja(std::max<int64_t>(c.nbu, 1)), fbja(std::max<int64_t>(c.nbf, 1));
  std::vector<int32_t> rank_fwd(std::max<int64_t>(c.nbu, 1)), rank_bwd(std::max<int64_t>(c.nbu, 1));
  std::vector<int64_t> lowfill(neq, 0);
  for (int64_t r = 0; r < neq; r++) {
    int64_t u = ubia[r];
    for (int64_t k = ia[r] + 1; k < ia[r + 1]; k++) {
      int64_t col = ja[k];
      ubja[u] = (int32_t)col;
      rank_fwd[u] = (int32_t)(nlow[r] + 1 + (u - ubia[r]));
      rank_bwd[u] = (int32_t)lowfill[col];
      fbja[fbia[col] + lowfill[col]] = (int32_t)r;
      lowfill[col]++;
      u++;
    }
    fbja[fbia[r] + nlow[r]] = (int32_t)r;
    for (int64_t k = ia[r] + 1; k < ia[r + 1]; k++) fbja[fbia[r] + nlow[r] + (k - ia[r])] = (int32_t)ja[k];
  }
  hipStream_t s = c.stream;
  PZG_TRY(dev_upload(c.d_ubia, ubia.data(), ubia.size(), s));
  PZG_TRY(dev_upload(c.d_ubja, ubja.data(), ubja.size(), s));
  PZG_TRY(dev_upload(c.d_fbia, fbia.data(), fbia.size(), s));
  PZG_TRY(dev_upload(c.d_fbja, fbja.data(), fbja.size(), s));
  PZG_TRY(dev_upload(c.d_rank_fwd, rank_fwd.data(), rank_fwd.size(), s));
  PZG_TRY(dev_upload(c.d_rank_bwd, rank_bwd.data(), rank_bwd.size(), s));
  c.lowsrc_ready = false;
  PZG_TRY(dev_upload(c.d_ia, ia, neq + 1, s));
  PZG_TRY(dev_upload(c.d_ja, ja, c.nnz_u, s));
  PZG_TRY(dev_upload(c.d_a, a, c.nnz_u, s));
  PZG_TRY(dev_alloc(c.d_rhs, sizeof(double) * neq));
  if (rhs) PZG_HIP(hipMemcpyAsync(c.d_rhs.p, rhs, sizeof(double) * neq, hipMemcpyHostToDevice, s));
  else PZG_HIP(hipMemsetAsync(c.d_rhs.p, 0, sizeof(double) * neq, s));
  PZG_TRY(dev_alloc(c.d_diag, sizeof(double) * neq));
  PZG_TRY(launch_diag(c));
  PZG_HIP(hipStreamSynchronize(s));
  c.has_struct = true;
  c.external_csr = true;
  c.mf_ready = c.mf_vals_valid = false;
  c.les_tile_ready = c.les_tile_ok = false;
  c.assembled = true;
  c.full_valid = c.bsr_valid = c.scal_valid = false;
  c.has_solution = c.f_valid = false;
  return PZG_OK;
}

int pzg_add_kel_batch(pzg_ctx *ctx, int64_t nel, int32_t ne, const int64_t *dest, const double *ek) {
  CTX_CHECK();
  Ctx &c = ctx->c;
  if (!c.has_struct) {
    set_error("pzg_add_kel_batch: no pattern (pzg_struct_create or pzg_csr_set first)");
    return PZG_ESTATE;
  }
  if (nel < 0 || ne <= 0 || ne > 1024 || (nel > 0 && (!dest || !ek))) {
    set_error("pzg_add_kel_batch: invalid arguments");
    return PZG_EINVAL;
  }
  if (nel == 0) return PZG_OK;
  const int64_t nrow = c.filt ? (int64_t)c.filt_active.size() : c.neq;
  for (int64_t i = 0; i < nel * ne; i++)
    if (dest[i] < 0 || dest[i] >= nrow) {
      set_error("pzg_add_kel_batch: destination equation out of range");
      return PZG_EINVAL;
    }
  std::vector<int64_t> dmesh;
  if (c.filt) {  // condensed destinations -> mesh numbering of the device system
    dmesh.resize(nel * ne);
    for (int64_t i = 0; i < nel * ne; i++) dmesh[i] = c.filt_active[dest[i]];
    dest = dmesh.data();
  }
  // colour by equations so that one launch never writes a row twice
  std::vector<uint64_t> used(c.neq, 0);
  std::vector<int> col(nel);
  int ncol = 0;
  for (int64_t e = 0; e < nel; e++) {
    uint64_t m = 0;
    for (int i = 0; i < ne; i++) m |= used[dest[e * ne + i]];
    if (m == ~0ULL) {
      set_error("pzg_add_kel_batch: more than 64 colours");
      return PZG_EUNSUP;
    }
    int k = __builtin_ctzll(~m);
    col[e] = k;
    ncol = std::max(ncol, k + 1);
    for (int i = 0; i < ne; i++) used[dest[e * ne + i]] |= 1ULL << k;
  }
  std::vector<int64_t> ptr(ncol + 1, 0);
  for (int64_t e = 0; e < nel; e++) ptr[col[e] + 1]++;
  for (int k = 0; k < ncol; k++) ptr[k + 1] += ptr[k];
  std::vector<int32_t> list(nel);
  std::vector<int64_t> fill(ptr.begin(), ptr.end() - 1);
  for (int64_t e = 0; e < nel; e++) list[fill[col[e]]++] = (int32_t)e;
  DevBuf d_dest, d_ek, d_list;
  int rc = dev_upload(d_dest, dest, nel * ne, c.stream);
  if (rc == PZG_OK) rc = dev_upload(d_ek, ek, nel * ne * (int64_t)ne, c.stream);
  if (rc == PZG_OK) rc = dev_upload(d_list, list.data(), list.size(), c.stream);
  for (int k = 0; k < ncol && rc == PZG_OK; k++)
    rc = launch_addkel_batch(c, nel, ne, d_dest.as<int64_t>(), d_ek.as<double>(),
                             d_list.as<int32_t>() + ptr[k], ptr[k + 1] - ptr[k]);
  if (rc == PZG_OK) rc = launch_diag(c);
  if (rc == PZG_OK && hipStreamSynchronize(c.stream) != hipSuccess) rc = PZG_EHIP;
  dev_free(d_dest);
  dev_free(d_ek);
  dev_free(d_list);
  c.assembled = true;
  c.kel_added = true;
  c.full_valid = c.bsr_valid = c.scal_valid = false;
  c.mg_ready = c.pen_ready = false;  // both are built from K
  return rc;
}

int pzg_multadd(pzg_ctx *ctx, const double *x, const double *y, double *z, double alpha, double beta) {
  CTX_CHECK();
  Ctx &c = ctx->c;
  if (!c.assembled) {
    set_error("pzg_multadd: no matrix");
    return PZG_ESTATE;
  }
  if (!x || !z || (beta != 0. && !y)) {
    set_error("pzg_multadd: null vector");
    return PZG_EINVAL;
  }
  int64_t n = c.neq;
  std::vector<double> xb, yb, zb;
  double *zout = z;
  if (c.filt) {  // condensed vectors
    filt_scatter(c, x, xb);
    x = xb.data();
    if (beta != 0.) {
      filt_scatter(c, y, yb);
      y = yb.data();
    }
    zb.resize(n);
    zout = zb.data();
  }
  PZG_TRY(dev_alloc(c.d_x, n * sizeof(double)));
  PZG_TRY(dev_alloc(c.d_w, n * sizeof(double)));
  PZG_TRY(dev_alloc(c.d_z, n * sizeof(double)));
  PZG_HIP(hipMemcpyAsync(c.d_x.p, x, n * sizeof(double), hipMemcpyHostToDevice, c.stream));
  if (beta != 0.) PZG_HIP(hipMemcpyAsync(c.d_w.p, y, n * sizeof(double), hipMemcpyHostToDevice, c.stream));
  PZG_TRY(launch_multadd_upper(c, c.d_x.as<double>(), beta != 0. ? c.d_w.as<double>() : nullptr,
                               c.d_z.as<double>(), alpha, beta));
  PZG_HIP(hipMemcpyAsync(zout, c.d_z.p, n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  PZG_HIP(hipStreamSynchronize(c.stream));
  if (c.filt) filt_gather(c, zout, z);
  return PZG_OK;
}

int pzg_solve(pzg_ctx *ctx, const pzg_solve_params *params, const double *rhs, double *u,
              pzg_solve_info *info) {
  CTX_CHECK();
  Ctx &c = ctx->c;
  if (!c.assembled) {
    set_error("pzg_solve: no assembled system");
    return PZG_ESTATE;
  }
  if (!params || params->tol < 0 || params->max_iter < 0) {
    set_error("pzg_solve: invalid parameters");
    return PZG_EINVAL;
  }
  pzg_solve_info tmp;
  pzg_solve_info &inf = info ? *info : tmp;
  std::memset(&inf, 0, sizeof(inf));
  int64_t n = c.neq;
  std::vector<double> rb;
  if (rhs && c.filt) {  // condensed load -> mesh numbering (inactive 0)
    filt_scatter(c, rhs, rb);
    rhs = rb.data();
  }
  PZG_TRY(dev_alloc(c.d_f, n * sizeof(double)));
  if (rhs) PZG_HIP(hipMemcpyAsync(c.d_f.p, rhs, n * sizeof(double), hipMemcpyHostToDevice, c.stream));
  else PZG_HIP(hipMemcpyAsync(c.d_f.p, c.d_rhs.p, n * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
  c.f_valid = true;
  int rc;
  if (params->solver == PZG_SOLVER_PCG64) rc = solve_pcg64(c, *params, inf);
  else if (params->solver == PZG_SOLVER_MIXED) rc = solve_mixed(c, *params, inf);
  else {
    set_error("pzg_solve: unknown solver");
    return PZG_EINVAL;
  }
  if (rc != PZG_OK && rc != PZG_ENOCONV) return rc;
  c.has_solution = true;
  if (u) {
    std::vector<double> ub(c.filt ? n : 0);
    double *dst = c.filt ? ub.data() : u;
    PZG_HIP(hipMemcpyAsync(dst, c.d_u.p, n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    PZG_HIP(hipStreamSynchronize(c.stream));
    if (c.filt) filt_gather(c, dst, u);
  }
  if (rc == PZG_ENOCONV) set_error("pzg_solve: no convergence within max_iter");
  return rc;
}

int pzg_true_residual(pzg_ctx *ctx, double *rel) {
  CTX_CHECK();
  Ctx &c = ctx->c;
  if (!c.has_solution || !rel) {
    set_error("pzg_true_residual: no solution on the device");
    return PZG_ESTATE;
  }
  if (!c.f_valid) {  // no solve of this system yet: the assembled load
    PZG_TRY(dev_alloc(c.d_f, c.neq * sizeof(double)));
    PZG_HIP(hipMemcpyAsync(c.d_f.p, c.d_rhs.p, c.neq * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
    c.f_valid = true;
  }
  return true_residual(c, c.d_u.as<double>(), rel);
}

int pzg_true_residual_noalloc(pzg_ctx *ctx, double *rel) {
  CTX_CHECK();
  Ctx &c = ctx->c;
  if (!c.has_solution || !rel) {
    set_error(!rel ? "pzg_true_residual_noalloc: null output"
                   : "pzg_true_residual_noalloc: no solution on the device");
    return !rel ? PZG_EINVAL : PZG_ESTATE;
  }
  return true_residual_noalloc(c, rel);
}

int pzg_load_solution(pzg_ctx *ctx, const double *u) {
  CTX_CHECK();
  Ctx &c = ctx->c;
  if (!c.has_mesh || !u) {
    set_error(!u ? "pzg_load_solution: null solution" : "pzg_load_solution: no mesh");
    return !u ? PZG_EINVAL : PZG_ESTATE;
  }
  PZG_TRY(dev_alloc(c.d_u, c.neq * sizeof(double)));
  PZG_HIP(hipMemcpyAsync(c.d_u.p, u, c.neq * sizeof(double), hipMemcpyHostToDevice, c.stream));
  PZG_HIP(hipStreamSynchronize(c.stream));
  c.has_solution = true;
  return PZG_OK;
}

int pzg_get_solution(pzg_ctx *ctx, double *u) {
  CTX_CHECK();
  Ctx &c = ctx->c;
  if (!c.has_solution || !u) {
    set_error(!u ? "pzg_get_solution: null output" : "pzg_get_solution: no solution loaded");
    return !u ? PZG_EINVAL : PZG_ESTATE;
  }
  PZG_HIP(hipMemcpyAsync(u, c.d_u.p, c.neq * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  PZG_HIP(hipStreamSynchronize(c.stream));
  return PZG_OK;
}

int pzg_error_points(pzg_ctx *ctx, int32_t porder, int64_t *npts, double *xq) {
  CTX_CHECK();
  return post_points(ctx->c, porder, npts, xq);
}

int pzg_evaluate_error(pzg_ctx *ctx, const pzg_exact *exact, double *el_errors, double *errors,
                       int32_t *nerr) {
  CTX_CHECK();
  if (!exact) {
    set_error("pzg_evaluate_error: null exact solution");
    return PZG_EINVAL;
  }
  int n = 0;
  PZG_TRY(post_errors(ctx->c, *exact, el_errors, errors, &n));
  if (nerr) *nerr = n;
  return PZG_OK;
}

int pzg_sloan_renumber(int64_t nnodes, int64_t nel, const int64_t *elgraphindex,
                       const int64_t *elgraph, int64_t *perm) {
  return sloan_renumber(nnodes, nel, elgraphindex, elgraph, perm);
}

int pzg_get_sizes(pzg_ctx *ctx, pzg_sizes *s) {
  CTX_CHECK();
  Ctx &c = ctx->c;
  if (!s) return PZG_EINVAL;
  s->neq = c.neq;
  s->nnz_upper = c.nnz_u;
  s->nnz_full = c.nnz_full;
  s->nblocks = c.nb;
  s->nnz_blocks = c.nbf;
  s->nvol = c.nvol;
  s->ncolors = c.vcolor_ptr.empty() ? 0
                                   : ((int64_t)c.vcolor_ptr.size() - 1) / (int64_t)std::max<size_t>(c.vol_rules.size(), 1);
  s->bs = c.bs;
  return PZG_OK;
}

int pzg_sync(pzg_ctx *ctx) {
  CTX_CHECK();
  PZG_HIP(hipStreamSynchronize(ctx->c.stream));
  return PZG_OK;
}

int pzg_time_kernel(pzg_ctx *ctx, int32_t which, int32_t nrep, double *avg_ms, double *bytes) {
  CTX_CHECK();
  Ctx &c = ctx->c;
  if (!c.assembled || nrep <= 0 || !avg_ms) {
    set_error("pzg_time_kernel: needs an assembled system");
    return PZG_ESTATE;
  }
  int64_t n = c.neq;
  double b = 0;
  if (which == 0) {
    if (!c.full_valid) PZG_TRY(launch_expand_full(c));
    PZG_TRY(dev_alloc(c.d_x, n * sizeof(double)));
    PZG_TRY(dev_alloc(c.d_w, n * sizeof(double)));
    PZG_HIP(hipMemsetAsync(c.d_x.p, 0, n * sizeof(double), c.stream));
    PZG_TRY(launch_spmv64(c, c.d_x.as<double>(), c.d_w.as<double>()));
    PZG_HIP(hipEventRecord(c.ev0, c.stream));
    for (int i = 0; i < nrep; i++) PZG_TRY(launch_spmv64(c, c.d_x.as<double>(), c.d_w.as<double>()));
    PZG_HIP(hipEventRecord(c.ev1, c.stream));
    b = (double)c.nbf * (c.bs * c.bs * 8.0 + 4.0) + (double)(c.nb + 1) * 8.0 + 2.0 * n * 8.0;
  } else if (which == 1) {
    if (!c.bsr_valid) PZG_TRY(launch_build_bsr(c));
    PZG_TRY(dev_alloc(c.d_xs, n * sizeof(float)));
    PZG_TRY(dev_alloc(c.d_qs, n * sizeof(float)));
    PZG_HIP(hipMemsetAsync(c.d_xs.p, 0, n * sizeof(float), c.stream));
    PZG_TRY(launch_spmv_bsr(c, c.d_xs.as<float>(), c.d_qs.as<float>()));
    PZG_HIP(hipEventRecord(c.ev0, c.stream));
    for (int i = 0; i < nrep; i++) PZG_TRY(launch_spmv_bsr(c, c.d_xs.as<float>(), c.d_qs.as<float>()));
    PZG_HIP(hipEventRecord(c.ev1, c.stream));
    b = (double)c.nbf * (c.bs * c.bs * 4.0 + 4.0) + (double)(c.nb + 1) * 8.0 + 2.0 * n * 4.0;
  } else if (which == 3) {
    if (!mf_available(c)) {
      set_error("pzg_time_kernel: matrix-free operator needs a mesh-built system");
      return PZG_ESTATE;
    }
    PZG_TRY(mf_prepare(c));
    PZG_TRY(dev_alloc(c.d_xs, n * sizeof(float)));
    PZG_TRY(dev_alloc(c.d_qs, n * sizeof(float)));
    PZG_HIP(hipMemsetAsync(c.d_xs.p, 0, n * sizeof(float), c.stream));
    PZG_TRY(mf_apply(c, c.d_xs.as<float>(), nullptr, c.d_qs.as<float>(), nullptr, nullptr, nullptr));
    PZG_HIP(hipEventRecord(c.ev0, c.stream));
    for (int i = 0; i < nrep; i++)
      PZG_TRY(mf_apply(c, c.d_xs.as<float>(), nullptr, c.d_qs.as<float>(), nullptr, nullptr, nullptr));
    PZG_HIP(hipEventRecord(c.ev1, c.stream));
    // compulsory traffic: element maps, corners, coefficients, the element
    // output buffer written and read once, row lists, x, D^-1/2 and y
    // per element: block map and output positions (int32 each), corners,
    // affine geometry, coefficient; the row lists written and read once; row
    // pointers; vectors p and D^-1/2 read twice, pt written and read, q written
    const int64_t nsh = c.tvol.nsh;
    b = (double)c.nvol * (nsh * 8.0 + c.vol_ncorner * 12.0 + 40.0 + 4.0) +
        2.0 * c.mf_nout * c.nstate * 4.0 + (double)(c.nb + 1) * 8.0 + 7.0 * n * 4.0;
  } else if (which == 4) {
    if (!mf_available(c)) {
      set_error("pzg_time_kernel: matrix-free operator needs a mesh-built system");
      return PZG_ESTATE;
    }
    PZG_TRY(mf_prepare(c));
    PZG_TRY(mf_launch_vol(c, nullptr));
    PZG_HIP(hipEventRecord(c.ev0, c.stream));
    for (int i = 0; i < nrep; i++) PZG_TRY(mf_launch_vol(c, nullptr));
    PZG_HIP(hipEventRecord(c.ev1, c.stream));
    // algorithmic flops (returned in 'bytes'): per element the two
    // contractions 2 * (2 nq nsh dim ns) and, per point, the gradient and
    // flux/stress maps 2 * (2 ns dim dim) plus the material law
    const double nq = c.tvol.nq, nsh = c.tvol.nsh, dim = c.dim, ns = c.nstate;
    const double law = ns == 3 ? 24.0 : 3.0;
    b = (double)c.nvol * (2.0 * (2.0 * nq * nsh * dim * ns) + nq * (2.0 * (2.0 * ns * dim * dim) + law));
  } else if (which == 2) {
    if (c.external_csr) {
      set_error("pzg_time_kernel: no mesh to assemble");
      return PZG_ESTATE;
    }
    PZG_HIP(hipEventRecord(c.ev0, c.stream));
    for (int i = 0; i < nrep; i++) PZG_TRY(launch_assemble(c));
    PZG_HIP(hipEventRecord(c.ev1, c.stream));
    b = 0;
  } else {
    set_error("pzg_time_kernel: unknown kernel");
    return PZG_EINVAL;
  }
  PZG_HIP(hipEventSynchronize(c.ev1));
  float ms = 0;
  PZG_HIP(hipEventElapsedTime(&ms, c.ev0, c.ev1));
  *avg_ms = ms / nrep;
  if (bytes) *bytes = b;
  return PZG_OK;
}

int pzg_scaled_apply(pzg_ctx *ctx, int32_t op, const double *x, double *y) {
  CTX_CHECK();
  Ctx &c = ctx->c;
  if (!c.assembled || !x || !y) {
    set_error("pzg_scaled_apply: needs an assembled system and host vectors");
    return PZG_ESTATE;
  }
  if (op != PZG_INNER_BSR && op != PZG_INNER_MATFREE) {
    set_error("pzg_scaled_apply: op must be PZG_INNER_BSR or PZG_INNER_MATFREE");
    return PZG_EINVAL;
  }
  if (op == PZG_INNER_MATFREE && !mf_available(c)) {
    set_error("pzg_scaled_apply: matrix-free operator needs a mesh-built system");
    return PZG_ESTATE;
  }
  const int64_t n = c.neq;
  if (!c.bsr_valid) PZG_TRY(launch_build_bsr(c));
  if (op == PZG_INNER_MATFREE) PZG_TRY(mf_prepare(c));
  std::vector<float> h(n);
  for (int64_t i = 0; i < n; i++) h[i] = (float)x[i];
  PZG_TRY(dev_alloc(c.d_xs, n * sizeof(float)));
  PZG_TRY(dev_alloc(c.d_qs, n * sizeof(float)));
  PZG_HIP(hipMemcpyAsync(c.d_xs.p, h.data(), n * sizeof(float), hipMemcpyHostToDevice, c.stream));
  if (op == PZG_INNER_BSR) PZG_TRY(launch_spmv_bsr(c, c.d_xs.as<float>(), c.d_qs.as<float>()));
  else PZG_TRY(mf_apply(c, c.d_xs.as<float>(), nullptr, c.d_qs.as<float>(), nullptr, nullptr, nullptr));
  PZG_HIP(hipMemcpyAsync(h.data(), c.d_qs.p, n * sizeof(float), hipMemcpyDeviceToHost, c.stream));
  PZG_HIP(hipStreamSynchronize(c.stream));
  for (int64_t i = 0; i < n; i++) y[i] = h[i];
  return PZG_OK;
}

int pzg_mg_apply(pzg_ctx *ctx, int32_t op, int32_t degree, double range, const double *r, double *z,
                 int32_t *nlev, int32_t *dims, double *lmax) {
  CTX_CHECK();
  Ctx &c = ctx->c;
  if (!c.assembled || !r || !z) {
    set_error("pzg_mg_apply: needs an assembled system and host vectors");
    return PZG_ESTATE;
  }
  const bool learned = (op & PZG_MG_LEARNED) != 0;
  op &= ~PZG_MG_LEARNED;
  if (op != PZG_INNER_BSR && op != PZG_INNER_MATFREE) {
    set_error("pzg_mg_apply: op must be PZG_INNER_BSR or PZG_INNER_MATFREE (| PZG_MG_LEARNED)");
    return PZG_EINVAL;
  }
  if (learned && !c.les_ready) {
    set_error("pzg_mg_apply: learned smoother not set up in this context");
    return PZG_ESTATE;
  }
  if (op == PZG_INNER_MATFREE && !mf_available(c)) {
    set_error("pzg_mg_apply: matrix-free operator needs a mesh-built system");
    return PZG_ESTATE;
  }
  const int64_t n = c.neq;
  if (!c.full_valid) PZG_TRY(launch_expand_full(c));
  if (!c.bsr_valid) PZG_TRY(launch_build_bsr(c));
  if (op == PZG_INNER_MATFREE) PZG_TRY(mf_prepare(c));
  PZG_TRY(ensure_reduce_pub(c));
  c.mg_degree = degree > 0 ? degree : 2;
  c.mg_range = range > 0 ? range : 30.;
  const bool use_mf = op == PZG_INNER_MATFREE;
  if (!c.mg_ready || c.mg_use_mf != use_mf || c.mg_les != learned) {
    c.mg_les = learned;  // level-0 smoothing: learned element smoother or block Jacobi
    PZG_TRY(mg_setup(c, use_mf));
  }
  std::vector<float> h(n);
  for (int64_t i = 0; i < n; i++) h[i] = (float)r[i];
  PZG_TRY(dev_alloc(c.d_rs, n * sizeof(float)));
  PZG_TRY(dev_alloc(c.d_zs, n * sizeof(float)));
  PZG_HIP(hipMemcpyAsync(c.d_rs.p, h.data(), n * sizeof(float), hipMemcpyHostToDevice, c.stream));
  PZG_TRY(mg_apply(c, c.d_rs.as<float>(), c.d_zs.as<float>(), nullptr, nullptr, nullptr, nullptr));
  PZG_HIP(hipMemcpyAsync(h.data(), c.d_zs.p, n * sizeof(float), hipMemcpyDeviceToHost, c.stream));
  PZG_HIP(hipStreamSynchronize(c.stream));
  for (int64_t i = 0; i < n; i++) z[i] = h[i];
  const int nl = (int)std::min<size_t>(c.mg.size(), 15);
  if (nlev) *nlev = (int32_t)c.mg.size();
  if (dims)
    for (int l = 0; l < nl; l++)
      for (int d = 0; d < 3; d++) dims[3 * l + d] = c.mg[l].n[d];
  if (lmax) {
    lmax[0] = c.mg_lmax0;
    for (int l = 0; l < nl; l++) lmax[1 + l] = c.mg[l].lmax;
  }
  return PZG_OK;
}

int pzg_mesh_set_coef(pzg_ctx *ctx, const double *vol_coef) {
  CTX_CHECK();
  Ctx &c = ctx->c;
  if (!c.has_mesh || !vol_coef) {
    set_error("pzg_mesh_set_coef: no mesh or null coefficients");
    return PZG_ESTATE;
  }
  c.vol_coef.assign(vol_coef, vol_coef + c.nvol);
  PZG_TRY(dev_upload(c.d_vcoef, c.vol_coef.data(), c.vol_coef.size(), c.stream));
  PZG_HIP(hipStreamSynchronize(c.stream));
  c.assembled = false;
  c.full_valid = c.bsr_valid = c.scal_valid = false;
  c.has_solution = c.f_valid = false;
  return PZG_OK;
}

int pzg_learned_setup(pzg_ctx *ctx, uint64_t seed) {
  CTX_CHECK();
  return les_setup(ctx->c, seed);
}

int pzg_learned_collect(pzg_ctx *ctx, int32_t nrhs, uint64_t seed) {
  CTX_CHECK();
  return les_collect(ctx->c, nrhs, seed);
}

int pzg_learned_train(pzg_ctx *ctx, int32_t steps, double lr, double *loss_first, double *loss_last) {
  CTX_CHECK();
  Ctx &c = ctx->c;
  if (steps < 0 || !(lr > 0.)) {
    set_error("pzg_learned_train: invalid arguments");
    return PZG_EINVAL;
  }
  double l = 0.;
  PZG_TRY(les_fit_G(c));  // G: least squares on every pair collected so far
  for (int t = 0; t < steps; t++) {
    PZG_TRY(les_loss_grad(c, t == 0 ? &l : nullptr));
    if (t == 0 && loss_first) *loss_first = l;
    PZG_TRY(les_adam_step(c, lr));
  }
  PZG_TRY(les_loss_grad(c, &l));
  if (loss_last) *loss_last = l;
  if (steps == 0 && loss_first) *loss_first = l;
  return PZG_OK;
}

int pzg_learned_weights(pzg_ctx *ctx, float *get, const float *set, int32_t *n) {
  CTX_CHECK();
  Ctx &c = ctx->c;
  if (!c.les_ready) {
    set_error("pzg_learned_weights: call pzg_learned_setup first");
    return PZG_ESTATE;
  }
  const int np = les_nparams(c.les_ne);
  if (n) *n = np;
  if (set) {
    PZG_HIP(hipMemcpyAsync(c.d_les_th.p, set, sizeof(float) * np, hipMemcpyHostToDevice, c.stream));
    c.mg_ready = false;
  }
  if (get) PZG_HIP(hipMemcpyAsync(get, c.d_les_th.p, sizeof(float) * np, hipMemcpyDeviceToHost, c.stream));
  PZG_HIP(hipStreamSynchronize(c.stream));
  return PZG_OK;
}

int pzg_learned_graph(pzg_ctx *ctx, int64_t *nel, int64_t *cia, int32_t *cja, float *features,
                      float *pen) {
  CTX_CHECK();
  Ctx &c = ctx->c;
  if (!c.les_ready || !c.assembled) {
    set_error("pzg_learned_graph: needs pzg_learned_setup and an assembled system");
    return PZG_ESTATE;
  }
  PZG_TRY(les_prepare(c));
  if (nel) *nel = c.nvol;
  if (cia) std::copy(c.les_cia.begin(), c.les_cia.end(), cia);
  if (cja) std::copy(c.les_cja.begin(), c.les_cja.end(), cja);
  if (pen) std::copy(c.les_pen.begin(), c.les_pen.end(), pen);
  if (features) {
    PZG_HIP(hipMemcpyAsync(features, c.d_les_X.p, sizeof(float) * 4 * c.nvol, hipMemcpyDeviceToHost, c.stream));
    PZG_HIP(hipStreamSynchronize(c.stream));
  }
  return PZG_OK;
}

int pzg_learned_samples(pzg_ctx *ctx, int32_t *ne, int32_t *nr, float *R, float *E, float *mask) {
  CTX_CHECK();
  Ctx &c = ctx->c;
  if (ne) *ne = c.les_ne;
  if (nr) *nr = (int32_t)c.les_nr;
  if (!c.les_ready || c.les_nr == 0) {
    if (R || E || mask) {
      set_error("pzg_learned_samples: no training pairs");
      return PZG_ESTATE;
    }
    return PZG_OK;
  }
  const size_t n = (size_t)c.nvol * c.les_nr * c.les_ne;
  if (R) PZG_HIP(hipMemcpyAsync(R, c.d_les_R.p, sizeof(float) * n, hipMemcpyDeviceToHost, c.stream));
  if (E) PZG_HIP(hipMemcpyAsync(E, c.d_les_E.p, sizeof(float) * n, hipMemcpyDeviceToHost, c.stream));
  if (mask)
    PZG_HIP(hipMemcpyAsync(mask, c.d_les_M.p, sizeof(float) * c.nvol * c.les_ne, hipMemcpyDeviceToHost,
                           c.stream));
  PZG_HIP(hipStreamSynchronize(c.stream));
  return PZG_OK;
}

int pzg_learned_loss_grad(pzg_ctx *ctx, double *loss, float *grad) {
  CTX_CHECK();
  Ctx &c = ctx->c;
  PZG_TRY(les_loss_grad(c, loss));
  if (grad) {
    PZG_HIP(hipMemcpyAsync(grad, c.d_les_g.p, sizeof(float) * les_nparams(c.les_ne), hipMemcpyDeviceToHost,
                           c.stream));
    PZG_HIP(hipStreamSynchronize(c.stream));
  }
  return PZG_OK;
}

int pzg_learned_apply(pzg_ctx *ctx, const double *r, double *z) {
  CTX_CHECK();
  Ctx &c = ctx->c;
  if (!c.les_ready || !c.assembled) {
    set_error("pzg_learned_apply: needs pzg_learned_setup and an assembled system");
    return PZG_ESTATE;
  }
  if (!r || !z) {
    set_error("pzg_learned_apply: null vector");
    return PZG_EINVAL;
  }
  PZG_TRY(les_prepare(c));
  const int64_t n = c.neq;
  std::vector<float> h(n);
  for (int64_t i = 0; i < n; i++) h[i] = (float)r[i];
  PZG_TRY(dev_alloc(c.d_xs, n * sizeof(float)));
  PZG_TRY(dev_alloc(c.d_zs, n * sizeof(float)));
  PZG_HIP(hipMemcpyAsync(c.d_xs.p, h.data(), n * sizeof(float), hipMemcpyHostToDevice, c.stream));
  PZG_TRY(les_apply(c, c.d_xs.as<float>(), c.d_zs.as<float>(), nullptr));
  PZG_HIP(hipMemcpyAsync(h.data(), c.d_zs.p, n * sizeof(float), hipMemcpyDeviceToHost, c.stream));
  PZG_HIP(hipStreamSynchronize(c.stream));
  for (int64_t i = 0; i < n; i++) z[i] = h[i];
  return PZG_OK;
}

int pzg_learned_scales(pzg_ctx *ctx, float *s) {
  CTX_CHECK();
  Ctx &c = ctx->c;
  if (!c.les_ready || !c.assembled) {
    set_error("pzg_learned_scales: needs pzg_learned_setup and an assembled system");
    return PZG_ESTATE;
  }
  PZG_TRY(les_prepare(c));
  if (s) {
    PZG_HIP(hipMemcpyAsync(s, c.d_les_s.p, sizeof(float) * c.nvol, hipMemcpyDeviceToHost, c.stream));
    PZG_HIP(hipStreamSynchronize(c.stream));
  }
  return PZG_OK;
}

}  // extern "C"
