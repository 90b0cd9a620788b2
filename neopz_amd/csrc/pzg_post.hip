// pzg_post.hip -- post-processing of the device solution.
//
//  * k_post<NS, POINTS>: the error norms of TPZInterpolationSpace::
//    EvaluateError (Mesh/pzinterpolationspace.cpp:1108-1207): per volume
//    element the weighted sum over the error rule's points of the material's
//    Errors values, then its square root.  A group of G lanes (G = the power
//    of two >= the rule's point count, at most 64) takes one element: the
//    element's solution coefficients, table columns and corner coordinates
//    go to LDS, each lane walks the points q = lane, lane + G, ... of tables
//    stored point-fastest (coalesced loads), and the group adds its lanes'
//    sums with cross-lane shuffles.  With POINTS = true the kernel writes the
//    physical coordinates of the points instead (the points at which a
//    caller evaluates its exact-solution function).
//
// Elements of one rule bucket share one table; a launch covers one bucket.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>

#include "pzg_internal.h"

namespace pzg {

constexpr int PTPB = 256;
constexpr int MAXERR = 6;

struct PostArgs {
  const double *xyz;
  const int64_t *nodes, *conn, *aidx, *off;
  const int32_t *list, *slot;  // slot: col_off[HP_MAXSH] | side[HP_MAXSH]
  const double *coef, *u, *tab, *uq, *duq;
  const int64_t *dep_ptr, *dep_blk;  // constrained connects (may be null)
  const double *dep_coef;
  double *out, *xq;
  int64_t nlist;
  int nq, ncol, nslot, nsv, nc, dim, bs, G, hp, tet, material, kind, nerr;
  double center[3], cst, E, nu, kconst;
};

// Scalar exact solutions of TLaplaceExample1::uxy (Pre/TPZAnalyticSolution.
// cpp:1254-1310) and their gradients (graduxy, :1927-1943, by automatic
// differentiation there, in closed form here).
__device__ inline void exact_dev(const PostArgs &a, const double x[3], double &ue, double ge[3]) {
  const double PI = 3.14159265358979323846;
  const double xl[3] = {x[0] - a.center[0], x[1] - a.center[1], x[2] - a.center[2]};
  ge[0] = ge[1] = ge[2] = 0.;
  ue = 0.;
  if (a.kind == PZG_EXACT_CONST) {
    ue = a.cst;
  } else if (a.kind == PZG_EXACT_X) {
    ue = xl[0];
    ge[0] = 1.;
  } else {
    const bool sn = a.kind == PZG_EXACT_SINSIN;
    double f[3], df[3];
    for (int i = 0; i < a.dim; i++) {
      double s, c;
      if (sn) {
        sincos(PI * xl[i], &s, &c);
        f[i] = s;
        df[i] = PI * c;
      } else {
        sincos(PI * xl[i] / 2, &s, &c);
        f[i] = c;
        df[i] = -(PI / 2) * s;
      }
    }
    ue = 1.;
    for (int i = 0; i < a.dim; i++) ue *= f[i];
    for (int j = 0; j < a.dim; j++) {
      double g = df[j];
      for (int i = 0; i < a.dim; i++)
        if (i != j) g *= f[i];
      ge[j] = g;
    }
  }
}

// The materials' Errors at one point; u/g: discrete value and gradient
// (g[i][k] = d u_k / d x_i), ue/ge the exact ones.
//   TPZMatPoisson::Errors (Material/Poisson/TPZMatPoisson.cpp:173-218)
//   TPZDarcyFlow::Errors (Material/DarcyFlow/TPZDarcyFlow.cpp:268-326): the
//     semi-norm term is (sum_i K |g_i - ge_i|)^2 and the flux terms compare
//     the exact derivative with -K g, as written there
//   TPZElasticity3D::Errors (Material/Elasticity/TPZElasticity3D.cpp:
//     1081-1133) with ComputeStressVector (:1178-1192), no prestress
template <int NS>
__device__ inline void err_values(const PostArgs &a, double coef, const double (&u)[NS],
                                  const double (&g)[3][NS], const double (&ue)[NS],
                                  const double (&ge)[3][NS], double (&v)[MAXERR]) {
  for (int m = 0; m < MAXERR; m++) v[m] = 0.;
  if constexpr (NS == 1) {
    if (a.material == PZG_MAT_POISSON) {
      const double d = u[0] - ue[0];
      v[1] = d * d;
      for (int i = 0; i < a.dim; i++) {
        const double t = g[i][0] - ge[i][0];
        v[2] += t * t;
      }
      v[0] = v[1] + v[2];
    } else {
      const double d = fabs(u[0] - ue[0]);
      v[1] = d * d;
      double s = 0.;
      for (int i = 0; i < a.dim; i++) s += coef * fabs(g[i][0] - ge[i][0]);
      v[2] = fabs(s * s);
      v[0] = v[1] + v[2];
      for (int i = 0; i < a.dim; i++) {
        const double t = fabs(ge[i][0] - (-coef * g[i][0]));
        v[3 + i] = t * t;
      }
    }
  } else {
    double l2 = 0., semi = 0., D[3][3];
    for (int k = 0; k < 3; k++) l2 += (u[k] - ue[k]) * (u[k] - ue[k]);
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) {
        D[i][j] = g[i][j] - ge[i][j];
        semi += D[i][j] * D[i][j];
      }
    const double E = coef, ni = a.nu;
    const double c1 = -1. + ni, c2 = -1. + ni + 2. * ni * ni, c3 = 2. * (1. + ni);
    const double s0 = E * (D[0][0] * c1 - (D[1][1] + D[2][2]) * ni) / c2;
    const double s1 = E * (D[1][1] * c1 - (D[0][0] + D[2][2]) * ni) / c2;
    const double s2 = E * (D[2][2] * c1 - (D[0][0] + D[1][1]) * ni) / c2;
    const double s3 = E * (D[1][0] + D[0][1]) / c3;
    const double s4 = E * (D[2][0] + D[0][2]) / c3;
    const double s5 = E * (D[2][1] + D[1][2]) / c3;
    const double S[3][3] = {{s0, s3, s4}, {s3, s1, s5}, {s4, s5, s2}};
    double en = 0.;
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) en += D[i][j] * S[i][j];
    v[0] = en;
    v[1] = l2;
    v[2] = semi;
  }
}

template <int NS, bool POINTS>
__global__ __launch_bounds__(PTPB) void k_post(const PostArgs a) {
  extern __shared__ double s_mem[];
  const int G = a.G, ngrp = PTPB / G;
  const int gl = threadIdx.x / G, lq = threadIdx.x - gl * G;
  const int64_t gi = (int64_t)blockIdx.x * ngrp + gl;
  const bool live = gi < a.nlist;
  const int64_t e = live ? a.list[gi] : 0;
  double *s_u = s_mem + (size_t)gl * a.nslot * NS;
  double *s_x = s_mem + (size_t)ngrp * a.nslot * NS + gl * 24;
  int *s_col = reinterpret_cast<int *>(s_mem + (size_t)ngrp * (a.nslot * NS + 24)) + gl * a.nslot;
  if (live) {
    int64_t id[4] = {0, 0, 0, 0};
    if (a.hp)
      for (int k = 0; k < a.nc; k++) id[k] = a.nodes[e * a.nc + k];
    for (int i = lq; i < a.nslot; i += G) {
      s_col[i] = a.hp ? a.slot[i] + hp_variant(a.tet, a.slot[HP_MAXSH + i], id) : i;
      if (!POINTS) {
        // a constrained connect's value follows from the connects it depends
        // on (TPZCompMesh::LoadSolution expands the solution)
        const int64_t cc = a.conn[e * a.nsv + i], b = a.aidx[cc];
        for (int k = 0; k < NS; k++) {
          double v = 0.;
          if (b >= 0) v = a.u[b * a.bs + k];
          else if (a.dep_ptr)
            for (int64_t d = a.dep_ptr[cc]; d < a.dep_ptr[cc + 1]; d++)
              if (a.dep_blk[d] >= 0) v += a.dep_coef[d] * a.u[a.dep_blk[d] * a.bs + k];
          s_u[i * NS + k] = v;
        }
      }
    }
    for (int k = lq; k < a.nc * 3; k += G) s_x[k] = a.xyz[3 * a.nodes[e * a.nc + k / 3] + k % 3];
  }
  __syncthreads();
  if (!live) return;

  const int nq = a.nq, dim = a.dim, nc = a.nc, ncol = a.ncol;
  const double *W = a.tab, *PH = W + nq, *DPH = PH + (size_t)ncol * nq;
  const double *CP = DPH + (size_t)dim * ncol * nq, *CD = CP + (size_t)nc * nq;
  const double coef = a.coef ? a.coef[e] : (a.material == PZG_MAT_ELAST3D ? a.E : a.kconst);
  double acc[MAXERR] = {0., 0., 0., 0., 0., 0.};
  for (int q = lq; q < nq; q += G) {
    // geometric map: x and J[i][j] = d x_i / d xi_j (TPZGeoEl::X / GradX)
    double x[3] = {0., 0., 0.}, J[3][3] = {{0., 0., 0.}, {0., 0., 0.}, {0., 0., 0.}};
    for (int n = 0; n < nc; n++) {
      const double cp = CP[(size_t)n * nq + q];
      for (int i = 0; i < 3; i++) x[i] += cp * s_x[3 * n + i];
      for (int j = 0; j < dim; j++) {
        const double cd = CD[((size_t)j * nc + n) * nq + q];
        for (int i = 0; i < 3; i++) J[i][j] += cd * s_x[3 * n + i];
      }
    }
    if constexpr (POINTS) {
      const int64_t p = a.off[e] + q;
      for (int i = 0; i < 3; i++) a.xq[p * 3 + i] = x[i];
      continue;
    }
    // discrete value and reference gradient (TPZInterpolationSpace::
    // ComputeSolution)
    double u[NS], gr[3][NS];
    for (int k = 0; k < NS; k++) {
      u[k] = 0.;
      for (int d = 0; d < 3; d++) gr[d][k] = 0.;
    }
    for (int i = 0; i < a.nslot; i++) {
      const int col = s_col[i];
      const double ph = PH[(size_t)col * nq + q];
      double dp[3] = {0., 0., 0.};
      for (int d = 0; d < dim; d++) dp[d] = DPH[((size_t)d * ncol + col) * nq + q];
      for (int k = 0; k < NS; k++) {
        const double cf = s_u[i * NS + k];
        u[k] += ph * cf;
        for (int d = 0; d < 3; d++) gr[d][k] += dp[d] * cf;
      }
    }
    // physical gradient (Axes2XYZ of the axes gradient, i.e. the gradient in
    // x, y, z) and |det J| (the element's volume / area / length measure)
    double g[3][NS], det;
    if (dim == 3) {
      det = -J[0][2] * J[1][1] * J[2][0] + J[0][1] * J[1][2] * J[2][0] + J[0][2] * J[1][0] * J[2][1] -
            J[0][0] * J[1][2] * J[2][1] - J[0][1] * J[1][0] * J[2][2] + J[0][0] * J[1][1] * J[2][2];
      double ji[3][3];  // inverse of J: ji[d][i] = d xi_d / d x_i
      ji[0][0] = (-J[1][2] * J[2][1] + J[1][1] * J[2][2]) / det;
      ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / det;
      ji[0][2] = (-J[0][2] * J[1][1] + J[0][1] * J[1][2]) / det;
      ji[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) / det;
      ji[1][1] = (-J[0][2] * J[2][0] + J[0][0] * J[2][2]) / det;
      ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / det;
      ji[2][0] = (-J[1][1] * J[2][0] + J[1][0] * J[2][1]) / det;
      ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / det;
      ji[2][2] = (-J[0][1] * J[1][0] + J[0][0] * J[1][1]) / det;
      for (int i = 0; i < 3; i++)
        for (int k = 0; k < NS; k++) g[i][k] = ji[0][i] * gr[0][k] + ji[1][i] * gr[1][k] + ji[2][i] * gr[2][k];
    } else if (dim == 2) {
      // surface element: g = J (J^T J)^-1 grad_xi, det = sqrt(det J^T J)
      const double g11 = J[0][0] * J[0][0] + J[1][0] * J[1][0] + J[2][0] * J[2][0];
      const double g12 = J[0][0] * J[0][1] + J[1][0] * J[1][1] + J[2][0] * J[2][1];
      const double g22 = J[0][1] * J[0][1] + J[1][1] * J[1][1] + J[2][1] * J[2][1];
      const double dg = g11 * g22 - g12 * g12;
      det = sqrt(dg);
      for (int k = 0; k < NS; k++) {
        const double h0 = (g22 * gr[0][k] - g12 * gr[1][k]) / dg;
        const double h1 = (-g12 * gr[0][k] + g11 * gr[1][k]) / dg;
        for (int i = 0; i < 3; i++) g[i][k] = J[i][0] * h0 + J[i][1] * h1;
      }
    } else {
      const double l2 = J[0][0] * J[0][0] + J[1][0] * J[1][0] + J[2][0] * J[2][0];
      det = sqrt(l2);
      for (int k = 0; k < NS; k++)
        for (int i = 0; i < 3; i++) g[i][k] = J[i][0] * gr[0][k] / l2;
    }
    // exact solution at x
    double ue[NS], ge[3][NS];
    if (a.kind == PZG_EXACT_POINTS) {
      const int64_t p = a.off[e] + q;
      for (int k = 0; k < NS; k++) {
        ue[k] = a.uq[p * NS + k];
        for (int d = 0; d < 3; d++) ge[d][k] = d < dim ? a.duq[(p * dim + d) * NS + k] : 0.;
      }
    } else {
      double g1[3], u1;
      exact_dev(a, x, u1, g1);
      for (int k = 0; k < NS; k++) {
        ue[k] = u1;
        for (int d = 0; d < 3; d++) ge[d][k] = g1[d];
      }
    }
    double v[MAXERR];
    err_values<NS>(a, coef, u, g, ue, ge, v);
    const double wdet = W[q] * fabs(det);
    for (int m = 0; m < MAXERR; m++) acc[m] += wdet * v[m];
  }
  if constexpr (POINTS) return;
  for (int o = G >> 1; o > 0; o >>= 1)
    for (int m = 0; m < MAXERR; m++) acc[m] += __shfl_xor(acc[m], o, 64);
  if (lq == 0)
    for (int m = 0; m < a.nerr; m++) a.out[e * a.nerr + m] = sqrt(acc[m]);
}

// ---------------------------------------------------------------------------
// Error-rule tables of every rule bucket for an exact solution of order
// porder, the element lists of the buckets and the point offsets.
static int post_prepare(Ctx &c, int porder) {
  if (!c.has_mesh) {
    set_error("post-processing: no mesh");
    return PZG_ESTATE;
  }
  if (porder < 0 || porder > 1000) {
    set_error("post-processing: exact-solution order out of range");
    return PZG_EINVAL;
  }
  if (c.post_porder == porder) return PZG_OK;
  const int nbk = (int)c.vol_rules.size();
  const bool hp = c.hp_simplex;
  c.post_tab.assign(nbk, PostTables());
  std::vector<int32_t> slot((size_t)nbk * 2 * HP_MAXSH, 0);
  c.post_tab_off.assign(nbk + 1, 0);
  for (int b = 0; b < nbk; b++) {
    const int order = post_rule_order(c.vol_etype, c.vol_rules[b], porder);
    PostTables &t = c.post_tab[b];
    PZG_TRY(build_post_tables(c.vol_etype, hp ? c.hp_order : c.porder, hp, order, t));
    if (hp && t.nsh != c.vol_nsides) {
      set_error("post-processing: virtual sides do not match the shape list");
      return PZG_EINVAL;
    }
    for (int i = 0; i < t.nsh; i++) {
      slot[(size_t)b * 2 * HP_MAXSH + i] = t.col_off[i];
      slot[(size_t)b * 2 * HP_MAXSH + HP_MAXSH + i] = t.shape_side[i];
    }
    c.post_tab_off[b + 1] = c.post_tab_off[b] + (int64_t)t.d.size();
  }
  std::vector<double> tab(c.post_tab_off[nbk]);
  for (int b = 0; b < nbk; b++)
    std::memcpy(tab.data() + c.post_tab_off[b], c.post_tab[b].d.data(),
                c.post_tab[b].d.size() * sizeof(double));
  std::vector<int64_t> off(c.nvol + 1, 0);
  c.post_lptr.assign(nbk + 1, 0);
  for (int64_t e = 0; e < c.nvol; e++) {
    off[e + 1] = off[e] + c.post_tab[c.vol_bucket[e]].nq;
    c.post_lptr[c.vol_bucket[e] + 1]++;
  }
  for (int b = 0; b < nbk; b++) c.post_lptr[b + 1] += c.post_lptr[b];
  std::vector<int32_t> list(c.nvol);
  {
    std::vector<int64_t> fill(c.post_lptr.begin(), c.post_lptr.end() - 1);
    for (int64_t e = 0; e < c.nvol; e++) list[fill[c.vol_bucket[e]]++] = (int32_t)e;
  }
  hipStream_t s = c.stream;
  PZG_TRY(dev_upload(c.d_post_tab, tab.data(), tab.size(), s));
  PZG_TRY(dev_upload(c.d_post_list, list.data(), list.size(), s));
  PZG_TRY(dev_upload(c.d_post_off, off.data(), off.size(), s));
  PZG_TRY(dev_upload(c.d_post_slot, slot.data(), slot.size(), s));
  PZG_HIP(hipStreamSynchronize(s));
  c.post_npts = off[c.nvol];
  c.post_porder = porder;
  return PZG_OK;
}

// One launch per rule bucket.
template <bool POINTS>
static int post_launch(Ctx &c, PostArgs a) {
  const bool hp = c.hp_simplex;
  for (size_t b = 0; b + 1 < c.post_lptr.size(); b++) {
    const PostTables &t = c.post_tab[b];
    a.nlist = c.post_lptr[b + 1] - c.post_lptr[b];
    if (a.nlist == 0) continue;
    a.list = c.d_post_list.as<int32_t>() + c.post_lptr[b];
    a.tab = c.d_post_tab.as<double>() + c.post_tab_off[b];
    a.slot = c.d_post_slot.as<int32_t>() + b * 2 * HP_MAXSH;
    a.nq = t.nq;
    a.ncol = t.ncol;
    a.nslot = hp ? t.nsh : (t.ncol < c.vol_nsides ? t.ncol : c.vol_nsides);
    // G lanes per element (a power of two up to a wavefront), enough for the
    // element's points, and enough for the workgroup's elements to fit in LDS
    const int ns = POINTS ? 0 : c.nstate;
    auto lds_of = [&](int g) {
      return (size_t)(PTPB / g) * ((size_t)a.nslot * (ns ? ns : 1) * 8 + 24 * 8 + a.nslot * 4);
    };
    int G = 1;
    while (G < 64 && (G < t.nq || lds_of(G) > 65536)) G <<= 1;
    a.G = G;
    const int ngrp = PTPB / G;
    const size_t lds = lds_of(G);
    if (lds > 65536) {
      set_error("post-processing: element too large for the error kernel");
      return PZG_EUNSUP;
    }
    const dim3 grid((unsigned)((a.nlist + ngrp - 1) / ngrp));
    if (POINTS)
      hipLaunchKernelGGL((k_post<1, true>), grid, dim3(PTPB), lds, c.stream, a);
    else if (c.nstate == 1)
      hipLaunchKernelGGL((k_post<1, false>), grid, dim3(PTPB), lds, c.stream, a);
    else
      hipLaunchKernelGGL((k_post<3, false>), grid, dim3(PTPB), lds, c.stream, a);
    PZG_HIP(hipGetLastError());
  }
  return PZG_OK;
}

static PostArgs post_args(const Ctx &c) {
  PostArgs a;
  std::memset(&a, 0, sizeof(a));
  a.xyz = c.d_xyz.as<double>();
  a.nodes = c.d_vnodes.as<int64_t>();
  a.conn = c.d_vconn.as<int64_t>();
  a.aidx = c.d_aidx.as<int64_t>();
  a.off = c.d_post_off.as<int64_t>();
  a.coef = c.vol_coef.empty() ? nullptr : c.d_vcoef.as<double>();
  a.nsv = c.vol_nsides;
  a.nc = c.vol_ncorner;
  a.dim = c.dim;
  a.bs = c.bs;
  a.hp = c.hp_simplex ? 1 : 0;
  a.tet = c.vol_etype == PZG_TET ? 1 : 0;
  a.material = c.material;
  a.E = c.E;
  a.nu = c.nu;
  a.kconst = c.scale;
  if (c.has_dep) {
    a.dep_ptr = c.d_dep_ptr.as<int64_t>();
    a.dep_blk = c.d_dep_blk.as<int64_t>();
    a.dep_coef = c.d_dep_coef.as<double>();
  }
  return a;
}

int post_points(Ctx &c, int porder, int64_t *npts, double *xq) {
  PZG_TRY(post_prepare(c, porder));
  if (npts) *npts = c.post_npts;
  if (!xq) return PZG_OK;
  PZG_TRY(dev_alloc(c.d_post_xq, (size_t)c.post_npts * 3 * sizeof(double)));
  PostArgs a = post_args(c);
  a.xq = c.d_post_xq.as<double>();
  PZG_TRY(post_launch<true>(c, a));
  PZG_HIP(hipMemcpyAsync(xq, c.d_post_xq.p, (size_t)c.post_npts * 3 * sizeof(double),
                         hipMemcpyDeviceToHost, c.stream));
  PZG_HIP(hipStreamSynchronize(c.stream));
  return PZG_OK;
}

int post_errors(Ctx &c, const pzg_exact &ex, double *el_errors, double *errors, int *nerr_out) {
  if (!c.has_mesh) {
    set_error("pzg_evaluate_error: no mesh");
    return PZG_ESTATE;
  }
  if (!c.has_solution) {
    set_error("pzg_evaluate_error: no solution loaded (pzg_solve or pzg_load_solution)");
    return PZG_ESTATE;
  }
  const int nerr = c.material == PZG_MAT_DARCY ? 6 : 3;
  if (ex.kind < PZG_EXACT_POINTS || ex.kind > PZG_EXACT_COSCOS) {
    set_error("pzg_evaluate_error: unknown exact-solution kind");
    return PZG_EINVAL;
  }
  if (ex.kind != PZG_EXACT_POINTS && c.nstate != 1) {
    set_error("pzg_evaluate_error: the device exact solutions are scalar; use PZG_EXACT_POINTS");
    return PZG_EINVAL;
  }
  PZG_TRY(post_prepare(c, ex.porder));
  PostArgs a = post_args(c);
  a.kind = ex.kind;
  a.nerr = nerr;
  for (int d = 0; d < 3; d++) a.center[d] = ex.center[d];
  a.cst = ex.constant;
  hipStream_t s = c.stream;
  if (ex.kind == PZG_EXACT_POINTS) {
    if (ex.npts != c.post_npts || !ex.u || !ex.du) {
      set_error("pzg_evaluate_error: exact values must cover the " + std::to_string(c.post_npts) +
                " points of pzg_error_points for this order");
      return PZG_EINVAL;
    }
    const size_t nu = (size_t)c.post_npts * c.nstate, ndu = nu * c.dim;
    PZG_TRY(dev_upload(c.d_post_uq, ex.u, nu, s));
    PZG_TRY(dev_upload(c.d_post_duq, ex.du, ndu, s));
    a.uq = c.d_post_uq.as<double>();
    a.duq = c.d_post_duq.as<double>();
  }
  a.u = c.d_u.as<double>();
  PZG_TRY(dev_alloc(c.d_post_out, (size_t)c.nvol * nerr * sizeof(double)));
  a.out = c.d_post_out.as<double>();
  PZG_TRY(post_launch<false>(c, a));
  std::vector<double> el((size_t)c.nvol * nerr);
  PZG_HIP(hipMemcpyAsync(el.data(), c.d_post_out.p, el.size() * sizeof(double), hipMemcpyDeviceToHost,
                         s));
  PZG_HIP(hipStreamSynchronize(s));
  // TPZCompMesh::EvaluateError: sum of the squared element norms in element
  // order, then the square root (Mesh/pzcmesh.cpp:1497-1552)
  if (errors) {
    double tot[MAXERR] = {0., 0., 0., 0., 0., 0.};
    for (int64_t e = 0; e < c.nvol; e++)
      for (int m = 0; m < nerr; m++) tot[m] += el[e * nerr + m] * el[e * nerr + m];
    for (int m = 0; m < nerr; m++) errors[m] = std::sqrt(tot[m]);
  }
  if (el_errors) std::memcpy(el_errors, el.data(), el.size() * sizeof(double));
  if (nerr_out) *nerr_out = nerr;
  return PZG_OK;
}

}  // namespace pzg
