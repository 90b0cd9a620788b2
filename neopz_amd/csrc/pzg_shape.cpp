// pzg_shape.cpp -- reference-element tables for the element kernels.
//
// H1 hierarchical shape functions of order p <= 2 on line/quadrilateral/
// hexahedron, evaluated once at the Gauss-Legendre points of the element's
// integration rule and uploaded to the device.
//
// Side numbering and side->corner tables must be NeoPZ's (they fix connect
// order and therefore equation numbering): Topology/tpzquadrilateral.h:323-324,
// Topology/tpzcube.h:271-275.  The shape functions are NeoPZ's p <= 2
// generating functions (Shape/TPZShapeH1.cpp:40-116 with
// Shape/pzshapequad.cpp:67-98, Shape/pzshapecube.cpp:93-160), written here in
// their closed tensor form: a side function is the product over coordinate
// directions of 4*l0*l1 (directions the side spans) or l0 / l1 (directions in
// which the side is fixed at -1 / +1), l0 = (1-xi)/2, l1 = (1+xi)/2.
#include <cmath>
#include <cstring>

#include "pzg_internal.h"

namespace pzg {

static const int QUAD_EDGE[4][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}};
static const int HEX_EDGE[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {0, 4}, {1, 5},
                                    {2, 6}, {3, 7}, {4, 5}, {5, 6}, {6, 7}, {7, 4}};
static const int HEX_FACE[6][4] = {{0, 1, 2, 3}, {0, 1, 5, 4}, {1, 2, 6, 5},
                                   {3, 2, 6, 7}, {0, 3, 7, 4}, {4, 5, 6, 7}};
// corner coordinates in {0,1}^dim (0 -> xi=-1, 1 -> xi=+1)
static const int CORNER_BITS[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                      {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};

// Triangle / tetrahedron sides (Topology/tpztriangle.h:290-291,
// Topology/tpztetrahedron.h:280-283).
static const int TRI_EDGE[3][2] = {{0, 1}, {1, 2}, {2, 0}};
static const int TET_EDGE[6][2] = {{0, 1}, {1, 2}, {2, 0}, {0, 3}, {1, 3}, {2, 3}};
static const int TET_FACE[4][3] = {{0, 1, 2}, {0, 1, 3}, {1, 2, 3}, {0, 2, 3}};

int etype_ncorner(int t) {
  return t == PZG_LINE ? 2 : t == PZG_QUAD ? 4 : t == PZG_HEX ? 8 : t == PZG_TRI ? 3 : 4;
}
int etype_nsides(int t) {
  return t == PZG_LINE ? 3 : t == PZG_QUAD ? 9 : t == PZG_HEX ? 27 : t == PZG_TRI ? 7 : 15;
}
static int etype_dim(int t) { return t == PZG_LINE ? 1 : (t == PZG_QUAD || t == PZG_TRI) ? 2 : 3; }

static int side_corner_list(int etype, int s, int *loc) {
  int nc = etype_ncorner(etype);
  if (s < nc) { loc[0] = s; return 1; }
  if (etype == PZG_TRI) {
    if (s < 6) { loc[0] = TRI_EDGE[s - 3][0]; loc[1] = TRI_EDGE[s - 3][1]; return 2; }
    for (int i = 0; i < 3; i++) loc[i] = i;
    return 3;
  }
  if (etype == PZG_TET) {
    if (s < 10) { loc[0] = TET_EDGE[s - 4][0]; loc[1] = TET_EDGE[s - 4][1]; return 2; }
    if (s < 14) { for (int i = 0; i < 3; i++) loc[i] = TET_FACE[s - 10][i]; return 3; }
    for (int i = 0; i < 4; i++) loc[i] = i;
    return 4;
  }
  if (etype == PZG_LINE) { loc[0] = 0; loc[1] = 1; return 2; }
  if (etype == PZG_QUAD) {
    if (s < 8) { loc[0] = QUAD_EDGE[s - 4][0]; loc[1] = QUAD_EDGE[s - 4][1]; return 2; }
    for (int i = 0; i < 4; i++) loc[i] = i;
    return 4;
  }
  if (s < 20) { loc[0] = HEX_EDGE[s - 8][0]; loc[1] = HEX_EDGE[s - 8][1]; return 2; }
  if (s < 26) { for (int i = 0; i < 4; i++) loc[i] = HEX_FACE[s - 20][i]; return 4; }
  for (int i = 0; i < 8; i++) loc[i] = i;
  return 8;
}

int side_corner_list_pub(int etype, int s, int *loc) { return side_corner_list(etype, s, loc); }

// Newton iteration on the Legendre recurrence in long double; points stored
// as (-z_i, +z_i) pairs from the largest root down, which is NeoPZ's order
// (Integral/tpzgaussrule.cpp:170-240); n = (int)(0.51*(order+2)).
int gauss_legendre(int order, double *loc, double *w) {
  int n = (int)(0.51 * (order + 2));
  if (n < 1) n = 1;
  const long double pi = 3.14159265358979323846264338327950288L;
  for (int i = 0; i < (n + 1) / 2; i++) {
    long double z = cosl(pi * ((long double)i + 0.75L) / ((long double)n + 0.5L)), dp = 1;
    for (int it = 0; it < 100; it++) {
      long double pm1 = 1.0L, pm2 = 0.0L;
      for (int j = 1; j <= n; j++) {
        long double pj = ((2.0L * j - 1.0L) * z * pm1 - (j - 1.0L) * pm2) / (long double)j;
        pm2 = pm1;
        pm1 = pj;
      }
      dp = (long double)n * (z * pm1 - pm2) / (z * z - 1.0L);
      long double dz = pm1 / dp;
      z -= dz;
      if (fabsl(dz) <= 1e-19L) break;
    }
    long double wt = 2.0L / ((1.0L - z * z) * dp * dp);
    loc[2 * i] = (double)(-z);
    w[2 * i] = (double)wt;
    if (2 * i + 1 < n) {
      loc[2 * i + 1] = (double)z;
      w[2 * i + 1] = (double)wt;
    }
  }
  return n;
}

// Symmetric rules on the reference triangle (area 1/2) and tetrahedron
// (volume 1/6): degree 2 with 3 / 4 points, degree 4 (triangle, 6 points) and
// 5 (tetrahedron, 14 points); orbit parameters from the moment equations.
// Affine simplices with p <= 2 and element-wise constant data have polynomial
// integrands of degree <= 2p, which these rules integrate exactly.
static int simplex_rule(int etype, int order, double (*pts)[3], double *w) {
  int n = 0;
  if (etype == PZG_TRI) {
    if (order <= 2) {
      const double a = 1. / 6., b = 2. / 3.;
      const double P[3][2] = {{a, a}, {b, a}, {a, b}};
      for (int i = 0; i < 3; i++) { pts[i][0] = P[i][0]; pts[i][1] = P[i][1]; pts[i][2] = 0; w[i] = 1. / 6.; }
      return 3;
    }
    if (order > 4) return -1;
    const double A[2] = {0.4459484909159655, 0.0915762135097718};
    const double W[2] = {0.1116907948390057, 0.05497587182766096};
    for (int o = 0; o < 2; o++) {
      const double a = A[o], b = 1. - 2. * a;
      const double L[3][3] = {{b, a, a}, {a, b, a}, {a, a, b}};
      for (int k = 0; k < 3; k++, n++) { pts[n][0] = L[k][1]; pts[n][1] = L[k][2]; pts[n][2] = 0; w[n] = W[o]; }
    }
    return n;
  }
  if (order <= 2) {
    const double a = 0.1381966011250105, b = 1. - 3. * a;
    for (int k = 0; k < 4; k++, n++) {
      double L[4] = {a, a, a, a};
      L[k] = b;
      pts[n][0] = L[1]; pts[n][1] = L[2]; pts[n][2] = L[3]; w[n] = 1. / 24.;
    }
    return 4;
  }
  if (order > 5) return -1;
  const double A[2] = {0.09273525031089165, 0.31088591926330106};
  const double W[2] = {0.0122488405193938, 0.018781320953003007};
  for (int o = 0; o < 2; o++)
    for (int k = 0; k < 4; k++, n++) {
      double L[4] = {A[o], A[o], A[o], A[o]};
      L[k] = 1. - 3. * A[o];
      pts[n][0] = L[1]; pts[n][1] = L[2]; pts[n][2] = L[3]; w[n] = W[o];
    }
  const double a3 = 0.045503704125647394, w3 = 0.007091003462846565, b3 = 0.5 - a3;
  for (int i = 0; i < 4; i++)
    for (int j = i + 1; j < 4; j++, n++) {
      double L[4] = {b3, b3, b3, b3};
      L[i] = a3;
      L[j] = a3;
      pts[n][0] = L[1]; pts[n][1] = L[2]; pts[n][2] = L[3]; w[n] = w3;
    }
  return n;
}

// Values and reference gradients of all p<=2 shapes at xi.
static int eval_shapes(int etype, int p, const double *xi, double *phi, double dphi[3][MAXSH]) {
  if (etype == PZG_TRI || etype == PZG_TET) {
    // barycentric corners; p = 2 adds 4 * la * lb per edge
    // (Shape/pzshapetriang.cpp:34-81, Shape/pzshapetetra.cpp:53-150)
    const int dim = etype_dim(etype), nc = dim + 1, ne = etype == PZG_TRI ? 3 : 6;
    double l[4], dl[3][4];
    l[0] = 1.;
    for (int d = 0; d < dim; d++) { l[0] -= xi[d]; l[d + 1] = xi[d]; }
    for (int d = 0; d < dim; d++)
      for (int k = 0; k < nc; k++) dl[d][k] = k == 0 ? -1. : (k == d + 1 ? 1. : 0.);
    for (int k = 0; k < nc; k++) {
      phi[k] = l[k];
      for (int d = 0; d < dim; d++) dphi[d][k] = dl[d][k];
    }
    if (p < 2) return nc;
    for (int e = 0; e < ne; e++) {
      const int a = etype == PZG_TRI ? TRI_EDGE[e][0] : TET_EDGE[e][0];
      const int b = etype == PZG_TRI ? TRI_EDGE[e][1] : TET_EDGE[e][1];
      phi[nc + e] = 4. * l[a] * l[b];
      for (int d = 0; d < dim; d++) dphi[d][nc + e] = 4. * (dl[d][a] * l[b] + l[a] * dl[d][b]);
    }
    return nc + ne;
  }
  int dim = etype == PZG_LINE ? 1 : etype == PZG_QUAD ? 2 : 3;
  int nc = etype == PZG_LINE ? 2 : etype == PZG_QUAD ? 4 : 8;
  int nsides = etype == PZG_LINE ? 3 : etype == PZG_QUAD ? 9 : 27;
  double f[3][3], df[3][3];  // per direction: [l0, l1, bubble]
  for (int d = 0; d < dim; d++) {
    double l0 = (1. - xi[d]) / 2., l1 = (1. + xi[d]) / 2.;
    f[d][0] = l0; df[d][0] = -0.5;
    f[d][1] = l1; df[d][1] = 0.5;
    f[d][2] = 4. * l0 * l1; df[d][2] = 4. * (-0.5 * l1 + 0.5 * l0);
  }
  int ns = 0;
  for (int s = 0; s < nsides; s++) {
    if (s >= nc && p < 2) break;
    int loc[8], nl = side_corner_list(etype, s, loc);
    int kind[3];
    for (int d = 0; d < dim; d++) {
      int b0 = CORNER_BITS[loc[0]][d], same = 1;
      for (int k = 1; k < nl; k++) same &= (CORNER_BITS[loc[k]][d] == b0);
      kind[d] = same ? b0 : 2;
    }
    double v = 1.;
    for (int d = 0; d < dim; d++) v *= f[d][kind[d]];
    phi[ns] = v;
    for (int g = 0; g < dim; g++) {
      double dv = 1.;
      for (int d = 0; d < dim; d++) dv *= (d == g) ? df[d][kind[d]] : f[d][kind[d]];
      dphi[g][ns] = dv;
    }
    ns++;
  }
  return ns;
}

int build_tables(int etype, int p, int rule_order, ElemTables &t) {
  std::memset(&t, 0, sizeof(t));
  t.etype = etype;
  t.dim = etype_dim(etype);
  t.ncorner = etype_ncorner(etype);
  t.nsides = etype_nsides(etype);
  const bool simplex = etype == PZG_TRI || etype == PZG_TET;
  double gl[64], gw[64], spts[64][3], sw[64];
  int ng = 0, nq = 1;
  if (simplex) {
    nq = simplex_rule(etype, rule_order, spts, sw);
    if (nq < 0) {
      set_error("integration rule order not available on simplices");
      return PZG_EUNSUP;
    }
  } else {
    ng = gauss_legendre(rule_order, gl, gw);
    for (int d = 0; d < t.dim; d++) nq *= ng;
  }
  if (nq > MAXQ) {
    set_error("integration rule too large for the element tables");
    return PZG_EUNSUP;
  }
  t.nq = nq;
  for (int q = 0; q < nq; q++) {
    double xi[3] = {0, 0, 0};
    if (simplex) {
      xi[0] = spts[q][0]; xi[1] = spts[q][1]; xi[2] = spts[q][2];
      t.w[q] = sw[q];
    } else if (t.dim == 1) {
      xi[0] = gl[q];
      t.w[q] = gw[q];
    } else if (t.dim == 2) {  // Integral/pzquad.cpp:159-169: ksi slow, eta fast
      int ik = q / ng, ie = q % ng;
      xi[0] = gl[ik]; xi[1] = gl[ie];
      t.w[q] = gw[ik] * gw[ie];
    } else {                  // Integral/pzquad.cpp:274-287: ksi fastest
      int ik = q % ng, ie = (q % (ng * ng)) / ng, iz = q / (ng * ng);
      xi[0] = gl[ik]; xi[1] = gl[ie]; xi[2] = gl[iz];
      t.w[q] = gw[ik] * gw[ie] * gw[iz];
    }
    double dphi[3][MAXSH], cd[3][MAXSH], cphi[MAXSH];
    t.nsh = eval_shapes(etype, p, xi, t.phi[q], dphi);
    for (int d = 0; d < t.dim; d++)
      for (int i = 0; i < t.nsh; i++) t.dphi[q][d][i] = dphi[d][i];
    eval_shapes(etype, 1, xi, cphi, cd);
    for (int n = 0; n < t.ncorner; n++) {
      t.cphi[q][n] = cphi[n];
      for (int d = 0; d < t.dim; d++) t.cdphi[q][d][n] = cd[d][n];
    }
  }
  for (int i = 0; i < t.nsh; i++) t.shape_side[i] = i;  // p<=2: shape i lives on side i
  return PZG_OK;
}

}  // namespace pzg
