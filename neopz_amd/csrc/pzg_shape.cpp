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
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <vector>

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

// ---------------------------------------------------------------------------
// High-order triangles and tetrahedra (TPZShapeH1::Shape, Shape/TPZShapeH1.cpp:
// 40-116, with the side transforms of ComputeTransforms,
// Shape/pzgenericshape.cpp:12-67).  In barycentric coordinates l_k:
//   blend of a side: 4 l_a l_b (edge), 27 l_a l_b l_c (face), 54 l_0 l_1 l_2 l_3
//   (tetrahedron interior)  (Shape/pzshapetetra.cpp:101-150,
//   Shape/pzshapetriang.cpp:53-81);
//   edge (a,b): s = l_b - l_a (TransformElementToSide), negated unless
//   id_a < id_b; functions blend * T_k(s), k = 0..n-1 (Chebyshev);
//   face (a,b,c): m_b, m_c with m_k = l_k + l_o/3 for the face opposite
//   corner 0 (o = 0), m_k = l_k otherwise, m_a = 1 - m_b - m_c; with u, v the
//   m of the face corners holding the middle and the largest id
//   (GetTransformId2dT + gTrans2dT), functions blend * T_i(2u-1) T_j(2v-1) by
//   i+j ascending, then j;
//   tetrahedron interior: blend * T_i(2x-1) T_j(2y-1) T_k(2z-1), i, j, k
//   ascending with i+j+k < P-3.
// Side numbering: Topology/tpztriangle.h:290-291, Topology/tpztetrahedron.h:
// 280-283.

int simplex_nshape(int etype, int s, int p) {
  const int nc = etype == PZG_TRI ? 3 : 4;
  if (s < nc) return 1;
  int nl = etype == PZG_TRI ? (s < 6 ? 2 : 3) : (s < 10 ? 2 : s < 14 ? 3 : 4);
  if (nl == 2) return p > 1 ? p - 1 : 0;
  if (nl == 3) return p > 2 ? (p - 1) * (p - 2) / 2 : 0;
  return p > 3 ? (p - 1) * (p - 2) * (p - 3) / 6 : 0;
}

int hp_shape_list(int etype, int P, int *side, int *loc) {
  int n = 0;
  for (int s = 0; s < etype_nsides(etype); s++)
    for (int k = 0; k < simplex_nshape(etype, s, P); k++, n++) {
      if (side) side[n] = s;
      if (loc) loc[n] = k;
    }
  return n;
}

static int side_nvariants(int etype, int s) {
  int loc[8], nl = side_corner_list(etype, s, loc);
  if (nl == 2) return 2;
  if (nl == 3) return 6;
  return 1;
}

// transform id of a triangle from its three corner ids
// (TPZShapeTriang::GetTransformId2dT, Shape/pzshapetriang.cpp:410-431)
static int tri_transform_id(int64_t i0, int64_t i1, int64_t i2) {
  const int64_t id[3] = {i0, i1, i2};
  int a = id[0] < id[1] ? 0 : 1;
  int mn = id[2] < id[a] ? 2 : a;
  int n1 = (mn + 1) % 3, n2 = (mn + 2) % 3;
  return 2 * mn + (id[n1] < id[n2] ? 0 : 1);
}

static void cheb(double x, int n, double *t, double *dt) {
  if (n <= 0) return;
  t[0] = 1.;
  dt[0] = 0.;
  if (n > 1) { t[1] = x; dt[1] = 1.; }
  for (int k = 2; k < n; k++) {
    t[k] = 2.0 * x * t[k - 1] - t[k - 2];
    dt[k] = 2.0 * x * dt[k - 1] + 2.0 * t[k - 1] - dt[k - 2];
  }
}

// all shapes of order P (every side at P) for corner ids `ids` at xi;
// returns the count, dphi[d][i]
static int eval_simplex_hp(int etype, int P, const int64_t *ids, const double *xi, double *phi,
                           double dphi[3][HP_MAXSH]) {
  const int dim = etype == PZG_TRI ? 2 : 3, nc = dim + 1, nsides = etype_nsides(etype);
  double l[4], g[4][3];
  l[0] = 1.;
  for (int d = 0; d < dim; d++) { l[0] -= xi[d]; l[d + 1] = xi[d]; }
  for (int k = 0; k < 4; k++)
    for (int d = 0; d < 3; d++) g[k][d] = (d >= dim) ? 0. : (k == 0 ? -1. : (k == d + 1 ? 1. : 0.));
  int n = 0;
  for (int s = 0; s < nsides; s++) {
    const int ns = simplex_nshape(etype, s, P);
    if (!ns) continue;
    int c[8];
    const int nl = side_corner_list(etype, s, c);
    if (nl == 1) {
      phi[n] = l[c[0]];
      for (int d = 0; d < dim; d++) dphi[d][n] = g[c[0]][d];
      n++;
      continue;
    }
    const double cst = nl == 2 ? 4. : nl == 3 ? 27. : 54.;
    double B = cst, dB[3];
    for (int k = 0; k < nl; k++) B *= l[c[k]];
    for (int d = 0; d < 3; d++) {
      double acc = 0.;
      for (int k = 0; k < nl; k++) {
        double pr = g[c[k]][d];
        for (int j = 0; j < nl; j++)
          if (j != k) pr *= l[c[j]];
        acc += pr;
      }
      dB[d] = cst * acc;
    }
    double f[HP_MAXSH], df[HP_MAXSH][3];
    if (nl == 2) {
      const double sg = ids[c[0]] < ids[c[1]] ? 1. : -1.;
      double T[8], dT[8];
      cheb(sg * (l[c[1]] - l[c[0]]), ns, T, dT);
      for (int k = 0; k < ns; k++) {
        f[k] = T[k];
        for (int d = 0; d < 3; d++) df[k][d] = dT[k] * sg * (g[c[1]][d] - g[c[0]][d]);
      }
    } else if (nl == 3) {
      const int opp = (etype == PZG_TET) ? 6 - c[0] - c[1] - c[2] : -1;
      const double wo = opp == 0 ? 1. / 3. : 0.;
      double m[3], gm[3][3];
      for (int k = 1; k < 3; k++) {
        m[k] = l[c[k]] + (opp >= 0 ? wo * l[opp] : 0.);
        for (int d = 0; d < 3; d++) gm[k][d] = g[c[k]][d] + (opp >= 0 ? wo * g[opp][d] : 0.);
      }
      m[0] = 1. - m[1] - m[2];
      for (int d = 0; d < 3; d++) gm[0][d] = -gm[1][d] - gm[2][d];
      // face corners ranked by id: u <- middle, v <- largest
      int r0 = 0, r1 = 1, r2 = 2;
      if (ids[c[r1]] < ids[c[r0]]) std::swap(r0, r1);
      if (ids[c[r2]] < ids[c[r1]]) std::swap(r1, r2);
      if (ids[c[r1]] < ids[c[r0]]) std::swap(r0, r1);
      const int deg = P - 2;  // degrees 0..P-3
      double Tu[8], dTu[8], Tv[8], dTv[8];
      cheb(2. * m[r1] - 1., deg, Tu, dTu);
      cheb(2. * m[r2] - 1., deg, Tv, dTv);
      int k = 0;
      for (int tot = 0; k < ns; tot++)
        for (int j = 0; j <= tot && k < ns; j++, k++) {
          const int i = tot - j;
          f[k] = Tu[i] * Tv[j];
          for (int d = 0; d < 3; d++)
            df[k][d] = 2. * (dTu[i] * Tv[j] * gm[r1][d] + Tu[i] * dTv[j] * gm[r2][d]);
        }
    } else {
      const int ord = P - 3;
      double T[3][8], dT[3][8];
      for (int a = 0; a < 3; a++) cheb(2. * xi[a] - 1., ord, T[a], dT[a]);
      int k = 0;
      for (int i = 0; i < ord; i++)
        for (int j = 0; j < ord; j++)
          for (int q = 0; q < ord; q++) {
            if (i + j + q >= ord) continue;
            f[k] = T[0][i] * T[1][j] * T[2][q];
            df[k][0] = 2. * dT[0][i] * T[1][j] * T[2][q];
            df[k][1] = 2. * T[0][i] * dT[1][j] * T[2][q];
            df[k][2] = 2. * T[0][i] * T[1][j] * dT[2][q];
            k++;
          }
    }
    for (int k = 0; k < ns; k++, n++) {
      phi[n] = k == 0 ? B : B * f[k];
      for (int d = 0; d < dim; d++) dphi[d][n] = k == 0 ? dB[d] : dB[d] * f[k] + B * df[k][d];
    }
  }
  return n;
}

// Gauss-Jacobi points and weights, weight (1-t)^alpha on [-1,1]: eigenpairs of
// the Jacobi matrix of the monic recurrence (Golub-Welsch) by cyclic Jacobi
// rotations; points ascending.
static void gauss_jacobi(int n, double alpha, double *t, double *w) {
  double A[16][16] = {}, V[16][16] = {};
  for (int i = 0; i < n; i++) V[i][i] = 1.;
  for (int k = 0; k < n; k++) {
    const double s = 2. * k + alpha;
    A[k][k] = alpha == 0. ? 0. : -alpha * alpha / (s * (s + 2.));
    if (k) A[k][k - 1] = A[k - 1][k] = std::sqrt(4. * k * k * (k + alpha) * (k + alpha) /
                                                 (s * s * (s + 1.) * (s - 1.)));
  }
  for (int sweep = 0; sweep < 64; sweep++) {
    double off = 0.;
    for (int p = 0; p < n; p++)
      for (int q = p + 1; q < n; q++) off += A[p][q] * A[p][q];
    if (off < 1e-300) break;
    for (int p = 0; p < n; p++)
      for (int q = p + 1; q < n; q++) {
        if (A[p][q] == 0.) continue;
        const double th = (A[q][q] - A[p][p]) / (2. * A[p][q]);
        const double tt = (th >= 0. ? 1. : -1.) / (std::fabs(th) + std::sqrt(th * th + 1.));
        const double c = 1. / std::sqrt(tt * tt + 1.), sn = tt * c;
        for (int k = 0; k < n; k++) {
          const double a = A[k][p], b = A[k][q];
          A[k][p] = c * a - sn * b;
          A[k][q] = sn * a + c * b;
        }
        for (int k = 0; k < n; k++) {
          const double a = A[p][k], b = A[q][k];
          A[p][k] = c * a - sn * b;
          A[q][k] = sn * a + c * b;
        }
        for (int k = 0; k < n; k++) {
          const double a = V[k][p], b = V[k][q];
          V[k][p] = c * a - sn * b;
          V[k][q] = sn * a + c * b;
        }
      }
  }
  int idx[16];
  for (int i = 0; i < n; i++) idx[i] = i;
  std::sort(idx, idx + n, [&](int a, int b) { return A[a][a] < A[b][b]; });
  const double mu0 = std::pow(2., alpha + 1.) / (alpha + 1.);
  for (int i = 0; i < n; i++) {
    t[i] = A[idx[i]][idx[i]];
    w[i] = mu0 * V[0][idx[i]] * V[0][idx[i]];
  }
}

// Rule for high orders on the reference simplex: the symmetric rules up to
// degree 4 (triangle) / 5 (tetrahedron), else the collapsed product of
// Gauss-Jacobi rules of weights 1, (1-b), (1-c)^2 with n = (order+2)/2 points
// per direction (degree 2n-1).  Exact for the polynomial integrands of affine
// simplices with element-wise constant data, so the element matrices equal
// the reference's (Integral/tpzintrulet3d.cpp tables) up to rounding.
static int simplex_rule_hp(int etype, int order, double (*pts)[3], double *w, int cap) {
  if ((etype == PZG_TRI && order <= 4) || (etype == PZG_TET && order <= 5))
    return simplex_rule(etype, order, pts, w);
  const int n = (order + 2) / 2, dim = etype == PZG_TRI ? 2 : 3;
  int tot = 1;
  for (int d = 0; d < dim; d++) tot *= n;
  if (n > 16 || tot > cap) return -1;
  double ta[16], wa[16], tb[16], wb[16], tc[16], wc[16];
  gauss_jacobi(n, 0., ta, wa);
  gauss_jacobi(n, 1., tb, wb);
  gauss_jacobi(n, 2., tc, wc);
  int q = 0;
  for (int k = 0; k < (dim == 3 ? n : 1); k++)
    for (int j = 0; j < n; j++)
      for (int i = 0; i < n; i++, q++) {
        const double a = 0.5 * (1. + ta[i]), b = 0.5 * (1. + tb[j]);
        if (dim == 2) {
          pts[q][0] = a * (1. - b); pts[q][1] = b; pts[q][2] = 0.;
          w[q] = 0.5 * wa[i] * 0.25 * wb[j];
        } else {
          const double cc = 0.5 * (1. + tc[k]);
          pts[q][0] = a * (1. - b) * (1. - cc); pts[q][1] = b * (1. - cc); pts[q][2] = cc;
          w[q] = 0.5 * wa[i] * 0.25 * wb[j] * 0.125 * wc[k];
        }
      }
  return q;
}

// Corner ids that realise variant v of side s: edge v = 0 <-> id_a < id_b;
// face: the id ranking whose transform id is v.
static int variant_ids(int etype, int s, int v, int64_t ids[4]) {
  int c[8];
  const int nl = side_corner_list(etype, s, c);
  for (int k = 0; k < 4; k++) ids[k] = 10 + k;
  if (nl == 2) {
    ids[c[0]] = v == 0 ? 0 : 1;
    ids[c[1]] = v == 0 ? 1 : 0;
  } else if (nl == 3) {
    int perm[3] = {0, 1, 2};
    bool found = false;
    do {
      if (tri_transform_id(perm[0], perm[1], perm[2]) == v) { found = true; break; }
    } while (std::next_permutation(perm, perm + 3));
    if (!found) return PZG_EINVAL;
    for (int k = 0; k < 3; k++) ids[c[k]] = perm[k];
  }
  return PZG_OK;
}

int build_hp_tables(int etype, int P, int rule_order, HpTables &t) {
  if ((etype != PZG_TRI && etype != PZG_TET) || P < 1 || P > HP_MAXP) {
    set_error("high-order tables: triangles and tetrahedra of order 1..5 only");
    return PZG_EUNSUP;
  }
  std::memset(&t, 0, sizeof(t));
  t.etype = etype;
  t.dim = etype == PZG_TRI ? 2 : 3;
  t.ncorner = t.dim + 1;
  t.order = P;
  t.nsh = hp_shape_list(etype, P, t.shape_side, t.shape_loc);
  for (int i = 0; i < t.nsh; i++) {
    t.col_off[i] = t.ncol;
    t.ncol += side_nvariants(etype, t.shape_side[i]);
  }
  if (t.nsh > HP_MAXSH || t.ncol > HP_MAXCOL) {
    set_error("high-order tables: too many shapes");
    return PZG_EUNSUP;
  }
  std::vector<double> ptbuf(3 * HP_MAXQ), w(HP_MAXQ);
  double(*pts)[3] = reinterpret_cast<double(*)[3]>(ptbuf.data());
  const int nq = simplex_rule_hp(etype, rule_order, pts, w.data(), HP_MAXQ);
  if (nq < 0) {
    set_error("integration rule order not available on simplices");
    return PZG_EUNSUP;
  }
  t.nq = nq;
  for (int q = 0; q < nq; q++) {
    t.w[q] = w[q];
    for (int d = 0; d < 3; d++) t.pt[q][d] = pts[q][d];
  }
  for (int s = 0; s < etype_nsides(etype); s++) {
    const int nv = side_nvariants(etype, s);
    for (int v = 0; v < nv; v++) {
      int64_t ids[4];
      PZG_TRY(variant_ids(etype, s, v, ids));
      for (int q = 0; q < nq; q++) {
        double phi[HP_MAXSH], dphi[3][HP_MAXSH];
        eval_simplex_hp(etype, P, ids, pts[q], phi, dphi);
        for (int i = 0; i < t.nsh; i++) {
          if (t.shape_side[i] != s) continue;
          const int col = t.col_off[i] + v;
          t.phi[q][col] = phi[i];
          for (int d = 0; d < t.dim; d++) t.dphi[q][d][col] = dphi[d][i];
        }
      }
    }
  }
  return PZG_OK;
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

// Highest rule order of the element's rule family (TPZIntPoints::GetMaxOrder:
// Integral/pzquad.h 36 for line/quadrilateral/hexahedron, TPZIntTriang 21
// (tpzintrulet.h:72), TPZIntTetra3D 14 (tpzintrulet3d.h:72)) and the order
// TPZInterpolationSpace::EvaluateError sets for an exact solution of order
// porder (Mesh/pzinterpolationspace.cpp:1140-1160).
int post_rule_order(int etype, int rule, int porder) {
  const int mx = etype == PZG_TRI ? 21 : etype == PZG_TET ? 14 : 36;
  if (mx > porder) return rule > porder ? rule : porder;
  return mx;
}

int build_post_tables(int etype, int p, bool hp, int rule_order, PostTables &t) {
  t = PostTables();
  t.etype = etype;
  t.dim = etype_dim(etype);
  t.nc = etype_ncorner(etype);
  const bool simplex = etype == PZG_TRI || etype == PZG_TET;
  if (hp && !simplex) return PZG_EINVAL;
  // points and weights
  std::vector<double> w;
  std::vector<std::array<double, 3>> xi;
  if (simplex) {
    const int cap = 4096;
    std::vector<double> pb(3 * cap), wb(cap);
    const int nq = simplex_rule_hp(etype, rule_order, reinterpret_cast<double(*)[3]>(pb.data()),
                                   wb.data(), cap);
    if (nq < 0) {
      set_error("error integration rule order not available on simplices");
      return PZG_EUNSUP;
    }
    for (int q = 0; q < nq; q++) {
      w.push_back(wb[q]);
      xi.push_back({pb[3 * q], pb[3 * q + 1], pb[3 * q + 2]});
    }
  } else {
    double gl[64], gw[64];
    const int ng = gauss_legendre(rule_order, gl, gw);
    int nq = 1;
    for (int d = 0; d < t.dim; d++) nq *= ng;
    for (int q = 0; q < nq; q++) {  // point order as in build_tables
      if (t.dim == 1) {
        w.push_back(gw[q]);
        xi.push_back({gl[q], 0., 0.});
      } else if (t.dim == 2) {
        const int ik = q / ng, ie = q % ng;
        w.push_back(gw[ik] * gw[ie]);
        xi.push_back({gl[ik], gl[ie], 0.});
      } else {
        const int ik = q % ng, ie = (q % (ng * ng)) / ng, iz = q / (ng * ng);
        w.push_back(gw[ik] * gw[ie] * gw[iz]);
        xi.push_back({gl[ik], gl[ie], gl[iz]});
      }
    }
  }
  const int nq = (int)w.size();
  t.nq = nq;
  // columns
  if (hp) {
    int loc[HP_MAXSH];
    t.nsh = hp_shape_list(etype, p, t.shape_side, loc);
    for (int i = 0; i < t.nsh; i++) {
      t.col_off[i] = t.ncol;
      t.ncol += side_nvariants(etype, t.shape_side[i]);
    }
  } else {
    double phi[MAXSH], dphi[3][MAXSH];
    const double x0[3] = {0., 0., 0.};
    t.nsh = eval_shapes(etype, p, x0, phi, dphi);
    for (int i = 0; i < t.nsh; i++) {
      t.col_off[i] = i;
      t.shape_side[i] = i;
    }
    t.ncol = t.nsh;
  }
  const int dim = t.dim, nc = t.nc, ncol = t.ncol;
  t.d.assign((size_t)nq * (1 + ncol * (1 + dim) + nc * (1 + dim)), 0.);
  double *W = t.d.data(), *PH = W + nq, *DPH = PH + (size_t)ncol * nq;
  double *CP = DPH + (size_t)dim * ncol * nq, *CD = CP + (size_t)nc * nq;
  for (int q = 0; q < nq; q++) {
    W[q] = w[q];
    double cphi[MAXSH], cd[3][MAXSH];
    eval_shapes(etype, 1, xi[q].data(), cphi, cd);
    for (int n = 0; n < nc; n++) {
      CP[(size_t)n * nq + q] = cphi[n];
      for (int d = 0; d < dim; d++) CD[((size_t)d * nc + n) * nq + q] = cd[d][n];
    }
    if (!hp) {
      double phi[MAXSH], dphi[3][MAXSH];
      eval_shapes(etype, p, xi[q].data(), phi, dphi);
      for (int i = 0; i < ncol; i++) {
        PH[(size_t)i * nq + q] = phi[i];
        for (int d = 0; d < dim; d++) DPH[((size_t)d * ncol + i) * nq + q] = dphi[d][i];
      }
    }
  }
  if (hp) {
    for (int s = 0; s < etype_nsides(etype); s++) {
      const int nv = side_nvariants(etype, s);
      for (int v = 0; v < nv; v++) {
        int64_t ids[4];
        PZG_TRY(variant_ids(etype, s, v, ids));
        for (int q = 0; q < nq; q++) {
          double phi[HP_MAXSH], dphi[3][HP_MAXSH];
          eval_simplex_hp(etype, p, ids, xi[q].data(), phi, dphi);
          for (int i = 0; i < t.nsh; i++) {
            if (t.shape_side[i] != s) continue;
            const int col = t.col_off[i] + v;
            PH[(size_t)col * nq + q] = phi[i];
            for (int d = 0; d < dim; d++) DPH[((size_t)d * ncol + col) * nq + q] = dphi[d][i];
          }
        }
      }
    }
  }
  return PZG_OK;
}

}  // namespace pzg
