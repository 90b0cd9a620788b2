/*
 * pz_oracle.c -- CPU restatement of NeoPZ's linear H1 FEM path (checker only).
 *
 * TEST INFRASTRUCTURE: loaded only by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py.  The product path never links or calls it.
 *
 * Every function cites the reference file:line whose algorithm it restates.
 * Pinning of each function (golden vectors / closed forms / invariants) is in
 * tests/test_oracle.py; see the header comment there.
 *
 * Partially pinned: orc_shape_h1 at orders >= 3.  Span, conformity under
 * random ids, gradients and element energies are pinned; the choice of basis
 * within the space (Chebyshev degrees, which face corner maps to u and v, the
 * 1/3 weight of the face opposite corner 0) is restated from the reference's
 * source and parity of these choices is unpinned by a reference run.
 */
#include "pz_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* Topology tables (Topology/tpzquadrilateral.h:323-324, Topology/tpzcube.h: */
/* 271-275, Topology/tpzcube.cpp:49-82).  These have to be the reference's:  */
/* they fix side order, hence connect creation order and shape order.        */
/* ------------------------------------------------------------------------ */
static const int QUAD_EDGE[4][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}};
static const int HEX_EDGE[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {0, 4}, {1, 5},
                                    {2, 6}, {3, 7}, {4, 5}, {5, 6}, {6, 7}, {7, 4}};
static const int HEX_FACE[6][4] = {{0, 1, 2, 3}, {0, 1, 5, 4}, {1, 2, 6, 5},
                                   {3, 2, 6, 7}, {0, 3, 7, 4}, {4, 5, 6, 7}};

/* Triangle and tetrahedron (Topology/tpztriangle.h:290-291,
 * Topology/tpztetrahedron.h:280-283). */
static const int TRI_EDGE[3][2] = {{0, 1}, {1, 2}, {2, 0}};
static const int TET_EDGE[6][2] = {{0, 1}, {1, 2}, {2, 0}, {0, 3}, {1, 3}, {2, 3}};
static const int TET_FACE[4][3] = {{0, 1, 2}, {0, 1, 3}, {1, 2, 3}, {0, 2, 3}};

static int etype_ncorner(int t) {
  return t == ORC_LINE ? 2 : t == ORC_QUAD ? 4 : t == ORC_HEX ? 8 : t == ORC_TRI ? 3 : 4;
}
static int etype_nsides(int t) {
  return t == ORC_LINE ? 3 : t == ORC_QUAD ? 9 : t == ORC_HEX ? 27 : t == ORC_TRI ? 7 : 15;
}
static int etype_dim(int t) { return t == ORC_LINE ? 1 : (t == ORC_QUAD || t == ORC_TRI) ? 2 : 3; }
static int is_simplex(int t) { return t == ORC_TRI || t == ORC_TET; }

/* dimension of side s of element type t */
static int side_dim(int t, int s) {
  int nc = etype_ncorner(t);
  if (s < nc) return 0;
  if (t == ORC_TRI) return s < 6 ? 1 : 2;
  if (t == ORC_TET) return s < 10 ? 1 : s < 14 ? 2 : 3;
  if (t == ORC_LINE) return 1;
  if (t == ORC_QUAD) return s < 8 ? 1 : 2;
  if (s < 20) return 1;
  if (s < 26) return 2;
  return 3;
}

/* TPZShape*::NConnectShapeF (Shape/pzshapequad.cpp:468-482,
 * Shape/pzshapecube.cpp:573-584): p-1 per edge, (p-1)^2 per quad face,
 * (p-1)^3 per hex interior. */
static int nconnect_shape(int t, int s, int p) {
  int d = side_dim(t, s);
  if (d == 0) return 1;
  if (is_simplex(t)) {
    /* TPZShapeTriang/TPZShapeTetra::NConnectShapeF: edge p-1, triangle
     * (p-1)(p-2)/2, tetrahedron (p-1)(p-2)(p-3)/6 */
    if (d == 1) return p - 1 > 0 ? p - 1 : 0;
    if (d == 2) return p > 2 ? (p - 1) * (p - 2) / 2 : 0;
    return p > 3 ? (p - 1) * (p - 2) * (p - 3) / 6 : 0;
  }
  int q = p - 1;
  if (q <= 0) return 0;
  if (d == 1) return q;
  if (d == 2) return q * q;
  return q * q * q;
}

/* local corner nodes of side s */
static int side_corners(int t, int s, int *loc) {
  int nc = etype_ncorner(t);
  if (s < nc) { loc[0] = s; return 1; }
  if (t == ORC_TRI) {
    if (s < 6) { loc[0] = TRI_EDGE[s - 3][0]; loc[1] = TRI_EDGE[s - 3][1]; return 2; }
    for (int i = 0; i < 3; i++) loc[i] = i;
    return 3;
  }
  if (t == ORC_TET) {
    if (s < 10) { loc[0] = TET_EDGE[s - 4][0]; loc[1] = TET_EDGE[s - 4][1]; return 2; }
    if (s < 14) { for (int i = 0; i < 3; i++) loc[i] = TET_FACE[s - 10][i]; return 3; }
    for (int i = 0; i < 4; i++) loc[i] = i;
    return 4;
  }
  if (t == ORC_LINE) { loc[0] = 0; loc[1] = 1; return 2; }
  if (t == ORC_QUAD) {
    if (s < 8) { loc[0] = QUAD_EDGE[s - 4][0]; loc[1] = QUAD_EDGE[s - 4][1]; return 2; }
    for (int i = 0; i < 4; i++) loc[i] = i;
    return 4;
  }
  if (s < 20) { loc[0] = HEX_EDGE[s - 8][0]; loc[1] = HEX_EDGE[s - 8][1]; return 2; }
  if (s < 26) { for (int i = 0; i < 4; i++) loc[i] = HEX_FACE[s - 20][i]; return 4; }
  for (int i = 0; i < 8; i++) loc[i] = i;
  return 8;
}

/* ------------------------------------------------------------------------ */
/* Gauss-Legendre (Integral/tpzgaussrule.cpp:170-240): n = (int)(0.51*(order */
/* +2)), Newton iteration in long double from cos(pi(i+0.75)/(n+0.5)),       */
/* stored as (-z_i, +z_i) pairs.                                              */
/* ------------------------------------------------------------------------ */
int orc_gauss_legendre(int order, double *loc, double *w) {
  int n = (int)(0.51 * (order + 2));
  if (n < 1) n = 1;
  long double tol = 1.0e-18L; /* below long double eps*|z|; Newton stalls at roundoff */
  int m = (n + 1) / 2;
  for (int i = 0; i < m; i++) {
    long double p1 = (long double)i + 0.75L, p2 = (long double)n + 0.5L, p3, pp = 1, z1, z;
    z = cosl((3.14159265358979323846264338327950288L * p1) / p2);
    for (int it = 0; it < 100; it++) {
      p1 = 1.0L;
      p2 = 0.0L;
      for (int j = 0; j < n; j++) {
        p3 = p2;
        p2 = p1;
        p1 = ((2.0L * (long double)j + 1.0L) * z * p2 - ((long double)j) * p3) / ((long double)j + 1.0L);
      }
      long double den = z * z - 1.0L;
      pp = (long double)n * (z * p1 - p2) / den;
      z1 = z;
      z = z1 - p1 / pp;
      if (fabsl(z - z1) <= tol) break;
    }
    long double weight = 2.0L / ((1.0L - z * z) * pp * pp);
    loc[2 * i] = (double)(-z);
    w[2 * i] = (double)weight;
    if (2 * i + 1 < n) {
      loc[2 * i + 1] = (double)z;
      w[2 * i + 1] = (double)weight;
    }
  }
  return n;
}

/* ------------------------------------------------------------------------ */
/* Shape functions, p <= 2 (Shape/TPZShapeH1.cpp:40-116 + ShapeCorner /     */
/* ShapeGenerating of Shape/pzshapelinear.cpp:203-274,                       */
/* Shape/pzshapequad.cpp:36-98, Shape/pzshapecube.cpp:36-160).               */
/* phi[ns], dphi[d*ns + i] (d-th reference derivative of shape i).           */
/* ------------------------------------------------------------------------ */
int orc_shape(int t, int p, const double *xi, double *phi, double *dphi) {
  int dim = etype_dim(t), nc = etype_ncorner(t), nsides = etype_nsides(t);
  if (is_simplex(t)) {
    /* TPZShapeTriang/TPZShapeTetra::ShapeCorner (barycentric coordinates) and
     * ShapeGenerating: edge = 4 * lambda_a * lambda_b (Shape/pzshapetriang.cpp:
     * 34-81, Shape/pzshapetetra.cpp:53-150).  p <= 2: faces and interior carry
     * no shapes, so the set is corners then edges. */
    double l[4], dl[3][4];
    l[0] = 1.;
    for (int d = 0; d < dim; d++) { l[0] -= xi[d]; l[d + 1] = xi[d]; }
    for (int d = 0; d < dim; d++)
      for (int k = 0; k < nc; k++) dl[d][k] = k == 0 ? -1. : (k == d + 1 ? 1. : 0.);
    int ne = t == ORC_TRI ? 3 : 6, ns = nc + (p >= 2 ? ne : 0);
    for (int k = 0; k < nc; k++) {
      phi[k] = l[k];
      for (int d = 0; d < dim; d++) dphi[d * ns + k] = dl[d][k];
    }
    for (int e = 0; p >= 2 && e < ne; e++) {
      int a = t == ORC_TRI ? TRI_EDGE[e][0] : TET_EDGE[e][0];
      int b = t == ORC_TRI ? TRI_EDGE[e][1] : TET_EDGE[e][1];
      phi[nc + e] = 4. * l[a] * l[b];
      for (int d = 0; d < dim; d++) dphi[d * ns + nc + e] = 4. * (dl[d][a] * l[b] + l[a] * dl[d][b]);
    }
    return ns;
  }
  /* raw (unscaled) generating functions on all sides, dims: [side], [d][side] */
  double ph[27], dph[3][27];
  if (t == ORC_LINE) {
    ph[0] = (1. - xi[0]) / 2.; ph[1] = (1. + xi[0]) / 2.;
    dph[0][0] = -0.5; dph[0][1] = 0.5;
  } else if (t == ORC_QUAD) {
    double x[2] = {(1. - xi[0]) / 2., (1. + xi[0]) / 2.}, dx[2] = {-0.5, 0.5};
    double y[2] = {(1. - xi[1]) / 2., (1. + xi[1]) / 2.}, dy[2] = {-0.5, 0.5};
    const int cx[4] = {0, 1, 1, 0}, cy[4] = {0, 0, 1, 1};
    for (int n = 0; n < 4; n++) {
      ph[n] = x[cx[n]] * y[cy[n]];
      dph[0][n] = dx[cx[n]] * y[cy[n]];
      dph[1][n] = x[cx[n]] * dy[cy[n]];
    }
  } else {
    double x[2] = {(1. - xi[0]) / 2., (1. + xi[0]) / 2.};
    double y[2] = {(1. - xi[1]) / 2., (1. + xi[1]) / 2.};
    double z[2] = {(1. - xi[2]) / 2., (1. + xi[2]) / 2.};
    const double dd[2] = {-0.5, 0.5};
    const int cx[8] = {0, 1, 1, 0, 0, 1, 1, 0}, cy[8] = {0, 0, 1, 1, 0, 0, 1, 1},
              cz[8] = {0, 0, 0, 0, 1, 1, 1, 1};
    for (int n = 0; n < 8; n++) {
      ph[n] = x[cx[n]] * y[cy[n]] * z[cz[n]];
      dph[0][n] = dd[cx[n]] * y[cy[n]] * z[cz[n]];
      dph[1][n] = x[cx[n]] * dd[cy[n]] * z[cz[n]];
      dph[2][n] = x[cx[n]] * y[cy[n]] * dd[cz[n]];
    }
  }
  int ns = nc;
  for (int i = 0; i < nc; i++) {
    phi[i] = ph[i];
  }
  if (p < 2) {
    for (int d = 0; d < dim; d++)
      for (int i = 0; i < nc; i++) dphi[d * nc + i] = dph[d][i];
    return nc;
  }
  /* raw products: edge = c_a*c_b, face = c0*c2 of the face, volume = c0*c6 */
  for (int s = nc; s < nsides; s++) {
    int a, b;
    if (t == ORC_LINE) { a = 0; b = 1; }
    else if (t == ORC_QUAD) {
      if (s < 8) { a = QUAD_EDGE[s - 4][0]; b = QUAD_EDGE[s - 4][1]; }
      else { a = 0; b = 2; }
    } else {
      if (s < 20) { a = HEX_EDGE[s - 8][0]; b = HEX_EDGE[s - 8][1]; }
      else if (s < 26) { a = HEX_FACE[s - 20][0]; b = HEX_FACE[s - 20][2]; }
      else { a = 0; b = 6; }
    }
    ph[s] = ph[a] * ph[b];
    for (int d = 0; d < dim; d++) dph[d][s] = dph[d][a] * ph[b] + ph[a] * dph[d][b];
  }
  /* "make the generating functions linear and unitary": add the raw values of
   * all higher-dimensional sides containing s (ascending side order, as in
   * Topology/tpzcube.cpp:49-82 highsides), then scale by 4/16/64. */
  for (int s = nc; s < nsides; s++) {
    int sd = side_dim(t, s);
    if (sd < dim) {
      int lc[8], nlc = side_corners(t, s, lc);
      for (int h = s + 1; h < nsides; h++) {
        if (side_dim(t, h) <= sd) continue;
        int hc[8], nhc = side_corners(t, h, hc), contained = 1;
        for (int k = 0; k < nlc && contained; k++) {
          int f = 0;
          for (int l = 0; l < nhc; l++) f |= (hc[l] == lc[k]);
          contained = f;
        }
        if (!contained) continue;
        ph[s] += ph[h];
        for (int d = 0; d < dim; d++) dph[d][s] += dph[d][h];
      }
    }
    double mult = sd == 1 ? 4. : sd == 2 ? 16. : 64.;
    ph[s] *= mult;
    for (int d = 0; d < dim; d++) dph[d][s] *= mult;
  }
  for (int s = nc; s < nsides; s++) phi[ns++] = ph[s];
  for (int d = 0; d < dim; d++)
    for (int i = 0; i < ns; i++) dphi[d * ns + i] = dph[d][i];
  return ns;
}

/* ------------------------------------------------------------------------ */
/* Hierarchical H1 shapes of any order on triangles and tetrahedra           */
/* (TPZShapeH1::Shape, Shape/TPZShapeH1.cpp:40-116).  Per side with n > 0    */
/* functions (NConnectShapeF of the connect's order): the blend function,   */
/* then blend * phin(i), i = 1..n-1, where phin are the side's internal      */
/* functions (ShapeInternal) at the point mapped to the side and reoriented  */
/* by the corner-node ids (ComputeTransforms, Shape/pzgenericshape.cpp:      */
/* 12-67).  Written with barycentric coordinates l_k of the element:         */
/*  * blends: edge 4 l_a l_b, face 27 l_a l_b l_c, tetrahedron interior 54   */
/*    l_0 l_1 l_2 l_3 (ShapeGenerating, Shape/pzshapetetra.cpp:101-150,      */
/*    Shape/pzshapetriang.cpp:53-81);                                        */
/*  * edge (a,b): TransformElementToSide gives s = l_b - l_a                 */
/*    (Topology/tpztetrahedron.cpp:460-545, tpztriangle.cpp:443-477), the    */
/*    parametric transform flips it unless id_a < id_b                       */
/*    (TPZShapeLinear::ParametricTransform, Shape/pzshapelinear.cpp:336-341);*/
/*    phin(i) = T_i(s) (TPZShapeLinear::ShapeInternal + Chebyshev,           */
/*    Shape/pzshapelinear.cpp:15-33,306-312);                                */
/*  * face (a,b,c): element-to-face coordinates (xi, eta) = (m_b, m_c) with   */
/*    m_k = l_k + w l_o (o the opposite corner, w = 1/3 for the face opposite */
/*    corner 0, else 0) and m_a = 1 - xi - eta; the parametric transform of  */
/*    GetTransformId2dT (Shape/pzshapetriang.cpp:18-27,410-431) turns them   */
/*    into (u, v) = (m of the face corner with the middle id, m of the one   */
/*    with the largest id); phin = T_i(2u-1) T_j(2v-1) enumerated by i+j     */
/*    ascending, then j ascending (TPZShapeTriang::ShapeInternal,            */
/*    Shape/pzshapetriang.cpp:279-311);                                      */
/*  * tetrahedron interior: identity transform; phin = T_i(2x-1) T_j(2y-1)    */
/*    T_k(2z-1) over i, j, k ascending with i+j+k < order-3                  */
/*    (TPZShapeTetra::ShapeInternal, Shape/pzshapetetra.cpp:345-371).        */
/* Gradients follow the reference's chain rule (dphiaux = mult^T dphin).     */
/* ------------------------------------------------------------------------ */
static void chebyshev(double x, int n, double *t, double *dt) {
  if (n <= 0) return;
  t[0] = 1.;
  dt[0] = 0.;
  if (n == 1) return;
  t[1] = x;
  dt[1] = 1.;
  for (int k = 2; k < n; k++) {
    t[k] = 2.0 * x * t[k - 1] - t[k - 2];
    dt[k] = 2.0 * x * dt[k - 1] + 2.0 * t[k - 1] - dt[k - 2];
  }
}

int orc_shape_h1(int t, const int64_t *ids, const int *orders, const double *xi, double *phi,
                 double *dphi) {
  if (!is_simplex(t)) return -1;
  const int dim = etype_dim(t), nc = dim + 1, nsides = etype_nsides(t);
  double l[4], dl[4][3];
  l[0] = 1.;
  for (int d = 0; d < dim; d++) { l[0] -= xi[d]; l[d + 1] = xi[d]; }
  for (int k = 0; k < nc; k++)
    for (int d = 0; d < 3; d++) dl[k][d] = d >= dim ? 0. : k == 0 ? -1. : (k == d + 1 ? 1. : 0.);
  for (int s = nc; s < nsides; s++)
    if (orders[s - nc] > 8) return -1; /* sizes of the local tables below */
  /* count first: dphi is laid out [d][ns] */
  int ns = nc;
  for (int s = nc; s < nsides; s++) ns += nconnect_shape(t, s, orders[s - nc]);
  int sh = 0;
  for (int k = 0; k < nc; k++, sh++) {
    phi[sh] = l[k];
    for (int d = 0; d < dim; d++) dphi[d * ns + sh] = dl[k][d];
  }
  for (int s = nc; s < nsides; s++) {
    int n = nconnect_shape(t, s, orders[s - nc]);
    if (n == 0) continue;
    int loc[4], nl = side_corners(t, s, loc);
    double B = nl == 2 ? 4. : nl == 3 ? 27. : 54., dB[3] = {0, 0, 0};
    for (int k = 0; k < nl; k++) B *= l[loc[k]];
    for (int d = 0; d < dim; d++) {
      double c = nl == 2 ? 4. : nl == 3 ? 27. : 54.;
      double acc = 0.;
      for (int k = 0; k < nl; k++) {
        double pr = dl[loc[k]][d];
        for (int j = 0; j < nl; j++)
          if (j != k) pr *= l[loc[j]];
        acc += pr;
      }
      dB[d] = c * acc;
    }
    double pin[64], dpin[64][3];
    pin[0] = 1.;
    dpin[0][0] = dpin[0][1] = dpin[0][2] = 0.;
    if (nl == 2) {
      int a = loc[0], b = loc[1];
      double sg = ids[a] < ids[b] ? 1. : -1.;
      double x = sg * (l[b] - l[a]), gx[3];
      for (int d = 0; d < 3; d++) gx[d] = sg * (dl[b][d] - dl[a][d]);
      double T[16], dT[16];
      chebyshev(x, n, T, dT);
      for (int i = 0; i < n; i++) {
        pin[i] = T[i];
        for (int d = 0; d < 3; d++) dpin[i][d] = dT[i] * gx[d];
      }
    } else if (nl == 3) {
      /* face corners in side order and the face coordinates m_k */
      double m[3], gm[3][3];
      double w = 0.;
      int o = -1;
      if (t == ORC_TET) {
        o = 0 + 1 + 2 + 3 - loc[0] - loc[1] - loc[2];
        w = o == 0 ? 1. / 3. : 0.;
      }
      for (int k = 1; k < 3; k++) {
        m[k] = l[loc[k]] + (o >= 0 ? w * l[o] : 0.);
        for (int d = 0; d < 3; d++) gm[k][d] = dl[loc[k]][d] + (o >= 0 ? w * dl[o][d] : 0.);
      }
      m[0] = 1. - m[1] - m[2];
      for (int d = 0; d < 3; d++) gm[0][d] = -gm[1][d] - gm[2][d];
      /* order the face corners by id: middle -> u, largest -> v */
      int r[3] = {0, 1, 2};
      for (int i = 0; i < 3; i++)
        for (int j = i + 1; j < 3; j++)
          if (ids[loc[r[j]]] < ids[loc[r[i]]]) { int tmp = r[i]; r[i] = r[j]; r[j] = tmp; }
      double u = 2. * m[r[1]] - 1., v = 2. * m[r[2]] - 1.;
      double Tu[16], dTu[16], Tv[16], dTv[16];
      chebyshev(u, orders[s - nc] - 2, Tu, dTu);
      chebyshev(v, orders[s - nc] - 2, Tv, dTv);
      int idx = 0;
      for (int ipj = 0; idx < n; ipj++)
        for (int j = 0; j <= ipj && idx < n; j++, idx++) {
          int i = ipj - j;
          pin[idx] = Tu[i] * Tv[j];
          for (int d = 0; d < 3; d++)
            dpin[idx][d] = 2. * dTu[i] * Tv[j] * gm[r[1]][d] + 2. * Tu[i] * dTv[j] * gm[r[2]][d];
        }
    } else {
      int ord = orders[s - nc] - 3;
      double T0[16], dT0[16], T1[16], dT1[16], T2[16], dT2[16];
      chebyshev(2. * xi[0] - 1., ord, T0, dT0);
      chebyshev(2. * xi[1] - 1., ord, T1, dT1);
      chebyshev(2. * xi[2] - 1., ord, T2, dT2);
      int idx = 0;
      for (int i = 0; i < ord; i++)
        for (int j = 0; j < ord; j++)
          for (int k = 0; k < ord; k++)
            if (i + j + k < ord) {
              pin[idx] = T0[i] * T1[j] * T2[k];
              dpin[idx][0] = 2. * dT0[i] * T1[j] * T2[k];
              dpin[idx][1] = 2. * T0[i] * dT1[j] * T2[k];
              dpin[idx][2] = 2. * T0[i] * T1[j] * dT2[k];
              idx++;
            }
    }
    for (int i = 0; i < n; i++, sh++) {
      phi[sh] = i == 0 ? B : B * pin[i];
      for (int d = 0; d < dim; d++)
        dphi[d * ns + sh] = i == 0 ? dB[d] : dB[d] * pin[i] + B * dpin[i][d];
    }
  }
  return ns;
}

/* ------------------------------------------------------------------------ */
/* Hash map: sorted corner-node tuple of a side -> connect index.            */
/* ------------------------------------------------------------------------ */
typedef struct {
  int64_t *keys; /* 4 per slot, -1 = empty */
  int64_t *val;
  int64_t cap;
} sidemap;

static uint64_t hash4(const int64_t *k) {
  uint64_t h = 1469598103934665603ULL;
  for (int i = 0; i < 4; i++) {
    h ^= (uint64_t)k[i] + 0x9e3779b97f4a7c15ULL + (h << 6) + (h >> 2);
    h *= 1099511628211ULL;
  }
  return h;
}

static int64_t sidemap_find_or_insert(sidemap *m, const int64_t *key, int64_t newval, int *inserted) {
  uint64_t h = hash4(key) & (uint64_t)(m->cap - 1);
  for (;;) {
    int64_t *k = m->keys + 4 * h;
    if (k[0] == -1) {
      memcpy(k, key, 4 * sizeof(int64_t));
      m->val[h] = newval;
      *inserted = 1;
      return newval;
    }
    if (k[0] == key[0] && k[1] == key[1] && k[2] == key[2] && k[3] == key[3]) {
      *inserted = 0;
      return m->val[h];
    }
    h = (h + 1) & (uint64_t)(m->cap - 1);
  }
}

static void sort_small(int64_t *a, int n) {
  for (int i = 1; i < n; i++) {
    int64_t v = a[i];
    int j = i - 1;
    while (j >= 0 && a[j] > v) { a[j + 1] = a[j]; j--; }
    a[j + 1] = v;
  }
}

/* ------------------------------------------------------------------------ */
/* Structured mesh + first-touch numbering.                                   */
/* Element order: volume elements (x fastest), then boundary elements face by */
/* face (faces 0..5 = x=0,x=1,y=0,y=1,z=0,z=1).  Connects are created in     */
/* element order, side order, when no earlier element shares the side:        */
/* TPZCreateApproximationSpace::BuildMesh (Mesh/pzcreateapproxspace.cpp:     */
/* 252-325), TPZCompElH1 ctor (Mesh/TPZCompElH1.cpp:17-24),                   */
/* TPZInterpolatedElement::CreateMidSideConnect (Mesh/pzintel.cpp:652-703),   */
/* TPZCompMesh::AllocateNewConnect (Mesh/pzcmesh.h:703-715).                 */
/* ------------------------------------------------------------------------ */
int orc_mesh_build(const orc_problem *pb, orc_mesh *m) {
  memset(m, 0, sizeof(*m));
  int dim = pb->dim, p = pb->p;
  int nx = pb->nx, ny = dim >= 2 ? pb->ny : 1, nz = dim >= 3 ? pb->nz : 1;
  m->dim = dim;
  m->p = p;
  m->nstate = pb->material == ORC_ELAST3D ? 3 : 1;
  int64_t NX = nx + 1, NY = dim >= 2 ? ny + 1 : 1, NZ = dim >= 3 ? nz + 1 : 1;
  m->nnodes = NX * NY * NZ;
  m->xyz = (double *)calloc(3 * m->nnodes, sizeof(double));
  for (int64_t k = 0; k < NZ; k++)
    for (int64_t j = 0; j < NY; j++)
      for (int64_t i = 0; i < NX; i++) {
        int64_t n = i + NX * (j + NY * k);
        m->xyz[3 * n + 0] = (double)i / nx;
        m->xyz[3 * n + 1] = dim >= 2 ? (double)j / ny : 0.;
        m->xyz[3 * n + 2] = dim >= 3 ? (double)k / nz : 0.;
      }
#define NODE(i, j, k) ((int64_t)(i) + NX * ((int64_t)(j) + NY * (int64_t)(k)))
  if (pb->simplex && dim == 3) {
    /* Kuhn split: 6 tetrahedra per cube around the diagonal corner 0 -> 6, one
     * per ordering of the axes (xyz, xzy, yxz, yzx, zxy, zyx); cubes x fastest.
     * Boundary triangles: faces 0..5 in turn, tetrahedra in order, tetrahedron
     * faces in side order, those lying on the face. */
    static const int KUHN[6][4] = {{0, 1, 2, 6}, {0, 1, 5, 6}, {0, 3, 2, 6},
                                   {0, 3, 7, 6}, {0, 4, 5, 6}, {0, 4, 7, 6}};
    int64_t ncube = (int64_t)nx * ny * nz, nvol = 6 * ncube, nbc = 0;
    for (int f = 0; f < 6; f++) {
      if (pb->bc_type[f] < 0) continue;
      int ax = f / 2;
      nbc += 2 * (int64_t)(ax == 0 ? ny * nz : ax == 1 ? nx * nz : nx * ny);
    }
    m->nvol = nvol;
    m->nel = nvol + nbc;
    m->etype = (int *)calloc(m->nel, sizeof(int));
    m->bcface = (int *)calloc(m->nel, sizeof(int));
    m->enodes = (int64_t *)calloc(8 * m->nel, sizeof(int64_t));
    m->econn = (int64_t *)calloc(27 * m->nel, sizeof(int64_t));
    int64_t e = 0;
    for (int iz = 0; iz < nz; iz++)
      for (int iy = 0; iy < ny; iy++)
        for (int ix = 0; ix < nx; ix++) {
          int64_t cn[8] = {NODE(ix, iy, iz), NODE(ix + 1, iy, iz), NODE(ix + 1, iy + 1, iz),
                           NODE(ix, iy + 1, iz), NODE(ix, iy, iz + 1), NODE(ix + 1, iy, iz + 1),
                           NODE(ix + 1, iy + 1, iz + 1), NODE(ix, iy + 1, iz + 1)};
          for (int k = 0; k < 6; k++, e++) {
            m->etype[e] = ORC_TET;
            m->bcface[e] = -1;
            for (int v = 0; v < 4; v++) m->enodes[8 * e + v] = cn[KUHN[k][v]];
          }
        }
    for (int f = 0; f < 6; f++) {
      if (pb->bc_type[f] < 0) continue;
      int ax = f / 2, side = f % 2, lim = ax == 0 ? nx : ax == 1 ? ny : nz;
      int64_t plane = side ? lim : 0;
      for (int64_t t = 0; t < nvol; t++)
        for (int fc = 0; fc < 4; fc++) {
          int on = 1;
          for (int v = 0; v < 3 && on; v++) {
            int64_t nd = m->enodes[8 * t + TET_FACE[fc][v]];
            int64_t ci = ax == 0 ? nd % NX : ax == 1 ? (nd / NX) % NY : nd / (NX * NY);
            on = ci == plane;
          }
          if (!on) continue;
          m->etype[e] = ORC_TRI;
          m->bcface[e] = f;
          for (int v = 0; v < 3; v++) m->enodes[8 * e + v] = m->enodes[8 * t + TET_FACE[fc][v]];
          e++;
        }
    }
    goto numbering;
  }
  int64_t nvol = (int64_t)nx * ny * nz;
  int64_t nbc = 0;
  if (dim >= 2)
    for (int f = 0; f < 2 * dim; f++) {
      if (pb->bc_type[f] < 0) continue;
      int ax = f / 2;
      int64_t cnt = 1;
      for (int a = 0; a < dim; a++)
        if (a != ax) cnt *= (a == 0 ? nx : a == 1 ? ny : nz);
      nbc += cnt;
    }
  m->nvol = nvol;
  m->nel = nvol + nbc;
  m->etype = (int *)calloc(m->nel, sizeof(int));
  m->bcface = (int *)calloc(m->nel, sizeof(int));
  m->enodes = (int64_t *)calloc(8 * m->nel, sizeof(int64_t));
  m->econn = (int64_t *)calloc(27 * m->nel, sizeof(int64_t));
  int64_t e = 0;
  int vt = dim == 1 ? ORC_LINE : dim == 2 ? ORC_QUAD : ORC_HEX;
  for (int iz = 0; iz < nz; iz++)
    for (int iy = 0; iy < ny; iy++)
      for (int ix = 0; ix < nx; ix++, e++) {
        m->etype[e] = vt;
        m->bcface[e] = -1;
        int64_t *en = m->enodes + 8 * e;
        if (dim == 1) { en[0] = ix; en[1] = ix + 1; }
        else if (dim == 2) {
          en[0] = NODE(ix, iy, 0); en[1] = NODE(ix + 1, iy, 0);
          en[2] = NODE(ix + 1, iy + 1, 0); en[3] = NODE(ix, iy + 1, 0);
        } else {
          en[0] = NODE(ix, iy, iz); en[1] = NODE(ix + 1, iy, iz);
          en[2] = NODE(ix + 1, iy + 1, iz); en[3] = NODE(ix, iy + 1, iz);
          en[4] = NODE(ix, iy, iz + 1); en[5] = NODE(ix + 1, iy, iz + 1);
          en[6] = NODE(ix + 1, iy + 1, iz + 1); en[7] = NODE(ix, iy + 1, iz + 1);
        }
      }
  if (dim >= 2)
    for (int f = 0; f < 2 * dim; f++) {
      if (pb->bc_type[f] < 0) continue;
      int ax = f / 2, side = f % 2;
      if (dim == 2) {
        int nn = ax == 0 ? ny : nx;
        for (int a = 0; a < nn; a++, e++) {
          m->etype[e] = ORC_LINE;
          m->bcface[e] = f;
          int64_t *en = m->enodes + 8 * e;
          if (ax == 0) { int x = side ? nx : 0; en[0] = NODE(x, a, 0); en[1] = NODE(x, a + 1, 0); }
          else { int y = side ? ny : 0; en[0] = NODE(a, y, 0); en[1] = NODE(a + 1, y, 0); }
        }
      } else {
        /* in-face axes (u,v): x-face (y,z), y-face (x,z), z-face (x,y) */
        int nu_ = ax == 0 ? ny : nx, nv_ = ax == 2 ? ny : nz;
        int c = side ? (ax == 0 ? nx : ax == 1 ? ny : nz) : 0;
        for (int b = 0; b < nv_; b++)
          for (int a = 0; a < nu_; a++, e++) {
            m->etype[e] = ORC_QUAD;
            m->bcface[e] = f;
            int64_t *en = m->enodes + 8 * e;
            if (ax == 0) {
              en[0] = NODE(c, a, b); en[1] = NODE(c, a + 1, b);
              en[2] = NODE(c, a + 1, b + 1); en[3] = NODE(c, a, b + 1);
            } else if (ax == 1) {
              en[0] = NODE(a, c, b); en[1] = NODE(a + 1, c, b);
              en[2] = NODE(a + 1, c, b + 1); en[3] = NODE(a, c, b + 1);
            } else {
              en[0] = NODE(a, b, c); en[1] = NODE(a + 1, b, c);
              en[2] = NODE(a + 1, b + 1, c); en[3] = NODE(a, b + 1, c);
            }
          }
      }
    }
numbering:;
#undef NODE
  /* first-touch numbering */
  sidemap sm;
  int64_t maxsides = 27 * m->nel;
  sm.cap = 1;
  while (sm.cap < 2 * maxsides) sm.cap <<= 1;
  sm.keys = (int64_t *)malloc(4 * sm.cap * sizeof(int64_t));
  memset(sm.keys, 0xff, 4 * sm.cap * sizeof(int64_t));
  sm.val = (int64_t *)malloc(sm.cap * sizeof(int64_t));
  /* Connects are created in element order, side order, the first time a side
   * is met (Mesh/pzintel.cpp:652-730: corners with order 1, other sides with
   * the default order).  Orders of higher-dimensional sides follow the hp rule
   * of TPZInterpolatedElement::IdentifySideOrder / ComputeSideOrder
   * (Mesh/pzintel.cpp:214-290, 1370-1390): the minimum preferred order of the
   * elements sharing the side.  Volume elements prefer el_porder[e] (or p);
   * boundary elements keep the default order p. */
  int64_t ccap = 1024, nconn = 0;
  int *cord = (int *)malloc(ccap * sizeof(int));
  for (e = 0; e < m->nel; e++) {
    int t = m->etype[e];
    int pe = (e < m->nvol && pb->el_porder) ? pb->el_porder[e] : p;
    for (int s = 0; s < etype_nsides(t); s++) {
      int ord = side_dim(t, s) == 0 ? 1 : pe;
      int lc[8], nlc = side_corners(t, s, lc);
      int64_t key[8];
      for (int k = 0; k < nlc; k++) key[k] = m->enodes[8 * e + lc[k]];
      int64_t idx;
      if (nlc == 8) {
        idx = nconn; /* hexahedron interior is never shared */
      } else {
        sort_small(key, nlc);
        for (int k = nlc; k < 4; k++) key[k] = -2;
        int ins;
        idx = sidemap_find_or_insert(&sm, key, nconn, &ins);
        if (!ins) {
          m->econn[27 * e + s] = idx;
          if (ord < cord[idx]) cord[idx] = ord;
          continue;
        }
      }
      if (nconn == ccap) { ccap *= 2; cord = (int *)realloc(cord, ccap * sizeof(int)); }
      cord[nconn] = ord;
      m->econn[27 * e + s] = nconn++;
    }
  }
  free(sm.keys);
  free(sm.val);
  int *cns = (int *)malloc((nconn > 0 ? nconn : 1) * sizeof(int));
  for (int64_t c = 0; c < nconn; c++) cns[c] = -1;
  for (e = 0; e < m->nel; e++) {
    int t = m->etype[e];
    for (int s = 0; s < etype_nsides(t); s++) {
      int64_t c = m->econn[27 * e + s];
      if (cns[c] < 0) cns[c] = nconnect_shape(t, s, cord[c]);
    }
  }
  m->nconn = nconn;
  m->conn_order = cord;
  m->conn_nshape = cns;
  m->conn_seq = (int64_t *)malloc(nconn * sizeof(int64_t));
  m->blk_pos = (int64_t *)malloc(nconn * sizeof(int64_t));
  m->blk_size = (int *)malloc(nconn * sizeof(int));
  int64_t pos = 0;
  for (int64_t c = 0; c < nconn; c++) {
    m->conn_seq[c] = c;
    m->blk_size[c] = cns[c] * m->nstate;
    m->blk_pos[c] = pos;
    pos += m->blk_size[c];
  }
  m->neq = pos;
  return 0;
}

/* Renumbers the connects: new sequence number perm[old seq] (the effect of
 * TPZCompMesh::Permute after a renumbering, Mesh/pzcmesh.cpp:1399-1459);
 * Block() positions follow the new sequence order. */
int orc_mesh_permute(orc_mesh *m, const int64_t *perm) {
  int64_t n = m->nconn;
  int *seen = (int *)calloc(n > 0 ? n : 1, sizeof(int));
  for (int64_t i = 0; i < n; i++) {
    if (perm[i] < 0 || perm[i] >= n || seen[perm[i]]) { free(seen); return -1; }
    seen[perm[i]] = 1;
  }
  free(seen);
  int *size = (int *)malloc((n > 0 ? n : 1) * sizeof(int));
  for (int64_t sq = 0; sq < n; sq++) size[perm[sq]] = m->blk_size[sq];
  int64_t pos = 0;
  for (int64_t sq = 0; sq < n; sq++) {
    m->blk_size[sq] = size[sq];
    m->blk_pos[sq] = pos;
    pos += size[sq];
  }
  free(size);
  for (int64_t c = 0; c < n; c++) m->conn_seq[c] = perm[m->conn_seq[c]];
  return 0;
}

void orc_mesh_free(orc_mesh *m) {
  free(m->xyz); free(m->etype); free(m->bcface); free(m->enodes); free(m->econn);
  free(m->conn_order); free(m->conn_nshape); free(m->conn_seq); free(m->blk_pos); free(m->blk_size);
  memset(m, 0, sizeof(*m));
}

int orc_elem_ndof(const orc_mesh *m, int64_t iel) {
  int t = m->etype[iel], n = 0;
  for (int s = 0; s < etype_nsides(t); s++) n += m->blk_size[m->conn_seq[m->econn[27 * iel + s]]];
  return n;
}

/* ------------------------------------------------------------------------ */
/* Jacobian (Mesh/pzgeoel.cpp:1168-1355).  gradx[i][j] = sum_n x_n[i]        */
/* dphi_corner[j][n].  Returns detjac (signed), fills jacinv[dim][dim] and    */
/* axes[dim][3].                                                             */
/* ------------------------------------------------------------------------ */
static double jacobian(int dim, const double gradx[3][3], double jacinv[3][3], double axes[3][3]) {
  double detjac = 0.;
  memset(axes, 0, 9 * sizeof(double));
  if (dim == 1) {
    double n1 = 0.;
    for (int i = 0; i < 3; i++) n1 += gradx[i][0] * gradx[i][0];
    n1 = sqrt(n1);
    detjac = n1;
    jacinv[0][0] = 1.0 / detjac;
    for (int i = 0; i < 3; i++) axes[0][i] = gradx[i][0] / n1;
  } else if (dim == 2) {
    double v1[3], v2[3], v1t[3], v2t[3], n1 = 0., n2 = 0., d12 = 0.;
    for (int i = 0; i < 3; i++) { v1[i] = gradx[i][0]; v2[i] = gradx[i][1]; }
    for (int i = 0; i < 3; i++) { n1 += v1[i] * v1[i]; d12 += v1[i] * v2[i]; }
    n1 = sqrt(n1);
    for (int i = 0; i < 3; i++) {
      v1t[i] = v1[i] / n1;
      v2t[i] = v2[i] - d12 * v1t[i] / n1;
      n2 += v2t[i] * v2t[i];
    }
    n2 = sqrt(n2);
    double j00 = n1, j01 = d12 / n1, j10 = 0., j11 = n2;
    detjac = j00 * j11 - j10 * j01;
    jacinv[0][0] = +j11 / detjac;
    jacinv[1][1] = +j00 / detjac;
    jacinv[0][1] = -j01 / detjac;
    jacinv[1][0] = -j10 / detjac;
    for (int i = 0; i < 3; i++) {
      v2t[i] /= n2;
      axes[0][i] = v1t[i];
      axes[1][i] = v2t[i];
    }
  } else {
    const double(*j)[3] = gradx;
    detjac -= j[0][2] * j[1][1] * j[2][0];
    detjac += j[0][1] * j[1][2] * j[2][0];
    detjac += j[0][2] * j[1][0] * j[2][1];
    detjac -= j[0][0] * j[1][2] * j[2][1];
    detjac -= j[0][1] * j[1][0] * j[2][2];
    detjac += j[0][0] * j[1][1] * j[2][2];
    jacinv[0][0] = (-j[1][2] * j[2][1] + j[1][1] * j[2][2]) / detjac;
    jacinv[0][1] = (j[0][2] * j[2][1] - j[0][1] * j[2][2]) / detjac;
    jacinv[0][2] = (-j[0][2] * j[1][1] + j[0][1] * j[1][2]) / detjac;
    jacinv[1][0] = (j[1][2] * j[2][0] - j[1][0] * j[2][2]) / detjac;
    jacinv[1][1] = (-j[0][2] * j[2][0] + j[0][0] * j[2][2]) / detjac;
    jacinv[1][2] = (j[0][2] * j[1][0] - j[0][0] * j[1][2]) / detjac;
    jacinv[2][0] = (-j[1][1] * j[2][0] + j[1][0] * j[2][1]) / detjac;
    jacinv[2][1] = (j[0][1] * j[2][0] - j[0][0] * j[2][1]) / detjac;
    jacinv[2][2] = (-j[0][1] * j[1][0] + j[0][0] * j[1][1]) / detjac;
    for (int i = 0; i < 3; i++) axes[i][i] = 1.;
  }
  return detjac;
}

static void forcing(const orc_problem *pb, const double *x, double *f) {
  f[0] = f[1] = f[2] = 0.;
  if (pb->force_kind == ORC_FORCE_CONST) {
    for (int k = 0; k < 3; k++) f[k] = pb->force_const[k];
  } else if (pb->force_kind == ORC_FORCE_BUMPS) {
    double w2 = 2. * pb->bump_width * pb->bump_width;
    for (int b = 0; b < pb->nbumps; b++) {
      const double *bp = pb->bumps + 6 * b;
      double r2 = 0.;
      for (int d = 0; d < 3; d++) r2 += (x[d] - bp[d]) * (x[d] - bp[d]);
      double g = exp(-r2 / w2);
      for (int k = 0; k < 3; k++) f[k] += bp[3 + k] * g;
    }
  }
}

/* Symmetric rules on the reference triangle {x,y >= 0, x+y <= 1} and
 * tetrahedron {x,y,z >= 0, x+y+z <= 1} (weights sum to the area 1/2 and the
 * volume 1/6).  Degree 2: 3 and 4 points; degree 4 (triangle, 6 points) and 5
 * (tetrahedron, 14 points), orbit parameters from the moment equations.  For
 * the polynomial integrands of affine simplices (p <= 2, element-wise constant
 * coefficients and loads) any rule of degree >= 2p integrates exactly, so the
 * element matrices equal the reference's (TPZIntTriang / TPZIntTetra3D,
 * Integral/pzquad.cpp:308-345) up to rounding. */
/* Gauss-Jacobi rule with weight (1-t)^alpha on [-1,1]: Golub-Welsch on the
 * symmetric tridiagonal matrix of the monic Jacobi recurrence (beta = 0),
 * eigenpairs by cyclic Jacobi rotations; nodes ascending. */
static void gauss_jacobi(int n, double alpha, double *t, double *w) {
  double A[16][16], V[16][16];
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) { A[i][j] = 0.; V[i][j] = i == j ? 1. : 0.; }
  for (int k = 0; k < n; k++) {
    double s = 2. * k + alpha;
    A[k][k] = (alpha == 0.) ? 0. : -alpha * alpha / (s * (s + 2.));
    if (k > 0) {
      double b = 4. * k * k * (k + alpha) * (k + alpha) / (s * s * (s + 1.) * (s - 1.));
      A[k - 1][k] = A[k][k - 1] = sqrt(b);
    }
  }
  for (int sweep = 0; sweep < 64; sweep++) {
    double off = 0.;
    for (int p = 0; p < n; p++)
      for (int q = p + 1; q < n; q++) off += A[p][q] * A[p][q];
    if (off < 1e-300) break;
    for (int p = 0; p < n; p++)
      for (int q = p + 1; q < n; q++) {
        if (A[p][q] == 0.) continue;
        double th = (A[q][q] - A[p][p]) / (2. * A[p][q]);
        double tt = (th >= 0. ? 1. : -1.) / (fabs(th) + sqrt(th * th + 1.));
        double c = 1. / sqrt(tt * tt + 1.), sn = tt * c;
        for (int k = 0; k < n; k++) {
          double akp = A[k][p], akq = A[k][q];
          A[k][p] = c * akp - sn * akq;
          A[k][q] = sn * akp + c * akq;
        }
        for (int k = 0; k < n; k++) {
          double apk = A[p][k], aqk = A[q][k];
          A[p][k] = c * apk - sn * aqk;
          A[q][k] = sn * apk + c * aqk;
        }
        for (int k = 0; k < n; k++) {
          double vkp = V[k][p], vkq = V[k][q];
          V[k][p] = c * vkp - sn * vkq;
          V[k][q] = sn * vkp + c * vkq;
        }
      }
  }
  double mu0 = pow(2., alpha + 1.) / (alpha + 1.);
  int ord[16];
  for (int i = 0; i < n; i++) ord[i] = i;
  for (int i = 0; i < n; i++)
    for (int j = i + 1; j < n; j++)
      if (A[ord[j]][ord[j]] < A[ord[i]][ord[i]]) { int tmp = ord[i]; ord[i] = ord[j]; ord[j] = tmp; }
  for (int i = 0; i < n; i++) {
    t[i] = A[ord[i]][ord[i]];
    w[i] = mu0 * V[0][ord[i]] * V[0][ord[i]];
  }
}

/* Collapsed (conical product) rule of degree 2n-1 on the reference simplex:
 * x = a(1-b)(1-c), y = b(1-c), z = c with Gauss-Jacobi rules of weights 1,
 * (1-b), (1-c)^2 on [0,1].  Any rule exact for the polynomial integrands of
 * affine simplices with element-wise constant data gives the reference's
 * element matrices (TPZIntTriang / TPZIntTetra3D tables, Integral/pzquad.cpp:
 * 308-345, Integral/tpzintrulet3d.cpp:447-505) up to rounding. */
int orc_collapsed_rule(int t, int n, double *pts, double *w) {
  if (n < 1 || n > 16) return -1;
  double ta[16], wa[16], tb[16], wb[16], tc[16], wc[16];
  gauss_jacobi(n, 0., ta, wa);
  gauss_jacobi(n, 1., tb, wb);
  gauss_jacobi(n, 2., tc, wc);
  int q = 0;
  if (t == ORC_TRI) {
    for (int j = 0; j < n; j++)
      for (int i = 0; i < n; i++, q++) {
        double a = 0.5 * (1. + ta[i]), b = 0.5 * (1. + tb[j]);
        pts[3 * q] = a * (1. - b); pts[3 * q + 1] = b; pts[3 * q + 2] = 0.;
        w[q] = (wa[i] / 2.) * (wb[j] / 4.);
      }
    return q;
  }
  for (int k = 0; k < n; k++)
    for (int j = 0; j < n; j++)
      for (int i = 0; i < n; i++, q++) {
        double a = 0.5 * (1. + ta[i]), b = 0.5 * (1. + tb[j]), c = 0.5 * (1. + tc[k]);
        pts[3 * q] = a * (1. - b) * (1. - c); pts[3 * q + 1] = b * (1. - c); pts[3 * q + 2] = c;
        w[q] = (wa[i] / 2.) * (wb[j] / 4.) * (wc[k] / 8.);
      }
  return q;
}

int orc_simplex_rule(int t, int order, double *pts, double *w) {
  int n = 0;
  if ((t == ORC_TRI && order > 4) || (t == ORC_TET && order > 5))
    return orc_collapsed_rule(t, (order + 2) / 2, pts, w);
  if (t == ORC_TRI) {
    if (order <= 2) {
      const double a = 1. / 6., b = 2. / 3.;
      const double P[3][2] = {{a, a}, {b, a}, {a, b}};
      for (int i = 0; i < 3; i++) { pts[3 * i] = P[i][0]; pts[3 * i + 1] = P[i][1]; pts[3 * i + 2] = 0; w[i] = 1. / 6.; }
      return 3;
    }
    const double A[2] = {0.4459484909159655, 0.0915762135097718};
    const double W[2] = {0.1116907948390057, 0.05497587182766096};
    for (int o = 0; o < 2; o++) {
      const double a = A[o], b = 1. - 2. * a;
      const double L[3][3] = {{b, a, a}, {a, b, a}, {a, a, b}};
      for (int k = 0; k < 3; k++, n++) {
        pts[3 * n] = L[k][1]; pts[3 * n + 1] = L[k][2]; pts[3 * n + 2] = 0; w[n] = W[o];
      }
    }
    return n;
  }
  if (order <= 2) {
    const double a = 0.1381966011250105, b = 1. - 3. * a;
    for (int k = 0; k < 4; k++, n++) {
      double L[4] = {a, a, a, a};
      L[k] = b;
      pts[3 * n] = L[1]; pts[3 * n + 1] = L[2]; pts[3 * n + 2] = L[3]; w[n] = 1. / 24.;
    }
    return 4;
  }
  const double A[2] = {0.09273525031089165, 0.31088591926330106};
  const double W[2] = {0.0122488405193938, 0.018781320953003007};
  for (int o = 0; o < 2; o++)
    for (int k = 0; k < 4; k++, n++) {
      double L[4] = {A[o], A[o], A[o], A[o]};
      L[k] = 1. - 3. * A[o];
      pts[3 * n] = L[1]; pts[3 * n + 1] = L[2]; pts[3 * n + 2] = L[3]; w[n] = W[o];
    }
  const double a3 = 0.045503704125647394, w3 = 0.007091003462846565, b3 = 0.5 - a3;
  for (int i = 0; i < 4; i++)
    for (int j = i + 1; j < 4; j++, n++) {
      double L[4] = {b3, b3, b3, b3};
      L[i] = a3;
      L[j] = a3;
      pts[3 * n] = L[1]; pts[3 * n + 1] = L[2]; pts[3 * n + 2] = L[3]; w[n] = w3;
    }
  return n;
}

/* TPZMatSingleSpaceT::IntegrationRuleOrder (Material/TPZMatSingleSpace.cpp:
 * 60-72) and the TPZElasticity3D override 2p (TPZElasticity3D.h:82-85). */
static int rule_order(const orc_problem *pb, int isbc, int p) {
  if (pb->material == ORC_ELAST3D || isbc) return 2 * p;
  int ff = pb->force_kind != ORC_FORCE_NONE ? pb->force_porder : 0;
  return p < ff ? ff + p : 2 * p;
}

/* fBigNumber = pow(10, max_digits10)*2/3 (Material/TPZMaterial.h:125) */
static double big_number(void) { return pow(10, 17) * 2 / 3; }

/* One integration point of the volume weak forms; ek row-major with leading
 * dimension ne.  kappa = fScale (Poisson) or permeability (Darcy).
 * TPZMatPoisson::Contribute (Material/Poisson/TPZMatPoisson.cpp:19-42),
 * TPZDarcyFlow::Contribute (Material/DarcyFlow/TPZDarcyFlow.cpp:19-46),
 * TPZElasticity3D::Contribute CODE3 (Material/Elasticity/TPZElasticity3D.cpp:
 * 269-372) with C1..C3 of TPZElasticity3D.h:183-188. */
void orc_contribute(int material, int dim, int nsh, int ne, double weight, const double *phi,
                    const double *dphix, const double *f, double kappa, double E, double nu,
                    double *ek, double *ef) {
  if (material == ORC_POISSON) {
    for (int i = 0; i < nsh; i++) {
      for (int j = 0; j < nsh; j++) {
        double dd = 0;
        for (int a = 0; a < dim; a++) dd += dphix[a * nsh + i] * dphix[a * nsh + j];
        ek[i * ne + j] += weight * kappa * dd;
      }
      ef[i] += weight * kappa * phi[i] * f[0];
    }
  } else if (material == ORC_DARCY) {
    double source = -f[0];
    for (int in = 0; in < nsh; in++) {
      ef[in] -= weight * source * phi[in];
      for (int jn = 0; jn < nsh; jn++)
        for (int kd = 0; kd < dim; kd++)
          ek[in * ne + jn] += weight * (dphix[kd * nsh + in] * kappa * dphix[kd * nsh + jn]);
    }
  } else if (material == ORC_ELAST3D) {
    double C1 = E / (2. + 2. * nu);
    double C2 = E * nu / (-1. + nu + 2. * nu * nu);
    double C3 = E * (nu - 1.) / (-1. + nu + 2. * nu * nu);
    for (int jn = 0; jn < nsh; jn++) {
      double dj[3] = {dphix[jn], dphix[nsh + jn], dphix[2 * nsh + jn]};
      for (int kd = 0; kd < 3; kd++) ef[3 * jn + kd] += weight * (f[kd] * phi[jn]);
      for (int in = 0; in < nsh; in++) {
        double D[3][3];
        for (int vd = 0; vd < 3; vd++)
          for (int ud = 0; ud < 3; ud++) D[vd][ud] = dphix[vd * nsh + in] * dj[ud];
        double *r0 = ek + (3 * in + 0) * ne + 3 * jn;
        double *r1 = ek + (3 * in + 1) * ne + 3 * jn;
        double *r2 = ek + (3 * in + 2) * ne + 3 * jn;
        r0[0] += weight * ((D[1][1] + D[2][2]) * C1 + D[0][0] * C3);
        r1[0] += weight * (D[0][1] * C1 - D[1][0] * C2);
        r2[0] += weight * (D[0][2] * C1 - D[2][0] * C2);
        r0[1] += weight * (D[1][0] * C1 - D[0][1] * C2);
        r1[1] += weight * ((D[0][0] + D[2][2]) * C1 + D[1][1] * C3);
        r2[1] += weight * (D[1][2] * C1 - D[2][1] * C2);
        r0[2] += weight * (D[2][0] * C1 - D[0][2] * C2);
        r1[2] += weight * (D[2][1] * C1 - D[1][2] * C2);
        r2[2] += weight * ((D[0][0] + D[1][1]) * C1 + D[2][2] * C3);
      }
    }
  }
}

/* ------------------------------------------------------------------------ */
/* Boundary conditions: TPZMatPoisson::ContributeBC (Material/Poisson/        */
/* TPZMatPoisson.cpp:45-121), TPZDarcyFlow::ContributeBC (Material/DarcyFlow/ */
/* TPZDarcyFlow.cpp:49-91), TPZElasticity3D::ContributeBC (Material/          */
/* Elasticity/TPZElasticity3D.cpp:616-773), one integration point.  Types     */
/* outside each switch are refused (-1): the reference prints an error and    */
/* adds nothing (Poisson, Elasticity3D) or stops (Darcy).                     */
/* ------------------------------------------------------------------------ */
static int bc_contribute(const orc_problem *pb, int face, int nsh, int ne, double weight,
                         const double *phi, const double *dphix, const double *normal, double *ek,
                         double *ef) {
  const int type = pb->bc_type[face];
  const double *v2 = pb->bc_val[face];
  const double *v1 = pb->bc_val1[face];
  if (pb->material == ORC_ELAST3D) {
    const double BIG = 1.e12;
    double v2l[3] = {v2[0], v2[1], v2[2]};
    /* components held by the penalty: 0 all, 5 x, 6 y, 7 z, 8 x and z */
    int pen[3] = {0, 0, 0};
    switch (type) {
      case 0: pen[0] = pen[1] = pen[2] = 1; break;
      case 5: pen[0] = 1; break;
      case 6: pen[1] = 1; break;
      case 7: pen[2] = 1; break;
      case 8: pen[0] = pen[2] = 1; break;
      case 1: case 2: case 3: case 4: break;
      default: return -1;
    }
    if (type == 4)
      for (int k = 0; k < 3; k++)
        v2l[k] = -(v1[3 * k + 0] * normal[0] + v1[3 * k + 1] * normal[1] + v1[3 * k + 2] * normal[2]);
    for (int in = 0; in < nsh; in++) {
      for (int k = 0; k < 3; k++) {
        if (pen[k]) ef[3 * in + k] += BIG * v2l[k] * phi[in] * weight;
        else if (type == 1 || type == 2 || type == 4) ef[3 * in + k] += v2l[k] * phi[in] * weight;
      }
      for (int jn = 0; jn < nsh; jn++) {
        for (int k = 0; k < 3; k++) {
          if (pen[k]) ek[(3 * in + k) * ne + 3 * jn + k] += BIG * phi[in] * phi[jn] * weight;
          if (type == 3)
            ek[(3 * in + k) * ne + 3 * jn + k] += BIG * phi[in] * phi[jn] * weight * v2l[k];
        }
        if (type == 2)
          for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++)
              ek[(3 * in + a) * ne + 3 * jn + b] += v1[3 * a + b] * weight * phi[in] * phi[jn];
      }
    }
    return 0;
  }
  const double big = big_number();
  if (pb->material == ORC_POISSON) {
    for (int in = 0; in < nsh; in++) {
      if (type == 0) {
        ef[in] += big * v2[0] * phi[in] * weight;
        for (int jn = 0; jn < nsh; jn++) ek[in * ne + jn] += big * phi[in] * phi[jn] * weight;
      } else if (type == 1) {
        ef[in] += v2[0] * pb->scale * phi[in] * weight;
      } else if (type == 2) {
        /* Robin as written there: the penalty times Val1 on the first
         * derivative (dphix row 0) of the boundary element */
        ef[in] += big * v2[0] * phi[in] * weight;
        for (int jn = 0; jn < nsh; jn++)
          ek[in * ne + jn] += big * v1[0] * dphix[in] * dphix[jn] * weight;
      } else {
        return -1;
      }
    }
    return 0;
  }
  for (int in = 0; in < nsh; in++) { /* Darcy */
    if (type == 0) {
      ef[in] += (big * phi[in] * weight) * v2[0];
      for (int jn = 0; jn < nsh; jn++) ek[in * ne + jn] += big * phi[in] * phi[jn] * weight;
    } else if (type == 1) {
      ef[in] += v2[0] * (phi[in] * weight);
    } else if (type == 2) {
      ef[in] += v2[0] * (phi[in] * weight);
      for (int jn = 0; jn < nsh; jn++) ek[in * ne + jn] += v1[0] * (phi[in] * phi[jn] * weight);
    } else {
      return -1;
    }
  }
  return 0;
}

/* Centre of a geometric element: X at TPZGeoEl::CenterPoint of its own side,
 * the mean of the corners for the linear / multilinear maps used here. */
static void elem_centre(const orc_mesh *m, int64_t iel, double *x) {
  int nc = etype_ncorner(m->etype[iel]);
  x[0] = x[1] = x[2] = 0.;
  for (int n = 0; n < nc; n++)
    for (int d = 0; d < 3; d++) x[d] += m->xyz[3 * m->enodes[8 * iel + n] + d] / nc;
}

/* vec of TPZInterpolationSpace::ComputeNormal (Mesh/pzinterpolationspace.cpp:
 * 301-345): centre of the boundary element minus the centre of its first
 * neighbour of higher dimension, here the volume element holding all its
 * corners (unique on the domain boundary). */
static int bc_neighbour_vec(const orc_mesh *m, int64_t iel, double *vec) {
  int nc = etype_ncorner(m->etype[iel]);
  for (int64_t v = 0; v < m->nvol; v++) {
    int nv = etype_ncorner(m->etype[v]), found = 0;
    for (int a = 0; a < nc; a++)
      for (int b = 0; b < nv; b++)
        if (m->enodes[8 * iel + a] == m->enodes[8 * v + b]) { found++; break; }
    if (found == nc) {
      double xb[3], xv[3];
      elem_centre(m, iel, xb);
      elem_centre(m, v, xv);
      for (int d = 0; d < 3; d++) vec[d] = -xv[d] + xb[d];
      return 0;
    }
  }
  return -1;
}

static void cross(const double *a, const double *b, double *c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = -a[0] * b[2] + a[2] * b[0];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

/* ComputeNormal (Mesh/pzinterpolationspace.cpp:347-383): the unit normal of
 * the boundary element from its axes, turned to have a non-negative product
 * with vec. */
static void bc_normal(int dim, const double axes[3][3], const double *vec, double *n) {
  double a1[3] = {axes[0][0], axes[0][1], axes[0][2]}, a2[3], t[3];
  if (dim == 1) {
    cross(a1, vec, a2);
    cross(a2, a1, t);
  } else {
    for (int i = 0; i < 3; i++) a2[i] = axes[1][i];
    cross(a1, a2, t);
  }
  double sz = sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
  double dot = 0.;
  for (int i = 0; i < 3; i++) { n[i] = t[i] / sz; dot += n[i] * vec[i]; }
  if (dot < 0.)
    for (int i = 0; i < 3; i++) n[i] = -n[i];
}

/* ------------------------------------------------------------------------ */
/* CalcStiff: TPZInterpolationSpace::CalcStiffInternal                        */
/* (Mesh/pzinterpolationspace.cpp:403-474) with ComputeRequiredDataT (:265-  */
/* 298) and TPZCompElH1::ComputeShape (Mesh/TPZCompElH1.cpp:139-149).         */
/* ek row-major, dofs ordered shape-major (index = nstate*shape + state).     */
/* ------------------------------------------------------------------------ */
int orc_calcstiff(const orc_problem *pb, const orc_mesh *m, int64_t iel, double *ek,
                  double *ef, int64_t *dest) {
  int t = m->etype[iel], dim = etype_dim(t), nc = etype_ncorner(t), ns_ = m->nstate;
  int isbc = m->bcface[iel] >= 0;
  /* destination indices (Mesh/pzelmat.cpp:37-70) */
  if (orc_elem_ndof(m, iel) > ORC_MAXNE) return -1;
  int ne = 0;
  for (int s = 0; s < etype_nsides(t); s++) {
    int64_t seq = m->conn_seq[m->econn[27 * iel + s]];
    for (int k = 0; k < m->blk_size[seq]; k++) dest[ne++] = m->blk_pos[seq] + k;
  }
  memset(ek, 0, sizeof(double) * ne * ne);
  memset(ef, 0, sizeof(double) * ne);
  if (pb->material == ORC_MOCK1D) {
    /* TPZMatTest in UnitTest_PZ/TestStruct/StructMatrixUnitTest.cpp:333-362 */
    if (!isbc && ne == 2) { ek[0] = 2; ek[1] = 1; ek[2] = 1; ek[3] = 2; }
    return ne;
  }
  double coord[8][3];
  for (int n = 0; n < nc; n++)
    for (int d = 0; d < 3; d++) coord[n][d] = m->xyz[3 * m->enodes[8 * iel + n] + d];
  /* TPZInterpolationSpace::MaxOrder (Mesh/pzinterpolationspace.cpp:53-66)
   * drives the rule (AdjustIntegrationRule, :69-83); shapes exist only for
   * connects with shape functions, in side order */
  int maxorder = -1, keep[27], nkeep = 0;
  for (int s = 0; s < etype_nsides(t); s++) {
    int64_t c = m->econn[27 * iel + s];
    if (m->conn_order[c] > maxorder) maxorder = m->conn_order[c];
    if (m->conn_nshape[c] > 0) keep[nkeep++] = s;
  }
  double gl[64], gw[64], spts[3 * 512], sw[512];
  int ng = 0, nq = 1;
  if (is_simplex(t)) {
    nq = orc_simplex_rule(t, rule_order(pb, isbc, maxorder), spts, sw);
    if (nq < 0) return -1;
  } else {
    ng = orc_gauss_legendre(rule_order(pb, isbc, maxorder), gl, gw);
    for (int d = 0; d < dim; d++) nq *= ng;
  }
  double coef = 1.;
  if (!isbc && pb->el_coef) coef = pb->el_coef[iel];
  double phi[64], dphi[192], dphix[192];
  /* simplices: corner ids and connect orders drive the side transforms */
  int64_t ids[4];
  int sord[27];
  for (int n = 0; n < nc; n++) ids[n] = m->enodes[8 * iel + n];
  for (int s = nc; s < etype_nsides(t); s++) sord[s - nc] = m->conn_order[m->econn[27 * iel + s]];
  double vec[3] = {0, 0, 0};
  if (isbc && pb->material == ORC_ELAST3D && pb->bc_type[m->bcface[iel]] == 4 &&
      bc_neighbour_vec(m, iel, vec) != 0)
    return -1;
  for (int q = 0; q < nq; q++) {
    double xi[3] = {0, 0, 0}, w;
    if (is_simplex(t)) {
      xi[0] = spts[3 * q]; xi[1] = spts[3 * q + 1]; xi[2] = spts[3 * q + 2]; w = sw[q];
    } else if (dim == 1) { xi[0] = gl[q]; w = gw[q]; }
    else if (dim == 2) { /* Integral/pzquad.cpp:159-169: ksi slow, eta fast */
      int ik = q / ng, ie = q % ng;
      xi[0] = gl[ik]; xi[1] = gl[ie]; w = gw[ik] * gw[ie];
    } else { /* Integral/pzquad.cpp:274-287: ksi fastest */
      int ik = q % ng, ie = (q % (ng * ng)) / ng, iz = q / (ng * ng);
      xi[0] = gl[ik]; xi[1] = gl[ie]; xi[2] = gl[iz]; w = gw[ik] * gw[ie] * gw[iz];
    }
    int nsh = 0;
    if (is_simplex(t)) {
      nsh = orc_shape_h1(t, ids, sord, xi, phi, dphi);
      if (nsh < 0) return -1;
    } else {
      double phiall[27], dphiall[81];
      int nall = orc_shape(t, m->p, xi, phiall, dphiall);
      /* p <= 2: shape i belongs to side i; keep those of sides with a shape */
      int sel[27];
      for (int k = 0; k < nkeep; k++)
        if (keep[k] < nall) sel[nsh++] = keep[k];
      for (int i = 0; i < nsh; i++) {
        phi[i] = phiall[sel[i]];
        for (int a = 0; a < dim; a++) dphi[a * nsh + i] = dphiall[a * nall + sel[i]];
      }
    }
    /* corner derivatives for GradX (Geom/TPZGeoCube.h:124, pzgeoquad.h:144) */
    double cphi[27], cdphi[81];
    orc_shape(t, 1, xi, cphi, cdphi);
    double gradx[3][3] = {{0}}, x[3] = {0, 0, 0};
    for (int n = 0; n < nc; n++)
      for (int i = 0; i < 3; i++) {
        x[i] += coord[n][i] * cphi[n];
        for (int j = 0; j < dim; j++) gradx[i][j] += coord[n][i] * cdphi[j * nc + n];
      }
    double jinv[3][3] = {{0}}, axes[3][3];
    double detjac = fabs(jacobian(dim, gradx, jinv, axes));
    /* dphix = jacinv^T dphi */
    for (int i = 0; i < nsh; i++)
      for (int a = 0; a < dim; a++) {
        double s = 0.;
        for (int b = 0; b < dim; b++) s += jinv[b][a] * dphi[b * nsh + i];
        dphix[a * nsh + i] = s;
      }
    double weight = w * detjac;
    double f[3];
    forcing(pb, x, f);
    if (isbc) {
      int face = m->bcface[iel];
      double nrm[3] = {0, 0, 0};
      if (pb->material == ORC_ELAST3D && pb->bc_type[face] == 4) bc_normal(dim, axes, vec, nrm);
      if (bc_contribute(pb, face, nsh, ne, weight, phi, dphix, nrm, ek, ef) != 0) return -1;
      continue;
    }
    orc_contribute(pb->material, dim, nsh, ne, weight, phi, dphix, f,
                   pb->material == ORC_DARCY ? (pb->el_coef ? coef : pb->scale) : pb->scale,
                   pb->el_coef ? coef : pb->E, pb->nu, ek, ef);
    /* CODE3 load term of the prestress (Material/Elasticity/TPZElasticity3D.cpp:271-278) */
    if (pb->material == ORC_ELAST3D)
      for (int jn = 0; jn < nsh; jn++)
        for (int kd = 0; kd < 3; kd++) ef[3 * jn + kd] -= weight * pb->prestress[kd] * dphix[kd * nsh + jn];
  }
  return ne;
}

/* ------------------------------------------------------------------------ */
/* SetupMatrixData (StrMatrix/TPZSSpStructMatrix.cpp:49-193) on the node     */
/* graph of TPZRenumbering::ConvertGraph (External/TPZRenumbering.cpp:76-110) */
/* built from TPZCompMesh::ComputeElGraph (Mesh/pzcmesh.cpp:1223-1266).       */
/* ------------------------------------------------------------------------ */
static int cmp_i64(const void *a, const void *b) {
  int64_t x = *(const int64_t *)a, y = *(const int64_t *)b;
  return x < y ? -1 : x > y;
}

/* Node graph of TPZRenumbering::ConvertGraph over explicit element lists
 * (element e holds the sequence numbers el[ep[e] .. ep[e+1])): per node the
 * sorted set of the other nodes of its elements. */
static void node_graph_lists(int64_t nb, int64_t nel, const int64_t *ep, const int64_t *el,
                             int64_t **ngp_out, int64_t **ng_out) {
  /* node -> element graph */
  int64_t *cnt = (int64_t *)calloc(nb + 1, sizeof(int64_t));
  for (int64_t e = 0; e < nel; e++)
    for (int64_t k = ep[e]; k < ep[e + 1]; k++) cnt[el[k] + 1]++;
  for (int64_t i = 0; i < nb; i++) cnt[i + 1] += cnt[i];
  int64_t *n2e = (int64_t *)malloc((cnt[nb] ? cnt[nb] : 1) * sizeof(int64_t));
  int64_t *fill = (int64_t *)calloc(nb ? nb : 1, sizeof(int64_t));
  for (int64_t e = 0; e < nel; e++)
    for (int64_t k = ep[e]; k < ep[e + 1]; k++) {
      int64_t c = el[k];
      n2e[cnt[c] + fill[c]++] = e;
    }
  free(fill);
  /* per node: sorted unique neighbour set (std::set semantics), without self */
  int64_t *ngp = (int64_t *)calloc(nb + 1, sizeof(int64_t));
  int64_t ngcap = 1 << 20, ngn = 0;
  int64_t *ng = (int64_t *)malloc(ngcap * sizeof(int64_t));
  int64_t *tmp = (int64_t *)malloc(8192 * sizeof(int64_t));
  for (int64_t i = 0; i < nb; i++) {
    int nt = 0;
    for (int64_t k = cnt[i]; k < cnt[i + 1]; k++) {
      int64_t e = n2e[k];
      for (int64_t j = ep[e]; j < ep[e + 1]; j++) {
        int64_t c = el[j];
        if (c != i) tmp[nt++] = c;
      }
    }
    qsort(tmp, nt, sizeof(int64_t), cmp_i64);
    int nu = 0;
    for (int k = 0; k < nt; k++)
      if (nu == 0 || tmp[k] != tmp[nu - 1]) tmp[nu++] = tmp[k];
    while (ngn + nu > ngcap) { ngcap *= 2; ng = (int64_t *)realloc(ng, ngcap * sizeof(int64_t)); }
    memcpy(ng + ngn, tmp, nu * sizeof(int64_t));
    ngn += nu;
    ngp[i + 1] = ngn;
  }
  free(tmp);
  free(n2e);
  free(cnt);
  *ngp_out = ngp;
  *ng_out = ng;
}

/* element lists of the mesh: every side's connect (sequence number) */
static void mesh_lists(const orc_mesh *m, int64_t **ep_out, int64_t **el_out) {
  int64_t *ep = (int64_t *)malloc((m->nel + 1) * sizeof(int64_t)), n = 0;
  ep[0] = 0;
  for (int64_t e = 0; e < m->nel; e++) n += etype_nsides(m->etype[e]), ep[e + 1] = n;
  int64_t *el = (int64_t *)malloc((n ? n : 1) * sizeof(int64_t));
  for (int64_t e = 0; e < m->nel; e++)
    for (int s = 0; s < etype_nsides(m->etype[e]); s++)
      el[ep[e] + s] = m->conn_seq[m->econn[27 * e + s]];
  *ep_out = ep;
  *el_out = el;
}

static void node_graph(const orc_mesh *m, int64_t **ngp_out, int64_t **ng_out) {
  int64_t *ep, *el;
  mesh_lists(m, &ep, &el);
  node_graph_lists(m->nconn, m->nel, ep, el, ngp_out, ng_out);
  free(ep);
  free(el);
}

/* The upper CSR of SetupMatrixData from a node graph; connects with skip[c]
 * set carry no equations (constrained connects, numbered after the others). */
static int64_t pattern_csr(const orc_mesh *m, const int64_t *ngp, const int64_t *ng, const char *skip,
                           int64_t **ia_out, int64_t **ja_out) {
  const int64_t nb = m->nconn;
#define BSZ(c) ((skip && skip[c]) ? 0 : (int64_t)m->blk_size[c])
  int64_t total = 0;
  for (int64_t i = 0; i < nb; i++) {
    int64_t bs = BSZ(i);
    if (!bs) continue;
    total += bs * (bs + 1) / 2;
    for (int64_t j = ngp[i]; j < ngp[i + 1]; j++)
      if (ng[j] > i) total += bs * BSZ(ng[j]);
  }
  int64_t *ia = (int64_t *)malloc((m->neq + 1) * sizeof(int64_t));
  int64_t *ja = (int64_t *)malloc((total ? total : 1) * sizeof(int64_t));
  int64_t pos = 0, ieq = 0;
  for (int64_t i = 0; i < nb; i++) {
    int64_t bs = BSZ(i), bp = m->blk_pos[i];
    for (int64_t k = 0; k < bs; k++, ieq++) {
      ia[ieq] = pos;
      int diag = 0;
      for (int64_t j = ngp[i]; j < ngp[i + 1]; j++) {
        int64_t col = ng[j];
        if (col < i) continue;
        if (!diag && col > i) {
          diag = 1;
          for (int64_t l = 0; l < bs; l++)
            if (bp + l >= ieq) ja[pos++] = bp + l;
        }
        for (int64_t l = 0; l < BSZ(col); l++)
          if (m->blk_pos[col] + l >= ieq) ja[pos++] = m->blk_pos[col] + l;
      }
      if (!diag)
        for (int64_t l = 0; l < bs; l++)
          if (bp + l >= ieq) ja[pos++] = bp + l;
    }
  }
#undef BSZ
  ia[ieq] = pos;
  *ia_out = ia;
  *ja_out = ja;
  return pos;
}

int64_t orc_setup_pattern(const orc_mesh *m, int64_t **ia_out, int64_t **ja_out) {
  int64_t *ngp, *ng;
  node_graph(m, &ngp, &ng);
  int64_t nnz = pattern_csr(m, ngp, ng, NULL, ia_out, ja_out);
  free(ng);
  free(ngp);
  return nnz;
}

/* SetupMatrixData with an active TPZEquationFilter (StrMatrix/
 * TPZSSpStructMatrix.cpp:49-193 with TPZEquationFilter::NumActive/Filter,
 * StrMatrix/TPZEquationFilter.h): blocks without an active equation are
 * skipped; every active equation of a block is a row of the condensed
 * numbering dest[] (original equation -> condensed index, -1 inactive;
 * increasing over the active equations); AddColEqs keeps the active
 * equations of each column block, renumbered, with index >= the row.  The
 * reference sizes Eq by whole blocks (Eq[totaleq] of the partially active
 * blocks is never written) and EqCol by whole block pairs; the used part,
 * returned here, is IA[0..nactive] and JA[0..IA[nactive]). */
int64_t orc_setup_pattern_filtered(const orc_mesh *m, const int64_t *dest, int64_t **ia_out,
                                   int64_t **ja_out) {
  int64_t *ngp, *ng;
  node_graph(m, &ngp, &ng);
  const int64_t nb = m->nconn;
  int64_t nact = 0, total = 0;
  for (int64_t e = 0; e < m->neq; e++) nact += dest[e] >= 0;
  for (int64_t i = 0; i < nb; i++) {
    int64_t bs = m->blk_size[i];
    total += bs * (bs + 1) / 2;
    for (int64_t j = ngp[i]; j < ngp[i + 1]; j++)
      if (ng[j] > i) total += bs * m->blk_size[ng[j]];
  }
  int64_t *ia = (int64_t *)malloc((nact + 1) * sizeof(int64_t));
  int64_t *ja = (int64_t *)malloc((total ? total : 1) * sizeof(int64_t));
  int64_t pos = 0, ieq = 0;
#define ADDCOLEQS(csize, cpos)                                   \
  for (int64_t l_ = 0; l_ < (csize); l_++) {                     \
    int64_t d_ = dest[(cpos) + l_];                              \
    if (d_ < 0 || d_ < ieq) continue;                            \
    ja[pos++] = d_;                                              \
  }
  for (int64_t i = 0; i < nb; i++) {
    int64_t bs = m->blk_size[i], bp = m->blk_pos[i], na = 0;
    for (int64_t l = 0; l < bs; l++) na += dest[bp + l] >= 0;
    if (!na) continue;
    for (int64_t k = 0; k < bs; k++) {
      if (dest[bp + k] < 0) continue;
      if (dest[bp + k] != ieq) { free(ia); free(ja); free(ng); free(ngp); return -1; }
      ia[ieq] = pos;
      int diag = 0;
      for (int64_t j = ngp[i]; j < ngp[i + 1]; j++) {
        int64_t col = ng[j];
        if (col < i) continue;
        if (!diag && col > i) {
          diag = 1;
          ADDCOLEQS(bs, bp);
        }
        int64_t cs = m->blk_size[col], cp = m->blk_pos[col], ca = 0;
        for (int64_t l = 0; l < cs; l++) ca += dest[cp + l] >= 0;
        if (!ca) continue;
        ADDCOLEQS(cs, cp);
      }
      if (!diag) { ADDCOLEQS(bs, bp); }
      ieq++;
    }
  }
#undef ADDCOLEQS
  ia[ieq] = pos;
  free(ng);
  free(ngp);
  *ia_out = ia;
  *ja_out = ja;
  return pos;
}

/* TPZStructMatrixOR::Assemble with an active filter (StrMatrix/
 * pzstrmatrixor.cpp:41-71, 226-233): every element matrix passes through
 * TPZEquationFilter::Filter(source, dest) -- inactive equations dropped, the
 * others renumbered -- before AddKel / AddFel into the condensed system;
 * the condensed load is then scattered into rhs_full[neq] (inactive 0). */
int orc_assemble_filtered(const orc_problem *pb, const orc_mesh *m, const int64_t *dest,
                          const int64_t *ia, const int64_t *ja, double *a, double *rhs_full) {
  int64_t nact = 0;
  for (int64_t e = 0; e < m->neq; e++) nact += dest[e] >= 0;
  memset(a, 0, sizeof(double) * ia[nact]);
  memset(rhs_full, 0, sizeof(double) * m->neq);
  double *ek = (double *)malloc(sizeof(double) * ORC_MAXNE * ORC_MAXNE);
  double *ekf = (double *)malloc(sizeof(double) * ORC_MAXNE * ORC_MAXNE);
  double *rhsloc = (double *)calloc(nact ? nact : 1, sizeof(double));
  double ef[ORC_MAXNE];
  int64_t dst[ORC_MAXNE], src[ORC_MAXNE], dstf[ORC_MAXNE];
  int rc = 0;
  for (int64_t e = 0; e < m->nel && !rc; e++) {
    int ne = orc_calcstiff(pb, m, e, ek, ef, dst);
    if (ne < 0) { rc = -1; break; }
    int nf = 0;
    for (int i = 0; i < ne; i++)
      if (dest[dst[i]] >= 0) { src[nf] = i; dstf[nf] = dest[dst[i]]; nf++; }
    /* AddKel(elmat, source, destination): the filtered sub-matrix */
    for (int i = 0; i < nf; i++)
      for (int j = 0; j < nf; j++) ekf[i * nf + j] = ek[src[i] * ne + src[j]];
    orc_addkel(ia, ja, a, nf, ekf, dstf);
    for (int i = 0; i < nf; i++) rhsloc[dstf[i]] += ef[src[i]];
  }
  /* TPZEquationFilter::Scatter */
  for (int64_t e = 0; e < m->neq; e++)
    if (dest[e] >= 0) rhs_full[e] = rhsloc[dest[e]];
  free(ek);
  free(ekf);
  free(rhsloc);
  return rc;
}

/* ------------------------------------------------------------------------ */
/* AddKel (Matrix/pzsysmp.cpp:369-411): upper entries only, skip            */
/* IsZero(v) = |v| < 1e-12 (Common/pzreal.h:736-741), sequential guess then  */
/* row search.                                                               */
/* ------------------------------------------------------------------------ */
/* TPZSpStructMatrix::SetupMatrixData (StrMatrix/TPZSpStructMatrix.cpp:53-190):
 * every equation of block i gets the equations of block i and of every
 * adjacent block (both triangles), then the row is stable-sorted. */
int64_t orc_setup_pattern_full(const orc_mesh *m, int64_t **ia_out, int64_t **ja_out) {
  const int64_t nb = m->nconn;
  int64_t *ngp, *ng;
  node_graph(m, &ngp, &ng);
  int64_t total = 0;
  for (int64_t i = 0; i < nb; i++) {
    int64_t bs = m->blk_size[i];
    if (!bs) continue;
    total += bs * bs;
    for (int64_t j = ngp[i]; j < ngp[i + 1]; j++) total += bs * m->blk_size[ng[j]];
  }
  int64_t *ia = (int64_t *)malloc((m->neq + 1) * sizeof(int64_t));
  int64_t *ja = (int64_t *)malloc((total ? total : 1) * sizeof(int64_t));
  int64_t pos = 0, ieq = 0;
  for (int64_t i = 0; i < nb; i++) {
    int64_t bs = m->blk_size[i], bp = m->blk_pos[i];
    for (int64_t k = 0; k < bs; k++, ieq++) {
      ia[ieq] = pos;
      for (int64_t l = 0; l < bs; l++) ja[pos++] = bp + l;
      for (int64_t j = ngp[i]; j < ngp[i + 1]; j++)
        for (int64_t l = 0; l < m->blk_size[ng[j]]; l++) ja[pos++] = m->blk_pos[ng[j]] + l;
      qsort(ja + ia[ieq], pos - ia[ieq], sizeof(int64_t), cmp_i64); /* distinct keys */
    }
  }
  ia[ieq] = pos;
  free(ng);
  free(ngp);
  *ia_out = ia;
  *ja_out = ja;
  return pos;
}

/* TPZFYsmpMatrix::AddKel (Matrix/pzysmp.cpp:135-176): every entry of ek with
 * |v| >= 1e-12, both triangles, sequential guess then row search. */
void orc_addkel_full(const int64_t *ia, const int64_t *ja, double *a, int ne, const double *ek,
                     const int64_t *dest) {
  int64_t k = 0;
  for (int i = 0; i < ne; i++)
    for (int j = 0; j < ne; j++) {
      int64_t ip = dest[i], jp = dest[j];
      double v = ek[i * ne + j];
      if (fabs(v) < 1.e-12) continue;
      k++;
      if (k >= ia[ip] && k < ia[ip + 1] && ja[k] == jp) {
        a[k] += v;
      } else {
        for (k = ia[ip]; k < ia[ip + 1]; k++)
          if (ja[k] == jp) { a[k] += v; break; }
      }
    }
}

int orc_assemble_full(const orc_problem *pb, const orc_mesh *m, const int64_t *ia,
                      const int64_t *ja, double *a, double *rhs) {
  memset(a, 0, sizeof(double) * ia[m->neq]);
  memset(rhs, 0, sizeof(double) * m->neq);
  double *ek = (double *)malloc(sizeof(double) * ORC_MAXNE * ORC_MAXNE);
  double ef[ORC_MAXNE];
  int64_t dest[ORC_MAXNE];
  for (int64_t e = 0; e < m->nel; e++) {
    int ne = orc_calcstiff(pb, m, e, ek, ef, dest);
    if (ne < 0) { free(ek); return -1; }
    orc_addkel_full(ia, ja, a, ne, ek, dest);
    for (int i = 0; i < ne; i++) rhs[dest[i]] += ef[i];
  }
  free(ek);
  return 0;
}

void orc_addkel(const int64_t *ia, const int64_t *ja, double *a, int ne, const double *ek,
                const int64_t *dest) {
  int64_t k = 0;
  for (int i = 0; i < ne; i++)
    for (int j = 0; j < ne; j++) {
      int64_t ip = dest[i], jp = dest[j];
      if (jp < ip) continue;
      double v = ek[i * ne + j];
      if (fabs(v) < 1.e-12) continue;
      k++;
      if (k >= ia[ip] && k < ia[ip + 1] && ja[k] == jp) {
        a[k] += v;
      } else {
        for (k = ia[ip]; k < ia[ip + 1]; k++)
          if (ja[k] == jp) { a[k] += v; break; }
      }
    }
}

int orc_assemble_range(const orc_problem *pb, const orc_mesh *m, const int64_t *ia,
                       const int64_t *ja, double *a, double *rhs, int64_t e0, int64_t e1) {
  double *ek = (double *)malloc(sizeof(double) * ORC_MAXNE * ORC_MAXNE);
  double ef[ORC_MAXNE];
  int64_t dest[ORC_MAXNE];
  for (int64_t e = e0; e < e1; e++) {
    int ne = orc_calcstiff(pb, m, e, ek, ef, dest);
    if (ne < 0) { free(ek); return -1; }
    orc_addkel(ia, ja, a, ne, ek, dest);
    for (int i = 0; i < ne; i++) rhs[dest[i]] += ef[i]; /* AddFel, Matrix/pzfmatrix.cpp:284-298 */
  }
  free(ek);
  return 0;
}

/* Elements (volume and boundary, ascending) whose destination indices hold a
 * row with rowmask[row] != 0 (ComputeDestinationIndices, Mesh/pzelmat.cpp:37-70,
 * without integration).  Returns the count written to elems (capacity cap). */
int64_t orc_elems_touching(const orc_mesh *m, const unsigned char *rowmask, int64_t *elems, int64_t cap) {
  int64_t n = 0;
  for (int64_t e = 0; e < m->nel; e++) {
    int t = m->etype[e], hit = 0;
    for (int s = 0; s < etype_nsides(t) && !hit; s++) {
      int64_t seq = m->conn_seq[m->econn[27 * e + s]];
      for (int k = 0; k < m->blk_size[seq]; k++)
        if (rowmask[m->blk_pos[seq] + k]) { hit = 1; break; }
    }
    if (hit) {
      if (n >= cap) return -1;
      elems[n++] = e;
    }
  }
  return n;
}

/* AddKel / AddFel (the drop rule of orc_addkel) restricted to the rows with
 * rowmap[row] >= 0, whose upper-CSR rows are given as a sub-CSR (ia_s over the
 * local rows, ja_s the global columns, ascending), over the elements elems[]
 * in ascending order: every entry of these rows gets the same terms in the same
 * order as in orc_assemble.  Element matrices are integrated in parallel in
 * chunks and added in element order.  Returns -1 when a column is missing from
 * its row. */
int orc_assemble_rows(const orc_problem *pb, const orc_mesh *m, const int64_t *rowmap,
                      const int64_t *ia_s, const int64_t *ja_s, double *a_s, double *rhs_s,
                      const int64_t *elems, int64_t nel, int nthreads) {
  enum { CH = 256 };
  double *ek = (double *)malloc(sizeof(double) * (size_t)CH * ORC_MAXNE * ORC_MAXNE);
  double *ef = (double *)malloc(sizeof(double) * (size_t)CH * ORC_MAXNE);
  int64_t *dest = (int64_t *)malloc(sizeof(int64_t) * (size_t)CH * ORC_MAXNE);
  int ne[CH];
  int rc = 0;
  for (int64_t c0 = 0; c0 < nel && rc == 0; c0 += CH) {
    const int nc = (int)(nel - c0 < CH ? nel - c0 : CH);
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 4)
    for (int q = 0; q < nc; q++)
      ne[q] = orc_calcstiff(pb, m, elems[c0 + q], ek + (size_t)q * ORC_MAXNE * ORC_MAXNE,
                            ef + (size_t)q * ORC_MAXNE, dest + (size_t)q * ORC_MAXNE);
    for (int q = 0; q < nc && rc == 0; q++) {
      const int n = ne[q];
      if (n < 0) { rc = -2; break; }
      const double *K = ek + (size_t)q * ORC_MAXNE * ORC_MAXNE;
      const int64_t *d = dest + (size_t)q * ORC_MAXNE;
      for (int i = 0; i < n && rc == 0; i++) {
        const int64_t lr = rowmap[d[i]];
        if (lr < 0) continue;
        rhs_s[lr] += ef[(size_t)q * ORC_MAXNE + i];
        for (int j = 0; j < n; j++) {
          if (d[j] < d[i]) continue;
          const double v = K[i * n + j];
          if (fabs(v) < 1.e-12) continue;
          int64_t lo = ia_s[lr], hi = ia_s[lr + 1];
          while (lo < hi) {
            const int64_t mid = (lo + hi) / 2;
            if (ja_s[mid] < d[j]) lo = mid + 1; else hi = mid;
          }
          if (lo == ia_s[lr + 1] || ja_s[lo] != d[j]) { rc = -1; break; }
          a_s[lo] += v;
        }
      }
    }
  }
  free(ek);
  free(ef);
  free(dest);
  return rc;
}

/* Serial_Assemble (StrMatrix/pzstrmatrixor.cpp:104-365), element order. */
/* The element loop of TPZStructMatrixOT (StrMatrix/pzstrmatrixot.cpp:
 * 92-553): elements are coloured so that no two of a colour share a connect
 * (first fit over the connects, as TPZStructMatrixOT::ElementColoring), and
 * the elements of one colour are integrated and added by nthreads threads
 * at once; colours run in order.  The sums per entry then follow the colour
 * order instead of the element order (a rounding difference).  CPU baseline
 * of the assembly only. */
int orc_assemble_colored(const orc_problem *pb, const orc_mesh *m, const int64_t *ia,
                         const int64_t *ja, double *a, double *rhs, int nthreads) {
  memset(a, 0, sizeof(double) * ia[m->neq]);
  memset(rhs, 0, sizeof(double) * m->neq);
  const int64_t nel = m->nel;
  int *color = (int *)malloc(sizeof(int) * (nel ? nel : 1));
  uint64_t *used = (uint64_t *)calloc(m->nconn ? m->nconn : 1, sizeof(uint64_t));
  int ncolor = 0;
  for (int64_t e = 0; e < nel; e++) {
    const int ns = etype_nsides(m->etype[e]);
    uint64_t msk = 0;
    for (int s = 0; s < ns; s++) msk |= used[m->econn[27 * e + s]];
    if (msk == ~0ULL) { free(color); free(used); return -1; }
    int c = __builtin_ctzll(~msk);
    color[e] = c;
    if (c + 1 > ncolor) ncolor = c + 1;
    for (int s = 0; s < ns; s++) used[m->econn[27 * e + s]] |= 1ULL << c;
  }
  free(used);
  int64_t *start = (int64_t *)calloc(ncolor + 1, sizeof(int64_t));
  int64_t *list = (int64_t *)malloc(sizeof(int64_t) * (nel ? nel : 1));
  for (int64_t e = 0; e < nel; e++) start[color[e] + 1]++;
  for (int c = 0; c < ncolor; c++) start[c + 1] += start[c];
  int64_t *fill = (int64_t *)malloc(sizeof(int64_t) * (ncolor ? ncolor : 1));
  for (int c = 0; c < ncolor; c++) fill[c] = start[c];
  for (int64_t e = 0; e < nel; e++) list[fill[color[e]]++] = e;
  free(fill);
  free(color);
  int rc = 0;
  for (int c = 0; c < ncolor; c++) {
#pragma omp parallel num_threads(nthreads)
    {
      double *ek = (double *)malloc(sizeof(double) * ORC_MAXNE * ORC_MAXNE);
      double ef[ORC_MAXNE];
      int64_t dest[ORC_MAXNE];
#pragma omp for schedule(dynamic, 64)
      for (int64_t k = start[c]; k < start[c + 1]; k++) {
        int ne = orc_calcstiff(pb, m, list[k], ek, ef, dest);
        if (ne < 0) {
#pragma omp atomic write
          rc = -1;
          continue;
        }
        orc_addkel(ia, ja, a, ne, ek, dest);
        for (int i = 0; i < ne; i++) rhs[dest[i]] += ef[i];
      }
      free(ek);
    }
  }
  free(start);
  free(list);
  return rc;
}

int orc_assemble(const orc_problem *pb, const orc_mesh *m, const int64_t *ia, const int64_t *ja,
                 double *a, double *rhs) {
  memset(a, 0, sizeof(double) * ia[m->neq]);
  memset(rhs, 0, sizeof(double) * m->neq);
  return orc_assemble_range(pb, m, ia, ja, a, rhs, 0, m->nel);
}

/* TPZSYsmpMatrix::MultAdd (Matrix/pzsysmp.cpp:189-232) with PrepareZ
 * (Matrix/pzmatrix.cpp:112-141): z = beta*y, then z += alpha*A*x mirrored. */
void orc_multadd(int64_t n, const int64_t *ia, const int64_t *ja, const double *a,
                 const double *x, const double *y, double *z, double alpha, double beta) {
  for (int64_t i = 0; i < n; i++) z[i] = (beta != 0. && y) ? beta * y[i] : 0.;
  for (int64_t r = 0; r < n; r++)
    for (int64_t iv = ia[r]; iv < ia[r + 1]; iv++) {
      int64_t c = ja[iv];
      z[r] += alpha * a[iv] * x[c];
      if (r != c) z[c] += alpha * a[iv] * x[r];
    }
}

static double dot(int64_t n, const double *x, const double *y) {
  double s = 0.;
  for (int64_t i = 0; i < n; i++) s += x[i] * y[i];
  return s;
}

static double getval_diag(const int64_t *ia, const int64_t *ja, const double *a, int64_t i) {
  for (int64_t k = ia[i]; k < ia[i + 1]; k++)
    if (ja[k] == i) return a[k];
  return 0.;
}

/* Preconditioner application z = M^-1 r of the CG below. */
typedef struct {
  int64_t n;
  const int64_t *ia, *ja;
  const double *a;
  int wasted_spmv;
  double *scr;
  int64_t nblk;          /* block form: consecutive blocks of bsize[] eqs */
  const int32_t *bsize;
  const int64_t *bpos;   /* first equation of each block                  */
  const int64_t *lpos;   /* offset of each block's LU factors in lu       */
  double *lu;
} pcg_pre;

/* EJacobi(1 iteration, tol 0) = TPZMatrix::SolveJacobi (Matrix/pzmatrix.cpp:
 * 642-667): z = r / GetVal(i,i), followed by a Residual whose result is
 * discarded (reproduced when wasted_spmv != 0). */
static void pre_jacobi(const pcg_pre *P, const double *r, double *z) {
  for (int64_t k = 0; k < P->n; k++) z[k] = r[k] / getval_diag(P->ia, P->ja, P->a, k);
  if (P->wasted_spmv) orc_multadd(P->n, P->ia, P->ja, P->a, z, r, P->scr, -1., 1.);
}

/* TPZBlockDiagonal after UpdateFrom (Matrix/pzblockdiag.cpp:567-590: the
 * consecutive diagonal blocks of the matrix) and Decompose_LU (each block by
 * TPZFMatrix::Decompose_LU, Matrix/pzfmatrix.cpp:1285-1318: no pivoting, a
 * zero pivot replaced by 1), applied by forward/back substitution. */
static void pre_block(const pcg_pre *P, const double *r, double *z) {
  for (int64_t b = 0; b < P->nblk; b++) {
    const int m = P->bsize[b];
    const int64_t e0 = P->bpos[b];
    const double *f = P->lu + P->lpos[b];  /* column-major m x m */
    double *y = z + e0;
    for (int i = 0; i < m; i++) {          /* L y = r, unit lower */
      double s = r[e0 + i];
      for (int k = 0; k < i; k++) s -= f[i + k * m] * y[k];
      y[i] = s;
    }
    for (int i = m - 1; i >= 0; i--) {     /* U z = y */
      double s = y[i];
      for (int k = i + 1; k < m; k++) s -= f[i + k * m] * y[k];
      y[i] = s / f[i + i * m];
    }
  }
}

static double getval_sym(const int64_t *ia, const int64_t *ja, const double *a, int64_t r, int64_t c) {
  if (r > c) { int64_t t = r; r = c; c = t; }
  for (int64_t k = ia[r]; k < ia[r + 1]; k++)
    if (ja[k] == c) return a[k];
  return 0.;
}

static void block_factor(pcg_pre *P) {
  for (int64_t b = 0; b < P->nblk; b++) {
    const int m = P->bsize[b];
    const int64_t e0 = P->bpos[b];
    double *f = P->lu + P->lpos[b];
    for (int i = 0; i < m; i++)
      for (int j = 0; j < m; j++) f[i + j * m] = getval_sym(P->ia, P->ja, P->a, e0 + i, e0 + j);
    for (int k = 0; k < m; k++) {
      double piv = f[k + k * m];
      if (piv == 0.) { f[k + k * m] = 1.; piv = 1.; }
      for (int i = k + 1; i < m; i++) f[i + k * m] /= piv;
      for (int j = k + 1; j < m; j++) {
        const double kj = f[k + j * m];
        for (int i = k + 1; i < m; i++) f[i + j * m] -= f[i + k * m] * kj;
      }
    }
  }
}

/* CG<> (Solvers/LinearSolvers/cg.h:29-132) with the preconditioner pre. */
static int pcg_core(int64_t n, const int64_t *ia, const int64_t *ja, const double *a, const double *b,
                    double *x, int64_t *iters, double *tol, double *res_history,
                    void (*pre)(const pcg_pre *, const double *, double *), const pcg_pre *P) {
  double *r = (double *)malloc(n * sizeof(double)), *z = (double *)malloc(n * sizeof(double));
  double *p = (double *)malloc(n * sizeof(double)), *q = (double *)malloc(n * sizeof(double));
  int64_t maxit = *iters;
  double normb = sqrt(dot(n, b, b)), resid, rho, rho1 = 0., alpha, beta;
  for (int64_t i = 0; i < n; i++) { x[i] = 0.; r[i] = b[i]; }
  if (normb == 0.) normb = 1.;
  int ret = 1;
  if ((resid = sqrt(dot(n, r, r)) / normb) <= *tol) {
    *tol = resid;
    *iters = 0;
    ret = 0;
    goto done;
  }
  int64_t i;
  for (i = 1; i <= maxit; i++) {
    pre(P, r, z);
    rho = dot(n, r, z);
    if (i == 1) memcpy(p, z, n * sizeof(double));
    else {
      beta = rho / rho1;
      for (int64_t k = 0; k < n; k++) p[k] = beta * p[k] + z[k];
    }
    orc_multadd(n, ia, ja, a, p, NULL, q, 1., 0.);
    alpha = rho / dot(n, p, q);
    for (int64_t k = 0; k < n; k++) x[k] += alpha * p[k];
    for (int64_t k = 0; k < n; k++) r[k] += -alpha * q[k];
    resid = sqrt(dot(n, r, r)) / normb;
    if (res_history) res_history[i - 1] = resid;
    if (resid <= *tol) {
      *tol = resid;
      *iters = i;
      ret = 0;
      goto done;
    }
    rho1 = rho;
  }
  *tol = resid;
  *iters = maxit;
done:
  free(r); free(z); free(p); free(q);
  return ret;
}

/* CG<> preconditioned by a TPZStepSolver EJacobi(1, 0, 0) (pre_jacobi). */
int orc_pcg(int64_t n, const int64_t *ia, const int64_t *ja, const double *a, const double *b,
            double *x, int64_t *iters, double *tol, int wasted_spmv, double *res_history) {
  pcg_pre P;
  memset(&P, 0, sizeof(P));
  P.n = n; P.ia = ia; P.ja = ja; P.a = a; P.wasted_spmv = wasted_spmv;
  P.scr = (double *)malloc((n > 0 ? n : 1) * sizeof(double));
  int ret = pcg_core(n, ia, ja, a, b, x, iters, tol, res_history, pre_jacobi, &P);
  free(P.scr);
  return ret;
}

/* CG<> preconditioned by the TPZStepSolver(TPZBlockDiagonal) with ELU that
 * TPZAnalysis::BuildPreconditioner(EBlockJacobi, true) returns
 * (Analysis/TPZAnalysis.cpp:1062-1070); the block sizes are those of the
 * connects in sequence order (TPZBlockDiagonalStructMatrix::BlockSizes,
 * StrMatrix/pzbdstrmatrix.cpp:59-93) and the blocks are taken from the
 * matrix when the solve starts (TPZStepSolver::Solve -> UpdateFrom,
 * Solvers/pzstepsolver.cpp:72-76). */
int orc_pcg_block(int64_t n, const int64_t *ia, const int64_t *ja, const double *a, const double *b,
                  int64_t nblk, const int32_t *bsize, double *x, int64_t *iters, double *tol,
                  double *res_history) {
  pcg_pre P;
  memset(&P, 0, sizeof(P));
  P.n = n; P.ia = ia; P.ja = ja; P.a = a; P.nblk = nblk; P.bsize = bsize;
  int64_t *bpos = (int64_t *)malloc((nblk + 1) * sizeof(int64_t));
  int64_t *lpos = (int64_t *)malloc((nblk + 1) * sizeof(int64_t));
  bpos[0] = lpos[0] = 0;
  for (int64_t k = 0; k < nblk; k++) {
    bpos[k + 1] = bpos[k] + bsize[k];
    lpos[k + 1] = lpos[k] + (int64_t)bsize[k] * bsize[k];
  }
  if (bpos[nblk] != n) {
    free(bpos); free(lpos);
    return -1;
  }
  P.bpos = bpos; P.lpos = lpos;
  P.lu = (double *)malloc((lpos[nblk] > 0 ? lpos[nblk] : 1) * sizeof(double));
  block_factor(&P);
  int ret = pcg_core(n, ia, ja, a, b, x, iters, tol, res_history, pre_block, &P);
  free(P.lu); free(bpos); free(lpos);
  return ret;
}

/* nnz_full = S^d s^2, S = 2(p+1) + (n-1)(2p+1) + n(p-1)(p+1) per direction
 * (1-D coupling count); nnz_upper = (nnz_full + N)/2. */
int64_t orc_nnz_closed_form(int dim, int p, int nstate, int nx, int ny, int nz) {
  int64_t n[3] = {nx, ny, nz}, full = 1, N = 1;
  for (int d = 0; d < dim; d++) {
    int64_t S = 2 * (p + 1) + (n[d] - 1) * (2 * p + 1) + n[d] * (p - 1) * (p + 1);
    full *= S;
    N *= p * n[d] + 1;
  }
  full *= (int64_t)nstate * nstate;
  N *= nstate;
  return (full + N) / 2;
}

/* ------------------------------------------------------------------------ */
/* Post-processing: TPZCompMesh::EvaluateError (Mesh/pzcmesh.cpp:1497-1552) */
/* over TPZInterpolationSpace::EvaluateError (Mesh/pzinterpolationspace.cpp: */
/* 1108-1207) with the materials' Errors.                                    */
/* ------------------------------------------------------------------------ */

/* Order of the error rule (EvaluateError, :1140-1160): the element rule's
 * order raised to the exact solution's order, unless that order reaches the
 * rule family's GetMaxOrder (36 for TPZInt1d / TPZIntQuad / TPZIntCube3D,
 * Integral/pzquad.h; 21 TPZIntTriang, Integral/tpzintrulet.h:72; 14
 * TPZIntTetra3D, Integral/tpzintrulet3d.h:72), which is then used. */
static int error_order(int t, int rule, int porder) {
  int mx = t == ORC_TRI ? 21 : t == ORC_TET ? 14 : 36;
  if (mx > porder) return rule > porder ? rule : porder;
  return mx;
}

/* Points (xi[3q]) and weights of element iel's error rule; malloc'ed. */
static int error_rule(const orc_problem *pb, const orc_mesh *m, int64_t iel, int porder,
                      double **xi_out, double **w_out) {
  int t = m->etype[iel], dim = etype_dim(t), maxorder = -1;
  for (int s = 0; s < etype_nsides(t); s++) {
    int64_t c = m->econn[27 * iel + s];
    if (m->conn_order[c] > maxorder) maxorder = m->conn_order[c];
  }
  int order = error_order(t, rule_order(pb, 0, maxorder), porder);
  double *xi = NULL, *w = NULL;
  int nq;
  if (is_simplex(t)) {
    xi = malloc(sizeof(double) * 3 * 4096);
    w = malloc(sizeof(double) * 4096);
    nq = orc_simplex_rule(t, order, xi, w);
    if (nq < 0) { free(xi); free(w); return -1; }
  } else {
    double gl[64], gw[64];
    int ng = orc_gauss_legendre(order, gl, gw);
    nq = 1;
    for (int d = 0; d < dim; d++) nq *= ng;
    xi = malloc(sizeof(double) * 3 * nq);
    w = malloc(sizeof(double) * nq);
    for (int q = 0; q < nq; q++) {
      double *x = xi + 3 * q;
      x[0] = x[1] = x[2] = 0.;
      if (dim == 1) { x[0] = gl[q]; w[q] = gw[q]; }
      else if (dim == 2) { /* Integral/pzquad.cpp:159-169 */
        int ik = q / ng, ie = q % ng;
        x[0] = gl[ik]; x[1] = gl[ie]; w[q] = gw[ik] * gw[ie];
      } else { /* Integral/pzquad.cpp:274-287 */
        int ik = q % ng, ie = (q % (ng * ng)) / ng, iz = q / (ng * ng);
        x[0] = gl[ik]; x[1] = gl[ie]; x[2] = gl[iz]; w[q] = gw[ik] * gw[ie] * gw[iz];
      }
    }
  }
  *xi_out = xi;
  *w_out = w;
  return nq;
}

/* Geometry at xi: x (TPZGeoEl::X), detjac, jacinv and axes (jacobian()). */
static double elem_geometry(const orc_mesh *m, int64_t iel, const double *xi, double *x,
                            double jinv[3][3], double axes[3][3]) {
  int t = m->etype[iel], dim = etype_dim(t), nc = etype_ncorner(t);
  double cphi[27], cdphi[81], gradx[3][3] = {{0}};
  orc_shape(t, 1, xi, cphi, cdphi);
  x[0] = x[1] = x[2] = 0.;
  for (int n = 0; n < nc; n++)
    for (int i = 0; i < 3; i++) {
      double c = m->xyz[3 * m->enodes[8 * iel + n] + i];
      x[i] += c * cphi[n];
      for (int j = 0; j < dim; j++) gradx[i][j] += c * cdphi[j * nc + n];
    }
  memset(jinv, 0, 9 * sizeof(double));
  return jacobian(dim, gradx, jinv, axes);
}

/* Shapes of element iel at xi as in orc_calcstiff; returns nshape. */
static int elem_shapes(const orc_mesh *m, int64_t iel, const double *xi, double *phi, double *dphi) {
  int t = m->etype[iel], dim = etype_dim(t), nc = etype_ncorner(t), nsh = 0;
  if (is_simplex(t)) {
    int64_t ids[4];
    int sord[27];
    for (int n = 0; n < nc; n++) ids[n] = m->enodes[8 * iel + n];
    for (int s = nc; s < etype_nsides(t); s++) sord[s - nc] = m->conn_order[m->econn[27 * iel + s]];
    return orc_shape_h1(t, ids, sord, xi, phi, dphi);
  }
  int keep[27], nkeep = 0;
  for (int s = 0; s < etype_nsides(t); s++)
    if (m->conn_nshape[m->econn[27 * iel + s]] > 0) keep[nkeep++] = s;
  double phiall[27], dphiall[81];
  int nall = orc_shape(t, m->p, xi, phiall, dphiall), sel[27];
  for (int k = 0; k < nkeep; k++)
    if (keep[k] < nall) sel[nsh++] = keep[k];
  for (int i = 0; i < nsh; i++) {
    phi[i] = phiall[sel[i]];
    for (int a = 0; a < dim; a++) dphi[a * nsh + i] = dphiall[a * nall + sel[i]];
  }
  return nsh;
}

/* Scalar exact solutions of TLaplaceExample1::uxy (Pre/TPZAnalyticSolution.
 * cpp:1254-1310: EConst, EX, ESinSin, ECosCos) over the first dim
 * coordinates, with their gradients (graduxy, :1927-1943). */
static void exact_laplace(const orc_exact *ex, int dim, const double *x, double *u, double *du) {
  const double pi = 3.14159265358979323846;
  double xl[3];
  for (int i = 0; i < 3; i++) { xl[i] = x[i] - ex->center[i]; du[i] = 0.; }
  if (ex->kind == ORC_EXACT_CONST) { *u = ex->constant; return; }
  if (ex->kind == ORC_EXACT_X) { *u = xl[0]; du[0] = 1.; return; }
  double f[3], df[3];
  for (int i = 0; i < dim; i++) {
    if (ex->kind == ORC_EXACT_SINSIN) { f[i] = sin(pi * xl[i]); df[i] = pi * cos(pi * xl[i]); }
    else { f[i] = cos(pi * xl[i] / 2); df[i] = -pi / 2 * sin(pi * xl[i] / 2); }
  }
  *u = 1.;
  for (int i = 0; i < dim; i++) *u *= f[i];
  for (int j = 0; j < dim; j++) {
    du[j] = df[j];
    for (int i = 0; i < dim; i++)
      if (i != j) du[j] *= f[i];
  }
}

/* TPZMatPoisson::Errors (Material/Poisson/TPZMatPoisson.cpp:173-218) */
static void errors_poisson(int dim, double u, const double *gradu, double ue, const double *due,
                           double *v) {
  double diff = u - ue;
  v[1] = diff * diff;
  v[2] = 0.;
  for (int i = 0; i < dim; i++) {
    diff = gradu[i] - due[i];
    v[2] += diff * diff;
  }
  v[0] = v[1] + v[2];
}

/* TPZDarcyFlow::Errors (Material/DarcyFlow/TPZDarcyFlow.cpp:268-326) */
static void errors_darcy(int dim, double perm, double p, const double *gradu, double pe,
                         const double *flux, double *v) {
  for (int k = 0; k < 6; k++) v[k] = 0.;
  double diff = fabs(p - pe);
  v[1] = diff * diff;
  double gd[3] = {0, 0, 0};
  for (int i = 0; i < dim; i++) gd[i] += fabs(gradu[i] - flux[i]);
  diff = 0;
  for (int i = 0; i < dim; i++) diff += perm * gd[i];
  v[2] += fabs(diff * diff);
  v[0] = v[1] + v[2];
  for (int i = 0; i < dim; i++) {
    double fs = -perm * gradu[i];
    diff = fabs(flux[i] - fs);
    v[3 + i] = diff * diff;
  }
}

/* TPZElasticity3D::Errors (Material/Elasticity/TPZElasticity3D.cpp:1081-1133)
 * with ComputeStressTensor / ComputeStressVector (:1152-1192), fPreStress 0.
 * dudx[i][j] = d u_j / d x_i. */
static void errors_elast(double E, double nu, const double *u, double dudx[3][3], const double *ue,
                         double due[3][3], double *v) {
  double L2 = 0., semi = 0., d[3][3];
  for (int i = 0; i < 3; i++) L2 += (u[i] - ue[i]) * (u[i] - ue[i]);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      d[i][j] = dudx[i][j] - due[i][j];
      semi += d[i][j] * d[i][j];
    }
  double const1 = -1. + nu, const2 = -1. + nu + 2. * nu * nu, const3 = 2. * (1. + nu), vec[6];
  vec[0] = E * (d[0][0] * const1 - (d[1][1] + d[2][2]) * nu) / const2;
  vec[1] = E * (d[1][1] * const1 - (d[0][0] + d[2][2]) * nu) / const2;
  vec[2] = E * (d[2][2] * const1 - (d[0][0] + d[1][1]) * nu) / const2;
  vec[3] = E * (d[1][0] + d[0][1]) / const3;
  vec[4] = E * (d[2][0] + d[0][2]) / const3;
  vec[5] = E * (d[2][1] + d[1][2]) / const3;
  double st[3][3] = {{vec[0], vec[3], vec[4]}, {vec[3], vec[1], vec[5]}, {vec[4], vec[5], vec[2]}};
  double energy = 0.;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) energy += d[i][j] * st[i][j];
  v[1] = L2;
  v[2] = semi;
  v[0] = energy;
}

int64_t orc_error_points(const orc_problem *pb, const orc_mesh *m, int porder, double *xq) {
  int64_t np = 0;
  for (int64_t iel = 0; iel < m->nvol; iel++) {
    double *xi, *w;
    int nq = error_rule(pb, m, iel, porder, &xi, &w);
    if (nq < 0) return -1;
    for (int q = 0; q < nq && xq; q++) {
      double jinv[3][3], axes[3][3];
      elem_geometry(m, iel, xi + 3 * q, xq + 3 * (np + q), jinv, axes);
    }
    np += nq;
    free(xi);
    free(w);
  }
  return np;
}

int orc_evaluate_error(const orc_problem *pb, const orc_mesh *m, const double *sol,
                       const orc_exact *ex, double *el_err, double *errors) {
  if (pb->material == ORC_MOCK1D) return -1;
  int ns = m->nstate, dim = m->dim, nerr = pb->material == ORC_DARCY ? 6 : 3;
  if (ex->kind != ORC_EXACT_POINTS && ns != 1) return -1;
  double sum[6] = {0, 0, 0, 0, 0, 0};
  int64_t np = 0;
  for (int64_t iel = 0; iel < m->nvol; iel++) {
    /* element solution (TPZInterpolationSpace::ComputeSolution): dofs in
     * connect order, shape-major (Mesh/pzelmat.cpp:37-70) */
    double uel[ORC_MAXNE];
    int t = m->etype[iel], ne = 0;
    for (int s = 0; s < etype_nsides(t); s++) {
      int64_t seq = m->conn_seq[m->econn[27 * iel + s]];
      for (int k = 0; k < m->blk_size[seq]; k++) uel[ne++] = sol[m->blk_pos[seq] + k];
    }
    double coef = pb->el_coef ? pb->el_coef[iel] : (pb->material == ORC_ELAST3D ? pb->E : pb->scale);
    double *xi, *w;
    int nq = error_rule(pb, m, iel, ex->porder, &xi, &w);
    if (nq < 0) return -1;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (int q = 0; q < nq; q++) {
      double phi[64], dphi[192], x[3], jinv[3][3], axes[3][3];
      int nsh = elem_shapes(m, iel, xi + 3 * q, phi, dphi);
      double detjac = elem_geometry(m, iel, xi + 3 * q, x, jinv, axes);
      /* sol and dsol in axes coordinates, then Axes2XYZ (Util/pzaxestools.h:
       * 61-86): dudx(i,k) = sum_a axes(a,i) dudaxes(a,k) */
      double u[3] = {0, 0, 0}, dax[3][3] = {{0}}, dudx[3][3] = {{0}};
      for (int i = 0; i < nsh; i++) {
        double dx[3] = {0, 0, 0};
        for (int a = 0; a < dim; a++)
          for (int b = 0; b < dim; b++) dx[a] += jinv[b][a] * dphi[b * nsh + i];
        for (int k = 0; k < ns; k++) {
          u[k] += phi[i] * uel[i * ns + k];
          for (int a = 0; a < dim; a++) dax[a][k] += dx[a] * uel[i * ns + k];
        }
      }
      for (int i = 0; i < 3; i++)
        for (int k = 0; k < ns; k++)
          for (int a = 0; a < dim; a++) dudx[i][k] += axes[a][i] * dax[a][k];
      double ue[3] = {0, 0, 0}, due[3][3] = {{0}};
      if (ex->kind == ORC_EXACT_POINTS) {
        int64_t p = np + q;
        for (int k = 0; k < ns; k++) {
          ue[k] = ex->u[p * ns + k];
          for (int d = 0; d < dim; d++) due[d][k] = ex->du[(p * dim + d) * ns + k];
        }
      } else {
        double g[3];
        exact_laplace(ex, dim, x, &ue[0], g);
        for (int d = 0; d < 3; d++) due[d][0] = g[d];
      }
      double v[6] = {0, 0, 0, 0, 0, 0};
      if (pb->material == ORC_POISSON) {
        double gu[3] = {dudx[0][0], dudx[1][0], dudx[2][0]}, ge[3] = {due[0][0], due[1][0], due[2][0]};
        errors_poisson(dim, u[0], gu, ue[0], ge, v);
      } else if (pb->material == ORC_DARCY) {
        double gu[3] = {dudx[0][0], dudx[1][0], dudx[2][0]}, ge[3] = {due[0][0], due[1][0], due[2][0]};
        errors_darcy(dim, coef, u[0], gu, ue[0], ge, v);
      } else {
        errors_elast(coef, pb->nu, u, dudx, ue, due, v);
      }
      double weight = w[q] * fabs(detjac);
      for (int k = 0; k < nerr; k++) acc[k] += weight * v[k];
    }
    free(xi);
    free(w);
    np += nq;
    for (int k = 0; k < nerr; k++) {
      double e = sqrt(acc[k]);
      if (el_err) el_err[iel * nerr + k] = e;
      sum[k] += e * e;
    }
  }
  for (int k = 0; k < nerr && errors; k++) errors[k] = sqrt(sum[k]);
  return nerr;
}

/* ------------------------------------------------------------------------ */
/* Bandwidth renumbering of TPZAnalysis::OptimizeBandwidth (Analysis/        */
/* TPZAnalysis.cpp:225-300) with the default TPZSloanRenumbering (RENUMBER,  */
/* :74-78): TPZSloanRenumbering::Resequence (External/                       */
/* TPZSloanRenumbering.cpp:23-161) on the ConvertGraph node graph, with the  */
/* SGraph helpers of External/TPZCutHillMcKee.cpp:16-240.                    */
/* ------------------------------------------------------------------------ */

static int64_t sg_degree(const int64_t *xadj, int64_t v) { return xadj[v + 1] - xadj[v]; }

/* RootedLevelStructure (TPZCutHillMcKee.cpp:93-122): breadth-first levels
 * from root; a node joins the first level that reaches it, in the order of
 * its parents and of their (sorted) adjacency.  lv[] holds the nodes level
 * after level, lptr[k] the start of level k; returns the number of levels. */
static int64_t sg_levels(int64_t n, const int64_t *xadj, const int64_t *adj, int64_t root,
                         int64_t *lv, int64_t *lptr, char *seen) {
  memset(seen, 0, n);
  int64_t nl = 0, head = 0, tail = 0;
  lv[tail++] = root;
  seen[root] = 1;
  while (head < tail) {
    int64_t end = tail;
    lptr[nl++] = head;
    for (; head < end; head++) {
      int64_t v = lv[head];
      for (int64_t k = xadj[v]; k < xadj[v + 1]; k++)
        if (!seen[adj[k]]) { seen[adj[k]] = 1; lv[tail++] = adj[k]; }
    }
  }
  lptr[nl] = tail;
  return nl;
}

/* PseudoPeripheralNodes (TPZCutHillMcKee.cpp:165-240) with SmallestDegree
 * (:16-31), SortNodes (:124-137, stable by degree) and ShrinkLastLevel
 * (:139-163, the last node of each degree, at most (size+2)/2 of them). */
static void sg_peripheral(int64_t n, const int64_t *xadj, const int64_t *adj, int64_t *start,
                          int64_t *endn) {
  int64_t *lv = malloc(sizeof(int64_t) * n), *lptr = malloc(sizeof(int64_t) * (n + 1));
  int64_t *lv2 = malloc(sizeof(int64_t) * n), *lptr2 = malloc(sizeof(int64_t) * (n + 1));
  int64_t *last = malloc(sizeof(int64_t) * n);
  char *seen = malloc(n);
  int64_t mindeg = 1000000, s = -1;
  for (int64_t i = 0; i < n; i++)
    if (sg_degree(xadj, i) < mindeg) { mindeg = sg_degree(xadj, i); s = i; }
  int64_t e = -1;
  int64_t nlev = sg_levels(n, xadj, adj, s, lv, lptr, seen);
  int count = 0;
  while (count < 10 || e == -1) {
    count++;
    int64_t cps = s, cpe = e;
    if (!nlev) break;
    int64_t nlast = lptr[nlev] - lptr[nlev - 1];
    memcpy(last, lv + lptr[nlev - 1], sizeof(int64_t) * nlast);
    /* stable sort by degree (insertion sort keeps equal keys in order) */
    for (int64_t i = 1; i < nlast; i++) {
      int64_t x = last[i], dx = sg_degree(xadj, x), j = i - 1;
      while (j >= 0 && sg_degree(xadj, last[j]) > dx) { last[j + 1] = last[j]; j--; }
      last[j + 1] = x;
    }
    /* one node per degree, the last of each run; then at most (n+2)/2 */
    int64_t nk = 0;
    for (int64_t i = 0; i < nlast; i++) {
      if (i + 1 < nlast && sg_degree(xadj, last[i + 1]) == sg_degree(xadj, last[i])) continue;
      last[nk++] = last[i];
    }
    if ((nlast + 2) / 2 < nk) nk = (nlast + 2) / 2;
    int64_t we = 1000000000L, hs = nlev;
    for (int64_t iq = 0; iq < nk; iq++) {
      int64_t q = last[iq];
      int64_t h = sg_levels(n, xadj, adj, q, lv2, lptr2, seen), w = 0;
      for (int64_t j = 0; j < h; j++)
        if (lptr2[j + 1] - lptr2[j] > w) w = lptr2[j + 1] - lptr2[j];
      if (h > hs && w < we) {
        s = q;
        memcpy(lv, lv2, sizeof(int64_t) * n);
        memcpy(lptr, lptr2, sizeof(int64_t) * (h + 1));
        nlev = h;
        break;
      } else if (w < we) {
        e = q;
        we = w;
      }
    }
    if (cps == s && cpe == e) break;
  }
  *start = s;
  *endn = e;
  free(lv); free(lptr); free(lv2); free(lptr2); free(last); free(seen);
}

int orc_sloan(int64_t n, const int64_t *xadj, const int64_t *adj, int64_t *perm) {
  enum { INACTIVE = 0, PREACTIVE = 1, ACTIVE = 2, POSTACTIVE = 3 };
  const int64_t W1 = 1, W2 = 2;
  if (n <= 0) return 0;
  int64_t s, e;
  sg_peripheral(n, xadj, adj, &s, &e);
  int64_t *lv = malloc(sizeof(int64_t) * n), *lptr = malloc(sizeof(int64_t) * (n + 1));
  char *seen = malloc(n);
  int64_t nlev = sg_levels(n, xadj, adj, e, lv, lptr, seen);
  int64_t *prio = malloc(sizeof(int64_t) * n), *dist = malloc(sizeof(int64_t) * n);
  for (int64_t i = 0; i < n; i++) dist[i] = -1;
  for (int64_t l = 0; l < nlev; l++)
    for (int64_t k = lptr[l]; k < lptr[l + 1]; k++) dist[lv[k]] = l;
  int ok = 1;
  for (int64_t i = 0; i < n; i++) {
    if (dist[i] < 0) ok = 0;
    prio[i] = W1 * dist[i] - W2 * (sg_degree(xadj, i) + 1);
  }
  /* SList: list without duplicates, a node enters at most once (history);
   * remove(i) moves the last entry into slot i */
  int64_t *Q = malloc(sizeof(int64_t) * n), qn = 0, nr = 0;
  char *hist = calloc(n, 1), *st = calloc(n, 1);
#define QPUSH(v) (hist[v] ? 0 : (hist[v] = 1, Q[qn++] = (v), 1))
  st[s] = PREACTIVE;
  QPUSH(s);
  while (ok && qn) {
    int64_t qi = 0, cur = Q[0];
    for (int64_t i = 1; i < qn; i++)
      if (prio[Q[i]] > prio[cur]) { cur = Q[i]; qi = i; }
    Q[qi] = Q[qn - 1];
    qn--;
    if (st[cur] == PREACTIVE)
      for (int64_t k = xadj[cur]; k < xadj[cur + 1]; k++) {
        int64_t a = adj[k];
        prio[a] += W2;
        if (QPUSH(a) && st[a] == INACTIVE) st[a] = PREACTIVE;
      }
    perm[cur] = nr++;
    st[cur] = POSTACTIVE;
    for (int64_t k = xadj[cur]; k < xadj[cur + 1]; k++) {
      int64_t a = adj[k];
      if (st[a] != PREACTIVE) continue;
      st[a] = ACTIVE;
      prio[a] += W2;
      for (int64_t j = xadj[a]; j < xadj[a + 1]; j++) {
        int64_t b = adj[j];
        if (st[b] != POSTACTIVE) prio[b] += W2;
        if (QPUSH(b) && st[b] == INACTIVE) st[b] = PREACTIVE;
      }
    }
  }
#undef QPUSH
  free(lv); free(lptr); free(seen); free(prio); free(dist); free(Q); free(hist); free(st);
  return (ok && nr == n) ? 0 : -1;
}

int orc_mesh_sloan(const orc_mesh *m, int64_t *perm) {
  int64_t *ngp, *ng;
  node_graph(m, &ngp, &ng);
  int rc = orc_sloan(m->nconn, ngp, ng, perm);
  free(ngp);
  free(ng);
  return rc;
}

/* ------------------------------------------------------------------------ */
/* Meshes handed over as arrays, and constrained connects (hanging sides):   */
/* TPZConnect dependencies (TPZConnect::TPZDepend, Mesh/pzconnect.h), the    */
/* element's constrained matrix of TPZElementMatrixT::ApplyConstraints       */
/* (Mesh/TPZElementMatrixT.cpp:134-322) and its assembly                     */
/* (StrMatrix/pzstrmatrixor.cpp:273-280).                                    */
/* ------------------------------------------------------------------------ */

int orc_mesh_from_arrays(int dim, int p, int nstate, int vt, int bt, int64_t nnodes,
                         const double *xyz, int64_t nvol, const int64_t *vol_nodes,
                         const int64_t *vol_conn, int64_t nbc, const int64_t *bc_nodes,
                         const int64_t *bc_conn, const int *bc_face, int64_t nconn,
                         const int64_t *blk_pos, const int *blk_size, const int *conn_order,
                         int64_t neq, orc_mesh *m) {
  memset(m, 0, sizeof(*m));
  int ncv = etype_ncorner(vt), nsv = etype_nsides(vt), ncb = etype_ncorner(bt), nsb = etype_nsides(bt);
  m->dim = dim; m->p = p; m->nstate = nstate;
  m->nnodes = nnodes;
  m->xyz = malloc(sizeof(double) * 3 * (nnodes ? nnodes : 1));
  memcpy(m->xyz, xyz, sizeof(double) * 3 * nnodes);
  m->nvol = nvol;
  m->nel = nvol + nbc;
  int64_t ne = m->nel ? m->nel : 1;
  m->etype = malloc(sizeof(int) * ne);
  m->bcface = malloc(sizeof(int) * ne);
  m->enodes = calloc(8 * ne, sizeof(int64_t));
  m->econn = calloc(27 * ne, sizeof(int64_t));
  for (int64_t e = 0; e < m->nel; e++) {
    int isbc = e >= nvol;
    int64_t k = isbc ? e - nvol : e;
    m->etype[e] = isbc ? bt : vt;
    m->bcface[e] = isbc ? bc_face[k] : -1;
    for (int n = 0; n < (isbc ? ncb : ncv); n++)
      m->enodes[8 * e + n] = isbc ? bc_nodes[k * ncb + n] : vol_nodes[k * ncv + n];
    for (int sd = 0; sd < (isbc ? nsb : nsv); sd++) {
      int64_t c = isbc ? bc_conn[k * nsb + sd] : vol_conn[k * nsv + sd];
      if (c < 0 || c >= nconn) return -1;
      m->econn[27 * e + sd] = c;
    }
  }
  m->nconn = nconn;
  int64_t nc = nconn ? nconn : 1;
  m->conn_order = malloc(sizeof(int) * nc);
  m->conn_nshape = malloc(sizeof(int) * nc);
  m->conn_seq = malloc(sizeof(int64_t) * nc);
  m->blk_pos = malloc(sizeof(int64_t) * nc);
  m->blk_size = malloc(sizeof(int) * nc);
  for (int64_t c = 0; c < nconn; c++) {
    m->conn_order[c] = conn_order[c];
    m->conn_nshape[c] = blk_size[c] / nstate;
    m->conn_seq[c] = c;
    m->blk_pos[c] = blk_pos[c];
    m->blk_size[c] = blk_size[c];
  }
  m->neq = neq;
  return 0;
}

static int cmp_i64b(const void *a, const void *b) {
  int64_t x = *(const int64_t *)a, y = *(const int64_t *)b;
  return x < y ? -1 : x > y;
}

/* TPZConnect::BuildConnectList: the element's connects and, recursively, the
 * connects they depend on; sorted, unique (std::set<int64_t>). */
static int constr_list(const orc_mesh *m, int64_t e, const int64_t *dep_ptr, const int64_t *dep_conn,
                       int64_t *out) {
  int n = 0;
  for (int sd = 0; sd < etype_nsides(m->etype[e]); sd++) out[n++] = m->conn_seq[m->econn[27 * e + sd]];
  for (int i = 0; i < n; i++) {  /* the list grows while it is scanned: closure */
    int64_t c = out[i];
    for (int64_t k = dep_ptr[c]; k < dep_ptr[c + 1]; k++) {
      int have = 0;
      for (int j = 0; j < n; j++) have |= out[j] == dep_conn[k];
      if (!have) out[n++] = dep_conn[k];
    }
  }
  qsort(out, n, sizeof(int64_t), cmp_i64b);
  int u = 0;
  for (int i = 0; i < n; i++)
    if (!u || out[i] != out[u - 1]) out[u++] = out[i];
  return u;
}

/* SetupMatrixData over TPZCompMesh::ComputeElGraph's lists (Mesh/pzcmesh.cpp:
 * 1223-1266): every element's BuildConnectList without the constrained
 * connects. */
int64_t orc_setup_pattern_constrained(const orc_mesh *m, const int64_t *dep_ptr,
                                      const int64_t *dep_conn, int64_t **ia_out, int64_t **ja_out) {
  int64_t *ep = malloc((m->nel + 1) * sizeof(int64_t)), n = 0, cap = 1024;
  int64_t *el = malloc(cap * sizeof(int64_t)), buf[512];
  char *skip = calloc(m->nconn ? m->nconn : 1, 1);
  for (int64_t c = 0; c < m->nconn; c++) {
    skip[c] = dep_ptr[c + 1] > dep_ptr[c];
    if (skip[c] && m->blk_size[c] && m->blk_pos[c] < m->neq) {  /* constrained blocks come last */
      free(ep); free(el); free(skip);
      return -1;
    }
  }
  ep[0] = 0;
  for (int64_t e = 0; e < m->nel; e++) {
    int k = constr_list(m, e, dep_ptr, dep_conn, buf);
    for (int i = 0; i < k; i++) {
      if (skip[buf[i]]) continue;
      if (n == cap) { cap *= 2; el = realloc(el, cap * sizeof(int64_t)); }
      el[n++] = buf[i];
    }
    ep[e + 1] = n;
  }
  int64_t *ngp, *ng;
  node_graph_lists(m->nconn, m->nel, ep, el, &ngp, &ng);
  int64_t nnz = pattern_csr(m, ngp, ng, skip, ia_out, ja_out);
  free(ngp); free(ng); free(ep); free(el); free(skip);
  return nnz;
}

/* TPZConnect::SetDependenceOrder (Mesh/pzconnect.cpp:558-600): the connects
 * a connect of order `order` depends on get at least order + 1. */
static void set_dep_order(int64_t c, int order, const int64_t *dep_ptr, const int64_t *dep_conn,
                          const int64_t *list, int n, int *dord) {
  for (int64_t k = dep_ptr[c]; k < dep_ptr[c + 1]; k++) {
    int64_t t = dep_conn[k];
    for (int i = 0; i < n; i++)
      if (list[i] == t) {
        if (dord[i] < order + 1) dord[i] = order + 1;
        set_dep_order(t, order + 1, dep_ptr, dep_conn, list, n, dord);
      }
  }
}

int orc_assemble_constrained(const orc_problem *pb, const orc_mesh *m, const int64_t *dep_ptr,
                             const int64_t *dep_conn, const double *dep_coef, const int64_t *ia,
                             const int64_t *ja, double *a, double *rhs) {
  memset(a, 0, sizeof(double) * ia[m->neq]);
  memset(rhs, 0, sizeof(double) * m->neq);
  const int ns = m->nstate;
  double *ek = malloc(sizeof(double) * ORC_MAXNE * ORC_MAXNE);
  double ef[ORC_MAXNE];
  int64_t dest[ORC_MAXNE];
  int64_t list[512], pos[513], kdest[ORC_MAXNE];
  int dord[512];
  double *kc = malloc(sizeof(double) * 4 * ORC_MAXNE * ORC_MAXNE);
  double *kr = malloc(sizeof(double) * 4 * ORC_MAXNE * ORC_MAXNE);
  double fc[4 * ORC_MAXNE], fr[4 * ORC_MAXNE];
  for (int64_t e = 0; e < m->nel; e++) {
    int ne = orc_calcstiff(pb, m, e, ek, ef, dest);
    if (ne < 0) { free(ek); free(kc); free(kr); return -1; }
    int t = m->etype[e], nsd = etype_nsides(t), hasdep = 0;
    for (int sd = 0; sd < nsd; sd++) {
      int64_t c = m->conn_seq[m->econn[27 * e + sd]];
      hasdep |= dep_ptr[c + 1] > dep_ptr[c] && m->blk_size[c] > 0;
    }
    if (!hasdep) {
      orc_addkel(ia, ja, a, ne, ek, dest);
      for (int i = 0; i < ne; i++) rhs[dest[i]] += ef[i];
      continue;
    }
    /* constrained block structure over the sorted connect list */
    int n = constr_list(m, e, dep_ptr, dep_conn, list);
    pos[0] = 0;
    for (int i = 0; i < n; i++) pos[i + 1] = pos[i] + m->blk_size[list[i]];
    int64_t tot = pos[n];
    if (tot > 4 * ORC_MAXNE) { free(ek); free(kc); free(kr); return -1; }
    memset(kc, 0, sizeof(double) * tot * tot);
    memset(fc, 0, sizeof(double) * tot);
    /* copy: element block of side sd -> block of its connect in the list */
    int epos[27], eidx[27], off = 0;
    for (int sd = 0; sd < nsd; sd++) {
      int64_t c = m->conn_seq[m->econn[27 * e + sd]];
      int i = 0;
      while (i < n && list[i] != c) i++;
      epos[sd] = off;
      eidx[sd] = i;
      off += m->blk_size[c];
    }
    for (int si = 0; si < nsd; si++) {
      int bi = m->blk_size[m->conn_seq[m->econn[27 * e + si]]];
      for (int r = 0; r < bi; r++) {
        fc[pos[eidx[si]] + r] = ef[epos[si] + r];
        for (int sj = 0; sj < nsd; sj++) {
          int bj = m->blk_size[m->conn_seq[m->econn[27 * e + sj]]];
          for (int q = 0; q < bj; q++)
            kc[(pos[eidx[si]] + r) * tot + pos[eidx[sj]] + q] = ek[(epos[si] + r) * ne + epos[sj] + q];
        }
      }
    }
    /* BuildDependencyOrder (Mesh/pzelmat.cpp:117-160) */
    for (int i = 0; i < n; i++) dord[i] = 0;
    for (int cur = 0, done = 1; done; cur++) {
      done = 0;
      for (int i = 0; i < n; i++)
        if (dep_ptr[list[i] + 1] > dep_ptr[list[i]] && dord[i] == cur) {
          set_dep_order(list[i], cur, dep_ptr, dep_conn, list, n, dord);
          done++;
        }
    }
    /* fold every constrained connect into the connects it depends on, in
     * dependency order: rows, then columns (one shape per connect) */
    int processed = 0;
    for (int cur = 0; processed < n; cur++)
      for (int i = 0; i < n; i++) {
        if (dord[i] != cur) continue;
        processed++;
        int64_t c = list[i];
        for (int64_t k = dep_ptr[c]; k < dep_ptr[c + 1]; k++) {
          int j = 0;
          while (j < n && list[j] != dep_conn[k]) j++;
          double coef = dep_coef[k];
          int64_t send = pos[i], recv = pos[j];
          if (m->blk_size[c] == 0 || m->blk_size[list[j]] == 0) continue;
          for (int64_t q = 0; q < tot; q++)
            for (int d = 0; d < ns; d++) kc[(recv + d) * tot + q] += coef * kc[(send + d) * tot + q];
          for (int d = 0; d < ns; d++) fc[recv + d] += coef * fc[send + d];
          for (int64_t q = 0; q < tot; q++)
            for (int d = 0; d < ns; d++) kc[q * tot + recv + d] += coef * kc[q * tot + send + d];
        }
      }
    /* AddKel / AddFel over the unconstrained connects of the list */
    int nr = 0;
    int64_t src[4 * ORC_MAXNE];
    for (int i = 0; i < n; i++) {
      int64_t c = list[i];
      if (dep_ptr[c + 1] > dep_ptr[c]) continue;
      for (int d = 0; d < m->blk_size[c]; d++) {
        src[nr] = pos[i] + d;
        kdest[nr] = m->blk_pos[c] + d;
        nr++;
      }
    }
    for (int i = 0; i < nr; i++) {
      fr[i] = fc[src[i]];
      for (int j = 0; j < nr; j++) kr[i * nr + j] = kc[src[i] * tot + src[j]];
    }
    orc_addkel(ia, ja, a, nr, kr, kdest);
    for (int i = 0; i < nr; i++) rhs[kdest[i]] += fr[i];
  }
  free(ek); free(kc); free(kr);
  return 0;
}
