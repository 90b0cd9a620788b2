// pzg_assemble.hip -- gfx950 kernels of the assembly phase.
//
//  * k_upper_ja      : equation-level JA of the reference upper CSR from the
//                      block graph (SetupMatrixData, StrMatrix/
//                      TPZSSpStructMatrix.cpp:134-188: diagonal block first,
//                      then every coupled block in ascending order).
//  * k_elem_maps     : per element and side pair, the index of the coupled
//                      block inside the upper block row (replaces AddKel's row
//                      search, Matrix/pzsysmp.cpp:369-411).
//  * k_vol<...>      : one workgroup per element: geometry at the quadrature
//                      points, gradient Gram blocks P^{ab} = sum_q w|J| dphi_a
//                      dphi_b^T, the material's element matrix from P, and the
//                      scatter of its upper part into the CSR with the
//                      |v| < 1e-12 drop rule.  Elements of one colour share no
//                      connect, so the read-modify-write needs no atomics and
//                      the result is deterministic.
//  * k_bc<...>       : boundary elements (penalty Dirichlet / Neumann loads).
//  * k_expand_full / k_build_bsr : FP64 full CSR (plain PCG path) and the
//                      FP32 scaled block CSR (mixed path) from the upper CSR.
#include <hip/hip_runtime.h>

#include "pzg_internal.h"

namespace pzg {

constexpr int TPB = 256;

// --------------------------------------------------------------------------
__global__ void k_upper_ja(int64_t nb, int bs, const int64_t *__restrict__ ubia,
                           const int32_t *__restrict__ ubja, const int64_t *__restrict__ ia,
                           int64_t *__restrict__ ja) {
  int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  int lane = threadIdx.x & 63;
  if (wave >= nb) return;
  int64_t a = wave, u0 = ubia[a], nup = ubia[a + 1] - u0;
  for (int k = 0; k < bs; k++) {
    int64_t r = a * bs + k, base = ia[r], len = (bs - k) + bs * nup;
    for (int64_t e = lane; e < len; e += 64) {
      int64_t col;
      if (e < bs - k) col = a * bs + k + e;
      else {
        int64_t e2 = e - (bs - k);
        col = (int64_t)ubja[u0 + e2 / bs] * bs + e2 % bs;
      }
      ja[base + e] = col;
    }
  }
}

int launch_expand_upper_ja(Ctx &c) {
  if (c.nb == 0) return PZG_OK;
  int64_t threads = c.nb * 64;
  hipLaunchKernelGGL(k_upper_ja, dim3((threads + TPB - 1) / TPB), dim3(TPB), 0, c.stream, c.nb,
                     c.bs, c.d_ubia.as<int64_t>(), c.d_ubja.as<int32_t>(), c.d_ia.as<int64_t>(),
                     c.d_ja.as<int64_t>());
  PZG_HIP(hipGetLastError());
  return PZG_OK;
}

// --------------------------------------------------------------------------
// map[e][s1][s2]: -2 inactive pair, -1 same block, t >= 0 index of the larger
// block inside the upper block row of the smaller one.
__global__ void k_elem_maps(int64_t nel, int ns, const int64_t *__restrict__ conn,
                            const int64_t *__restrict__ aidx, const int64_t *__restrict__ ubia,
                            const int32_t *__restrict__ ubja, int32_t *__restrict__ map) {
  int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t total = nel * ns * ns;
  if (gid >= total) return;
  int64_t e = gid / (ns * ns);
  int rem = (int)(gid % (ns * ns)), s1 = rem / ns, s2 = rem % ns;
  int64_t a = aidx[conn[e * ns + s1]], b = aidx[conn[e * ns + s2]];
  int32_t v;
  if (a < 0 || b < 0) v = -2;
  else if (a == b) v = -1;
  else {
    int64_t lo = a < b ? a : b, hi = a < b ? b : a;
    int64_t l = ubia[lo], h = ubia[lo + 1] - 1, pos = -1;
    while (l <= h) {
      int64_t m = (l + h) >> 1;
      int64_t vm = ubja[m];
      if (vm == hi) { pos = m; break; }
      if (vm < hi) l = m + 1;
      else h = m - 1;
    }
    v = pos < 0 ? -3 : (int32_t)(pos - ubia[lo]);
  }
  map[gid] = v;
}

int launch_elem_maps(Ctx &c) {
  int ns = c.vol_nsides;
  PZG_TRY(dev_alloc(c.d_vmap, sizeof(int32_t) * c.nvol * ns * ns));
  int64_t tot = c.nvol * ns * ns;
  hipLaunchKernelGGL(k_elem_maps, dim3((tot + TPB - 1) / TPB), dim3(TPB), 0, c.stream, c.nvol, ns,
                     c.d_vconn.as<int64_t>(), c.d_aidx.as<int64_t>(), c.d_ubia.as<int64_t>(),
                     c.d_ubja.as<int32_t>(), c.d_vmap.as<int32_t>());
  PZG_HIP(hipGetLastError());
  if (c.nbc > 0) {
    int nb = c.bc_nsides;
    PZG_TRY(dev_alloc(c.d_bmap, sizeof(int32_t) * c.nbc * nb * nb));
    int64_t t2 = c.nbc * nb * nb;
    hipLaunchKernelGGL(k_elem_maps, dim3((t2 + TPB - 1) / TPB), dim3(TPB), 0, c.stream, c.nbc, nb,
                       c.d_bconn.as<int64_t>(), c.d_aidx.as<int64_t>(), c.d_ubia.as<int64_t>(),
                       c.d_ubja.as<int32_t>(), c.d_bmap.as<int32_t>());
    PZG_HIP(hipGetLastError());
  }
  return PZG_OK;
}

// --------------------------------------------------------------------------
template <class T> __global__ void k_fill(T *p, int64_t n, T v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

template <class T> static void fill(Ctx &c, T *p, int64_t n, T v) {
  if (n <= 0) return;
  int64_t blocks = (n + TPB - 1) / TPB;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(k_fill<T>, dim3(blocks), dim3(TPB), 0, c.stream, p, n, v);
}

struct ElemArgs {
  const double *xyz;
  const int64_t *nodes;
  const int64_t *conn;
  const double *coef;
  const int64_t *aidx;
  const int32_t *map;
  const int64_t *ia;
  double *a;
  double *rhs;
  const ElemTables *tab;
  const double *bumps;
  const int32_t *bcid;
  const int32_t *bctype;
  const double *bcval;
  const double *bccoef;  // BCC coefficients per bc id
  const double *bcvec;   // per boundary element: neighbour-to-element centre vector
  int nbumps;
  double bw;
  double fconst[3];
  int force_kind;
  double scale, E, nu, big;
  double prestress[3];  // TPZElasticity3D::fPreStress
  int material;
  int bs;
  // element-matrix output (constrained elements): shapes are active when the
  // side's connect has shapes (aidx_ek >= 0 unconstrained, -2 constrained);
  // the element matrix / load vector of list entry k go to ek_out / ef_out
  const int64_t *aidx_ek;
  double *ek_out, *ef_out;
};

__device__ inline void forcing(const ElemArgs &g, const double *x, double *f) {
  f[0] = f[1] = f[2] = 0.;
  if (g.force_kind == PZG_FORCE_CONST) {
    f[0] = g.fconst[0];
    f[1] = g.fconst[1];
    f[2] = g.fconst[2];
  } else if (g.force_kind == PZG_FORCE_BUMPS) {
    double w2 = 2. * g.bw * g.bw;
    for (int b = 0; b < g.nbumps; b++) {
      const double *bp = g.bumps + 6 * b;
      double r2 = (x[0] - bp[0]) * (x[0] - bp[0]) + (x[1] - bp[1]) * (x[1] - bp[1]) +
                  (x[2] - bp[2]) * (x[2] - bp[2]);
      double e = exp(-r2 / w2);
      f[0] += bp[3] * e;
      f[1] += bp[4] * e;
      f[2] += bp[5] * e;
    }
  }
}

// Volume elements.  DIM: 2 (quad) / 3 (hex); NC corners; NSD sides; NSH shapes;
// NS states.  One workgroup per element of the current colour.
template <int DIM, int NC, int NSD, int NSH, int NS, bool EK = false>
__global__ __launch_bounds__(TPB) void k_vol(ElemArgs g, const int32_t *__restrict__ list) {
  constexpr int NE = NSH * NS;
  constexpr int NP = (NS == 3) ? 6 : 1;  // Gram blocks P^{ab}, a <= b
  __shared__ double s_coord[NC * 3];
  __shared__ double s_jinv[MAXQ][DIM][DIM];
  __shared__ double s_wdet[MAXQ];
  __shared__ double s_f[MAXQ][3];
  __shared__ double s_G[MAXQ][DIM][NSH];
  __shared__ double s_Gw[MAXQ][DIM][NSH];
  __shared__ double s_P[NP][NSH][NSH];
  __shared__ int64_t s_row[NE];
  __shared__ int64_t s_blk[NSH];
  __shared__ int32_t s_map[NSD * NSD];
  const int tid = threadIdx.x;
  const int64_t e = list[blockIdx.x];
  const ElemTables &T = *g.tab;
  const int nq = T.nq;

  if (tid < NC * 3) s_coord[tid] = g.xyz[3 * g.nodes[e * NC + tid / 3] + tid % 3];
  if (!EK)
    for (int k = tid; k < NSD * NSD; k += TPB) s_map[k] = g.map[e * NSD * NSD + k];
  if (tid < NSH) s_blk[tid] = (EK ? g.aidx_ek : g.aidx)[g.conn[e * NSD + T.shape_side[tid]]];
  __syncthreads();
  // geometry per quadrature point (TPZGeoEl::Jacobian, Mesh/pzgeoel.cpp:1168-1355)
  if (tid < nq) {
    const int q = tid;
    double gx[3][DIM], x[3] = {0, 0, 0};
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < DIM; j++) gx[i][j] = 0.;
    for (int n = 0; n < NC; n++)
      for (int i = 0; i < 3; i++) {
        x[i] += s_coord[3 * n + i] * T.cphi[q][n];
        for (int j = 0; j < DIM; j++) gx[i][j] += s_coord[3 * n + i] * T.cdphi[q][j][n];
      }
    double det;
    if (DIM == 3) {
      double(*j)[DIM] = gx;
      det = -j[0][2] * j[1][1] * j[2][0] + j[0][1] * j[1][2] * j[2][0] + j[0][2] * j[1][0] * j[2][1] -
            j[0][0] * j[1][2] * j[2][1] - j[0][1] * j[1][0] * j[2][2] + j[0][0] * j[1][1] * j[2][2];
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
      for (int a = 0; a < DIM; a++)
        for (int b = 0; b < DIM; b++) s_jinv[q][a][b] = inv[a][b];
    } else {
      // 2D element in 3D space: Gram-Schmidt frame, jac upper triangular
      double n1 = 0, d12 = 0, n2 = 0;
      for (int i = 0; i < 3; i++) { n1 += gx[i][0] * gx[i][0]; d12 += gx[i][0] * gx[i][1]; }
      n1 = sqrt(n1);
      for (int i = 0; i < 3; i++) {
        double v2 = gx[i][1] - d12 * (gx[i][0] / n1) / n1;
        n2 += v2 * v2;
      }
      n2 = sqrt(n2);
      double j01 = d12 / n1;
      det = n1 * n2;
      s_jinv[q][0][0] = n2 / det;
      s_jinv[q][1][1] = n1 / det;
      s_jinv[q][0][1] = -j01 / det;
      s_jinv[q][1][0] = 0.;
    }
    s_wdet[q] = T.w[q] * fabs(det);
    double f[3];
    forcing(g, x, f);
    s_f[q][0] = f[0];
    s_f[q][1] = f[1];
    s_f[q][2] = f[2];
  }
  __syncthreads();
  // dphix = jacinv^T dphi (Mesh/TPZCompElH1.cpp:139-149)
  for (int k = tid; k < nq * NSH; k += TPB) {
    int q = k / NSH, i = k % NSH;
    for (int a = 0; a < DIM; a++) {
      double s = 0.;
      for (int b = 0; b < DIM; b++) s += s_jinv[q][b][a] * T.dphi[q][b][i];
      s_G[q][a][i] = s;
      s_Gw[q][a][i] = s * s_wdet[q];
    }
  }
  // row starts of the element's equations in the upper CSR
  if (!EK)
    for (int k = tid; k < NE; k += TPB) {
      int64_t b = s_blk[k / NS];
      s_row[k] = b >= 0 ? g.ia[b * g.bs + k % NS] : -1;
    }
  __syncthreads();
  // Gram blocks
  for (int k = tid; k < NP * NSH * NSH; k += TPB) {
    int pb = k / (NSH * NSH), ij = k % (NSH * NSH), i = ij / NSH, j = ij % NSH;
    double s = 0.;
    if (NS == 3) {
      const int PA[6] = {0, 0, 0, 1, 1, 2}, PB[6] = {0, 1, 2, 1, 2, 2};
      int a = PA[pb], b = PB[pb];
      for (int q = 0; q < nq; q++) s += s_Gw[q][a][i] * s_G[q][b][j];
    } else {
      for (int q = 0; q < nq; q++)
        for (int a = 0; a < DIM; a++) s += s_Gw[q][a][i] * s_G[q][a][j];
    }
    s_P[pb][i][j] = s;
  }
  __syncthreads();
  // material constants
  double coef = g.coef ? g.coef[e] : (g.material == PZG_MAT_ELAST3D ? g.E : g.scale);
  double C1 = 0, C2 = 0, C3 = 0;
  if (NS == 3) {
    double E = coef, nu = g.nu;
    C1 = E / (2. + 2. * nu);
    C2 = E * nu / (-1. + nu + 2. * nu * nu);
    C3 = E * (nu - 1.) / (-1. + nu + 2. * nu * nu);
  }
  auto entry = [&](int i, int ci, int j, int cj) -> double {
    if (NS == 3) {
      // P^{xy}[i][j] = sum_q w dphi_x(i) dphi_y(j)
      auto P = [&](int x, int y, int ii, int jj) -> double {
        const int idx[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
        return x <= y ? s_P[idx[x][y]][ii][jj] : s_P[idx[y][x]][jj][ii];
      };
      if (ci == cj) {
        double tr = 0.;
        for (int cc = 0; cc < 3; cc++)
          if (cc != ci) tr += P(cc, cc, i, j);
        return tr * C1 + P(ci, ci, i, j) * C3;
      }
      return P(cj, ci, i, j) * C1 - P(ci, cj, i, j) * C2;
    }
    return coef * s_P[0][i][j];
  };
  if (EK) {
    // the whole element matrix and load vector of list entry blockIdx.x
    // (shapes of every side with shapes, constrained or not)
    double *ek = g.ek_out + (int64_t)blockIdx.x * NE * NE;
    for (int k = tid; k < NE * NE; k += TPB) {
      const int I = k / NE, J = k % NE, i = I / NS, ci = I % NS, j = J / NS, cj = J % NS;
      ek[k] = (s_blk[i] == -1 || s_blk[j] == -1) ? 0. : entry(i, ci, j, cj);
    }
    for (int I = tid; I < NE; I += TPB) {
      const int i = I / NS, ci = I % NS;
      double sf = 0.;
      if (s_blk[i] != -1) {
        for (int q = 0; q < nq; q++) sf += s_wdet[q] * T.phi[q][i] * s_f[q][ci];
        if (g.material == PZG_MAT_POISSON) sf *= g.scale;
        if (NS == 3 && g.prestress[ci] != 0.) {  // - w prestress_k dphix_k (fPreStress)
          double sg = 0.;
          for (int q = 0; q < nq; q++) sg += s_Gw[q][ci][i];
          sf -= g.prestress[ci] * sg;
        }
      }
      g.ef_out[(int64_t)blockIdx.x * NE + I] = sf;
    }
    return;
  }
  // element matrix entries, upper part by global equation, scatter; SU
  // entries per thread at a time, their read-modify-writes issued together
  constexpr int SU = 8;
  for (int k0 = tid; k0 < NE * NE; k0 += TPB * SU) {
   int64_t slot[SU];
   double val[SU];
   bool on[SU];
#pragma unroll
   for (int u = 0; u < SU; u++) {
    on[u] = false;
    slot[u] = 0;
    val[u] = 0.;
    const int k = k0 + u * TPB;
    if (k >= NE * NE) continue;
    int I = k / NE, J = k % NE;
    int i = I / NS, ci = I % NS, j = J / NS, cj = J % NS;
    int64_t bi = s_blk[i], bj = s_blk[j];
    if (bi < 0 || bj < 0) continue;
    if (bi > bj || (bi == bj && (ci > cj || (ci == cj && i > j)))) continue;
    const double v = entry(i, ci, j, cj);
    if (fabs(v) < 1.e-12) continue;  // IsZero drop rule of AddKel
    if (bi == bj) slot[u] = s_row[I] + (cj - ci);
    else {
      int t = s_map[T.shape_side[i] * NSD + T.shape_side[j]];
      slot[u] = s_row[I] + g.bs + (int64_t)t * g.bs - ci + cj;
    }
    val[u] = v;
    on[u] = true;
   }
   double old[SU];
#pragma unroll
   for (int u = 0; u < SU; u++) old[u] = g.a[slot[u]];  // slot 0 for inactive entries
#pragma unroll
   for (int u = 0; u < SU; u++)
     if (on[u]) g.a[slot[u]] = old[u] + val[u];
  }
  // load vector
  for (int I = tid; I < NE; I += TPB) {
    int i = I / NS, ci = I % NS;
    if (s_blk[i] < 0) continue;
    double s = 0.;
    for (int q = 0; q < nq; q++) s += s_wdet[q] * T.phi[q][i] * s_f[q][ci];
    if (g.material == PZG_MAT_POISSON) s *= g.scale;
    if (NS == 3 && g.prestress[ci] != 0.) {  // - w prestress_k dphix_k (fPreStress)
      double sg = 0.;
      for (int q = 0; q < nq; q++) sg += s_Gw[q][ci][i];
      s -= g.prestress[ci] * sg;
    }
    g.rhs[s_blk[i] * g.bs + ci] += s;
  }
}

// Affine volume elements (parallelepipeds / parallelograms: constant
// Jacobian), several per workgroup.  With R_mn(i,j) = sum_q w_q dphi_m(i)
// dphi_n(j) the reference Gram matrices of the rule (the same for every
// element, staged in LDS) the gradient Gram blocks are
//     P^{ab}(i,j) = |det J| sum_{m,n} Jinv[m][a] Jinv[n][b] R_mn(i,j),
// exactly the quadrature sum of k_vol for a constant Jacobian (rounding
// aside), so an element costs 81 FMA per shape pair instead of a pass over
// its quadrature points.  One thread per (element, shape pair) of the upper
// part builds the pair's nstate x nstate block and adds it with AddKel's drop
// rule; the load vector keeps the quadrature (the body load varies).
constexpr int AFF_EPB = 4;  // elements per workgroup
template <int DIM, int NC, int NSD, int NSH, int NS, bool MAT = true>
__global__ __launch_bounds__(TPB) void k_vol_aff(ElemArgs g, const double *__restrict__ R,
                                                 const int32_t *__restrict__ list, int64_t n) {
  constexpr int NR = DIM * (DIM + 1) / 2;  // R_mn, m <= n
  constexpr int NE = NSH * NS;
  __shared__ double s_R[NR][NSH][NSH];
  __shared__ double s_jinv[AFF_EPB][DIM][DIM];
  __shared__ double s_det[AFF_EPB];
  __shared__ double s_coef[AFF_EPB];
  __shared__ double s_coord[AFF_EPB][NC * 3];
  __shared__ double s_f[AFF_EPB][MAXQ][3];
  __shared__ int64_t s_blk[AFF_EPB][NSH];
  __shared__ int64_t s_row[AFF_EPB][NE];
  __shared__ int32_t s_map[AFF_EPB][NSD * NSD];
  const int tid = threadIdx.x;
  const ElemTables &T = *g.tab;
  const int nq = T.nq;
  const int64_t e0 = (int64_t)blockIdx.x * AFF_EPB;
  const int nel = (int)(n - e0 < AFF_EPB ? n - e0 : AFF_EPB);
  if (MAT)
    for (int t = tid; t < NR * NSH * NSH; t += TPB) (&s_R[0][0][0])[t] = R[t];
  for (int t = tid; t < nel * NC * 3; t += TPB) {
    const int el = t / (NC * 3), k = t % (NC * 3);
    const int64_t e = list[e0 + el];
    s_coord[el][k] = g.xyz[3 * g.nodes[e * NC + k / 3] + k % 3];
  }
  if (MAT)
    for (int t = tid; t < nel * NSD * NSD; t += TPB) {
      const int el = t / (NSD * NSD), k = t % (NSD * NSD);
      s_map[el][k] = g.map[(int64_t)list[e0 + el] * NSD * NSD + k];
    }
  for (int t = tid; t < nel * NSH; t += TPB) {
    const int el = t / NSH, k = t % NSH;
    s_blk[el][k] = g.aidx[g.conn[(int64_t)list[e0 + el] * NSD + T.shape_side[k]]];
  }
  __syncthreads();
  // the constant Jacobian (TPZGeoEl::Jacobian at the first point) and the
  // row starts of the element's equations in the upper CSR
  if (tid < nel) {
    const int el = tid;
    const int64_t e = list[e0 + el];
    double gx[3][DIM];
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < DIM; j++) {
        double v = 0.;
        for (int c = 0; c < NC; c++) v += s_coord[el][3 * c + i] * T.cdphi[0][j][c];
        gx[i][j] = v;
      }
    double det;
    if (DIM == 3) {
      const double(*j)[DIM] = gx;
      det = -j[0][2] * j[1][1] * j[2][0] + j[0][1] * j[1][2] * j[2][0] + j[0][2] * j[1][0] * j[2][1] -
            j[0][0] * j[1][2] * j[2][1] - j[0][1] * j[1][0] * j[2][2] + j[0][0] * j[1][1] * j[2][2];
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
      for (int a = 0; a < DIM; a++)
        for (int b = 0; b < DIM; b++) s_jinv[el][a][b] = inv[a][b];
    } else {
      double n1 = 0, d12 = 0, n2 = 0;
      for (int i = 0; i < 3; i++) { n1 += gx[i][0] * gx[i][0]; d12 += gx[i][0] * gx[i][1]; }
      n1 = sqrt(n1);
      for (int i = 0; i < 3; i++) {
        const double v2 = gx[i][1] - d12 * (gx[i][0] / n1) / n1;
        n2 += v2 * v2;
      }
      n2 = sqrt(n2);
      const double j01 = d12 / n1;
      det = n1 * n2;
      s_jinv[el][0][0] = n2 / det;
      s_jinv[el][1][1] = n1 / det;
      s_jinv[el][0][1] = -j01 / det;
      s_jinv[el][1][0] = 0.;
    }
    s_det[el] = fabs(det);
    s_coef[el] = g.coef ? g.coef[e] : (g.material == PZG_MAT_ELAST3D ? g.E : g.scale);
  }
  for (int t = tid; t < nel * NE; t += TPB) {
    const int el = t / NE, k = t % NE;
    const int64_t b = s_blk[el][k / NS];
    s_row[el][k] = b >= 0 ? g.ia[b * g.bs + k % NS] : -1;
  }
  // the load at the quadrature points
  for (int t = tid; t < nel * nq; t += TPB) {
    const int el = t / nq, q = t % nq;
    double x[3] = {0., 0., 0.}, f[3];
    for (int c = 0; c < NC; c++)
      for (int i = 0; i < 3; i++) x[i] += s_coord[el][3 * c + i] * T.cphi[q][c];
    forcing(g, x, f);
    for (int i = 0; i < 3; i++) s_f[el][q][i] = f[i];
  }
  __syncthreads();
  for (int t = tid; MAT && t < nel * NSH * NSH; t += TPB) {
    const int el = t / (NSH * NSH), ij = t % (NSH * NSH), i = ij / NSH, j = ij % NSH;
    const int64_t bi = s_blk[el][i], bj = s_blk[el][j];
    if (bi < 0 || bj < 0 || bi > bj || (bi == bj && i > j)) continue;
    // P[a][b] = |det| Jinv^T R(i,j) Jinv
    double Q[DIM][DIM], P[DIM][DIM];
    for (int m = 0; m < DIM; m++)
      for (int b = 0; b < DIM; b++) {
        double v = 0.;
        for (int nn = 0; nn < DIM; nn++) {
          const int lo = m < nn ? m : nn, hi = m < nn ? nn : m;
          const int r = lo * DIM - lo * (lo - 1) / 2 + (hi - lo);
          const double rv = m <= nn ? s_R[r][i][j] : s_R[r][j][i];
          v += rv * s_jinv[el][nn][b];
        }
        Q[m][b] = v;
      }
    for (int a = 0; a < DIM; a++)
      for (int b = 0; b < DIM; b++) {
        double v = 0.;
        for (int m = 0; m < DIM; m++) v += s_jinv[el][m][a] * Q[m][b];
        P[a][b] = v * s_det[el];
      }
    const double coef = s_coef[el];
    if (NS == 3) {
      const double E = coef, nu = g.nu;
      const double C1 = E / (2. + 2. * nu);
      const double C2 = E * nu / (-1. + nu + 2. * nu * nu);
      const double C3 = E * (nu - 1.) / (-1. + nu + 2. * nu * nu);
      const int tb = bi == bj ? 0 : s_map[el][T.shape_side[i] * NSD + T.shape_side[j]];
      // the block's nine entries and slots first, then all read-modify-writes
      // issued together (entries below the diagonal and dropped ones read their
      // row's first slot and write nothing)
      int64_t slot[9];
      double val[9];
      bool on[9];
#pragma unroll
      for (int ci = 0; ci < 3; ci++)
#pragma unroll
        for (int cj = 0; cj < 3; cj++) {
          const int u = 3 * ci + cj;
          double v;
          if (ci == cj) {
            double tr = 0.;
#pragma unroll
            for (int cc = 0; cc < 3; cc++)
              if (cc != ci) tr += P[cc][cc];
            v = tr * C1 + P[ci][ci] * C3;
          } else {
            v = P[cj][ci] * C1 - P[ci][cj] * C2;
          }
          const int64_t row = s_row[el][i * NS + ci];
          on[u] = !(bi == bj && cj < ci) && !(fabs(v) < 1.e-12);  // upper part; IsZero drop rule
          slot[u] = on[u] ? (bi == bj ? row + (cj - ci) : row + g.bs + (int64_t)tb * g.bs - ci + cj) : row;
          val[u] = v;
        }
      double old[9];
#pragma unroll
      for (int u = 0; u < 9; u++) old[u] = g.a[slot[u]];
#pragma unroll
      for (int u = 0; u < 9; u++)
        if (on[u]) g.a[slot[u]] = old[u] + val[u];
    } else {
      double tr = 0.;
      for (int a = 0; a < DIM; a++) tr += P[a][a];
      const double v = coef * tr;
      if (fabs(v) >= 1.e-12) {
        int64_t slot;
        if (bi == bj) slot = s_row[el][i];
        else slot = s_row[el][i] + g.bs + (int64_t)s_map[el][T.shape_side[i] * NSD + T.shape_side[j]] * g.bs;
        g.a[slot] += v;
      }
    }
  }
  // load vector
  for (int t = tid; t < nel * NE; t += TPB) {
    const int el = t / NE, I = t % NE, i = I / NS, ci = I % NS;
    if (s_blk[el][i] < 0) continue;
    double sf = 0.;
    for (int q = 0; q < nq; q++) sf += T.w[q] * s_det[el] * T.phi[q][i] * s_f[el][q][ci];
    if (g.material == PZG_MAT_POISSON) sf *= g.scale;
    if (NS == 3 && g.prestress[ci] != 0.) {  // - w prestress_k dphix_k (fPreStress)
      double sg = 0.;
      for (int q = 0; q < nq; q++) {
        double d = 0.;
        for (int m = 0; m < DIM; m++) d += s_jinv[el][m][ci] * T.dphi[q][m][i];
        sg += T.w[q] * s_det[el] * d;
      }
      sf -= g.prestress[ci] * sg;
    }
    g.rhs[s_blk[el][i] * g.bs + ci] += sf;
  }
}

// Volume elements that are all one affine hexahedron (one Jacobian, one
// Poisson ratio): K_e = coef_e K_ref with K_ref the element matrix of
// coefficient 1 (FP64, built on the host from the same reference Gram
// matrices as k_vol_aff).  Instead of scattering every element's matrix, each
// stored block (A, B), B >= A, is formed once from the elements that share A
// and B (the intersection of the two connects' element lists, ascending
// element order, as AddKel adds them) and written: no read-modify-write and
// no colours.  Each element term keeps AddKel's |v| < 1e-12 drop rule.
// One wavefront per block row (grid-stride), one lane per block; the
// incidence entries pack (element << 5 | local shape).
constexpr int BLK_TPB = 1024;
template <int NS, int NE>
__global__ __launch_bounds__(BLK_TPB) void k_vol_blk(int64_t nb, const int64_t *__restrict__ ia,
                                                     const int64_t *__restrict__ ubia,
                                                     const int32_t *__restrict__ ubja,
                                                     const int64_t *__restrict__ iptr,
                                                     const int32_t *__restrict__ inc,
                                                     const double *__restrict__ coefv, double coef0,
                                                     const double *__restrict__ kref, double *__restrict__ a) {
  constexpr int NW = BLK_TPB / 64, MAXI = 8;
  __shared__ double sK[NE * NE];
  __shared__ int32_t sI[NW][MAXI];
  for (int t = threadIdx.x; t < NE * NE; t += BLK_TPB) sK[t] = kref[t];
  __syncthreads();
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int64_t A = (int64_t)blockIdx.x * NW + w; A < nb; A += (int64_t)gridDim.x * NW) {
    const int64_t a0 = iptr[A];
    const int na = (int)(iptr[A + 1] - a0);  // <= MAXI (checked on the host)
    if (lane < na) sI[w][lane] = inc[a0 + lane];
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    int64_t ra[NS];
#pragma unroll
    for (int k = 0; k < NS; k++) ra[k] = ia[A * NS + k];
    const int64_t u0 = ubia[A], nup = ubia[A + 1] - u0;
    for (int64_t j = lane; j <= nup; j += 64) {
      const int64_t B = j == 0 ? A : (int64_t)ubja[u0 + j - 1];
      double acc[NS][NS];
#pragma unroll
      for (int k = 0; k < NS; k++)
#pragma unroll
        for (int l = 0; l < NS; l++) acc[k][l] = 0.;
      const int64_t b0 = iptr[B], b1 = iptr[B + 1];
      int pa = 0;
      for (int64_t q = b0; q < b1; q++) {
        const int32_t vb = inc[q];
        const int eb = vb >> 5;
        while (pa < na && (sI[w][pa] >> 5) < eb) pa++;
        if (pa == na) break;
        const int32_t va = sI[w][pa];
        if ((va >> 5) != eb) continue;
        const double cf = coefv ? coefv[eb] : coef0;
        const double *kr = sK + (va & 31) * NS * NE + (vb & 31) * NS;
#pragma unroll
        for (int k = 0; k < NS; k++)
#pragma unroll
          for (int l = 0; l < NS; l++) {
            const double v = cf * kr[k * NE + l];
            if (!(fabs(v) < 1.e-12)) acc[k][l] += v;  // IsZero drop rule of AddKel
          }
      }
      if (j == 0) {
#pragma unroll
        for (int k = 0; k < NS; k++)
#pragma unroll
          for (int l = k; l < NS; l++) a[ra[k] + (l - k)] = acc[k][l];
      } else {
#pragma unroll
        for (int k = 0; k < NS; k++)
#pragma unroll
          for (int l = 0; l < NS; l++) a[ra[k] + (NS - k) + (j - 1) * NS + l] = acc[k][l];
      }
    }
    __builtin_amdgcn_wave_barrier();  // sI[w] is rewritten by the next row
  }
}

// Boundary elements (dimension DIMB = DIM-1 embedded in 3D).  The material's
// ContributeBC (Material/Poisson/TPZMatPoisson.cpp:45-121, Material/DarcyFlow/
// TPZDarcyFlow.cpp:49-91, Material/Elasticity/TPZElasticity3D.cpp:616-773)
// enters as the coefficients of the bc id (BCC layout): ek(I,J) =
// ck[ci][cj] M_ij + cd D_ij, ef(I) = cf[ci] m_i (+ the stress-field load).
template <int DIMB, int NC, int NSD, int NSH, int NS, bool EK = false>
__global__ __launch_bounds__(TPB) void k_bc(ElemArgs g, const int32_t *__restrict__ list) {
  constexpr int NE = NSH * NS;
  __shared__ double s_coord[NC * 3];
  __shared__ double s_wdet[MAXQ];
  __shared__ double s_ia0[MAXQ];   // 1 / |dx/dxi_0|: dphix row 0 = dphi row 0 * s_ia0
  __shared__ double s_g[MAXQ][3];  // stress-field load -val1 . n at the points
  __shared__ double s_M[NSH][NSH];
  __shared__ double s_D[NSH][NSH];
  __shared__ double s_cf[BCC];
  __shared__ int64_t s_row[NE];
  __shared__ int64_t s_blk[NSH];
  __shared__ int32_t s_map[NSD * NSD];
  const int tid = threadIdx.x;
  const int64_t e = list[blockIdx.x];
  const ElemTables &T = *g.tab;
  const int nq = T.nq;
  const int id = g.bcid[e];
  if (tid < NC * 3) s_coord[tid] = g.xyz[3 * g.nodes[e * NC + tid / 3] + tid % 3];
  if (tid < BCC) s_cf[tid] = g.bccoef[(int64_t)id * BCC + tid];
  if (!EK)
    for (int k = tid; k < NSD * NSD; k += TPB) s_map[k] = g.map[e * NSD * NSD + k];
  if (tid < NSH) s_blk[tid] = (EK ? g.aidx_ek : g.aidx)[g.conn[e * NSD + T.shape_side[tid]]];
  __syncthreads();
  const bool stress = s_cf[BCC_STRESS] != 0.;
  const double cd = s_cf[BCC_CD];
  if (tid < nq) {
    const int q = tid;
    double gx[3][2] = {{0, 0}, {0, 0}, {0, 0}};
    for (int n = 0; n < NC; n++)
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < DIMB; j++) gx[i][j] += s_coord[3 * n + i] * T.cdphi[q][j][n];
    // TPZGeoEl::Jacobian of a line / face in 3D space (Gram-Schmidt axes)
    double det, n1 = 0, a0[3], a1[3] = {0, 0, 0};
    for (int i = 0; i < 3; i++) n1 += gx[i][0] * gx[i][0];
    n1 = sqrt(n1);
    for (int i = 0; i < 3; i++) a0[i] = gx[i][0] / n1;
    if (DIMB == 1) {
      det = n1;
    } else {
      double d12 = 0, n2 = 0;
      for (int i = 0; i < 3; i++) d12 += gx[i][0] * gx[i][1];
      for (int i = 0; i < 3; i++) {
        a1[i] = gx[i][1] - d12 * a0[i] / n1;
        n2 += a1[i] * a1[i];
      }
      n2 = sqrt(n2);
      for (int i = 0; i < 3; i++) a1[i] /= n2;
      det = n1 * n2;
    }
    s_wdet[q] = T.w[q] * fabs(det);
    s_ia0[q] = 1.0 / n1;
    if (stress) {
      // TPZInterpolationSpace::ComputeNormal: unit normal from the axes,
      // turned to point away from the volume neighbour
      const double *v = g.bcvec + 3 * e;
      double t[3];
      if (DIMB == 1) {
        const double b[3] = {a0[1] * v[2] - a0[2] * v[1], -a0[0] * v[2] + a0[2] * v[0],
                             a0[0] * v[1] - a0[1] * v[0]};
        t[0] = b[1] * a0[2] - b[2] * a0[1];
        t[1] = -b[0] * a0[2] + b[2] * a0[0];
        t[2] = b[0] * a0[1] - b[1] * a0[0];
      } else {
        t[0] = a0[1] * a1[2] - a0[2] * a1[1];
        t[1] = -a0[0] * a1[2] + a0[2] * a1[0];
        t[2] = a0[0] * a1[1] - a0[1] * a1[0];
      }
      const double sz = sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
      double nrm[3], dot = 0.;
      for (int i = 0; i < 3; i++) { nrm[i] = t[i] / sz; dot += nrm[i] * v[i]; }
      if (dot < 0.)
        for (int i = 0; i < 3; i++) nrm[i] = -nrm[i];
      for (int k = 0; k < 3; k++)
        s_g[q][k] = -(s_cf[BCC_VAL1 + 3 * k] * nrm[0] + s_cf[BCC_VAL1 + 3 * k + 1] * nrm[1] +
                      s_cf[BCC_VAL1 + 3 * k + 2] * nrm[2]);
    }
  }
  if (!EK)
    for (int k = tid; k < NE; k += TPB) {
      int64_t b = s_blk[k / NS];
      s_row[k] = b >= 0 ? g.ia[b * g.bs + k % NS] : -1;
    }
  __syncthreads();
  for (int k = tid; k < NSH * NSH; k += TPB) {
    int i = k / NSH, j = k % NSH;
    double s = 0., sd = 0.;
    for (int q = 0; q < nq; q++) s += T.phi[q][i] * T.phi[q][j] * s_wdet[q];
    if (cd != 0.)
      for (int q = 0; q < nq; q++)
        sd += T.dphi[q][0][i] * s_ia0[q] * T.dphi[q][0][j] * s_ia0[q] * s_wdet[q];
    s_M[i][j] = s;
    s_D[i][j] = sd;
  }
  __syncthreads();
  auto ek_entry = [&](int i, int ci, int j, int cj) {
    return s_cf[BCC_CK + NS * ci + cj] * s_M[i][j] + (ci == cj ? cd * s_D[i][j] : 0.);
  };
  auto ef_entry = [&](int i, int ci) {
    double sq = 0., st = 0.;
    for (int q = 0; q < nq; q++) sq += T.phi[q][i] * s_wdet[q];
    if (stress)
      for (int q = 0; q < nq; q++) st += T.phi[q][i] * s_wdet[q] * s_g[q][ci];
    return s_cf[BCC_CF + ci] * sq + st;
  };
  if (EK) {
    // the whole element matrix and load vector of list entry blockIdx.x
    double *ek = g.ek_out + (int64_t)blockIdx.x * NE * NE;
    for (int k = tid; k < NE * NE; k += TPB) {
      const int I = k / NE, J = k % NE, i = I / NS, ci = I % NS, j = J / NS, cj = J % NS;
      ek[k] = (s_blk[i] != -1 && s_blk[j] != -1) ? ek_entry(i, ci, j, cj) : 0.;
    }
    for (int I = tid; I < NE; I += TPB) {
      const int i = I / NS, ci = I % NS;
      g.ef_out[(int64_t)blockIdx.x * NE + I] = s_blk[i] != -1 ? ef_entry(i, ci) : 0.;
    }
    return;
  }
  for (int k = tid; k < NE * NE; k += TPB) {
    int I = k / NE, J = k % NE;
    int i = I / NS, ci = I % NS, j = J / NS, cj = J % NS;
    int64_t bi = s_blk[i], bj = s_blk[j];
    if (bi < 0 || bj < 0) continue;
    // upper part of the equation numbering: block order, then component
    if (bi > bj || (bi == bj && (i != j || cj < ci))) continue;
    double v = ek_entry(i, ci, j, cj);
    if (fabs(v) < 1.e-12) continue;  // AddKel's drop rule
    int64_t slot;
    if (bi == bj) slot = s_row[I] + (cj - ci);
    else {
      int t = s_map[T.shape_side[i] * NSD + T.shape_side[j]];
      slot = s_row[I] + (g.bs - ci) + (int64_t)t * g.bs + cj;
    }
    g.a[slot] += v;
  }
  for (int I = tid; I < NE; I += TPB) {
    int i = I / NS, ci = I % NS;
    if (s_blk[i] < 0) continue;
    g.rhs[s_blk[i] * g.bs + ci] += ef_entry(i, ci);
  }
}

// ---------------------------------------------------------------------------
// High-order triangles / tetrahedra (scalar materials).  The element's shapes
// are the virtual sides of its connect list; shape i reads table column
// col_off[i] + (transform variant of its real side, from the corner ids).
// The Gram matrix sum_q w|J| grad phi_i . grad phi_j is accumulated over
// chunks of QC quadrature points, each thread keeping its own (i, j) entries.
// (tri_tid / hp_variant: pzg_internal.h)

// Shapes of one rule bucket: the sub-list sub[0..NSUB) of the element's NSV
// virtual sides (the shapes an element of the bucket's highest order can
// carry).  The Gram matrix is symmetric: the TSxTS tiles (ti, tj), ti <= tj,
// of sub-list pairs are dealt to the threads, which scatter every pair once.
// QC quadrature points are staged per pass.  (QC, TS) is picked per bucket by
// timing (hp_tune).
template <int NSUB, int QC, int TS>
__global__ __launch_bounds__(TPB) void k_vol_hp(ElemArgs g, int nsv, const HpTables *__restrict__ tab,
                                                const int32_t *__restrict__ sub,
                                                const int32_t *__restrict__ list) {
  constexpr int NT = (NSUB + TS - 1) / TS, NSP = TS * NT;
  constexpr int NTILE = NT * (NT + 1) / 2;
  constexpr int MT = (NTILE + TPB - 1) / TPB;  // tiles per thread
  __shared__ double s_coord[12];
  __shared__ double s_jinv[3][3];
  __shared__ double s_det;
  __shared__ double s_G[QC][3][NSP];
  __shared__ double s_Gw[QC][3][NSP];
  __shared__ double s_pw[QC][NSP];
  __shared__ double s_f[QC];
  __shared__ int s_col[NSUB];
  __shared__ int64_t s_blk[NSUB], s_row[NSUB];
  __shared__ int32_t s_map[NSUB * NSUB];
  const int tid = threadIdx.x;
  const int64_t e = list[blockIdx.x];
  const HpTables &T = *tab;
  if (tid < 12) s_coord[tid] = g.xyz[3 * g.nodes[e * 4 + tid / 3] + tid % 3];
  for (int k = tid; k < NSUB * NSUB; k += TPB) {
    const int i = k / NSUB, j = k % NSUB;
    s_map[k] = g.map[((int64_t)e * nsv + sub[i]) * nsv + sub[j]];
  }
  if (tid < NSUB) {
    int64_t id[4];
    for (int k = 0; k < 4; k++) id[k] = g.nodes[e * 4 + k];
    const int gi = sub[tid];
    s_col[tid] = T.col_off[gi] + hp_variant(true, T.shape_side[gi], id);
    const int64_t b = g.aidx[g.conn[e * nsv + gi]];
    s_blk[tid] = b;
    s_row[tid] = b >= 0 ? g.ia[b] : -1;
  }
  // padding columns of the tiles stay zero
  if constexpr (NSP > NSUB) {
    for (int k = tid; k < QC * 3 * (NSP - NSUB); k += TPB) {
      const int qa = k / (NSP - NSUB), i = NSUB + k % (NSP - NSUB);
      s_G[qa / 3][qa % 3][i] = 0.;
      s_Gw[qa / 3][qa % 3][i] = 0.;
    }
  }
  if (tid == 0) {
    // GradX of the straight tetrahedron: column k = x_{k+1} - x_0
    double j[3][3];
    for (int i = 0; i < 3; i++)
      for (int k = 0; k < 3; k++) j[i][k] = g.xyz[3 * g.nodes[e * 4 + k + 1] + i] - g.xyz[3 * g.nodes[e * 4] + i];
    const double det = -j[0][2] * j[1][1] * j[2][0] + j[0][1] * j[1][2] * j[2][0] +
                       j[0][2] * j[1][0] * j[2][1] - j[0][0] * j[1][2] * j[2][1] -
                       j[0][1] * j[1][0] * j[2][2] + j[0][0] * j[1][1] * j[2][2];
    s_jinv[0][0] = (-j[1][2] * j[2][1] + j[1][1] * j[2][2]) / det;
    s_jinv[0][1] = (j[0][2] * j[2][1] - j[0][1] * j[2][2]) / det;
    s_jinv[0][2] = (-j[0][2] * j[1][1] + j[0][1] * j[1][2]) / det;
    s_jinv[1][0] = (j[1][2] * j[2][0] - j[1][0] * j[2][2]) / det;
    s_jinv[1][1] = (-j[0][2] * j[2][0] + j[0][0] * j[2][2]) / det;
    s_jinv[1][2] = (j[0][2] * j[1][0] - j[0][0] * j[1][2]) / det;
    s_jinv[2][0] = (-j[1][1] * j[2][0] + j[1][0] * j[2][1]) / det;
    s_jinv[2][1] = (j[0][1] * j[2][0] - j[0][0] * j[2][1]) / det;
    s_jinv[2][2] = (-j[0][1] * j[1][0] + j[0][0] * j[1][1]) / det;
    s_det = fabs(det);
  }
  // tiles of this thread: t = tid + m*TPB
  int ti[MT], tj[MT];
#pragma unroll
  for (int m = 0; m < MT; m++) {
    int t = tid + m * TPB, a = 0;
    if (t < NTILE) {
      while (t >= NT - a) { t -= NT - a; a++; }
      ti[m] = a;
      tj[m] = a + t;
    } else {
      ti[m] = tj[m] = -1;
    }
  }
  __syncthreads();
  double acc[MT][TS][TS];
#pragma unroll
  for (int m = 0; m < MT; m++)
#pragma unroll
    for (int r = 0; r < TS; r++)
#pragma unroll
      for (int q = 0; q < TS; q++) acc[m][r][q] = 0.;
  double accf = 0.;
  const int nq = T.nq;
  for (int q0 = 0; q0 < nq; q0 += QC) {
    const int nqc = nq - q0 < QC ? nq - q0 : QC;
    if (tid < nqc) {
      const double *xi = T.pt[q0 + tid];
      const double l[4] = {1. - xi[0] - xi[1] - xi[2], xi[0], xi[1], xi[2]};
      double x[3] = {0., 0., 0.}, f[3];
      for (int n = 0; n < 4; n++)
        for (int i = 0; i < 3; i++) x[i] += s_coord[3 * n + i] * l[n];
      forcing(g, x, f);
      s_f[tid] = f[0];
    }
    for (int k = tid; k < nqc * NSUB; k += TPB) {
      const int qq = k / NSUB, i = k % NSUB, q = q0 + qq, col = s_col[i];
      const double wd = T.w[q] * s_det;
      for (int a = 0; a < 3; a++) {
        double sum = 0.;
        for (int b = 0; b < 3; b++) sum += s_jinv[b][a] * T.dphi[q][b][col];
        s_G[qq][a][i] = sum;
        s_Gw[qq][a][i] = sum * wd;
      }
      s_pw[qq][i] = wd * T.phi[q][col];
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < MT; m++) {
      if (ti[m] < 0) continue;
      for (int qq = 0; qq < nqc; qq++)
#pragma unroll
        for (int a = 0; a < 3; a++) {
          double u[TS], v[TS];
#pragma unroll
          for (int r = 0; r < TS; r++) {
            u[r] = s_Gw[qq][a][TS * ti[m] + r];
            v[r] = s_G[qq][a][TS * tj[m] + r];
          }
#pragma unroll
          for (int r = 0; r < TS; r++)
#pragma unroll
            for (int q = 0; q < TS; q++) acc[m][r][q] += u[r] * v[q];
        }
    }
    if (tid < NSUB)
      for (int qq = 0; qq < nqc; qq++) accf += s_pw[qq][tid] * s_f[qq];
    __syncthreads();
  }
  const double coef = g.coef ? g.coef[e] : g.scale;
#pragma unroll
  for (int m = 0; m < MT; m++) {
    if (ti[m] < 0) continue;
#pragma unroll
    for (int r = 0; r < TS; r++)
#pragma unroll
      for (int q = 0; q < TS; q++) {
        const int i = TS * ti[m] + r, j = TS * tj[m] + q;
        if (i >= NSUB || j >= NSUB || (ti[m] == tj[m] && q < r)) continue;
        const int64_t bi = s_blk[i], bj = s_blk[j];
        if (bi < 0 || bj < 0) continue;
        const double v = coef * acc[m][r][q];
        if (fabs(v) < 1.e-12) continue;  // IsZero drop rule of AddKel
        int64_t slot;
        if (bi == bj) slot = s_row[i];
        else if (bi < bj) slot = s_row[i] + 1 + s_map[i * NSUB + j];
        else slot = s_row[j] + 1 + s_map[j * NSUB + i];
        g.a[slot] += v;
      }
  }
  if (tid < NSUB && s_blk[tid] >= 0)
    g.rhs[s_blk[tid]] += g.material == PZG_MAT_POISSON ? accf * g.scale : accf;
}

// TPZElasticity3D on high-order tetrahedra: the same tiles of shape pairs,
// each keeping the nine directional sums Q_xy(i,j) = sum_q w|J| dphi_x(i)
// dphi_y(j), from which the pair's 3x3 block is formed as in k_vol (C1, C2,
// C3 of TPZElasticity3D.h:183-188).  A pair (i, j), i <= j in shape order, is
// scattered to the upper triangle of the equation numbering: the block of
// row i when block(i) < block(j), the transposed block of row j otherwise
// (K is symmetric); the diagonal block keeps its upper part.
template <int NSUB, int QC, int TS>
__global__ __launch_bounds__(TPB) void k_vol_hp3(ElemArgs g, int nsv, const HpTables *__restrict__ tab,
                                                 const int32_t *__restrict__ sub,
                                                 const int32_t *__restrict__ list) {
  constexpr int NT = (NSUB + TS - 1) / TS, NSP = TS * NT;
  constexpr int NTILE = NT * (NT + 1) / 2;
  constexpr int MT = (NTILE + TPB - 1) / TPB;
  __shared__ double s_coord[12];
  __shared__ double s_jinv[3][3];
  __shared__ double s_det;
  __shared__ double s_G[QC][3][NSP];
  __shared__ double s_Gw[QC][3][NSP];
  __shared__ double s_pw[QC][NSP];
  __shared__ double s_f[QC][3];
  __shared__ int s_col[NSUB];
  __shared__ int64_t s_blk[NSUB], s_row[NSUB][3];
  __shared__ int32_t s_map[NSUB * NSUB];
  const int tid = threadIdx.x;
  const int64_t e = list[blockIdx.x];
  const HpTables &T = *tab;
  if (tid < 12) s_coord[tid] = g.xyz[3 * g.nodes[e * 4 + tid / 3] + tid % 3];
  for (int k = tid; k < NSUB * NSUB; k += TPB) {
    const int i = k / NSUB, j = k % NSUB;
    s_map[k] = g.map[((int64_t)e * nsv + sub[i]) * nsv + sub[j]];
  }
  if (tid < NSUB) {
    int64_t id[4];
    for (int k = 0; k < 4; k++) id[k] = g.nodes[e * 4 + k];
    const int gi = sub[tid];
    s_col[tid] = T.col_off[gi] + hp_variant(true, T.shape_side[gi], id);
    const int64_t b = g.aidx[g.conn[e * nsv + gi]];
    s_blk[tid] = b;
    for (int c = 0; c < 3; c++) s_row[tid][c] = b >= 0 ? g.ia[b * 3 + c] : -1;
  }
  if constexpr (NSP > NSUB) {
    for (int k = tid; k < QC * 3 * (NSP - NSUB); k += TPB) {
      const int qa = k / (NSP - NSUB), i = NSUB + k % (NSP - NSUB);
      s_G[qa / 3][qa % 3][i] = 0.;
      s_Gw[qa / 3][qa % 3][i] = 0.;
    }
  }
  if (tid == 0) {
    double j[3][3];
    for (int i = 0; i < 3; i++)
      for (int k = 0; k < 3; k++) j[i][k] = g.xyz[3 * g.nodes[e * 4 + k + 1] + i] - g.xyz[3 * g.nodes[e * 4] + i];
    const double det = -j[0][2] * j[1][1] * j[2][0] + j[0][1] * j[1][2] * j[2][0] +
                       j[0][2] * j[1][0] * j[2][1] - j[0][0] * j[1][2] * j[2][1] -
                       j[0][1] * j[1][0] * j[2][2] + j[0][0] * j[1][1] * j[2][2];
    s_jinv[0][0] = (-j[1][2] * j[2][1] + j[1][1] * j[2][2]) / det;
    s_jinv[0][1] = (j[0][2] * j[2][1] - j[0][1] * j[2][2]) / det;
    s_jinv[0][2] = (-j[0][2] * j[1][1] + j[0][1] * j[1][2]) / det;
    s_jinv[1][0] = (j[1][2] * j[2][0] - j[1][0] * j[2][2]) / det;
    s_jinv[1][1] = (-j[0][2] * j[2][0] + j[0][0] * j[2][2]) / det;
    s_jinv[1][2] = (j[0][2] * j[1][0] - j[0][0] * j[1][2]) / det;
    s_jinv[2][0] = (-j[1][1] * j[2][0] + j[1][0] * j[2][1]) / det;
    s_jinv[2][1] = (j[0][1] * j[2][0] - j[0][0] * j[2][1]) / det;
    s_jinv[2][2] = (-j[0][1] * j[1][0] + j[0][0] * j[1][1]) / det;
    s_det = fabs(det);
  }
  int ti[MT], tj[MT];
#pragma unroll
  for (int m = 0; m < MT; m++) {
    int t = tid + m * TPB, a = 0;
    if (t < NTILE) {
      while (t >= NT - a) { t -= NT - a; a++; }
      ti[m] = a;
      tj[m] = a + t;
    } else {
      ti[m] = tj[m] = -1;
    }
  }
  __syncthreads();
  double acc[MT][TS][TS][9];
#pragma unroll
  for (int m = 0; m < MT; m++)
#pragma unroll
    for (int r = 0; r < TS; r++)
#pragma unroll
      for (int q = 0; q < TS; q++)
#pragma unroll
        for (int x = 0; x < 9; x++) acc[m][r][q][x] = 0.;
  double accf[3] = {0., 0., 0.}, accp[3] = {0., 0., 0.};
  const int nq = T.nq;
  for (int q0 = 0; q0 < nq; q0 += QC) {
    const int nqc = nq - q0 < QC ? nq - q0 : QC;
    if (tid < nqc) {
      const double *xi = T.pt[q0 + tid];
      const double l[4] = {1. - xi[0] - xi[1] - xi[2], xi[0], xi[1], xi[2]};
      double x[3] = {0., 0., 0.}, f[3];
      for (int n = 0; n < 4; n++)
        for (int i = 0; i < 3; i++) x[i] += s_coord[3 * n + i] * l[n];
      forcing(g, x, f);
      s_f[tid][0] = f[0];
      s_f[tid][1] = f[1];
      s_f[tid][2] = f[2];
    }
    for (int k = tid; k < nqc * NSUB; k += TPB) {
      const int qq = k / NSUB, i = k % NSUB, q = q0 + qq, col = s_col[i];
      const double wd = T.w[q] * s_det;
      for (int a = 0; a < 3; a++) {
        double sum = 0.;
        for (int b = 0; b < 3; b++) sum += s_jinv[b][a] * T.dphi[q][b][col];
        s_G[qq][a][i] = sum;
        s_Gw[qq][a][i] = sum * wd;
      }
      s_pw[qq][i] = wd * T.phi[q][col];
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < MT; m++) {
      if (ti[m] < 0) continue;
      for (int qq = 0; qq < nqc; qq++) {
        double u[3][TS], v[3][TS];
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
          for (int r = 0; r < TS; r++) {
            u[a][r] = s_Gw[qq][a][TS * ti[m] + r];
            v[a][r] = s_G[qq][a][TS * tj[m] + r];
          }
#pragma unroll
        for (int r = 0; r < TS; r++)
#pragma unroll
          for (int q = 0; q < TS; q++)
#pragma unroll
            for (int a = 0; a < 3; a++)
#pragma unroll
              for (int b = 0; b < 3; b++) acc[m][r][q][3 * a + b] += u[a][r] * v[b][q];
      }
    }
    if (tid < NSUB)
      for (int qq = 0; qq < nqc; qq++) {
        for (int c = 0; c < 3; c++) {
          accf[c] += s_pw[qq][tid] * s_f[qq][c];
          accp[c] += s_Gw[qq][c][tid];
        }
      }
    __syncthreads();
  }
  const double E = g.coef ? g.coef[e] : g.E, nu = g.nu;
  const double C1 = E / (2. + 2. * nu), C2 = E * nu / (-1. + nu + 2. * nu * nu),
               C3 = E * (nu - 1.) / (-1. + nu + 2. * nu * nu);
#pragma unroll
  for (int m = 0; m < MT; m++) {
    if (ti[m] < 0) continue;
#pragma unroll
    for (int r = 0; r < TS; r++)
#pragma unroll
      for (int q = 0; q < TS; q++) {
        const int i = TS * ti[m] + r, j = TS * tj[m] + q;
        if (i >= NSUB || j >= NSUB || (ti[m] == tj[m] && q < r)) continue;
        const int64_t bi = s_blk[i], bj = s_blk[j];
        if (bi < 0 || bj < 0) continue;
        const double *Q = acc[m][r][q];  // Q[3x+y] = sum w dphi_x(i) dphi_y(j)
        for (int ci = 0; ci < 3; ci++)
          for (int cj = 0; cj < 3; cj++) {
            if (bi == bj && cj < ci) continue;
            double v;
            if (ci == cj) {
              double tr = 0.;
              for (int cc = 0; cc < 3; cc++)
                if (cc != ci) tr += Q[4 * cc];
              v = tr * C1 + Q[4 * ci] * C3;
            } else {
              v = Q[3 * cj + ci] * C1 - Q[3 * ci + cj] * C2;
            }
            if (fabs(v) < 1.e-12) continue;  // IsZero drop rule of AddKel
            int64_t slot;
            if (bi == bj) slot = s_row[i][ci] + (cj - ci);
            else if (bi < bj) slot = s_row[i][ci] + (3 - ci) + 3 * (int64_t)s_map[i * NSUB + j] + cj;
            else slot = s_row[j][cj] + (3 - cj) + 3 * (int64_t)s_map[j * NSUB + i] + ci;
            g.a[slot] += v;
          }
      }
  }
  if (tid < NSUB && s_blk[tid] >= 0)
    for (int c = 0; c < 3; c++)  // body load, minus w prestress_c dphix_c (fPreStress)
      g.rhs[s_blk[tid] * 3 + c] += accf[c] - g.prestress[c] * accp[c];
}

// NS states: ek block (i, j) = ck (NS x NS) M_ij, ef_i = (cf + g) m_i with the
// boundary mass M and moment m; g = -Val1 . n for the TPZElasticity3D stress
// field (constant on the flat triangle), n the unit normal turned away from
// the volume neighbour as in k_bc (TPZMatPoisson type 2 is refused on these
// meshes by pzg_mesh_set).
template <int NSUB, int NS>
__global__ __launch_bounds__(TPB) void k_bc_hp(ElemArgs g, int nsv, const HpTables *__restrict__ tab,
                                               const int32_t *__restrict__ sub,
                                               const int32_t *__restrict__ list) {
  constexpr int QC = 32;
  constexpr int NPAIR = NSUB * NSUB;
  constexpr int PER = (NPAIR + TPB - 1) / TPB;
  __shared__ double s_det, s_g[3];
  __shared__ double s_p[QC][NSUB];
  __shared__ double s_pw[QC][NSUB];
  __shared__ int s_col[NSUB];
  __shared__ int64_t s_blk[NSUB], s_row[NSUB][NS];
  __shared__ int32_t s_map[NPAIR];
  const int tid = threadIdx.x;
  const int64_t e = list[blockIdx.x];
  const HpTables &T = *tab;
  const int id = g.bcid[e];
  const double *cfs = g.bccoef + (int64_t)id * BCC;
  for (int k = tid; k < NPAIR; k += TPB) {
    const int i = k / NSUB, j = k % NSUB;
    s_map[k] = g.map[((int64_t)e * nsv + sub[i]) * nsv + sub[j]];
  }
  if (tid < NSUB) {
    int64_t ids[3];
    for (int k = 0; k < 3; k++) ids[k] = g.nodes[e * 3 + k];
    const int gi = sub[tid];
    s_col[tid] = T.col_off[gi] + hp_variant(false, T.shape_side[gi], ids);
    const int64_t b = g.aidx[g.conn[e * nsv + gi]];
    s_blk[tid] = b;
    for (int c = 0; c < NS; c++) s_row[tid][c] = b >= 0 ? g.ia[b * NS + c] : -1;
  }
  if (tid == 0) {
    // 2D element in 3D space (TPZGeoEl::Jacobian): |det| = |g0| * |g1 - (g0.g1/|g0|^2) g0|
    double gx[3][2];
    for (int i = 0; i < 3; i++)
      for (int k = 0; k < 2; k++) gx[i][k] = g.xyz[3 * g.nodes[e * 3 + k + 1] + i] - g.xyz[3 * g.nodes[e * 3] + i];
    double n1 = 0, d12 = 0, n2 = 0, a1[3];
    for (int i = 0; i < 3; i++) { n1 += gx[i][0] * gx[i][0]; d12 += gx[i][0] * gx[i][1]; }
    n1 = sqrt(n1);
    for (int i = 0; i < 3; i++) {
      a1[i] = gx[i][1] - d12 * (gx[i][0] / n1) / n1;
      n2 += a1[i] * a1[i];
    }
    n2 = sqrt(n2);
    s_det = fabs(n1 * n2);
    s_g[0] = s_g[1] = s_g[2] = 0.;
    if (NS == 3 && cfs[BCC_STRESS] != 0.) {
      // TPZInterpolationSpace::ComputeNormal: axes a0 x a1, turned away from
      // the volume neighbour (g.bcvec: neighbour centre -> element centre)
      double a0[3], t[3];
      for (int i = 0; i < 3; i++) { a0[i] = gx[i][0] / n1; a1[i] /= n2; }
      t[0] = a0[1] * a1[2] - a0[2] * a1[1];
      t[1] = -a0[0] * a1[2] + a0[2] * a1[0];
      t[2] = a0[0] * a1[1] - a0[1] * a1[0];
      const double *v = g.bcvec + 3 * e;
      const double sz = sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
      double nrm[3], dot = 0.;
      for (int i = 0; i < 3; i++) { nrm[i] = t[i] / sz; dot += nrm[i] * v[i]; }
      if (dot < 0.)
        for (int i = 0; i < 3; i++) nrm[i] = -nrm[i];
      for (int k = 0; k < 3; k++)
        s_g[k] = -(cfs[BCC_VAL1 + 3 * k] * nrm[0] + cfs[BCC_VAL1 + 3 * k + 1] * nrm[1] +
                   cfs[BCC_VAL1 + 3 * k + 2] * nrm[2]);
    }
  }
  __syncthreads();
  double acc[PER];
#pragma unroll
  for (int k = 0; k < PER; k++) acc[k] = 0.;
  double accf = 0.;
  const int nq = T.nq;
  for (int q0 = 0; q0 < nq; q0 += QC) {
    const int nqc = nq - q0 < QC ? nq - q0 : QC;
    for (int k = tid; k < nqc * NSUB; k += TPB) {
      const int qq = k / NSUB, i = k % NSUB, q = q0 + qq;
      const double ph = T.phi[q][s_col[i]];
      s_p[qq][i] = ph;
      s_pw[qq][i] = ph * T.w[q] * s_det;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < PER; k++) {
      const int idx = tid + k * TPB;
      if (idx < NPAIR) {
        const int i = idx / NSUB, j = idx % NSUB;
        double sum = 0.;
        for (int qq = 0; qq < nqc; qq++) sum += s_pw[qq][i] * s_p[qq][j];
        acc[k] += sum;
      }
    }
    if (tid < NSUB)
      for (int qq = 0; qq < nqc; qq++) accf += s_pw[qq][tid];
    __syncthreads();
  }
#pragma unroll
  for (int k = 0; k < PER; k++) {
    const int idx = tid + k * TPB;
    if (idx >= NPAIR) continue;
    const int i = idx / NSUB, j = idx % NSUB;
    const int64_t bi = s_blk[i], bj = s_blk[j];
    if (bi < 0 || bj < 0 || bi > bj) continue;
    for (int ci = 0; ci < NS; ci++)
      for (int cj = bi == bj ? ci : 0; cj < NS; cj++) {
        const double v = cfs[BCC_CK + NS * ci + cj] * acc[k];
        if (fabs(v) < 1.e-12) continue;  // AddKel's drop rule
        g.a[bi == bj ? s_row[i][ci] + (cj - ci) : s_row[i][ci] + (NS - ci) + NS * (int64_t)s_map[i * NSUB + j] + cj] += v;
      }
  }
  if (tid < NSUB && s_blk[tid] >= 0)
    for (int c = 0; c < NS; c++) g.rhs[s_blk[tid] * NS + c] += (cfs[BCC_CF + c] + s_g[c]) * accf;
}

static ElemArgs make_args(Ctx &c, bool vol) {
  ElemArgs g;
  g.xyz = c.d_xyz.as<double>();
  g.nodes = vol ? c.d_vnodes.as<int64_t>() : c.d_bnodes.as<int64_t>();
  g.conn = vol ? c.d_vconn.as<int64_t>() : c.d_bconn.as<int64_t>();
  g.coef = (vol && !c.vol_coef.empty()) ? c.d_vcoef.as<double>() : nullptr;
  g.aidx = c.d_aidx.as<int64_t>();
  g.map = vol ? c.d_vmap.as<int32_t>() : c.d_bmap.as<int32_t>();
  g.ia = c.d_ia.as<int64_t>();
  g.a = c.d_a.as<double>();
  g.rhs = c.d_rhs.as<double>();
  g.tab = vol ? c.d_tvol.as<ElemTables>() : c.d_tbc.as<ElemTables>();
  g.bumps = c.d_bumps.as<double>();
  g.bcid = c.d_bcid.as<int32_t>();
  g.bctype = c.d_bctype.as<int32_t>();
  g.bcval = c.d_bcval.as<double>();
  g.bccoef = c.d_bccoef.as<double>();
  g.bcvec = c.d_bcvec.as<double>();
  g.nbumps = c.nbumps;
  g.bw = c.bump_width;
  for (int k = 0; k < 3; k++) g.fconst[k] = c.force_const[k];
  g.force_kind = c.force_kind;
  g.scale = c.scale;
  g.E = c.E;
  g.nu = c.nu;
  for (int k = 0; k < 3; k++) g.prestress[k] = c.prestress[k];
  g.big = c.big;
  g.material = c.material;
  g.bs = c.bs;
  return g;
}

template <int DIM, int NC, int NSD, int NSH, int NS>
static int run_vol(Ctx &c) {
  ElemArgs g = make_args(c, true);
  const size_t nbk = c.vol_rules.size();
  // affine elements first in each colour (c.vcolor_aff of them), when the
  // reference Gram matrices of the single rule are on the device
  const bool aff = !c.vcolor_aff.empty() && c.d_vol_R.p && c.aff_on;
  // identical affine elements: the stiffness by blocks (k_vol_blk, which
  // writes every stored entry), the load vector by the affine kernel
  const bool blk = aff && c.blk_asm && NSH * NS <= 81;
  if (blk) {
    if (c.blk_grid <= 0) {
      int per_cu = 0, ncu = 0;
      PZG_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c.device));
      PZG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_vol_blk<NS, NSH * NS>, BLK_TPB, 0));
      c.blk_grid = std::max(1, per_cu) * std::max(1, ncu);
    }
    const int64_t rows = (c.nb + BLK_TPB / 64 - 1) / (BLK_TPB / 64);
    hipLaunchKernelGGL((k_vol_blk<NS, NSH * NS>), dim3((unsigned)std::max<int64_t>(1, std::min<int64_t>(rows, c.blk_grid))),
                       dim3(BLK_TPB), 0, c.stream, c.nb, c.d_ia.as<int64_t>(), c.d_ubia.as<int64_t>(),
                       c.d_ubja.as<int32_t>(), c.d_blk_iptr.as<int64_t>(), c.d_blk_inc.as<int32_t>(),
                       g.coef, c.material == PZG_MAT_ELAST3D ? c.E : c.scale, c.d_blk_kref.as<double>(),
                       c.d_a.as<double>());
    PZG_HIP(hipGetLastError());
  }
  for (size_t k = 0; k + 1 < c.vcolor_ptr.size(); k++) {
    int64_t n = c.vcolor_ptr[k + 1] - c.vcolor_ptr[k];
    if (!n) continue;
    g.tab = c.d_tvol.as<ElemTables>() + k % nbk;
    const int32_t *lst = c.d_vcolor_list.as<int32_t>() + c.vcolor_ptr[k];
    int64_t na = aff ? c.vcolor_aff[k] : 0;
    if (na > 0) {
      if (blk)
        hipLaunchKernelGGL((k_vol_aff<DIM, NC, NSD, NSH, NS, false>),
                           dim3((unsigned)((na + AFF_EPB - 1) / AFF_EPB)), dim3(TPB), 0, c.stream, g,
                           c.d_vol_R.as<double>(), lst, na);
      else
        hipLaunchKernelGGL((k_vol_aff<DIM, NC, NSD, NSH, NS>), dim3((unsigned)((na + AFF_EPB - 1) / AFF_EPB)),
                           dim3(TPB), 0, c.stream, g, c.d_vol_R.as<double>(), lst, na);
      PZG_HIP(hipGetLastError());
    }
    if (n > na) {
      hipLaunchKernelGGL((k_vol<DIM, NC, NSD, NSH, NS>), dim3((unsigned)(n - na)), dim3(TPB), 0, c.stream, g,
                         lst + na);
      PZG_HIP(hipGetLastError());
    }
  }
  return PZG_OK;
}

// Constrained elements (TPZElementMatrixT::ApplyConstraints, Mesh/
// TPZElementMatrixT.cpp:134-322, and the assembly of StrMatrix/
// pzstrmatrixor.cpp:273-280): with T[i][a] the weight of list entry a in
// shape i (1 for the shape's own unconstrained connect, the dependency
// coefficients for a constrained one), K_c = T^T K T and F_c = T^T F over the
// element's unconstrained blocks, added with AddKel's upper-part rule and
// drop rule |v| < 1e-12, and AddFel.  T is sparse (a list entry collects a
// few shapes), so each entry of K_c is a short double sum over the shapes of
// its row and column entries.  One workgroup per element of a colour
// (elements of a colour share no block).
template <int NSH, int NS>
__global__ __launch_bounds__(TPB) void k_constrain(const int32_t *__restrict__ slots,
                                                   const double *__restrict__ ek,
                                                   const double *__restrict__ ef,
                                                   const double *__restrict__ Tm, int nsv,
                                                   const int64_t *__restrict__ blk,
                                                   const int32_t *__restrict__ ncl, const int64_t *__restrict__ ia,
                                                   const int64_t *__restrict__ ja, double *__restrict__ a,
                                                   double *__restrict__ rhs, int bs) {
  constexpr int NE = NSH * NS;
  __shared__ int s_cnt[CEL_NC];
  __shared__ int s_i[CEL_NC][NSH];
  __shared__ double s_c[CEL_NC][NSH];
  __shared__ int64_t s_b[CEL_NC];
  const int tid = threadIdx.x;
  const int64_t k = slots[blockIdx.x];
  const int n = ncl[k];
  const double *K = ek + k * NE * NE;
  for (int e = tid; e < n; e += TPB) {
    int cnt = 0;
    for (int i = 0; i < NSH; i++) {
      const double t = Tm[(k * nsv + i) * CEL_NC + e];
      if (t != 0.) {
        s_i[e][cnt] = i;
        s_c[e][cnt] = t;
        cnt++;
      }
    }
    s_cnt[e] = cnt;
    s_b[e] = blk[k * CEL_NC + e];
  }
  __syncthreads();
  const int nn = n * NS;
  for (int x = tid; x < nn * nn; x += TPB) {
    const int A = x / nn, B = x % nn, ea = A / NS, ca = A % NS, eb = B / NS, cb = B % NS;
    const int64_t ra = s_b[ea] * bs + ca, rb = s_b[eb] * bs + cb;
    if (rb < ra) continue;
    double v = 0.;
    for (int p = 0; p < s_cnt[ea]; p++) {
      const double *Kr = K + (s_i[ea][p] * NS + ca) * NE + cb;
      double w = 0.;
      for (int q = 0; q < s_cnt[eb]; q++) w += s_c[eb][q] * Kr[s_i[eb][q] * NS];
      v += s_c[ea][p] * w;
    }
    if (fabs(v) < 1.e-12) continue;
    int64_t l = ia[ra], h = ia[ra + 1] - 1;
    while (l <= h) {
      const int64_t m = (l + h) >> 1;
      if (ja[m] == rb) {
        a[m] += v;
        break;
      }
      if (ja[m] < rb) l = m + 1;
      else h = m - 1;
    }
  }
  for (int A = tid; A < nn; A += TPB) {
    const int ea = A / NS, ca = A % NS;
    double f = 0.;
    for (int p = 0; p < s_cnt[ea]; p++) f += s_c[ea][p] * ef[k * NE + s_i[ea][p] * NS + ca];
    rhs[s_b[ea] * bs + ca] += f;
  }
}

// K_c = T^T K T of a set of constrained elements whose element matrices are
// in cs.d_ek / cs.d_ef, colour by colour.
template <int NSH, int NS>
static int constrain_set(Ctx &c, CelSet &cs, int nsx) {
  for (size_t k = 0; k + 1 < cs.cptr.size(); k++) {
    const int64_t n = cs.cptr[k + 1] - cs.cptr[k];
    if (!n) continue;
    hipLaunchKernelGGL((k_constrain<NSH, NS>), dim3(n), dim3(TPB), 0, c.stream,
                       cs.d_cslot.as<int32_t>() + cs.cptr[k], cs.d_ek.as<double>(), cs.d_ef.as<double>(),
                       cs.d_T.as<double>(), nsx, cs.d_blk.as<int64_t>(), cs.d_n.as<int32_t>(),
                       c.d_ia.as<int64_t>(), c.d_ja.as<int64_t>(), c.d_a.as<double>(),
                       c.d_rhs.as<double>(), c.bs);
    PZG_HIP(hipGetLastError());
  }
  return PZG_OK;
}

// constrained volume elements: element matrices (k_vol, EK), then K_c
template <int DIM, int NC, int NSD, int NSH, int NS>
static int run_constrained(Ctx &c) {
  CelSet &cs = c.cv;
  if (!c.has_dep || cs.cel.empty()) return PZG_OK;
  constexpr int NE = NSH * NS;
  const int64_t ncel = (int64_t)cs.cel.size();
  PZG_TRY(dev_alloc(cs.d_ek, sizeof(double) * ncel * NE * NE));
  PZG_TRY(dev_alloc(cs.d_ef, sizeof(double) * ncel * NE));
  ElemArgs g = make_args(c, true);
  g.aidx_ek = c.d_aidx_ek.as<int64_t>();
  for (size_t b = 0; b + 1 < cs.bptr.size(); b++) {
    const int64_t n = cs.bptr[b + 1] - cs.bptr[b];
    if (!n) continue;
    g.tab = c.d_tvol.as<ElemTables>() + b;
    g.ek_out = cs.d_ek.as<double>() + cs.bptr[b] * NE * NE;
    g.ef_out = cs.d_ef.as<double>() + cs.bptr[b] * NE;
    hipLaunchKernelGGL((k_vol<DIM, NC, NSD, NSH, NS, true>), dim3(n), dim3(TPB), 0, c.stream, g,
                       cs.d_cel.as<int32_t>() + cs.bptr[b]);
    PZG_HIP(hipGetLastError());
  }
  return constrain_set<NSH, NS>(c, cs, c.vol_nsides);
}

// constrained boundary elements: element matrices (k_bc, EK), then K_c
template <int DIMB, int NC, int NSD, int NSH, int NS>
static int run_constrained_bc(Ctx &c) {
  CelSet &cs = c.cb;
  if (!c.has_dep || cs.cel.empty()) return PZG_OK;
  constexpr int NE = NSH * NS;
  const int64_t ncel = (int64_t)cs.cel.size();
  PZG_TRY(dev_alloc(cs.d_ek, sizeof(double) * ncel * NE * NE));
  PZG_TRY(dev_alloc(cs.d_ef, sizeof(double) * ncel * NE));
  ElemArgs g = make_args(c, false);
  g.aidx_ek = c.d_aidx_ek.as<int64_t>();
  for (size_t b = 0; b + 1 < cs.bptr.size(); b++) {
    const int64_t n = cs.bptr[b + 1] - cs.bptr[b];
    if (!n) continue;
    g.tab = c.d_tbc.as<ElemTables>() + b;
    g.ek_out = cs.d_ek.as<double>() + cs.bptr[b] * NE * NE;
    g.ef_out = cs.d_ef.as<double>() + cs.bptr[b] * NE;
    hipLaunchKernelGGL((k_bc<DIMB, NC, NSD, NSH, NS, true>), dim3(n), dim3(TPB), 0, c.stream, g,
                       cs.d_cel.as<int32_t>() + cs.bptr[b]);
    PZG_HIP(hipGetLastError());
  }
  return constrain_set<NSH, NS>(c, cs, c.bc_nsides);
}

template <int DIMB, int NC, int NSD, int NSH, int NS>
static int run_bc(Ctx &c) {
  if (c.nbc == 0) return PZG_OK;
  ElemArgs g = make_args(c, false);
  const size_t nbk = c.bc_rules.size();
  for (size_t k = 0; k + 1 < c.bcolor_ptr.size(); k++) {
    int64_t n = c.bcolor_ptr[k + 1] - c.bcolor_ptr[k];
    if (!n) continue;
    g.tab = c.d_tbc.as<ElemTables>() + k % nbk;
    hipLaunchKernelGGL((k_bc<DIMB, NC, NSD, NSH, NS>), dim3(n), dim3(TPB), 0, c.stream, g,
                       c.d_bcolor_list.as<int32_t>() + c.bcolor_ptr[k]);
    PZG_HIP(hipGetLastError());
  }
  return PZG_OK;
}

// shape counts of the order-P sets: tetrahedron 4/10/20/35/56, triangle 3/6/10/15/21.
// Volume kernel variants (QC quadrature points per pass, TSxTS tiles):
// 0: (16,4)  1: (8,4)  2: (32,4)  3: (16,2)  4: (32,2)
constexpr int HP_NVARIANT = 5;
template <int NV>
static void launch_vol_variant(Ctx &c, const ElemArgs &g, int var, int nsv, const HpTables *tab,
                               const int32_t *sub, const int32_t *list, int64_t n) {
  switch (var) {
    case 1: hipLaunchKernelGGL((k_vol_hp<NV, 8, 4>), dim3(n), dim3(TPB), 0, c.stream, g, nsv, tab, sub, list); break;
    case 2: hipLaunchKernelGGL((k_vol_hp<NV, 32, 4>), dim3(n), dim3(TPB), 0, c.stream, g, nsv, tab, sub, list); break;
    case 3: hipLaunchKernelGGL((k_vol_hp<NV, 16, 2>), dim3(n), dim3(TPB), 0, c.stream, g, nsv, tab, sub, list); break;
    case 4: hipLaunchKernelGGL((k_vol_hp<NV, 32, 2>), dim3(n), dim3(TPB), 0, c.stream, g, nsv, tab, sub, list); break;
    default: hipLaunchKernelGGL((k_vol_hp<NV, 16, 4>), dim3(n), dim3(TPB), 0, c.stream, g, nsv, tab, sub, list); break;
  }
}

// elasticity (nine sums per shape pair): 0: (16,2)  1: (8,2)  2: (32,2)  3: (16,1)  4: (8,1)
template <int NV>
static void launch_vol_variant3(Ctx &c, const ElemArgs &g, int var, int nsv, const HpTables *tab,
                                const int32_t *sub, const int32_t *list, int64_t n) {
  switch (var) {
    case 1: hipLaunchKernelGGL((k_vol_hp3<NV, 8, 2>), dim3(n), dim3(TPB), 0, c.stream, g, nsv, tab, sub, list); break;
    case 2: hipLaunchKernelGGL((k_vol_hp3<NV, 32, 2>), dim3(n), dim3(TPB), 0, c.stream, g, nsv, tab, sub, list); break;
    case 3: hipLaunchKernelGGL((k_vol_hp3<NV, 16, 1>), dim3(n), dim3(TPB), 0, c.stream, g, nsv, tab, sub, list); break;
    case 4: hipLaunchKernelGGL((k_vol_hp3<NV, 8, 1>), dim3(n), dim3(TPB), 0, c.stream, g, nsv, tab, sub, list); break;
    default: hipLaunchKernelGGL((k_vol_hp3<NV, 16, 2>), dim3(n), dim3(TPB), 0, c.stream, g, nsv, tab, sub, list); break;
  }
}

template <bool VOL>
static int launch_hp_bucket(Ctx &c, const ElemArgs &g, int P, int var, const HpTables *tab,
                            const int32_t *sub, const int32_t *list, int64_t n) {
  const int nsv = VOL ? c.vol_nsides : c.bc_nsides;
  const bool el = c.bs == 3;
#define PZG_HP_CASE(PP, NV, NB)                                                                        \
  case PP:                                                                                             \
    if (VOL && el) launch_vol_variant3<NV>(c, g, var, nsv, tab, sub, list, n);                         \
    else if (VOL) launch_vol_variant<NV>(c, g, var, nsv, tab, sub, list, n);                           \
    else if (el) hipLaunchKernelGGL((k_bc_hp<NB, 3>), dim3(n), dim3(TPB), 0, c.stream, g, nsv, tab, sub, list); \
    else hipLaunchKernelGGL((k_bc_hp<NB, 1>), dim3(n), dim3(TPB), 0, c.stream, g, nsv, tab, sub, list); \
    break;
  switch (P) {
    PZG_HP_CASE(1, 4, 3)
    PZG_HP_CASE(2, 10, 6)
    PZG_HP_CASE(3, 20, 10)
    PZG_HP_CASE(4, 35, 15)
    PZG_HP_CASE(5, 56, 21)
    default: set_error("pzg_assemble: unsupported order"); return PZG_EUNSUP;
  }
#undef PZG_HP_CASE
  PZG_HIP(hipGetLastError());
  return PZG_OK;
}

// Autotuner of the volume kernels: per rule bucket, every variant runs on the
// bucket's largest colour list (once to warm up, then timed over three
// launches with HIP events) and the fastest is kept for the context.  The
// launches scatter into the system, which launch_assemble zeroes afterwards.
static int hp_tune(Ctx &c, const ElemArgs &g) {
  const size_t nbk = c.vol_rules.size();
  c.hp_vol_var.assign(nbk, 0);
  c.hp_vol_ms.assign(nbk * HP_NVARIANT, 0.f);
  hipEvent_t e0, e1;
  PZG_HIP(hipEventCreate(&e0));
  PZG_HIP(hipEventCreate(&e1));
  int rc = PZG_OK;
  for (size_t b = 0; b < nbk && rc == PZG_OK; b++) {
    int64_t best = -1, nbest = 0;
    for (size_t k = b; k + 1 < c.vcolor_ptr.size(); k += nbk) {
      const int64_t n = c.vcolor_ptr[k + 1] - c.vcolor_ptr[k];
      if (n > nbest) { nbest = n; best = (int64_t)k; }
    }
    if (best < 0) continue;
    const int32_t *list = c.d_vcolor_list.as<int32_t>() + c.vcolor_ptr[best];
    const HpTables *tab = c.d_thp_vol.as<HpTables>() + b;
    const int32_t *sub = c.d_hp_vsub.as<int32_t>() + b * HP_MAXSH;
    float tbest = 0.f;
    for (int v = 0; v < HP_NVARIANT && rc == PZG_OK; v++) {
      rc = launch_hp_bucket<true>(c, g, c.hp_vol_P[b], v, tab, sub, list, nbest);
      if (rc != PZG_OK) break;
      if (hipEventRecord(e0, c.stream) != hipSuccess) { rc = PZG_EHIP; break; }
      for (int r = 0; r < 3 && rc == PZG_OK; r++)
        rc = launch_hp_bucket<true>(c, g, c.hp_vol_P[b], v, tab, sub, list, nbest);
      if (rc != PZG_OK) break;
      float ms = 0.f;
      if (hipEventRecord(e1, c.stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
          hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
        set_error("hp_tune: event timing failed");
        rc = PZG_EHIP;
        break;
      }
      c.hp_vol_ms[b * HP_NVARIANT + v] = ms / 3.f;
      if (v == 0 || ms < tbest) { tbest = ms; c.hp_vol_var[b] = v; }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc == PZG_OK) c.hp_tuned = true;
  return rc;
}

static int run_hp(Ctx &c) {
  if (c.bs != 1 && !(c.bs == 3 && c.material == PZG_MAT_ELAST3D)) {
    set_error("pzg_assemble: high-order tetrahedra: block size other than the material's states");
    return PZG_EUNSUP;
  }
  if (c.thp_vol.size() != c.vol_rules.size() || c.hp_vol_P.size() != c.vol_rules.size() ||
      (c.nbc > 0 && (c.thp_bc.size() != c.bc_rules.size() || c.hp_bc_P.size() != c.bc_rules.size()))) {
    set_error("pzg_assemble: high-order tables do not match the mesh");
    return PZG_EINVAL;
  }
  ElemArgs g = make_args(c, true);
  size_t nbk = c.vol_rules.size();
  if (!c.hp_tuned) {
    PZG_TRY(hp_tune(c, g));
    fill<double>(c, c.d_a.as<double>(), c.nnz_u, 0.);
    fill<double>(c, c.d_rhs.as<double>(), c.neq, 0.);
  }
  for (size_t k = 0; k + 1 < c.vcolor_ptr.size(); k++) {
    const int64_t n = c.vcolor_ptr[k + 1] - c.vcolor_ptr[k];
    if (!n) continue;
    const size_t b = k % nbk;
    PZG_TRY(launch_hp_bucket<true>(c, g, c.hp_vol_P[b], c.hp_vol_var[b], c.d_thp_vol.as<HpTables>() + b,
                                   c.d_hp_vsub.as<int32_t>() + b * HP_MAXSH,
                                   c.d_vcolor_list.as<int32_t>() + c.vcolor_ptr[k], n));
  }
  if (c.nbc == 0) return PZG_OK;
  g = make_args(c, false);
  nbk = c.bc_rules.size();
  for (size_t k = 0; k + 1 < c.bcolor_ptr.size(); k++) {
    const int64_t n = c.bcolor_ptr[k + 1] - c.bcolor_ptr[k];
    if (!n) continue;
    const size_t b = k % nbk;
    PZG_TRY(launch_hp_bucket<false>(c, g, c.hp_bc_P[b], 0, c.d_thp_bc.as<HpTables>() + b,
                                    c.d_hp_bsub.as<int32_t>() + b * HP_MAXSH,
                                    c.d_bcolor_list.as<int32_t>() + c.bcolor_ptr[k], n));
  }
  return PZG_OK;
}

int launch_assemble(Ctx &c) {
  // (the block assembly of identical elements writes every stored entry)
  if (!(c.blk_asm && c.aff_on)) fill<double>(c, c.d_a.as<double>(), c.nnz_u, 0.);
  fill<double>(c, c.d_rhs.as<double>(), c.neq, 0.);
  if (c.hp_simplex) return run_hp(c);
  for (const ElemTables &t : c.tvol_all)
    if (t.nq > MAXQ) return PZG_EUNSUP;
  for (const ElemTables &t : c.tbc_all)
    if (t.nq > MAXQ) return PZG_EUNSUP;
  int p = c.porder;
  if (c.vol_etype == PZG_TET) {
    // tetrahedra (4 corners, 15 sides; p <= 2: 4 or 10 shapes) with boundary
    // triangles (3 corners, 7 sides; 3 or 6 shapes)
    if (c.nstate == 3) {
      if (p == 2) { PZG_TRY((run_vol<3, 4, 15, 10, 3>(c))); PZG_TRY((run_bc<2, 3, 7, 6, 3>(c))); }
      else { PZG_TRY((run_vol<3, 4, 15, 4, 3>(c))); PZG_TRY((run_bc<2, 3, 7, 3, 3>(c))); }
    } else {
      if (p == 2) { PZG_TRY((run_vol<3, 4, 15, 10, 1>(c))); PZG_TRY((run_bc<2, 3, 7, 6, 1>(c))); }
      else { PZG_TRY((run_vol<3, 4, 15, 4, 1>(c))); PZG_TRY((run_bc<2, 3, 7, 3, 1>(c))); }
    }
    return PZG_OK;
  }
  if (c.dim == 3 && c.nstate == 3) {
    if (p == 2) {
      PZG_TRY((run_vol<3, 8, 27, 27, 3>(c)));
      PZG_TRY((run_constrained<3, 8, 27, 27, 3>(c)));
      PZG_TRY((run_bc<2, 4, 9, 9, 3>(c)));
      PZG_TRY((run_constrained_bc<2, 4, 9, 9, 3>(c)));
    } else {
      PZG_TRY((run_vol<3, 8, 27, 8, 3>(c)));
      PZG_TRY((run_constrained<3, 8, 27, 8, 3>(c)));
      PZG_TRY((run_bc<2, 4, 9, 4, 3>(c)));
      PZG_TRY((run_constrained_bc<2, 4, 9, 4, 3>(c)));
    }
  } else if (c.dim == 3 && c.nstate == 1) {
    if (p == 2) {
      PZG_TRY((run_vol<3, 8, 27, 27, 1>(c)));
      PZG_TRY((run_constrained<3, 8, 27, 27, 1>(c)));
      PZG_TRY((run_bc<2, 4, 9, 9, 1>(c)));
      PZG_TRY((run_constrained_bc<2, 4, 9, 9, 1>(c)));
    } else {
      PZG_TRY((run_vol<3, 8, 27, 8, 1>(c)));
      PZG_TRY((run_constrained<3, 8, 27, 8, 1>(c)));
      PZG_TRY((run_bc<2, 4, 9, 4, 1>(c)));
      PZG_TRY((run_constrained_bc<2, 4, 9, 4, 1>(c)));
    }
  } else if (c.dim == 2 && c.nstate == 1) {
    if (p == 2) {
      PZG_TRY((run_vol<2, 4, 9, 9, 1>(c)));
      PZG_TRY((run_constrained<2, 4, 9, 9, 1>(c)));
      PZG_TRY((run_bc<1, 2, 3, 3, 1>(c)));
      PZG_TRY((run_constrained_bc<1, 2, 3, 3, 1>(c)));
    } else {
      PZG_TRY((run_vol<2, 4, 9, 4, 1>(c)));
      PZG_TRY((run_constrained<2, 4, 9, 4, 1>(c)));
      PZG_TRY((run_bc<1, 2, 3, 2, 1>(c)));
      PZG_TRY((run_constrained_bc<1, 2, 3, 2, 1>(c)));
    }
  } else {
    set_error("pzg_assemble: unsupported dimension/state combination");
    return PZG_EUNSUP;
  }
  return PZG_OK;
}

// --------------------------------------------------------------------------
__global__ void k_diag(int64_t n, const int64_t *__restrict__ ia, const double *__restrict__ a,
                       double *__restrict__ d) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = a[ia[i]];  // diagonal is the first entry of every row
}

// Equation filter: the inactive equations of the mesh numbering become
// decoupled unit rows (diagonal 1, the rest of the row, the column entries
// held in other rows and the load 0), so the device system equals the
// condensed TPZSYsmpMatrix on the active equations.  One thread per row.
__global__ void k_filter_mask(int64_t n, const int64_t *__restrict__ ia, const int64_t *__restrict__ ja,
                              const int64_t *__restrict__ dest, double *__restrict__ a,
                              double *__restrict__ rhs) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int64_t k0 = ia[r], k1 = ia[r + 1];
  if (dest[r] < 0) {
    a[k0] = 1.0;
    for (int64_t k = k0 + 1; k < k1; k++) a[k] = 0.0;
    rhs[r] = 0.0;
  } else {
    for (int64_t k = k0 + 1; k < k1; k++)
      if (dest[ja[k]] < 0) a[k] = 0.0;
  }
}

int launch_filter_mask(Ctx &c) {
  if (!c.filt || !c.neq) return PZG_OK;
  hipLaunchKernelGGL(k_filter_mask, dim3((c.neq + TPB - 1) / TPB), dim3(TPB), 0, c.stream, c.neq,
                     c.d_ia.as<int64_t>(), c.d_ja.as<int64_t>(), c.d_filt_dest.as<int64_t>(),
                     c.d_a.as<double>(), c.d_rhs.as<double>());
  PZG_HIP(hipGetLastError());
  return PZG_OK;
}

int launch_diag(Ctx &c) {
  if (!c.neq) return PZG_OK;
  hipLaunchKernelGGL(k_diag, dim3((c.neq + TPB - 1) / TPB), dim3(TPB), 0, c.stream, c.neq,
                     c.d_ia.as<int64_t>(), c.d_a.as<double>(), c.d_diag.as<double>());
  PZG_HIP(hipGetLastError());
  return PZG_OK;
}

// --------------------------------------------------------------------------
// Full (both triangles) FP64 block CSR of K, from the reference's upper CSR
// (k_bsr_gather).  Values at bsr_pos (chunks of BSR_CHUNK blocks per row,
// component-major inside a chunk).
__global__ void k_nlow(int64_t nb, const int64_t *__restrict__ fbia, const int64_t *__restrict__ ubia,
                       int32_t *__restrict__ nlow) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nb) nlow[i] = (int32_t)((fbia[i + 1] - fbia[i]) - 1 - (ubia[i + 1] - ubia[i]));
}

// nlow[A] = number of lower blocks of block row A (= rank of its diagonal block)


// Row-gather form of the expansion: one wavefront per block row A writes the
// row's storage in order (coalesced), each lane one value.  Upper blocks
// come from row A's scalar rows, lower blocks (A, B), B < A, from the upper
// block (B, A) in row B (lowsrc: its index among the upper blocks).  SCALED:
// the FP32 values of D^-1/2 K D^-1/2, diagonal blocks left to k_bsr_diag.
template <int BS, typename T, bool SCALED>
__global__ __launch_bounds__(256) void k_bsr_gather(int64_t nb, const int64_t *__restrict__ ia,
                                                    const double *__restrict__ a,
                                                    const int64_t *__restrict__ ubia,
                                                    const int64_t *__restrict__ fbia,
                                                    const int32_t *__restrict__ fbja,
                                                    const int32_t *__restrict__ lowsrc,
                                                    const double *__restrict__ dis, T *__restrict__ v) {
  const int64_t A = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (A >= nb) return;
  constexpr int BB = BS * BS;
  const int64_t b0 = fbia[A], u0 = ubia[A];
  const int n = (int)(fbia[A + 1] - b0), nl = n - 1 - (int)(ubia[A + 1] - u0);
  const int64_t lb = b0 - A - u0;  // first lower block of row A in lowsrc
  int64_t ra[BS];                  // row starts of A's equations in the upper CSR
  double da[BS];
#pragma unroll
  for (int k = 0; k < BS; k++) {
    ra[k] = ia[A * BS + k];
    da[k] = SCALED ? dis[A * BS + k] : 1.;
  }
  // one lane per block: its indices once, its BB values loaded together; the
  // value kl of block j sits at chunk offset kl*m + (j - j0), so the lanes of
  // a chunk write each kl contiguously
  for (int j = lane; j < n; j += 64) {
    if (SCALED && j == nl) continue;
    const int j0 = j / BSR_CHUNK * BSR_CHUNK, m = min(BSR_CHUNK, n - j0);
    T *out = v + (b0 + j0) * BB + (j - j0);
    double x[BB];
    if (j == nl) {
#pragma unroll
      for (int k = 0; k < BS; k++)
#pragma unroll
        for (int l = 0; l < BS; l++) x[k * BS + l] = k <= l ? a[ra[k] + (l - k)] : a[ra[l] + (k - l)];
    } else if (j > nl) {
      const int64_t t = j - nl - 1, B = fbja[b0 + j];
#pragma unroll
      for (int k = 0; k < BS; k++)
#pragma unroll
        for (int l = 0; l < BS; l++) {
          x[k * BS + l] = a[ra[k] + BS + t * BS - k + l];
          if (SCALED) x[k * BS + l] *= da[k] * dis[B * BS + l];
        }
    } else {
      // the transpose of the upper block (B, A) of row B
      const int64_t u = lowsrc[lb + j], B = fbja[b0 + j], t = u - ubia[B];
#pragma unroll
      for (int l = 0; l < BS; l++) {
        const int64_t r = ia[B * BS + l] + BS + t * BS - l;
        const double dl = SCALED ? dis[B * BS + l] : 1.;
#pragma unroll
        for (int k = 0; k < BS; k++) {
          x[k * BS + l] = a[r + k];
          if (SCALED) x[k * BS + l] *= dl * da[k];
        }
      }
    }
#pragma unroll
    for (int kl = 0; kl < BB; kl++) out[kl * m] = (T)x[kl];
  }
}

// lowsrc[lower block of (A, B)] = index of the upper block (B, A)
__global__ void k_lowsrc(int64_t nb, const int64_t *__restrict__ ubia, const int32_t *__restrict__ ubja,
                         const int32_t *__restrict__ rbwd, const int64_t *__restrict__ fbia,
                         int32_t *__restrict__ lowsrc) {
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (wave >= nb) return;
  const int64_t B = wave;
  for (int64_t u = ubia[B] + lane; u < ubia[B + 1]; u += 64) {
    const int64_t A = ubja[u];
    lowsrc[fbia[A] - A - ubia[A] + rbwd[u]] = (int32_t)u;
  }
}

static int ensure_lowsrc(Ctx &c) {
  if (c.lowsrc_ready) return PZG_OK;
  if (c.nbu >= ((int64_t)1 << 31)) {
    set_error("block CSR: more than 2^31 upper blocks");
    return PZG_EUNSUP;
  }
  PZG_TRY(dev_alloc(c.d_lowsrc, sizeof(int32_t) * std::max<int64_t>(c.nbu, 1)));
  const int64_t threads = c.nb * 64;
  hipLaunchKernelGGL(k_lowsrc, dim3((unsigned)((threads + TPB - 1) / TPB)), dim3(TPB), 0, c.stream, c.nb,
                     c.d_ubia.as<int64_t>(), c.d_ubja.as<int32_t>(), c.d_rank_bwd.as<int32_t>(),
                     c.d_fbia.as<int64_t>(), c.d_lowsrc.as<int32_t>());
  PZG_HIP(hipGetLastError());
  c.lowsrc_ready = true;
  return PZG_OK;
}

int launch_expand_full(Ctx &c) {
  const int bs = c.bs;
  PZG_TRY(dev_alloc(c.d_fa, sizeof(double) * c.nbf * bs * bs));
  PZG_TRY(ensure_lowsrc(c));
  const int64_t threads = c.nb * 64;
  const dim3 g((unsigned)((threads + TPB - 1) / TPB));
  if (bs == 3)
    hipLaunchKernelGGL((k_bsr_gather<3, double, false>), g, dim3(TPB), 0, c.stream, c.nb, c.d_ia.as<int64_t>(),
                       c.d_a.as<double>(), c.d_ubia.as<int64_t>(), c.d_fbia.as<int64_t>(),
                       c.d_fbja.as<int32_t>(), c.d_lowsrc.as<int32_t>(), (const double *)nullptr,
                       c.d_fa.as<double>());
  else if (bs == 1)
    hipLaunchKernelGGL((k_bsr_gather<1, double, false>), g, dim3(TPB), 0, c.stream, c.nb, c.d_ia.as<int64_t>(),
                       c.d_a.as<double>(), c.d_ubia.as<int64_t>(), c.d_fbia.as<int64_t>(),
                       c.d_fbja.as<int32_t>(), c.d_lowsrc.as<int32_t>(), (const double *)nullptr,
                       c.d_fa.as<double>());
  else {
    set_error("block CSR: unsupported block size");
    return PZG_EUNSUP;
  }
  PZG_HIP(hipGetLastError());
  PZG_HIP(hipStreamSynchronize(c.stream));
  c.full_valid = true;
  return PZG_OK;
}

// --------------------------------------------------------------------------
// Equation-level full CSR (TPZFYsmpMatrix layout) from the FP64 block CSR: one
// wavefront per block row A; row A*BS+k holds, for every block (A, B) in
// ascending B, the columns B*BS .. B*BS+BS-1.
template <int BS>
__global__ void k_full_csr(int64_t nb, const int64_t *__restrict__ fbia,
                           const int32_t *__restrict__ fbja, const double *__restrict__ fa,
                           int64_t *__restrict__ ia, int64_t *__restrict__ ja,
                           double *__restrict__ a) {
  const int64_t A = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (A >= nb) return;
  constexpr int BB = BS * BS;
  const int64_t b0 = fbia[A], n = fbia[A + 1] - b0;
  if (ia && lane < BS) ia[A * BS + lane] = b0 * BB + lane * n * BS;
  for (int64_t e = lane; e < n * BB; e += 64) {
    const int k = (int)(e / (n * BS));
    const int64_t r = e % (n * BS), i = r / BS;
    const int l = (int)(r % BS);
    const int64_t out = b0 * BB + e;
    if (ja) ja[out] = (int64_t)fbja[b0 + i] * BS + l;
    if (a) a[out] = fa[bsr_pos<BB>(b0, n, i, k * BS + l)];
  }
}

int launch_full_csr(Ctx &c, int64_t *d_ia, int64_t *d_ja, double *d_a) {
  if (!c.full_valid) PZG_TRY(launch_expand_full(c));
  const int64_t threads = c.nb * 64;
  const dim3 g((unsigned)((threads + TPB - 1) / TPB));
  if (c.bs == 3)
    hipLaunchKernelGGL(k_full_csr<3>, g, dim3(TPB), 0, c.stream, c.nb, c.d_fbia.as<int64_t>(),
                       c.d_fbja.as<int32_t>(), c.d_fa.as<double>(), d_ia, d_ja, d_a);
  else if (c.bs == 1)
    hipLaunchKernelGGL(k_full_csr<1>, g, dim3(TPB), 0, c.stream, c.nb, c.d_fbia.as<int64_t>(),
                       c.d_fbja.as<int32_t>(), c.d_fa.as<double>(), d_ia, d_ja, d_a);
  else {
    set_error("full CSR: unsupported block size");
    return PZG_EUNSUP;
  }
  PZG_HIP(hipGetLastError());
  return PZG_OK;
}

// --------------------------------------------------------------------------
// Scaled FP32 block CSR S = D^-1/2 K D^-1/2 (unit diagonal) and the inverses
// of its diagonal blocks.
__global__ void k_dinvsqrt(int64_t n, const double *__restrict__ d, double *__restrict__ s) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) s[i] = 1.0 / sqrt(fabs(d[i]));
}

// Values at bsr_pos, as the FP64 block CSR.
template <int BS>
__global__ void k_bsr_diag(int64_t nb, int64_t nbf, const int64_t *__restrict__ ia,
                           const double *__restrict__ a, const int32_t *__restrict__ nlow,
                           const int64_t *__restrict__ fbia, const double *__restrict__ dis,
                           float *__restrict__ bsr, float *__restrict__ binv) {
  int64_t A = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (A >= nb) return;
  double m[BS][BS];
  for (int k = 0; k < BS; k++)
    for (int l = k; l < BS; l++) {
      double v = a[ia[A * BS + k] + (l - k)] * dis[A * BS + k] * dis[A * BS + l];
      m[k][l] = v;
      m[l][k] = v;
    }
  const int64_t b0 = fbia[A], n = fbia[A + 1] - b0;
  if (bsr)
    for (int k = 0; k < BS; k++)
      for (int l = 0; l < BS; l++) bsr[bsr_pos<BS * BS>(b0, n, nlow[A], k * BS + l)] = (float)m[k][l];
  float *inv = binv + A * BS * BS;
  if (BS == 1) {
    inv[0] = (float)(1.0 / m[0][0]);
  } else {
    // 3x3 inverse by cofactors
    double c00 = m[1][1] * m[2][2] - m[1][2] * m[2][1];
    double c01 = m[1][2] * m[2][0] - m[1][0] * m[2][2];
    double c02 = m[1][0] * m[2][1] - m[1][1] * m[2][0];
    double det = m[0][0] * c00 + m[0][1] * c01 + m[0][2] * c02;
    double id = 1.0 / det;
    double r[3][3];
    r[0][0] = c00 * id;
    r[1][0] = c01 * id;
    r[2][0] = c02 * id;
    r[0][1] = (m[0][2] * m[2][1] - m[0][1] * m[2][2]) * id;
    r[1][1] = (m[0][0] * m[2][2] - m[0][2] * m[2][0]) * id;
    r[2][1] = (m[0][1] * m[2][0] - m[0][0] * m[2][1]) * id;
    r[0][2] = (m[0][1] * m[1][2] - m[0][2] * m[1][1]) * id;
    r[1][2] = (m[0][2] * m[1][0] - m[0][0] * m[1][2]) * id;
    r[2][2] = (m[0][0] * m[1][1] - m[0][1] * m[1][0]) * id;
    for (int k = 0; k < 3; k++)
      for (int l = 0; l < 3; l++) inv[k * 3 + l] = (float)r[k][l];
  }
}


int launch_build_bsr(Ctx &c, bool values) {
  const int bs = c.bs;
  if (values) PZG_TRY(dev_alloc(c.d_bsr, sizeof(float) * c.nbf * bs * bs));
  float *bsr = values ? c.d_bsr.as<float>() : nullptr;
  PZG_TRY(dev_alloc(c.d_bjinv, sizeof(float) * c.nb * bs * bs));
  PZG_TRY(dev_alloc(c.d_dinvsqrt, sizeof(double) * c.neq));
  if (values) PZG_TRY(ensure_lowsrc(c));
  DevBuf nl;
  PZG_TRY(dev_alloc(nl, sizeof(int32_t) * c.nb));
  hipLaunchKernelGGL(k_nlow, dim3((c.nb + TPB - 1) / TPB), dim3(TPB), 0, c.stream, c.nb,
                     c.d_fbia.as<int64_t>(), c.d_ubia.as<int64_t>(), nl.as<int32_t>());
  hipLaunchKernelGGL(k_dinvsqrt, dim3((c.neq + TPB - 1) / TPB), dim3(TPB), 0, c.stream, c.neq,
                     c.d_diag.as<double>(), c.d_dinvsqrt.as<double>());
  int64_t threads = c.nb * 64;
  if (bs == 3) {
    hipLaunchKernelGGL(k_bsr_diag<3>, dim3((c.nb + TPB - 1) / TPB), dim3(TPB), 0, c.stream, c.nb,
                       c.nbf, c.d_ia.as<int64_t>(), c.d_a.as<double>(), nl.as<int32_t>(),
                       c.d_fbia.as<int64_t>(), c.d_dinvsqrt.as<double>(), bsr, c.d_bjinv.as<float>());
    if (values)
      hipLaunchKernelGGL((k_bsr_gather<3, float, true>), dim3((threads + TPB - 1) / TPB), dim3(TPB), 0, c.stream,
                         c.nb, c.d_ia.as<int64_t>(), c.d_a.as<double>(), c.d_ubia.as<int64_t>(),
                         c.d_fbia.as<int64_t>(), c.d_fbja.as<int32_t>(), c.d_lowsrc.as<int32_t>(),
                         c.d_dinvsqrt.as<double>(), c.d_bsr.as<float>());
  } else if (bs == 1) {
    hipLaunchKernelGGL(k_bsr_diag<1>, dim3((c.nb + TPB - 1) / TPB), dim3(TPB), 0, c.stream, c.nb,
                       c.nbf, c.d_ia.as<int64_t>(), c.d_a.as<double>(), nl.as<int32_t>(),
                       c.d_fbia.as<int64_t>(), c.d_dinvsqrt.as<double>(), bsr, c.d_bjinv.as<float>());
    if (values)
      hipLaunchKernelGGL((k_bsr_gather<1, float, true>), dim3((threads + TPB - 1) / TPB), dim3(TPB), 0, c.stream,
                         c.nb, c.d_ia.as<int64_t>(), c.d_a.as<double>(), c.d_ubia.as<int64_t>(),
                         c.d_fbia.as<int64_t>(), c.d_fbja.as<int32_t>(), c.d_lowsrc.as<int32_t>(),
                         c.d_dinvsqrt.as<double>(), c.d_bsr.as<float>());
  } else {
    dev_free(nl);
    set_error("block CSR: unsupported block size");
    return PZG_EUNSUP;
  }
  PZG_HIP(hipGetLastError());
  PZG_HIP(hipStreamSynchronize(c.stream));
  dev_free(nl);
  c.scal_valid = true;
  if (values) c.bsr_valid = true;
  c.mf_vals_valid = false;  // D^-1/2 recomputed
  c.mg_ready = false;       // multigrid levels are built from K and D
  c.pen_ready = false;      // so is the penalty block of the learned path
  return PZG_OK;
}

// --------------------------------------------------------------------------
// AddKel of caller-given element matrices (one workgroup per element).
__global__ void k_addkel(int ne, const int64_t *__restrict__ dest, const double *__restrict__ ek,
                         const int32_t *__restrict__ list, const int64_t *__restrict__ ia,
                         const int64_t *__restrict__ ja, double *__restrict__ a) {
  int64_t e = list[blockIdx.x];
  for (int k = threadIdx.x; k < ne * ne; k += blockDim.x) {
    int i = k / ne, j = k % ne;
    int64_t ip = dest[e * ne + i], jp = dest[e * ne + j];
    if (jp < ip) continue;
    double v = ek[(e * ne + i) * (int64_t)ne + j];
    if (fabs(v) < 1.e-12) continue;
    int64_t l = ia[ip], h = ia[ip + 1] - 1;
    while (l <= h) {
      int64_t m = (l + h) >> 1;
      if (ja[m] == jp) { a[m] += v; break; }
      if (ja[m] < jp) l = m + 1;
      else h = m - 1;
    }
  }
}

int launch_addkel_batch(Ctx &c, int64_t nel, int ne, const int64_t *d_dest, const double *d_ek,
                        const int32_t *d_list, int64_t nlist) {
  (void)nel;
  if (!nlist) return PZG_OK;
  hipLaunchKernelGGL(k_addkel, dim3(nlist), dim3(TPB), 0, c.stream, ne, d_dest, d_ek, d_list,
                     c.d_ia.as<int64_t>(), c.d_ja.as<int64_t>(), c.d_a.as<double>());
  PZG_HIP(hipGetLastError());
  return PZG_OK;
}

}  // namespace pzg
