// pzg_matfree.hip -- FP32 matrix-free application of the scaled operator
// S = D^-1/2 K D^-1/2 for the inner iterations of the mixed solver.
//
// K is the sum of the element matrices (TPZStructMatrixOR::Serial_Assemble,
// StrMatrix/pzstrmatrixor.cpp:104-365, with CalcStiff of
// Mesh/pzinterpolationspace.cpp:403-474).  Instead of streaming the stored
// matrix, every application recomputes K_e u_e at the element's quadrature
// points from the corner coordinates, in FP32:
//   volume:   y_e = sum_q w_q |det J_q| B_q^T C_e B_q u_e, evaluated as
//             gradient -> flux/stress -> weighted back-projection, with the
//             material laws of TPZMatPoisson / TPZDarcyFlow
//             (Material/Poisson/TPZMatPoisson.cpp:19-42,
//             Material/DarcyFlow/TPZDarcyFlow.cpp:19-46) and the CODE3 form of
//             TPZElasticity3D (Material/Elasticity/TPZElasticity3D.cpp:269-372):
//             sigma_kk = C3 e_kk - C2 (sum of the other e_jj), sigma_kd =
//             C1 (du_k/dx_d + du_d/dx_k);
//   boundary: Dirichlet penalty big * (phi_i, phi_j) per component
//             (TPZMatPoisson.cpp:45-121, TPZElasticity3D.cpp:616-773), whose
//             element matrices depend on geometry only and are stored.
// Element outputs go to a per-(element, shape) buffer; a second kernel sums
// them per block row in a fixed order (no atomics: the result is
// deterministic) and fuses the dot products of the flexible CG step.
//
// The operator is an FP32 approximation of S of the same order as the stored
// FP32 block CSR; the FP64 true residual of the refinement loop (CSR, FP64)
// still decides convergence.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "pzg_device.h"
#include "pzg_internal.h"

namespace pzg {

namespace {

constexpr int MF_TPB = 256;
// k_mf_hex27_mfma: sixteen wavefronts per workgroup share one copy of the
// tables in LDS (4 waves/SIMD; fewer copies staged per launch)
constexpr int MFK_TPB = 1024;

struct MfTab {
  float w[MAXQ];
  float4 dphi[MAXQ][MAXSH];  // reference derivatives (d/dxi, d/deta, d/dzeta, 0)
  float4 cd[MAXQ][8];        // corner (geometric) derivatives
};

struct MfArgs {
  int64_t nvol;
  int nq;
  float nu;
  // elasticity: C1 = E*lam1, C2 = E*lam2, C3 = E*lam3 (TPZElasticity3D.h:183-188)
  float lam1, lam2, lam3;
  int zblk;             // block whose pt entries are zero (pt holds one extra block)
  const MfTab *tab;
  const int32_t *blk;   // [nvol][NSH] active block of each shape (-1: none)
  const float *xe;      // [nvol][NC*3] corner coordinates
  const float *coef;    // [nvol] E (elasticity), permeability or scale
  const float *geo;     // [nvol][10] affine elements: J^-1 (row m: d xi_m/d x), |det J|; else [9] = -1
  const float *pt;      // input, already multiplied by D^-1/2
  const int32_t *pos;   // [nvol][NSH] position of (element, shape) in its row's output list
  float *ye;            // [position][NS] element outputs K_e pt_e, grouped by block row
};

__device__ inline float4 f4(float a, float b, float c) { return make_float4(a, b, c, 0.f); }

// LPE lanes per element (a power of two >= max(nq, NSH)); 256/LPE elements
// per workgroup; grid-stride loop over element groups.
template <int DIM, int NC, int NSH, int NS, int LPE>
__global__ __launch_bounds__(MF_TPB) void k_mf_vol(MfArgs a, const SolveScalars *sc) {
  if (sc && sc->done) return;
  constexpr int EPB = MF_TPB / LPE;
  __shared__ float4 s_dphi[LPE][NSH];
  __shared__ float4 s_cd[LPE][NC];
  __shared__ float s_w[LPE];
  __shared__ float4 s_u[EPB][NSH];
  __shared__ float s_x[EPB][NC * 3];
  __shared__ float4 s_P[EPB][LPE][NS];
  const int tid = threadIdx.x, el = tid / LPE, l = tid % LPE;
  const int nq = a.nq;
  for (int t = tid; t < nq * NSH; t += MF_TPB) s_dphi[t / NSH][t % NSH] = a.tab->dphi[t / NSH][t % NSH];
  for (int t = tid; t < nq * NC; t += MF_TPB) s_cd[t / NC][t % NC] = a.tab->cd[t / NC][t % NC];
  if (tid < nq) s_w[tid] = a.tab->w[tid];
  for (int64_t g0 = (int64_t)blockIdx.x * EPB; g0 < a.nvol; g0 += (int64_t)gridDim.x * EPB) {
    const int64_t e = g0 + el;
    const bool live = e < a.nvol;
    int32_t b = -1;
    if (live && l < NSH) {
      b = a.blk[e * NSH + l];
      float u[3] = {0.f, 0.f, 0.f};
      if (b >= 0)
        for (int k = 0; k < NS; k++) u[k] = a.pt[(int64_t)b * NS + k];
      s_u[el][l] = f4(u[0], u[1], u[2]);
    }
    if (live)
      for (int t = l; t < NC * 3; t += LPE) s_x[el][t] = a.xe[e * (NC * 3) + t];
    __syncthreads();
    if (live && l < nq) {
      // J[i][j] = dx_i / dxi_j (TPZGeoEl::Jacobian, Mesh/pzgeoel.cpp:1168-1355)
      float J[3][3] = {{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}};
#pragma unroll
      for (int c = 0; c < NC; c++) {
        const float4 d = s_cd[l][c];
        const float dd[3] = {d.x, d.y, d.z};
#pragma unroll
        for (int i = 0; i < 3; i++) {
          const float x = s_x[el][3 * c + i];
#pragma unroll
          for (int j = 0; j < DIM; j++) J[i][j] += x * dd[j];
        }
      }
      // reference gradient gx[k][m] = d u_k / d xi_m
      float gx[NS][3];
#pragma unroll
      for (int k = 0; k < NS; k++) gx[k][0] = gx[k][1] = gx[k][2] = 0.f;
#pragma unroll 9
      for (int s = 0; s < NSH; s++) {
        const float4 d = s_dphi[l][s];
        const float4 u4 = s_u[el][s];
        const float u[3] = {u4.x, u4.y, u4.z};
#pragma unroll
        for (int k = 0; k < NS; k++) {
          gx[k][0] += u[k] * d.x;
          gx[k][1] += u[k] * d.y;
          if (DIM == 3) gx[k][2] += u[k] * d.z;
        }
      }
      const float coef = a.coef[e];
      float P[NS][3];
      if constexpr (DIM == 3) {
        const float det = -J[0][2] * J[1][1] * J[2][0] + J[0][1] * J[1][2] * J[2][0] +
                          J[0][2] * J[1][0] * J[2][1] - J[0][0] * J[1][2] * J[2][1] -
                          J[0][1] * J[1][0] * J[2][2] + J[0][0] * J[1][1] * J[2][2];
        const float id = 1.f / det;
        float inv[3][3];  // inv[m][d] = d xi_m / d x_d
        inv[0][0] = (-J[1][2] * J[2][1] + J[1][1] * J[2][2]) * id;
        inv[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
        inv[0][2] = (-J[0][2] * J[1][1] + J[0][1] * J[1][2]) * id;
        inv[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) * id;
        inv[1][1] = (-J[0][2] * J[2][0] + J[0][0] * J[2][2]) * id;
        inv[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
        inv[2][0] = (-J[1][1] * J[2][0] + J[1][0] * J[2][1]) * id;
        inv[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
        inv[2][2] = (-J[0][1] * J[1][0] + J[0][0] * J[1][1]) * id;
        const float wdet = s_w[l] * fabsf(det);
        float g[NS][3];  // physical gradient
#pragma unroll
        for (int k = 0; k < NS; k++)
#pragma unroll
          for (int d = 0; d < 3; d++) g[k][d] = gx[k][0] * inv[0][d] + gx[k][1] * inv[1][d] + gx[k][2] * inv[2][d];
        float S[NS][3];
        if constexpr (NS == 3) {
          const float nu = a.nu;
          const float C1 = coef / (2.f + 2.f * nu);
          const float den = -1.f + nu + 2.f * nu * nu;
          const float C2 = coef * nu / den;
          const float C3 = coef * (nu - 1.f) / den;
          const float tr = g[0][0] + g[1][1] + g[2][2];
#pragma unroll
          for (int k = 0; k < 3; k++)
#pragma unroll
            for (int d = 0; d < 3; d++)
              S[k][d] = d == k ? C3 * g[k][k] - C2 * (tr - g[k][k]) : C1 * (g[k][d] + g[d][k]);
        } else {
#pragma unroll
          for (int d = 0; d < 3; d++) S[0][d] = coef * g[0][d];
        }
#pragma unroll
        for (int k = 0; k < NS; k++)
#pragma unroll
          for (int m = 0; m < 3; m++)
            P[k][m] = wdet * (S[k][0] * inv[m][0] + S[k][1] * inv[m][1] + S[k][2] * inv[m][2]);
      } else {
        // 2D element in 3D space: grad(a).grad(b) = da^T (J^T J)^-1 db
        float G00 = 0.f, G01 = 0.f, G11 = 0.f;
#pragma unroll
        for (int i = 0; i < 3; i++) {
          G00 += J[i][0] * J[i][0];
          G01 += J[i][0] * J[i][1];
          G11 += J[i][1] * J[i][1];
        }
        const float detG = G00 * G11 - G01 * G01;
        const float wdet = s_w[l] * sqrtf(detG);
        const float f = wdet * coef / detG;
#pragma unroll
        for (int k = 0; k < NS; k++) {
          P[k][0] = f * (G11 * gx[k][0] - G01 * gx[k][1]);
          P[k][1] = f * (-G01 * gx[k][0] + G00 * gx[k][1]);
          P[k][2] = 0.f;
        }
      }
#pragma unroll
      for (int k = 0; k < NS; k++) s_P[el][l][k] = f4(P[k][0], P[k][1], P[k][2]);
    }
    __syncthreads();
    if (live && l < NSH) {
      float y[NS];
#pragma unroll
      for (int k = 0; k < NS; k++) y[k] = 0.f;
#pragma unroll 9
      for (int q = 0; q < nq; q++) {
        const float4 d = s_dphi[q][l];
#pragma unroll
        for (int k = 0; k < NS; k++) {
          const float4 P = s_P[el][q][k];
          y[k] += d.x * P.x + d.y * P.y + (DIM == 3 ? d.z * P.z : 0.f);
        }
      }
      const int32_t t = a.pos[e * NSH + l];
      if (t >= 0)
#pragma unroll
        for (int k = 0; k < NS; k++) a.ye[(int64_t)t * NS + k] = y[k];
    }
    __syncthreads();
  }
}

// Hexahedron p=2 (27 shapes, 27 points) on the matrix cores.  The two
// contractions of the element product are GEMMs with the constant shape
// derivative tables as the left operand:
//   phase 1: G_m[q][(e,k)] = sum_s dphi_m(q,s) u_e[s][k]      (m = d/dxi_m)
//   phase 3: y[s][(e,k)]  = sum_(m,q) dphi_m(q,s) P[q][(e,k)][m]
// as v_mfma_f32_16x16x4f32 tiles (FP32 in, FP32 accumulate): rows = points
// or shapes (27 -> two 16-row tiles), columns = (element, component) pairs of
// 16/NS elements per wavefront, k = shapes (7 steps of 4) or (m, point)
// (21 steps of 4).  The point-wise part (Jacobian, inverse, flux/stress) runs
// on the vector ALUs between the two, through LDS.
// Operand layouts of the 16x16x4 instruction (wave64): A: lane l holds
// A[l%16][l/16]; B: lane l holds B[l/16][l%16]; D: lane l, component r holds
// D[4*(l/16)+r][l%16].
#ifndef PZG_MF_MFMA
#define PZG_MF_MFMA 1
#endif
#ifndef PZG_MF_XCD
#define PZG_MF_XCD 1
#endif

typedef float mf_f4 __attribute__((ext_vector_type(4)));

// s_g and s_x below belong to one wavefront each: between its phases a wave
// only needs its own LDS stores ordered before its loads (no workgroup barrier)
__device__ inline void mf_wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

struct MfMfmaTab {
  float a1[3][2][7][64];  // phase 1 A operands per (m, row tile, k step, lane)
  float a3[2][21][64];    // phase 3 A operands per (row tile, k step, lane)
};

template <int NS>
__global__ __launch_bounds__(MFK_TPB) void k_mf_hex27_mfma(MfArgs a, const MfMfmaTab *__restrict__ mt,
                                                         const SolveScalars *sc) {
  if (sc && sc->done) return;
  constexpr int EPT = 16 / NS;  // elements per wavefront tile
  constexpr int NCOL = EPT * NS;
  constexpr int NW = MFK_TPB / 64;
  constexpr int NG = NS * 3;
  __shared__ float4 s_cd[27][8];
  __shared__ float s_w[27];
  __shared__ float s_g[NW][EPT][27][NG];
  // J^-1 (row-major), |det J| ([9] < 0: not affine), coefficient [10]
  __shared__ float s_x[NW][EPT][11];
  __shared__ __align__(16) MfMfmaTab s_mt;
  const int tid = threadIdx.x, wv = tid >> 6, l = tid & 63, lr = l >> 4;
  for (int t = tid; t < 27 * 8; t += MFK_TPB) s_cd[t / 8][t % 8] = a.tab->cd[t / 8][t % 8];
  if (tid < 27) s_w[tid] = a.tab->w[tid];
  {
    // every load of the thread issued before its first store
    const float4 *src = reinterpret_cast<const float4 *>(mt);
    float4 *dst = reinterpret_cast<float4 *>(&s_mt);
    constexpr int N4 = (int)(sizeof(MfMfmaTab) / 16), NPT = (N4 + MFK_TPB - 1) / MFK_TPB;
    float4 v[NPT];
#pragma unroll
    for (int u = 0; u < NPT; u++) {
      const int t = tid + MFK_TPB * u;
      v[u] = src[t < N4 ? t : 0];
    }
#pragma unroll
    for (int u = 0; u < NPT; u++) {
      const int t = tid + MFK_TPB * u;
      if (t < N4) dst[t] = v[u];
    }
  }
  __syncthreads();
  const float(&A1)[3][2][7][64] = s_mt.a1;
  const float(&A3)[2][21][64] = s_mt.a3;
  const int j = l & 15, ee = j / NS, kc = j % NS;  // column -> (element, component)
  const bool colv = j < NCOL;
#if PZG_MF_XCD
  int64_t elo, ehi;
  int xi, xg;
  xcd_slab(a.nvol, elo, ehi, xi, xg);
  const int64_t gfirst = elo + (int64_t)xi * (NW * EPT), gstep = (int64_t)xg * (NW * EPT);
  const int64_t eend = ehi;
#else
  const int64_t gfirst = (int64_t)blockIdx.x * (NW * EPT), gstep = (int64_t)gridDim.x * (NW * EPT);
  const int64_t eend = a.nvol;
#endif
  // B operands of phase 1 for the group starting at g: u[s][(e,k)], s = 4t + lr
  // element, shape and output offsets fit in 32 bits (checked on the host).
  // Branch-free: the seven block indices are loaded together (clamped
  // addresses), then the seven inputs, absent ones from the zero block, so
  // that each batch is one round trip to memory.
  auto gather = [&](int64_t g, float (&u)[7]) {
    const int64_t e = g + (int64_t)wv * EPT + ee;
    const bool ok = colv && e < eend;
    const int eo = ok ? (int)e * 27 : 0;
    int32_t bb[7];
#pragma unroll
    for (int t = 0; t < 7; t++) {
      const int s = 4 * t + lr;
      bb[t] = a.blk[eo + (s < 27 ? s : 26)];
    }
#pragma unroll
    for (int t = 0; t < 7; t++) {
      const int s = 4 * t + lr;
      const int b = (ok && s < 27 && bb[t] >= 0) ? bb[t] : a.zblk;
      u[t] = a.pt[b * NS + kc];
    }
  };
  // per-element geometry and coefficient of a group: entry t = l + 64*h
  // (t < EPT*11) is entry t % 11 of element t / 11 (absent elements read
  // element 0)
  constexpr int XG = (EPT * 11 + 63) / 64;
  auto geo_load = [&](int64_t g, float (&xv)[XG]) {
#pragma unroll
    for (int h = 0; h < XG; h++) {
      const int t = l + 64 * h, el = t / 11, k = t % 11;
      const int64_t e = g + (int64_t)wv * EPT + el;
      const int ec = (t < EPT * 11 && e < eend) ? (int)e : 0;
      xv[h] = k < 10 ? a.geo[ec * 10 + k] : a.coef[ec];
    }
  };
  float ubn[7], xgn[XG];
  if (gfirst < eend) {
    gather(gfirst, ubn);
    geo_load(gfirst, xgn);
  }
  for (int64_t g0 = gfirst; g0 < eend; g0 += gstep) {
    const int64_t eb = g0 + (int64_t)wv * EPT;
#pragma unroll
    for (int h = 0; h < XG; h++) {
      const int t = l + 64 * h, el = t / 11;
      if (t < EPT * 11) s_x[wv][el][t % 11] = eb + el < eend ? xgn[h] : 0.f;
    }
    const int64_t e = eb + ee;
    const bool ev = colv && e < eend;
    float ub[7];
#pragma unroll
    for (int t = 0; t < 7; t++) ub[t] = ubn[t];
    // the next group's gathers are in flight while this group computes
    if (g0 + gstep < eend) {
      gather(g0 + gstep, ubn);
      geo_load(g0 + gstep, xgn);
    }
    // output positions (raw; validity applied at the stores)
    int32_t opos[2][4];
    const int epo = ev ? (int)e * 27 : 0;
#pragma unroll
    for (int r = 0; r < 2; r++)
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int sh = 16 * r + 4 * lr + i;
        opos[r][i] = a.pos[epo + (sh < 27 ? sh : 26)];
      }
    mf_f4 acc[3][2];
#pragma unroll
    for (int m = 0; m < 3; m++)
#pragma unroll
      for (int r = 0; r < 2; r++) {
        mf_f4 c = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int t = 0; t < 7; t++) c = __builtin_amdgcn_mfma_f32_16x16x4f32(A1[m][r][t][l], ub[t], c, 0, 0, 0);
        acc[m][r] = c;
      }
    if (colv) {
#pragma unroll
      for (int m = 0; m < 3; m++)
#pragma unroll
        for (int r = 0; r < 2; r++)
#pragma unroll
          for (int i = 0; i < 4; i++) {
            const int q = 16 * r + 4 * lr + i;
            if (q < 27) s_g[wv][ee][q][kc * 3 + m] = acc[m][r][i];
          }
    }
    mf_wave_sync();
    // point-wise: Jacobian, physical gradient, flux/stress, weighted back-map
    for (int pi = l; pi < EPT * 27; pi += 64) {
      const int el = pi / 27, q = pi % 27;
      if (eb + el >= eend) continue;
      const float *xg = s_x[wv][el];
      float inv[3][3], adet;
      if (xg[9] >= 0.f) {  // affine element: constant J^-1 and |det J|
#pragma unroll
        for (int m = 0; m < 3; m++)
#pragma unroll
          for (int d = 0; d < 3; d++) inv[m][d] = xg[3 * m + d];
        adet = xg[9];
      } else {
        const float *xc = a.xe + (eb + el) * 24;
        float J[3][3] = {{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}};
#pragma unroll
        for (int c = 0; c < 8; c++) {
          const float4 d = s_cd[q][c];
#pragma unroll
          for (int i = 0; i < 3; i++) {
            const float x = xc[3 * c + i];
            J[i][0] += x * d.x;
            J[i][1] += x * d.y;
            J[i][2] += x * d.z;
          }
        }
        const float det = -J[0][2] * J[1][1] * J[2][0] + J[0][1] * J[1][2] * J[2][0] +
                          J[0][2] * J[1][0] * J[2][1] - J[0][0] * J[1][2] * J[2][1] -
                          J[0][1] * J[1][0] * J[2][2] + J[0][0] * J[1][1] * J[2][2];
        const float id = 1.f / det;
        inv[0][0] = (-J[1][2] * J[2][1] + J[1][1] * J[2][2]) * id;
        inv[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
        inv[0][2] = (-J[0][2] * J[1][1] + J[0][1] * J[1][2]) * id;
        inv[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) * id;
        inv[1][1] = (-J[0][2] * J[2][0] + J[0][0] * J[2][2]) * id;
        inv[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
        inv[2][0] = (-J[1][1] * J[2][0] + J[1][0] * J[2][1]) * id;
        inv[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
        inv[2][2] = (-J[0][1] * J[1][0] + J[0][0] * J[1][1]) * id;
        adet = fabsf(det);
      }
      const float wdet = s_w[q] * adet;
      const float coef = xg[10];
      float *gp = s_g[wv][el][q];
      float g[NS][3];
#pragma unroll
      for (int k = 0; k < NS; k++)
#pragma unroll
        for (int d = 0; d < 3; d++)
          g[k][d] = gp[k * 3 + 0] * inv[0][d] + gp[k * 3 + 1] * inv[1][d] + gp[k * 3 + 2] * inv[2][d];
      float S[NS][3];
      if constexpr (NS == 3) {
        const float C1 = coef * a.lam1, C2 = coef * a.lam2, C3 = coef * a.lam3;
        const float tr = g[0][0] + g[1][1] + g[2][2];
#pragma unroll
        for (int k = 0; k < 3; k++)
#pragma unroll
          for (int d = 0; d < 3; d++)
            S[k][d] = d == k ? C3 * g[k][k] - C2 * (tr - g[k][k]) : C1 * (g[k][d] + g[d][k]);
      } else {
#pragma unroll
        for (int d = 0; d < 3; d++) S[0][d] = coef * g[0][d];
      }
#pragma unroll
      for (int k = 0; k < NS; k++)
#pragma unroll
        for (int m = 0; m < 3; m++)
          gp[k * 3 + m] = wdet * (S[k][0] * inv[m][0] + S[k][1] * inv[m][1] + S[k][2] * inv[m][2]);
    }
    mf_wave_sync();
    mf_f4 y[2];
#pragma unroll
    for (int r = 0; r < 2; r++) y[r] = mf_f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int t = 0; t < 21; t++) {
      const int kk = 4 * t + lr, m = kk / 28, q = kk % 28;
      const float bv = (colv && q < 27) ? s_g[wv][ee][q][kc * 3 + m] : 0.f;
#pragma unroll
      for (int r = 0; r < 2; r++) y[r] = __builtin_amdgcn_mfma_f32_16x16x4f32(A3[r][t][l], bv, y[r], 0, 0, 0);
    }
    if (ev) {
#pragma unroll
      for (int r = 0; r < 2; r++)
#pragma unroll
        for (int i = 0; i < 4; i++) {
          const int s = 16 * r + 4 * lr + i;
          if (s < 27 && opos[r][i] >= 0) a.ye[opos[r][i] * NS + kc] = y[r][i];  // ev: enclosing if
        }
    }
    mf_wave_sync();
  }
}

// Dirichlet boundary elements: y_i,k = big * sum_j M_ij u_j,k with the stored
// FP32 matrices big*M; one thread per (element, shape).
template <int NSHB, int NS>
__global__ __launch_bounds__(MF_TPB) void k_mf_bc(int64_t nbcd, const int32_t *__restrict__ bblk,
                                                  const float *__restrict__ bm,
                                                  const float *__restrict__ pt,
                                                  const int32_t *__restrict__ bpos,
                                                  float *__restrict__ ye, const SolveScalars *sc) {
  if (sc && sc->done) return;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nbcd * NSHB) return;
  const int64_t e = t / NSHB;
  const int i = (int)(t % NSHB);
  float y[NS];
  for (int k = 0; k < NS; k++) y[k] = 0.f;
  for (int j = 0; j < NSHB; j++) {
    const int32_t bj = bblk[e * NSHB + j];
    if (bj < 0) continue;
    const float m = bm[(e * NSHB + i) * NSHB + j];
    for (int k = 0; k < NS; k++) y[k] += m * pt[(int64_t)bj * NS + k];
  }
  const int32_t o = bpos[t];
  if (o >= 0)
    for (int k = 0; k < NS; k++) ye[(int64_t)o * NS + k] = y[k];
}

// q[A] = sum of the element outputs of block row A (fixed order); with
// partials: slots 0/1 <- per-block sums of p.q and p.r (flexible CG alpha).
template <int NS>
__global__ __launch_bounds__(MF_TPB) void k_mf_gather(int64_t nb, const int64_t *__restrict__ ptr,
                                                      const float *__restrict__ ye,
                                                      const float *__restrict__ dis,
                                                      const float *__restrict__ p,
                                                      const float *__restrict__ r,
                                                      float *__restrict__ q, const SolveScalars *sc,
                                                      double *partials) {
  if (sc && sc->done) return;
  double pq = 0., pr = 0.;
#if PZG_MF_XCD
  int64_t lo, hi;
  int xi, xg;
  xcd_slab(nb, lo, hi, xi, xg);
  for (int64_t A = lo + (int64_t)xi * MF_TPB + threadIdx.x; A < hi; A += (int64_t)xg * MF_TPB) {
#else
  for (int64_t A = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; A < nb;
       A += (int64_t)gridDim.x * blockDim.x) {
#endif
    // every load of the row is issued before the first use: the vectors, then
    // the row's first MF_GE entries (clamped addresses, masked sums in entry
    // order), entries beyond MF_GE in a loop
    constexpr int MF_GE = 8;
    const int64_t t0 = ptr[A], n = ptr[A + 1] - t0;
    // without partials r may be absent: read p in its place (unused)
    const float *rs = partials ? r : p;
    float dv[NS], pv[NS], rv[NS];
#pragma unroll
    for (int k = 0; k < NS; k++) {
      const int64_t i = A * NS + k;
      dv[k] = dis[i];
      pv[k] = p[i];
      rv[k] = rs[i];
    }
    float v[MF_GE][NS];
#pragma unroll
    for (int j = 0; j < MF_GE; j++) {
      const int64_t t = t0 + (j < n ? j : 0);
#pragma unroll
      for (int k = 0; k < NS; k++) v[j][k] = ye[t * NS + k];
    }
    // keep every use below the loads (one wait for the whole batch)
    __builtin_amdgcn_sched_barrier(0);
    float acc[NS];
#pragma unroll
    for (int k = 0; k < NS; k++) acc[k] = 0.f;
#pragma unroll
    for (int j = 0; j < MF_GE; j++)
      if (j < n)
#pragma unroll
        for (int k = 0; k < NS; k++) acc[k] += v[j][k];
    for (int64_t t = t0 + MF_GE; t < t0 + n; t++)
      for (int k = 0; k < NS; k++) acc[k] += ye[t * NS + k];
#pragma unroll
    for (int k = 0; k < NS; k++) {
      const float qi = acc[k] * dv[k];
      q[A * NS + k] = qi;
      if (partials) {
        pq += (double)qi * pv[k];
        pr += (double)pv[k] * rv[k];
      }
    }
  }
  if (partials) {
    double vv[2] = {pq, pr};
    block_partials<2>(vv, partials);
  }
}

__global__ void k_mf_scale(int64_t n, const float *__restrict__ p, const float *__restrict__ dis,
                           float *__restrict__ pt, const SolveScalars *sc) {
  if (sc && sc->done) return;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) pt[i] = p[i] * dis[i];
}

__global__ void k_d2f(int64_t n, const double *__restrict__ x, float *__restrict__ y) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = (float)x[i];
}

int pow2_at_least(int v) {
  int p = 4;
  while (p < v) p <<= 1;
  return p;
}

template <int DIM, int NC, int NSH, int NS>
int launch_vol_lpe(Ctx &c, const MfArgs &a, const SolveScalars *sc) {
  const int lpe = pow2_at_least(std::max(a.nq, NSH));
  auto go = [&](auto kern, int lpe_) {
    int64_t groups = (a.nvol + (MF_TPB / lpe_) - 1) / (MF_TPB / lpe_);
    int grid = (int)std::min<int64_t>(std::max<int64_t>(groups, 1), 4096);
    hipLaunchKernelGGL(kern, dim3(grid), dim3(MF_TPB), 0, c.stream, a, sc);
    return hipGetLastError() == hipSuccess ? PZG_OK : PZG_EHIP;
  };
  if constexpr (NSH <= 4) {
    if (lpe <= 4) return go(k_mf_vol<DIM, NC, NSH, NS, 4>, 4);
  }
  if constexpr (NSH <= 8) {
    if (lpe <= 8) return go(k_mf_vol<DIM, NC, NSH, NS, 8>, 8);
  }
  if constexpr (NSH <= 16) {
    if (lpe <= 16) return go(k_mf_vol<DIM, NC, NSH, NS, 16>, 16);
  }
  if (lpe <= 32) return go(k_mf_vol<DIM, NC, NSH, NS, 32>, 32);
  set_error("matrix-free operator: too many quadrature points");
  return PZG_EUNSUP;
}

}  // namespace

// Geometry-dependent set-up: element block maps, corner coordinates, float
// tables, Dirichlet boundary element matrices, row incidence lists.
int mf_setup(Ctx &c) {
  if (c.mf_ready) return PZG_OK;
  if (!c.has_mesh || !c.has_struct || c.external_csr) {
    set_error("matrix-free operator: needs a mesh-built system");
    return PZG_ESTATE;
  }
  const ElemTables &T = c.tvol;
  const int NSH = T.nsh, NC = c.vol_ncorner, NSD = c.vol_nsides, ns = c.nstate;
  if (T.nq > 32 || NSH > 27) {
    set_error("matrix-free operator: element too large");
    return PZG_EUNSUP;
  }
  const int64_t nvol = c.nvol, nb = c.nb;
  std::vector<int32_t> blk(nvol * NSH);
  std::vector<float> xe(nvol * NC * 3), geo(nvol * 10, 0.f);
  for (int64_t e = 0; e < nvol; e++) {
    for (int s = 0; s < NSH; s++) blk[e * NSH + s] = (int32_t)c.aidx[c.vol_conn[e * NSD + T.shape_side[s]]];
    for (int k = 0; k < NC; k++)
      for (int i = 0; i < 3; i++) xe[(e * NC + k) * 3 + i] = (float)c.xyz[3 * c.vol_nodes[e * NC + k] + i];
    geo[e * 10 + 9] = -1.f;
    if (c.dim == 3 && NC == 8) {
      // affine hexahedron: x_c = x_0 + sum_j b_j(c) (x_{e_j} - x_0), e_j = corners 1, 3, 4
      const double *x[8];
      for (int k = 0; k < 8; k++) x[k] = &c.xyz[3 * c.vol_nodes[e * 8 + k]];
      static const int bits[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                     {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
      static const int axis_corner[3] = {1, 3, 4};
      double J[3][3], scale = 0, dev = 0;
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
          J[i][j] = 0.5 * (x[axis_corner[j]][i] - x[0][i]);
          scale = std::max(scale, std::fabs(J[i][j]));
        }
      for (int k = 0; k < 8; k++)
        for (int i = 0; i < 3; i++) {
          double v = x[0][i];
          for (int j = 0; j < 3; j++) v += 2.0 * J[i][j] * bits[k][j];
          dev = std::max(dev, std::fabs(v - x[k][i]));
        }
      const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                         J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                         J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
      if (dev <= 1e-12 * scale && det != 0.0) {
        double inv[3][3];  // inv[m][d] = d xi_m / d x_d
        inv[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) / det;
        inv[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / det;
        inv[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det;
        inv[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) / det;
        inv[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det;
        inv[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / det;
        inv[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) / det;
        inv[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / det;
        inv[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det;
        for (int m = 0; m < 3; m++)
          for (int d = 0; d < 3; d++) geo[e * 10 + 3 * m + d] = (float)inv[m][d];
        geo[e * 10 + 9] = (float)std::fabs(det);
      }
    }
  }
  // Dirichlet boundary elements and their penalty matrices (FP64, then FP32)
  const ElemTables &B = c.tbc;
  const int NSHB = B.nsh, NCB = c.bc_ncorner, NSDB = c.bc_nsides, dimb = c.dim - 1;
  std::vector<int64_t> bcd;
  for (int64_t e = 0; e < c.nbc; e++)
    if (c.bc_type[c.bc_id[e]] == PZG_BC_DIRICHLET) bcd.push_back(e);
  const int64_t nbcd = (int64_t)bcd.size();
  std::vector<int32_t> bblk(nbcd * NSHB);
  std::vector<float> bm(nbcd * NSHB * NSHB);
  for (int64_t t = 0; t < nbcd; t++) {
    const int64_t e = bcd[t];
    for (int s = 0; s < NSHB; s++) bblk[t * NSHB + s] = (int32_t)c.aidx[c.bc_conn[e * NSDB + B.shape_side[s]]];
    double wdet[MAXQ];
    for (int q = 0; q < B.nq; q++) {
      double gx[3][2] = {{0, 0}, {0, 0}, {0, 0}};
      for (int k = 0; k < NCB; k++)
        for (int i = 0; i < 3; i++)
          for (int j = 0; j < dimb; j++) gx[i][j] += c.xyz[3 * c.bc_nodes[e * NCB + k] + i] * B.cdphi[q][j][k];
      double det;
      if (dimb == 1) {
        det = std::sqrt(gx[0][0] * gx[0][0] + gx[1][0] * gx[1][0] + gx[2][0] * gx[2][0]);
      } else {
        double n1 = 0, d12 = 0, n2 = 0;
        for (int i = 0; i < 3; i++) { n1 += gx[i][0] * gx[i][0]; d12 += gx[i][0] * gx[i][1]; }
        n1 = std::sqrt(n1);
        for (int i = 0; i < 3; i++) {
          double v2 = gx[i][1] - d12 * (gx[i][0] / n1) / n1;
          n2 += v2 * v2;
        }
        det = n1 * std::sqrt(n2);
      }
      wdet[q] = B.w[q] * std::fabs(det);
    }
    for (int i = 0; i < NSHB; i++)
      for (int j = 0; j < NSHB; j++) {
        double s = 0;
        for (int q = 0; q < B.nq; q++) s += B.phi[q][i] * B.phi[q][j] * wdet[q];
        bm[(t * NSHB + i) * NSHB + j] = (float)(c.big * s);
      }
  }
  // row incidence lists: volume slots e*NSH+s, then boundary slots
  const int64_t slot0 = nvol * NSH, nslots = slot0 + nbcd * NSHB;
  if (nslots >= ((int64_t)1 << 31)) {
    set_error("matrix-free operator: too many element slots");
    return PZG_EUNSUP;
  }
  std::vector<int64_t> ptr(nb + 1, 0);
  for (int32_t b : blk) if (b >= 0) ptr[b + 1]++;
  for (int32_t b : bblk) if (b >= 0) ptr[b + 1]++;
  for (int64_t i = 0; i < nb; i++) ptr[i + 1] += ptr[i];
  // each (element, shape) writes its output at its place in the row's list:
  // volume elements in element order, then boundary elements (fixed order)
  if (ptr[nb] >= ((int64_t)1 << 31)) {
    set_error("matrix-free operator: too many element contributions");
    return PZG_EUNSUP;
  }
  std::vector<int32_t> pos(slot0, -1), bpos(nbcd * NSHB, -1);
  std::vector<int64_t> fill(ptr.begin(), ptr.end() - 1);
  for (int64_t s = 0; s < slot0; s++) if (blk[s] >= 0) pos[s] = (int32_t)fill[blk[s]]++;
  for (int64_t s = 0; s < nbcd * NSHB; s++) if (bblk[s] >= 0) bpos[s] = (int32_t)fill[bblk[s]]++;
  MfTab tab;
  std::memset(&tab, 0, sizeof(tab));
  for (int q = 0; q < T.nq; q++) {
    tab.w[q] = (float)T.w[q];
    for (int s = 0; s < NSH; s++)
      tab.dphi[q][s] = make_float4((float)T.dphi[q][0][s], (float)T.dphi[q][1][s],
                                   c.dim == 3 ? (float)T.dphi[q][2][s] : 0.f, 0.f);
    for (int k = 0; k < NC; k++)
      tab.cd[q][k] = make_float4((float)T.cdphi[q][0][k], (float)T.cdphi[q][1][k],
                                 c.dim == 3 ? (float)T.cdphi[q][2][k] : 0.f, 0.f);
  }
  std::vector<MfMfmaTab> mtab(1);
  std::memset(mtab.data(), 0, sizeof(MfMfmaTab));
  if (c.dim == 3 && NSH == 27 && T.nq == 27) {
    MfMfmaTab &M = mtab[0];
    for (int l = 0; l < 64; l++) {
      for (int m = 0; m < 3; m++)
        for (int r = 0; r < 2; r++)
          for (int t = 0; t < 7; t++) {
            const int q = 16 * r + l % 16, sh = 4 * t + l / 16;
            M.a1[m][r][t][l] = (q < 27 && sh < 27) ? (float)T.dphi[q][m][sh] : 0.f;
          }
      for (int r = 0; r < 2; r++)
        for (int t = 0; t < 21; t++) {
          const int sh = 16 * r + l % 16, kk = 4 * t + l / 16, m = kk / 28, q = kk % 28;
          M.a3[r][t][l] = (sh < 27 && q < 27) ? (float)T.dphi[q][m][sh] : 0.f;
        }
    }
  }
  hipStream_t s = c.stream;
  PZG_TRY(dev_upload(c.d_mf_mfma, mtab.data(), 1, s));
  PZG_TRY(dev_upload(c.d_mf_blk, blk.data(), blk.size(), s));
  PZG_TRY(dev_upload(c.d_mf_xe, xe.data(), xe.size(), s));
  PZG_TRY(dev_upload(c.d_mf_geo, geo.data(), geo.size(), s));
  // pt: nb blocks + one zero block read for absent shapes
  PZG_TRY(dev_alloc(c.d_mf_pt, sizeof(float) * (c.nb + 1) * ns));
  PZG_HIP(hipMemsetAsync(c.d_mf_pt.p, 0, sizeof(float) * (c.nb + 1) * ns, s));
  PZG_TRY(dev_upload(c.d_mf_bblk, bblk.data(), bblk.size(), s));
  PZG_TRY(dev_upload(c.d_mf_bm, bm.data(), bm.size(), s));
  PZG_TRY(dev_upload(c.d_mf_ptr, ptr.data(), ptr.size(), s));
  PZG_TRY(dev_upload(c.d_mf_pos, pos.data(), pos.size(), s));
  PZG_TRY(dev_upload(c.d_mf_bpos, bpos.data(), bpos.size(), s));
  PZG_TRY(dev_upload(c.d_mf_tab, &tab, 1, s));
  PZG_TRY(dev_alloc(c.d_mf_ye, sizeof(float) * std::max<int64_t>(ptr[nb], 1) * ns));
  PZG_TRY(dev_alloc(c.d_mf_coef, sizeof(float) * std::max<int64_t>(nvol, 1)));
  PZG_TRY(dev_alloc(c.d_mf_dis, sizeof(float) * std::max<int64_t>(c.neq, 1)));
  PZG_HIP(hipStreamSynchronize(s));
  c.mf_nbcd = nbcd;
  c.mf_nout = ptr[nb];
  c.mf_ready = true;
  c.mf_vals_valid = false;
  return PZG_OK;
}

// Value-dependent set-up (after each assembly): coefficients and D^-1/2 in FP32.
int mf_prepare(Ctx &c) {
  PZG_TRY(mf_setup(c));
  // D^-1/2 and the block-Jacobi inverses of the current assembly (a rebuild
  // clears mf_vals_valid, so the coefficients and D^-1/2 below follow it)
  if (!c.scal_valid) PZG_TRY(launch_build_bsr(c, false));
  if (c.mf_vals_valid) return PZG_OK;
  std::vector<float> coef(c.nvol);
  const double dflt = c.material == PZG_MAT_ELAST3D ? c.E : c.scale;
  for (int64_t e = 0; e < c.nvol; e++) coef[e] = (float)(c.vol_coef.empty() ? dflt : c.vol_coef[e]);
  if (c.nvol)
    PZG_HIP(hipMemcpyAsync(c.d_mf_coef.p, coef.data(), coef.size() * sizeof(float),
                           hipMemcpyHostToDevice, c.stream));
  if (c.neq) {
    hipLaunchKernelGGL(k_d2f, dim3((c.neq + MF_TPB - 1) / MF_TPB), dim3(MF_TPB), 0, c.stream, c.neq,
                       c.d_dinvsqrt.as<double>(), c.d_mf_dis.as<float>());
    PZG_HIP(hipGetLastError());
  }
  PZG_HIP(hipStreamSynchronize(c.stream));
  c.mf_vals_valid = true;
  return PZG_OK;
}

// q = S p.  With partials: per-block sums of p.q (slot 0) and p.r (slot 1);
// *gather_grid receives the number of blocks that wrote partials.
int mf_scale(Ctx &c, const float *p, const SolveScalars *sc) {
  if (c.neq > 0) {
    hipLaunchKernelGGL(k_mf_scale, dim3((unsigned)((c.neq + MF_TPB - 1) / MF_TPB)), dim3(MF_TPB), 0,
                       c.stream, c.neq, p, c.d_mf_dis.as<float>(), c.d_mf_pt.as<float>(), sc);
    PZG_HIP(hipGetLastError());
  }
  return PZG_OK;
}

// Volume element kernel only (K_e pt_e into the row lists).
int mf_launch_vol(Ctx &c, const SolveScalars *sc) {
  MfArgs a;
  a.nvol = c.nvol;
  a.nq = c.tvol.nq;
  a.nu = (float)c.nu;
  {
    const double nu = c.nu, den = -1. + nu + 2. * nu * nu;
    a.lam1 = (float)(1. / (2. + 2. * nu));
    a.lam2 = (float)(nu / den);
    a.lam3 = (float)((nu - 1.) / den);
  }
  a.zblk = (int)c.nb;
  a.tab = c.d_mf_tab.as<MfTab>();
  a.blk = c.d_mf_blk.as<int32_t>();
  a.xe = c.d_mf_xe.as<float>();
  a.coef = c.d_mf_coef.as<float>();
  a.geo = c.d_mf_geo.as<float>();
  a.pt = c.d_mf_pt.as<float>();
  a.pos = c.d_mf_pos.as<int32_t>();
  a.ye = c.d_mf_ye.as<float>();
  const int nsh = c.tvol.nsh;
  if (c.nvol > 0) {
    const bool mfma = PZG_MF_MFMA && c.dim == 3 && nsh == 27 && c.tvol.nq == 27;
    if (mfma && (c.nvol * 27 >= INT32_MAX || c.mf_nout * c.nstate >= INT32_MAX ||
                 c.nb * c.nstate >= INT32_MAX)) {
      set_error("matrix-free operator: mesh too large for 32-bit element offsets");
      return PZG_EUNSUP;
    }
    if (mfma) {
      // persistent grid: as many workgroups as fit on the device at once
      const int64_t per_wg = (MFK_TPB / 64) * (16 / c.nstate);  // elements per workgroup pass
      int64_t groups = (c.nvol + per_wg - 1) / per_wg;
      if (c.mf_resident <= 0) {
        int per_cu = 0, ncu = 0;
        PZG_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c.device));
        if (c.nstate == 3)
          PZG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_mf_hex27_mfma<3>, MFK_TPB, 0));
        else
          PZG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_mf_hex27_mfma<1>, MFK_TPB, 0));
        c.mf_resident = std::max(1, per_cu) * std::max(1, ncu);
      }
      dim3 grid((unsigned)std::min<int64_t>(std::max<int64_t>(groups, 1), c.mf_resident));
      const MfMfmaTab *mt = c.d_mf_mfma.as<MfMfmaTab>();
      if (c.nstate == 3) hipLaunchKernelGGL(k_mf_hex27_mfma<3>, grid, dim3(MFK_TPB), 0, c.stream, a, mt, sc);
      else hipLaunchKernelGGL(k_mf_hex27_mfma<1>, grid, dim3(MFK_TPB), 0, c.stream, a, mt, sc);
      PZG_HIP(hipGetLastError());
    } else if (c.dim == 3 && c.nstate == 3) {
      if (nsh == 27) PZG_TRY((launch_vol_lpe<3, 8, 27, 3>(c, a, sc)));
      else PZG_TRY((launch_vol_lpe<3, 8, 8, 3>(c, a, sc)));
    } else if (c.dim == 3 && c.nstate == 1) {
      if (nsh == 27) PZG_TRY((launch_vol_lpe<3, 8, 27, 1>(c, a, sc)));
      else PZG_TRY((launch_vol_lpe<3, 8, 8, 1>(c, a, sc)));
    } else if (c.dim == 2 && c.nstate == 1) {
      if (nsh == 9) PZG_TRY((launch_vol_lpe<2, 4, 9, 1>(c, a, sc)));
      else PZG_TRY((launch_vol_lpe<2, 4, 4, 1>(c, a, sc)));
    } else {
      set_error("matrix-free operator: unsupported dimension/state combination");
      return PZG_EUNSUP;
    }
  }
  return PZG_OK;
}

int mf_apply(Ctx &c, const float *p, const float *r, float *q, const SolveScalars *sc,
             double *partials, int *gather_grid, bool scaled_ready) {
  const float *dis = c.d_mf_dis.as<float>();
  if (!scaled_ready) PZG_TRY(mf_scale(c, p, sc));
  PZG_TRY(mf_launch_vol(c, sc));
  const int nsh = c.tvol.nsh;
  const int64_t nbcd = c.mf_nbcd;
  const int nshb = c.tbc.nsh;
  if (nbcd > 0) {
    const int64_t th = nbcd * nshb;
    dim3 g((unsigned)((th + MF_TPB - 1) / MF_TPB));
    const int32_t *bb = c.d_mf_bblk.as<int32_t>();
    const float *bm = c.d_mf_bm.as<float>();
    float *ye = c.d_mf_ye.as<float>();
    const float *pt = c.d_mf_pt.as<float>();
    const int32_t *bpos = c.d_mf_bpos.as<int32_t>();
    if (c.nstate == 3 && nshb == 9)
      hipLaunchKernelGGL((k_mf_bc<9, 3>), g, dim3(MF_TPB), 0, c.stream, nbcd, bb, bm, pt, bpos, ye, sc);
    else if (c.nstate == 3 && nshb == 4)
      hipLaunchKernelGGL((k_mf_bc<4, 3>), g, dim3(MF_TPB), 0, c.stream, nbcd, bb, bm, pt, bpos, ye, sc);
    else if (c.nstate == 1 && nshb == 9)
      hipLaunchKernelGGL((k_mf_bc<9, 1>), g, dim3(MF_TPB), 0, c.stream, nbcd, bb, bm, pt, bpos, ye, sc);
    else if (c.nstate == 1 && nshb == 4)
      hipLaunchKernelGGL((k_mf_bc<4, 1>), g, dim3(MF_TPB), 0, c.stream, nbcd, bb, bm, pt, bpos, ye, sc);
    else if (c.nstate == 1 && nshb == 3)
      hipLaunchKernelGGL((k_mf_bc<3, 1>), g, dim3(MF_TPB), 0, c.stream, nbcd, bb, bm, pt, bpos, ye, sc);
    else if (c.nstate == 1 && nshb == 2)
      hipLaunchKernelGGL((k_mf_bc<2, 1>), g, dim3(MF_TPB), 0, c.stream, nbcd, bb, bm, pt, bpos, ye, sc);
    else {
      set_error("matrix-free operator: unsupported boundary element");
      return PZG_EUNSUP;
    }
    PZG_HIP(hipGetLastError());
  }
  int64_t gg = (c.nb + MF_TPB - 1) / MF_TPB;
  const int grid = (int)std::max<int64_t>(1, std::min<int64_t>(gg, MAXBLK));
  if (c.nstate == 3)
    hipLaunchKernelGGL(k_mf_gather<3>, dim3(grid), dim3(MF_TPB), 0, c.stream, c.nb,
                       c.d_mf_ptr.as<int64_t>(), c.d_mf_ye.as<float>(), dis, p, r,
                       q, sc, partials);
  else
    hipLaunchKernelGGL(k_mf_gather<1>, dim3(grid), dim3(MF_TPB), 0, c.stream, c.nb,
                       c.d_mf_ptr.as<int64_t>(), c.d_mf_ye.as<float>(), dis, p, r,
                       q, sc, partials);
  PZG_HIP(hipGetLastError());
  if (gather_grid) *gather_grid = grid;
  return PZG_OK;
}

}  // namespace pzg
