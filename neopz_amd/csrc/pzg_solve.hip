// pzg_solve.hip -- gfx950 kernels and drivers of the solve phase.
//
//  PCG64 (plain path, reference semantics).  CG<> of
//  Solvers/LinearSolvers/cg.h:29-132 with the EJacobi TPZStepSolver
//  (z = r / diag, Matrix/pzmatrix.cpp:642-667), x0 = 0, stop on the recursive
//  ||r||/||b|| <= tol.  The reference's Jacobi step also runs a residual SpMV
//  whose result it discards; it has no effect on the iterates and is not run.
//  Matrix: full (both triangles) block CSR of nstate x nstate blocks, FP64
//  values, int32 block columns, expanded once per assembly from the
//  reference upper CSR.
//
//  MIXED.  FP32 flexible CG (Notay's FCG(1)) on S = D^-1/2 K D^-1/2 stored as
//  FP32 block CSR, preconditioned by the inverse diagonal blocks, with FP64
//  reliable updates: whenever the FP32 recursive residual has dropped by a
//  factor delta, the FP32 correction is folded into the FP64 solution
//  (u += D^-1/2 x), the true residual f - K u is recomputed in FP64 and
//  replaces the recursive one, and the iteration continues along the same
//  search direction.  This is iterative refinement (the shape of IR<>,
//  Solvers/LinearSolvers/ir.h:23-66) without restarting the Krylov
//  iteration; the stop test is the true FP64 residual.
//
//  All scalars live on the device; reductions are two-level and
//  deterministic (per-block partials, the last block sums them in a fixed
//  order), so iterations need no host round trip.  The host only polls the
//  'done' flag once per batch of iterations; a batch can be captured once in
//  a hipGraph and replayed.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>

#include "pzg_device.h"
#include "pzg_internal.h"

namespace pzg {

constexpr int TPB = RED_TPB;

static int grid_for(int64_t n, int per_block) {
  int64_t g = (n + per_block - 1) / per_block;
  if (g > MAXBLK) g = MAXBLK;
  if (g < 1) g = 1;
  return (int)g;
}

// ----------------------------------------------------------------------------
// Block CSR SpMV, shared by the FP64 path (unscaled K) and the FP32 path
// (scaled S).  Every block row is handled by BSR_L lanes; each lane takes
// BSR_U blocks per pass and issues all their loads before the products, so a
// wavefront keeps 64/BSR_L rows x BSR_U blocks of loads in flight.  Values at
// bsr_pos: lane l of a row reads block l of each chunk of BSR_L blocks, the
// chunk component-major, so every value load of the row's lanes is one
// contiguous segment.
#ifndef PZG_BSR_L
#define PZG_BSR_L 16
#endif
#ifndef PZG_BSR_U
#define PZG_BSR_U 4
#endif
constexpr int BSR_L = PZG_BSR_L;
constexpr int BSR_U = PZG_BSR_U;
static_assert(BSR_L == BSR_CHUNK, "lanes per block row must match the value chunks");

template <class T, int BS>
__device__ inline void bsr_row(int64_t A, const int64_t *__restrict__ bia,
                               const int32_t *__restrict__ bja, const T *__restrict__ v,
                               const T *__restrict__ x, int sub, T (&out)[BS]) {
  constexpr int BB = BS * BS;
  T acc[BS];
  for (int k = 0; k < BS; k++) acc[k] = T(0);
  const int64_t b0 = bia[A], n = bia[A + 1] - b0;
  for (int64_t base = sub; base < n; base += BSR_L * BSR_U) {
    T m[BSR_U][BB];
    int64_t col[BSR_U];
#pragma unroll
    for (int u = 0; u < BSR_U; u++) {
      const int64_t i = base + BSR_L * u;  // block of the row; i % BSR_L == sub
      if (i < n) {
        col[u] = bja[b0 + i];
        const int64_t j = i - sub, mc = n - j < BSR_L ? n - j : BSR_L;
        const T *vp = v + (b0 + j) * BB + sub;
#pragma unroll
        for (int c = 0; c < BB; c++) m[u][c] = vp[c * mc];
      } else {
        col[u] = 0;
#pragma unroll
        for (int c = 0; c < BB; c++) m[u][c] = T(0);
      }
    }
    T xv[BSR_U][BS];
#pragma unroll
    for (int u = 0; u < BSR_U; u++)
#pragma unroll
      for (int l = 0; l < BS; l++) xv[u][l] = x[col[u] * BS + l];
#pragma unroll
    for (int u = 0; u < BSR_U; u++)
#pragma unroll
      for (int k = 0; k < BS; k++)
#pragma unroll
        for (int l = 0; l < BS; l++) acc[k] += m[u][k * BS + l] * xv[u][l];
  }
#pragma unroll
  for (int k = 0; k < BS; k++) {
    T s = acc[k];
    for (int o = BSR_L / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, BSR_L);
    out[k] = s;
  }
}

// lane 'sub' (< BS) of a row group picks component sub of out
template <class T, int BS>
__device__ inline T pick(const T (&out)[BS], int sub) {
  T o = out[0];
#pragma unroll
  for (int k = 1; k < BS; k++)
    if (sub == k) o = out[k];
  return o;
}

// y = A x over block rows; the rows of a workgroup are a contiguous slice
// (xcd_block_range) so that gathers of x stay in one XCD's L2.
template <class T, int BS>
__global__ __launch_bounds__(TPB) void k_spmv_bsr(int64_t nb, const int64_t *__restrict__ bia,
                                                  const int32_t *__restrict__ bja,
                                                  const T *__restrict__ v,
                                                  const T *__restrict__ x, T *__restrict__ y) {
  const int sub = threadIdx.x % BSR_L;
  int64_t lo, hi;
  int xi, xg;
  constexpr int RPB = TPB / BSR_L;  // rows per workgroup tile
#if PZG_XCD_INTERLEAVE
  xcd_slab(nb, lo, hi, xi, xg);
  const int64_t first = lo + (int64_t)xi * RPB, step = (int64_t)xg * RPB;
#else
  xcd_block_range(nb, lo, hi);
  xi = xg = 0;
  const int64_t first = lo, step = RPB;
#endif
  for (int64_t A = first + threadIdx.x / BSR_L; A < hi; A += step) {
    T out[BS];
    bsr_row<T, BS>(A, bia, bja, v, x, sub, out);
    if (sub < BS) y[A * BS + sub] = pick<T, BS>(out, sub);
  }
}

// PCG64: q = K p and per-block partial sums of p.q (slot 0)
template <int BS>
__global__ __launch_bounds__(TPB) void k_spmv64_pq(int64_t nb, const int64_t *__restrict__ bia,
                                                   const int32_t *__restrict__ bja,
                                                   const double *__restrict__ v,
                                                   const double *__restrict__ p,
                                                   double *__restrict__ q, SolveScalars *sc,
                                                   double *partials) {
  if (sc->done) return;
  const int sub = threadIdx.x % BSR_L;
  int64_t lo, hi;
  int xi, xg;
  constexpr int RPB = TPB / BSR_L;  // rows per workgroup tile
#if PZG_XCD_INTERLEAVE
  xcd_slab(nb, lo, hi, xi, xg);
  const int64_t first = lo + (int64_t)xi * RPB, step = (int64_t)xg * RPB;
#else
  xcd_block_range(nb, lo, hi);
  xi = xg = 0;
  const int64_t first = lo, step = RPB;
#endif
  double acc = 0.;
  for (int64_t A = first + threadIdx.x / BSR_L; A < hi; A += step) {
    double out[BS];
    bsr_row<double, BS>(A, bia, bja, v, p, sub, out);
    if (sub < BS) {
      const int64_t i = A * BS + sub;
      const double qi = pick<double, BS>(out, sub);
      q[i] = qi;
      acc += qi * p[i];
    }
  }
  double vv[1] = {acc};
  block_partials<1>(vv, partials);
}

static int bsr_grid(int64_t nb) {
  int64_t g = (nb + TPB / BSR_L - 1) / (TPB / BSR_L);
  if (g > MAXBLK) g = MAXBLK;
  return (int)std::max<int64_t>(g, 1);
}

// alpha = rho / (p.q)   (one block)
__global__ void k_fin_alpha64(SolveScalars *sc, const double *partials, int nblk) {
  if (sc->done) return;
  double pq = sum_partials(partials, 0, nblk);
  if (threadIdx.x == 0) {
    sc->pq = pq;
    sc->alpha = sc->rho / pq;
  }
}

// p = z + beta p  (p = z on the first iteration)
__global__ void k_p_update64(int64_t n, const double *__restrict__ z, double *__restrict__ p,
                             const SolveScalars *sc) {
  if (sc->done) return;
  const double beta = sc->iter == 0 ? 0. : sc->beta;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t st = (int64_t)gridDim.x * blockDim.x;
  if (sc->iter == 0)
    for (; i < n; i += st) p[i] = z[i];
  else
    for (; i < n; i += st) p[i] = beta * p[i] + z[i];
}

// x += alpha p; r -= alpha q; resid = |r|/|b|; z = r/diag; beta = (r.z)/rho
__global__ __launch_bounds__(TPB) void k_cg_update64(int64_t n, double *__restrict__ x,
                                                     double *__restrict__ r,
                                                     double *__restrict__ z,
                                                     const double *__restrict__ p,
                                                     const double *__restrict__ q,
                                                     const double *__restrict__ diag,
                                                     SolveScalars *sc, double *hist,
                                                     double *partials, unsigned *counter) {
  if (sc->done) return;
  const double alpha = sc->alpha;
  double rr = 0., rz = 0.;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t st = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += st) {
    x[i] += alpha * p[i];
    double ri = r[i] + (-alpha) * q[i];
    r[i] = ri;
    rr += ri * ri;
    if (diag) {  // point Jacobi; the block form follows in k_bj_apply64
      double zi = ri / diag[i];
      z[i] = zi;
      rz += ri * zi;
    }
  }
  double v[2] = {rr, rz};
  block_partials<2>(v, partials);
}

// Connect-block Jacobi (TPZBlockDiagonal with ELU blocks, the preconditioner
// TPZAnalysis::BuildPreconditioner(EBlockJacobi, true) returns,
// Analysis/TPZAnalysis.cpp:1062-1070): z = Binv r per block row, r.z into
// partials slot 1 (same grid as k_cg_update64).
template <int BS>
__global__ __launch_bounds__(TPB) void k_bj_apply64(int64_t nb, const double *__restrict__ binv,
                                                    const double *__restrict__ r, double *__restrict__ z,
                                                    const SolveScalars *sc, double *partials) {
  if (sc && sc->done) return;
  double rz = 0.;
  for (int64_t A = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; A < nb; A += (int64_t)gridDim.x * blockDim.x) {
    double rv[BS];
#pragma unroll
    for (int k = 0; k < BS; k++) rv[k] = r[A * BS + k];
    const double *bi = binv + A * BS * BS;
#pragma unroll
    for (int k = 0; k < BS; k++) {
      double s = 0.;
#pragma unroll
      for (int l = 0; l < BS; l++) s += bi[k * BS + l] * rv[l];
      z[A * BS + k] = s;
      rz += rv[k] * s;
    }
  }
  __shared__ double sm[TPB / 64];
  const double t = block_sum(rz, sm);
  if (threadIdx.x == 0) partials[1 * MAXBLK + blockIdx.x] = t;
}

// rho = r.z from slot 1 after the block apply of the initial residual
__global__ void k_bj_rho64(SolveScalars *sc, const double *partials, int nblk) {
  const double rz = sum_partials(partials, 1, nblk);
  if (threadIdx.x == 0) sc->rho = rz;
}

// inverse of each diagonal block of the FP64 full block CSR (cofactors for
// 3 x 3, the reciprocal for 1 x 1); a singular block sets *fail
template <int BS>
__global__ void k_bj_inv64(int64_t nb, const int64_t *__restrict__ fbia, const int32_t *__restrict__ fbja,
                           const double *__restrict__ fa, double *__restrict__ binv, int *fail) {
  const int64_t A = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (A >= nb) return;
  const int64_t b0 = fbia[A], n = fbia[A + 1] - b0;
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t m = (lo + hi) / 2;
    if (fbja[b0 + m] < A) lo = m + 1;
    else hi = m;
  }
  double a[BS * BS];
#pragma unroll
  for (int k = 0; k < BS * BS; k++) a[k] = fa[bsr_pos<BS * BS>(b0, n, lo, k)];
  static_assert(BS == 1 || BS == 3, "connect blocks of 1 or 3 equations");
  double *o = binv + A * BS * BS;
  if (BS == 1) {
    if (a[0] == 0.) atomicExch(fail, 1);
    o[0] = 1. / a[0];
  } else {
    const double c00 = a[4] * a[8] - a[5] * a[7], c01 = a[5] * a[6] - a[3] * a[8], c02 = a[3] * a[7] - a[4] * a[6];
    const double det = a[0] * c00 + a[1] * c01 + a[2] * c02;
    if (det == 0.) atomicExch(fail, 1);
    const double id = 1. / det;
    o[0] = c00 * id;
    o[1] = (a[2] * a[7] - a[1] * a[8]) * id;
    o[2] = (a[1] * a[5] - a[2] * a[4]) * id;
    o[3] = c01 * id;
    o[4] = (a[0] * a[8] - a[2] * a[6]) * id;
    o[5] = (a[2] * a[3] - a[0] * a[5]) * id;
    o[6] = c02 * id;
    o[7] = (a[1] * a[6] - a[0] * a[7]) * id;
    o[8] = (a[0] * a[4] - a[1] * a[3]) * id;
  }
}

// Connect-block Jacobi on high-order meshes: a connect of several shapes
// is one block of its consecutive equations [pos, pos + m), m <= VBJ_MAX.
// Inverse: one wavefront per block, the block (from the upper CSR, whose
// row i starts with the columns i, i+1, ... of its own connect) and the
// identity in LDS, Gauss-Jordan with lane r holding row r; a missing entry
// or a zero pivot sets *fail.
__global__ __launch_bounds__(256) void k_vbj_inv64(int64_t nblk, const int64_t *__restrict__ bpos,
                                                   const int32_t *__restrict__ bm,
                                                   const int64_t *__restrict__ boff,
                                                   const int64_t *__restrict__ ia,
                                                   const int64_t *__restrict__ ja,
                                                   const double *__restrict__ a, double *__restrict__ binv,
                                                   int *fail) {
  __shared__ double sA[4][VBJ_MAX][VBJ_MAX], sI[4][VBJ_MAX][VBJ_MAX];
  const int w = threadIdx.x >> 6, r = threadIdx.x & 63;
  const int64_t b = (int64_t)blockIdx.x * 4 + w;
  if (b >= nblk) return;  // whole wavefronts
  const int64_t p = bpos[b];
  const int m = bm[b];
  auto sync = [] {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
  };
  if (r < m)
    for (int c = 0; c < m; c++) {
      const int lo = r < c ? r : c, hi = r < c ? c : r;
      const int64_t q = ia[p + lo] + (hi - lo);
      if (q >= ia[p + lo + 1] || ja[q] != p + hi) atomicExch(fail, 1);
      sA[w][r][c] = q < ia[p + lo + 1] ? a[q] : 0.;
      sI[w][r][c] = r == c ? 1. : 0.;
    }
  sync();
  for (int k = 0; k < m; k++) {
    const double piv = sA[w][k][k];
    if (piv == 0.) {
      if (r == 0) atomicExch(fail, 1);
      return;
    }
    sync();
    if (r == k)
      for (int c = 0; c < m; c++) {
        sA[w][k][c] /= piv;
        sI[w][k][c] /= piv;
      }
    sync();
    if (r < m && r != k) {
      const double f = sA[w][r][k];
      for (int c = 0; c < m; c++) {
        sA[w][r][c] -= f * sA[w][k][c];
        sI[w][r][c] -= f * sI[w][k][c];
      }
    }
    sync();
  }
  if (r < m)
    for (int c = 0; c < m; c++) binv[boff[b] + (int64_t)r * m + c] = sI[w][r][c];
}

// z = Binv r on the blocks (one thread per equation), r.z into partials
// slot 1 (same grid as k_cg_update64)
__global__ __launch_bounds__(TPB) void k_vbj_apply64(int64_t n, const int32_t *__restrict__ eqb,
                                                     const int64_t *__restrict__ bpos,
                                                     const int32_t *__restrict__ bm,
                                                     const int64_t *__restrict__ boff,
                                                     const double *__restrict__ binv,
                                                     const double *__restrict__ r, double *__restrict__ z,
                                                     const SolveScalars *sc, double *partials) {
  if (sc && sc->done) return;
  double rz = 0.;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t b = eqb[i];
    const int64_t p = bpos[b];
    const int m = bm[b];
    const double *row = binv + boff[b] + (i - p) * m;
    double s = 0.;
    for (int c = 0; c < m; c++) s += row[c] * r[p + c];
    z[i] = s;
    rz += r[i] * s;
  }
  __shared__ double sm[TPB / 64];
  const double t = block_sum(rz, sm);
  if (threadIdx.x == 0) partials[1 * MAXBLK + blockIdx.x] = t;
}

// The connect blocks of the scaled FP32 operator S = D^-1/2 K D^-1/2:
// S_b^-1 = D^1/2 K_b^-1 D^1/2, from the FP64 inverses (one thread per row)
__global__ void k_vbj_scale32(int64_t n, const int32_t *__restrict__ eqb, const int64_t *__restrict__ bpos,
                              const int32_t *__restrict__ bm, const int64_t *__restrict__ boff,
                              const double *__restrict__ inv64, const double *__restrict__ dis,
                              float *__restrict__ inv32) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t b = eqb[i];
    const int64_t p = bpos[b], o = boff[b] + (i - p) * bm[b];
    for (int c = 0; c < bm[b]; c++) inv32[o + c] = (float)(inv64[o + c] / (dis[i] * dis[p + c]));
  }
}

// FP32 FCG with the connect blocks: z = S_b^-1 r (one thread per equation);
// p = z at a restart; z.q replaces slot 1 of the FCG update's partials
// (launched on the same grid)
__global__ __launch_bounds__(TPB) void k_vbj_apply32(int64_t n, const int32_t *__restrict__ eqb,
                                                     const int64_t *__restrict__ bpos,
                                                     const int32_t *__restrict__ bm,
                                                     const int64_t *__restrict__ boff,
                                                     const float *__restrict__ inv, const float *__restrict__ r,
                                                     float *__restrict__ z, float *__restrict__ p,
                                                     const float *__restrict__ q, const SolveScalars *sc,
                                                     double *partials) {
  if (sc && sc->done) return;
  double zq = 0.;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t b = eqb[i];
    const int64_t s0 = bpos[b];
    const int m = bm[b];
    const float *row = inv + boff[b] + (i - s0) * m;
    float v = 0.f;
    for (int c = 0; c < m; c++) v += row[c] * r[s0 + c];
    z[i] = v;
    if (p) p[i] = v;
    if (q) zq += (double)v * q[i];
  }
  if (!q) return;
  __shared__ double sm[TPB / 64];
  const double t = block_sum(zq, sm);
  if (threadIdx.x == 0) partials[1 * MAXBLK + blockIdx.x] = t;
}

// resid = |r|/|b|; stop tests; beta = (r.z)/rho   (one block)
__global__ void k_fin_cg64(SolveScalars *sc, const double *partials, int nblk, double *hist) {
  if (sc->done) return;
  double rr = sum_partials(partials, 0, nblk);
  double rz = sum_partials(partials, 1, nblk);
  if (threadIdx.x == 0) {
    int64_t it = sc->iter + 1;
    sc->iter = it;
    double resid = sqrt(rr) / sc->normb;
    sc->resid = resid;
    if (hist) hist[it - 1] = resid;
    if (resid <= sc->tol) sc->done = 1;
    else if (it >= sc->maxit) sc->done = 2;
    sc->beta = rz / sc->rho;
    sc->rho_old = sc->rho;
    sc->rho = rz;
  }
}

// x = 0; r = b; z = r/diag; rho = r.z; normb = |b|
__global__ __launch_bounds__(TPB) void k_cg_init64(int64_t n, const double *__restrict__ b,
                                                   double *__restrict__ x, double *__restrict__ r,
                                                   double *__restrict__ z, double *__restrict__ p,
                                                   const double *__restrict__ diag,
                                                   SolveScalars *sc, double *partials,
                                                   unsigned *counter) {
  double bb = 0., rz = 0.;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t st = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += st) {
    double bi = b[i];
    x[i] = 0.;
    r[i] = bi;
    p[i] = 0.;
    bb += bi * bi;
    if (diag) {
      double zi = bi / diag[i];
      z[i] = zi;
      rz += bi * zi;
    }
  }
  double v[2] = {bb, rz}, tot[2];
  if (reduce_last<2>(v, partials, counter, tot) && threadIdx.x == 0) {
    double normb = sqrt(tot[0]);
    double resid = normb == 0. ? 0. : 1.0;
    if (normb == 0.) normb = 1.;
    sc->normb = normb;
    sc->resid = resid;
    sc->iter = 0;
    sc->rho = tot[1];
    sc->beta = 0.;
    sc->done = resid <= sc->tol ? 1 : 0;
  }
}

// Continuation from the current x with its residual r = f - K x already in
// r (CG<> with fromCurrent, Solvers/LinearSolvers/cg.h:29-132): z = r/diag,
// p = 0, rho = r.z, normb = |f|, resid = |r|/|f|.
__global__ __launch_bounds__(TPB) void k_cg_restart64(int64_t n, const double *__restrict__ f,
                                                      const double *__restrict__ r,
                                                      double *__restrict__ z, double *__restrict__ p,
                                                      const double *__restrict__ diag,
                                                      SolveScalars *sc, double *partials,
                                                      unsigned *counter) {
  double ff = 0., rr = 0., rz = 0.;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t st = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += st) {
    const double fi = f[i], ri = r[i], zi = ri / diag[i];
    p[i] = 0.;
    z[i] = zi;
    ff += fi * fi;
    rr += ri * ri;
    rz += ri * zi;
  }
  double v[3] = {ff, rr, rz}, tot[3];
  if (reduce_last<3>(v, partials, counter, tot) && threadIdx.x == 0) {
    double normb = sqrt(tot[0]);
    if (normb == 0.) normb = 1.;
    const double resid = sqrt(tot[1]) / normb;
    sc->normb = normb;
    sc->resid = resid;
    sc->rho = tot[2];
    sc->beta = 0.;
    sc->done = resid <= sc->tol ? 1 : 0;
  }
}

// r = f - K u (FP64 block CSR), |r|^2 and |f|^2
template <int BS>
__global__ __launch_bounds__(TPB) void k_residual64(int64_t nb, const int64_t *__restrict__ bia,
                                                    const int32_t *__restrict__ bja,
                                                    const double *__restrict__ v,
                                                    const double *__restrict__ u,
                                                    const double *__restrict__ f,
                                                    double *__restrict__ r, double *out,
                                                    double *partials, unsigned *counter) {
  const int sub = threadIdx.x % BSR_L;
  int64_t lo, hi;
  int xi, xg;
  constexpr int RPB = TPB / BSR_L;  // rows per workgroup tile
#if PZG_XCD_INTERLEAVE
  xcd_slab(nb, lo, hi, xi, xg);
  const int64_t first = lo + (int64_t)xi * RPB, step = (int64_t)xg * RPB;
#else
  xcd_block_range(nb, lo, hi);
  xi = xg = 0;
  const int64_t first = lo, step = RPB;
#endif
  double rr = 0., ff = 0.;
  for (int64_t A = first + threadIdx.x / BSR_L; A < hi; A += step) {
    double o[BS];
    bsr_row<double, BS>(A, bia, bja, v, u, sub, o);
    if (sub < BS) {
      const int64_t i = A * BS + sub;
      const double ri = f[i] - pick<double, BS>(o, sub);
      if (r) r[i] = ri;
      rr += ri * ri;
      ff += f[i] * f[i];
    }
  }
  double vv[2] = {rr, ff}, tot[2];
  if (reduce_last<2>(vv, partials, counter, tot) && threadIdx.x == 0) {
    out[0] = tot[0];
    out[1] = tot[1];
  }
}

// the residual of u = 0: r = f, out = (|f|^2, |f|^2)
__global__ __launch_bounds__(TPB) void k_residual_zero(int64_t n, const double *__restrict__ f,
                                                       double *__restrict__ r, double *out,
                                                       double *partials, unsigned *counter) {
  double ff = 0.;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double v = f[i];
    r[i] = v;
    ff += v * v;
  }
  double vv[1] = {ff}, tot[1];
  if (reduce_last<1>(vv, partials, counter, tot) && threadIdx.x == 0) {
    out[0] = tot[0];
    out[1] = tot[0];
  }
}

__global__ void k_axpby64(int64_t n, double alpha, const double *__restrict__ x, double beta,
                          const double *__restrict__ y, double *__restrict__ z) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t st = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += st) z[i] = alpha * x[i] + (y ? beta * y[i] : 0.);
}

#define DISPATCH_BS(BSV, CALL)                         \
  switch (BSV) {                                       \
    case 3: { constexpr int BS = 3; CALL; } break;     \
    default: { constexpr int BS = 1; CALL; } break;    \
  }

int launch_spmv64(Ctx &c, const double *x, double *y) {
  DISPATCH_BS(c.bs, hipLaunchKernelGGL((k_spmv_bsr<double, BS>), dim3(bsr_grid(c.nb)), dim3(TPB), 0,
                                       c.stream, c.nb, c.d_fbia.as<int64_t>(),
                                       c.d_fbja.as<int32_t>(), c.d_fa.as<double>(), x, y));
  PZG_HIP(hipGetLastError());
  return PZG_OK;
}

static int ensure_reduce(Ctx &c) {
  if (!c.d_partials.p) {
    PZG_TRY(dev_alloc(c.d_partials, sizeof(double) * NSLOT * MAXBLK));
    PZG_TRY(dev_alloc(c.d_counter, sizeof(unsigned) * 16));
    PZG_HIP(hipMemsetAsync(c.d_counter.p, 0, sizeof(unsigned) * 16, c.stream));
  }
  PZG_TRY(dev_alloc(c.d_scal, sizeof(SolveScalars) * 2));
  PZG_TRY(dev_alloc(c.d_res2, 2 * sizeof(double)));
  return PZG_OK;
}
int ensure_reduce_pub(Ctx &c) { return ensure_reduce(c); }

int launch_multadd_upper(Ctx &c, const double *x, const double *y, double *z, double alpha,
                         double beta) {
  if (!c.full_valid) PZG_TRY(launch_expand_full(c));
  PZG_TRY(launch_spmv64(c, x, z));
  int64_t n = c.neq;
  hipLaunchKernelGGL(k_axpby64, dim3(grid_for(n, TPB)), dim3(TPB), 0, c.stream, n, alpha, z, beta,
                     y, z);
  PZG_HIP(hipGetLastError());
  return PZG_OK;
}

int true_residual(Ctx &c, const double *u, double *rel) {
  if (!c.f_valid || !c.d_f.p || !u) {  // never launch on a missing load or solution
    set_error("true residual: no load or solution of the current system on the device");
    return PZG_ESTATE;
  }
  if (!c.full_valid) PZG_TRY(launch_expand_full(c));
  PZG_TRY(ensure_reduce(c));
  DISPATCH_BS(c.bs, hipLaunchKernelGGL(k_residual64<BS>, dim3(bsr_grid(c.nb)), dim3(TPB), 0,
                                       c.stream, c.nb, c.d_fbia.as<int64_t>(),
                                       c.d_fbja.as<int32_t>(), c.d_fa.as<double>(), u,
                                       c.d_f.as<double>(), nullptr, c.d_res2.as<double>(),
                                       c.d_partials.as<double>(), c.d_counter.as<unsigned>()));
  PZG_HIP(hipGetLastError());
  double h[2];
  PZG_HIP(hipMemcpyAsync(h, c.d_res2.p, sizeof(h), hipMemcpyDeviceToHost, c.stream));
  PZG_HIP(hipStreamSynchronize(c.stream));
  *rel = h[1] > 0 ? std::sqrt(h[0] / h[1]) : std::sqrt(h[0]);
  return PZG_OK;
}

// ---- allocation-free true residual ---------------------------------------
// Fixed device scratch of pzg_true_residual_noalloc, one copy per GPU, held
// until the process ends: the reduction partials (2 x MAXBLK doubles, one
// slot per workgroup of the capped residual grid), the two sums and the
// arrival counter.  32.8 KB whatever the system size, because the residual
// kernel strides over block rows with at most MAXBLK workgroups.
constexpr int NOALLOC_PART = 2 * MAXBLK;           // partials, doubles
constexpr int NOALLOC_OUT = NOALLOC_PART;          // |r|^2, |f|^2
constexpr int NOALLOC_CNT = NOALLOC_PART + 4;      // counter (as unsigned)
constexpr int NOALLOC_WORDS = NOALLOC_PART + 8;
static __device__ __align__(256) double scratch[NOALLOC_WORDS];

constexpr int NOALLOC_MAXDEV = 64;
static std::mutex g_scratch_mu;
static double *g_scratch_ptr[NOALLOC_MAXDEV] = {};

// Looks up this device's address of 'scratch' (the host-side name is only a
// stand-in).  Called by pzg_create with the context's device current.
int noalloc_scratch_setup(int device) {
  if (device < 0 || device >= NOALLOC_MAXDEV) {
    set_error("pzg_create: device index beyond the allocation-free scratch table");
    return PZG_EINVAL;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return PZG_OK;
  void *p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    return PZG_EHIP;
  }
  g_scratch_ptr[device] = static_cast<double *>(p);
  return PZG_OK;
}

__global__ void k_noalloc_clear(double *s) {
  if (threadIdx.x < 2) s[NOALLOC_OUT + threadIdx.x] = 0.;
  if (threadIdx.x == 0) *reinterpret_cast<unsigned *>(s + NOALLOC_CNT) = 0u;
}

// The residual of true_residual on the full block CSR already on the device,
// reducing in the per-device scratch: no allocation of any kind.
int true_residual_noalloc(Ctx &c, double *rel) {
  if (!c.full_valid || !c.d_fbia.p || !c.f_valid || !c.d_f.p) {
    set_error("pzg_true_residual_noalloc: no full block CSR or load of the current system on the "
              "device (call a solve or pzg_true_residual after assembling first)");
    return PZG_ESTATE;
  }
  double *s = (c.device >= 0 && c.device < NOALLOC_MAXDEV) ? g_scratch_ptr[c.device] : nullptr;
  if (!s) {
    set_error("pzg_true_residual_noalloc: scratch of this device not set up");
    return PZG_ESTATE;
  }
  const int grid = bsr_grid(c.nb);
  if (grid > NOALLOC_PART / 2) {
    set_error("pzg_true_residual_noalloc: residual grid of " + std::to_string(grid) +
              " workgroups exceeds the scratch's " + std::to_string(NOALLOC_PART / 2));
    return PZG_EINVAL;
  }
  hipLaunchKernelGGL(k_noalloc_clear, dim3(1), dim3(64), 0, c.stream, s);
  DISPATCH_BS(c.bs, hipLaunchKernelGGL(k_residual64<BS>, dim3(grid), dim3(TPB), 0, c.stream, c.nb,
                                       c.d_fbia.as<int64_t>(), c.d_fbja.as<int32_t>(),
                                       c.d_fa.as<double>(), c.d_u.as<double>(),
                                       c.d_f.as<double>(), nullptr, s + NOALLOC_OUT, s,
                                       reinterpret_cast<unsigned *>(s + NOALLOC_CNT)));
  PZG_HIP(hipGetLastError());
  double h[2];
  PZG_HIP(hipMemcpyAsync(h, s + NOALLOC_OUT, sizeof(h), hipMemcpyDeviceToHost, c.stream));
  PZG_HIP(hipStreamSynchronize(c.stream));
  *rel = h[1] > 0 ? std::sqrt(h[0] / h[1]) : std::sqrt(h[0]);
  return PZG_OK;
}

// Runs 'batch' iterations of the given enqueue function, optionally through a
// captured hipGraph, until the device 'done' flag is set.
template <class F>
static int run_batches(Ctx &c, int batch, bool use_graph, F enqueue_iter, SolveScalars &hs) {
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
  if (use_graph) {
    PZG_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
    for (int k = 0; k < batch; k++) enqueue_iter();
    PZG_HIP(hipStreamEndCapture(c.stream, &graph));
    PZG_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
  }
  for (;;) {
    if (use_graph) PZG_HIP(hipGraphLaunch(exec, c.stream));
    else
      for (int k = 0; k < batch; k++) enqueue_iter();
    PZG_HIP(hipMemcpyAsync(&hs, c.d_scal.p, sizeof(SolveScalars), hipMemcpyDeviceToHost, c.stream));
    PZG_HIP(hipStreamSynchronize(c.stream));
    if (hs.done) break;
  }
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  return PZG_OK;
}

// FP64 Jacobi-PCG iterations (the EJacobi CG<> of solve_pcg64) continued from
// the current solution u, whose true residual f - K u is in d_r: the mixed
// solver's fallback when the FP32 iterations stall.  Runs until the
// recursive residual reaches 0.9 tol, then checks the true residual; at most
// four such restarts.  iters accumulates the iterations; rel receives the
// true relative residual.
static int pcg64_continue(Ctx &c, const pzg_solve_params &prm, int64_t &iters, double &rel) {
  const int64_t n = c.neq;
  for (DevBuf *b : {&c.d_z, &c.d_p, &c.d_q}) PZG_TRY(dev_alloc(*b, n * sizeof(double)));
  SolveScalars *sc = c.d_scal.as<SolveScalars>();
  double *part = c.d_partials.as<double>();
  unsigned *cnt = c.d_counter.as<unsigned>();
  const int gv = grid_for(n, TPB), gs = bsr_grid(c.nb);
  for (int round = 0; round < 4 && rel > prm.tol && iters < prm.max_iter; round++) {
    SolveScalars hs;
    std::memset(&hs, 0, sizeof(hs));
    hs.tol = 0.9 * prm.tol;
    hs.maxit = prm.max_iter;
    hs.iter = iters;
    PZG_HIP(hipMemcpyAsync(c.d_scal.p, &hs, sizeof(hs), hipMemcpyHostToDevice, c.stream));
    hipLaunchKernelGGL(k_cg_restart64, dim3(gv), dim3(TPB), 0, c.stream, n, c.d_f.as<double>(),
                       c.d_r.as<double>(), c.d_z.as<double>(), c.d_p.as<double>(),
                       c.d_diag.as<double>(), sc, part, cnt);
    PZG_HIP(hipGetLastError());
    auto iter = [&]() {
      hipLaunchKernelGGL(k_p_update64, dim3(gv), dim3(TPB), 0, c.stream, n, c.d_z.as<double>(),
                         c.d_p.as<double>(), sc);
      DISPATCH_BS(c.bs, hipLaunchKernelGGL(k_spmv64_pq<BS>, dim3(gs), dim3(TPB), 0, c.stream, c.nb,
                                           c.d_fbia.as<int64_t>(), c.d_fbja.as<int32_t>(),
                                           c.d_fa.as<double>(), c.d_p.as<double>(),
                                           c.d_q.as<double>(), sc, part));
      hipLaunchKernelGGL(k_fin_alpha64, dim3(1), dim3(TPB), 0, c.stream, sc, (const double *)part, gs);
      hipLaunchKernelGGL(k_cg_update64, dim3(gv), dim3(TPB), 0, c.stream, n, c.d_u.as<double>(),
                         c.d_r.as<double>(), c.d_z.as<double>(), c.d_p.as<double>(),
                         c.d_q.as<double>(), c.d_diag.as<double>(), sc, (double *)nullptr, part,
                         cnt);
      hipLaunchKernelGGL(k_fin_cg64, dim3(1), dim3(TPB), 0, c.stream, sc, (const double *)part, gv,
                         (double *)nullptr);
    };
    PZG_HIP(hipMemcpyAsync(&hs, c.d_scal.p, sizeof(hs), hipMemcpyDeviceToHost, c.stream));
    PZG_HIP(hipStreamSynchronize(c.stream));
    if (!hs.done) PZG_TRY(run_batches(c, prm.check_every > 0 ? prm.check_every : 50, prm.use_graph != 0, iter, hs));
    iters = hs.iter;
    // true residual into d_r (the next round restarts from it)
    DISPATCH_BS(c.bs, hipLaunchKernelGGL(k_residual64<BS>, dim3(gs), dim3(TPB), 0, c.stream, c.nb,
                                         c.d_fbia.as<int64_t>(), c.d_fbja.as<int32_t>(),
                                         c.d_fa.as<double>(), c.d_u.as<double>(), c.d_f.as<double>(),
                                         c.d_r.as<double>(), c.d_res2.as<double>(), part, cnt));
    PZG_HIP(hipGetLastError());
    double h[2];
    PZG_HIP(hipMemcpyAsync(h, c.d_res2.p, sizeof(h), hipMemcpyDeviceToHost, c.stream));
    PZG_HIP(hipStreamSynchronize(c.stream));
    rel = h[1] > 0 ? std::sqrt(h[0] / h[1]) : std::sqrt(h[0]);
    if (hs.done == 2) break;
  }
  return PZG_OK;
}

int solve_pcg64(Ctx &c, const pzg_solve_params &prm, pzg_solve_info &info) {
  const int64_t n = c.neq;
  if (!c.full_valid) PZG_TRY(launch_expand_full(c));
  PZG_TRY(ensure_reduce(c));
  for (DevBuf *b : {&c.d_u, &c.d_r, &c.d_z, &c.d_p, &c.d_q}) PZG_TRY(dev_alloc(*b, n * sizeof(double)));
  SolveScalars hs;
  std::memset(&hs, 0, sizeof(hs));
  hs.tol = prm.tol;
  hs.maxit = prm.max_iter;
  PZG_HIP(hipMemcpyAsync(c.d_scal.p, &hs, sizeof(hs), hipMemcpyHostToDevice, c.stream));
  PZG_HIP(hipEventRecord(c.ev0, c.stream));
  SolveScalars *sc = c.d_scal.as<SolveScalars>();
  double *part = c.d_partials.as<double>();
  unsigned *cnt = c.d_counter.as<unsigned>();
  const int gv = grid_for(n, TPB);
  // EJacobi (reference default) or the connect-block Jacobi of
  // BuildPreconditioner(EBlockJacobi): blocks of one connect's nstate
  // equations (the connects of order <= 2 carry one shape)
  const bool block = prm.precond == PZG_PRECOND_BLOCK_JACOBI;
  if (!block && prm.precond != PZG_PRECOND_JACOBI) {
    set_error("pzg_solve: PCG64 takes the point or the connect-block Jacobi preconditioner");
    return PZG_EINVAL;
  }
  // high-order meshes: blocks of the real connects (several shapes each)
  const bool vblock = block && c.hp_simplex;
  if (vblock) {
    if (c.filt) {
      set_error("pzg_solve: connect-block Jacobi of a filtered high-order system not built");
      return PZG_EUNSUP;
    }
    if (c.vbj_nblk == 0) {
      set_error("pzg_solve: connect blocks of the high-order mesh not recorded");
      return PZG_ESTATE;
    }
    PZG_TRY(dev_alloc(c.d_bjinv64, sizeof(double) * std::max<int64_t>(c.vbj_ninv, 1)));
    int *fail = reinterpret_cast<int *>(c.d_counter.as<unsigned>() + 8);
    PZG_HIP(hipMemsetAsync(fail, 0, sizeof(int), c.stream));
    hipLaunchKernelGGL(k_vbj_inv64, dim3((unsigned)((c.vbj_nblk + 3) / 4)), dim3(256), 0, c.stream, c.vbj_nblk,
                       c.d_vbj_pos.as<int64_t>(), c.d_vbj_m.as<int32_t>(), c.d_vbj_off.as<int64_t>(),
                       c.d_ia.as<int64_t>(), c.d_ja.as<int64_t>(), c.d_a.as<double>(), c.d_bjinv64.as<double>(),
                       fail);
    int hf = 0;
    PZG_HIP(hipMemcpyAsync(&hf, fail, sizeof(int), hipMemcpyDeviceToHost, c.stream));
    PZG_HIP(hipStreamSynchronize(c.stream));
    if (hf) {
      set_error("pzg_solve: a connect block of K is singular or not stored whole");
      return PZG_EUNSUP;
    }
  } else if (block) {
    if (c.external_csr) {
      set_error("pzg_solve: connect-block Jacobi needs the mesh's connects");
      return PZG_EUNSUP;
    }
    PZG_TRY(dev_alloc(c.d_bjinv64, sizeof(double) * c.nb * c.bs * c.bs));
    int *fail = reinterpret_cast<int *>(c.d_counter.as<unsigned>() + 8);
    PZG_HIP(hipMemsetAsync(fail, 0, sizeof(int), c.stream));
    DISPATCH_BS(c.bs, hipLaunchKernelGGL(k_bj_inv64<BS>, dim3((unsigned)((c.nb + TPB - 1) / TPB)), dim3(TPB), 0,
                                         c.stream, c.nb, c.d_fbia.as<int64_t>(), c.d_fbja.as<int32_t>(),
                                         c.d_fa.as<double>(), c.d_bjinv64.as<double>(), fail));
    int hf = 0;
    PZG_HIP(hipMemcpyAsync(&hf, fail, sizeof(int), hipMemcpyDeviceToHost, c.stream));
    PZG_HIP(hipStreamSynchronize(c.stream));
    if (hf) {
      set_error("pzg_solve: a connect block of K is singular");
      return PZG_EUNSUP;
    }
  }
  const double *dg = block ? nullptr : c.d_diag.as<double>();
  hipLaunchKernelGGL(k_cg_init64, dim3(gv), dim3(TPB), 0, c.stream, n, c.d_f.as<double>(),
                     c.d_u.as<double>(), c.d_r.as<double>(), c.d_z.as<double>(),
                     c.d_p.as<double>(), dg, sc, part, cnt);
  auto vbj_apply = [&](const SolveScalars *s2) {
    hipLaunchKernelGGL(k_vbj_apply64, dim3(gv), dim3(TPB), 0, c.stream, n, c.d_vbj_eqb.as<int32_t>(),
                       c.d_vbj_pos.as<int64_t>(), c.d_vbj_m.as<int32_t>(), c.d_vbj_off.as<int64_t>(),
                       (const double *)c.d_bjinv64.as<double>(), (const double *)c.d_r.as<double>(),
                       c.d_z.as<double>(), s2, part);
  };
  if (vblock) {
    vbj_apply(nullptr);
    hipLaunchKernelGGL(k_bj_rho64, dim3(1), dim3(TPB), 0, c.stream, sc, (const double *)part, gv);
  } else if (block) {
    DISPATCH_BS(c.bs, hipLaunchKernelGGL(k_bj_apply64<BS>, dim3(gv), dim3(TPB), 0, c.stream, c.nb,
                                         c.d_bjinv64.as<double>(), c.d_r.as<double>(), c.d_z.as<double>(),
                                         (const SolveScalars *)nullptr, part));
    hipLaunchKernelGGL(k_bj_rho64, dim3(1), dim3(TPB), 0, c.stream, sc, (const double *)part, gv);
  }
  PZG_HIP(hipGetLastError());
  const int gs = bsr_grid(c.nb);
  auto iter = [&]() {
    hipLaunchKernelGGL(k_p_update64, dim3(gv), dim3(TPB), 0, c.stream, n, c.d_z.as<double>(),
                       c.d_p.as<double>(), sc);
    DISPATCH_BS(c.bs, hipLaunchKernelGGL(k_spmv64_pq<BS>, dim3(gs), dim3(TPB), 0, c.stream, c.nb,
                                         c.d_fbia.as<int64_t>(), c.d_fbja.as<int32_t>(),
                                         c.d_fa.as<double>(), c.d_p.as<double>(),
                                         c.d_q.as<double>(), sc, part));
    hipLaunchKernelGGL(k_fin_alpha64, dim3(1), dim3(TPB), 0, c.stream, sc, (const double *)part, gs);
    hipLaunchKernelGGL(k_cg_update64, dim3(gv), dim3(TPB), 0, c.stream, n, c.d_u.as<double>(),
                       c.d_r.as<double>(), c.d_z.as<double>(), c.d_p.as<double>(),
                       c.d_q.as<double>(), dg, sc, (double *)nullptr, part,
                       cnt);
    if (vblock)
      vbj_apply(sc);
    else if (block)
      DISPATCH_BS(c.bs, hipLaunchKernelGGL(k_bj_apply64<BS>, dim3(gv), dim3(TPB), 0, c.stream, c.nb,
                                           c.d_bjinv64.as<double>(), c.d_r.as<double>(), c.d_z.as<double>(),
                                           (const SolveScalars *)sc, part));
    hipLaunchKernelGGL(k_fin_cg64, dim3(1), dim3(TPB), 0, c.stream, sc, (const double *)part, gv,
                       (double *)nullptr);
  };
  int batch = prm.check_every > 0 ? prm.check_every : 50;
  PZG_TRY(run_batches(c, batch, prm.use_graph != 0, iter, hs));
  PZG_HIP(hipEventRecord(c.ev1, c.stream));
  PZG_HIP(hipEventSynchronize(c.ev1));
  float ms = 0;
  PZG_HIP(hipEventElapsedTime(&ms, c.ev0, c.ev1));
  info.solve_ms = ms;
  info.iterations = hs.iter;
  info.outer = 1;
  info.achieved = hs.resid;
  info.converged = hs.done == 1;
  double tr = 0;
  PZG_TRY(true_residual(c, c.d_u.as<double>(), &tr));
  info.true_residual = tr;
  return hs.done == 1 ? PZG_OK : PZG_ENOCONV;
}

// ----------------------------------------------------------------------------
// Mixed path.

// q = S p; pq = p.q; pr = p.r; alpha = pr/pq
template <int BS>
__global__ __launch_bounds__(TPB) void k_fcg_spmv(int64_t nb, const int64_t *__restrict__ bia,
                                                  const int32_t *__restrict__ bja,
                                                  const float *__restrict__ v,
                                                  const float *__restrict__ p,
                                                  const float *__restrict__ r,
                                                  float *__restrict__ q, SolveScalars *sc,
                                                  double *partials) {
  if (sc->done) return;
  const int sub = threadIdx.x % BSR_L;
  int64_t lo, hi;
  int xi, xg;
  constexpr int RPB = TPB / BSR_L;  // rows per workgroup tile
#if PZG_XCD_INTERLEAVE
  xcd_slab(nb, lo, hi, xi, xg);
  const int64_t first = lo + (int64_t)xi * RPB, step = (int64_t)xg * RPB;
#else
  xcd_block_range(nb, lo, hi);
  xi = xg = 0;
  const int64_t first = lo, step = RPB;
#endif
  double pq = 0., pr = 0.;
  for (int64_t A = first + threadIdx.x / BSR_L; A < hi; A += step) {
    float out[BS];
    bsr_row<float, BS>(A, bia, bja, v, p, sub, out);
    if (sub < BS) {
      const float o = pick<float, BS>(out, sub);
      const int64_t i = A * BS + sub;
      q[i] = o;
      pq += (double)o * p[i];
      pr += (double)p[i] * r[i];
    }
  }
  double vv[2] = {pq, pr};
  block_partials<2>(vv, partials);
}

// alpha = (p.r)/(p.q)   (one block).  A pending update request (4) raised by
// the previous iteration becomes 5 here, so that the completion kernels of that
// iteration (prolongation, p update), which run on 4, do not run again.
__global__ void k_fin_fcg_alpha(SolveScalars *sc, const double *partials, int nblk) {
  if (sc->done == 4) {
    __syncthreads();
    if (threadIdx.x == 0) sc->done = 5;
    return;
  }
  if (sc->done) return;
  double pq = sum_partials(partials, 0, nblk);
  double pr = sum_partials(partials, 1, nblk);
  if (threadIdx.x == 0) {
    sc->pq = pq;
    sc->pr = pr;
    sc->alpha = pr / pq;
  }
}

// x += alpha p; r -= alpha q; z = Binv r (block Jacobi); rr, zq; beta = -zq/pq
template <int BS>
__global__ __launch_bounds__(TPB) void k_fcg_update(int64_t nb, float *__restrict__ x,
                                                    float *__restrict__ r, float *__restrict__ z,
                                                    const float *__restrict__ p,
                                                    const float *__restrict__ q,
                                                    const float *__restrict__ binv,
                                                    SolveScalars *sc, double *partials,
                                                    unsigned *counter) {
  if (sc->done) return;
  const float alpha = (float)sc->alpha;
  double rr = 0., zq = 0.;
  int64_t A = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t st = (int64_t)gridDim.x * blockDim.x;
  for (; A < nb; A += st) {
    float rv[BS];
    for (int k = 0; k < BS; k++) {
      const int64_t i = A * BS + k;
      x[i] += alpha * p[i];
      rv[k] = r[i] - alpha * q[i];
      r[i] = rv[k];
      rr += (double)rv[k] * rv[k];
    }
    const float *bi = binv + A * BS * BS;
    for (int k = 0; k < BS; k++) {
      float s = 0.f;
      for (int l = 0; l < BS; l++) s += bi[k * BS + l] * rv[l];
      z[A * BS + k] = s;
      zq += (double)s * q[A * BS + k];
    }
  }
  double vv[2] = {rr, zq};
  block_partials<2>(vv, partials);
}

// stop tests on |r|; beta = -(z.q)/(p.q)   (one block; slot 1 may have been
// rewritten by the learned preconditioner's prolongation)
__global__ void k_fin_fcg_beta(SolveScalars *sc, const double *partials, int nblk) {
  if (sc->done) return;
  double rr = sum_partials(partials, 0, nblk);
  double zq = sum_partials(partials, 1, nblk);
  if (threadIdx.x == 0) {
    int64_t it = sc->iter + 1;
    sc->iter = it;
    double rn = sqrt(rr);
    double resid = rn / sc->normb;
    // divergence guard: leave the loop, the host restarts from the true residual
    if (!(resid == resid) || resid > 1e3) sc->done = 3;
    else if (rn < sc->delta * sc->rlast) sc->done = 4;
    else if (it >= sc->maxit) sc->done = 2;
    sc->resid = resid;
    sc->beta = -zq / sc->pq;
  }
}

// runs when the iteration completed normally (done 0) or completed and asked
// for a reliable update (done 4)
// p = z + beta p; with pt: also pt = D^-1/2 p (input of the matrix-free operator)
__global__ void k_p_update32(int64_t n, const float *__restrict__ z, float *__restrict__ p,
                             const SolveScalars *sc, const float *__restrict__ dis = nullptr,
                             float *__restrict__ pt = nullptr) {
  if (sc->done != 0 && sc->done != 4) return;
  const float beta = (float)sc->beta;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t st = (int64_t)gridDim.x * blockDim.x;
  if (pt) {
    for (; i < n; i += st) {
      const float v = z[i] + beta * p[i];
      p[i] = v;
      pt[i] = v * dis[i];
    }
  } else {
    for (; i < n; i += st) p[i] = z[i] + beta * p[i];
  }
}

// b = D^-1/2 r64 (fp32); x = 0; r = b; z = Binv r; p = z; normb = |b|
template <int BS>
__global__ __launch_bounds__(TPB) void k_fcg_init(int64_t nb, const double *__restrict__ r64,
                                                  const double *__restrict__ dis,
                                                  float *__restrict__ x, float *__restrict__ r,
                                                  float *__restrict__ z, float *__restrict__ p,
                                                  const float *__restrict__ binv,
                                                  SolveScalars *sc, double *partials,
                                                  unsigned *counter) {
  double bb = 0.;
  int64_t A = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t st = (int64_t)gridDim.x * blockDim.x;
  for (; A < nb; A += st) {
    float rv[BS];
    for (int k = 0; k < BS; k++) {
      const int64_t i = A * BS + k;
      double bi = r64[i] * dis[i];
      rv[k] = (float)bi;
      r[i] = rv[k];
      x[i] = 0.f;
      bb += bi * bi;
    }
    if (!binv) continue;  // z and p come from the multigrid cycle
    const float *bi = binv + A * BS * BS;
    for (int k = 0; k < BS; k++) {
      float s = 0.f;
      for (int l = 0; l < BS; l++) s += bi[k * BS + l] * rv[l];
      z[A * BS + k] = s;
      p[A * BS + k] = s;
    }
  }
  double vv[1] = {bb}, tot[1];
  if (reduce_last<1>(vv, partials, counter, tot) && threadIdx.x == 0) {
    double nbv = sqrt(tot[0]);
    sc->normb = nbv > 0. ? nbv : 1.;
    sc->rlast = nbv;
    sc->resid = nbv > 0. ? 1. : 0.;
    sc->done = nbv > 0. ? 0 : 1;
  }
}

// u += D^-1/2 y; y = 0
__global__ void k_ir_update(int64_t n, float *__restrict__ y, const double *__restrict__ dis,
                            double *__restrict__ u) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t st = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += st) {
    u[i] += dis[i] * (double)y[i];
    y[i] = 0.f;
  }
}

// reliable update: r = D^-1/2 r64 replaces the recursive residual; the search
// direction p is kept
__global__ __launch_bounds__(TPB) void k_fcg_reset(int64_t n, const double *__restrict__ r64,
                                                   const double *__restrict__ dis,
                                                   float *__restrict__ r, double delta,
                                                   SolveScalars *sc, double *partials,
                                                   unsigned *counter) {
  double bb = 0.;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t st = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += st) {
    float v = (float)(r64[i] * dis[i]);
    r[i] = v;
    bb += (double)v * v;
  }
  double vv[1] = {bb}, tot[1];
  if (reduce_last<1>(vv, partials, counter, tot) && threadIdx.x == 0) {
    sc->rlast = sqrt(tot[0]);
    sc->delta = delta;
    sc->done = 0;
  }
}



int launch_spmv_bsr(Ctx &c, const float *x, float *y) {
  DISPATCH_BS(c.bs, hipLaunchKernelGGL((k_spmv_bsr<float, BS>), dim3(bsr_grid(c.nb)), dim3(TPB), 0,
                                       c.stream, c.nb, c.d_fbia.as<int64_t>(),
                                       c.d_fbja.as<int32_t>(), c.d_bsr.as<float>(), x, y));
  PZG_HIP(hipGetLastError());
  return PZG_OK;
}

constexpr int64_t MG_INNER_CAP = 2000;  // FP32 iterations per refinement step, multigrid paths

template <int BS>
static int mixed_impl(Ctx &c, const pzg_solve_params &prm, pzg_solve_info &info) {
  const int64_t n = c.neq, nb = c.nb;
  if (!c.full_valid) PZG_TRY(launch_expand_full(c));
  // operator of the FP32 iterations
  bool use_mf = false;
  if (prm.inner_op == PZG_INNER_MATFREE) {
    if (!mf_available(c)) {
      set_error("pzg_solve: the matrix-free inner operator needs a mesh-built system");
      return PZG_ESTATE;
    }
    use_mf = true;
  } else if (prm.inner_op == PZG_INNER_AUTO) {
    use_mf = mf_available(c);
  } else if (prm.inner_op != PZG_INNER_BSR) {
    set_error("pzg_solve: unknown inner operator");
    return PZG_EINVAL;
  }
  // the stored FP32 values of S only for the stored operator (the
  // matrix-free one needs D^-1/2 and the block-Jacobi inverses alone)
  if (use_mf) PZG_TRY(mf_prepare(c));
  else if (!c.bsr_valid) PZG_TRY(launch_build_bsr(c, true));
  PZG_TRY(ensure_reduce(c));
  // the learned preconditioner is the multigrid cycle with the learned
  // element smoother on the p level
  const bool multigrid = prm.precond == PZG_PRECOND_MULTIGRID || prm.precond == PZG_PRECOND_LEARNED;
  for (DevBuf *b : {&c.d_u, &c.d_r}) PZG_TRY(dev_alloc(*b, n * sizeof(double)));
  for (DevBuf *b : {&c.d_xs, &c.d_rs, &c.d_zs, &c.d_ps, &c.d_qs}) PZG_TRY(dev_alloc(*b, n * sizeof(float)));
  double *out = c.d_res2.as<double>();
  SolveScalars *sc = c.d_scal.as<SolveScalars>();
  double *part = c.d_partials.as<double>();
  unsigned *cnt = c.d_counter.as<unsigned>();
  const int gv = grid_for(n, TPB), gb = grid_for(nb, TPB), gs = bsr_grid(nb);
  PZG_HIP(hipEventRecord(c.ev0, c.stream));
  hipLaunchKernelGGL(k_axpby64, dim3(gv), dim3(TPB), 0, c.stream, n, 0.0, c.d_f.as<double>(), 0.0,
                     (const double *)nullptr, c.d_u.as<double>());
  int64_t total_it = 0;
  int outer = 0;
  double rel = 1.0, fnorm = 0.;
  bool conv = false;
  // a multigrid iteration costs several operator applications: short batches
  const int batch = prm.check_every > 0 ? prm.check_every : (multigrid ? 4 : 50);
  auto true_res = [&](bool u_zero = false) -> int {
    if (u_zero)  // the starting guess: no product needed
      hipLaunchKernelGGL(k_residual_zero, dim3(gv), dim3(TPB), 0, c.stream, n, c.d_f.as<double>(),
                         c.d_r.as<double>(), out, part, cnt);
    else
      hipLaunchKernelGGL(k_residual64<BS>, dim3(bsr_grid(nb)), dim3(TPB), 0, c.stream, nb,
                         c.d_fbia.as<int64_t>(), c.d_fbja.as<int32_t>(), c.d_fa.as<double>(),
                         c.d_u.as<double>(), c.d_f.as<double>(), c.d_r.as<double>(), out,
                         part, cnt);
    double h[2];
    PZG_HIP(hipMemcpyAsync(h, out, sizeof(h), hipMemcpyDeviceToHost, c.stream));
    PZG_HIP(hipStreamSynchronize(c.stream));
    rel = h[1] > 0 ? std::sqrt(h[0] / h[1]) : std::sqrt(h[0]);
    fnorm = std::sqrt(h[1]);
    return PZG_OK;
  };
  auto next_delta = [&]() {
    // reduction asked of the FP32 recursion before the next true residual
    double d = prm.inner_tol > 0 ? prm.inner_tol : 1e-4;
    double need = 0.5 * prm.tol / rel;
    if (need > d) d = need;
    return d > 0.5 ? 0.5 : d;
  };
  PZG_TRY(true_res(true));
  if (rel <= prm.tol) conv = true;
  // block Jacobi on a high-order mesh: the real connects' blocks (several
  // shapes each), inverted in FP64 and scaled into the FP32 system
  const bool vbj = prm.precond == PZG_PRECOND_BLOCK_JACOBI && c.hp_simplex && c.vbj_nblk > 0 && !c.filt;
  if (vbj && !conv) {
    PZG_TRY(dev_alloc(c.d_vbjinv64, sizeof(double) * std::max<int64_t>(c.vbj_ninv, 1)));
    PZG_TRY(dev_alloc(c.d_vbjinv32, sizeof(float) * std::max<int64_t>(c.vbj_ninv, 1)));
    int *fail = reinterpret_cast<int *>(c.d_counter.as<unsigned>() + 8);
    PZG_HIP(hipMemsetAsync(fail, 0, sizeof(int), c.stream));
    hipLaunchKernelGGL(k_vbj_inv64, dim3((unsigned)((c.vbj_nblk + 3) / 4)), dim3(256), 0, c.stream, c.vbj_nblk,
                       c.d_vbj_pos.as<int64_t>(), c.d_vbj_m.as<int32_t>(), c.d_vbj_off.as<int64_t>(),
                       c.d_ia.as<int64_t>(), c.d_ja.as<int64_t>(), c.d_a.as<double>(), c.d_vbjinv64.as<double>(),
                       fail);
    hipLaunchKernelGGL(k_vbj_scale32, dim3(gv), dim3(TPB), 0, c.stream, n, c.d_vbj_eqb.as<int32_t>(),
                       c.d_vbj_pos.as<int64_t>(), c.d_vbj_m.as<int32_t>(), c.d_vbj_off.as<int64_t>(),
                       (const double *)c.d_vbjinv64.as<double>(), c.d_dinvsqrt.as<double>(),
                       c.d_vbjinv32.as<float>());
    int hf = 0;
    PZG_HIP(hipMemcpyAsync(&hf, fail, sizeof(int), hipMemcpyDeviceToHost, c.stream));
    PZG_HIP(hipStreamSynchronize(c.stream));
    if (hf) {
      set_error("pzg_solve: a connect block of K is singular or not stored whole");
      return PZG_EUNSUP;
    }
  }
  auto vbj_apply = [&](float *z, float *p, const float *q) {
    hipLaunchKernelGGL(k_vbj_apply32, dim3(gb), dim3(TPB), 0, c.stream, n, c.d_vbj_eqb.as<int32_t>(),
                       c.d_vbj_pos.as<int64_t>(), c.d_vbj_m.as<int32_t>(), c.d_vbj_off.as<int64_t>(),
                       (const float *)c.d_vbjinv32.as<float>(), (const float *)c.d_rs.as<float>(), z, p, q,
                       (const SolveScalars *)sc, part);
  };
  if (multigrid && !conv) {
    // levels and eigenvalue estimates (timed with the solve)
    const bool les = prm.precond == PZG_PRECOND_LEARNED;
    // the learned smoother covers a wider interval (64^3 elasticity: 146 ms
    // per step at 60 against 152 ms at 30, the same at 100)
    c.mg_degree = prm.mg_degree > 0 ? prm.mg_degree : 2;
    c.mg_range = prm.mg_range > 0 ? prm.mg_range : les ? 60. : 30.;
    if (!c.mg_ready || c.mg_use_mf != use_mf || c.mg_les != les) {
      c.mg_les = les;
      PZG_TRY(mg_setup(c, use_mf));
    }
    if (les && !c.pen_ready) PZG_TRY(pen_setup(c));
  }
  bool restart = true;
  bool learned = prm.precond == PZG_PRECOND_LEARNED;
  // PZG_TRACE=1: one line per refinement step on stderr
  const char *tr = std::getenv("PZG_TRACE");
  const bool trace = tr && tr[0] == '1';
  // learned path: the penalty rows are solved in FP64 here, after every fold,
  // to a tenth of the target residual; the FP32 iterations leave them alone
  const bool pen_rows = learned && !conv && c.pen_ready && c.pen_n > 0;
  // The penalty solve keeps d_r the exact residual f - K u (it subtracts
  // K delta on the rows that the penalized unknowns reach) and returns its
  // norm; rel_exact turns false only where an estimate stands in for it.
  int64_t pen_its = 0;
  bool rel_exact = true;
  auto pen_fix = [&]() -> int {
    pen_its = 0;
    if (!pen_rows || rel <= prm.tol) return PZG_OK;
    double p0 = 0., p1 = 0., rrf = -1.;
    // to a tenth of the target, or of what the other rows still carry
    const double tot2 = rel * rel * fnorm * fnorm;
    PZG_TRY(pen_solve(c, c.d_r.as<double>(), c.d_u.as<double>(), 0.1 * prm.tol * fnorm, prm.use_graph != 0,
                      &pen_its, &p0, &p1, 0.1, tot2, &rrf));
    if (pen_its == 0) return PZG_OK;  // u unchanged
    // without the exact update: the estimate with the -K_ip d change left out
    const double rr = rrf >= 0. ? rrf : std::max(tot2 - p0 + p1, p1);
    rel = std::sqrt(rr) / (fnorm > 0 ? fnorm : 1.);
    rel_exact = rrf >= 0.;
    return PZG_OK;
  };
  // the test on the estimate, confirmed by a true residual
  auto converged = [&](bool &cv) -> int {
    cv = false;
    if (rel > prm.tol) return PZG_OK;
    if (!rel_exact) {
      PZG_TRY(true_res());
      rel_exact = true;
    }
    cv = rel <= prm.tol;
    return PZG_OK;
  };
  PZG_TRY(pen_fix());
  PZG_TRY(converged(conv));
  int fallback = 0, stall = 0;
  double prev_rel = rel;
  SolveScalars hs;
  std::memset(&hs, 0, sizeof(hs));
  while (!conv) {
    if (restart) {
      hs.done = 0;
      hs.iter = total_it;
      // a multigrid-preconditioned inner solve takes tens of iterations: one
      // that runs past MG_INNER_CAP has stalled (the refinement step below
      // then sees no progress and falls back)
      hs.maxit = multigrid ? std::min<int64_t>(prm.max_iter, total_it + MG_INNER_CAP) : prm.max_iter;
      hs.delta = next_delta();
      PZG_HIP(hipMemcpyAsync(c.d_scal.p, &hs, sizeof(hs), hipMemcpyHostToDevice, c.stream));
      hipLaunchKernelGGL(k_fcg_init<BS>, dim3(gb), dim3(TPB), 0, c.stream, nb, c.d_r.as<double>(),
                         c.d_dinvsqrt.as<double>(), c.d_xs.as<float>(), c.d_rs.as<float>(),
                         c.d_zs.as<float>(), c.d_ps.as<float>(),
                         multigrid ? (const float *)nullptr : c.d_bjinv.as<float>(), sc, part, cnt);
      PZG_HIP(hipGetLastError());
      if (pen_rows) PZG_TRY(pen_zero32(c, c.d_rs.as<float>()));
      if (multigrid)
        PZG_TRY(mg_apply(c, c.d_rs.as<float>(), c.d_zs.as<float>(), c.d_ps.as<float>(), nullptr, sc, part));
      if (vbj) vbj_apply(c.d_zs.as<float>(), c.d_ps.as<float>(), nullptr);
      if (use_mf) PZG_TRY(mf_scale(c, c.d_ps.as<float>(), sc));
      restart = false;
    }
    auto iter = [&]() {
      int ga = gs;
      if (use_mf) {
        (void)mf_apply(c, c.d_ps.as<float>(), c.d_rs.as<float>(), c.d_qs.as<float>(), sc, part, &ga,
                       true);
      } else {
        hipLaunchKernelGGL(k_fcg_spmv<BS>, dim3(gs), dim3(TPB), 0, c.stream, nb,
                           c.d_fbia.as<int64_t>(), c.d_fbja.as<int32_t>(), c.d_bsr.as<float>(),
                           c.d_ps.as<float>(), c.d_rs.as<float>(), c.d_qs.as<float>(), sc, part);
      }
      hipLaunchKernelGGL(k_fin_fcg_alpha, dim3(1), dim3(TPB), 0, c.stream, sc, (const double *)part, ga);
      hipLaunchKernelGGL(k_fcg_update<BS>, dim3(gb), dim3(TPB), 0, c.stream, nb,
                         c.d_xs.as<float>(), c.d_rs.as<float>(), c.d_zs.as<float>(),
                         c.d_ps.as<float>(), c.d_qs.as<float>(), c.d_bjinv.as<float>(), sc, part,
                         cnt);
      // S couples the penalty rows to the others (at the 1/sqrt(big) level):
      // keep their residual at zero rather than let it drift
      if (pen_rows) (void)pen_zero32(c, c.d_rs.as<float>());
      if (multigrid)
        (void)mg_apply(c, c.d_rs.as<float>(), c.d_zs.as<float>(), nullptr, c.d_qs.as<float>(), sc, part);
      if (vbj) vbj_apply(c.d_zs.as<float>(), nullptr, c.d_qs.as<float>());
      hipLaunchKernelGGL(k_fin_fcg_beta, dim3(1), dim3(TPB), 0, c.stream, sc, (const double *)part, gb);
      hipLaunchKernelGGL(k_p_update32, dim3(gv), dim3(TPB), 0, c.stream, n, c.d_zs.as<float>(),
                         c.d_ps.as<float>(), sc, use_mf ? c.d_mf_dis.as<float>() : nullptr,
                         use_mf ? c.d_mf_pt.as<float>() : nullptr);
    };
    PZG_TRY(run_batches(c, batch, prm.use_graph != 0, iter, hs));
    total_it = hs.iter;
    // fold the FP32 correction into the FP64 solution and recompute the true residual
    hipLaunchKernelGGL(k_ir_update, dim3(gv), dim3(TPB), 0, c.stream, n, c.d_xs.as<float>(),
                       c.d_dinvsqrt.as<double>(), c.d_u.as<double>());
    PZG_HIP(hipGetLastError());
    PZG_TRY(true_res());
    rel_exact = true;
    const double rel_fold = rel;
    PZG_TRY(pen_fix());
    outer++;
    if (trace)
      std::fprintf(stderr, "[pzg] refinement %d: FP32 iterations %lld (stop %d), true residual %.3e"
                   " after the fold, %.3e after the penalty rows (%lld CG iterations)\n",
                   outer, (long long)total_it, hs.done, rel_fold, rel, (long long)pen_its);
    PZG_TRY(converged(conv));
    if (conv) break;
    if (hs.done == 2 && total_it >= prm.max_iter) break;
    // Stall of the FP32 iterations (the true residual no longer halves over
    // three refinement steps, or the refinement budget is spent; with the
    // learned preconditioner already when one step gains less than 10%):
    // continue with plain FP64 Jacobi-PCG from the current solution, judged
    // by the true FP64 residual.
    stall = rel > 0.5 * prev_rel ? stall + 1 : 0;
    if (stall >= 3 || outer >= prm.max_outer ||
        (learned && rel > 0.9 * prev_rel)) {
      fallback = 1;
      if (!rel_exact) {  // the PCG restarts from d_r: make it the true residual
        PZG_TRY(true_res());
        rel_exact = true;
      }
      if (trace) std::fprintf(stderr, "[pzg] stalled: FP64 Jacobi-PCG from residual %.3e\n", rel);
      PZG_TRY(pcg64_continue(c, prm, total_it, rel));
      conv = rel <= prm.tol;
      break;
    }
    prev_rel = rel;
    if (hs.done == 3 || prm.refine == 0) { restart = true; continue; }
    hipLaunchKernelGGL(k_fcg_reset, dim3(gv), dim3(TPB), 0, c.stream, n, c.d_r.as<double>(),
                       c.d_dinvsqrt.as<double>(), c.d_rs.as<float>(), next_delta(), sc, part, cnt);
    PZG_HIP(hipGetLastError());
    if (pen_rows) PZG_TRY(pen_zero32(c, c.d_rs.as<float>()));
  }
  if (!rel_exact) PZG_TRY(true_res());  // the reported residual is a true one
  PZG_HIP(hipEventRecord(c.ev1, c.stream));
  PZG_HIP(hipEventSynchronize(c.ev1));
  float ms = 0;
  PZG_HIP(hipEventElapsedTime(&ms, c.ev0, c.ev1));
  info.solve_ms = ms;
  info.iterations = total_it;
  info.outer = outer;
  info.true_residual = rel;
  info.achieved = rel;
  info.converged = conv ? 1 : 0;
  info.fallback = fallback;
  return conv ? PZG_OK : PZG_ENOCONV;
}

int solve_mixed(Ctx &c, const pzg_solve_params &prm, pzg_solve_info &info) {
  if (prm.precond == PZG_PRECOND_LEARNED && !c.les_ready) {
    set_error("pzg_solve: learned preconditioner not set up in this context");
    return PZG_ESTATE;
  }
  if (c.bs == 3) return mixed_impl<3>(c, prm, info);
  if (c.bs == 1) return mixed_impl<1>(c, prm, info);
  set_error("pzg_solve: unsupported block size for the mixed solver");
  return PZG_EUNSUP;
}

}  // namespace pzg
