// pzg_mg.hip -- multigrid preconditioner of the mixed FCG path on hexahedral
// and quadrilateral meshes whose vertices form a lattice (p <= 2; a plane
// lattice has one node in z, and every kernel treats that direction as the
// identity).
//
// Levels.  Level 0 is the FCG system S = D^-1/2 K D^-1/2 (matrix-free or
// stored, as the FCG iteration uses it).  Level 1 is the trilinear subspace:
// the hierarchical shape functions of H1 (TPZShapeCube: vertex functions are
// the trilinear ones, edge/face/interior functions vanish at the vertices)
// make it the span of the vertex functions, so its prolongation is the
// injection of vertex values and its Galerkin operator P^T K P is the
// vertex-vertex part K_vv of the assembled K, read from the FP64 block CSR.
// In level 0's scaled unknowns the transfer carries D^1/2 at the vertex
// equations.  Level 1 is stored in lattice order as a 27-point block
// stencil.  Levels 2.. halve the lattice with trilinear prolongation and the
// Galerkin product A_c = P^T A P formed on the device, while every element
// count is even and the level is too large for the dense solve; the coarsest
// level is solved with a dense inverse when it has at most MG_MAXDENSE
// unknowns; a larger one (an odd element count stopped the halving) gets a
// Chebyshev iteration of degree MG_COARSE_DEGREE from zero instead, a fixed
// SPD polynomial in M^-1 A, so the cycle stays a fixed SPD operator.
//
// Smoother: Chebyshev iteration (Saad, Iterative Methods for Sparse Linear
// Systems, 2nd ed., Alg. 12.1) of degree m on M^-1 A, M = the diagonal
// nstate x nstate blocks, on the interval [lhi/range, lhi], lhi = 1.1 x the
// largest eigenvalue estimated by CG-Lanczos; m = the requested degree on
// level 0 and 1 on the lattice levels (a second step there does not lower
// the iteration count).  Each level takes m pre- and m post-smoothing steps
// with the same polynomial and the coarsest solve is exact, so the
// preconditioner is a fixed symmetric positive definite operator.  Every kernel of the cycle leaves at once when the FCG
// iteration has stopped (graph batches run past the stop).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <thread>
#include <vector>

#include "pzg_device.h"
#include "pzg_internal.h"

namespace pzg {

namespace {

constexpr int TPB = RED_TPB;
constexpr int64_t MG_DENSE = 128;      // coarsening stops at this many unknowns
constexpr int64_t MG_MAXDENSE = 1536;  // largest dense coarsest solve
constexpr int MG_COARSE_DEGREE = 8;    // Chebyshev degree of a coarsest level beyond it
constexpr double MG_COARSE_RANGE = 100.;
constexpr int MG_LANCZOS_IT = 20;      // CG-Lanczos steps per eigenvalue estimate
constexpr int MG_LATTICE_DEGREE = 1;   // Chebyshev degree on the lattice levels

int grid1(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>((n + TPB - 1) / TPB, 1 << 20)); }
// one thread per item (kernels without a grid-stride loop)
int gridx(int64_t n) { return (int)std::max<int64_t>(1, (n + TPB - 1) / TPB); }
int gridr(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>((n + TPB - 1) / TPB, MAXBLK)); }

__device__ inline bool stopped(const SolveScalars *sc) { return sc && sc->done; }

// y = M^-1 x at block A: entry (A, k) of the inverse blocks at m[A*sa + k*sk]
template <int BS>
__device__ inline void minv_mul(const float *__restrict__ m, int64_t sa, int64_t sk, int64_t A,
                                const float (&x)[BS], float (&y)[BS]) {
#pragma unroll
  for (int i = 0; i < BS; i++) {
    float s = 0.f;
#pragma unroll
    for (int j = 0; j < BS; j++) s += m[A * sa + (i * BS + j) * sk] * x[j];
    y[i] = s;
  }
}

// ---- Chebyshev steps (vectors [nb][BS]) --------------------------------------
// Level-0 variants also write pt = D^-1/2 d (pt != null), the input of the
// matrix-free operator's next application.
// r = b; d = c0 M^-1 b; x = d
template <int BS>
__global__ void k_cheb_init(int64_t nb, const float *__restrict__ m, int64_t sa, int64_t sk,
                            const float *__restrict__ b, float *__restrict__ x, float *__restrict__ r,
                            float *__restrict__ d, float c0, const float *__restrict__ dis,
                            float *__restrict__ pt, const SolveScalars *sc) {
  if (stopped(sc)) return;
  for (int64_t A = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; A < nb;
       A += (int64_t)gridDim.x * blockDim.x) {
    float bv[BS], y[BS];
#pragma unroll
    for (int k = 0; k < BS; k++) bv[k] = b[A * BS + k];
    minv_mul<BS>(m, sa, sk, A, bv, y);
#pragma unroll
    for (int k = 0; k < BS; k++) {
      r[A * BS + k] = bv[k];
      d[A * BS + k] = c0 * y[k];
      x[A * BS + k] = c0 * y[k];
      if (pt) pt[A * BS + k] = c0 * y[k] * dis[A * BS + k];
    }
  }
}

// r = b - t (t = A x); d = c0 M^-1 r; x += d
template <int BS>
__global__ void k_cheb_restart(int64_t nb, const float *__restrict__ m, int64_t sa, int64_t sk,
                               const float *__restrict__ b, const float *__restrict__ t,
                               float *__restrict__ x, float *__restrict__ r, float *__restrict__ d,
                               float c0, const float *__restrict__ dis, float *__restrict__ pt,
                               const SolveScalars *sc) {
  if (stopped(sc)) return;
  for (int64_t A = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; A < nb;
       A += (int64_t)gridDim.x * blockDim.x) {
    float rv[BS], y[BS];
#pragma unroll
    for (int k = 0; k < BS; k++) {
      rv[k] = b[A * BS + k] - t[A * BS + k];
      r[A * BS + k] = rv[k];
    }
    minv_mul<BS>(m, sa, sk, A, rv, y);
#pragma unroll
    for (int k = 0; k < BS; k++) {
      d[A * BS + k] = c0 * y[k];
      x[A * BS + k] += c0 * y[k];
      if (pt) pt[A * BS + k] = c0 * y[k] * dis[A * BS + k];
    }
  }
}

// r -= t; d = c1 d + c2 M^-1 r; x += d
template <int BS>
__global__ void k_cheb_step(int64_t nb, const float *__restrict__ m, int64_t sa, int64_t sk,
                            const float *__restrict__ t, float *__restrict__ r, float *__restrict__ d,
                            float *__restrict__ x, float c1, float c2, const float *__restrict__ dis,
                            float *__restrict__ pt, const SolveScalars *sc) {
  if (stopped(sc)) return;
  for (int64_t A = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; A < nb;
       A += (int64_t)gridDim.x * blockDim.x) {
    float rv[BS], y[BS];
#pragma unroll
    for (int k = 0; k < BS; k++) {
      rv[k] = r[A * BS + k] - t[A * BS + k];
      r[A * BS + k] = rv[k];
    }
    minv_mul<BS>(m, sa, sk, A, rv, y);
#pragma unroll
    for (int k = 0; k < BS; k++) {
      const float dv = c1 * d[A * BS + k] + c2 * y[k];
      d[A * BS + k] = dv;
      x[A * BS + k] += dv;
      if (pt) pt[A * BS + k] = dv * dis[A * BS + k];
    }
  }
}

// Learned smoother on level 0 (PZG_PRECOND_LEARNED): M r is an operator of
// its own (les_apply), so a Chebyshev step splits into the residual update and
// y = M r, whose gather also applies the direction / iterate update.
// mode 0: r = b;  1: r -= t;  2: r = b - t
__global__ void k_les_res(int64_t n, int mode, const float *__restrict__ b, const float *__restrict__ t,
                          float *__restrict__ r, const SolveScalars *sc) {
  if (stopped(sc)) return;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    r[i] = mode == 0 ? b[i] : mode == 1 ? r[i] - t[i] : b[i] - t[i];
}
// ---- lattice operators ---------------------------------------------------------
struct Dims {
  int n[3];
};
__device__ inline void node_ijk(const Dims &D, int64_t v, int &i, int &j, int &k) {
  i = (int)(v % D.n[0]);
  const int64_t q = v / D.n[0];
  j = (int)(q % D.n[1]);
  k = (int)(q / D.n[1]);
}
__device__ inline int64_t node_id(const Dims &D, int i, int j, int k) {
  return i + (int64_t)D.n[0] * (j + (int64_t)D.n[1] * k);
}

// y = A x over the 27-point block stencil; stencil entry s = (dz+1)*9 +
// (dy+1)*3 + (dx+1), value (s, e) of node v at A[(s*BS*BS + e)*nn + v].
// One thread per (node, row component a), node fastest: the small levels
// are latency-bound and take BS times the threads.
template <int BS>
__global__ void k_stencil(Dims D, int64_t nn, const float *__restrict__ A, const float *__restrict__ x,
                          float *__restrict__ y, const SolveScalars *sc) {
  if (stopped(sc)) return;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nn * BS;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t v = t % nn;
    const int a = (int)(t / nn);
    int i, j, k;
    node_ijk(D, v, i, j, k);
    float acc = 0.f;
    for (int s = 0; s < 27; s++) {
      const int ii = i + s % 3 - 1, jj = j + (s / 3) % 3 - 1, kk = k + s / 9 - 1;
      if (ii < 0 || jj < 0 || kk < 0 || ii >= D.n[0] || jj >= D.n[1] || kk >= D.n[2]) continue;
      const int64_t w = node_id(D, ii, jj, kk);
#pragma unroll
      for (int b = 0; b < BS; b++) acc += A[((int64_t)s * BS * BS + a * BS + b) * nn + v] * x[w * BS + b];
    }
    y[v * BS + a] = acc;
  }
}

// trilinear restriction of the residual r_f - t_f (the pre-smoothing leaves
// its last product t_f unsubtracted): b_c(N) = sum over the fine nodes
// 2N + d, d in {-1,0,1}^3, of prod_i (1 - |d_i|/2) (r_f - t_f)
template <int BS>
__global__ void k_restrict(Dims C, Dims F, const float *__restrict__ rf, const float *__restrict__ tf,
                           float *__restrict__ bc, const SolveScalars *sc) {
  if (stopped(sc)) return;
  const int64_t nc = (int64_t)C.n[0] * C.n[1] * C.n[2];
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nc;
       v += (int64_t)gridDim.x * blockDim.x) {
    int I, J, K;
    node_ijk(C, v, I, J, K);
    float acc[BS];
#pragma unroll
    for (int e = 0; e < BS; e++) acc[e] = 0.f;
    for (int dz = -1; dz <= 1; dz++) {
      const int k = 2 * K + dz;
      if (k < 0 || k >= F.n[2]) continue;
      for (int dy = -1; dy <= 1; dy++) {
        const int j = 2 * J + dy;
        if (j < 0 || j >= F.n[1]) continue;
        for (int dx = -1; dx <= 1; dx++) {
          const int i = 2 * I + dx;
          if (i < 0 || i >= F.n[0]) continue;
          const float w = (dx ? 0.5f : 1.f) * (dy ? 0.5f : 1.f) * (dz ? 0.5f : 1.f);
          const int64_t f = node_id(F, i, j, k);
#pragma unroll
          for (int e = 0; e < BS; e++) acc[e] += w * (rf[f * BS + e] - tf[f * BS + e]);
        }
      }
    }
#pragma unroll
    for (int e = 0; e < BS; e++) bc[v * BS + e] = acc[e];
  }
}

// trilinear prolongation, added: x_f += P x_c
template <int BS>
__global__ void k_prolong(Dims C, Dims F, const float *__restrict__ xc, float *__restrict__ xf,
                          const SolveScalars *sc) {
  if (stopped(sc)) return;
  const int64_t nf = (int64_t)F.n[0] * F.n[1] * F.n[2];
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nf;
       v += (int64_t)gridDim.x * blockDim.x) {
    int i, j, k;
    node_ijk(F, v, i, j, k);
    const int i0 = i >> 1, j0 = j >> 1, k0 = k >> 1;
    const int ni = (i & 1) ? 2 : 1, nj = (j & 1) ? 2 : 1, nk = (k & 1) ? 2 : 1;
    const float wi = (i & 1) ? 0.5f : 1.f, wj = (j & 1) ? 0.5f : 1.f, wk = (k & 1) ? 0.5f : 1.f;
    float acc[BS];
#pragma unroll
    for (int e = 0; e < BS; e++) acc[e] = 0.f;
    for (int c = 0; c < nk; c++)
      for (int b = 0; b < nj; b++)
        for (int a = 0; a < ni; a++) {
          const int64_t w = node_id(C, i0 + a, j0 + b, k0 + c);
#pragma unroll
          for (int e = 0; e < BS; e++) acc[e] += xc[w * BS + e];
        }
    const float s = wi * wj * wk;
#pragma unroll
    for (int e = 0; e < BS; e++) xf[v * BS + e] += s * acc[e];
  }
}

// level 0 <-> level 1: b1(v) = D^1/2 (r0 - t0) at the vertex block vmap[v];
// x0(vmap[v]) += D^1/2 x1(v)
template <int BS>
__global__ void k_restrict0(int64_t nn, const int32_t *__restrict__ vmap, const float *__restrict__ dsq,
                            const float *__restrict__ r0, const float *__restrict__ t0,
                            float *__restrict__ b1, const SolveScalars *sc) {
  if (stopped(sc)) return;
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nn;
       v += (int64_t)gridDim.x * blockDim.x) {
    const int64_t A = vmap[v];
#pragma unroll
    for (int e = 0; e < BS; e++) b1[v * BS + e] = dsq[v * BS + e] * (r0[A * BS + e] - t0[A * BS + e]);
  }
}
template <int BS>
__global__ void k_prolong0(int64_t nn, const int32_t *__restrict__ vmap, const float *__restrict__ dsq,
                           const float *__restrict__ x1, float *__restrict__ x0, const SolveScalars *sc) {
  if (stopped(sc)) return;
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nn;
       v += (int64_t)gridDim.x * blockDim.x) {
    const int64_t A = vmap[v];
#pragma unroll
    for (int e = 0; e < BS; e++) x0[A * BS + e] += dsq[v * BS + e] * x1[v * BS + e];
  }
}

// x = Ainv b (dense, symmetric, nd unknowns): one wavefront per row
__global__ void k_dense(int nd, const float *__restrict__ ainv, const float *__restrict__ b,
                        float *__restrict__ x, const SolveScalars *sc) {
  if (stopped(sc)) return;
  const int i = (int)(((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int lane = threadIdx.x & 63;
  if (i >= nd) return;
  float s = 0.f;
  for (int j = lane; j < nd; j += 64) s += ainv[(int64_t)i * nd + j] * b[j];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if (lane == 0) x[i] = s;
}

// z stays the preconditioned residual; p = z (restart); z.q replaces slot 1
// of the FCG update kernel's partials (same grid)
template <int BS>
__global__ __launch_bounds__(TPB) void k_mg_finish(int64_t nb, const float *__restrict__ z,
                                                   float *__restrict__ p, const float *__restrict__ q,
                                                   const SolveScalars *sc, double *partials) {
  if (stopped(sc)) return;
  double zq = 0.;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nb * BS;
       i += (int64_t)gridDim.x * blockDim.x) {
    const float v = z[i];
    if (p) p[i] = v;
    if (q) zq += (double)v * q[i];
  }
  if (!q) return;
  __shared__ double sm[TPB / 64];
  const double t = block_sum(zq, sm);
  if (threadIdx.x == 0) partials[1 * MAXBLK + blockIdx.x] = t;
}

// ---- set-up kernels --------------------------------------------------------------
// level 1 stencil from the FP64 full block CSR (values at bsr_pos): block
// (vmap[v], vmap[w]) for the lattice neighbour w of v in direction s (zero
// off the lattice or when the row does not hold the column).  One wavefront
// per vertex v scans its block row once: vinv maps a block to its vertex (-1
// for the other connects); the directions found are gathered in a mask so
// that the zero entries and the found ones are written by disjoint stores.
template <int BS>
__global__ void k_mg_l1(Dims D, int64_t nn, const int32_t *__restrict__ vmap, const int32_t *__restrict__ vinv,
                        const int64_t *__restrict__ fbia, const int32_t *__restrict__ fbja,
                        const double *__restrict__ fa, float *__restrict__ A) {
  constexpr int BB = BS * BS;
  const int64_t v = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (v >= nn) return;
  int i, j, k;
  node_ijk(D, v, i, j, k);
  const int64_t row = vmap[v];
  const int64_t b0 = row >= 0 ? fbia[row] : 0, nr = row >= 0 ? fbia[row + 1] - b0 : 0;
  int found = 0;
  for (int64_t q = lane; q < nr; q += 64) {
    const int32_t w = vinv[fbja[b0 + q]];
    if (w < 0) continue;
    int ii, jj, kk;
    node_ijk(D, w, ii, jj, kk);
    const int di = ii - i, dj = jj - j, dk = kk - k;
    if (di < -1 || di > 1 || dj < -1 || dj > 1 || dk < -1 || dk > 1) continue;
    const int sd = (di + 1) + 3 * (dj + 1) + 9 * (dk + 1);
    found |= 1 << sd;
    for (int e = 0; e < BB; e++) A[((int64_t)sd * BB + e) * nn + v] = (float)fa[bsr_pos<BB>(b0, nr, q, e)];
  }
  for (int o = 32; o > 0; o >>= 1) found |= __shfl_xor(found, o, 64);
  for (int t = lane; t < 27 * BB; t += 64) {
    const int sd = t / BB;
    if (!((found >> sd) & 1)) A[(int64_t)t * nn + v] = 0.f;
  }
}

// Galerkin product A_c = P^T A_f P for trilinear P, one thread per coarse
// (node, stencil entry).  Along each direction the fine node f = 2N + a
// (a in {-1,0,1}, weight 1 - |a|/2) couples to the fine node g = 2M + b of
// the neighbour M = N + o when |g - f| <= 1; the pairs (a, b) are enumerated
// per direction.
template <int BS>
__global__ void k_mg_rap(Dims C, Dims F, const float *__restrict__ Af, float *__restrict__ Ac) {
  const int64_t nc = (int64_t)C.n[0] * C.n[1] * C.n[2];
  const int64_t nf = (int64_t)F.n[0] * F.n[1] * F.n[2];
  const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  // one thread per (coarse node, stencil entry, block entry): the coarse
  // levels are small, so the block entries spread the work over more threads
  if (x >= nc * 27 * BS * BS) return;
  const int64_t v = x % nc;
  const int s = (int)((x / nc) % 27), e = (int)(x / (nc * 27));
  int N[3];
  node_ijk(C, v, N[0], N[1], N[2]);
  const int o[3] = {s % 3 - 1, (s / 3) % 3 - 1, s / 9 - 1};
  double acc = 0.;
  bool inside = true;
  for (int d = 0; d < 3; d++) inside &= N[d] + o[d] >= 0 && N[d] + o[d] < C.n[d];
  if (inside) {
    // per direction: fine pairs (f, g) with weights and the stencil offset g - f
    int pf[3][9], pg[3][9], np[3];
    float pw[3][9];
    for (int d = 0; d < 3; d++) {
      np[d] = 0;
      const int M = N[d] + o[d];
      for (int a = -1; a <= 1; a++) {
        const int f = 2 * N[d] + a;
        if (f < 0 || f >= F.n[d]) continue;
        for (int b = -1; b <= 1; b++) {
          const int g = 2 * M + b;
          if (g < 0 || g >= F.n[d] || g - f < -1 || g - f > 1) continue;
          pf[d][np[d]] = f;
          pg[d][np[d]] = g - f;
          pw[d][np[d]] = (a ? 0.5f : 1.f) * (b ? 0.5f : 1.f);
          np[d]++;
        }
      }
    }
    for (int c = 0; c < np[2]; c++)
      for (int b = 0; b < np[1]; b++)
        for (int a = 0; a < np[0]; a++) {
          const int64_t f = pf[0][a] + (int64_t)F.n[0] * (pf[1][b] + (int64_t)F.n[1] * pf[2][c]);
          const int sf = (pg[2][c] + 1) * 9 + (pg[1][b] + 1) * 3 + (pg[0][a] + 1);
          const double w = (double)pw[0][a] * pw[1][b] * pw[2][c];
          acc += w * Af[((int64_t)sf * BS * BS + e) * nf + f];
        }
  }
  Ac[((int64_t)s * BS * BS + e) * nc + v] = (float)acc;
}

// inverse of the diagonal block (stencil entry 13), stored [e*nn + v]
template <int BS>
__global__ void k_mg_minv(int64_t nn, const float *__restrict__ A, float *__restrict__ minv) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nn) return;
  double a[BS * BS];
  for (int e = 0; e < BS * BS; e++) a[e] = A[((int64_t)13 * BS * BS + e) * nn + v];
  double r[BS * BS];
  if (BS == 1) {
    r[0] = a[0] != 0. ? 1. / a[0] : 0.;
  } else {
    const double c00 = a[4] * a[8] - a[5] * a[7], c01 = a[5] * a[6] - a[3] * a[8],
                 c02 = a[3] * a[7] - a[4] * a[6];
    const double det = a[0] * c00 + a[1] * c01 + a[2] * c02;
    const double id = det != 0. ? 1. / det : 0.;
    r[0] = c00 * id;
    r[1] = (a[2] * a[7] - a[1] * a[8]) * id;
    r[2] = (a[1] * a[5] - a[2] * a[4]) * id;
    r[3] = c01 * id;
    r[4] = (a[0] * a[8] - a[2] * a[6]) * id;
    r[5] = (a[2] * a[3] - a[0] * a[5]) * id;
    r[6] = c02 * id;
    r[7] = (a[1] * a[6] - a[0] * a[7]) * id;
    r[8] = (a[0] * a[4] - a[1] * a[3]) * id;
  }
  for (int e = 0; e < BS * BS; e++) minv[(int64_t)e * nn + v] = (float)r[e];
}

// dsq(v) = 1 / dinvsqrt at the vertex block vmap[v]
template <int BS>
__global__ void k_mg_dsq(int64_t nn, const int32_t *__restrict__ vmap, const double *__restrict__ dis,
                         float *__restrict__ dsq) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nn) return;
  for (int e = 0; e < BS; e++) dsq[v * BS + e] = (float)(1.0 / dis[(int64_t)vmap[v] * BS + e]);
}

__global__ void k_hash_fill(int64_t n, float *__restrict__ v, uint64_t seed) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    uint64_t h = (uint64_t)i * 0x9E3779B97F4A7C15ULL + seed;
    h ^= h >> 31;
    h *= 0xBF58476D1CE4E5B9ULL;
    h ^= h >> 29;
    v[i] = (float)((double)(h >> 11) / 9007199254740992.0) - 0.5f;
  }
}

// Chebyshev coefficients of one level
struct Cheb {
  float c0;
  std::vector<float> c1, c2;  // steps 1 .. m-1
};
Cheb cheb_coeffs(double lmax, double range, int m) {
  const double hi = 1.1 * lmax, lo = hi / range;
  const double theta = 0.5 * (hi + lo), delta = 0.5 * (hi - lo), sigma = theta / delta;
  Cheb ch;
  ch.c0 = (float)(1.0 / theta);
  double rho = 1.0 / sigma;
  for (int j = 1; j < m; j++) {
    const double rn = 1.0 / (2.0 * sigma - rho);
    ch.c1.push_back((float)(rn * rho));
    ch.c2.push_back((float)(2.0 * rn / delta));
    rho = rn;
  }
  return ch;
}

// A level for the smoother: operator, inverse diagonal blocks, vectors
struct LevelView {
  int64_t nb;              // blocks
  const float *m;          // inverse diagonal blocks
  int64_t sa, sk;          // their layout
  float *r, *d, *t;        // work vectors
  const float *dis = nullptr;  // level 0, matrix-free: D^-1/2 and the operator's
  float *pt = nullptr;         // pre-scaled input, written with each new d
  float *y = nullptr;          // level 0, learned smoother: M r (else the blocks m)
  Cheb ch;
};

// op(x, y, prescaled): y = A x; prescaled: L.pt already holds D^-1/2 x.
// Leaves x and, as r - t, the residual b - A x (the restriction subtracts).
// the learned smoother's steps (LevelView::y set)
template <int BS, class Op>
int smooth_les(Ctx &c, const LevelView &L, const float *b, float *x, Op op, const SolveScalars *sc,
               bool pre) {
  const int64_t n = L.nb * BS;
  const int g = grid1(n);
  const bool ps = L.pt != nullptr;
  if (pre) {
    hipLaunchKernelGGL(k_les_res, dim3(g), dim3(TPB), 0, c.stream, n, 0, b, (const float *)nullptr, L.r, sc);
  } else {
    PZG_TRY(op(x, L.t, false));
    hipLaunchKernelGGL(k_les_res, dim3(g), dim3(TPB), 0, c.stream, n, 2, b, L.t, L.r, sc);
  }
  LesCheb ch{pre ? 0 : 1, 0.f, (float)L.ch.c0, L.d, x, L.pt, L.dis};
  PZG_TRY(les_apply(c, L.r, nullptr, sc, &ch));
  for (size_t j = 0; j < L.ch.c1.size(); j++) {
    PZG_TRY(op(L.d, L.t, ps));
    hipLaunchKernelGGL(k_les_res, dim3(g), dim3(TPB), 0, c.stream, n, 1, b, L.t, L.r, sc);
    LesCheb chj{2, (float)L.ch.c1[j], (float)L.ch.c2[j], L.d, x, L.pt, L.dis};
    PZG_TRY(les_apply(c, L.r, nullptr, sc, &chj));
  }
  return pre ? op(L.d, L.t, ps) : PZG_OK;
}

template <int BS, class Op>
int smooth_pre(Ctx &c, const LevelView &L, const float *b, float *x, Op op, const SolveScalars *sc) {
  if (L.y) return smooth_les<BS>(c, L, b, x, op, sc, true);
  const int g = grid1(L.nb);
  const bool ps = L.pt != nullptr;
  hipLaunchKernelGGL(k_cheb_init<BS>, dim3(g), dim3(TPB), 0, c.stream, L.nb, L.m, L.sa, L.sk, b, x, L.r,
                     L.d, L.ch.c0, L.dis, L.pt, sc);
  PZG_HIP(hipGetLastError());
  for (size_t j = 0; j < L.ch.c1.size(); j++) {
    PZG_TRY(op(L.d, L.t, ps));
    hipLaunchKernelGGL(k_cheb_step<BS>, dim3(g), dim3(TPB), 0, c.stream, L.nb, L.m, L.sa, L.sk, L.t, L.r,
                       L.d, x, L.ch.c1[j], L.ch.c2[j], L.dis, L.pt, sc);
    PZG_HIP(hipGetLastError());
  }
  return op(L.d, L.t, ps);
}

template <int BS, class Op>
int smooth_post(Ctx &c, const LevelView &L, const float *b, float *x, Op op, const SolveScalars *sc) {
  if (L.y) return smooth_les<BS>(c, L, b, x, op, sc, false);
  const int g = grid1(L.nb);
  const bool ps = L.pt != nullptr;
  PZG_TRY(op(x, L.t, false));
  hipLaunchKernelGGL(k_cheb_restart<BS>, dim3(g), dim3(TPB), 0, c.stream, L.nb, L.m, L.sa, L.sk, b, L.t, x,
                     L.r, L.d, L.ch.c0, L.dis, L.pt, sc);
  PZG_HIP(hipGetLastError());
  for (size_t j = 0; j < L.ch.c1.size(); j++) {
    PZG_TRY(op(L.d, L.t, ps));
    hipLaunchKernelGGL(k_cheb_step<BS>, dim3(g), dim3(TPB), 0, c.stream, L.nb, L.m, L.sa, L.sk, L.t, L.r,
                       L.d, x, L.ch.c1[j], L.ch.c2[j], L.dis, L.pt, sc);
    PZG_HIP(hipGetLastError());
  }
  return PZG_OK;
}

Dims dims_of(const MgLevel &l) { return Dims{{l.n[0], l.n[1], l.n[2]}}; }

// level-0 operator y = S x (matrix-free or stored FP32); prescaled: the
// matrix-free input D^-1/2 x is already in place
int op0(Ctx &c, const float *x, float *y, const SolveScalars *sc, bool prescaled) {
  if (c.mg_use_mf) return mf_apply(c, x, nullptr, y, sc, nullptr, nullptr, prescaled);
  return launch_spmv_bsr(c, x, y);
}

template <int BS>
int op_lattice(Ctx &c, const MgLevel &l, const float *x, float *y, const SolveScalars *sc) {
  hipLaunchKernelGGL(k_stencil<BS>, dim3(grid1(l.nn * BS)), dim3(TPB), 0, c.stream, dims_of(l), l.nn,
                     l.A.as<float>(), x, y, sc);
  PZG_HIP(hipGetLastError());
  return PZG_OK;
}

LevelView view0(Ctx &c) {
  LevelView v;
  v.nb = c.nb;
  v.m = c.d_bjinv.as<float>();
  v.sa = (int64_t)c.bs * c.bs;
  v.sk = 1;
  v.r = c.d_mg_r0.as<float>();
  v.d = c.d_mg_d0.as<float>();
  v.t = c.d_mg_t0.as<float>();
  if (c.mg_use_mf) {
    v.dis = c.d_mf_dis.as<float>();
    v.pt = c.d_mf_pt.as<float>();
  }
  if (c.mg_les) v.y = c.d_mg_y0.as<float>();
  v.ch = cheb_coeffs(c.mg_lmax0, c.mg_range, c.mg_degree);
  return v;
}
LevelView view_l(Ctx &c, MgLevel &l) {
  LevelView v;
  v.nb = l.nn;
  v.m = l.minv.as<float>();
  v.sa = 1;
  v.sk = l.nn;
  v.r = l.r.as<float>();
  v.d = l.d.as<float>();
  v.t = l.t.as<float>();
  v.ch = cheb_coeffs(l.lmax, c.mg_range, MG_LATTICE_DEGREE);
  return v;
}

// largest eigenvalue of the symmetric tridiagonal (a, b): bisection on the
// Sturm count (eigenvalues below x = negative pivots of T - x I)
double tridiag_max(const std::vector<double> &a, const std::vector<double> &b) {
  const int k = (int)a.size();
  double hi = 0.;
  for (int i = 0; i < k; i++)
    hi = std::max(hi, a[i] + (i > 0 ? std::fabs(b[i - 1]) : 0.) + (i + 1 < k ? std::fabs(b[i]) : 0.));
  double lo = 0.;
  auto below = [&](double x) {
    int cnt = 0;
    double d = 1.;
    for (int i = 0; i < k; i++) {
      d = a[i] - x - (i > 0 ? b[i - 1] * b[i - 1] / d : 0.);
      if (d == 0.) d = -1e-300;
      if (d < 0.) cnt++;
    }
    return cnt;
  };
  for (int it = 0; it < 100; it++) {
    const double mid = 0.5 * (lo + hi);
    if (below(mid) == k) hi = mid;
    else lo = mid;
  }
  return hi;
}

// CG-Lanczos on the device.  lz[0] = r.z of the current step, lz[1 + 2k] =
// alpha_k, lz[2 + 2k] = beta_k, lz[LZ_ZERO] = 0; a step whose p.q or r.z is
// not positive leaves alpha = 0 (and every later step with it).
constexpr int LZ_ZERO = 2 * MG_LANCZOS_IT + 1;

// r -= alpha q; z = M^-1 r; partial sums of r.z (alpha = *al)
template <int BS>
__global__ __launch_bounds__(TPB) void k_lz_update(int64_t nb, const float *__restrict__ m, int64_t sa,
                                                   int64_t sk, const double *__restrict__ al,
                                                   const float *__restrict__ q, float *__restrict__ r,
                                                   float *__restrict__ z, double *partials) {
  const float alpha = (float)*al;
  double rz = 0.;
  for (int64_t A = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; A < nb;
       A += (int64_t)gridDim.x * blockDim.x) {
    float rv[BS], y[BS];
    for (int k = 0; k < BS; k++) {
      rv[k] = r[A * BS + k] - alpha * q[A * BS + k];
      r[A * BS + k] = rv[k];
    }
    minv_mul<BS>(m, sa, sk, A, rv, y);
    for (int k = 0; k < BS; k++) {
      z[A * BS + k] = y[k];
      rz += (double)rv[k] * y[k];
    }
  }
  double vv[1] = {rz};
  block_partials<1>(vv, partials);
}
__global__ __launch_bounds__(TPB) void k_dot(int64_t n, const float *__restrict__ a,
                                             const float *__restrict__ b, double *partials) {
  double s = 0.;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    s += (double)a[i] * b[i];
  double vv[1] = {s};
  block_partials<1>(vv, partials);
}
// one block: what = 0: lz[0] = r.z; 1: alpha_k = lz[0] / p.q; 2: beta_k =
// r.z / lz[0], lz[0] = r.z
__global__ void k_lz_scalar(const double *partials, int nblk, double *lz, int k, int what) {
  const double s = sum_partials(partials, 0, nblk);
  if (threadIdx.x != 0) return;
  if (what == 0) {
    lz[0] = s;
  } else if (what == 1) {
    const bool live = k == 0 || lz[1 + 2 * (k - 1)] > 0.;
    lz[1 + 2 * k] = live && s > 0. && lz[0] > 0. ? lz[0] / s : 0.;
  } else {
    const bool live = lz[1 + 2 * k] > 0. && s > 0.;
    lz[2 + 2 * k] = live ? s / lz[0] : 0.;
    if (live) lz[0] = s;
  }
}
// p = z + beta p (beta = *be)
__global__ void k_xpby(int64_t n, const float *__restrict__ z, const double *__restrict__ be,
                       float *__restrict__ p) {
  const float beta = (float)*be;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = z[i] + beta * p[i];
}

// largest eigenvalue of M^-1 A: MG_LANCZOS_IT steps of preconditioned CG
// from a pseudo-random right-hand side, all on the device; the CG
// coefficients give the Lanczos tridiagonal (diagonal 1/alpha_k +
// beta_{k-1}/alpha_{k-1}, off-diagonal sqrt(beta_k)/alpha_k), whose largest
// eigenvalue approximates lmax from below.  Work vectors: r, z, p, q.
// r -= alpha q (alpha = *al)
__global__ void k_lz_rsub(int64_t n, const double *__restrict__ al, const float *__restrict__ q,
                          float *__restrict__ r) {
  const float alpha = (float)*al;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    r[i] -= alpha * q[i];
}

template <int BS, class Op>
int lanczos_lmax(Ctx &c, int64_t nb, const float *m, int64_t sa, int64_t sk, float *r, float *z, float *p,
                 float *q, Op op, double &lmax, bool les = false) {
  const int64_t n = nb * BS;
  double *part = c.d_partials.as<double>();
  double *lz = c.d_mg_lz.as<double>();
  const int gr = gridr(n), gb = gridr(nb);
  // r -= alpha q; z = M r; partials of r.z; returns the number of partial blocks
  auto mstep = [&](const double *al, const float *qq) -> int {
    if (!les) {
      hipLaunchKernelGGL(k_lz_update<BS>, dim3(gb), dim3(TPB), 0, c.stream, nb, m, sa, sk, al, qq, r, z, part);
      return gb;
    }
    hipLaunchKernelGGL(k_lz_rsub, dim3(grid1(n)), dim3(TPB), 0, c.stream, n, al, qq, r);
    if (les_apply(c, r, z, nullptr) != PZG_OK) return -1;
    hipLaunchKernelGGL(k_dot, dim3(gr), dim3(TPB), 0, c.stream, n, r, z, part);
    return gr;
  };
  PZG_HIP(hipMemsetAsync(lz, 0, sizeof(double) * (LZ_ZERO + 1), c.stream));
  hipLaunchKernelGGL(k_hash_fill, dim3(grid1(n)), dim3(TPB), 0, c.stream, n, r, (uint64_t)0x5DEECE66DULL);
  // z = M^-1 r and r.z (the update with alpha = 0)
  int nblk = mstep(lz + LZ_ZERO, r);
  if (nblk < 0) return PZG_EHIP;
  hipLaunchKernelGGL(k_lz_scalar, dim3(1), dim3(TPB), 0, c.stream, (const double *)part, nblk, lz, 0, 0);
  PZG_HIP(hipGetLastError());
  PZG_HIP(hipMemcpyAsync(p, z, sizeof(float) * n, hipMemcpyDeviceToDevice, c.stream));
  for (int k = 0; k < MG_LANCZOS_IT; k++) {
    PZG_TRY(op(p, q));
    hipLaunchKernelGGL(k_dot, dim3(gr), dim3(TPB), 0, c.stream, n, p, q, part);
    hipLaunchKernelGGL(k_lz_scalar, dim3(1), dim3(TPB), 0, c.stream, (const double *)part, gr, lz, k, 1);
    nblk = mstep(lz + 1 + 2 * k, q);
    if (nblk < 0) return PZG_EHIP;
    hipLaunchKernelGGL(k_lz_scalar, dim3(1), dim3(TPB), 0, c.stream, (const double *)part, nblk, lz, k, 2);
    hipLaunchKernelGGL(k_xpby, dim3(grid1(n)), dim3(TPB), 0, c.stream, n, z, lz + 2 + 2 * k, p);
    PZG_HIP(hipGetLastError());
  }
  std::vector<double> h(LZ_ZERO + 1);
  PZG_HIP(hipMemcpyAsync(h.data(), lz, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c.stream));
  PZG_HIP(hipStreamSynchronize(c.stream));
  std::vector<double> ta, tb;
  for (int k = 0; k < MG_LANCZOS_IT; k++) {
    const double alpha = h[1 + 2 * k], beta = h[2 + 2 * k];
    if (!(alpha > 0.)) break;
    ta.push_back(1.0 / alpha + (k > 0 ? h[2 * k] / h[2 * k - 1] : 0.));
    if (!(beta > 0.)) break;
    tb.push_back(std::sqrt(beta) / alpha);
  }
  if (ta.empty()) {
    set_error("multigrid: operator not positive definite on a level");
    return PZG_EUNSUP;
  }
  tb.resize(ta.size());
  lmax = tridiag_max(ta, tb);
  return PZG_OK;
}

// inverse of a symmetric positive definite matrix (Cholesky, in place)
// In-place inverse of a symmetric positive definite matrix (FP64, n x n) by
// Gauss-Jordan sweeps without pivoting, one pivot per launch pair: k_gj_save
// keeps row and column k, k_gj_step applies the sweep
//   a_kk = 1/p, a_kj = a_kj/p, a_ik = -a_ik/p, a_ij -= a_ik a_kj / p
// (p = a_kk; after every pivot the array holds A^-1).  A pivot <= 0 sets
// *fail (A not positive definite).
__global__ void k_gj_save(int n, const double *__restrict__ m, int k, double *__restrict__ prow,
                          double *__restrict__ pcol) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) {
    prow[t] = m[(size_t)k * n + t];
    pcol[t] = m[(size_t)t * n + k];
  }
}
__global__ void k_gj_step(int n, double *__restrict__ m, int k, const double *__restrict__ prow,
                          const double *__restrict__ pcol, int *fail) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)n * n) return;
  const int i = (int)(t / n), j = (int)(t % n);
  const double p = prow[k];
  if (t == 0 && !(p > 0.)) *fail = 1;
  const double ip = 1. / p;
  double v;
  if (i == k) v = j == k ? ip : prow[j] * ip;
  else if (j == k) v = -pcol[i] * ip;
  else v = m[t] - pcol[i] * prow[j] * ip;
  m[t] = v;
}
__global__ void k_d2f_n(int64_t n, const double *__restrict__ a, float *__restrict__ b) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) b[t] = (float)a[t];
}

// Mesh-level part: the vertex lattice, the levels and their buffers (kept
// until the mesh changes).
template <int BS>
int structure_impl(Ctx &c) {
  mg_free(c);
  // vertex lattice: the corner nodes of the volume elements and their blocks
  const bool hex = c.dim == 3 && c.vol_etype == PZG_HEX, quad = c.dim == 2 && c.vol_etype == PZG_QUAD;
  if (!c.has_mesh || !(hex || quad) || c.hp_simplex || c.has_dep || c.external_csr) {
    set_error("multigrid preconditioner: needs a hexahedral or quadrilateral mesh-built system without "
              "constrained connects");
    return PZG_EUNSUP;
  }
  const int nc = c.vol_ncorner, nsd = c.vol_nsides;
  std::vector<int64_t> nblk(c.nnodes, -1);
  int64_t nvert = 0;
  for (int64_t e = 0; e < c.nvol; e++)
    for (int k = 0; k < nc; k++) {  // sides 0 .. nc-1: the corners
      const int64_t node = c.vol_nodes[e * nc + k];
      const int64_t b = c.aidx[c.vol_conn[e * nsd + k]];
      if (b < 0 || c.blk_size[c.vol_conn[e * nsd + k]] != BS) {
        set_error("multigrid preconditioner: a vertex connect without its own equations");
        return PZG_EUNSUP;
      }
      if (nblk[node] == -1) {
        nblk[node] = b;
        nvert++;
      } else if (nblk[node] != b) {
        set_error("multigrid preconditioner: a mesh node with two vertex connects");
        return PZG_EUNSUP;
      }
    }
  // distinct coordinates per direction (to 1e-9 of the extent)
  int n[3];
  std::vector<std::vector<double>> u(3);
  for (int d = 0; d < 3; d++) {
    std::vector<double> x;
    for (int64_t p = 0; p < c.nnodes; p++)
      if (nblk[p] >= 0) x.push_back(c.xyz[3 * p + d]);
    std::sort(x.begin(), x.end());
    const double tol = 1e-9 * std::max(1.0, std::fabs(x.back() - x.front()));
    for (double v : x)
      if (u[d].empty() || v - u[d].back() > tol) u[d].push_back(v);
    n[d] = (int)u[d].size();
  }
  const int64_t nn1 = (int64_t)n[0] * n[1] * n[2];
  if (nn1 != nvert || nn1 > INT32_MAX) {
    set_error("multigrid preconditioner: the mesh vertices do not form a lattice");
    return PZG_EUNSUP;
  }
  std::vector<int32_t> vmap(nn1, -1);
  for (int64_t p = 0; p < c.nnodes; p++) {
    if (nblk[p] < 0) continue;
    int id[3];
    for (int d = 0; d < 3; d++) {
      const double v = c.xyz[3 * p + d];
      auto it = std::lower_bound(u[d].begin(), u[d].end(), v - 1e-9 * std::max(1.0, std::fabs(u[d].back() - u[d].front())));
      id[d] = (int)(it - u[d].begin());
      if (id[d] >= n[d]) id[d] = n[d] - 1;
    }
    const int64_t l = id[0] + (int64_t)n[0] * (id[1] + (int64_t)n[1] * id[2]);
    if (vmap[l] != -1) {
      set_error("multigrid preconditioner: the mesh vertices do not form a lattice");
      return PZG_EUNSUP;
    }
    vmap[l] = (int32_t)nblk[p];
  }
  // levels: halve while the element counts are even and the level is too
  // large for the dense solve
  {
    MgLevel l;
    for (int d = 0; d < 3; d++) l.n[d] = n[d];
    l.nn = nn1;
    c.mg.push_back(std::move(l));
  }
  for (;;) {
    const MgLevel &f = c.mg.back();
    if (f.nn * BS <= MG_DENSE) break;
    // a direction of one node (the z of a plane lattice) stays as it is
    bool even = true;
    for (int d = 0; d < 3; d++) even &= f.n[d] == 1 || ((f.n[d] - 1) % 2 == 0 && f.n[d] >= 3);
    if (!even) break;
    MgLevel l;
    for (int d = 0; d < 3; d++) l.n[d] = f.n[d] == 1 ? 1 : (f.n[d] - 1) / 2 + 1;
    l.nn = (int64_t)l.n[0] * l.n[1] * l.n[2];
    c.mg.push_back(std::move(l));
  }
  c.mg_dense = c.mg.back().nn * BS <= MG_MAXDENSE;
  // device arrays
  PZG_TRY(dev_upload(c.d_mg_vmap, vmap.data(), vmap.size(), c.stream));
  {
    std::vector<int32_t> vinv(c.nb, -1);  // block -> vertex of the level-1 lattice
    for (size_t w = 0; w < vmap.size(); w++)
      if (vmap[w] >= 0) vinv[vmap[w]] = (int32_t)w;
    PZG_TRY(dev_upload(c.d_mg_vinv, vinv.data(), vinv.size(), c.stream));
  }
  PZG_TRY(dev_alloc(c.d_mg_dsq, sizeof(float) * nn1 * BS));
  PZG_TRY(dev_alloc(c.d_mg_lz, sizeof(double) * (LZ_ZERO + 1)));
  for (DevBuf *b : {&c.d_mg_r0, &c.d_mg_d0, &c.d_mg_t0, &c.d_mg_w0}) PZG_TRY(dev_alloc(*b, sizeof(float) * c.neq));
  for (MgLevel &l : c.mg) {
    PZG_TRY(dev_alloc(l.A, sizeof(float) * 27 * BS * BS * l.nn));
    PZG_TRY(dev_alloc(l.minv, sizeof(float) * BS * BS * l.nn));
    for (DevBuf *b : {&l.x, &l.b, &l.r, &l.d, &l.t}) PZG_TRY(dev_alloc(*b, sizeof(float) * BS * l.nn));
  }
  c.mg_struct = true;
  return PZG_OK;
}

// Matrix-level part (after every assembly): D^1/2 at the vertices, the
// level operators, the coarsest inverse and the eigenvalue estimates.
template <int BS>
int setup_impl(Ctx &c) {
  if (!c.mg_struct) PZG_TRY(structure_impl<BS>(c));
  const int64_t nn1 = c.mg[0].nn;
  hipLaunchKernelGGL(k_mg_dsq<BS>, dim3(gridx(nn1)), dim3(TPB), 0, c.stream, nn1, c.d_mg_vmap.as<int32_t>(),
                     c.d_dinvsqrt.as<double>(), c.d_mg_dsq.as<float>());
  hipLaunchKernelGGL(k_mg_l1<BS>, dim3(gridx(64 * nn1)), dim3(TPB), 0, c.stream, dims_of(c.mg[0]), nn1,
                     c.d_mg_vmap.as<int32_t>(), c.d_mg_vinv.as<int32_t>(), c.d_fbia.as<int64_t>(),
                     c.d_fbja.as<int32_t>(), c.d_fa.as<double>(), c.mg[0].A.as<float>());
  PZG_HIP(hipGetLastError());
  for (size_t k = 1; k < c.mg.size(); k++) {
    hipLaunchKernelGGL(k_mg_rap<BS>, dim3(gridx(27 * BS * BS * c.mg[k].nn)), dim3(TPB), 0, c.stream, dims_of(c.mg[k]),
                       dims_of(c.mg[k - 1]), c.mg[k - 1].A.as<float>(), c.mg[k].A.as<float>());
    PZG_HIP(hipGetLastError());
  }
  for (MgLevel &l : c.mg) {
    hipLaunchKernelGGL(k_mg_minv<BS>, dim3(gridx(l.nn)), dim3(TPB), 0, c.stream, l.nn, l.A.as<float>(),
                       l.minv.as<float>());
    PZG_HIP(hipGetLastError());
  }
  // dense inverse of the coarsest level
  if (c.mg_dense) {
    const MgLevel &l = c.mg.back();
    const int nd = (int)(l.nn * BS);
    std::vector<float> h((size_t)27 * BS * BS * l.nn);
    PZG_HIP(hipMemcpyAsync(h.data(), l.A.p, h.size() * sizeof(float), hipMemcpyDeviceToHost, c.stream));
    PZG_HIP(hipStreamSynchronize(c.stream));
    std::vector<double> a((size_t)nd * nd, 0.);
    for (int64_t v = 0; v < l.nn; v++) {
      const int i = (int)(v % l.n[0]), j = (int)((v / l.n[0]) % l.n[1]), k = (int)(v / ((int64_t)l.n[0] * l.n[1]));
      for (int s = 0; s < 27; s++) {
        const int ii = i + s % 3 - 1, jj = j + (s / 3) % 3 - 1, kk = k + s / 9 - 1;
        if (ii < 0 || jj < 0 || kk < 0 || ii >= l.n[0] || jj >= l.n[1] || kk >= l.n[2]) continue;
        const int64_t w = ii + (int64_t)l.n[0] * (jj + (int64_t)l.n[1] * kk);
        for (int e = 0; e < BS * BS; e++)
          a[(size_t)(v * BS + e / BS) * nd + w * BS + e % BS] = h[((size_t)s * BS * BS + e) * l.nn + v];
      }
    }
    for (int i = 0; i < nd; i++)  // symmetric part (the stencil holds both triangles)
      for (int j = i + 1; j < nd; j++) {
        const double m = 0.5 * (a[(size_t)i * nd + j] + a[(size_t)j * nd + i]);
        a[(size_t)i * nd + j] = a[(size_t)j * nd + i] = m;
      }
    // inverse on the device (a host inverse of 1536 unknowns took 0.2-0.4 s)
    DevBuf dm, dp;
    PZG_TRY(dev_upload(dm, a.data(), a.size(), c.stream));
    if (int rc = dev_alloc(dp, sizeof(double) * 2 * nd + 64)) {
      dev_free(dm);
      return rc;
    }
    double *prow = dp.as<double>(), *pcol = prow + nd;
    int *fail = reinterpret_cast<int *>(pcol + nd);
    PZG_HIP(hipMemsetAsync(fail, 0, sizeof(int), c.stream));
    const int64_t nn2 = (int64_t)nd * nd;
    for (int k = 0; k < nd; k++) {
      hipLaunchKernelGGL(k_gj_save, dim3((nd + TPB - 1) / TPB), dim3(TPB), 0, c.stream, nd, dm.as<double>(), k,
                         prow, pcol);
      hipLaunchKernelGGL(k_gj_step, dim3((unsigned)((nn2 + TPB - 1) / TPB)), dim3(TPB), 0, c.stream, nd,
                         dm.as<double>(), k, (const double *)prow, (const double *)pcol, fail);
    }
    PZG_TRY(dev_alloc(c.d_mg_cinv, sizeof(float) * nn2));
    hipLaunchKernelGGL(k_d2f_n, dim3((unsigned)((nn2 + TPB - 1) / TPB)), dim3(TPB), 0, c.stream, nn2,
                       dm.as<double>(), c.d_mg_cinv.as<float>());
    PZG_HIP(hipGetLastError());
    int hf = 0;
    PZG_HIP(hipMemcpyAsync(&hf, fail, sizeof(int), hipMemcpyDeviceToHost, c.stream));
    PZG_HIP(hipStreamSynchronize(c.stream));
    dev_free(dm);
    dev_free(dp);
    if (hf) {
      set_error("multigrid preconditioner: the coarsest operator is not positive definite");
      return PZG_EUNSUP;
    }
    c.mg_ncoarse = nd;
  }
  // largest eigenvalues of M^-1 A: level 0, then every smoothed lattice level
  {
    if (c.mg_les) {  // the learned smoother of the current assembly
      PZG_TRY(les_prepare(c));
      PZG_TRY(dev_alloc(c.d_mg_y0, sizeof(float) * std::max<int64_t>(c.neq, 1)));
    }
    auto op = [&](const float *x, float *y) { return op0(c, x, y, nullptr, false); };
    PZG_TRY(lanczos_lmax<BS>(c, c.nb, c.d_bjinv.as<float>(), (int64_t)BS * BS, 1, c.d_mg_r0.as<float>(),
                             c.d_mg_d0.as<float>(), c.d_mg_t0.as<float>(), c.d_mg_w0.as<float>(), op,
                             c.mg_lmax0, c.mg_les));
  }
  for (size_t k = 0; k + (c.mg_dense ? 1 : 0) < c.mg.size(); k++) {
    MgLevel &l = c.mg[k];
    auto op = [&](const float *x, float *y) { return op_lattice<BS>(c, l, x, y, nullptr); };
    PZG_TRY(lanczos_lmax<BS>(c, l.nn, l.minv.as<float>(), 1, l.nn, l.x.as<float>(), l.b.as<float>(),
                             l.d.as<float>(), l.t.as<float>(), op, l.lmax));
  }
  PZG_HIP(hipStreamSynchronize(c.stream));
  c.mg_ready = true;
  return PZG_OK;
}

// z = V-cycle applied to r; then p = z (when given) and the z.q partials
// (when q is given), as the block-Jacobi update leaves them
template <int BS>
int apply_impl(Ctx &c, const float *r, float *z, float *p, const float *q, SolveScalars *sc,
               double *partials) {
  const LevelView L0 = view0(c);
  auto op_0 = [&](const float *x, float *y, bool ps) { return op0(c, x, y, sc, ps); };
  const int nl = (int)c.mg.size();
  PZG_TRY(smooth_pre<BS>(c, L0, r, z, op_0, sc));
  hipLaunchKernelGGL(k_restrict0<BS>, dim3(grid1(c.mg[0].nn)), dim3(TPB), 0, c.stream, c.mg[0].nn,
                     c.d_mg_vmap.as<int32_t>(), c.d_mg_dsq.as<float>(), L0.r, L0.t, c.mg[0].b.as<float>(), sc);
  PZG_HIP(hipGetLastError());
  for (int k = 0; k + 1 < nl; k++) {
    MgLevel &l = c.mg[k];
    const LevelView V = view_l(c, l);
    auto op = [&](const float *x, float *y, bool) { return op_lattice<BS>(c, l, x, y, sc); };
    PZG_TRY(smooth_pre<BS>(c, V, l.b.as<float>(), l.x.as<float>(), op, sc));
    hipLaunchKernelGGL(k_restrict<BS>, dim3(grid1(c.mg[k + 1].nn)), dim3(TPB), 0, c.stream,
                       dims_of(c.mg[k + 1]), dims_of(l), V.r, V.t, c.mg[k + 1].b.as<float>(), sc);
    PZG_HIP(hipGetLastError());
  }
  if (c.mg_dense) {
    hipLaunchKernelGGL(k_dense, dim3(gridx(c.mg_ncoarse * 64)), dim3(TPB), 0, c.stream, (int)c.mg_ncoarse,
                       c.d_mg_cinv.as<float>(), c.mg.back().b.as<float>(), c.mg.back().x.as<float>(), sc);
    PZG_HIP(hipGetLastError());
  } else {  // x = q(M^-1 A) M^-1 b, a Chebyshev iteration from zero
    MgLevel &l = c.mg.back();
    LevelView V = view_l(c, l);
    V.ch = cheb_coeffs(l.lmax, MG_COARSE_RANGE, MG_COARSE_DEGREE);
    auto op = [&](const float *x, float *y, bool) { return op_lattice<BS>(c, l, x, y, sc); };
    PZG_TRY(smooth_pre<BS>(c, V, l.b.as<float>(), l.x.as<float>(), op, sc));
  }
  for (int k = nl - 2; k >= 0; k--) {
    MgLevel &l = c.mg[k];
    hipLaunchKernelGGL(k_prolong<BS>, dim3(grid1(l.nn)), dim3(TPB), 0, c.stream, dims_of(c.mg[k + 1]),
                       dims_of(l), c.mg[k + 1].x.as<float>(), l.x.as<float>(), sc);
    PZG_HIP(hipGetLastError());
    const LevelView V = view_l(c, l);
    auto op = [&](const float *x, float *y, bool) { return op_lattice<BS>(c, l, x, y, sc); };
    PZG_TRY(smooth_post<BS>(c, V, l.b.as<float>(), l.x.as<float>(), op, sc));
  }
  hipLaunchKernelGGL(k_prolong0<BS>, dim3(grid1(c.mg[0].nn)), dim3(TPB), 0, c.stream, c.mg[0].nn,
                     c.d_mg_vmap.as<int32_t>(), c.d_mg_dsq.as<float>(), c.mg[0].x.as<float>(), z, sc);
  PZG_HIP(hipGetLastError());
  PZG_TRY(smooth_post<BS>(c, L0, r, z, op_0, sc));
  // the learned path leaves the penalty rows to the FP64 refinement
  if ((p || q) && c.mg_les) PZG_TRY(pen_zero32(c, z));
  if (p || q) {
    // same grid as the FCG update kernel, whose partials slot 1 this replaces
    hipLaunchKernelGGL(k_mg_finish<BS>, dim3(gridr(c.nb)), dim3(TPB), 0, c.stream, c.nb, z, p, q, sc, partials);
    PZG_HIP(hipGetLastError());
  }
  return PZG_OK;
}

}  // namespace

void mg_free(Ctx &c) {
  pen_free(c);
  for (MgLevel &l : c.mg)
    for (DevBuf *b : {&l.A, &l.minv, &l.x, &l.b, &l.r, &l.d, &l.t}) dev_free(*b);
  c.mg.clear();
  c.mg_struct = false;
  for (DevBuf *b : {&c.d_mg_vmap, &c.d_mg_vinv, &c.d_mg_dsq, &c.d_mg_cinv, &c.d_mg_r0, &c.d_mg_d0, &c.d_mg_t0, &c.d_mg_w0,
                    &c.d_mg_lz, &c.d_mg_y0})
    dev_free(*b);
  c.mg_ready = false;
}

int mg_setup(Ctx &c, bool use_mf) {
  c.mg_ready = false;
  c.mg_use_mf = use_mf;
  if (c.mg_degree < 1 || c.mg_degree > 8 || !(c.mg_range > 1.)) {
    set_error("multigrid preconditioner: degree must be 1..8 and the range above 1");
    return PZG_EINVAL;
  }
  if (c.bs == 3) return setup_impl<3>(c);
  if (c.bs == 1) return setup_impl<1>(c);
  set_error("multigrid preconditioner: nstate must be 1 or 3");
  return PZG_EUNSUP;
}

int mg_apply(Ctx &c, const float *r, float *z, float *p, const float *q, SolveScalars *sc, double *partials) {
  if (!c.mg_ready) {
    set_error("multigrid preconditioner: not set up");
    return PZG_ESTATE;
  }
  if (c.bs == 3) return apply_impl<3>(c, r, z, p, q, sc, partials);
  return apply_impl<1>(c, r, z, p, q, sc, partials);
}

}  // namespace pzg
