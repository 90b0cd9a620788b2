// pzg_learned.hip -- the learned element smoother of the mixed solver's
// multigrid V-cycle (PZG_PRECOND_LEARNED).  NeoPZ has no learned component;
// this serves the FP32 iterations only (the FP64 true residual still decides
// convergence).
//
// Operator.  For a scaled residual r (S = D^-1/2 K D^-1/2) the smoother returns
//     z = sum_e s_e P_e^T G P_e r   on the unpenalized equations,
//     z_i = r_i                     on the Dirichlet-penalty equations,
// where P_e gathers the element's ne = nshape x nstate dofs (shape-major, the
// order of the element matrix), G is one learned ne x ne symmetric matrix and
// s_e > 0 is the output of a message-passing network on the element graph
// (elements adjacent through a face), 3 layers of 32 features:
//     H0 = X W0;  H_l = lrelu(H_{l-1} A_l + (Ahat H_{l-1}) B_l);  s = exp(H3 w)
// (Ahat = mean over the face neighbours).  X holds four features of the
// element: log c_e - mean of the neighbours' log c, the largest |log c_n -
// log c_e| over the neighbours (c: the element's E or permeability), the
// fraction of penalized dofs, and 1 - neighbours / faces.  It replaces the
// block-Jacobi M^-1 of the Chebyshev smoothing on the p level; the cycle
// stays a fixed SPD operator for a fixed field.
//
// Training data are (residual, error) pairs of the local problems the solver
// sees: for every element of a training system, the element's block S_e of
// the assembled scaled matrix (penalized dofs decoupled), random residuals r
// (zero on penalized dofs) and the errors e = S_e^-1 r from a batched
// Cholesky solve.  The loss
//     L = sum_e sum_k |mask (s_e G r_ek - e_ek)|^2 / sum_e sum_k |e_ek|^2
// is minimised in two parts.  G: the least-squares fit G = B A^-1 with
// A = sum r r^T, B = sum e r^T over the pairs of the elements without
// penalized dofs (sums formed on the device, accumulated over every
// training system; for isotropic r this is the mean of the local inverses
// S_e^-1), symmetrised and projected onto the positive definite matrices
// (eigenvalues below 1e-3 of the largest raised), so the smoother stays SPD.
// The network: Adam on the loss with G fixed.  G starts as diag(1 / number of
// elements sharing the dof's side) (a point-Jacobi-like smoother), the output
// layer at zero (s = 1).  Forward, backward and Adam are the kernels below;
// the gradient (of every parameter, G included) is checked against a numpy
// statement.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <random>

#include "pzg_device.h"
#include "pzg_internal.h"

namespace pzg {

constexpr int TPB = RED_TPB;
constexpr int NBLK_RED = 512;  // blocks of the weight-gradient reductions
constexpr int LES_FIN = 4;     // element features
typedef float floatx4 __attribute__((ext_vector_type(4)));

__device__ inline float lrelu(float x) { return x > 0.f ? x : NN_SLOPE * x; }
__device__ inline float dlrelu(float x) { return x > 0.f ? 1.f : NN_SLOPE; }

// H0 = X W0   (rows = samples * nc)
template <int BS>
__global__ void k_nn_in(int64_t rows, const float *__restrict__ X, const float *__restrict__ W0,
                        float *__restrict__ H) {
  int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= rows * NNF) return;
  int64_t r = g / NNF;
  int f = (int)(g % NNF);
  float s = 0.f;
  for (int k = 0; k < BS; k++) s += X[r * BS + k] * W0[k * NNF + f];
  H[g] = s;
}

// M[s][I] = mean_{J in N(I)} H[s][J]
__global__ void k_nn_agg(int64_t rows, int64_t nc, const int64_t *__restrict__ cia,
                         const int32_t *__restrict__ cja, const float *__restrict__ H,
                         float *__restrict__ M) {
  int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= rows * NNF) return;
  int64_t r = g / NNF;
  int f = (int)(g % NNF);
  int64_t s = r / nc, I = r % nc;
  const int64_t j0 = cia[I], j1 = cia[I + 1];
  float acc = 0.f;
  for (int64_t k = j0; k < j1; k++) acc += H[(s * nc + cja[k]) * NNF + f];
  M[g] = j1 > j0 ? acc / (float)(j1 - j0) : 0.f;
}

// out[s][J] (+)= sum_{I in N(J)} Y[s][I] / deg(I)   (transpose of the mean)
__global__ void k_nn_agg_t(int64_t rows, int64_t nc, const int64_t *__restrict__ cia,
                           const int32_t *__restrict__ cja, const float *__restrict__ Y,
                           float *__restrict__ out) {
  int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= rows * NNF) return;
  int64_t r = g / NNF;
  int f = (int)(g % NNF);
  int64_t s = r / nc, J = r % nc;
  float acc = 0.f;
  for (int64_t k = cia[J]; k < cia[J + 1]; k++) {
    const int64_t I = cja[k];
    acc += Y[(s * nc + I) * NNF + f] / (float)(cia[I + 1] - cia[I]);
  }
  out[g] += acc;
}

// P = H A + M B;  Hout = lrelu(P)   (weights staged in LDS)
__global__ __launch_bounds__(TPB) void k_nn_layer(int64_t rows, const float *__restrict__ H,
                                                  const float *__restrict__ M,
                                                  const float *__restrict__ A,
                                                  const float *__restrict__ B,
                                                  float *__restrict__ P, float *__restrict__ Hout) {
  __shared__ float sA[NNF * NNF], sB[NNF * NNF];
  for (int k = threadIdx.x; k < NNF * NNF; k += blockDim.x) {
    sA[k] = A[k];
    sB[k] = B[k];
  }
  __syncthreads();
  int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= rows * NNF) return;
  int64_t r = g / NNF;
  int f = (int)(g % NNF);
  const float *h = H + r * NNF, *m = M + r * NNF;
  float s = 0.f;
  for (int k = 0; k < NNF; k++) s += h[k] * sA[k * NNF + f] + m[k] * sB[k * NNF + f];
  if (P) P[g] = s;
  Hout[g] = lrelu(s);
}

// Z = H Wout
template <int BS>
__global__ void k_nn_out(int64_t rows, const float *__restrict__ H, const float *__restrict__ Wout,
                         float *__restrict__ Z) {
  int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= rows * BS) return;
  int64_t r = g / BS;
  int k = (int)(g % BS);
  float s = 0.f;
  for (int f = 0; f < NNF; f++) s += H[r * NNF + f] * Wout[f * BS + k];
  Z[g] = s;
}

// ---- training kernels -------------------------------------------------------
// dH = dZ Wout^T
template <int BS>
__global__ void k_nn_back_out(int64_t rows, const float *__restrict__ dZ,
                              const float *__restrict__ Wout, float *__restrict__ dH) {
  int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= rows * NNF) return;
  int64_t r = g / NNF;
  int f = (int)(g % NNF);
  float s = 0.f;
  for (int k = 0; k < BS; k++) s += dZ[r * BS + k] * Wout[f * BS + k];
  dH[g] = s;
}

// G = dH * lrelu'(P);  dHprev = G A^T;  GBt = G B^T
__global__ __launch_bounds__(TPB) void k_nn_back_layer(int64_t rows, const float *__restrict__ dH,
                                                       const float *__restrict__ P,
                                                       const float *__restrict__ A,
                                                       const float *__restrict__ B,
                                                       float *__restrict__ G,
                                                       float *__restrict__ dHprev,
                                                       float *__restrict__ GBt) {
  __shared__ float sA[NNF * NNF], sB[NNF * NNF];
  __shared__ float sG[TPB];
  for (int k = threadIdx.x; k < NNF * NNF; k += blockDim.x) {
    sA[k] = A[k];
    sB[k] = B[k];
  }
  int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  float gv = 0.f;
  if (g < rows * NNF) gv = dH[g] * dlrelu(P[g]);
  sG[threadIdx.x] = gv;
  __syncthreads();
  if (g >= rows * NNF) return;
  G[g] = gv;
  const int f = (int)(g % NNF);
  const float *gr = sG + (threadIdx.x - f);  // the row's 32 gradients (rows never straddle blocks)
  float sa = 0.f, sb = 0.f;
  for (int k = 0; k < NNF; k++) {
    sa += gr[k] * sA[f * NNF + k];
    sb += gr[k] * sB[f * NNF + k];
  }
  dHprev[g] = sa;
  GBt[g] = sb;
}

// dX = dH W0^T is not needed (inputs are data); dW0 etc. via outer products.
// partial[blk][a][b] = sum_{rows in blk} L[row][a] * R[row][b]
template <int FA, int FB>
__global__ __launch_bounds__(TPB) void k_nn_outer(int64_t rows, const float *__restrict__ Lm,
                                                  const float *__restrict__ Rm,
                                                  float *__restrict__ partial) {
  constexpr int NE = FA * FB;
  float acc[(NE + TPB - 1) / TPB];
  for (int k = 0; k < (NE + TPB - 1) / TPB; k++) acc[k] = 0.f;
  const int64_t per = (rows + gridDim.x - 1) / gridDim.x;
  const int64_t r0 = blockIdx.x * per, r1 = (rows < r0 + per) ? rows : r0 + per;
  __shared__ float sL[32][FA], sR[32][FB];
  for (int64_t rb = r0; rb < r1; rb += 32) {
    int nr = (int)((r1 - rb) < 32 ? (r1 - rb) : 32);
    __syncthreads();
    for (int k = threadIdx.x; k < 32 * FA; k += blockDim.x) {
      int rr = k / FA, a = k % FA;
      sL[rr][a] = rr < nr ? Lm[(rb + rr) * FA + a] : 0.f;
    }
    for (int k = threadIdx.x; k < 32 * FB; k += blockDim.x) {
      int rr = k / FB, b = k % FB;
      sR[rr][b] = rr < nr ? Rm[(rb + rr) * FB + b] : 0.f;
    }
    __syncthreads();
    for (int k = 0; k < (NE + TPB - 1) / TPB; k++) {
      int e = threadIdx.x + k * TPB;
      if (e < NE) {
        int a = e / FB, b = e % FB;
        float s = 0.f;
        for (int rr = 0; rr < 32; rr++) s += sL[rr][a] * sR[rr][b];
        acc[k] += s;
      }
    }
  }
  for (int k = 0; k < (NE + TPB - 1) / TPB; k++) {
    int e = threadIdx.x + k * TPB;
    if (e < NE) partial[(int64_t)blockIdx.x * NE + e] = acc[k];
  }
}

// g[off + e] = sum_blk partial[blk][e]  (fixed order)
__global__ void k_nn_sum_partials(int nblk, int ne, const float *__restrict__ partial,
                                  float *__restrict__ g) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  double s = 0.;
  for (int b = 0; b < nblk; b++) s += partial[(int64_t)b * ne + e];
  g[e] = (float)s;
}

__global__ void k_adam(int n, float *__restrict__ th, const float *__restrict__ g,
                       float *__restrict__ m, float *__restrict__ v, float lr, float b1, float b2,
                       float c1, float c2, float eps) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float gi = g[i];
  float mi = b1 * m[i] + (1.f - b1) * gi;
  float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  th[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}


// ---- parameters ---------------------------------------------------------------
struct LesOff {
  int g, w0, a[NNL], b[NNL], wout, n;
};
static LesOff les_offsets(int ne) {
  LesOff o;
  int k = 0;
  o.g = k; k += ne * ne;
  o.w0 = k; k += LES_FIN * NNF;
  for (int l = 0; l < NNL; l++) {
    o.a[l] = k; k += NNF * NNF;
    o.b[l] = k; k += NNF * NNF;
  }
  o.wout = k; k += NNF;
  o.n = k;
  return o;
}
int les_nparams(int ne) { return les_offsets(ne).n; }

// s = exp(z), the exponent clamped to [-4, 4]
__global__ void k_les_exp(int64_t n, const float *__restrict__ z, float *__restrict__ s) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) s[i] = expf(fminf(4.f, fmaxf(-4.f, z[i])));
}

// ---- application ---------------------------------------------------------------
// Elements with NE = NSH*NS <= 96 dofs: z_e = s_e G r_e for 16 elements per
// wavefront as v_mfma_f32_16x16x4f32 tiles: rows = the element's output dofs
// padded to NT*16, columns = elements, k = input dofs padded to NK*4.  G is
// staged in LDS in the A operand order (tile ti, k-step kk, lane l:
// G[16 ti + l%16][4 kk + l/16]), so every LDS read of a wavefront is one
// contiguous segment.  Lane (element l%16, quarter l/16) gathers its NK
// inputs: the element's block of each input shape (blkm: block index, bit
// 28 + component set on penalized dofs, -1 when the side has no shapes),
// then the values, both as batches so each is one round trip to memory; the
// next group's batches are in flight while the current group multiplies.
// The outputs are stored element by element (NE floats each); the gather
// below sums them per block row.
// G (row-major, ne x ne) into LDS in A-operand order of the 16 x 16 x 4
// tiles: entry (row, k) at ((row/16) NK + k/4) 64 + row%16 + 16 (k%4), the
// padding zero.  G is read in order (coalesced); the caller's next barrier
// publishes sA.
template <int NE, int NT, int NK, int NTH = 256>
__device__ inline void les_stage_G(const float *__restrict__ G, float *sA, int tid) {
  for (int t = tid; t < NT * NK * 64; t += NTH) {
    const int ll = t & 63, kk = (t >> 6) % NK, ti = (t >> 6) / NK;
    const int row = 16 * ti + (ll & 15), k = 4 * kk + (ll >> 4);
    if (row >= NE || k >= NE) sA[t] = 0.f;
  }
  // all of a thread's loads issued before its first store (a load-store
  // pair per iteration would wait out the memory latency every time)
  constexpr int NPT = (NE * NE + NTH - 1) / NTH;
  float v[NPT];
#pragma unroll
  for (int u = 0; u < NPT; u++) {
    const int t = tid + NTH * u;
    v[u] = G[t < NE * NE ? t : 0];
  }
#pragma unroll
  for (int u = 0; u < NPT; u++) {
    const int t = tid + NTH * u;
    if (t < NE * NE) {
      const int row = t / NE, k = t % NE;
      sA[((row >> 4) * NK + (k >> 2)) * 64 + (row & 15) + 16 * (k & 3)] = v[u];
    }
  }
}

template <int NSH, int NS>
__global__ __launch_bounds__(256) void k_les_mfma(int64_t nvol, const float *__restrict__ G,
                                                  const float *__restrict__ scale,
                                                  const int32_t *__restrict__ blkm,
                                                  const float *__restrict__ r, float *__restrict__ ye,
                                                  const SolveScalars *sc) {
  if (sc && sc->done) return;
  constexpr int NE = NSH * NS, NT = (NE + 15) / 16, NK = (NE + 3) / 4;
  static_assert(NT * 16 <= 96, "element too large for the tile set");
  __shared__ float sA[NT * NK * 64];
  const int tid = threadIdx.x, wv = tid >> 6, l = tid & 63, col = l & 15, kq = l >> 4;
  les_stage_G<NE, NT, NK>(G, sA, tid);
  __syncthreads();
  int64_t lo, hi;
  int xi, xg;
  xcd_slab(nvol, lo, hi, xi, xg);
  const int64_t gfirst = lo + (int64_t)xi * 64, gstep = (int64_t)xg * 64;
  auto gather = [&](int64_t g, float (&u)[NK]) {
    const int64_t e = g + wv * 16 + col;
    const bool ev = e < hi;
    const int64_t eo = ev ? e * NSH : 0;
    int32_t bk[NK];
#pragma unroll
    for (int kk = 0; kk < NK; kk++) {
      const int k = 4 * kk + kq;
      bk[kk] = blkm[eo + (k < NE ? k / NS : NSH - 1)];
    }
#pragma unroll
    for (int kk = 0; kk < NK; kk++) {
      const int k = 4 * kk + kq, cp = k % NS;
      const int32_t bm = bk[kk];
      const bool ok = ev && k < NE && bm >= 0 && !((bm >> (28 + cp)) & 1);
      const float v = r[ok ? (int64_t)(bm & 0x0FFFFFFF) * NS + cp : 0];
      u[kk] = ok ? v : 0.f;
    }
  };
  float un[NK];
  if (gfirst < hi) gather(gfirst, un);
  for (int64_t g0 = gfirst; g0 < hi; g0 += gstep) {
    float u[NK];
#pragma unroll
    for (int kk = 0; kk < NK; kk++) u[kk] = un[kk];
    const int64_t e = g0 + wv * 16 + col;
    const bool ev = e < hi;
    const float s = scale[ev ? e : 0];
    float *yo = ye + (ev ? e : 0) * NE;
    if (g0 + gstep < hi) gather(g0 + gstep, un);
    // one row tile at a time (registers: the A operands of a tile come from
    // LDS as the products issue); the element's NE outputs are contiguous
#pragma unroll 1
    for (int ti = 0; ti < NT; ti++) {
      floatx4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kk = 0; kk < NK; kk++)
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(sA[(ti * NK + kk) * 64 + l], u[kk], acc, 0, 0, 0);
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int o = 16 * ti + 4 * kq + i;
        if (ev && o < NE) yo[o] = s * acc[i];
      }
    }
  }
}

// blkm = the element's block per shape with the penalized components as bits
// 28 + k (-1 where the side has no shapes)
__global__ void k_les_blkm(int64_t n, int ns, const int32_t *__restrict__ blk, const float *__restrict__ pen,
                           int32_t *__restrict__ blkm) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int32_t b = blk[t];
  int32_t m = b;
  if (b >= 0)
    for (int k = 0; k < ns; k++)
      if (pen[(int64_t)b * ns + k] != 0.f) m |= 1 << (28 + k);
  blkm[t] = m;
}

// Any other element: one thread per (element, output dof).
template <int NSH, int NS>
__global__ __launch_bounds__(256) void k_les_gen(int64_t nvol, const float *__restrict__ G,
                                                 const float *__restrict__ scale,
                                                 const int32_t *__restrict__ blk,
                                                 const float *__restrict__ pen,
                                                 const float *__restrict__ r, float *__restrict__ ye,
                                                 const SolveScalars *sc) {
  if (sc && sc->done) return;
  constexpr int NE = NSH * NS;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nvol * NE;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t e = t / NE;
    const int o = (int)(t % NE);
    float acc = 0.f;
    for (int k = 0; k < NE; k++) {
      const int32_t b = blk[e * NSH + k / NS];
      if (b < 0) continue;
      const int64_t i = (int64_t)b * NS + k % NS;
      acc += G[o * NE + k] * r[i] * (1.f - pen[i]);
    }
    ye[t] = scale[e] * acc;
  }
}

// eoff[slot] = e*NE + sh*NS for the volume element shape (e, sh) that a
// row-ordered output slot belongs to (-1: a boundary element's slot)
__global__ void k_les_eoff(int64_t n, int nsh, int ns, const int32_t *__restrict__ pos,
                           int32_t *__restrict__ eoff) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int32_t p = pos[t];
  if (p >= 0) eoff[p] = (int32_t)((t / nsh) * nsh * ns + (t % nsh) * ns);
}
__global__ void k_les_fill_i32(int64_t n, int32_t v, int32_t *__restrict__ a) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x)
    a[t] = v;
}

// z = the sum of a block row's element outputs (fixed order of the row's
// slots); r on the penalized equations
template <int NS>
__global__ __launch_bounds__(256) void k_les_gather(int64_t nb, const int64_t *__restrict__ ptr,
                                                    const int32_t *__restrict__ eoff,
                                                    const float *__restrict__ ye,
                                                    const float *__restrict__ pen,
                                                    const float *__restrict__ r, float *__restrict__ z,
                                                    const SolveScalars *sc) {
  if (sc && sc->done) return;
  for (int64_t A = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; A < nb;
       A += (int64_t)gridDim.x * blockDim.x) {
    const int64_t t0 = ptr[A], t1 = ptr[A + 1];
    float acc[NS];
#pragma unroll
    for (int k = 0; k < NS; k++) acc[k] = 0.f;
    for (int64_t t = t0; t < t1; t++) {
      const int32_t o = eoff[t];
      if (o < 0) continue;
#pragma unroll
      for (int k = 0; k < NS; k++) acc[k] += ye[(int64_t)o + k];
    }
#pragma unroll
    for (int k = 0; k < NS; k++) {
      const int64_t i = A * NS + k;
      z[i] = pen[i] != 0.f ? r[i] : acc[k];
    }
  }
}

// The same gather fused with the Chebyshev step of the smoothing that uses it
// (y = M r never stored): mode 0: d = c2 y, x = d;  1: d = c2 y, x += d;
// 2: d = c1 d + c2 y, x += d;  pt = D^-1/2 d when given.
template <int NS>
__global__ __launch_bounds__(256) void k_les_gather_cheb(int64_t nb, const int64_t *__restrict__ ptr,
                                                         const int32_t *__restrict__ eoff,
                                                         const float *__restrict__ ye,
                                                         const float *__restrict__ pen,
                                                         const float *__restrict__ r, LesCheb ch,
                                                         const SolveScalars *sc) {
  if (sc && sc->done) return;
  for (int64_t A = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; A < nb;
       A += (int64_t)gridDim.x * blockDim.x) {
    const int64_t t0 = ptr[A], t1 = ptr[A + 1];
    float acc[NS];
#pragma unroll
    for (int k = 0; k < NS; k++) acc[k] = 0.f;
    for (int64_t t = t0; t < t1; t++) {
      const int32_t o = eoff[t];
      if (o < 0) continue;
#pragma unroll
      for (int k = 0; k < NS; k++) acc[k] += ye[(int64_t)o + k];
    }
#pragma unroll
    for (int k = 0; k < NS; k++) {
      const int64_t i = A * NS + k;
      const float y = pen[i] != 0.f ? r[i] : acc[k];
      const float dv = ch.mode == 2 ? ch.c1 * ch.d[i] + ch.c2 * y : ch.c2 * y;
      ch.d[i] = dv;
      ch.x[i] = ch.mode == 0 ? dv : ch.x[i] + dv;
      if (ch.pt) ch.pt[i] = dv * ch.dis[i];
    }
  }
}

// ---- tiled application on hexahedral lattice meshes -------------------------------
// Elements of order P (1 or 2) on a lattice of cells, all with the same
// shape -> node table: shape k of the element in cell (I, J, K) sits at the
// fine-lattice node (P I, P J, P K) + off(k).  A workgroup takes a tile of
// LT_TC^3 cells at a time: it stages r at the tile's (P LT_TC + 1)^3 fine
// nodes in LDS (zero at penalized dofs), runs the 64 element products
// s_e G r_e as 16 x 16 x 4 FP32 MFMA tiles with B operands from LDS (16
// elements per wavefront), keeps the scaled outputs per element in LDS and
// sums them per node in a fixed order.  Nodes inside the tile box are
// complete and leave through the output step; nodes on the box faces leave
// as per-tile partial sums, which k_les_halo adds up in tile order.  Global
// loads are issued a phase ahead: the next tile's node map and r during this
// tile's products, the output operands of this tile's inner nodes before them.
constexpr int LT_TC = 4;
// tiles in flight per workgroup (k_les_tile): three workgroups of one tile
// per CU ran the 64^3 application in 72-83 us, one of three tiles (one G
// copy per CU, barriers over twelve waves) in 100-121 us
constexpr int LES_NGRP = 1;

struct LesTileArgs {
  int n[3], nt[3];          // cells, tiles per direction
  const int32_t *cmap;      // cell -> element (-1: none)
  const float *tscale;      // per tile and element slot: the element's scale (0: no element)
  const int32_t *tmap;      // per tile and box node: block | penalized components as bits 28 + k (-1: none)
  float *part;              // per tile and box node: partial sums of face nodes
  int64_t nhalo;            // face nodes of the lattice (block order)
  const int32_t *hmap;      // their block | penalty bits
  const int32_t *hptr, *hsrc;  // their partial sums: part[(tile NB + q) NS ...] in tile order
  int16_t off[27];          // shape -> box offset ox + BN (oy + BN oz)
  int16_t inv[27];          // local node (lx + (P+1)(ly + (P+1) lz)) -> shape
};

// Output of a row: z = y (mode -1) or the Chebyshev step of LesCheb.
struct LesOutArgs {
  int mode;
  float c1, c2;
  float *z, *d, *x, *pt;
  const float *dis;
};

// MODE: out.mode (-1: z = y; 0, 1, 2: the Chebyshev steps), fixed per launch.
// Loads are branch-free (clamped addresses, results selected afterwards), so
// each batch of them is one round trip to memory.
// NGRP tiles at a time per workgroup (NGRP groups of four wavefronts sharing
// the staged G; every group runs the same number of rounds, the barriers are
// the workgroup's).
template <int P, int NS, int MODE, int NGRP>
__global__ __launch_bounds__(256 * NGRP) void k_les_tile(LesTileArgs a, const float *__restrict__ G,
                                                  const float *__restrict__ scale,
                                                  const float *__restrict__ r, LesOutArgs out,
                                                  const SolveScalars *sc) {
  if (sc && sc->done) return;
  constexpr int NSH = (P + 1) * (P + 1) * (P + 1), NE = NSH * NS, NT = (NE + 15) / 16, NK = (NE + 3) / 4;
  constexpr int W = P * LT_TC, BN = W + 1, NB = BN * BN * BN, NEL = LT_TC * LT_TC * LT_TC;
  constexpr int NQ = (NB + 255) / 256;  // box nodes per thread
  static_assert(NEL == 64, "four wavefronts of 16 elements");
  // r at the box nodes (sR) and the scaled element outputs (sY) share one
  // array: the B operands are in registers before the outputs are written
  constexpr int NRY = NB * NS > NEL * NE ? NB * NS : NEL * NE;
  __shared__ float sA[NT * NK * 64];
  __shared__ float sRY[NGRP][NRY];
  __shared__ int16_t sInv[NSH];
  const int grp = threadIdx.x >> 8;  // this thread's tile group
  float *sR = sRY[grp], *sY = sRY[grp];
  const int tid = threadIdx.x & 255, wv = tid >> 6, l = tid & 63, col = l & 15, kq = l >> 4;
  les_stage_G<NE, NT, NK, 256 * NGRP>(G, sA, threadIdx.x);
  if (threadIdx.x < NSH) sInv[threadIdx.x] = a.inv[threadIdx.x];
  // this lane's B-operand offsets into sR relative to its element's cell
  int uo[NK];
#pragma unroll
  for (int kk = 0; kk < NK; kk++) {
    const int k = 4 * kk + kq;
    uo[kk] = k < NE ? a.off[k / NS] * NS + k % NS : -1;
  }
  const int le = wv * 16 + col;  // this lane's element of the tile
  const int cx = le % LT_TC, cy = (le / LT_TC) % LT_TC, cz = le / (LT_TC * LT_TC);
  const int cb = (P * cx + BN * (P * cy + BN * P * cz)) * NS;
  // this thread's box nodes: whether they are inner, and the sY offsets of
  // the (element, shape) outputs summed into each (cells in z, y, x order;
  // the same for every tile), packed two per register
  __syncthreads();  // sInv staged
  bool inner[NQ];
  int ncnt[NQ];
  uint32_t noff[NQ][4];
#pragma unroll
  for (int j = 0; j < NQ; j++) {
    const int q = tid + 256 * j;
    const int b[3] = {q % BN, (q / BN) % BN, q / (BN * BN)};
    inner[j] = q < NB && b[0] > 0 && b[0] < W && b[1] > 0 && b[1] < W && b[2] > 0 && b[2] < W;
    int c0[3], c1[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
      c1[d] = min(b[d] / P, LT_TC - 1);
      c0[d] = (b[d] % P == 0 && b[d] > 0) ? b[d] / P - 1 : c1[d];
    }
    int cnt = 0;
#pragma unroll
    for (int h = 0; h < 4; h++) noff[j][h] = 0u;
    if (q < NB)
      for (int z = c0[2]; z <= c1[2]; z++)
        for (int y = c0[1]; y <= c1[1]; y++)
          for (int x = c0[0]; x <= c1[0]; x++) {
            const int lx = b[0] - P * x, ly = b[1] - P * y, lz = b[2] - P * z;
            const int sh = sInv[lx + (P + 1) * (ly + (P + 1) * lz)];
            const uint32_t o = (uint32_t)((x + LT_TC * (y + LT_TC * z)) * NE + sh * NS);
            noff[j][cnt >> 1] |= o << (16 * (cnt & 1));
            cnt++;
          }
    ncnt[j] = cnt;
  }
  const int64_t ntile = (int64_t)a.nt[0] * a.nt[1] * a.nt[2];
  int64_t lo, hi;
  int xi, xg;
  xcd_slab(ntile, lo, hi, xi, xg);
  // loads run one tile ahead of their use: at the top of a tile, r of the
  // next tile (its node map came one tile earlier), the node map of the tile
  // after that and the next tile's element scales are issued.  Loaded values
  // are selected only where they are used, so no load is waited for early.
  auto load_map = [&](int64_t tile, int32_t (&mm)[NQ], bool (&vv)[NQ]) {
#pragma unroll
    for (int j = 0; j < NQ; j++) {
      const int q = tid + 256 * j;
      vv[j] = tile < hi && q < NB;
      mm[j] = a.tmap[vv[j] ? tile * NB + q : 0];
    }
  };
  auto load_r = [&](const int32_t (&mm)[NQ], const bool (&vv)[NQ], float (&rr)[NQ][NS]) {
#pragma unroll
    for (int j = 0; j < NQ; j++)
#pragma unroll
      for (int k = 0; k < NS; k++)
        rr[j][k] = r[(int64_t)(vv[j] && mm[j] >= 0 ? mm[j] & 0x0FFFFFFF : 0) * NS + k];
  };
  auto scale_of = [&](int64_t tile) {  // this lane's element's scale (0: no element)
    return a.tscale[(tile < hi ? tile : 0) * NEL + le];
  };
  int32_t mq[NQ], mn[NQ];
  bool vq[NQ], vn[NQ];
  float rq[NQ][NS];
  const int64_t first = lo + (int64_t)xi * NGRP + grp, step = (int64_t)xg * NGRP;
  load_map(first, mq, vq);
  load_r(mq, vq, rq);
  load_map(first + step, mn, vn);
  float s_cur = scale_of(first);
  // rounds: every group takes its tile of the round (or none past the slab)
  for (int64_t t0 = lo + (int64_t)xi * NGRP; t0 < hi; t0 += step) {
    const int64_t tile = t0 + grp;
    const bool active = tile < hi;
    __syncthreads();  // the previous tile's LDS reads are done (and sA is staged)
    int32_t m[NQ];
    float rv[NQ][NS];
#pragma unroll
    for (int j = 0; j < NQ; j++) {
      const int q = tid + 256 * j;
      m[j] = vq[j] ? mq[j] : -1;
#pragma unroll
      for (int k = 0; k < NS; k++) {
        rv[j][k] = m[j] >= 0 ? rq[j][k] : 0.f;
        if (q < NB) sR[q * NS + k] = ((m[j] >> (28 + k)) & 1) ? 0.f : rv[j][k];
      }
    }
    const float s_el = s_cur;
    load_r(mn, vn, rq);  // next tile
#pragma unroll
    for (int j = 0; j < NQ; j++) {
      mq[j] = mn[j];
      vq[j] = vn[j];
    }
    load_map(tile + 2 * step, mn, vn);
    s_cur = scale_of(tile + step);
    // output operands of the inner nodes
    float od[NQ][NS], ox[NQ][NS], os[NQ][NS];
#pragma unroll
    for (int j = 0; j < NQ; j++)
#pragma unroll
      for (int k = 0; k < NS; k++) {
        const int64_t i = (int64_t)(inner[j] && m[j] >= 0 ? m[j] & 0x0FFFFFFF : 0) * NS + k;
        od[j][k] = MODE == 2 ? out.d[i] : 0.f;
        ox[j][k] = MODE >= 1 ? out.x[i] : 0.f;
        os[j][k] = MODE >= 0 && out.pt ? out.dis[i] : 0.f;
      }
    __syncthreads();
    {
      const float s = s_el;
      float u[NK];
#pragma unroll
      for (int kk = 0; kk < NK; kk++) u[kk] = uo[kk] >= 0 ? sR[cb + uo[kk]] : 0.f;
      __syncthreads();  // every B operand read before sY overwrites sR
#pragma unroll 1
      for (int ti = 0; ti < NT; ti++) {
        floatx4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kk = 0; kk < NK; kk++)
          acc = __builtin_amdgcn_mfma_f32_16x16x4f32(sA[(ti * NK + kk) * 64 + l], u[kk], acc, 0, 0, 0);
#pragma unroll
        for (int i = 0; i < 4; i++) {
          const int o = 16 * ti + 4 * kq + i;
          if (o < NE) sY[le * NE + o] = s * acc[i];
        }
      }
    }
    __syncthreads();
    // per box node: the sum over the tile's elements holding it (cells in
    // z, y, x order)
#pragma unroll
    for (int j = 0; j < NQ; j++) {
      const int q = tid + 256 * j;
      if (q >= NB) continue;
      float acc[NS];
#pragma unroll
      for (int k = 0; k < NS; k++) acc[k] = 0.f;
#pragma unroll
      for (int h = 0; h < 8; h++) {
        if (h >= ncnt[j]) break;
        const float *yy = sY + ((noff[j][h >> 1] >> (16 * (h & 1))) & 0xFFFFu);
#pragma unroll
        for (int k = 0; k < NS; k++) acc[k] += yy[k];
      }
      if (inner[j]) {
        if (m[j] >= 0) {
          const int64_t B = m[j] & 0x0FFFFFFF;
#pragma unroll
          for (int k = 0; k < NS; k++) {
            const int64_t i = B * NS + k;
            const float y = ((m[j] >> (28 + k)) & 1) ? rv[j][k] : acc[k];
            if (MODE < 0) {
              out.z[i] = y;
            } else {
              const float dv = MODE == 2 ? out.c1 * od[j][k] + out.c2 * y : out.c2 * y;
              out.d[i] = dv;
              out.x[i] = MODE == 0 ? dv : ox[j][k] + dv;
              if (out.pt) out.pt[i] = dv * os[j][k];
            }
          }
        }
      } else if (active) {
        float *pp = a.part + (tile * NB + q) * NS;
#pragma unroll
        for (int k = 0; k < NS; k++) pp[k] = acc[k];
      }
    }
  }
}

// Fine nodes on tile faces, in block order: the partial sums of the tiles
// holding them, in tile order, then the output step (MODE as in k_les_tile;
// the output operands are loaded before the partial sums, with them in flight).
template <int NS, int MODE>
__global__ __launch_bounds__(256) void k_les_halo(LesTileArgs a, const float *__restrict__ r, LesOutArgs out,
                                                  const SolveScalars *sc) {
  if (sc && sc->done) return;
  for (int64_t h = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; h < a.nhalo;
       h += (int64_t)gridDim.x * blockDim.x) {
    const int32_t m = a.hmap[h];
    const int32_t p0 = a.hptr[h], p1 = a.hptr[h + 1];
    const int64_t B = m & 0x0FFFFFFF;
    float od[NS], ox[NS], os[NS], orr[NS];
#pragma unroll
    for (int k = 0; k < NS; k++) {
      const int64_t i = B * NS + k;
      od[k] = MODE == 2 ? out.d[i] : 0.f;
      ox[k] = MODE >= 1 ? out.x[i] : 0.f;
      os[k] = MODE >= 0 && out.pt ? out.dis[i] : 0.f;
      orr[k] = r[i];
    }
    float acc[NS];
#pragma unroll
    for (int k = 0; k < NS; k++) acc[k] = 0.f;
    for (int32_t p = p0; p < p1; p++) {
      const float *pp = a.part + (int64_t)a.hsrc[p] * NS;
#pragma unroll
      for (int k = 0; k < NS; k++) acc[k] += pp[k];
    }
#pragma unroll
    for (int k = 0; k < NS; k++) {
      const int64_t i = B * NS + k;
      const float y = ((m >> (28 + k)) & 1) ? orr[k] : acc[k];
      if (MODE < 0) {
        out.z[i] = y;
      } else {
        const float dv = MODE == 2 ? out.c1 * od[k] + out.c2 * y : out.c2 * y;
        out.d[i] = dv;
        out.x[i] = MODE == 0 ? dv : ox[k] + dv;
        if (out.pt) out.pt[i] = dv * os[k];
      }
    }
  }
}

// tscale[tile NEL + slot] = the scale of the element in that cell of the tile
__global__ void k_les_tscale(int64_t ntile, int nt0, int nt1, int n0, int n1, int n2,
                             const int32_t *__restrict__ cmap, const float *__restrict__ s,
                             float *__restrict__ ts) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntile * 64) return;
  const int64_t tile = t / 64;
  const int le = (int)(t % 64);
  const int tx = (int)(tile % nt0), ty = (int)((tile / nt0) % nt1), tz = (int)(tile / ((int64_t)nt0 * nt1));
  const int ex = LT_TC * tx + le % LT_TC, ey = LT_TC * ty + (le / LT_TC) % LT_TC, ez = LT_TC * tz + le / (LT_TC * LT_TC);
  int32_t e = -1;
  if (ex < n0 && ey < n1 && ez < n2) e = cmap[ex + (int64_t)n0 * (ey + (int64_t)n1 * ez)];
  ts[t] = e >= 0 ? s[e] : 0.f;
}

// tmap / hmap = the node maps with the penalized components of each block
// as bits 28 + k
__global__ void k_les_penmap(int64_t n, int ns, const int32_t *__restrict__ src, const float *__restrict__ pen,
                             int32_t *__restrict__ dst) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int32_t b = src[t];
  int32_t m = b;
  if (b >= 0)
    for (int k = 0; k < ns; k++)
      if (pen[(int64_t)b * ns + k] != 0.f) m |= 1 << (28 + k);
  dst[t] = m;
}


// ---- training kernels ------------------------------------------------------------
// S_e (FP64, ne x ne, row-major) of every element from the FP64 full block
// CSR of K scaled by D^-1/2; rows / columns of penalized dofs and of absent
// shapes become unit rows.  One thread per (element, shape a, shape b).
template <int NS>
__global__ void k_les_local(int64_t nvol, int nsh, const int32_t *__restrict__ blk,
                            const int64_t *__restrict__ fbia, const int32_t *__restrict__ fbja,
                            const double *__restrict__ fa, const double *__restrict__ dis,
                            const float *__restrict__ pen, double *__restrict__ Se) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nvol * nsh * nsh) return;
  const int64_t e = t / (nsh * nsh);
  const int a = (int)(t % (nsh * nsh)) / nsh, b = (int)(t % nsh);
  const int ne = nsh * NS;
  double *out = Se + e * ne * ne;
  const int32_t A = blk[e * nsh + a], B = blk[e * nsh + b];
  int64_t idx = -1, n = 0, r0 = 0;
  if (A >= 0 && B >= 0) {
    r0 = fbia[A];
    n = fbia[A + 1] - r0;
    int64_t lo = 0, hi = n;  // columns of a full block row are ascending
    while (lo < hi) {
      const int64_t m = (lo + hi) / 2;
      if (fbja[r0 + m] < B) lo = m + 1;
      else hi = m;
    }
    if (lo < n && fbja[r0 + lo] == B) idx = lo;
  }
  for (int ka = 0; ka < NS; ka++)
    for (int kb = 0; kb < NS; kb++) {
      const int i = a * NS + ka, j = b * NS + kb;
      double v = 0.;
      bool unit = A < 0 || B < 0;
      if (!unit) {
        const int64_t ia = (int64_t)A * NS + ka, ib = (int64_t)B * NS + kb;
        unit = pen[ia] != 0.f || pen[ib] != 0.f;
        if (!unit && idx >= 0) v = dis[ia] * fa[bsr_pos<NS * NS>(r0, n, idx, ka * NS + kb)] * dis[ib];
      }
      if (unit) v = i == j ? 1. : 0.;
      out[(int64_t)i * ne + j] = v;
    }
}

// Cholesky of S_e in LDS and E = S_e^-1 R for the element's nr residuals
// (one workgroup per element; ne <= LES_MAXNE).
constexpr int LES_MAXNE = 81;
__global__ __launch_bounds__(256) void k_les_solve(int ne, int nr, const double *__restrict__ Se,
                                                   const float *__restrict__ R, float *__restrict__ E,
                                                   int *__restrict__ fail) {
  __shared__ double L[LES_MAXNE * LES_MAXNE];
  __shared__ double y[8][LES_MAXNE];
  const int64_t e = blockIdx.x;
  const int tid = threadIdx.x;
  for (int t = tid; t < ne * ne; t += blockDim.x) L[t] = Se[e * ne * ne + t];
  __syncthreads();
  for (int j = 0; j < ne; j++) {
    if (tid == 0) {
      const double d = L[j * ne + j];
      if (!(d > 0.)) atomicExch(fail, 1);
      L[j * ne + j] = d > 0. ? sqrt(d) : 1.;
    }
    __syncthreads();
    const double pj = L[j * ne + j];
    for (int i = j + 1 + tid; i < ne; i += blockDim.x) L[i * ne + j] /= pj;
    __syncthreads();
    const int m = ne - j - 1;
    for (int t = tid; t < m * m; t += blockDim.x) {
      const int i = j + 1 + t / m, k = j + 1 + t % m;
      if (k <= i) L[i * ne + k] -= L[i * ne + j] * L[k * ne + j];
    }
    __syncthreads();
  }
  // forward / back substitution, one thread per right-hand side
  if (tid < nr && tid < 8) {
    const float *rr = R + (e * nr + tid) * ne;
    double *yy = y[tid];
    for (int i = 0; i < ne; i++) {
      double s = rr[i];
      for (int k = 0; k < i; k++) s -= L[i * ne + k] * yy[k];
      yy[i] = s / L[i * ne + i];
    }
    for (int i = ne - 1; i >= 0; i--) {
      double s = yy[i];
      for (int k = i + 1; k < ne; k++) s -= L[k * ne + i] * yy[k];
      yy[i] = s / L[i * ne + i];
    }
    float *ee = E + (e * nr + tid) * ne;
    for (int i = 0; i < ne; i++) ee[i] = (float)yy[i];
  }
}

// per element: y_k = G r_k, d_k = mask (s y_k - e_k); loss_e = sum |d|^2,
// ds_e = sum d.y; Dsc = 2 s d / den (the G gradient rows).  One workgroup
// per element.
__global__ __launch_bounds__(128) void k_les_fwd(int ne, int nr, const float *__restrict__ G,
                                                 const float *__restrict__ scale,
                                                 const float *__restrict__ M, const float *__restrict__ R,
                                                 const float *__restrict__ E, double den,
                                                 float *__restrict__ Dsc, double *__restrict__ loss_e,
                                                 double *__restrict__ ds_e) {
  __shared__ float sr[LES_MAXNE];
  __shared__ double red[2][128];
  const int64_t e = blockIdx.x;
  const int tid = threadIdx.x;
  const float s = scale[e];
  double lacc = 0., dacc = 0.;
  for (int k = 0; k < nr; k++) {
    __syncthreads();
    for (int t = tid; t < ne; t += blockDim.x) sr[t] = R[(e * nr + k) * ne + t];
    __syncthreads();
    for (int i = tid; i < ne; i += blockDim.x) {
      float y = 0.f;
      for (int j = 0; j < ne; j++) y += G[i * ne + j] * sr[j];
      const float m = M[e * ne + i];
      const double d = m * ((double)s * y - (double)E[(e * nr + k) * ne + i]);
      lacc += d * d;
      dacc += d * y;
      Dsc[(e * nr + k) * ne + i] = (float)(2.0 * s * d / den);
    }
  }
  red[0][tid] = lacc;
  red[1][tid] = dacc;
  __syncthreads();
  for (int w = 64; w > 0; w >>= 1) {
    if (tid < w) {
      red[0][tid] += red[0][tid + w];
      red[1][tid] += red[1][tid + w];
    }
    __syncthreads();
  }
  if (tid == 0) {
    loss_e[e] = red[0][0];
    ds_e[e] = red[1][0];
  }
}

// dZ_e = s_e * 2 ds_e / den (the scale is exp(z) inside the clamp)
__global__ void k_les_dz(int64_t n, const float *__restrict__ z, const float *__restrict__ s,
                         const double *__restrict__ ds, double den, float *__restrict__ dz) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const bool inside = z[i] > -4.f && z[i] < 4.f;
  dz[i] = inside ? (float)(2.0 * (double)s[i] * ds[i] / den) : 0.f;
}

// symmetric part of the G gradient (G stays symmetric under Adam)
__global__ void k_les_sym(int ne, float *__restrict__ g) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ne * ne) return;
  const int i = t / ne, j = t % ne;
  if (j <= i) return;
  const float m = 0.5f * (g[i * ne + j] + g[j * ne + i]);
  g[i * ne + j] = m;
  g[j * ne + i] = m;
}

// fixed-order sum of per-element doubles (one block)
__global__ void k_les_sum(int64_t n, const double *__restrict__ v, double *__restrict__ out) {
  __shared__ double red[256];
  double a = 0.;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) a += v[i];
  red[threadIdx.x] = a;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

// ============================================================================
// host side
// PZG_LES_TILES=0 keeps the element-major kernels on lattice meshes too
static bool les_tiles_on() {
  static const bool on = [] {
    const char *v = std::getenv("PZG_LES_TILES");
    return !(v && v[0] == '0');
  }();
  return on;
}

static int grid1(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>((n + TPB - 1) / TPB, 1 << 20)); }

static bool les_supported(const Ctx &c) {
  return c.vol_etype != PZG_TET && !c.hp_simplex && c.vol_rules.size() == 1 && !c.has_dep && !c.filt &&
         c.tvol.nsh * c.bs <= LES_MAXNE;
}

// Element graph (elements sharing a face: a side with as many corners as a
// boundary element), the penalized-equation mask and the per-element
// penalized fraction; kept per mesh.
static int les_graph(Ctx &c) {
  if (c.les_graph) return PZG_OK;
  const int nsd = c.vol_nsides, ncv = c.vol_ncorner, nfc = c.bc_ncorner;
  std::vector<int> faces;
  for (int sd = 0; sd < nsd; sd++) {
    int loc[8];
    if (side_corner_list_pub(c.vol_etype, sd, loc) == nfc) faces.push_back(sd);
  }
  // face connect -> the elements holding it
  std::map<int64_t, std::vector<int64_t>> byface;
  for (int64_t e = 0; e < c.nvol; e++)
    for (int sd : faces) byface[c.vol_conn[e * nsd + sd]].push_back(e);
  std::vector<std::vector<int32_t>> nb(c.nvol);
  for (auto &kv : byface)
    for (size_t i = 0; i < kv.second.size(); i++)
      for (size_t j = 0; j < kv.second.size(); j++)
        if (i != j) nb[kv.second[i]].push_back((int32_t)kv.second[j]);
  c.les_cia.assign(c.nvol + 1, 0);
  c.les_cja.clear();
  for (int64_t e = 0; e < c.nvol; e++) {
    auto &v = nb[e];
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
    c.les_cja.insert(c.les_cja.end(), v.begin(), v.end());
    c.les_cia[e + 1] = (int64_t)c.les_cja.size();
  }
  c.les_nfaces = (int)faces.size();
  // penalized equations: the components a boundary element holds with a
  // penalty (a diagonal coefficient of the penalty size)
  const int ns = c.bs;
  c.les_pen.assign(c.neq, 0.f);
  const ElemTables &B = c.tbc;
  for (int64_t e = 0; e < c.nbc; e++) {
    const double *ck = &c.bc_coef[(size_t)BCC * c.bc_id[e] + BCC_CK];
    for (int k = 0; k < ns; k++) {
      if (!(std::fabs(ck[(ns + 1) * k]) >= 1e6)) continue;
      for (int s = 0; s < B.nsh; s++) {
        const int64_t a = c.aidx[c.bc_conn[e * c.bc_nsides + B.shape_side[s]]];
        if (a >= 0) c.les_pen[a * ns + k] = 1.f;
      }
    }
  }
  c.les_penfrac.assign(c.nvol, 0.f);
  const ElemTables &V = c.tvol;
  for (int64_t e = 0; e < c.nvol; e++) {
    int np = 0, nt = 0;
    for (int s = 0; s < V.nsh; s++) {
      const int64_t a = c.aidx[c.vol_conn[e * nsd + V.shape_side[s]]];
      if (a < 0) continue;
      for (int k = 0; k < ns; k++, nt++) np += c.les_pen[a * ns + k] != 0.f;
    }
    c.les_penfrac[e] = nt ? (float)np / nt : 0.f;
  }
  hipStream_t s = c.stream;
  PZG_TRY(dev_upload(c.d_les_cia, c.les_cia.data(), c.les_cia.size(), s));
  std::vector<int32_t> cja(c.les_cja);
  if (cja.empty()) cja.push_back(0);
  PZG_TRY(dev_upload(c.d_les_cja, cja.data(), cja.size(), s));
  PZG_TRY(dev_upload(c.d_les_pen, c.les_pen.data(), c.les_pen.size(), s));
  PZG_HIP(hipStreamSynchronize(s));
  c.les_graph = true;
  return PZG_OK;
}

// features of the current coefficient field (host) -> d_les_X
static int les_features(Ctx &c) {
  std::vector<float> X((size_t)c.nvol * LES_FIN, 0.f);
  const double dflt = c.material == PZG_MAT_ELAST3D ? c.E : c.scale;
  auto lc = [&](int64_t e) { return std::log(c.vol_coef.empty() ? dflt : c.vol_coef[e]); };
  for (int64_t e = 0; e < c.nvol; e++) {
    const int64_t j0 = c.les_cia[e], j1 = c.les_cia[e + 1];
    const double le = lc(e);
    double mean = 0., mx = 0.;
    for (int64_t j = j0; j < j1; j++) {
      const double ln = lc(c.les_cja[j]);
      mean += ln;
      mx = std::max(mx, std::fabs(ln - le));
    }
    const int64_t nn = j1 - j0;
    X[e * LES_FIN + 0] = nn ? (float)(le - mean / nn) : 0.f;
    X[e * LES_FIN + 1] = (float)mx;
    X[e * LES_FIN + 2] = c.les_penfrac[e];
    X[e * LES_FIN + 3] = c.les_nfaces ? 1.f - (float)nn / c.les_nfaces : 0.f;
  }
  PZG_TRY(dev_upload(c.d_les_X, X.data(), X.size(), c.stream));
  return PZG_OK;
}

// GNN forward over the element graph; keep: the intermediate tensors for
// the backward pass (training)
static int les_forward(Ctx &c, bool keep) {
  const int64_t rows = c.nvol;
  const float *th = c.d_les_th.as<float>();
  LesOff o = les_offsets(c.les_ne);
  const size_t need = (size_t)rows * NNF;
  PZG_TRY(dev_alloc(c.d_nn_trH, sizeof(float) * need * (NNL + 1)));
  PZG_TRY(dev_alloc(c.d_nn_trM, sizeof(float) * need * NNL));
  PZG_TRY(dev_alloc(c.d_nn_trP, sizeof(float) * need * NNL));
  PZG_TRY(dev_alloc(c.d_les_z, sizeof(float) * std::max<int64_t>(rows, 1)));
  PZG_TRY(dev_alloc(c.d_les_s, sizeof(float) * std::max<int64_t>(rows, 1)));
  float *H = c.d_nn_trH.as<float>(), *M = c.d_nn_trM.as<float>(), *P = c.d_nn_trP.as<float>();
  const int g = grid1(rows * NNF);
  hipLaunchKernelGGL(k_nn_in<LES_FIN>, dim3(g), dim3(TPB), 0, c.stream, rows, c.d_les_X.as<float>(),
                     th + o.w0, H);
  for (int l = 0; l < NNL; l++) {
    hipLaunchKernelGGL(k_nn_agg, dim3(g), dim3(TPB), 0, c.stream, rows, rows, c.d_les_cia.as<int64_t>(),
                       c.d_les_cja.as<int32_t>(), H + (int64_t)l * need, M + (int64_t)l * need);
    hipLaunchKernelGGL(k_nn_layer, dim3(g), dim3(TPB), 0, c.stream, rows, H + (int64_t)l * need,
                       M + (int64_t)l * need, th + o.a[l], th + o.b[l], keep ? P + (int64_t)l * need : nullptr,
                       H + (int64_t)(l + 1) * need);
  }
  hipLaunchKernelGGL(k_nn_out<1>, dim3(grid1(rows)), dim3(TPB), 0, c.stream, rows, H + (int64_t)NNL * need,
                     th + o.wout, c.d_les_z.as<float>());
  hipLaunchKernelGGL(k_les_exp, dim3(grid1(rows)), dim3(TPB), 0, c.stream, rows, c.d_les_z.as<float>(),
                     c.d_les_s.as<float>());
  if (c.les_tile_ok) {  // the scales in tile order for k_les_tile
    int nt[3];
    for (int d = 0; d < 3; d++) nt[d] = (c.les_tn[d] + LT_TC - 1) / LT_TC;
    const int64_t ntile = (int64_t)nt[0] * nt[1] * nt[2];
    PZG_TRY(dev_alloc(c.d_les_tscale, sizeof(float) * ntile * 64));
    hipLaunchKernelGGL(k_les_tscale, dim3(grid1(ntile * 64)), dim3(TPB), 0, c.stream, ntile, nt[0], nt[1],
                       c.les_tn[0], c.les_tn[1], c.les_tn[2], c.d_les_cmap.as<int32_t>(), c.d_les_s.as<float>(),
                       c.d_les_tscale.as<float>());
  }
  PZG_HIP(hipGetLastError());
  return PZG_OK;
}

int les_setup(Ctx &c, uint64_t seed) {
  if (!c.has_struct || c.external_csr || !c.has_mesh) {
    set_error("learned preconditioner: needs a mesh pattern (pzg_struct_create)");
    return PZG_ESTATE;
  }
  if (!les_supported(c)) {
    set_error("learned preconditioner: quadrilateral / hexahedral meshes with one rule, no constraints "
              "and no equation filter");
    return PZG_EUNSUP;
  }
  const int ne = c.tvol.nsh * c.bs;
  c.les_ne = ne;
  LesOff o = les_offsets(ne);
  std::vector<float> th(o.n, 0.f);
  // G = diag(1 / elements sharing the side of the dof's shape), the sharing
  // of an interior element of a structured mesh: 2^(dim - side dimension)
  for (int s = 0; s < c.tvol.nsh; s++) {
    int loc[8];
    const int nc = side_corner_list_pub(c.vol_etype, c.tvol.shape_side[s], loc);
    const int sdim = nc == 1 ? 0 : nc == 2 ? 1 : nc == 4 ? 2 : 3;
    const float mult = (float)(1 << std::max(0, c.dim - sdim));
    for (int k = 0; k < c.bs; k++) th[o.g + (s * c.bs + k) * ne + s * c.bs + k] = 1.f / mult;
  }
  std::mt19937_64 rng(seed);
  std::normal_distribution<float> nd(0.f, 1.f);
  for (int k = 0; k < LES_FIN * NNF; k++) th[o.w0 + k] = nd(rng) / std::sqrt((float)LES_FIN);
  for (int l = 0; l < NNL; l++)
    for (int k = 0; k < NNF * NNF; k++) {
      th[o.a[l] + k] = nd(rng) / std::sqrt(2.f * NNF);
      th[o.b[l] + k] = nd(rng) / std::sqrt(2.f * NNF);
    }
  hipStream_t s = c.stream;
  PZG_TRY(dev_upload(c.d_les_th, th.data(), th.size(), s));
  std::vector<float> zero(o.n, 0.f);
  PZG_TRY(dev_upload(c.d_les_m, zero.data(), zero.size(), s));
  PZG_TRY(dev_upload(c.d_les_v, zero.data(), zero.size(), s));
  PZG_TRY(dev_alloc(c.d_les_g, sizeof(float) * o.n));
  PZG_HIP(hipStreamSynchronize(s));
  c.les_adam_t = 0;
  c.les_nr = 0;
  c.les_A.clear();
  c.les_B.clear();
  c.les_fit_rows = 0;
  c.les_fit_stale = false;
  c.les_ready = true;
  c.mg_ready = false;  // a smoother of the cycle changed
  return PZG_OK;
}

// per-element scales of the current assembly and the apply buffers
// Lattice structure for k_les_tile (once per mesh): every volume element a
// hexahedron of order 1 or 2 filling one cell of the lattice of its corner
// coordinates, all with the same shape -> node table, connects shared
// consistently.  Anything else keeps the element-major kernels.
static int les_tile_structure(Ctx &c) {
  c.les_tile_ready = true;
  c.les_tile_ok = false;
  const int nsh = c.tvol.nsh;
  const int P = nsh == 27 ? 2 : nsh == 8 ? 1 : 0;
  if (c.dim != 3 || c.vol_etype != PZG_HEX || P == 0 || c.vol_ncorner != 8 || (c.bs != 1 && c.bs != 3) ||
      c.nvol == 0)
    return PZG_OK;
  std::vector<double> u[3];
  for (int d = 0; d < 3; d++) {
    std::vector<double> x;
    x.reserve(c.nvol * 8);
    for (int64_t e = 0; e < c.nvol; e++)
      for (int k = 0; k < 8; k++) x.push_back(c.xyz[3 * c.vol_nodes[e * 8 + k] + d]);
    std::sort(x.begin(), x.end());
    const double tol = 1e-9 * std::max(1.0, std::fabs(x.back() - x.front()));
    for (double v : x)
      if (u[d].empty() || v - u[d].back() > tol) u[d].push_back(v);
    if (u[d].size() < 2) return PZG_OK;
  }
  int n[3], fn[3];
  for (int d = 0; d < 3; d++) {
    n[d] = (int)u[d].size() - 1;
    fn[d] = P * n[d] + 1;
  }
  const int64_t ncell = (int64_t)n[0] * n[1] * n[2], nfine = (int64_t)fn[0] * fn[1] * fn[2];
  if (ncell != c.nvol || nfine >= ((int64_t)1 << 31)) return PZG_OK;
  auto index = [&](int d, double v) {
    const double tol = 1e-9 * std::max(1.0, std::fabs(u[d].back() - u[d].front()));
    auto it = std::lower_bound(u[d].begin(), u[d].end(), v - tol);
    return (int)std::min<int64_t>(it - u[d].begin(), (int64_t)u[d].size() - 1);
  };
  std::vector<int32_t> cmap(ncell, -1), fmapb(nfine, -2);
  int16_t off[27], inv[27];
  for (int k = 0; k < 27; k++) inv[k] = -1;
  const int BN = P * LT_TC + 1, NSD = c.vol_nsides;
  for (int64_t e = 0; e < c.nvol; e++) {
    int lo[3], hi[3];
    for (int d = 0; d < 3; d++) {
      lo[d] = INT32_MAX;
      hi[d] = -1;
    }
    for (int k = 0; k < 8; k++)
      for (int d = 0; d < 3; d++) {
        const int i = index(d, c.xyz[3 * c.vol_nodes[e * 8 + k] + d]);
        lo[d] = std::min(lo[d], i);
        hi[d] = std::max(hi[d], i);
      }
    for (int d = 0; d < 3; d++)
      if (hi[d] != lo[d] + 1) return PZG_OK;
    const int64_t cell = lo[0] + (int64_t)n[0] * (lo[1] + (int64_t)n[1] * lo[2]);
    if (cmap[cell] != -1) return PZG_OK;
    cmap[cell] = (int32_t)e;
    for (int s = 0; s < nsh; s++) {
      const int side = c.tvol.shape_side[s];
      int loc[8];
      const int nl = side_corner_list_pub(PZG_HEX, side, loc);
      int o[3];
      for (int d = 0; d < 3; d++) {
        double m = 0.;
        for (int k = 0; k < nl; k++) m += c.xyz[3 * c.vol_nodes[e * 8 + loc[k]] + d];
        m /= nl;
        const double rel = (m - u[d][lo[d]]) / (u[d][lo[d] + 1] - u[d][lo[d]]) * P;
        o[d] = (int)std::lround(rel);
        if (std::fabs(rel - o[d]) > 1e-6 || o[d] < 0 || o[d] > P) return PZG_OK;
      }
      const int16_t bo = (int16_t)(o[0] + BN * (o[1] + BN * o[2]));
      const int16_t lp = (int16_t)(o[0] + (P + 1) * (o[1] + (P + 1) * o[2]));
      if (e == 0) {
        off[s] = bo;
        if (inv[lp] != -1) return PZG_OK;
        inv[lp] = (int16_t)s;
      } else if (off[s] != bo) {
        return PZG_OK;
      }
      const int64_t f = (P * lo[0] + o[0]) + (int64_t)fn[0] * ((P * lo[1] + o[1]) + (int64_t)fn[1] * (P * lo[2] + o[2]));
      const int64_t conn = c.vol_conn[e * NSD + side];
      const int32_t b = c.aidx[conn] >= 0 ? (int32_t)c.aidx[conn] : -1;  // as the matrix-free block map
      if (fmapb[f] == -2) fmapb[f] = b;
      else if (fmapb[f] != b) return PZG_OK;
    }
  }
  for (int k = 0; k < nsh; k++)
    if (inv[k] < 0) return PZG_OK;
  for (int64_t f = 0; f < nfine; f++)
    if (fmapb[f] == -2) fmapb[f] = -1;
  int nt[3];
  for (int d = 0; d < 3; d++) nt[d] = (n[d] + LT_TC - 1) / LT_TC;
  const int W = P * LT_TC, NB = BN * BN * BN;
  const int64_t ntile = (int64_t)nt[0] * nt[1] * nt[2];
  if (ntile * NB >= ((int64_t)1 << 31)) return PZG_OK;
  // per tile and box node: the block (-1 outside the lattice)
  std::vector<int32_t> tmapb(ntile * NB, -1);
  for (int64_t t = 0; t < ntile; t++) {
    const int tx = (int)(t % nt[0]), ty = (int)((t / nt[0]) % nt[1]), tz = (int)(t / ((int64_t)nt[0] * nt[1]));
    for (int q = 0; q < NB; q++) {
      const int fx = W * tx + q % BN, fy = W * ty + (q / BN) % BN, fz = W * tz + q / (BN * BN);
      if (fx < fn[0] && fy < fn[1] && fz < fn[2])
        tmapb[t * NB + q] = fmapb[fx + (int64_t)fn[0] * (fy + (int64_t)fn[1] * fz)];
    }
  }
  // nodes on tile faces (a coordinate a multiple of W), in block order, with
  // the (tile, box node) slots of their partial sums in tile order
  std::vector<std::pair<int32_t, int64_t>> face;
  for (int64_t f = 0; f < nfine; f++) {
    const int fx = (int)(f % fn[0]), fy = (int)((f / fn[0]) % fn[1]), fz = (int)(f / ((int64_t)fn[0] * fn[1]));
    if ((fx % W) && (fy % W) && (fz % W)) continue;
    if (fmapb[f] >= 0) face.push_back({fmapb[f], f});
  }
  std::sort(face.begin(), face.end());
  std::vector<int32_t> hmapb(face.size()), hptr(face.size() + 1, 0), hsrc;
  hsrc.reserve(face.size() * 3);
  for (size_t h = 0; h < face.size(); h++) {
    const int64_t f = face[h].second;
    const int fc[3] = {(int)(f % fn[0]), (int)((f / fn[0]) % fn[1]), (int)(f / ((int64_t)fn[0] * fn[1]))};
    int t0[3], t1[3];
    for (int d = 0; d < 3; d++) {
      const int t = fc[d] / W;
      if (fc[d] % W) {
        t0[d] = t1[d] = t;
      } else {
        t0[d] = t > 0 ? t - 1 : t;
        t1[d] = t < nt[d] ? t : t - 1;
      }
    }
    for (int z = t0[2]; z <= t1[2]; z++)
      for (int y = t0[1]; y <= t1[1]; y++)
        for (int x = t0[0]; x <= t1[0]; x++) {
          const int q = (fc[0] - W * x) + BN * ((fc[1] - W * y) + BN * (fc[2] - W * z));
          hsrc.push_back((int32_t)((((int64_t)z * nt[1] + y) * nt[0] + x) * NB + q));
        }
    hmapb[h] = face[h].first;
    hptr[h + 1] = (int32_t)hsrc.size();
  }
  const int64_t nh = (int64_t)face.size();
  PZG_TRY(dev_alloc(c.d_les_cmap, sizeof(int32_t) * ncell));
  PZG_TRY(dev_alloc(c.d_les_tmapb, sizeof(int32_t) * ntile * NB));
  PZG_TRY(dev_alloc(c.d_les_tmap, sizeof(int32_t) * ntile * NB));
  PZG_TRY(dev_alloc(c.d_les_part, sizeof(float) * ntile * NB * c.bs));
  PZG_TRY(dev_alloc(c.d_les_hmapb, sizeof(int32_t) * std::max<int64_t>(nh, 1)));
  PZG_TRY(dev_alloc(c.d_les_hmap, sizeof(int32_t) * std::max<int64_t>(nh, 1)));
  PZG_TRY(dev_alloc(c.d_les_hptr, sizeof(int32_t) * (nh + 1)));
  PZG_TRY(dev_alloc(c.d_les_hsrc, sizeof(int32_t) * std::max<size_t>(hsrc.size(), 1)));
  PZG_HIP(hipMemcpyAsync(c.d_les_cmap.p, cmap.data(), sizeof(int32_t) * ncell, hipMemcpyHostToDevice, c.stream));
  PZG_HIP(hipMemcpyAsync(c.d_les_tmapb.p, tmapb.data(), sizeof(int32_t) * ntile * NB, hipMemcpyHostToDevice,
                         c.stream));
  if (nh) {
    PZG_HIP(hipMemcpyAsync(c.d_les_hmapb.p, hmapb.data(), sizeof(int32_t) * nh, hipMemcpyHostToDevice, c.stream));
    PZG_HIP(hipMemcpyAsync(c.d_les_hsrc.p, hsrc.data(), sizeof(int32_t) * hsrc.size(), hipMemcpyHostToDevice,
                           c.stream));
  }
  PZG_HIP(hipMemcpyAsync(c.d_les_hptr.p, hptr.data(), sizeof(int32_t) * (nh + 1), hipMemcpyHostToDevice, c.stream));
  PZG_HIP(hipStreamSynchronize(c.stream));
  c.les_nhalo = nh;
  c.les_tP = P;
  for (int d = 0; d < 3; d++) c.les_tn[d] = n[d];
  std::memcpy(c.les_toff, off, sizeof(off));
  std::memcpy(c.les_tinv, inv, sizeof(inv));
  c.les_tile_ok = true;
  return PZG_OK;
}

int les_prepare(Ctx &c) {
  if (!c.les_ready) {
    set_error("learned preconditioner: call pzg_learned_setup first");
    return PZG_ESTATE;
  }
  if (c.tvol.nsh * c.bs != c.les_ne || !les_supported(c)) {
    set_error("learned preconditioner: the mesh's element differs from the trained one");
    return PZG_EUNSUP;
  }
  PZG_TRY(mf_prepare(c));  // element -> block map and row-ordered output slots
  PZG_TRY(les_graph(c));
  if (c.nb >= (1 << 28)) {
    set_error("learned preconditioner: more than 2^28 blocks");
    return PZG_EUNSUP;
  }
  {
    const int64_t nt = c.nvol * (int64_t)c.tvol.nsh;
    PZG_TRY(dev_alloc(c.d_les_blkm, sizeof(int32_t) * std::max<int64_t>(nt, 1)));
    hipLaunchKernelGGL(k_les_blkm, dim3((unsigned)std::max<int64_t>(1, (nt + TPB - 1) / TPB)), dim3(TPB), 0,
                       c.stream, nt, c.bs, c.d_mf_blk.as<int32_t>(), c.d_les_pen.as<float>(),
                       c.d_les_blkm.as<int32_t>());
    PZG_HIP(hipGetLastError());
  }
  if (!c.les_tile_ready) PZG_TRY(les_tile_structure(c));
  if (c.les_tile_ok) {  // the node maps with this system's penalized components
    const int BN = c.les_tP * LT_TC + 1;
    int64_t ntile = 1;
    for (int d = 0; d < 3; d++) ntile *= (c.les_tn[d] + LT_TC - 1) / LT_TC;
    const int64_t nq = ntile * BN * BN * BN;
    hipLaunchKernelGGL(k_les_penmap, dim3((unsigned)((nq + TPB - 1) / TPB)), dim3(TPB), 0, c.stream, nq, c.bs,
                       c.d_les_tmapb.as<int32_t>(), c.d_les_pen.as<float>(), c.d_les_tmap.as<int32_t>());
    if (c.les_nhalo)
      hipLaunchKernelGGL(k_les_penmap, dim3((unsigned)((c.les_nhalo + TPB - 1) / TPB)), dim3(TPB), 0, c.stream,
                         c.les_nhalo, c.bs, c.d_les_hmapb.as<int32_t>(), c.d_les_pen.as<float>(),
                         c.d_les_hmap.as<int32_t>());
    PZG_HIP(hipGetLastError());
  }
  PZG_TRY(les_features(c));
  PZG_TRY(les_forward(c, false));
  // element outputs, element-major, and the slot -> output offset map
  const int64_t nte = c.nvol * (int64_t)c.tvol.nsh;
  if (nte * c.bs >= ((int64_t)1 << 31)) {
    set_error("learned preconditioner: more than 2^31 element outputs");
    return PZG_EUNSUP;
  }
  PZG_TRY(dev_alloc(c.d_les_ye, sizeof(float) * std::max<int64_t>(nte * c.bs, 1)));
  PZG_TRY(dev_alloc(c.d_les_eoff, sizeof(int32_t) * std::max<int64_t>(c.mf_nout, 1)));
  hipLaunchKernelGGL(k_les_fill_i32, dim3(grid1(c.mf_nout)), dim3(TPB), 0, c.stream, c.mf_nout, -1,
                     c.d_les_eoff.as<int32_t>());
  hipLaunchKernelGGL(k_les_eoff, dim3(grid1(nte)), dim3(TPB), 0, c.stream, nte, c.tvol.nsh, c.bs,
                     c.d_mf_pos.as<int32_t>(), c.d_les_eoff.as<int32_t>());
  PZG_HIP(hipGetLastError());
  PZG_HIP(hipStreamSynchronize(c.stream));
  return PZG_OK;
}

// z = M r (learned smoother), both FP32 device vectors of neq entries
int les_apply(Ctx &c, const float *r, float *z, const SolveScalars *sc, const LesCheb *ch) {
  const float *G = c.d_les_th.as<float>() + les_offsets(c.les_ne).g;
  const int nsh = c.tvol.nsh;
  const int32_t *blk = c.d_mf_blk.as<int32_t>(), *pos = c.d_mf_pos.as<int32_t>();
  const int32_t *blkm = c.d_les_blkm.as<int32_t>(), *eoff = c.d_les_eoff.as<int32_t>();
  const float *pen = c.d_les_pen.as<float>(), *s = c.d_les_s.as<float>();
  float *ye = c.d_les_ye.as<float>();
  const int64_t nv = c.nvol;
  if (c.les_tile_ok && les_tiles_on()) {
    LesTileArgs a;
    for (int d = 0; d < 3; d++) {
      a.n[d] = c.les_tn[d];
      a.nt[d] = (c.les_tn[d] + LT_TC - 1) / LT_TC;
    }
    a.cmap = c.d_les_cmap.as<int32_t>();
    a.tscale = c.d_les_tscale.as<float>();
    a.tmap = c.d_les_tmap.as<int32_t>();
    a.part = c.d_les_part.as<float>();
    a.nhalo = c.les_nhalo;
    a.hmap = c.d_les_hmap.as<int32_t>();
    a.hptr = c.d_les_hptr.as<int32_t>();
    a.hsrc = c.d_les_hsrc.as<int32_t>();
    std::memcpy(a.off, c.les_toff, sizeof(a.off));
    std::memcpy(a.inv, c.les_tinv, sizeof(a.inv));
    LesOutArgs o;
    std::memset(&o, 0, sizeof(o));
    o.mode = -1;
    o.z = z;
    if (ch) {
      o.mode = ch->mode;
      o.c1 = ch->c1;
      o.c2 = ch->c2;
      o.d = ch->d;
      o.x = ch->x;
      o.pt = ch->pt;
      o.dis = ch->dis;
    }
    const int64_t ntile = (int64_t)a.nt[0] * a.nt[1] * a.nt[2];
    // 768 / LES_NGRP workgroups: three per CU at one tile each
    const int gt = (int)std::max<int64_t>(1, std::min<int64_t>((ntile + LES_NGRP - 1) / LES_NGRP, 768 / LES_NGRP));
    const int gh = grid1(std::max<int64_t>(a.nhalo, 1));
#define LES_TILE_M(PP, NSV, MD)                                                                             \
  do {                                                                                                    \
    hipLaunchKernelGGL((k_les_tile<PP, NSV, MD, LES_NGRP>), dim3(gt), dim3(256 * LES_NGRP), 0, c.stream, a, G, s, \
                       r, o, sc);                                                                         \
    hipLaunchKernelGGL((k_les_halo<NSV, MD>), dim3(gh), dim3(256), 0, c.stream, a, r, o, sc);               \
  } while (0)
#define LES_TILE(PP, NSV)                                                                                   \
  do {                                                                                                    \
    if (o.mode < 0) LES_TILE_M(PP, NSV, -1);                                                              \
    else if (o.mode == 0) LES_TILE_M(PP, NSV, 0);                                                         \
    else if (o.mode == 1) LES_TILE_M(PP, NSV, 1);                                                         \
    else LES_TILE_M(PP, NSV, 2);                                                                          \
  } while (0)
    if (c.les_tP == 2 && c.bs == 3) LES_TILE(2, 3);
    else if (c.les_tP == 2) LES_TILE(2, 1);
    else if (c.bs == 3) LES_TILE(1, 3);
    else LES_TILE(1, 1);
#undef LES_TILE
#undef LES_TILE_M
    PZG_HIP(hipGetLastError());
    return PZG_OK;
  }
  const int gm = (int)std::max<int64_t>(1, std::min<int64_t>((nv + 63) / 64, 4096));
  if (nsh == 27 && c.bs == 3)
    hipLaunchKernelGGL((k_les_mfma<27, 3>), dim3(gm), dim3(256), 0, c.stream, nv, G, s, blkm, r, ye, sc);
  else if (nsh == 27 && c.bs == 1)
    hipLaunchKernelGGL((k_les_mfma<27, 1>), dim3(gm), dim3(256), 0, c.stream, nv, G, s, blkm, r, ye, sc);
  else if (nsh == 8 && c.bs == 3)
    hipLaunchKernelGGL((k_les_mfma<8, 3>), dim3(gm), dim3(256), 0, c.stream, nv, G, s, blkm, r, ye, sc);
  else if (nsh == 8 && c.bs == 1)
    hipLaunchKernelGGL((k_les_gen<8, 1>), dim3(grid1(nv * 8)), dim3(256), 0, c.stream, nv, G, s, blk, pen, r, ye, sc);
  else if (nsh == 9 && c.bs == 1)
    hipLaunchKernelGGL((k_les_gen<9, 1>), dim3(grid1(nv * 9)), dim3(256), 0, c.stream, nv, G, s, blk, pen, r, ye, sc);
  else if (nsh == 4 && c.bs == 1)
    hipLaunchKernelGGL((k_les_gen<4, 1>), dim3(grid1(nv * 4)), dim3(256), 0, c.stream, nv, G, s, blk, pen, r, ye, sc);
  else {
    set_error("learned preconditioner: element not supported");
    return PZG_EUNSUP;
  }
  const int gg = grid1(c.nb);
  if (ch) {
    if (c.bs == 3)
      hipLaunchKernelGGL(k_les_gather_cheb<3>, dim3(gg), dim3(256), 0, c.stream, c.nb, c.d_mf_ptr.as<int64_t>(), eoff,
                         ye, pen, r, *ch, sc);
    else
      hipLaunchKernelGGL(k_les_gather_cheb<1>, dim3(gg), dim3(256), 0, c.stream, c.nb, c.d_mf_ptr.as<int64_t>(), eoff,
                         ye, pen, r, *ch, sc);
  } else if (c.bs == 3)
    hipLaunchKernelGGL(k_les_gather<3>, dim3(gg), dim3(256), 0, c.stream, c.nb, c.d_mf_ptr.as<int64_t>(), eoff, ye, pen, r, z, sc);
  else
    hipLaunchKernelGGL(k_les_gather<1>, dim3(gg), dim3(256), 0, c.stream, c.nb, c.d_mf_ptr.as<int64_t>(), eoff, ye, pen, r, z, sc);
  PZG_HIP(hipGetLastError());
  return PZG_OK;
}

// (residual, error) pairs of every element's local problem of the assembled
// system: nrhs Gaussian residuals per element (zero on penalized dofs) and
// their local solutions
int les_collect(Ctx &c, int nrhs, uint64_t seed) {
  if (!c.les_ready || !c.assembled) {
    set_error("learned preconditioner: call pzg_learned_setup on an assembled system first");
    return PZG_ESTATE;
  }
  if (nrhs < 1 || nrhs > 8) {
    set_error("learned preconditioner: 1..8 residuals per element");
    return PZG_EINVAL;
  }
  PZG_TRY(les_prepare(c));
  if (!c.full_valid) PZG_TRY(launch_expand_full(c));
  if (!c.scal_valid) PZG_TRY(launch_build_bsr(c, false));
  const int ne = c.les_ne, nsh = c.tvol.nsh;
  const int64_t nv = c.nvol;
  PZG_TRY(dev_alloc(c.d_les_Se, sizeof(double) * nv * ne * ne));
  const int64_t nt = nv * nsh * nsh;
  if (c.bs == 3)
    hipLaunchKernelGGL(k_les_local<3>, dim3(grid1(nt)), dim3(TPB), 0, c.stream, nv, nsh, c.d_mf_blk.as<int32_t>(),
                       c.d_fbia.as<int64_t>(), c.d_fbja.as<int32_t>(), c.d_fa.as<double>(),
                       c.d_dinvsqrt.as<double>(), c.d_les_pen.as<float>(), c.d_les_Se.as<double>());
  else
    hipLaunchKernelGGL(k_les_local<1>, dim3(grid1(nt)), dim3(TPB), 0, c.stream, nv, nsh, c.d_mf_blk.as<int32_t>(),
                       c.d_fbia.as<int64_t>(), c.d_fbja.as<int32_t>(), c.d_fa.as<double>(),
                       c.d_dinvsqrt.as<double>(), c.d_les_pen.as<float>(), c.d_les_Se.as<double>());
  PZG_HIP(hipGetLastError());
  // residuals (host, seeded) and the per-element dof mask
  std::vector<int32_t> blk((size_t)nv * nsh);
  PZG_HIP(hipMemcpyAsync(blk.data(), c.d_mf_blk.p, blk.size() * sizeof(int32_t), hipMemcpyDeviceToHost, c.stream));
  PZG_HIP(hipStreamSynchronize(c.stream));
  std::vector<float> mask((size_t)nv * ne), R((size_t)nv * nrhs * ne);
  std::mt19937_64 rng(seed);
  std::normal_distribution<float> nd(0.f, 1.f);
  for (int64_t e = 0; e < nv; e++)
    for (int i = 0; i < ne; i++) {
      const int32_t b = blk[e * nsh + i / c.bs];
      mask[e * ne + i] = (b >= 0 && c.les_pen[(int64_t)b * c.bs + i % c.bs] == 0.f) ? 1.f : 0.f;
    }
  for (int64_t e = 0; e < nv; e++)
    for (int k = 0; k < nrhs; k++)
      for (int i = 0; i < ne; i++) R[(e * nrhs + k) * ne + i] = nd(rng) * mask[e * ne + i];
  hipStream_t s = c.stream;
  PZG_TRY(dev_upload(c.d_les_M, mask.data(), mask.size(), s));
  PZG_TRY(dev_upload(c.d_les_R, R.data(), R.size(), s));
  PZG_TRY(dev_alloc(c.d_les_E, sizeof(float) * R.size()));
  PZG_TRY(dev_alloc(c.d_les_fail, sizeof(int)));
  PZG_HIP(hipMemsetAsync(c.d_les_fail.p, 0, sizeof(int), s));
  hipLaunchKernelGGL(k_les_solve, dim3(nv), dim3(256), 0, s, ne, nrhs, c.d_les_Se.as<double>(),
                     c.d_les_R.as<float>(), c.d_les_E.as<float>(), c.d_les_fail.as<int>());
  PZG_HIP(hipGetLastError());
  int fail = 0;
  std::vector<float> E(R.size());
  PZG_HIP(hipMemcpyAsync(&fail, c.d_les_fail.p, sizeof(int), hipMemcpyDeviceToHost, s));
  PZG_HIP(hipMemcpyAsync(E.data(), c.d_les_E.p, E.size() * sizeof(float), hipMemcpyDeviceToHost, s));
  PZG_HIP(hipStreamSynchronize(s));
  if (fail) {
    set_error("learned preconditioner: an element block is not positive definite");
    return PZG_EUNSUP;
  }
  double den = 0.;
  for (size_t i = 0; i < E.size(); i++) den += (double)E[i] * E[i];
  c.les_den = den > 0. ? den : 1.;
  c.les_nr = nrhs;
  dev_free(c.d_les_Se);
  // least-squares sums of the G fit over the elements without penalized dofs
  std::vector<float> Rf, Ef;
  for (int64_t e = 0; e < nv; e++) {
    bool full = true;
    for (int i = 0; i < ne && full; i++) full = mask[e * ne + i] != 0.f;
    if (!full) continue;
    Rf.insert(Rf.end(), R.begin() + e * nrhs * ne, R.begin() + (e + 1) * nrhs * ne);
    Ef.insert(Ef.end(), E.begin() + e * nrhs * ne, E.begin() + (e + 1) * nrhs * ne);
  }
  const int64_t rows = (int64_t)Rf.size() / ne;
  if (c.les_A.size() != (size_t)ne * ne) {
    c.les_A.assign((size_t)ne * ne, 0.);
    c.les_B.assign((size_t)ne * ne, 0.);
  }
  if (rows > 0) {
    DevBuf dR, dE;
    PZG_TRY(dev_upload(dR, Rf.data(), Rf.size(), s));
    PZG_TRY(dev_upload(dE, Ef.data(), Ef.size(), s));
    PZG_TRY(dev_alloc(c.d_nn_part, sizeof(float) * NBLK_RED * std::max(ne * ne, NNF * NNF)));
    PZG_TRY(dev_alloc(c.d_les_g, sizeof(float) * les_offsets(ne).n));
    const int nred = (int)std::min<int64_t>(NBLK_RED, std::max<int64_t>(1, rows / 64));
    std::vector<float> h((size_t)ne * ne);
    for (int which = 0; which < 2; which++) {
      const float *Lm = which ? dE.as<float>() : dR.as<float>();
      float *part = c.d_nn_part.as<float>();
      switch (ne) {
        case 81: hipLaunchKernelGGL((k_nn_outer<81, 81>), dim3(nred), dim3(TPB), 0, s, rows, Lm, dR.as<float>(), part); break;
        case 27: hipLaunchKernelGGL((k_nn_outer<27, 27>), dim3(nred), dim3(TPB), 0, s, rows, Lm, dR.as<float>(), part); break;
        case 24: hipLaunchKernelGGL((k_nn_outer<24, 24>), dim3(nred), dim3(TPB), 0, s, rows, Lm, dR.as<float>(), part); break;
        case 9: hipLaunchKernelGGL((k_nn_outer<9, 9>), dim3(nred), dim3(TPB), 0, s, rows, Lm, dR.as<float>(), part); break;
        case 8: hipLaunchKernelGGL((k_nn_outer<8, 8>), dim3(nred), dim3(TPB), 0, s, rows, Lm, dR.as<float>(), part); break;
        case 4: hipLaunchKernelGGL((k_nn_outer<4, 4>), dim3(nred), dim3(TPB), 0, s, rows, Lm, dR.as<float>(), part); break;
        default:
          set_error("learned preconditioner: element not supported");
          return PZG_EUNSUP;
      }
      hipLaunchKernelGGL(k_nn_sum_partials, dim3(grid1(ne * ne)), dim3(TPB), 0, s, nred, ne * ne, part,
                         c.d_les_g.as<float>());
      PZG_HIP(hipGetLastError());
      PZG_HIP(hipMemcpyAsync(h.data(), c.d_les_g.p, sizeof(float) * ne * ne, hipMemcpyDeviceToHost, s));
      PZG_HIP(hipStreamSynchronize(s));
      std::vector<double> &acc = which ? c.les_B : c.les_A;
      for (int k = 0; k < ne * ne; k++) acc[k] += h[k];
    }
    dev_free(dR);
    dev_free(dE);
    c.les_fit_rows += rows;
    c.les_fit_stale = true;
  }
  return PZG_OK;
}

// eigen-decomposition of a symmetric n x n matrix (cyclic Jacobi rotations):
// a is destroyed, w receives the eigenvalues, v the eigenvectors (columns)
static void sym_eig(std::vector<double> &a, int n, std::vector<double> &w, std::vector<double> &v) {
  v.assign((size_t)n * n, 0.);
  for (int i = 0; i < n; i++) v[(size_t)i * n + i] = 1.;
  for (int sweep = 0; sweep < 60; sweep++) {
    double off = 0.;
    for (int i = 0; i < n; i++)
      for (int j = i + 1; j < n; j++) off += a[(size_t)i * n + j] * a[(size_t)i * n + j];
    if (off < 1e-30) break;
    for (int p = 0; p < n; p++)
      for (int q = p + 1; q < n; q++) {
        const double apq = a[(size_t)p * n + q];
        if (std::fabs(apq) < 1e-300) continue;
        const double app = a[(size_t)p * n + p], aqq = a[(size_t)q * n + q];
        const double th = 0.5 * (aqq - app) / apq;
        const double t = (th >= 0 ? 1. : -1.) / (std::fabs(th) + std::sqrt(th * th + 1.));
        const double cs = 1. / std::sqrt(t * t + 1.), sn = t * cs;
        for (int k = 0; k < n; k++) {
          const double akp = a[(size_t)k * n + p], akq = a[(size_t)k * n + q];
          a[(size_t)k * n + p] = cs * akp - sn * akq;
          a[(size_t)k * n + q] = sn * akp + cs * akq;
        }
        for (int k = 0; k < n; k++) {
          const double apk = a[(size_t)p * n + k], aqk = a[(size_t)q * n + k];
          a[(size_t)p * n + k] = cs * apk - sn * aqk;
          a[(size_t)q * n + k] = sn * apk + cs * aqk;
        }
        for (int k = 0; k < n; k++) {
          const double vkp = v[(size_t)k * n + p], vkq = v[(size_t)k * n + q];
          v[(size_t)k * n + p] = cs * vkp - sn * vkq;
          v[(size_t)k * n + q] = sn * vkp + cs * vkq;
        }
      }
  }
  w.resize(n);
  for (int i = 0; i < n; i++) w[i] = a[(size_t)i * n + i];
}

// G = B A^-1 from the accumulated sums, symmetrised and made positive definite
int les_fit_G(Ctx &c) {
  if (!c.les_fit_stale) return PZG_OK;
  const int ne = c.les_ne;
  if (c.les_fit_rows < ne) {
    set_error("learned preconditioner: too few interior pairs for the element fit");
    return PZG_EINVAL;
  }
  // A = V diag(w) V^T  ->  G = B V diag(1/w) V^T
  std::vector<double> a(c.les_A), w, V;
  sym_eig(a, ne, w, V);
  double wmax = 0.;
  for (double x : w) wmax = std::max(wmax, x);
  std::vector<double> BV((size_t)ne * ne, 0.), G((size_t)ne * ne, 0.);
  for (int i = 0; i < ne; i++)
    for (int k = 0; k < ne; k++) {
      double t = 0.;
      for (int j = 0; j < ne; j++) t += c.les_B[(size_t)i * ne + j] * V[(size_t)j * ne + k];
      BV[(size_t)i * ne + k] = t / std::max(w[k], 1e-12 * wmax);
    }
  for (int i = 0; i < ne; i++)
    for (int j = 0; j < ne; j++) {
      double t = 0.;
      for (int k = 0; k < ne; k++) t += BV[(size_t)i * ne + k] * V[(size_t)j * ne + k];
      G[(size_t)i * ne + j] = t;
    }
  for (int i = 0; i < ne; i++)
    for (int j = i + 1; j < ne; j++) {
      const double m = 0.5 * (G[(size_t)i * ne + j] + G[(size_t)j * ne + i]);
      G[(size_t)i * ne + j] = G[(size_t)j * ne + i] = m;
    }
  // positive definite projection
  std::vector<double> g(G), u, U;
  sym_eig(g, ne, u, U);
  double umax = 0.;
  for (double x : u) umax = std::max(umax, x);
  if (!(umax > 0.)) {
    set_error("learned preconditioner: the element fit has no positive eigenvalue");
    return PZG_EUNSUP;
  }
  for (double &x : u) x = std::max(x, 1e-3 * umax);
  std::vector<float> Gf((size_t)ne * ne);
  for (int i = 0; i < ne; i++)
    for (int j = 0; j < ne; j++) {
      double t = 0.;
      for (int k = 0; k < ne; k++) t += U[(size_t)i * ne + k] * u[k] * U[(size_t)j * ne + k];
      Gf[(size_t)i * ne + j] = (float)t;
    }
  PZG_HIP(hipMemcpyAsync(c.d_les_th.as<float>() + les_offsets(ne).g, Gf.data(), sizeof(float) * ne * ne,
                         hipMemcpyHostToDevice, c.stream));
  PZG_HIP(hipStreamSynchronize(c.stream));
  c.les_fit_stale = false;
  c.mg_ready = false;
  return PZG_OK;
}

// loss and gradient (into d_les_g) on the stored pairs
int les_loss_grad(Ctx &c, double *loss_out) {
  if (!c.les_ready || c.les_nr == 0) {
    set_error("learned preconditioner: no training pairs (pzg_learned_collect)");
    return PZG_ESTATE;
  }
  const int ne = c.les_ne, nr = (int)c.les_nr;
  const int64_t nv = c.nvol, rows = nv, need = rows * NNF;
  LesOff o = les_offsets(ne);
  PZG_TRY(les_features(c));
  PZG_TRY(les_forward(c, true));
  PZG_TRY(dev_alloc(c.d_les_D, sizeof(float) * nv * nr * ne));
  PZG_TRY(dev_alloc(c.d_les_le, sizeof(double) * nv));
  PZG_TRY(dev_alloc(c.d_les_ds, sizeof(double) * nv));
  PZG_TRY(dev_alloc(c.d_les_loss, sizeof(double)));
  PZG_TRY(dev_alloc(c.d_nn_dZ, sizeof(float) * rows));
  PZG_TRY(dev_alloc(c.d_nn_dH, sizeof(float) * need));
  PZG_TRY(dev_alloc(c.d_nn_G, sizeof(float) * need));
  PZG_TRY(dev_alloc(c.d_nn_tmp, sizeof(float) * need));
  PZG_TRY(dev_alloc(c.d_nn_part, sizeof(float) * NBLK_RED * std::max(ne * ne, NNF * NNF)));
  const float *th = c.d_les_th.as<float>();
  float *gr = c.d_les_g.as<float>();
  hipStream_t s = c.stream;
  hipLaunchKernelGGL(k_les_fwd, dim3(nv), dim3(128), 0, s, ne, nr, th + o.g, c.d_les_s.as<float>(),
                     c.d_les_M.as<float>(), c.d_les_R.as<float>(), c.d_les_E.as<float>(), c.les_den,
                     c.d_les_D.as<float>(), c.d_les_le.as<double>(), c.d_les_ds.as<double>());
  hipLaunchKernelGGL(k_les_sum, dim3(1), dim3(256), 0, s, nv, c.d_les_le.as<double>(), c.d_les_loss.as<double>());
  hipLaunchKernelGGL(k_les_dz, dim3(grid1(nv)), dim3(TPB), 0, s, nv, c.d_les_z.as<float>(),
                     c.d_les_s.as<float>(), c.d_les_ds.as<double>(), c.les_den, c.d_nn_dZ.as<float>());
  PZG_HIP(hipGetLastError());
  // gradient of G: sum over (element, pair) rows of Dsc (x) r, symmetrised
  const int64_t grow = nv * nr;
  const int nred = (int)std::min<int64_t>(NBLK_RED, std::max<int64_t>(1, grow / 64));
  auto outerG = [&]() -> int {
    float *part = c.d_nn_part.as<float>();
    const float *Lm = c.d_les_D.as<float>(), *Rm = c.d_les_R.as<float>();
    switch (ne) {
      case 81: hipLaunchKernelGGL((k_nn_outer<81, 81>), dim3(nred), dim3(TPB), 0, s, grow, Lm, Rm, part); break;
      case 27: hipLaunchKernelGGL((k_nn_outer<27, 27>), dim3(nred), dim3(TPB), 0, s, grow, Lm, Rm, part); break;
      case 24: hipLaunchKernelGGL((k_nn_outer<24, 24>), dim3(nred), dim3(TPB), 0, s, grow, Lm, Rm, part); break;
      case 9: hipLaunchKernelGGL((k_nn_outer<9, 9>), dim3(nred), dim3(TPB), 0, s, grow, Lm, Rm, part); break;
      case 8: hipLaunchKernelGGL((k_nn_outer<8, 8>), dim3(nred), dim3(TPB), 0, s, grow, Lm, Rm, part); break;
      case 4: hipLaunchKernelGGL((k_nn_outer<4, 4>), dim3(nred), dim3(TPB), 0, s, grow, Lm, Rm, part); break;
      default:
        set_error("learned preconditioner: element not supported");
        return PZG_EUNSUP;
    }
    hipLaunchKernelGGL(k_nn_sum_partials, dim3(grid1(ne * ne)), dim3(TPB), 0, s, nred, ne * ne, part, gr + o.g);
    hipLaunchKernelGGL(k_les_sym, dim3(grid1(ne * ne)), dim3(TPB), 0, s, ne, gr + o.g);
    return PZG_OK;
  };
  PZG_TRY(outerG());
  // network backward (rows = elements)
  float *H = c.d_nn_trH.as<float>(), *M = c.d_nn_trM.as<float>(), *P = c.d_nn_trP.as<float>();
  const int gnode = grid1(need);
  const int nrn = (int)std::min<int64_t>(NBLK_RED, std::max<int64_t>(1, rows / 64));
  auto outer = [&](const float *Lm, const float *Rm, int fa, int fb, int off) {
    float *part = c.d_nn_part.as<float>();
    if (fa == NNF && fb == NNF)
      hipLaunchKernelGGL((k_nn_outer<NNF, NNF>), dim3(nrn), dim3(TPB), 0, s, rows, Lm, Rm, part);
    else if (fa == NNF && fb == 1)
      hipLaunchKernelGGL((k_nn_outer<NNF, 1>), dim3(nrn), dim3(TPB), 0, s, rows, Lm, Rm, part);
    else
      hipLaunchKernelGGL((k_nn_outer<LES_FIN, NNF>), dim3(nrn), dim3(TPB), 0, s, rows, Lm, Rm, part);
    hipLaunchKernelGGL(k_nn_sum_partials, dim3(grid1(fa * fb)), dim3(TPB), 0, s, nrn, fa * fb, part, gr + off);
  };
  hipLaunchKernelGGL(k_nn_back_out<1>, dim3(gnode), dim3(TPB), 0, s, rows, c.d_nn_dZ.as<float>(), th + o.wout,
                     c.d_nn_dH.as<float>());
  outer(H + (int64_t)NNL * need, c.d_nn_dZ.as<float>(), NNF, 1, o.wout);
  for (int l = NNL - 1; l >= 0; l--) {
    hipLaunchKernelGGL(k_nn_back_layer, dim3(gnode), dim3(TPB), 0, s, rows, c.d_nn_dH.as<float>(),
                       P + (int64_t)l * need, th + o.a[l], th + o.b[l], c.d_nn_G.as<float>(),
                       c.d_nn_dH.as<float>(), c.d_nn_tmp.as<float>());
    outer(H + (int64_t)l * need, c.d_nn_G.as<float>(), NNF, NNF, o.a[l]);
    outer(M + (int64_t)l * need, c.d_nn_G.as<float>(), NNF, NNF, o.b[l]);
    hipLaunchKernelGGL(k_nn_agg_t, dim3(gnode), dim3(TPB), 0, s, rows, rows, c.d_les_cia.as<int64_t>(),
                       c.d_les_cja.as<int32_t>(), c.d_nn_tmp.as<float>(), c.d_nn_dH.as<float>());
  }
  outer(c.d_les_X.as<float>(), c.d_nn_dH.as<float>(), LES_FIN, NNF, o.w0);
  PZG_HIP(hipGetLastError());
  if (loss_out) {
    PZG_HIP(hipMemcpyAsync(loss_out, c.d_les_loss.p, sizeof(double), hipMemcpyDeviceToHost, s));
    PZG_HIP(hipStreamSynchronize(s));
    *loss_out /= c.les_den;
  }
  return PZG_OK;
}

// Adam on the network weights (G is the least-squares fit)
int les_adam_step(Ctx &c, double lr) {
  LesOff o = les_offsets(c.les_ne);
  c.les_adam_t++;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - std::pow(b1, (float)c.les_adam_t), c2 = 1.f - std::pow(b2, (float)c.les_adam_t);
  const int n = o.n - o.w0;
  hipLaunchKernelGGL(k_adam, dim3(grid1(n)), dim3(TPB), 0, c.stream, n, c.d_les_th.as<float>() + o.w0,
                     c.d_les_g.as<float>() + o.w0, c.d_les_m.as<float>() + o.w0,
                     c.d_les_v.as<float>() + o.w0, (float)lr, b1, b2, c1, c2, 1e-8f);
  PZG_HIP(hipGetLastError());
  c.mg_ready = false;  // the smoother changed: the cycle's eigenvalue estimate too
  return PZG_OK;
}

}  // namespace pzg
