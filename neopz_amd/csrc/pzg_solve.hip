// pzg_solve.hip -- gfx950 kernels and drivers of the solve phase.
//
//  PCG64 (plain path, reference semantics).  CG<> of
//  Solvers/LinearSolvers/cg.h:29-132 with the EJacobi TPZStepSolver
//  (z = r / diag, Matrix/pzmatrix.cpp:642-667), x0 = 0, stop on the recursive
//  ||r||/||b|| <= tol.  The reference's Jacobi step also runs a residual SpMV
//  whose result it discards; it has no effect on the iterates and is not run.
//  Matrix: full CSR, FP64 values, int32 columns (expanded once per assembly
//  from the reference upper CSR).
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
#include <cstring>

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
#ifndef PZG_NT_LOADS
#define PZG_NT_LOADS 0
#endif
template <class T> __device__ inline T ld_stream0(const T *p) {
#if PZG_NT_LOADS
  return __builtin_nontemporal_load(p);
#else
  return *p;
#endif
}
// FP64 full CSR SpMV; VL lanes per row, U entries per lane loaded before use
// so that a whole row (up to VL*U entries) is in flight at once.
template <int VL>
__device__ inline double row_dot64(int64_t r, const int64_t *__restrict__ ia,
                                   const int32_t *__restrict__ ja, const double *__restrict__ a,
                                   const double *__restrict__ x, int sub) {
  constexpr int U = 4;
  double s = 0.;
  const int64_t k1 = ia[r + 1];
  for (int64_t base = ia[r] + sub; base < k1; base += VL * U) {
    double av[U];
    int32_t cv[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int64_t k = base + u * VL;
      const bool ok = k < k1;
      av[u] = ok ? ld_stream0(&a[k]) : 0.;
      cv[u] = ok ? ld_stream0(&ja[k]) : 0;
    }
    double xv[U];
#pragma unroll
    for (int u = 0; u < U; u++) xv[u] = x[cv[u]];
#pragma unroll
    for (int u = 0; u < U; u++) s += av[u] * xv[u];
  }
  for (int o = VL / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, VL);
  return s;
}

template <int VL>
__global__ __launch_bounds__(TPB) void k_spmv64(int64_t n, const int64_t *__restrict__ ia,
                                                const int32_t *__restrict__ ja,
                                                const double *__restrict__ a,
                                                const double *__restrict__ x,
                                                double *__restrict__ y) {
  const int sub = threadIdx.x % VL;
  int64_t g = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / VL;
  const int64_t ng = (int64_t)gridDim.x * blockDim.x / VL;
  for (int64_t r = g; r < n; r += ng) {
    double s = row_dot64<VL>(r, ia, ja, a, x, sub);
    if (sub == 0) y[r] = s;
  }
}

// q = A p, then alpha = rho / (p.q)
template <int VL>
__global__ __launch_bounds__(TPB) void k_spmv64_pq(int64_t n, const int64_t *__restrict__ ia,
                                                   const int32_t *__restrict__ ja,
                                                   const double *__restrict__ a,
                                                   const double *__restrict__ p,
                                                   double *__restrict__ q, SolveScalars *sc,
                                                   double *partials, unsigned *counter) {
  if (sc->done) return;
  const int sub = threadIdx.x % VL;
  int64_t g = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / VL;
  const int64_t ng = (int64_t)gridDim.x * blockDim.x / VL;
  double acc = 0.;
  for (int64_t r = g; r < n; r += ng) {
    double s = row_dot64<VL>(r, ia, ja, a, p, sub);
    if (sub == 0) {
      q[r] = s;
      acc += s * p[r];
    }
  }
  double v[1] = {acc};
  block_partials<1>(v, partials);
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
    double zi = ri / diag[i];
    z[i] = zi;
    rr += ri * ri;
    rz += ri * zi;
  }
  double v[2] = {rr, rz};
  block_partials<2>(v, partials);
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
    double zi = bi / diag[i];
    z[i] = zi;
    bb += bi * bi;
    rz += bi * zi;
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

// r = f - A u (FP64), |r|^2 and |f|^2
template <int VL>
__global__ __launch_bounds__(TPB) void k_residual64(int64_t n, const int64_t *__restrict__ ia,
                                                    const int32_t *__restrict__ ja,
                                                    const double *__restrict__ a,
                                                    const double *__restrict__ u,
                                                    const double *__restrict__ f,
                                                    double *__restrict__ r, double *out,
                                                    double *partials, unsigned *counter) {
  const int sub = threadIdx.x % VL;
  int64_t g = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / VL;
  const int64_t ng = (int64_t)gridDim.x * blockDim.x / VL;
  double rr = 0., ff = 0.;
  for (int64_t i = g; i < n; i += ng) {
    double s = row_dot64<VL>(i, ia, ja, a, u, sub);
    if (sub == 0) {
      double ri = f[i] - s;
      if (r) r[i] = ri;
      rr += ri * ri;
      ff += f[i] * f[i];
    }
  }
  double v[2] = {rr, ff}, tot[2];
  if (reduce_last<2>(v, partials, counter, tot) && threadIdx.x == 0) {
    out[0] = tot[0];
    out[1] = tot[1];
  }
}

__global__ void k_axpby64(int64_t n, double alpha, const double *__restrict__ x, double beta,
                          const double *__restrict__ y, double *__restrict__ z) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t st = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += st) z[i] = alpha * x[i] + (y ? beta * y[i] : 0.);
}

static int pick_vl(int64_t nnz, int64_t n) {
  double avg = n ? (double)nnz / n : 1;
  if (avg > 96) return 64;
  if (avg > 40) return 32;
  if (avg > 20) return 16;
  if (avg > 10) return 8;
  return 4;
}

#define DISPATCH_VL(VLV, CALL)   \
  switch (VLV) {                 \
    case 64: { constexpr int VL = 64; CALL; } break; \
    case 32: { constexpr int VL = 32; CALL; } break; \
    case 16: { constexpr int VL = 16; CALL; } break; \
    case 8:  { constexpr int VL = 8;  CALL; } break; \
    default: { constexpr int VL = 4;  CALL; } break; \
  }

static int spmv_grid(int64_t n, int vl) {
  int rows_per_block = TPB / vl;
  int64_t g = (n + rows_per_block - 1) / rows_per_block;
  int64_t cap = 8192;
  if (g > cap) g = cap;
  return (int)std::max<int64_t>(g, 1);
}

int launch_spmv64(Ctx &c, const double *x, double *y) {
  int vl = pick_vl(c.nnz_full, c.neq);
  int g = spmv_grid(c.neq, vl);
  DISPATCH_VL(vl, hipLaunchKernelGGL(k_spmv64<VL>, dim3(g), dim3(TPB), 0, c.stream, c.neq,
                                     c.d_fia.as<int64_t>(), c.d_fja.as<int32_t>(),
                                     c.d_fa.as<double>(), x, y));
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
  return PZG_OK;
}

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
  if (!c.full_valid) PZG_TRY(launch_expand_full(c));
  PZG_TRY(ensure_reduce(c));
  DevBuf out;
  PZG_TRY(dev_alloc(out, 2 * sizeof(double)));
  int vl = pick_vl(c.nnz_full, c.neq);
  int g = grid_for(c.neq, TPB / vl);
  DISPATCH_VL(vl, hipLaunchKernelGGL(k_residual64<VL>, dim3(g), dim3(TPB), 0, c.stream, c.neq,
                                     c.d_fia.as<int64_t>(), c.d_fja.as<int32_t>(),
                                     c.d_fa.as<double>(), u, c.d_f.as<double>(), nullptr,
                                     out.as<double>(), c.d_partials.as<double>(),
                                     c.d_counter.as<unsigned>()));
  PZG_HIP(hipGetLastError());
  double h[2];
  PZG_HIP(hipMemcpyAsync(h, out.p, sizeof(h), hipMemcpyDeviceToHost, c.stream));
  PZG_HIP(hipStreamSynchronize(c.stream));
  dev_free(out);
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
  hipLaunchKernelGGL(k_cg_init64, dim3(gv), dim3(TPB), 0, c.stream, n, c.d_f.as<double>(),
                     c.d_u.as<double>(), c.d_r.as<double>(), c.d_z.as<double>(),
                     c.d_p.as<double>(), c.d_diag.as<double>(), sc, part, cnt);
  PZG_HIP(hipGetLastError());
  const int vl = pick_vl(c.nnz_full, n);
  const int gs = grid_for(n, TPB / vl);
  auto iter = [&]() {
    hipLaunchKernelGGL(k_p_update64, dim3(gv), dim3(TPB), 0, c.stream, n, c.d_z.as<double>(),
                       c.d_p.as<double>(), sc);
    DISPATCH_VL(vl, hipLaunchKernelGGL(k_spmv64_pq<VL>, dim3(gs), dim3(TPB), 0, c.stream, n,
                                       c.d_fia.as<int64_t>(), c.d_fja.as<int32_t>(),
                                       c.d_fa.as<double>(), c.d_p.as<double>(),
                                       c.d_q.as<double>(), sc, part, cnt));
    hipLaunchKernelGGL(k_fin_alpha64, dim3(1), dim3(TPB), 0, c.stream, sc, (const double *)part, gs);
    hipLaunchKernelGGL(k_cg_update64, dim3(gv), dim3(TPB), 0, c.stream, n, c.d_u.as<double>(),
                       c.d_r.as<double>(), c.d_z.as<double>(), c.d_p.as<double>(),
                       c.d_q.as<double>(), c.d_diag.as<double>(), sc, (double *)nullptr, part,
                       cnt);
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
// Mixed path: FP32 block CSR SpMV (block size BS, one wave per block row).
#ifndef PZG_BSR_VEC
#define PZG_BSR_VEC 0
#endif
template <class T> __device__ inline T ld_stream(const T *p) {
#if PZG_NT_LOADS
  return __builtin_nontemporal_load(p);
#else
  return *p;
#endif
}

template <int BS>
__device__ inline void bsr_load(int64_t t, const float *__restrict__ v, int64_t nbf,
                                float (&m)[BS * BS]) {
#if PZG_BSR_SOA
  for (int c = 0; c < BS * BS; c++) m[c] = ld_stream(&v[c * nbf + t]);
#else
  (void)nbf;
  for (int c = 0; c < BS * BS; c++) m[c] = ld_stream(&v[t * (BS * BS) + c]);
#endif
}

// One wavefront per block row; each lane takes U blocks per pass, all loads
// of a pass issued before the products.
template <int BS>
__device__ inline void bsr_row(int64_t A, const int64_t *__restrict__ bia,
                               const int32_t *__restrict__ bja, const float *__restrict__ v,
                               int64_t nbf, const float *__restrict__ x, int lane,
                               float (&out)[BS]) {
  constexpr int U = 2;
  float acc[BS];
  for (int k = 0; k < BS; k++) acc[k] = 0.f;
  const int64_t b0 = bia[A], b1 = bia[A + 1];
  for (int64_t base = b0 + lane; base < b1; base += 64 * U) {
    float m[U][BS * BS];
    int64_t col[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int64_t t = base + 64 * u;
      if (t < b1) {
        col[u] = ld_stream(&bja[t]);
        bsr_load<BS>(t, v, nbf, m[u]);
      } else {
        col[u] = 0;
        for (int c = 0; c < BS * BS; c++) m[u][c] = 0.f;
      }
    }
    float xv[U][BS];
#pragma unroll
    for (int u = 0; u < U; u++)
      for (int l = 0; l < BS; l++) xv[u][l] = x[col[u] * BS + l];
#pragma unroll
    for (int u = 0; u < U; u++)
      for (int k = 0; k < BS; k++)
        for (int l = 0; l < BS; l++) acc[k] += m[u][k * BS + l] * xv[u][l];
  }
  for (int k = 0; k < BS; k++) {
    float s = acc[k];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    out[k] = s;
  }
}

template <int BS>
__global__ __launch_bounds__(TPB) void k_spmv_bsr(int64_t nb, const int64_t *__restrict__ bia,
                                                  const int32_t *__restrict__ bja,
                                                  const float *__restrict__ v, int64_t nbf,
                                                  const float *__restrict__ x,
                                                  float *__restrict__ y) {
  const int lane = threadIdx.x & 63;
  int64_t w = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t A = w; A < nb; A += nw) {
    float out[BS];
    bsr_row<BS>(A, bia, bja, v, nbf, x, lane, out);
    if (lane < BS) {
      float o = out[0];
      for (int k = 1; k < BS; k++)
        if (lane == k) o = out[k];
      y[A * BS + lane] = o;
    }
  }
}

// q = S p; pq = p.q; pr = p.r; alpha = pr/pq
template <int BS>
__global__ __launch_bounds__(TPB) void k_fcg_spmv(int64_t nb, const int64_t *__restrict__ bia,
                                                  const int32_t *__restrict__ bja,
                                                  const float *__restrict__ v, int64_t nbf,
                                                  const float *__restrict__ p,
                                                  const float *__restrict__ r,
                                                  float *__restrict__ q, SolveScalars *sc,
                                                  double *partials, unsigned *counter) {
  if (sc->done) return;
  const int lane = threadIdx.x & 63;
  int64_t w = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
  double pq = 0., pr = 0.;
  for (int64_t A = w; A < nb; A += nw) {
    float out[BS];
    bsr_row<BS>(A, bia, bja, v, nbf, p, lane, out);
    if (lane < BS) {
      float o = out[0];
      for (int k = 1; k < BS; k++)
        if (lane == k) o = out[k];
      const int64_t i = A * BS + lane;
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
__global__ void k_p_update32(int64_t n, const float *__restrict__ z, float *__restrict__ p,
                             const SolveScalars *sc) {
  if (sc->done != 0 && sc->done != 4) return;
  const float beta = (float)sc->beta;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t st = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += st) p[i] = z[i] + beta * p[i];
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

// x = 0; r = b (fp32); z = Binv r; p = z; normb = |b|
template <int BS>
__global__ __launch_bounds__(TPB) void k_fcg_start(int64_t nb, const float *__restrict__ b,
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
      rv[k] = b[i];
      r[i] = rv[k];
      x[i] = 0.f;
      bb += (double)rv[k] * rv[k];
    }
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
    sc->resid = 1.;
    sc->done = 0;
  }
}

// out = a - b - c
__global__ void k_sub3(int64_t n, const float *__restrict__ a, const float *__restrict__ b,
                       const float *__restrict__ c, float *__restrict__ out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t st = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += st) out[i] = a[i] - b[i] - c[i];
}

int launch_spmv_bsr(Ctx &c, const float *x, float *y) {
  int g = grid_for(c.nb, TPB / 64);
  if (g < 1) g = 1;
  if (c.bs == 3)
    hipLaunchKernelGGL(k_spmv_bsr<3>, dim3(g), dim3(TPB), 0, c.stream, c.nb,
                       c.d_fbia.as<int64_t>(), c.d_fbja.as<int32_t>(), c.d_bsr.as<float>(), c.nbf, x, y);
  else
    hipLaunchKernelGGL(k_spmv_bsr<1>, dim3(g), dim3(TPB), 0, c.stream, c.nb,
                       c.d_fbia.as<int64_t>(), c.d_fbja.as<int32_t>(), c.d_bsr.as<float>(), c.nbf, x, y);
  PZG_HIP(hipGetLastError());
  return PZG_OK;
}

template <int BS>
static int mixed_impl(Ctx &c, const pzg_solve_params &prm, pzg_solve_info &info) {
  const int64_t n = c.neq, nb = c.nb;
  if (!c.full_valid) PZG_TRY(launch_expand_full(c));
  if (!c.bsr_valid) PZG_TRY(launch_build_bsr(c));
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
  if (use_mf) PZG_TRY(mf_prepare(c));
  PZG_TRY(ensure_reduce(c));
  for (DevBuf *b : {&c.d_u, &c.d_r}) PZG_TRY(dev_alloc(*b, n * sizeof(double)));
  for (DevBuf *b : {&c.d_xs, &c.d_rs, &c.d_zs, &c.d_ps, &c.d_qs}) PZG_TRY(dev_alloc(*b, n * sizeof(float)));
  DevBuf out;
  PZG_TRY(dev_alloc(out, 2 * sizeof(double)));
  SolveScalars *sc = c.d_scal.as<SolveScalars>();
  double *part = c.d_partials.as<double>();
  unsigned *cnt = c.d_counter.as<unsigned>();
  const int gv = grid_for(n, TPB), gb = grid_for(nb, TPB), gs = grid_for(nb, TPB / 64);
  const int vl = pick_vl(c.nnz_full, n), gr = grid_for(n, TPB / vl);
  PZG_HIP(hipEventRecord(c.ev0, c.stream));
  hipLaunchKernelGGL(k_axpby64, dim3(gv), dim3(TPB), 0, c.stream, n, 0.0, c.d_f.as<double>(), 0.0,
                     (const double *)nullptr, c.d_u.as<double>());
  int64_t total_it = 0;
  int outer = 0;
  double rel = 1.0;
  bool conv = false;
  const int batch = prm.check_every > 0 ? prm.check_every : 50;
  auto true_res = [&]() -> int {
    DISPATCH_VL(vl, hipLaunchKernelGGL(k_residual64<VL>, dim3(gr), dim3(TPB), 0, c.stream, n,
                                       c.d_fia.as<int64_t>(), c.d_fja.as<int32_t>(),
                                       c.d_fa.as<double>(), c.d_u.as<double>(),
                                       c.d_f.as<double>(), c.d_r.as<double>(), out.as<double>(),
                                       part, cnt));
    double h[2];
    PZG_HIP(hipMemcpyAsync(h, out.p, sizeof(h), hipMemcpyDeviceToHost, c.stream));
    PZG_HIP(hipStreamSynchronize(c.stream));
    rel = h[1] > 0 ? std::sqrt(h[0] / h[1]) : std::sqrt(h[0]);
    return PZG_OK;
  };
  auto next_delta = [&]() {
    // reduction asked of the FP32 recursion before the next true residual
    double d = prm.inner_tol > 0 ? prm.inner_tol : 1e-4;
    double need = 0.5 * prm.tol / rel;
    if (need > d) d = need;
    return d > 0.5 ? 0.5 : d;
  };
  PZG_TRY(true_res());
  if (rel <= prm.tol) conv = true;
  bool restart = true;
  bool learned = prm.precond == PZG_PRECOND_LEARNED;
  int fallback = 0;
  double prev_rel = rel;
  SolveScalars hs;
  std::memset(&hs, 0, sizeof(hs));
  while (!conv) {
    if (restart) {
      hs.done = 0;
      hs.iter = total_it;
      hs.maxit = prm.max_iter;
      hs.delta = next_delta();
      PZG_HIP(hipMemcpyAsync(c.d_scal.p, &hs, sizeof(hs), hipMemcpyHostToDevice, c.stream));
      hipLaunchKernelGGL(k_fcg_init<BS>, dim3(gb), dim3(TPB), 0, c.stream, nb, c.d_r.as<double>(),
                         c.d_dinvsqrt.as<double>(), c.d_xs.as<float>(), c.d_rs.as<float>(),
                         c.d_zs.as<float>(), c.d_ps.as<float>(), c.d_bjinv.as<float>(), sc, part,
                         cnt);
      PZG_HIP(hipGetLastError());
      if (learned)
        PZG_TRY(nn_apply(c, c.d_rs.as<float>(), c.d_zs.as<float>(), c.d_ps.as<float>(), nullptr, sc,
                         part, cnt));
      restart = false;
    }
    auto iter = [&]() {
      int ga = gs;
      if (use_mf) {
        (void)mf_apply(c, c.d_ps.as<float>(), c.d_rs.as<float>(), c.d_qs.as<float>(), sc, part, &ga);
      } else {
        hipLaunchKernelGGL(k_fcg_spmv<BS>, dim3(gs), dim3(TPB), 0, c.stream, nb,
                           c.d_fbia.as<int64_t>(), c.d_fbja.as<int32_t>(), c.d_bsr.as<float>(),
                           c.nbf, c.d_ps.as<float>(), c.d_rs.as<float>(), c.d_qs.as<float>(), sc,
                           part, cnt);
      }
      hipLaunchKernelGGL(k_fin_fcg_alpha, dim3(1), dim3(TPB), 0, c.stream, sc, (const double *)part, ga);
      hipLaunchKernelGGL(k_fcg_update<BS>, dim3(gb), dim3(TPB), 0, c.stream, nb,
                         c.d_xs.as<float>(), c.d_rs.as<float>(), c.d_zs.as<float>(),
                         c.d_ps.as<float>(), c.d_qs.as<float>(), c.d_bjinv.as<float>(), sc, part,
                         cnt);
      if (learned)
        (void)nn_apply(c, c.d_rs.as<float>(), c.d_zs.as<float>(), nullptr, c.d_qs.as<float>(), sc,
                       part, cnt);
      hipLaunchKernelGGL(k_fin_fcg_beta, dim3(1), dim3(TPB), 0, c.stream, sc, (const double *)part, gb);
      hipLaunchKernelGGL(k_p_update32, dim3(gv), dim3(TPB), 0, c.stream, n, c.d_zs.as<float>(),
                         c.d_ps.as<float>(), sc);
    };
    PZG_TRY(run_batches(c, batch, prm.use_graph != 0, iter, hs));
    total_it = hs.iter;
    // fold the FP32 correction into the FP64 solution and recompute the true residual
    hipLaunchKernelGGL(k_ir_update, dim3(gv), dim3(TPB), 0, c.stream, n, c.d_xs.as<float>(),
                       c.d_dinvsqrt.as<double>(), c.d_u.as<double>());
    PZG_HIP(hipGetLastError());
    PZG_TRY(true_res());
    outer++;
    if (rel <= prm.tol) { conv = true; break; }
    if (hs.done == 2 || outer >= prm.max_outer) break;
    // stall of the learned preconditioner: fall back to block Jacobi and
    // restart from the true residual
    if (learned && (hs.done == 3 || rel > 0.9 * prev_rel)) {
      learned = false;
      fallback = 1;
      restart = true;
      prev_rel = rel;
      continue;
    }
    prev_rel = rel;
    if (hs.done == 3 || prm.refine == 0) { restart = true; continue; }
    hipLaunchKernelGGL(k_fcg_reset, dim3(gv), dim3(TPB), 0, c.stream, n, c.d_r.as<double>(),
                       c.d_dinvsqrt.as<double>(), c.d_rs.as<float>(), next_delta(), sc, part, cnt);
    PZG_HIP(hipGetLastError());
  }
  PZG_HIP(hipEventRecord(c.ev1, c.stream));
  PZG_HIP(hipEventSynchronize(c.ev1));
  float ms = 0;
  PZG_HIP(hipEventElapsedTime(&ms, c.ev0, c.ev1));
  dev_free(out);
  info.solve_ms = ms;
  info.iterations = total_it;
  info.outer = outer;
  info.true_residual = rel;
  info.achieved = rel;
  info.converged = conv ? 1 : 0;
  info.fallback = fallback;
  return conv ? PZG_OK : PZG_ENOCONV;
}

// (residual, error) pairs from the solver itself: random scaled solutions y*,
// b = S y*, block-Jacobi FCG from zero; at the iterations in rec the pair
// (R r_k, R(y* - y_k - B^-1 r_k)) is stored.
template <int BS>
static int collect_impl(Ctx &c, int nrhs, const int32_t *rec, int nrec, uint64_t seed) {
  const int64_t n = c.neq, nb = c.nb;
  if (!c.full_valid) PZG_TRY(launch_expand_full(c));
  if (!c.bsr_valid) PZG_TRY(launch_build_bsr(c));
  PZG_TRY(ensure_reduce(c));
  for (DevBuf *b : {&c.d_xs, &c.d_rs, &c.d_zs, &c.d_ps, &c.d_qs, &c.d_bs, &c.d_nn_ystar, &c.d_nn_e})
    PZG_TRY(dev_alloc(*b, n * sizeof(float)));
  int kmax = 0;
  for (int k = 0; k < nrec; k++) kmax = rec[k] > kmax ? rec[k] : kmax;
  std::vector<float> ys(n);
  uint64_t st = seed * 6364136223846793005ULL + 1442695040888963407ULL;
  SolveScalars *sc = c.d_scal.as<SolveScalars>();
  double *part = c.d_partials.as<double>();
  unsigned *cnt = c.d_counter.as<unsigned>();
  const int gv = grid_for(n, TPB), gb = grid_for(nb, TPB), gs = grid_for(nb, TPB / 64);
  for (int j = 0; j < nrhs; j++) {
    for (int64_t i = 0; i < n; i++) {  // Box-Muller on a 64-bit LCG
      st = st * 6364136223846793005ULL + 1442695040888963407ULL;
      double u1 = ((st >> 11) + 1.0) / 9007199254740993.0;
      st = st * 6364136223846793005ULL + 1442695040888963407ULL;
      double u2 = (st >> 11) / 9007199254740992.0;
      ys[i] = (float)(std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2));
    }
    PZG_HIP(hipMemcpyAsync(c.d_nn_ystar.p, ys.data(), n * sizeof(float), hipMemcpyHostToDevice, c.stream));
    PZG_TRY(launch_spmv_bsr(c, c.d_nn_ystar.as<float>(), c.d_bs.as<float>()));
    SolveScalars hs;
    std::memset(&hs, 0, sizeof(hs));
    hs.maxit = (int64_t)1 << 40;
    hs.delta = 0.;
    PZG_HIP(hipMemcpyAsync(c.d_scal.p, &hs, sizeof(hs), hipMemcpyHostToDevice, c.stream));
    hipLaunchKernelGGL(k_fcg_start<BS>, dim3(gb), dim3(TPB), 0, c.stream, nb, c.d_bs.as<float>(),
                       c.d_xs.as<float>(), c.d_rs.as<float>(), c.d_zs.as<float>(), c.d_ps.as<float>(),
                       c.d_bjinv.as<float>(), sc, part, cnt);
    for (int k = 1; k <= kmax; k++) {
      hipLaunchKernelGGL(k_fcg_spmv<BS>, dim3(gs), dim3(TPB), 0, c.stream, nb, c.d_fbia.as<int64_t>(),
                         c.d_fbja.as<int32_t>(), c.d_bsr.as<float>(), c.nbf, c.d_ps.as<float>(),
                         c.d_rs.as<float>(), c.d_qs.as<float>(), sc, part, cnt);
      hipLaunchKernelGGL(k_fin_fcg_alpha, dim3(1), dim3(TPB), 0, c.stream, sc, (const double *)part, gs);
      hipLaunchKernelGGL(k_fcg_update<BS>, dim3(gb), dim3(TPB), 0, c.stream, nb, c.d_xs.as<float>(),
                         c.d_rs.as<float>(), c.d_zs.as<float>(), c.d_ps.as<float>(), c.d_qs.as<float>(),
                         c.d_bjinv.as<float>(), sc, part, cnt);
      hipLaunchKernelGGL(k_fin_fcg_beta, dim3(1), dim3(TPB), 0, c.stream, sc, (const double *)part, gb);
      hipLaunchKernelGGL(k_p_update32, dim3(gv), dim3(TPB), 0, c.stream, n, c.d_zs.as<float>(),
                         c.d_ps.as<float>(), sc);
      for (int t = 0; t < nrec; t++)
        if (rec[t] == k) {
          if (c.nn_nsamp >= c.nn_maxsamp) {
            set_error("learned preconditioner: sample capacity exceeded");
            return PZG_EINVAL;
          }
          hipLaunchKernelGGL(k_sub3, dim3(gv), dim3(TPB), 0, c.stream, n, c.d_nn_ystar.as<float>(),
                             c.d_xs.as<float>(), c.d_zs.as<float>(), c.d_nn_e.as<float>());
          PZG_TRY(nn_store_pair(c, c.d_rs.as<float>(), c.d_nn_e.as<float>(), c.nn_nsamp));
          c.nn_nsamp++;
        }
    }
    PZG_HIP(hipGetLastError());
    PZG_HIP(hipStreamSynchronize(c.stream));
    SolveScalars chk;
    PZG_HIP(hipMemcpy(&chk, c.d_scal.p, sizeof(chk), hipMemcpyDeviceToHost));
    if (chk.done == 3) {
      set_error("learned preconditioner: pair generation diverged");
      return PZG_EINVAL;
    }
  }
  return PZG_OK;
}

int nn_collect(Ctx &c, int nrhs, const int32_t *rec, int nrec, uint64_t seed) {
  if (!c.nn_ready) {
    set_error("learned preconditioner: call pzg_learned_setup first");
    return PZG_ESTATE;
  }
  if (c.bs == 3) return collect_impl<3>(c, nrhs, rec, nrec, seed);
  return collect_impl<1>(c, nrhs, rec, nrec, seed);
}

int solve_mixed(Ctx &c, const pzg_solve_params &prm, pzg_solve_info &info) {
  if (prm.precond == PZG_PRECOND_LEARNED && !c.nn_ready) {
    set_error("pzg_solve: learned preconditioner not set up in this context");
    return PZG_ESTATE;
  }
  if (c.bs == 3) return mixed_impl<3>(c, prm, info);
  if (c.bs == 1) return mixed_impl<1>(c, prm, info);
  set_error("pzg_solve: unsupported block size for the mixed solver");
  return PZG_EUNSUP;
}

}  // namespace pzg
