// pzg_penalty.hip -- the Dirichlet penalty rows of the learned solve path.
//
// Dirichlet data enter as a penalty (TPZBndCondT type 0 and the directional
// types of ContributeBC): fBigNumber times a boundary mass matrix added to the
// rows of the boundary equations.  Scaled by D^-1/2 these rows form a block
// S_pp ~ D_p^-1/2 (big M_face) D_p^-1/2 that couples to the other rows only at
// the 1/sqrt(big) level, while the stopping test on the unscaled residual
// ||f - K u|| weighs them by sqrt(big).  The block-Jacobi scaled S_pp is badly
// conditioned (hierarchical face mass matrix), and the learned cycle converges
// much faster on the other rows than on it.  The penalty rows are therefore
// taken out of the FP32 inner iterations of the learned path (zero residual and
// zero preconditioned residual there) and solved in FP64 by the outer
// refinement: after every fold of the FP32 correction
//     K_pp d = (f - K u)_p,   u_p += d
// by conjugate gradients on K_pp (the rows and columns of the penalized
// equations of the FP64 full block CSR), preconditioned by the additive Schwarz
// sum over the penalized boundary elements of the inverses of K_pp restricted
// to each element's penalized equations.  The FP64 true residual of the whole
// system, recomputed afterwards, stays the stopping test.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "pzg_device.h"
#include "pzg_internal.h"

namespace pzg {

namespace {

constexpr int TPB = RED_TPB;
constexpr int PEN_BATCH = 16;    // CG iterations between host checks
constexpr int PEN_MAXIT = 1000;  // CG iterations per solve at most

struct PenScal {
  double rz, pq, alpha, beta, rr, stop2, rr0;
  int64_t it;
  int done;
};

int pen_grid(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>((n + TPB - 1) / TPB, MAXBLK)); }

__global__ void k_pen_zero_int(int *p) { *p = 0; }

// entries of K_pp per penalty row: columns of its block row that are penalized
__global__ void k_pen_count(int64_t np, int bs, const int64_t *__restrict__ idx,
                            const int64_t *__restrict__ fbia, const int32_t *__restrict__ fbja,
                            const int32_t *__restrict__ pmap, int64_t *__restrict__ cnt) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= np) return;
  const int64_t A = idx[k] / bs;
  int64_t m = 0;
  for (int64_t q = fbia[A]; q < fbia[A + 1]; q++)
    for (int l = 0; l < bs; l++) m += pmap[(int64_t)fbja[q] * bs + l] >= 0;
  cnt[k] = m;
}

// K_pp in CSR over the penalty numbering (columns ascending: blocks ascend
// along the full block row and the numbering follows the equations)
template <int BS>
__global__ void k_pen_fill(int64_t np, const int64_t *__restrict__ idx, const int64_t *__restrict__ fbia,
                           const int32_t *__restrict__ fbja, const double *__restrict__ fa,
                           const int32_t *__restrict__ pmap, const int64_t *__restrict__ ia,
                           int32_t *__restrict__ ja, double *__restrict__ a) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= np) return;
  const int64_t A = idx[k] / BS;
  const int ci = (int)(idx[k] % BS);
  const int64_t b0 = fbia[A], n = fbia[A + 1] - b0;
  int64_t o = ia[k];
  for (int64_t q = 0; q < n; q++)
    for (int l = 0; l < BS; l++) {
      const int32_t pj = pmap[(int64_t)fbja[b0 + q] * BS + l];
      if (pj < 0) continue;
      ja[o] = pj;
      a[o] = fa[bsr_pos<BS * BS>(b0, n, q, ci * BS + l)];
      o++;
    }
}

// Inverse of K_pp restricted to each Schwarz set (values at the CSR positions
// sval, -1 = structurally zero), symmetrised, by Gauss-Jordan elimination
// without pivoting (symmetric positive definite), one workgroup per set.
constexpr int PEN_MAXNS = 64;
__global__ __launch_bounds__(64) void k_pen_inv(const int64_t *__restrict__ sptr, const int64_t *__restrict__ soff,
                                                const int32_t *__restrict__ sval, const double *__restrict__ a,
                                                double *__restrict__ sinv, int *__restrict__ fail) {
  __shared__ double M[PEN_MAXNS * PEN_MAXNS];
  const int64_t set = blockIdx.x;
  const int ns = (int)(sptr[set + 1] - sptr[set]);
  const int64_t o = soff[set];
  const int tid = threadIdx.x;
  for (int t = tid; t < ns * ns; t += 64) {
    const int32_t q = sval[o + t];
    M[t] = q >= 0 ? a[q] : 0.;
  }
  __syncthreads();
  for (int t = tid; t < ns * ns; t += 64) {  // symmetric part
    const int i = t / ns, j = t % ns;
    if (i < j) {
      const double v = 0.5 * (M[i * ns + j] + M[j * ns + i]);
      M[i * ns + j] = v;
      M[j * ns + i] = v;
    }
  }
  __syncthreads();
  for (int k = 0; k < ns; k++) {
    const double piv = M[k * ns + k];
    __syncthreads();
    if (!(piv > 0.)) {
      if (tid == 0) atomicExch(fail, 1);
      return;
    }
    // row k /= piv (M[k][k] becomes 1/piv); then eliminate column k elsewhere
    for (int j = tid; j < ns; j += 64) M[k * ns + j] = j == k ? 1. / piv : M[k * ns + j] / piv;
    __syncthreads();
    for (int t = tid; t < ns * ns; t += 64) {
      const int i = t / ns, j = t % ns;
      if (i == k) continue;
      const double f = M[i * ns + k];
      if (j == k) continue;
      M[i * ns + j] -= f * M[k * ns + j];
    }
    __syncthreads();
    for (int i = tid; i < ns; i += 64)
      if (i != k) M[i * ns + k] = -M[i * ns + k] * M[k * ns + k];
    __syncthreads();
  }
  for (int t = tid; t < ns * ns; t += 64) sinv[o + t] = M[t];
}

// r = r64[idx]; x = 0; partial r.r
__global__ __launch_bounds__(TPB) void k_pen_gather(int64_t np, const int64_t *__restrict__ idx,
                                                    const double *__restrict__ r64, double *__restrict__ r,
                                                    double *__restrict__ x, double *partials) {
  double s = 0.;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < np; k += (int64_t)gridDim.x * blockDim.x) {
    const double v = r64[idx[k]];
    r[k] = v;
    x[k] = 0.;
    s += v * v;
  }
  double vv[1] = {s};
  block_partials<1>(vv, partials);
}

// stop: |r_p| <= max(stop_abs, frac |r_other|), |r_other|^2 = tot2 - |r_p|^2
__global__ void k_pen_start(PenScal *sc, const double *partials, int nblk, double stop_abs, double frac,
                            double tot2) {
  const double rr = sum_partials(partials, 0, nblk);
  if (threadIdx.x == 0) {
    sc->rr = rr;
    sc->rr0 = rr;
    sc->stop2 = std::max(std::max(stop_abs * stop_abs, frac * frac * (tot2 - rr)), 1e-30 * rr);
    sc->it = 0;
    sc->rz = 0.;
    sc->done = rr <= sc->stop2 ? 1 : 0;
  }
}

// y[slot] = (K_e^-1 r_e)[i] for slot = (set, i)
__global__ void k_pen_asm(int64_t nslot, const int32_t *__restrict__ sset, const int64_t *__restrict__ sptr,
                          const int32_t *__restrict__ sidx, const int64_t *__restrict__ soff,
                          const double *__restrict__ sinv, const double *__restrict__ r, double *__restrict__ y,
                          const PenScal *sc) {
  if (sc->done) return;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nslot) return;
  const int32_t s = sset[t];
  const int64_t b = sptr[s];
  const int ns = (int)(sptr[s + 1] - b), i = (int)(t - b);
  const double *inv = sinv + soff[s] + (int64_t)i * ns;
  double v = 0.;
  for (int j = 0; j < ns; j++) v += inv[j] * r[sidx[b + j]];
  y[t] = v;
}

// z = sum of the slots of each row; partial r.z
__global__ __launch_bounds__(TPB) void k_pen_gat(int64_t np, const int64_t *__restrict__ gptr,
                                                 const int64_t *__restrict__ gslot, const double *__restrict__ y,
                                                 const double *__restrict__ r, double *__restrict__ z,
                                                 const PenScal *sc, double *partials) {
  if (sc->done) return;
  double s = 0.;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < np; k += (int64_t)gridDim.x * blockDim.x) {
    double v = 0.;
    for (int64_t q = gptr[k]; q < gptr[k + 1]; q++) v += y[gslot[q]];
    z[k] = v;
    s += r[k] * v;
  }
  double vv[1] = {s};
  block_partials<1>(vv, partials);
}

// beta = r.z / previous r.z (0 on the first direction)
__global__ void k_pen_beta(PenScal *sc, const double *partials, int nblk) {
  if (sc->done) return;
  const double rz = sum_partials(partials, 0, nblk);
  if (threadIdx.x == 0) {
    sc->beta = sc->rz != 0. ? rz / sc->rz : 0.;
    sc->rz = rz;
  }
}

// p = z + beta p  (p = z on the first direction: p holds no values yet)
__global__ void k_pen_p(int64_t np, const double *__restrict__ z, double *__restrict__ p, const PenScal *sc,
                        int first) {
  if (sc->done) return;
  const double beta = sc->beta;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < np; k += (int64_t)gridDim.x * blockDim.x)
    p[k] = first ? z[k] : z[k] + beta * p[k];
}

// q = K_pp p; partial p.q  (PEN_L lanes per row)
constexpr int PEN_L = 8;
__global__ __launch_bounds__(TPB) void k_pen_spmv(int64_t np, const int64_t *__restrict__ ia,
                                                  const int32_t *__restrict__ ja, const double *__restrict__ a,
                                                  const double *__restrict__ p, double *__restrict__ q,
                                                  const PenScal *sc, double *partials) {
  if (sc->done) return;
  const int sub = threadIdx.x % PEN_L;
  double s = 0.;
  for (int64_t k = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / PEN_L; k < np;
       k += (int64_t)gridDim.x * blockDim.x / PEN_L) {
    double v = 0.;
    for (int64_t o = ia[k] + sub; o < ia[k + 1]; o += PEN_L) v += a[o] * p[ja[o]];
    for (int m = PEN_L / 2; m > 0; m >>= 1) v += __shfl_xor(v, m, PEN_L);
    if (sub == 0) {
      q[k] = v;
      s += p[k] * v;
    }
  }
  double vv[1] = {s};
  block_partials<1>(vv, partials);
}

__global__ void k_pen_alpha(PenScal *sc, const double *partials, int nblk) {
  if (sc->done) return;
  const double pq = sum_partials(partials, 0, nblk);
  if (threadIdx.x == 0) {
    sc->pq = pq;
    sc->alpha = pq > 0. ? sc->rz / pq : 0.;
    if (!(pq > 0.)) sc->done = 3;  // breakdown: K_pp not positive definite on p
  }
}

// x += alpha p; r -= alpha q; partial r.r
__global__ __launch_bounds__(TPB) void k_pen_update(int64_t np, double *__restrict__ x, double *__restrict__ r,
                                                    const double *__restrict__ p, const double *__restrict__ q,
                                                    const PenScal *sc, double *partials) {
  if (sc->done) return;
  const double al = sc->alpha;
  double s = 0.;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < np; k += (int64_t)gridDim.x * blockDim.x) {
    x[k] += al * p[k];
    const double v = r[k] - al * q[k];
    r[k] = v;
    s += v * v;
  }
  double vv[1] = {s};
  block_partials<1>(vv, partials);
}

__global__ void k_pen_check(PenScal *sc, const double *partials, int nblk, int64_t maxit) {
  if (sc->done) return;
  const double rr = sum_partials(partials, 0, nblk);
  if (threadIdx.x == 0) {
    sc->rr = rr;
    sc->it += 1;
    if (rr <= sc->stop2) sc->done = 1;
    else if (sc->it >= maxit) sc->done = 2;
  }
}

// u[idx] += x; r64[idx] = r (the penalty rows' residual after the solve)
__global__ void k_pen_scatter(int64_t np, const int64_t *__restrict__ idx, const double *__restrict__ x,
                              const double *__restrict__ r, double *__restrict__ u, double *__restrict__ r64) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < np; k += (int64_t)gridDim.x * blockDim.x) {
    u[idx[k]] += x[k];
    r64[idx[k]] = r[k];
  }
}

// u_p += x; delta_p = x (delta is zero off the penalty rows)
__global__ void k_pen_scatter_d(int64_t np, const int64_t *__restrict__ idx, const double *__restrict__ x,
                                double *__restrict__ u, double *__restrict__ delta) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < np; k += (int64_t)gridDim.x * blockDim.x) {
    u[idx[k]] += x[k];
    delta[idx[k]] = x[k];
  }
}

// the rows whose block row holds a penalized block: flag[A] = 1
__global__ void k_pen_mark(int64_t npb, const int32_t *__restrict__ pblk, const int64_t *__restrict__ fbia,
                           const int32_t *__restrict__ fbja, unsigned char *__restrict__ flag) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= npb) return;
  const int64_t B = pblk[k];
  for (int64_t q = fbia[B]; q < fbia[B + 1]; q++) flag[fbja[q]] = 1;  // symmetric pattern
}

// r -= K delta on the listed block rows; delta is zero off the penalty rows,
// so this is the exact change of the residual f - K u from u_p += delta_p.
// Sixteen lanes per row, lane sub taking the row's blocks i = sub mod 16
// (one value chunk of the block CSR per pass, as the block SpMV reads it).
template <int BS>
__global__ void k_pen_rupd(int64_t nr, const int32_t *__restrict__ rows, const int64_t *__restrict__ fbia,
                           const int32_t *__restrict__ fbja, const double *__restrict__ fa,
                           const double *__restrict__ delta, double *__restrict__ r) {
  constexpr int BB = BS * BS;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t gi = t >> 4;
  const int sub = (int)(t & 15);
  if (gi >= nr) return;  // whole 16-lane groups
  const int64_t A = rows[gi], b0 = fbia[A], n = fbia[A + 1] - b0;
  double acc[BS];
#pragma unroll
  for (int k = 0; k < BS; k++) acc[k] = 0.;
  for (int64_t i = sub; i < n; i += BSR_CHUNK) {
    const int64_t j = i - sub, mc = n - j < BSR_CHUNK ? n - j : BSR_CHUNK;
    const double *vp = fa + (b0 + j) * BB + sub;
    const int64_t B = fbja[b0 + i];
    double dl[BS];
#pragma unroll
    for (int l = 0; l < BS; l++) dl[l] = delta[B * BS + l];
#pragma unroll
    for (int k = 0; k < BS; k++)
#pragma unroll
      for (int l = 0; l < BS; l++) acc[k] += vp[(k * BS + l) * mc] * dl[l];
  }
  double mine = 0.;
#pragma unroll
  for (int k = 0; k < BS; k++) {
    double s = acc[k];
    for (int o = BSR_CHUNK / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, BSR_CHUNK);
    if (sub == k) mine = s;
  }
  if (sub < BS) r[A * BS + sub] -= mine;
}

// ||r||^2 over n entries: per-block partials, then one block sums them
__global__ __launch_bounds__(TPB) void k_pen_norm2(int64_t n, const double *__restrict__ r, double *partials) {
  double s = 0.;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    s += r[i] * r[i];
  double v[1] = {s};
  block_partials<1>(v, partials);
}
__global__ void k_pen_norm2_fin(const double *partials, int nblk, double *out) {
  const double s = sum_partials(partials, 0, nblk);
  if (threadIdx.x == 0) *out = s;
}

// r (float, scaled residual of the inner iterations) = 0 on the penalty rows
__global__ void k_pen_zero32(int64_t np, const int64_t *__restrict__ idx, float *__restrict__ r) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < np; k += (int64_t)gridDim.x * blockDim.x)
    r[idx[k]] = 0.f;
}

}  // namespace

static int pen_fill(Ctx &c);

// Structure of the penalty block, once per mesh (the full block pattern and
// the penalty flags of the learned element graph): the penalized equations,
// K_pp's CSR pattern, the Schwarz sets of the penalized boundary elements with
// the CSR position of every entry of their dense blocks, and the row -> slot
// gather.
static int pen_structure(Ctx &c) {
  const int bs = c.bs;
  std::vector<int64_t> idx;
  std::vector<int32_t> pmap(c.neq, -1);
  for (int64_t i = 0; i < c.neq; i++)
    if (c.les_pen[i] != 0.f) {
      pmap[i] = (int32_t)idx.size();
      idx.push_back(i);
    }
  const int64_t np = (int64_t)idx.size();
  c.pen_n = np;
  if (np == 0) {
    c.pen_struct = true;
    return PZG_OK;
  }
  hipStream_t s = c.stream;
  PZG_TRY(dev_upload(c.d_pen_idx, idx.data(), idx.size(), s));
  PZG_TRY(dev_upload(c.d_pen_pmap, pmap.data(), pmap.size(), s));
  DevBuf d_cnt;
  PZG_TRY(dev_alloc(d_cnt, sizeof(int64_t) * np));
  std::vector<int64_t> ia(np + 1, 0);
  const int g1 = (int)((np + TPB - 1) / TPB);
  hipLaunchKernelGGL(k_pen_count, dim3(g1), dim3(TPB), 0, s, np, bs, c.d_pen_idx.as<int64_t>(),
                     c.d_fbia.as<int64_t>(), c.d_fbja.as<int32_t>(), c.d_pen_pmap.as<int32_t>(), d_cnt.as<int64_t>());
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(ia.data() + 1, d_cnt.p, sizeof(int64_t) * np, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  dev_free(d_cnt);
  if (e != hipSuccess) {
    set_error(std::string("penalty rows: counting the K_pp entries: ") + hipGetErrorString(e));
    return PZG_EHIP;
  }
  for (int64_t k = 0; k < np; k++) ia[k + 1] += ia[k];
  const int64_t nnz = ia[np];
  if (nnz >= ((int64_t)1 << 31)) {
    set_error("penalty rows: K_pp too large");
    return PZG_EUNSUP;
  }
  c.pen_nnz = nnz;
  PZG_TRY(dev_upload(c.d_pen_ia, ia.data(), ia.size(), s));
  PZG_TRY(dev_alloc(c.d_pen_ja, sizeof(int32_t) * std::max<int64_t>(nnz, 1)));
  PZG_TRY(dev_alloc(c.d_pen_a, sizeof(double) * std::max<int64_t>(nnz, 1)));
  PZG_TRY(pen_fill(c));
  std::vector<int32_t> ja(std::max<int64_t>(nnz, 1));
  PZG_HIP(hipMemcpyAsync(ja.data(), c.d_pen_ja.p, sizeof(int32_t) * nnz, hipMemcpyDeviceToHost, s));
  PZG_HIP(hipStreamSynchronize(s));
  // Schwarz sets: the penalized equations of every boundary element
  const ElemTables &B = c.tbc;
  std::vector<int64_t> sptr(1, 0), soff(1, 0);
  std::vector<int32_t> sidx, sset, sval;
  std::vector<int32_t> loc;
  for (int64_t el = 0; el < c.nbc; el++) {
    loc.clear();
    for (int sh = 0; sh < B.nsh; sh++) {
      const int64_t blk = c.aidx[c.bc_conn[el * c.bc_nsides + B.shape_side[sh]]];
      if (blk < 0) continue;
      for (int k = 0; k < bs; k++)
        if (pmap[blk * bs + k] >= 0) loc.push_back(pmap[blk * bs + k]);
    }
    std::sort(loc.begin(), loc.end());
    loc.erase(std::unique(loc.begin(), loc.end()), loc.end());
    const int ns = (int)loc.size();
    if (ns == 0) continue;
    if (ns > PEN_MAXNS) {
      set_error("penalty rows: a boundary element holds more than 64 penalized equations");
      return PZG_EUNSUP;
    }
    const int32_t sid = (int32_t)(sptr.size() - 1);
    for (int i = 0; i < ns; i++) {
      sidx.push_back(loc[i]);
      sset.push_back(sid);
      const int32_t *rb = ja.data() + ia[loc[i]], *re = ja.data() + ia[loc[i] + 1];
      for (int j = 0; j < ns; j++) {
        const int32_t *f = std::lower_bound(rb, re, loc[j]);
        sval.push_back(f != re && *f == loc[j] ? (int32_t)(ia[loc[i]] + (f - rb)) : -1);
      }
    }
    sptr.push_back((int64_t)sidx.size());
    soff.push_back((int64_t)sval.size());
  }
  const int64_t nslot = (int64_t)sidx.size();
  // every penalized equation lies in some boundary element (ContributeBC adds
  // the penalty there); one outside would keep z = 0: refuse
  std::vector<int64_t> gptr(np + 1, 0), gslot(std::max<int64_t>(nslot, 1));
  for (int64_t t = 0; t < nslot; t++) gptr[sidx[t] + 1]++;
  for (int64_t k = 0; k < np; k++) {
    if (gptr[k + 1] == 0) {
      set_error("penalty rows: a penalized equation outside every boundary element");
      return PZG_EUNSUP;
    }
    gptr[k + 1] += gptr[k];
  }
  {
    std::vector<int64_t> fill(gptr.begin(), gptr.end() - 1);
    for (int64_t t = 0; t < nslot; t++) gslot[fill[sidx[t]]++] = t;
  }
  c.pen_nset = (int64_t)sptr.size() - 1;
  c.pen_nslot = nslot;
  PZG_TRY(dev_upload(c.d_pen_sptr, sptr.data(), sptr.size(), s));
  PZG_TRY(dev_upload(c.d_pen_sidx, sidx.data(), sidx.size(), s));
  PZG_TRY(dev_upload(c.d_pen_sset, sset.data(), sset.size(), s));
  PZG_TRY(dev_upload(c.d_pen_soff, soff.data(), soff.size(), s));
  PZG_TRY(dev_upload(c.d_pen_sval, sval.data(), sval.size(), s));
  PZG_TRY(dev_alloc(c.d_pen_sinv, sizeof(double) * std::max<size_t>(sval.size(), 1)));
  PZG_TRY(dev_upload(c.d_pen_gptr, gptr.data(), gptr.size(), s));
  PZG_TRY(dev_upload(c.d_pen_gslot, gslot.data(), gslot.size(), s));
  PZG_TRY(dev_alloc(c.d_pen_w, sizeof(double) * (5 * np + nslot)));
  PZG_TRY(dev_alloc(c.d_pen_scal, sizeof(PenScal) + 16));
  // the block rows that a change of the penalized unknowns reaches (the
  // columns of the penalized block rows; the pattern is symmetric), and the
  // FP64 change itself, zero off the penalized rows
  {
    std::vector<int32_t> pblk;
    for (int64_t k = 0; k < np; k++)
      if (pblk.empty() || pblk.back() != (int32_t)(idx[k] / bs)) pblk.push_back((int32_t)(idx[k] / bs));
    const int64_t nbr = c.neq / bs;
    DevBuf d_pblk, d_flag;
    PZG_TRY(dev_upload(d_pblk, pblk.data(), pblk.size(), s));
    PZG_TRY(dev_alloc(d_flag, (size_t)nbr));
    PZG_HIP(hipMemsetAsync(d_flag.p, 0, (size_t)nbr, s));
    const int64_t npb = (int64_t)pblk.size();
    hipLaunchKernelGGL(k_pen_mark, dim3((unsigned)((npb + TPB - 1) / TPB)), dim3(TPB), 0, s, npb,
                       d_pblk.as<int32_t>(), c.d_fbia.as<int64_t>(), c.d_fbja.as<int32_t>(),
                       d_flag.as<unsigned char>());
    PZG_HIP(hipGetLastError());
    std::vector<unsigned char> flag(nbr);
    PZG_HIP(hipMemcpyAsync(flag.data(), d_flag.p, (size_t)nbr, hipMemcpyDeviceToHost, s));
    PZG_HIP(hipStreamSynchronize(s));
    std::vector<int32_t> arow;
    for (int64_t A = 0; A < nbr; A++)
      if (flag[A]) arow.push_back((int32_t)A);
    c.pen_narow = (int64_t)arow.size();
    PZG_TRY(dev_upload(c.d_pen_arow, arow.data(), arow.size(), s));
    PZG_TRY(dev_alloc(c.d_pen_delta, sizeof(double) * c.neq));
    PZG_HIP(hipMemsetAsync(c.d_pen_delta.p, 0, sizeof(double) * c.neq, s));
  }
  PZG_HIP(hipStreamSynchronize(s));
  c.pen_struct = true;
  return PZG_OK;
}

// K_pp values of the current assembly (pattern of pen_structure)
static int pen_fill(Ctx &c) {
  const int64_t np = c.pen_n;
  const int g1 = (int)((np + TPB - 1) / TPB);
  if (c.bs == 3)
    hipLaunchKernelGGL(k_pen_fill<3>, dim3(g1), dim3(TPB), 0, c.stream, np, c.d_pen_idx.as<int64_t>(),
                       c.d_fbia.as<int64_t>(), c.d_fbja.as<int32_t>(), c.d_fa.as<double>(),
                       c.d_pen_pmap.as<int32_t>(), c.d_pen_ia.as<int64_t>(), c.d_pen_ja.as<int32_t>(),
                       c.d_pen_a.as<double>());
  else
    hipLaunchKernelGGL(k_pen_fill<1>, dim3(g1), dim3(TPB), 0, c.stream, np, c.d_pen_idx.as<int64_t>(),
                       c.d_fbia.as<int64_t>(), c.d_fbja.as<int32_t>(), c.d_fa.as<double>(),
                       c.d_pen_pmap.as<int32_t>(), c.d_pen_ia.as<int64_t>(), c.d_pen_ja.as<int32_t>(),
                       c.d_pen_a.as<double>());
  PZG_HIP(hipGetLastError());
  return PZG_OK;
}

// The penalty block of the current assembly: structure once per mesh, then
// K_pp and the Schwarz inverses on the device.  Needs les_prepare (penalty
// flags) and the full block CSR of the current assembly.
int pen_setup(Ctx &c) {
  c.pen_ready = false;
  if (!c.les_graph || (int64_t)c.les_pen.size() != c.neq || !c.full_valid) {
    set_error("penalty rows: learned element graph or full block CSR missing");
    return PZG_ESTATE;
  }
  if (c.bs != 1 && c.bs != 3) {
    set_error("penalty rows: nstate must be 1 or 3");
    return PZG_EUNSUP;
  }
  if (!c.pen_struct) {
    PZG_TRY(pen_structure(c));
  } else if (c.pen_n > 0) {
    PZG_TRY(pen_fill(c));
  }
  if (c.pen_n == 0) {
    c.pen_ready = true;
    return PZG_OK;
  }
  int *fail = reinterpret_cast<int *>(reinterpret_cast<char *>(c.d_pen_scal.p) + sizeof(PenScal));
  hipLaunchKernelGGL(k_pen_zero_int, dim3(1), dim3(1), 0, c.stream, fail);
  hipLaunchKernelGGL(k_pen_inv, dim3((unsigned)c.pen_nset), dim3(64), 0, c.stream, c.d_pen_sptr.as<int64_t>(),
                     c.d_pen_soff.as<int64_t>(), c.d_pen_sval.as<int32_t>(), c.d_pen_a.as<double>(),
                     c.d_pen_sinv.as<double>(), fail);
  PZG_HIP(hipGetLastError());
  int hf = 0;
  PZG_HIP(hipMemcpyAsync(&hf, fail, sizeof(int), hipMemcpyDeviceToHost, c.stream));
  PZG_HIP(hipStreamSynchronize(c.stream));
  if (hf) {
    set_error("penalty rows: a boundary element's penalty block is not positive definite");
    return PZG_EUNSUP;
  }
  c.pen_ready = true;
  return PZG_OK;
}

// u_p += d = K_pp^-1 r_p (r64: the FP64 residual f - K u of the whole system),
// by Schwarz-preconditioned CG until ||r_p - K_pp d|| <= max(stop_abs, frac
// ||r_other||) (tot2 = ||r||^2 of the whole residual: the other rows' part is
// what the next FP32 iterations leave, and solving past it gains nothing).
// With rr_full, r64 becomes the exact residual of the new u (r64 -= K delta on
// the block rows that the penalized columns reach, delta = d on the penalized
// rows and zero elsewhere) and *rr_full its squared norm (-1 when u did not
// change); without it, r64's penalty rows receive r_p - K_pp d and its other
// rows keep the -K_ip d change for the next true residual.  *its: the CG
// iterations taken; *rr0, *rr1: |r_p|^2 before and after.
int pen_solve(Ctx &c, double *r64, double *u64, double stop_abs, bool use_graph, int64_t *its, double *rr0,
              double *rr1, double frac, double tot2, double *rr_full) {
  if (its) *its = 0;
  if (rr0) *rr0 = 0.;
  if (rr1) *rr1 = 0.;
  if (!c.pen_ready) {
    set_error("penalty rows: not set up");
    return PZG_ESTATE;
  }
  const int64_t np = c.pen_n;
  if (np == 0) return PZG_OK;
  hipStream_t s = c.stream;
  double *w = c.d_pen_w.as<double>();
  double *r = w, *x = w + np, *z = w + 2 * np, *p = w + 3 * np, *q = w + 4 * np, *y = w + 5 * np;
  PenScal *sc = c.d_pen_scal.as<PenScal>();
  double *part = c.d_partials.as<double>();
  const int g = pen_grid(np);
  const int gq = pen_grid(np * PEN_L);  // the product: PEN_L lanes per row
  const int gs = (int)((c.pen_nslot + TPB - 1) / TPB);
  const int64_t *idx = c.d_pen_idx.as<int64_t>();
  auto precond = [&](int first) {
    hipLaunchKernelGGL(k_pen_asm, dim3(std::max(gs, 1)), dim3(TPB), 0, s, c.pen_nslot, c.d_pen_sset.as<int32_t>(),
                       c.d_pen_sptr.as<int64_t>(), c.d_pen_sidx.as<int32_t>(), c.d_pen_soff.as<int64_t>(),
                       c.d_pen_sinv.as<double>(), (const double *)r, y, (const PenScal *)sc);
    hipLaunchKernelGGL(k_pen_gat, dim3(g), dim3(TPB), 0, s, np, c.d_pen_gptr.as<int64_t>(),
                       c.d_pen_gslot.as<int64_t>(), (const double *)y, (const double *)r, z, (const PenScal *)sc,
                       part);
    hipLaunchKernelGGL(k_pen_beta, dim3(1), dim3(TPB), 0, s, sc, (const double *)part, g);
    hipLaunchKernelGGL(k_pen_p, dim3(g), dim3(TPB), 0, s, np, (const double *)z, p, (const PenScal *)sc, first);
  };
  hipLaunchKernelGGL(k_pen_gather, dim3(g), dim3(TPB), 0, s, np, idx, r64, r, x, part);
  hipLaunchKernelGGL(k_pen_start, dim3(1), dim3(TPB), 0, s, sc, (const double *)part, g, stop_abs, frac, tot2);
  precond(1);
  PZG_HIP(hipGetLastError());
  PenScal hs{};
  auto batch = [&]() {
    for (int k = 0; k < PEN_BATCH; k++) {
      hipLaunchKernelGGL(k_pen_spmv, dim3(gq), dim3(TPB), 0, s, np, c.d_pen_ia.as<int64_t>(),
                         c.d_pen_ja.as<int32_t>(), c.d_pen_a.as<double>(), (const double *)p, q,
                         (const PenScal *)sc, part);
      hipLaunchKernelGGL(k_pen_alpha, dim3(1), dim3(TPB), 0, s, sc, (const double *)part, gq);
      hipLaunchKernelGGL(k_pen_update, dim3(g), dim3(TPB), 0, s, np, x, r, (const double *)p, (const double *)q,
                         (const PenScal *)sc, part);
      hipLaunchKernelGGL(k_pen_check, dim3(1), dim3(TPB), 0, s, sc, (const double *)part, g, (int64_t)PEN_MAXIT);
      precond(0);
    }
  };
  // a batch is one graph, captured once per mesh (its buffers live as long)
  if (use_graph && !c.pen_graph) {
    hipGraph_t gr = nullptr;
    PZG_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    batch();
    PZG_HIP(hipStreamEndCapture(s, &gr));
    hipGraphExec_t ex = nullptr;
    const hipError_t e = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gr);
    if (e != hipSuccess) {
      set_error(std::string("penalty rows: hipGraphInstantiate: ") + hipGetErrorString(e));
      return PZG_EHIP;
    }
    c.pen_graph = ex;
  }
  for (int64_t done_it = 0; done_it < PEN_MAXIT; done_it += PEN_BATCH) {
    if (use_graph) PZG_HIP(hipGraphLaunch(c.pen_graph, s));
    else batch();
    PZG_HIP(hipGetLastError());
    PZG_HIP(hipMemcpyAsync(&hs, sc, sizeof(hs), hipMemcpyDeviceToHost, s));
    PZG_HIP(hipStreamSynchronize(s));
    if (hs.done) break;
  }
  if (hs.done == 3) {
    set_error("penalty rows: K_pp not positive definite");
    return PZG_EUNSUP;
  }
  if (hs.it > 0 && rr_full && c.pen_narow > 0) {
    // the exact residual: r -= K delta on the rows delta reaches, then its norm
    hipLaunchKernelGGL(k_pen_scatter_d, dim3(g), dim3(TPB), 0, s, np, idx, (const double *)x, u64,
                       c.d_pen_delta.as<double>());
    const int64_t th = c.pen_narow * BSR_CHUNK;
    if (c.bs == 3)
      hipLaunchKernelGGL(k_pen_rupd<3>, dim3((unsigned)((th + TPB - 1) / TPB)), dim3(TPB), 0, s, c.pen_narow,
                         c.d_pen_arow.as<int32_t>(), c.d_fbia.as<int64_t>(), c.d_fbja.as<int32_t>(),
                         c.d_fa.as<double>(), (const double *)c.d_pen_delta.as<double>(), r64);
    else
      hipLaunchKernelGGL(k_pen_rupd<1>, dim3((unsigned)((th + TPB - 1) / TPB)), dim3(TPB), 0, s, c.pen_narow,
                         c.d_pen_arow.as<int32_t>(), c.d_fbia.as<int64_t>(), c.d_fbja.as<int32_t>(),
                         c.d_fa.as<double>(), (const double *)c.d_pen_delta.as<double>(), r64);
    const int gn = pen_grid(c.neq);
    hipLaunchKernelGGL(k_pen_norm2, dim3(gn), dim3(TPB), 0, s, c.neq, (const double *)r64, part);
    hipLaunchKernelGGL(k_pen_norm2_fin, dim3(1), dim3(TPB), 0, s, (const double *)part, gn, &sc->rr0);
    PZG_HIP(hipGetLastError());
    double h = 0.;
    PZG_HIP(hipMemcpyAsync(&h, &sc->rr0, sizeof(double), hipMemcpyDeviceToHost, s));
    PZG_HIP(hipStreamSynchronize(s));
    *rr_full = h;
  } else if (hs.it > 0) {
    hipLaunchKernelGGL(k_pen_scatter, dim3(g), dim3(TPB), 0, s, np, idx, (const double *)x, (const double *)r, u64,
                       r64);
    PZG_HIP(hipGetLastError());
  }
  if (rr_full && hs.it == 0) *rr_full = -1.;  // nothing changed: the caller's norm stands
  if (its) *its = hs.it;
  if (rr0) *rr0 = hs.rr0;
  if (rr1) *rr1 = hs.rr;
  return PZG_OK;
}

// r32[penalty rows] = 0 (the inner iterations leave those rows to pen_solve)
int pen_zero32(Ctx &c, float *r32) {
  if (!c.pen_ready || c.pen_n == 0) return PZG_OK;
  hipLaunchKernelGGL(k_pen_zero32, dim3(pen_grid(c.pen_n)), dim3(TPB), 0, c.stream, c.pen_n,
                     c.d_pen_idx.as<int64_t>(), r32);
  PZG_HIP(hipGetLastError());
  return PZG_OK;
}

void pen_free(Ctx &c) {
  for (DevBuf *b : {&c.d_pen_idx, &c.d_pen_pmap, &c.d_pen_ia, &c.d_pen_ja, &c.d_pen_a, &c.d_pen_sptr,
                    &c.d_pen_sidx, &c.d_pen_sset, &c.d_pen_soff, &c.d_pen_sval, &c.d_pen_sinv, &c.d_pen_gptr,
                    &c.d_pen_gslot, &c.d_pen_w, &c.d_pen_scal, &c.d_pen_arow, &c.d_pen_delta})
    dev_free(*b);
  if (c.pen_graph) (void)hipGraphExecDestroy(c.pen_graph);
  c.pen_graph = nullptr;
  c.pen_ready = c.pen_struct = false;
  c.pen_n = c.pen_nnz = c.pen_nset = c.pen_nslot = c.pen_narow = 0;
}

}  // namespace pzg
