// pzg_device.h -- device helpers shared by the solver and learned kernels:
// wavefront/block sums and the deterministic two-level reduction.
#pragma once
#include <hip/hip_runtime.h>

namespace pzg {

constexpr int RED_TPB = 256;

// Workgroups are dispatched round-robin over the XCDs (workgroup b runs on XCD
// b % NXCD), and each XCD has its own L2.  xcd_block_range gives workgroup
// blockIdx.x a contiguous slice [lo, hi) of [0, n) such that the slices of the
// workgroups of one XCD are adjacent: XCD k covers about [k n/NXCD,
// (k+1) n/NXCD), so the gathered operands of neighbouring rows stay in that
// XCD's L2.
constexpr int NXCD = 8;
__device__ inline void xcd_block_range(int64_t n, int64_t &lo, int64_t &hi) {
  const int G = (int)gridDim.x, b = (int)blockIdx.x;
  const int x = b % NXCD, i = b / NXCD;
  const int per = G / NXCD, rem = G % NXCD;
  const int64_t vb = (int64_t)x * per + (x < rem ? x : rem) + i;
  lo = n * vb / G;
  hi = n * (vb + 1) / G;
}

// The slab [lo, hi) of [0, n) owned by this workgroup's XCD, the index i of
// the workgroup among that XCD's gk workgroups.  Interleaving the XCD's
// workgroups over its slab (tile i, i + gk, ...) keeps the rows in flight on
// one XCD in a narrow band, so the gathered operands fit in its L2.
#ifndef PZG_XCD_INTERLEAVE
#define PZG_XCD_INTERLEAVE 1
#endif
__device__ inline void xcd_slab(int64_t n, int64_t &lo, int64_t &hi, int &i, int &gk) {
  const int G = (int)gridDim.x, b = (int)blockIdx.x;
  const int x = b % NXCD;
  const int per = G / NXCD, rem = G % NXCD;
  i = b / NXCD;
  gk = per + (x < rem ? 1 : 0);
  const int64_t first = (int64_t)x * per + (x < rem ? x : rem);
  lo = n * first / G;
  hi = n * (first + gk) / G;
}
constexpr int MAXBLK = 2048;  // partial-sum slots per reduction
constexpr int NSLOT = 4;

__device__ inline double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Sum of v over the block; result valid in thread 0.
__device__ inline double block_sum(double v, double *sm) {
  v = wave_sum(v);
  int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) sm[w] = v;
  __syncthreads();
  double s = 0.;
  if (threadIdx.x == 0)
    for (int k = 0; k < (int)(blockDim.x >> 6); k++) s += sm[k];
  return s;
}

// Stores this block's partials; returns true in the (single) last block once
// every block has stored, after which 'tot' holds the fixed-order totals.
// Launches must use at most MAXBLK blocks of RED_TPB threads.
template <int NV>
__device__ inline bool reduce_last(const double (&v)[NV], double *partials, unsigned *counter,
                                   double (&tot)[NV]) {
  __shared__ double sm[RED_TPB / 64];
  __shared__ bool last;
  for (int k = 0; k < NV; k++) {
    double s = block_sum(v[k], sm);
    if (threadIdx.x == 0) partials[k * MAXBLK + blockIdx.x] = s;
  }
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned prev = atomicAdd(counter, 1u);
    last = (prev == gridDim.x - 1);
  }
  __syncthreads();
  if (!last) return false;
  __threadfence();
  for (int k = 0; k < NV; k++) {
    double s = 0.;
    for (int b = threadIdx.x; b < (int)gridDim.x; b += blockDim.x)
      s += __builtin_nontemporal_load(&partials[k * MAXBLK + b]);
    s = block_sum(s, sm);
    if (threadIdx.x == 0) tot[k] = s;
  }
  if (threadIdx.x == 0) *counter = 0u;
  return true;
}

// Stores this block's sums of v into partials[k*MAXBLK + blockIdx.x] (plain
// stores: a following kernel reads them, ordered by the kernel boundary).
template <int NV>
__device__ inline void block_partials(const double (&v)[NV], double *partials) {
  __shared__ double sm[RED_TPB / 64];
  for (int k = 0; k < NV; k++) {
    double s = block_sum(v[k], sm);
    if (threadIdx.x == 0) partials[k * MAXBLK + blockIdx.x] = s;
  }
}

// Fixed-order sum of nblk partials of one slot, run by a single block;
// result valid in thread 0.
__device__ inline double sum_partials(const double *partials, int slot, int nblk) {
  __shared__ double sm[RED_TPB / 64];
  double s = 0.;
  for (int b = threadIdx.x; b < nblk; b += blockDim.x) s += partials[slot * MAXBLK + b];
  return block_sum(s, sm);
}

}  // namespace pzg
