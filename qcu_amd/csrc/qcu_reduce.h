// qcu_reduce.h -- deterministic FP64 block reductions with a fused epilogue.
//
// Each block reduces its lanes with 64-wide wavefront shuffles, the waves of a
// block through LDS, and writes one partial per block.  The block that takes
// the last ticket of an agent-scope counter sums all partials in a fixed order
// (so the result does not depend on scheduling) and runs the epilogue (scalar
// updates of the solver) in one thread, then re-arms the counter.  No second
// kernel launch and no host round trip are needed for alpha, omega, beta.
#pragma once
#include "qcu_common.h"

namespace qcu {

constexpr int kMaxWaves = 16;  // up to 1024 threads per block

QCU_D double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

QCU_D double load_agent(const double *p) {
  unsigned long long u =
      __hip_atomic_load((const unsigned long long *)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return __longlong_as_double((long long)u);
}

// Reduces NR values over the grid.  Returns true in thread 0 of the last block,
// with the grid totals in tot.  Must be called by every thread of the block.
// nblocks is the number of blocks taking part (gridDim.x for 1-D grids) and
// slot the index of this block's partial.
template <int NR>
QCU_D bool grid_reduce(double (&v)[NR], double *partials, unsigned *counter, int slot, int nblocks,
                       double (&tot)[NR]) {
  __shared__ double sh[kMaxWaves][NR];
  __shared__ int last;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nw = (blockDim.x + 63) >> 6;
#pragma unroll
  for (int k = 0; k < NR; ++k) {
    double w = wave_sum(v[k]);
    if (lane == 0) sh[wave][k] = w;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NR; ++k) {
      double s = 0;
      for (int w = 0; w < nw; ++w) s += sh[w][k];
      partials[(size_t)slot * NR + k] = s;
    }
    __threadfence();
    unsigned t = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    last = (t == (unsigned)(nblocks - 1));
  }
  __syncthreads();
  if (!last) return false;
  __threadfence();
  // fixed-order sum of the partials by the last block
  double acc[NR];
#pragma unroll
  for (int k = 0; k < NR; ++k) acc[k] = 0;
  for (int b = threadIdx.x; b < nblocks; b += blockDim.x) {
#pragma unroll
    for (int k = 0; k < NR; ++k) acc[k] += load_agent(partials + (size_t)b * NR + k);
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NR; ++k) {
    double w = wave_sum(acc[k]);
    if (lane == 0) sh[wave][k] = w;
  }
  __syncthreads();
  if (threadIdx.x != 0) return false;
#pragma unroll
  for (int k = 0; k < NR; ++k) {
    double s = 0;
    for (int w = 0; w < nw; ++w) s += sh[w][k];
    tot[k] = s;
  }
  __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return true;
}

}  // namespace qcu
