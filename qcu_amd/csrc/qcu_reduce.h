// qcu_reduce.h -- deterministic FP64 grid reductions.
//
// Each block reduces its lanes with 64-wide wavefront shuffles and the waves
// of the block through LDS, and writes one partial per block.  A one-block
// finishing kernel (finish() in blas.hip) then sums the partials in a fixed
// order -- so results do not depend on scheduling -- and runs the scalar
// epilogue of the solver (alpha, omega, rho/beta updates) on the device.  The
// kernel boundary orders the two; no device-wide atomics are used (a counter
// shared by thousands of blocks serialises at the memory side of an MI355X).
#pragma once
#include "qcu_common.h"

namespace qcu {

constexpr int kMaxWaves = 16;  // up to 1024 threads per block

QCU_D double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Sums v over the block and stores it as partials[slot * NR + k].  Must be
// called by every thread of the block.
template <int NR>
QCU_D void block_partial(const double (&v)[NR], double *partials, int slot) {
  __shared__ double sh[kMaxWaves][NR];
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
  }
}

// epilogues run by the finishing kernel on the grid totals
enum FinKindAll {
  FK_NONE = 0,
  FK_INIT = 1,      // BiCGSTAB start: r~ = r, rho = |r|^2
  FK_STORE = 3,     // out[k] = tot[k]
  FK_ACCUM = 4,     // out[k] += tot[k]
  FK_ALPHA = 5,     // alpha = rho / <r~, v>
  FK_OMEGA = 6,     // omega = <t, s> / <t, t>
  FK_RHO = 7,       // rho, |r|^2, beta
  FK_RELIABLE = 8,  // after a reliable update
  FK_STORE_TMP = 9, // st->tmp[k] = tot[k]
  FK_CG_INIT = 10,  // CG on the normal equations: rho = |z|^2, beta = 0
  FK_CG_ALPHA = 11, // alpha = rho / |q|^2
  FK_CG_RHO = 12,   // rho' = |z|^2, beta = rho'/rho, rr = tmp[0] (|res|^2), stop test
};

}  // namespace qcu
