// qcu_batch.h -- segment tables: many systems in one launch.
//
// A batch is a list of segments.  A segment is one lattice (geometry, resident
// links and clover, kappa) carrying nrhs right-hand sides that are interleaved
// site-fastest: the spinor element (component k, site i, rhs r) of a segment
// sits at virtual site i*nrhs + r of a field with component stride
// vh*nrhs.  Threads map one to one onto virtual sites, so the nrhs lanes of a
// site read the same link address (one fetch shared by all right-hand sides)
// while their spinor loads stay contiguous.  Every system (segment, rhs) has
// its own solver state.  One system on one lattice is the one-segment,
// one-rhs case of the same kernels.
//
// Batch vectors concatenate the segments: segment s starts at virtual site
// voff[s]; its payload is ncomp 16-byte words per virtual site (12 FP64, 6 FP32,
// 3 FP16) and the FP16 per-site scale array is indexed by virtual site.
#pragma once
#include "qcu_common.h"
#include "qcu_reduce.h"

namespace qcu {

constexpr int kMaxRhs = 16;  // right-hand sides per segment
// virtual sites per segment: the kernels address a field with a 32-bit
// per-lane byte offset of 16 bytes per virtual site
constexpr long long kMaxVirtualSites = 1LL << 28;

struct SegDesc {
  Geom g;
  int nrhs;      // interleaved right-hand sides
  int sys0;      // solver state index of rhs 0
  int nthr;      // vh * nrhs virtual sites
  int pcap;      // block-partial slots per right-hand side (the same for every segment of a batch)
  long long voff;  // virtual-site offset of this segment in batch vectors
  double kappa;
  LinkPtr U[3];          // links per precision
  CloverPtr Cinv[3][2];  // inverse clover per precision and parity (p == null: Wilson)
  CloverPtr A[2];        // FP64 clover per parity (true residual)
};

// the segment and its first local thread for logical block blk of size bsz
struct BlockSeg {
  int seg;
  int t0;
};

QCU_D BlockSeg find_seg(const SegDesc *d, int nseg, int blk, int bsz) {
  int acc = 0;
  for (int s = 0; s < nseg; ++s) {
    const int nb = (d[s].nthr + bsz - 1) / bsz;
    if (blk < acc + nb) {
      BlockSeg r;
      r.seg = s;
      r.t0 = (blk - acc) * bsz;
      return r;
    }
    acc += nb;
  }
  BlockSeg r;
  r.seg = -1;
  r.t0 = 0;
  return r;
}

QCU_HD int words_per_site(int prec) { return prec == P64 ? 12 : (prec == P32 ? 6 : 3); }

// view of a batch vector restricted to one segment
QCU_HD SpinorPtr seg_view(const SpinorPtr &f, int prec, long long voff, int nthr) {
  SpinorPtr v;
  v.p = (char *)f.p + voff * (long long)words_per_site(prec) * 16;
  v.norm = f.norm ? f.norm + voff : nullptr;
  v.stride = nthr;
  return v;
}

// Per-system block partials: partials[(r * pcap + blk) * NR + k], so that the
// partials of one system are contiguous for the finishing kernel.  Threads of
// a block belong to one segment; rhs r of thread tid is (t0 + tid) % nrhs.
// Deterministic: a fixed-order sum per rhs.
template <int NR>
QCU_D void block_partial_rhs(const double (&v)[NR], double *partials, int blk, int t0, int nrhs, int pcap) {
  if (nrhs == 1) {
    block_partial<NR>(v, partials, blk);
    return;
  }
  __shared__ double sh[256][NR];
#pragma unroll
  for (int k = 0; k < NR; ++k) sh[threadIdx.x][k] = v[k];
  __syncthreads();
  const int r = threadIdx.x;
  if (r < nrhs) {
    const int j0 = ((r - t0 % nrhs) % nrhs + nrhs) % nrhs;
    double acc[NR];
#pragma unroll
    for (int k = 0; k < NR; ++k) acc[k] = 0;
    for (int j = j0; j < (int)blockDim.x; j += nrhs)
#pragma unroll
      for (int k = 0; k < NR; ++k) acc[k] += sh[j][k];
#pragma unroll
    for (int k = 0; k < NR; ++k) partials[((size_t)r * pcap + blk) * NR + k] = acc[k];
  }
}

}  // namespace qcu
