// dslash.hip -- even-odd Wilson / clover hopping term for gfx950.
//
// out_p(x) = sum_mu (1 - g_mu) U_mu(x) psi(x+mu) + (1 + g_mu) U_mu(x-mu)^dag psi(x-mu)
// on the sites of output parity p, psi on parity 1-p: the operator of
// src/wilson_dslash.cu:9-245 with the spin projectors of the DeGrand-Rossi
// basis written out as half-spinor (two 3-vectors) products, the 12-real link
// reconstruction of include/define.h:198-206, periodic wrap in every direction
// (the single-GPU case of include/lattice_dslash.h:11-197) and an optional -1
// phase on the hops that cross the time boundary.
//
// One thread per output site, site-fastest SoA loads (coalesced 16-byte lanes),
// eight directions fused, optional epilogues: packed clover inverse (two
// Hermitian 6x6 blocks, src/clover_dslash.cu:857-882 restricted to its
// non-zero chiral blocks), x - k2 * (...) for the even-odd Schur operator
// (src/bistabcg.cu:85-96), and fused FP64 grid reductions feeding the BiCGSTAB
// scalar epilogues.  The stencil is HBM bound (about 0.5 flop per byte), so
// no matrix cores are involved.
#include "qcu_batch.h"
#include "qcu_internal.h"
#include "qcu_reduce.h"
#include "qcu_state.h"

namespace qcu {

template <class T>
struct Half {
  cplx<T> h0[3], h1[3];
};

// spin projection of psi(neighbour) for direction MU, FWD = 1: (1 - g_mu), 0: (1 + g_mu)
template <int MU, int FWD, class T>
QCU_D void project(const cplx<T> (&s)[12], cplx<T> (&h0)[3], cplx<T> (&h1)[3]) {
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const cplx<T> s0 = s[c], s1 = s[3 + c], s2 = s[6 + c], s3 = s[9 + c];
    if (MU == 0) {
      h0[c] = FWD ? s0 + mul_mi(s3) : s0 + mul_i(s3);
      h1[c] = FWD ? s1 + mul_mi(s2) : s1 + mul_i(s2);
    } else if (MU == 1) {
      h0[c] = FWD ? s0 + s3 : s0 - s3;
      h1[c] = FWD ? s1 - s2 : s1 + s2;
    } else if (MU == 2) {
      h0[c] = FWD ? s0 + mul_mi(s2) : s0 + mul_i(s2);
      h1[c] = FWD ? s1 + mul_i(s3) : s1 + mul_mi(s3);
    } else {
      h0[c] = FWD ? s0 - s2 : s0 + s2;
      h1[c] = FWD ? s1 - s3 : s1 + s3;
    }
  }
}

// add the reconstructed 4-spinor of the transported half spinor (a0, a1)
template <int MU, int FWD, class T>
QCU_D void accumulate(cplx<T> (&d)[12], const cplx<T> (&a0)[3], const cplx<T> (&a1)[3]) {
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    d[c] = d[c] + a0[c];
    d[3 + c] = d[3 + c] + a1[c];
    if (MU == 0) {
      d[6 + c] = FWD ? d[6 + c] + mul_i(a1[c]) : d[6 + c] + mul_mi(a1[c]);
      d[9 + c] = FWD ? d[9 + c] + mul_i(a0[c]) : d[9 + c] + mul_mi(a0[c]);
    } else if (MU == 1) {
      d[6 + c] = FWD ? d[6 + c] - a1[c] : d[6 + c] + a1[c];
      d[9 + c] = FWD ? d[9 + c] + a0[c] : d[9 + c] - a0[c];
    } else if (MU == 2) {
      d[6 + c] = FWD ? d[6 + c] + mul_i(a0[c]) : d[6 + c] + mul_mi(a0[c]);
      d[9 + c] = FWD ? d[9 + c] + mul_mi(a1[c]) : d[9 + c] + mul_i(a1[c]);
    } else {
      d[6 + c] = FWD ? d[6 + c] - a0[c] : d[6 + c] + a0[c];
      d[9 + c] = FWD ? d[9 + c] - a1[c] : d[9 + c] + a1[c];
    }
  }
}

// a = U h (FWD) or U^dag h
template <int FWD, class T>
QCU_D void transport(const cplx<T> (&u)[9], const cplx<T> (&h)[3], cplx<T> (&a)[3]) {
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    cplx<T> acc = mk<T>(0, 0);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      if (FWD)
        cmad(acc, u[r * 3 + c], h[c]);
      else
        cmadj(acc, u[c * 3 + r], h[c]);
    }
    a[r] = acc;
  }
}

template <int P, int MU, int FWD>
QCU_D void hop_dir(cplx<typename real_of<P>::type> (&d)[12], const LinkPtr &U, int pm, int li,
                   const SpinorPtr &in, int si, typename real_of<P>::type ph) {
  typedef typename real_of<P>::type T;
  cplx<T> s[12];
  SpinorIO<P>::load(in, si, s);
  cplx<T> h0[3], h1[3];
  project<MU, FWD>(s, h0, h1);
  cplx<T> u[9];
  LinkIO<P>::load(U, pm, li, u);
  cplx<T> a0[3], a1[3];
  transport<FWD>(u, h0, a0);
  transport<FWD>(u, h1, a1);
  if (MU == 3) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      a0[c] = ph * a0[c];
      a1[c] = ph * a1[c];
    }
  }
  accumulate<MU, FWD>(d, a0, a1);
}

struct HopKArgs {
  SpinorPtr out, in, x, y;  // batch vectors
  const SegDesc *segs;
  int nseg;
  int parity;
  int remap;
  SysState *st;  // per-system states, or null
  double *partials;
};

// G: scheduling groups.  0 lets the compiler hoist every load of the eight
// directions (most memory-level parallelism, most registers); 1 and 2 fence the
// instruction schedule after every direction / every pair of directions, which
// caps the live registers and raises occupancy (the FP64 kernel needs this).
#define QCU_GROUP_FENCE(G, k)                                                      \
  do {                                                                             \
    if ((G) == 1 || ((G) == 2 && ((k) & 1))) __builtin_amdgcn_sched_barrier(0);    \
  } while (0)

template <int P, int CLOVER, int EPI, int RED, int G>
__global__ void __launch_bounds__(256) k_hop(HopKArgs a) {
  typedef typename real_of<P>::type T;
  const int blk = remap_block(blockIdx.x, gridDim.x, a.remap);
  const BlockSeg bs = find_seg(a.segs, a.nseg, blk, blockDim.x);
  if (bs.seg < 0) return;  // uniform: a block past the last segment
  const SegDesc &sd = a.segs[bs.seg];
  const Geom g = sd.g;
  const int nrhs = sd.nrhs;
  const int t = bs.t0 + threadIdx.x;
  const int site = t / nrhs;
  const int rhs = t - site * nrhs;
  bool active = t < sd.nthr;
  // a solver state, when given, stops the work of its system once it has stopped
  if (a.st && active && a.st[sd.sys0 + rhs].done) active = false;
  cplx<T> d[12];
  double red[3] = {0, 0, 0};
  if (active) {
    const SpinorPtr in = seg_view(a.in, P, sd.voff, sd.nthr);
    const LinkPtr U = sd.U[P];
    const int i = site;
    const Site s = decode(g, i, a.parity);
    const int p = a.parity, q = 1 - p;
    const int sy = g.lxh, sz = g.lxh * g.ly, st_ = sz * g.lz;
#pragma unroll
    for (int k = 0; k < 12; ++k) d[k] = mk<T>(0, 0);
    {
      // branch-free index arithmetic keeps the body one basic block, so the
      // scheduling fences below bound the live registers; a site with odd
      // full x has its x-1 neighbour at the same half index
      const int xe = 1 - s.xodd;
      const int jb = i - xe + xe * (int)(s.xh == 0) * g.lxh;
      const int jf = i + s.xodd - s.xodd * (int)(s.xh == g.lxh - 1) * g.lxh;
      hop_dir<P, 0, 0>(d, U, q * 4 + 0, jb, in, jb * nrhs + rhs, (T)1);
      QCU_GROUP_FENCE(G, 0);
      hop_dir<P, 0, 1>(d, U, p * 4 + 0, i, in, jf * nrhs + rhs, (T)1);
      QCU_GROUP_FENCE(G, 1);
    }
    {
      const int jb = i - sy + (int)(s.y == 0) * g.ly * sy;
      const int jf = i + sy - (int)(s.y == g.ly - 1) * g.ly * sy;
      hop_dir<P, 1, 0>(d, U, q * 4 + 1, jb, in, jb * nrhs + rhs, (T)1);
      QCU_GROUP_FENCE(G, 2);
      hop_dir<P, 1, 1>(d, U, p * 4 + 1, i, in, jf * nrhs + rhs, (T)1);
      QCU_GROUP_FENCE(G, 3);
    }
    {
      const int jb = i - sz + (int)(s.z == 0) * g.lz * sz;
      const int jf = i + sz - (int)(s.z == g.lz - 1) * g.lz * sz;
      hop_dir<P, 2, 0>(d, U, q * 4 + 2, jb, in, jb * nrhs + rhs, (T)1);
      QCU_GROUP_FENCE(G, 4);
      hop_dir<P, 2, 1>(d, U, p * 4 + 2, i, in, jf * nrhs + rhs, (T)1);
      QCU_GROUP_FENCE(G, 5);
    }
    {
      const int jb = i - st_ + (int)(s.t == 0) * g.lt * st_;
      const int jf = i + st_ - (int)(s.t == g.lt - 1) * g.lt * st_;
      const int flip = (g.tbc < 0);
      const T phb = (T)(1 - 2 * (flip & (int)(s.t == 0)));
      const T phf = (T)(1 - 2 * (flip & (int)(s.t == g.lt - 1)));
      hop_dir<P, 3, 0>(d, U, q * 4 + 3, jb, in, jb * nrhs + rhs, phb);
      QCU_GROUP_FENCE(G, 6);
      hop_dir<P, 3, 1>(d, U, p * 4 + 3, i, in, jf * nrhs + rhs, phf);
      QCU_GROUP_FENCE(G, 7);
    }
    if (CLOVER) {
      const CloverPtr C = sd.Cinv[P][p];
      if (P == P16) {
        T c[72];
        CloverIO<P>::load(C, i, c);
        clover_apply_site(c, d);
      } else {
        // one chiral block at a time bounds the live registers
        T c[36];
        clover_load_block(C, i, 0, c);
        clover_apply_block(c, d, 0);
        QCU_GROUP_FENCE(G, 1);
        clover_load_block(C, i, 1, c);
        clover_apply_block(c, d, 1);
      }
    }
    if (EPI == EPI_XPAY) {
      cplx<T> x[12];
      SpinorIO<P>::load(seg_view(a.x, P, sd.voff, sd.nthr), t, x);
      const T k2 = (T)(sd.kappa * sd.kappa);
#pragma unroll
      for (int k = 0; k < 12; ++k) d[k] = x[k] - k2 * d[k];
      if (RED == RED_OUT_X) {
        // <out, x> and |out|^2
#pragma unroll
        for (int k = 0; k < 12; ++k) {
          red[0] += (double)d[k].re * (double)x[k].re + (double)d[k].im * (double)x[k].im;
          red[1] += (double)d[k].re * (double)x[k].im - (double)d[k].im * (double)x[k].re;
          red[2] += (double)d[k].re * (double)d[k].re + (double)d[k].im * (double)d[k].im;
        }
      }
    }
    if (RED == RED_DOT_Y_OUT) {
      cplx<T> y[12];
      SpinorIO<P>::load(seg_view(a.y, P, sd.voff, sd.nthr), t, y);
#pragma unroll
      for (int k = 0; k < 12; ++k) {
        red[0] += (double)y[k].re * (double)d[k].re + (double)y[k].im * (double)d[k].im;
        red[1] += (double)y[k].re * (double)d[k].im - (double)y[k].im * (double)d[k].re;
      }
    }
    SpinorIO<P>::store(seg_view(a.out, P, sd.voff, sd.nthr), t, d);
  }
  if (RED == RED_DOT_Y_OUT) {
    double v[2] = {red[0], red[1]};
    block_partial_rhs<2>(v, a.partials, blk, bs.t0, nrhs, sd.pcap);
  } else if (RED == RED_OUT_X) {
    block_partial_rhs<3>(red, a.partials, blk, bs.t0, nrhs, sd.pcap);
  }
}

static int total_blocks(const std::vector<SegDesc> &segs, int bsz) {
  int nb = 0;
  for (const SegDesc &d : segs) nb += (d.nthr + bsz - 1) / bsz;
  return nb;
}

template <int P, int CLOVER, int EPI, int RED>
static int launch(const HopKArgs &a, int nb, const LaunchCfg &cfg, hipStream_t s) {
  switch (cfg.group) {
    case 1: hipLaunchKernelGGL((k_hop<P, CLOVER, EPI, RED, 1>), dim3(nb), dim3(cfg.block), 0, s, a); break;
    case 2: hipLaunchKernelGGL((k_hop<P, CLOVER, EPI, RED, 2>), dim3(nb), dim3(cfg.block), 0, s, a); break;
    default: hipLaunchKernelGGL((k_hop<P, CLOVER, EPI, RED, 0>), dim3(nb), dim3(cfg.block), 0, s, a); break;
  }
  QCU_HIP(hipGetLastError());
  return 0;
}

template <int P>
static int dispatch(const HopKArgs &a, int nb, int clover, int epi, int red, const LaunchCfg &cfg, hipStream_t s) {
  if (epi == EPI_PLAIN) {
    return clover ? launch<P, 1, EPI_PLAIN, RED_NONE>(a, nb, cfg, s)
                  : launch<P, 0, EPI_PLAIN, RED_NONE>(a, nb, cfg, s);
  }
  if (red == RED_NONE)
    return clover ? launch<P, 1, EPI_XPAY, RED_NONE>(a, nb, cfg, s)
                  : launch<P, 0, EPI_XPAY, RED_NONE>(a, nb, cfg, s);
  if (red == RED_DOT_Y_OUT)
    return clover ? launch<P, 1, EPI_XPAY, RED_DOT_Y_OUT>(a, nb, cfg, s)
                  : launch<P, 0, EPI_XPAY, RED_DOT_Y_OUT>(a, nb, cfg, s);
  return clover ? launch<P, 1, EPI_XPAY, RED_OUT_X>(a, nb, cfg, s)
                : launch<P, 0, EPI_XPAY, RED_OUT_X>(a, nb, cfg, s);
}

int hop(const BatchOp &op, int prec, SpinorPtr out, SpinorPtr in, int parity, int epi, SpinorPtr x, int red,
        SpinorPtr y, SysState *st, double *partials, const LaunchCfg &cfg, hipStream_t s, bool use_clover) {
  QCU_CHECK(parity == 0 || parity == 1, QCU_ERR_ARG, "parity must be 0 or 1");
  QCU_CHECK(prec >= P64 && prec <= P16, QCU_ERR_ARG, "unknown precision");
  QCU_CHECK(cfg.block >= 64 && cfg.block <= 256 && cfg.block % 64 == 0, QCU_ERR_ARG,
            "block size must be a multiple of 64 in [64, 256]");
  QCU_CHECK(red == RED_NONE || (epi == EPI_XPAY && st && partials), QCU_ERR_ARG,
            "fused reductions need the xpay epilogue and solver state");
  HopKArgs a;
  a.out = out;
  a.in = in;
  a.x = x;
  a.y = y;
  a.segs = op.dev_segs;
  a.nseg = (int)op.segs.size();
  a.parity = parity;
  a.remap = cfg.remap;
  a.st = st;
  a.partials = partials;
  const int clover = use_clover && op.has_clover;
  const int nb = total_blocks(op.segs, cfg.block);
  int rc;
  switch (prec) {
    case P64: rc = dispatch<P64>(a, nb, clover, epi, red, cfg, s); break;
    case P32: rc = dispatch<P32>(a, nb, clover, epi, red, cfg, s); break;
    default: rc = dispatch<P16>(a, nb, clover, epi, red, cfg, s); break;
  }
  if (rc || red == RED_NONE) return rc;
  return red == RED_DOT_Y_OUT ? finish(op, partials, cfg.block, 2, FK_ALPHA, st, nullptr, s)
                              : finish(op, partials, cfg.block, 3, FK_OMEGA, st, nullptr, s);
}

}  // namespace qcu
