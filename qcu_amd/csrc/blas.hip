// blas.hip -- fused BLAS-1 site kernels of the BiCGSTAB recurrence over a
// batch of systems (segment tables, qcu_batch.h).
//
// Each kernel touches every vector once; reductions are per system, in FP64,
// in two deterministic stages (block partials, then the one-block-per-system
// finishing kernel k_finish), and the scalar epilogues of qcu_state.h keep
// rho, alpha, beta, omega on the device (the reference reads them back
// through cuBLAS and 1-thread kernels, src/bistabcg.cu:4-180,
// include/lattice_bistabcg.h:149-163).  Vectors may be FP64, FP32 or FP16 with
// a per-site scale; arithmetic is FP32 for the sloppy precisions and FP64
// otherwise.  A system whose state says it has stopped is skipped.
//
// Every site kernel runs in 256-thread blocks and says so (__launch_bounds__):
// without it the compiler assumes 1024-thread blocks, caps a thread at 128
// VGPRs, and the FP64 kernels that hold five 12-component spinors spill.
#include "qcu_batch.h"
#include "qcu_internal.h"
#include "qcu_reduce.h"
#include "qcu_solver.h"
#include "qcu_state.h"

namespace qcu {

template <class T>
QCU_D cplx<T> cs(const double *c) {
  return mk<T>((T)c[0], (T)c[1]);
}

template <int P>
using RT = typename real_of<P>::type;

template <int Pd, int Ps>
QCU_D void cvt(const cplx<RT<Ps>> (&a)[12], cplx<RT<Pd>> (&b)[12]) {
#pragma unroll
  for (int k = 0; k < 12; ++k) b[k] = mk<RT<Pd>>((RT<Pd>)a[k].re, (RT<Pd>)a[k].im);
}

template <int P>
QCU_D void zero12(cplx<RT<P>> (&v)[12]) {
#pragma unroll
  for (int k = 0; k < 12; ++k) v[k] = mk<RT<P>>(0, 0);
}

// per-thread view of a batch: segment, local virtual site, system
struct TCtx {
  const SegDesc *sd;
  int t, t0, sys, nrhs;
  bool valid;   // inside a segment
  bool active;  // valid and its system still running (or no state given)
};

QCU_D TCtx tctx(const SegDesc *segs, int nseg, const SysState *st) {
  TCtx c;
  const BlockSeg bs = find_seg(segs, nseg, blockIdx.x, blockDim.x);
  c.valid = false;
  c.active = false;
  c.sd = nullptr;
  c.t0 = bs.t0;
  c.nrhs = 1;
  c.t = 0;
  c.sys = 0;
  if (bs.seg < 0) return c;
  c.sd = segs + bs.seg;
  c.nrhs = c.sd->nrhs;
  c.t = bs.t0 + threadIdx.x;
  c.sys = c.sd->sys0 + c.t % c.nrhs;
  c.valid = c.t < c.sd->nthr;
  c.active = c.valid && (!st || st[c.sys].done == 0);
  return c;
}

template <int P>
QCU_D SpinorPtr view(const SpinorPtr &f, const TCtx &c) {
  return seg_view(f, P, c.sd->voff, c.sd->nthr);
}

// ---------------------------------------------------------------- epilogues
// start of a solve: r~ = r = source, rho = |r|^2, alpha = omega = rho_prev = 1
QCU_D void fin_init(SysState *st, double rr) {
  if (st->done == 4) return;  // a system left out of this solve
  st->rho_prev[0] = 1;
  st->rho_prev[1] = 0;
  st->alpha[0] = 1;
  st->alpha[1] = 0;
  st->omega[0] = 1;
  st->omega[1] = 0;
  st->beta[0] = 0;
  st->beta[1] = 0;
  st->iter = 0;
  st->rho[0] = rr;
  st->rho[1] = 0;
  st->rr = rr;
  st->rr_ref = rr;
  st->done = (rr < st->rr_stop) ? 1 : (isfinite(rr) ? 0 : 2);
}

// after a reliable update replaced r by the recomputed residual:
// rho = <r~, r>, beta = (rho / rho_prev) (alpha / omega), rr_ref = |r|^2
QCU_D void fin_reliable(SysState *st, double rre, double rim, double rr) {
  st->rho[0] = rre;
  st->rho[1] = rim;
  st->rr = rr;
  st->rr_ref = rr;
  if (!isfinite(rr)) {
    st->done = 2;
    return;
  }
  double q1[2], q2[2];
  cdiv(st->rho, st->rho_prev, q1);
  cdiv(st->alpha, st->omega, q2);
  cmul(q1, q2, st->beta);
  if (!isfinite(st->beta[0]) || !isfinite(st->beta[1])) {
    st->beta[0] = 0;  // restart the search direction
    st->beta[1] = 0;
  }
  st->done = (rr < st->rr_stop) ? 1 : 0;
}

// ---------------------------------------------------------------- kernels
// r = b (converted), r~ = r, p = v = 0, x = 0; |r|^2 per system
template <int Pb, int P, int Px>
__global__ void __launch_bounds__(256) k_bicg_init(SpinorPtr r, SpinorPtr rt, SpinorPtr p, SpinorPtr v, SpinorPtr x, SpinorPtr b,
                            const SegDesc *segs, int nseg, double *partials) {
  const TCtx c = tctx(segs, nseg, nullptr);
  if (!c.sd) return;
  double acc[1] = {0};
  if (c.valid) {
    cplx<RT<Pb>> bb[12];
    SpinorIO<Pb>::load(view<Pb>(b, c), c.t, bb);
    cplx<RT<P>> rr[12], z[12];
    cvt<P, Pb>(bb, rr);
    zero12<P>(z);
#pragma unroll
    for (int k = 0; k < 12; ++k) acc[0] += (double)bb[k].re * bb[k].re + (double)bb[k].im * bb[k].im;
    SpinorIO<P>::store(view<P>(r, c), c.t, rr);
    SpinorIO<P>::store(view<P>(rt, c), c.t, rr);
    SpinorIO<P>::store(view<P>(p, c), c.t, z);
    SpinorIO<P>::store(view<P>(v, c), c.t, z);
    cplx<RT<Px>> zx[12];
    zero12<Px>(zx);
    SpinorIO<Px>::store(view<Px>(x, c), c.t, zx);
  }
  block_partial_rhs<1>(acc, partials, blockIdx.x, c.t0, c.nrhs, c.sd->pcap);
}

// p = r + beta (p - omega v)
template <int P>
__global__ void __launch_bounds__(256) k_p_update(SpinorPtr p, SpinorPtr r, SpinorPtr v, const SegDesc *segs, int nseg, const SysState *st) {
  const TCtx c = tctx(segs, nseg, st);
  if (!c.active) return;
  typedef RT<P> T;
  const cplx<T> beta = cs<T>(st[c.sys].beta), omega = cs<T>(st[c.sys].omega);
  cplx<T> pp[12], rr[12], vv[12];
  SpinorIO<P>::load(view<P>(p, c), c.t, pp);
  SpinorIO<P>::load(view<P>(r, c), c.t, rr);
  SpinorIO<P>::load(view<P>(v, c), c.t, vv);
#pragma unroll
  for (int k = 0; k < 12; ++k) pp[k] = rr[k] + beta * (pp[k] - omega * vv[k]);
  SpinorIO<P>::store(view<P>(p, c), c.t, pp);
}

// s = r - alpha v
template <int P>
__global__ void __launch_bounds__(256) k_s_update(SpinorPtr s, SpinorPtr r, SpinorPtr v, const SegDesc *segs, int nseg, const SysState *st) {
  const TCtx c = tctx(segs, nseg, st);
  if (!c.active) return;
  typedef RT<P> T;
  const cplx<T> alpha = cs<T>(st[c.sys].alpha);
  cplx<T> rr[12], vv[12];
  SpinorIO<P>::load(view<P>(r, c), c.t, rr);
  SpinorIO<P>::load(view<P>(v, c), c.t, vv);
#pragma unroll
  for (int k = 0; k < 12; ++k) rr[k] = rr[k] - alpha * vv[k];
  SpinorIO<P>::store(view<P>(s, c), c.t, rr);
}

// x += alpha P + omega S ; r = s - omega t ; <r~, r>, |r|^2 per system
// (P, S) = (p, s) for the plain method, (K p, K s) with a right preconditioner K
template <int Px, int P, int PC>
__global__ void __launch_bounds__(256) k_xr_update(SpinorPtr x, SpinorPtr r, SpinorPtr rt, SpinorPtr p, SpinorPtr sx, SpinorPtr s,
                            SpinorPtr t, const SegDesc *segs, int nseg, SysState *st, double *partials) {
  const TCtx c = tctx(segs, nseg, st);
  if (!c.sd) return;
  double acc[3] = {0, 0, 0};
  if (c.active) {
    typedef RT<P> T;
    typedef RT<Px> TX;
    cplx<T> pp[12], ss[12], tt[12];
    SpinorIO<P>::load(view<P>(p, c), c.t, pp);
    SpinorIO<P>::load(view<P>(s, c), c.t, ss);
    SpinorIO<P>::load(view<P>(t, c), c.t, tt);
    cplx<TX> xx[12];
    SpinorIO<Px>::load(view<Px>(x, c), c.t, xx);
    const cplx<TX> al = cs<TX>(st[c.sys].alpha), omx = cs<TX>(st[c.sys].omega);
    if (PC) {
      cplx<T> sh[12];
      SpinorIO<P>::load(view<P>(sx, c), c.t, sh);
#pragma unroll
      for (int k = 0; k < 12; ++k) {
        cplx<TX> pk = mk<TX>((TX)pp[k].re, (TX)pp[k].im), sk = mk<TX>((TX)sh[k].re, (TX)sh[k].im);
        xx[k] = xx[k] + al * pk + omx * sk;
      }
    } else {
#pragma unroll
      for (int k = 0; k < 12; ++k) {
        cplx<TX> pk = mk<TX>((TX)pp[k].re, (TX)pp[k].im), sk = mk<TX>((TX)ss[k].re, (TX)ss[k].im);
        xx[k] = xx[k] + al * pk + omx * sk;
      }
    }
    SpinorIO<Px>::store(view<Px>(x, c), c.t, xx);
    const cplx<T> om = cs<T>(st[c.sys].omega);
    cplx<T> rtt[12];
    SpinorIO<P>::load(view<P>(rt, c), c.t, rtt);
#pragma unroll
    for (int k = 0; k < 12; ++k) {
      ss[k] = ss[k] - om * tt[k];
      acc[0] += (double)rtt[k].re * ss[k].re + (double)rtt[k].im * ss[k].im;
      acc[1] += (double)rtt[k].re * ss[k].im - (double)rtt[k].im * ss[k].re;
      acc[2] += (double)ss[k].re * ss[k].re + (double)ss[k].im * ss[k].im;
    }
    SpinorIO<P>::store(view<P>(r, c), c.t, ss);
  }
  block_partial_rhs<3>(acc, partials, blockIdx.x, c.t0, c.nrhs, c.sd->pcap);
}

// two-vector reductions: <a, b> with |a|^2 (omega of the preconditioned method)
// or with |b|^2 (rho after a reliable update)
template <int P, int NORM_OF_A>
__global__ void __launch_bounds__(256) k_dot2(SpinorPtr a_, SpinorPtr b_, const SegDesc *segs, int nseg, const SysState *st,
                       double *partials) {
  const TCtx c = tctx(segs, nseg, st);
  if (!c.sd) return;
  double acc[3] = {0, 0, 0};
  if (c.active) {
    cplx<RT<P>> a[12], b[12];
    SpinorIO<P>::load(view<P>(a_, c), c.t, a);
    SpinorIO<P>::load(view<P>(b_, c), c.t, b);
#pragma unroll
    for (int k = 0; k < 12; ++k) {
      acc[0] += (double)a[k].re * b[k].re + (double)a[k].im * b[k].im;
      acc[1] += (double)a[k].re * b[k].im - (double)a[k].im * b[k].re;
      acc[2] += NORM_OF_A ? (double)a[k].re * a[k].re + (double)a[k].im * a[k].im
                          : (double)b[k].re * b[k].re + (double)b[k].im * b[k].im;
    }
  }
  block_partial_rhs<3>(acc, partials, blockIdx.x, c.t0, c.nrhs, c.sd->pcap);
}

// y = a x + b y (all systems)
template <int Px, int Py>
__global__ void __launch_bounds__(256) k_axpby(double a, SpinorPtr x, double b, SpinorPtr y, const SegDesc *segs, int nseg) {
  const TCtx c = tctx(segs, nseg, nullptr);
  if (!c.valid) return;
  typedef RT<Py> T;
  cplx<RT<Px>> xx[12];
  cplx<T> yy[12];
  if (a != 0.0) SpinorIO<Px>::load(view<Px>(x, c), c.t, xx);
  else zero12<Px>(xx);
  if (b != 0.0) SpinorIO<Py>::load(view<Py>(y, c), c.t, yy);
  else zero12<Py>(yy);
#pragma unroll
  for (int k = 0; k < 12; ++k) yy[k] = mk<T>((T)(a * xx[k].re + b * yy[k].re), (T)(a * xx[k].im + b * yy[k].im));
  SpinorIO<Py>::store(view<Py>(y, c), c.t, yy);
}

// r = b - r (FP64), |r|^2 per system
__global__ void __launch_bounds__(256) k_xmy_norm(SpinorPtr b, SpinorPtr r, const SegDesc *segs, int nseg, double *partials) {
  const TCtx c = tctx(segs, nseg, nullptr);
  if (!c.sd) return;
  double acc[1] = {0};
  if (c.valid) {
    cplx<double> bb[12], rr[12];
    SpinorIO<P64>::load(view<P64>(b, c), c.t, bb);
    SpinorIO<P64>::load(view<P64>(r, c), c.t, rr);
#pragma unroll
    for (int k = 0; k < 12; ++k) {
      rr[k] = bb[k] - rr[k];
      acc[0] += norm2(rr[k]);
    }
    SpinorIO<P64>::store(view<P64>(r, c), c.t, rr);
  }
  block_partial_rhs<1>(acc, partials, blockIdx.x, c.t0, c.nrhs, c.sd->pcap);
}

template <int P>
__global__ void __launch_bounds__(256) k_norm2(SpinorPtr x, const SegDesc *segs, int nseg, double *partials) {
  const TCtx c = tctx(segs, nseg, nullptr);
  if (!c.sd) return;
  double acc[1] = {0};
  if (c.valid) {
    cplx<RT<P>> xx[12];
    SpinorIO<P>::load(view<P>(x, c), c.t, xx);
#pragma unroll
    for (int k = 0; k < 12; ++k) acc[0] += (double)xx[k].re * xx[k].re + (double)xx[k].im * xx[k].im;
  }
  block_partial_rhs<1>(acc, partials, blockIdx.x, c.t0, c.nrhs, c.sd->pcap);
}

// out = C (x + a y) in FP64; a = kappa of the segment when a_is_kappa;
// which: 0 identity, 1 inverse clover of the parity, 2 clover of the parity
__global__ void __launch_bounds__(256) k_cl_axpy(SpinorPtr out, SpinorPtr x, double a, int a_is_kappa, SpinorPtr y, int which, int parity,
                          const SegDesc *segs, int nseg) {
  const TCtx c = tctx(segs, nseg, nullptr);
  if (!c.valid) return;
  cplx<double> xx[12];
  SpinorIO<P64>::load(view<P64>(x, c), c.t, xx);
  const double aa = a_is_kappa ? c.sd->kappa : a;
  if (aa != 0.0) {
    cplx<double> yy[12];
    SpinorIO<P64>::load(view<P64>(y, c), c.t, yy);
#pragma unroll
    for (int k = 0; k < 12; ++k) xx[k] = xx[k] + aa * yy[k];
  }
  CloverPtr C{};
  if (which == 1) C = c.sd->Cinv[P64][parity];
  if (which == 2) C = c.sd->A[parity];
  if (C.p) {
    const int site = c.t / c.nrhs;
    double cb[36];
    clover_load_block(C, site, 0, cb);
    clover_apply_block(cb, xx, 0);
    clover_load_block(C, site, 1, cb);
    clover_apply_block(cb, xx, 1);
  }
  SpinorIO<P64>::store(view<P64>(out, c), c.t, xx);
}

// |b - (A x - kappa t)|^2 and |b|^2 per system on one parity
__global__ void __launch_bounds__(256) k_residual(SpinorPtr b, SpinorPtr x, SpinorPtr t, int parity, const SegDesc *segs, int nseg,
                           double *partials) {
  const TCtx c = tctx(segs, nseg, nullptr);
  if (!c.sd) return;
  double acc[2] = {0, 0};
  if (c.valid) {
    cplx<double> bb[12], xx[12], tt[12];
    SpinorIO<P64>::load(view<P64>(b, c), c.t, bb);
    SpinorIO<P64>::load(view<P64>(x, c), c.t, xx);
    SpinorIO<P64>::load(view<P64>(t, c), c.t, tt);
    const CloverPtr A = c.sd->A[parity];
    if (A.p) {
      const int site = c.t / c.nrhs;
      double cb[36];
      clover_load_block(A, site, 0, cb);
      clover_apply_block(cb, xx, 0);
      clover_load_block(A, site, 1, cb);
      clover_apply_block(cb, xx, 1);
    }
    const double kappa = c.sd->kappa;
#pragma unroll
    for (int k = 0; k < 12; ++k) {
      cplx<double> r = bb[k] - (xx[k] - kappa * tt[k]);
      acc[0] += norm2(r);
      acc[1] += norm2(bb[k]);
    }
  }
  block_partial_rhs<2>(acc, partials, blockIdx.x, c.t0, c.nrhs, c.sd->pcap);
}

// ---------------------------------------------------------------- CG on the normal equations (FP64)
// x += alpha p ; res -= alpha q ; |res|^2 per system
__global__ void __launch_bounds__(256) k_cg_xr(SpinorPtr x, SpinorPtr r, SpinorPtr p, SpinorPtr q, const SegDesc *segs, int nseg,
                        const SysState *st, double *partials) {
  const TCtx c = tctx(segs, nseg, st);
  if (!c.sd) return;
  double acc[1] = {0};
  if (c.active) {
    const double al = st[c.sys].alpha[0];
    cplx<double> xx[12], rr[12], pp[12], qq[12];
    SpinorIO<P64>::load(view<P64>(x, c), c.t, xx);
    SpinorIO<P64>::load(view<P64>(p, c), c.t, pp);
    SpinorIO<P64>::load(view<P64>(r, c), c.t, rr);
    SpinorIO<P64>::load(view<P64>(q, c), c.t, qq);
#pragma unroll
    for (int k = 0; k < 12; ++k) {
      xx[k] = xx[k] + al * pp[k];
      rr[k] = rr[k] - al * qq[k];
      acc[0] += norm2(rr[k]);
    }
    SpinorIO<P64>::store(view<P64>(x, c), c.t, xx);
    SpinorIO<P64>::store(view<P64>(r, c), c.t, rr);
  }
  block_partial_rhs<1>(acc, partials, blockIdx.x, c.t0, c.nrhs, c.sd->pcap);
}

// p = z + beta p
__global__ void __launch_bounds__(256) k_cg_p(SpinorPtr p, SpinorPtr z, const SegDesc *segs, int nseg, const SysState *st) {
  const TCtx c = tctx(segs, nseg, st);
  if (!c.active) return;
  const double be = st[c.sys].beta[0];
  cplx<double> pp[12], zz[12];
  SpinorIO<P64>::load(view<P64>(z, c), c.t, zz);
  if (be != 0.0) {
    SpinorIO<P64>::load(view<P64>(p, c), c.t, pp);
#pragma unroll
    for (int k = 0; k < 12; ++k) zz[k] = zz[k] + be * pp[k];
  }
  SpinorIO<P64>::store(view<P64>(p, c), c.t, zz);
}

// out = gamma5 Ao^-1 in on the odd sites (gamma5 = diag(1, 1, -1, -1) in this basis)
__global__ void __launch_bounds__(256) k_g5_clinv(SpinorPtr out, SpinorPtr in, const SegDesc *segs, int nseg, const SysState *st) {
  const TCtx c = tctx(segs, nseg, st);
  if (!c.active) return;
  cplx<double> v[12];
  SpinorIO<P64>::load(view<P64>(in, c), c.t, v);
  const CloverPtr C = c.sd->Cinv[P64][1];
  if (C.p) {
    const int site = c.t / c.nrhs;
    double cb[36];
    clover_load_block(C, site, 0, cb);
    clover_apply_block(cb, v, 0);
    clover_load_block(C, site, 1, cb);
    clover_apply_block(cb, v, 1);
  }
#pragma unroll
  for (int k = 6; k < 12; ++k) v[k] = mk(-v[k].re, -v[k].im);
  SpinorIO<P64>::store(view<P64>(out, c), c.t, v);
}

// z = r - kappa^2 gamma5 y ; |z|^2 per system
__global__ void __launch_bounds__(256) k_cg_z(SpinorPtr z, SpinorPtr r, SpinorPtr y, const SegDesc *segs, int nseg, const SysState *st,
                       double *partials) {
  const TCtx c = tctx(segs, nseg, st);
  if (!c.sd) return;
  double acc[1] = {0};
  if (c.active) {
    const double k2 = c.sd->kappa * c.sd->kappa;
    cplx<double> rr[12], yy[12];
    SpinorIO<P64>::load(view<P64>(r, c), c.t, rr);
    SpinorIO<P64>::load(view<P64>(y, c), c.t, yy);
#pragma unroll
    for (int k = 0; k < 12; ++k) {
      const double sg = k < 6 ? k2 : -k2;
      rr[k] = rr[k] - sg * yy[k];
      acc[0] += norm2(rr[k]);
    }
    SpinorIO<P64>::store(view<P64>(z, c), c.t, rr);
  }
  block_partial_rhs<1>(acc, partials, blockIdx.x, c.t0, c.nrhs, c.sd->pcap);
}

// |q|^2 per system of the running systems
__global__ void __launch_bounds__(256) k_cg_norm(SpinorPtr q, const SegDesc *segs, int nseg, const SysState *st, double *partials) {
  const TCtx c = tctx(segs, nseg, st);
  if (!c.sd) return;
  double acc[1] = {0};
  if (c.active) {
    cplx<double> v[12];
    SpinorIO<P64>::load(view<P64>(q, c), c.t, v);
#pragma unroll
    for (int k = 0; k < 12; ++k) acc[0] += norm2(v[k]);
  }
  block_partial_rhs<1>(acc, partials, blockIdx.x, c.t0, c.nrhs, c.sd->pcap);
}

// ---------------------------------------------------------------- finishing kernel
// One block per system: fixed-order sum of the system's block partials, then
// the epilogue.
template <int NR>
__global__ void __launch_bounds__(1024) k_finish(const double *partials, const SegDesc *segs, int nseg, int bsz,
                                                int kind, SysState *st, double *out) {
  const int sys = blockIdx.x;
  int seg = -1, r = 0, blk0 = 0, nb = 0, acc_b = 0;
  for (int k = 0; k < nseg; ++k) {
    const int n = (segs[k].nthr + bsz - 1) / bsz;
    if (sys >= segs[k].sys0 && sys < segs[k].sys0 + segs[k].nrhs) {
      seg = k;
      r = sys - segs[k].sys0;
      blk0 = acc_b;
      nb = n;
      break;
    }
    acc_b += n;
  }
  if (seg < 0) return;
  if (st && (kind == FK_ALPHA || kind == FK_OMEGA || kind == FK_RHO || kind == FK_RELIABLE || kind == FK_CG_ALPHA ||
             kind == FK_CG_RHO) &&
      st[sys].done)
    return;
  // this system's partials are contiguous: slots r * pcap + blk0 .. + nb - 1;
  // up to 8 independent loads in flight per thread (the kernel is latency bound)
  const double *part = partials + ((size_t)r * segs[seg].pcap + blk0) * NR;
  double acc4[4][NR];
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int k = 0; k < NR; ++k) acc4[u][k] = 0;
  const int bs = blockDim.x;
  int b = threadIdx.x;
  for (; b + 7 * bs < nb; b += 8 * bs) {
    double v[8][NR];
#pragma unroll
    for (int u = 0; u < 8; ++u)
#pragma unroll
      for (int k = 0; k < NR; ++k) v[u][k] = part[(size_t)(b + u * bs) * NR + k];
#pragma unroll
    for (int u = 0; u < 8; ++u)
#pragma unroll
      for (int k = 0; k < NR; ++k) acc4[u & 3][k] += v[u][k];
  }
  for (; b < nb; b += bs) {
#pragma unroll
    for (int k = 0; k < NR; ++k) acc4[0][k] += part[(size_t)b * NR + k];
  }
  double acc[NR];
#pragma unroll
  for (int k = 0; k < NR; ++k) acc[k] = (acc4[0][k] + acc4[1][k]) + (acc4[2][k] + acc4[3][k]);
  __shared__ double tot_sh[NR];
  block_partial<NR>(acc, tot_sh, 0);
  __syncthreads();
  if (threadIdx.x != 0) return;
  double tot[3] = {0, 0, 0};
#pragma unroll
  for (int k = 0; k < NR; ++k) tot[k] = tot_sh[k];
  SysState *s = st ? st + sys : nullptr;
  switch (kind) {
    case FK_INIT: fin_init(s, tot[0]); break;
    case FK_STORE:
      for (int k = 0; k < NR; ++k) out[(size_t)sys * NR + k] = tot[k];
      break;
    case FK_ACCUM:
      for (int k = 0; k < NR; ++k) out[(size_t)sys * NR + k] += tot[k];
      break;
    case FK_STORE_TMP:
      for (int k = 0; k < NR; ++k) s->tmp[k] = tot[k];
      break;
    case FK_ALPHA: fin_alpha(s, tot[0], tot[1]); break;
    case FK_OMEGA: fin_omega(s, tot[0], tot[1], tot[2]); break;
    case FK_RHO: fin_rho(s, tot[0], tot[1], tot[2]); break;
    case FK_RELIABLE: fin_reliable(s, tot[0], tot[1], tot[2]); break;
    case FK_CG_INIT: fin_cg_init(s, tot[0], s->tmp[0]); break;
    case FK_CG_ALPHA: fin_cg_alpha(s, tot[0]); break;
    case FK_CG_RHO: fin_cg_rho(s, tot[0]); break;
    default: break;
  }
}

int finish(const BatchOp &op, const double *partials, int bsz, int nr, int kind, SysState *st, double *out,
           hipStream_t s) {
  if (kind == FK_NONE) return 0;
  const int nsys = op.nsys;
  const SegDesc *d = op.dev_segs;
  const int nseg = (int)op.segs.size();
  switch (nr) {
    case 1: hipLaunchKernelGGL(k_finish<1>, dim3(nsys), dim3(1024), 0, s, partials, d, nseg, bsz, kind, st, out); break;
    case 2: hipLaunchKernelGGL(k_finish<2>, dim3(nsys), dim3(1024), 0, s, partials, d, nseg, bsz, kind, st, out); break;
    case 3: hipLaunchKernelGGL(k_finish<3>, dim3(nsys), dim3(1024), 0, s, partials, d, nseg, bsz, kind, st, out); break;
    default: set_error(QCU_ERR_ARG, "finish: bad reduction width"); return QCU_ERR_ARG;
  }
  QCU_HIP(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------- launchers
int batch_blocks(const BatchOp &op, int bsz) {
  int nb = 0;
  for (const SegDesc &d : op.segs) nb += (d.nthr + bsz - 1) / bsz;
  return nb;
}

#define QCU_SEGS op.dev_segs, (int)op.segs.size()

int bicg_init(const BatchOp &op, SpinorPtr r, SpinorPtr rt, SpinorPtr p, SpinorPtr v, SpinorPtr x, int prec,
              SpinorPtr b, SysState *st, double *partials, hipStream_t s) {
  const int bs = 256, nb = batch_blocks(op, bs);
  // x precision: FP64 for the plain solver, FP32 for the sloppy accumulator
  if (prec == P64)
    hipLaunchKernelGGL((k_bicg_init<P64, P64, P64>), dim3(nb), dim3(bs), 0, s, r, rt, p, v, x, b, QCU_SEGS, partials);
  else if (prec == P32)
    hipLaunchKernelGGL((k_bicg_init<P64, P32, P32>), dim3(nb), dim3(bs), 0, s, r, rt, p, v, x, b, QCU_SEGS, partials);
  else
    hipLaunchKernelGGL((k_bicg_init<P64, P16, P32>), dim3(nb), dim3(bs), 0, s, r, rt, p, v, x, b, QCU_SEGS, partials);
  QCU_HIP(hipGetLastError());
  return finish(op, partials, bs, 1, FK_INIT, st, nullptr, s);
}

int bicg_p_update(const BatchOp &op, SpinorPtr p, SpinorPtr r, SpinorPtr v, int prec, SysState *st, hipStream_t s) {
  const int bs = 256, nb = batch_blocks(op, bs);
  switch (prec) {
    case P64: hipLaunchKernelGGL(k_p_update<P64>, dim3(nb), dim3(bs), 0, s, p, r, v, QCU_SEGS, st); break;
    case P32: hipLaunchKernelGGL(k_p_update<P32>, dim3(nb), dim3(bs), 0, s, p, r, v, QCU_SEGS, st); break;
    default: hipLaunchKernelGGL(k_p_update<P16>, dim3(nb), dim3(bs), 0, s, p, r, v, QCU_SEGS, st); break;
  }
  QCU_HIP(hipGetLastError());
  return 0;
}

int bicg_s_update(const BatchOp &op, SpinorPtr s_, SpinorPtr r, SpinorPtr v, int prec, SysState *st, hipStream_t s) {
  const int bs = 256, nb = batch_blocks(op, bs);
  switch (prec) {
    case P64: hipLaunchKernelGGL(k_s_update<P64>, dim3(nb), dim3(bs), 0, s, s_, r, v, QCU_SEGS, st); break;
    case P32: hipLaunchKernelGGL(k_s_update<P32>, dim3(nb), dim3(bs), 0, s, s_, r, v, QCU_SEGS, st); break;
    default: hipLaunchKernelGGL(k_s_update<P16>, dim3(nb), dim3(bs), 0, s, s_, r, v, QCU_SEGS, st); break;
  }
  QCU_HIP(hipGetLastError());
  return 0;
}

template <int PC>
static int xr_launch(const BatchOp &op, SpinorPtr x, int px, SpinorPtr r, SpinorPtr rt, SpinorPtr p, SpinorPtr sx,
                     SpinorPtr s_, SpinorPtr t, int prec, SysState *st, double *partials, hipStream_t s) {
  const int bs = 256, nb = batch_blocks(op, bs);
  if (px == P64 && prec == P64)
    hipLaunchKernelGGL((k_xr_update<P64, P64, PC>), dim3(nb), dim3(bs), 0, s, x, r, rt, p, sx, s_, t, QCU_SEGS, st,
                       partials);
  else if (px == P32 && prec == P32)
    hipLaunchKernelGGL((k_xr_update<P32, P32, PC>), dim3(nb), dim3(bs), 0, s, x, r, rt, p, sx, s_, t, QCU_SEGS, st,
                       partials);
  else if (px == P32 && prec == P16)
    hipLaunchKernelGGL((k_xr_update<P32, P16, PC>), dim3(nb), dim3(bs), 0, s, x, r, rt, p, sx, s_, t, QCU_SEGS, st,
                       partials);
  else {
    set_error(QCU_ERR_ARG, "bicg_xr_update: unsupported precision pair");
    return QCU_ERR_ARG;
  }
  QCU_HIP(hipGetLastError());
  return finish(op, partials, bs, 3, FK_RHO, st, nullptr, s);
}

int bicg_xr_update(const BatchOp &op, SpinorPtr x, int px, SpinorPtr r, SpinorPtr rt, SpinorPtr p, SpinorPtr s_,
                   SpinorPtr t, int prec, SysState *st, double *partials, hipStream_t s) {
  return xr_launch<0>(op, x, px, r, rt, p, s_, s_, t, prec, st, partials, s);
}

int bicg_xr_update_pc(const BatchOp &op, SpinorPtr x, int px, SpinorPtr r, SpinorPtr rt, SpinorPtr ph, SpinorPtr sh,
                      SpinorPtr s_, SpinorPtr t, int prec, SysState *st, double *partials, hipStream_t s) {
  return xr_launch<1>(op, x, px, r, rt, ph, sh, s_, t, prec, st, partials, s);
}

template <int NA>
static int dot2(const BatchOp &op, SpinorPtr a, SpinorPtr b, int prec, SysState *st, double *partials, int kind,
                hipStream_t s) {
  const int bs = 256, nb = batch_blocks(op, bs);
  switch (prec) {
    case P64: hipLaunchKernelGGL((k_dot2<P64, NA>), dim3(nb), dim3(bs), 0, s, a, b, QCU_SEGS, st, partials); break;
    case P32: hipLaunchKernelGGL((k_dot2<P32, NA>), dim3(nb), dim3(bs), 0, s, a, b, QCU_SEGS, st, partials); break;
    default: hipLaunchKernelGGL((k_dot2<P16, NA>), dim3(nb), dim3(bs), 0, s, a, b, QCU_SEGS, st, partials); break;
  }
  QCU_HIP(hipGetLastError());
  return finish(op, partials, bs, 3, kind, st, nullptr, s);
}

int omega_dot(const BatchOp &op, SpinorPtr t, SpinorPtr s_, int prec, SysState *st, double *partials, hipStream_t s) {
  return dot2<1>(op, t, s_, prec, st, partials, FK_OMEGA, s);
}

int reliable_dot(const BatchOp &op, SpinorPtr rt, SpinorPtr r, int prec, SysState *st, double *partials,
                 hipStream_t s) {
  return dot2<0>(op, rt, r, prec, st, partials, FK_RELIABLE, s);
}

int axpby(const BatchOp &op, double a, SpinorPtr x, int px, double b, SpinorPtr y, int py, hipStream_t s) {
  const int bs = 256, nb = batch_blocks(op, bs);
#define QCU_AXPBY(A, B)                                                                         \
  if (px == A && py == B) {                                                                   \
    hipLaunchKernelGGL((k_axpby<A, B>), dim3(nb), dim3(bs), 0, s, a, x, b, y, QCU_SEGS);      \
    QCU_HIP(hipGetLastError());                                                               \
    return 0;                                                                                 \
  }
  QCU_AXPBY(P64, P64)
  QCU_AXPBY(P32, P64)
  QCU_AXPBY(P16, P64)
  QCU_AXPBY(P64, P32)
  QCU_AXPBY(P32, P32)
  QCU_AXPBY(P16, P32)
  QCU_AXPBY(P64, P16)
  QCU_AXPBY(P32, P16)
  QCU_AXPBY(P16, P16)
#undef QCU_AXPBY
  set_error(QCU_ERR_ARG, "axpby: bad precision");
  return QCU_ERR_ARG;
}

int spinor_convert(const BatchOp &op, SpinorPtr dst, int pdst, SpinorPtr src, int psrc, hipStream_t s) {
  return axpby(op, 1.0, src, psrc, 0.0, dst, pdst, s);
}

int zero_spinor(const BatchOp &op, SpinorPtr x, int px, hipStream_t s) { return axpby(op, 0.0, x, px, 0.0, x, px, s); }

int xmy_norm(const BatchOp &op, SpinorPtr b, SpinorPtr r, SysState *st, double *partials, int store,
             hipStream_t s) {
  const int bs = 256, nb = batch_blocks(op, bs);
  hipLaunchKernelGGL(k_xmy_norm, dim3(nb), dim3(bs), 0, s, b, r, QCU_SEGS, partials);
  QCU_HIP(hipGetLastError());
  return finish(op, partials, bs, 1, store ? FK_STORE_TMP : FK_NONE, st, nullptr, s);
}

int norm2_dev(const BatchOp &op, SpinorPtr x, int px, double *out_dev, double *partials, hipStream_t s) {
  const int bs = 256, nb = batch_blocks(op, bs);
  switch (px) {
    case P64: hipLaunchKernelGGL(k_norm2<P64>, dim3(nb), dim3(bs), 0, s, x, QCU_SEGS, partials); break;
    case P32: hipLaunchKernelGGL(k_norm2<P32>, dim3(nb), dim3(bs), 0, s, x, QCU_SEGS, partials); break;
    default: hipLaunchKernelGGL(k_norm2<P16>, dim3(nb), dim3(bs), 0, s, x, QCU_SEGS, partials); break;
  }
  QCU_HIP(hipGetLastError());
  return finish(op, partials, bs, 1, FK_STORE, nullptr, out_dev, s);
}

int cl_axpy(const BatchOp &op, SpinorPtr out, SpinorPtr x, double a, bool a_is_kappa, SpinorPtr y, int which,
            int parity, hipStream_t s) {
  const int bs = 256, nb = batch_blocks(op, bs);
  hipLaunchKernelGGL(k_cl_axpy, dim3(nb), dim3(bs), 0, s, out, x, a, (int)a_is_kappa, y, which, parity, QCU_SEGS);
  QCU_HIP(hipGetLastError());
  return 0;
}

int cg_xr(const BatchOp &op, SpinorPtr x, SpinorPtr r, SpinorPtr p, SpinorPtr q, SysState *st, double *partials,
          hipStream_t s) {
  const int bs = 256, nb = batch_blocks(op, bs);
  hipLaunchKernelGGL(k_cg_xr, dim3(nb), dim3(bs), 0, s, x, r, p, q, QCU_SEGS, st, partials);
  QCU_HIP(hipGetLastError());
  return finish(op, partials, bs, 1, FK_STORE_TMP, st, nullptr, s);
}

int cg_p(const BatchOp &op, SpinorPtr p, SpinorPtr z, SysState *st, hipStream_t s) {
  const int bs = 256, nb = batch_blocks(op, bs);
  hipLaunchKernelGGL(k_cg_p, dim3(nb), dim3(bs), 0, s, p, z, QCU_SEGS, st);
  QCU_HIP(hipGetLastError());
  return 0;
}

int g5_clinv(const BatchOp &op, SpinorPtr out, SpinorPtr in, SysState *st, hipStream_t s) {
  const int bs = 256, nb = batch_blocks(op, bs);
  hipLaunchKernelGGL(k_g5_clinv, dim3(nb), dim3(bs), 0, s, out, in, QCU_SEGS, st);
  QCU_HIP(hipGetLastError());
  return 0;
}

int cg_z(const BatchOp &op, SpinorPtr z, SpinorPtr r, SpinorPtr y, int kind, SysState *st, double *partials,
         hipStream_t s) {
  const int bs = 256, nb = batch_blocks(op, bs);
  hipLaunchKernelGGL(k_cg_z, dim3(nb), dim3(bs), 0, s, z, r, y, QCU_SEGS, st, partials);
  QCU_HIP(hipGetLastError());
  return finish(op, partials, bs, 1, kind, st, nullptr, s);
}

int cg_qnorm(const BatchOp &op, SpinorPtr q, SysState *st, double *partials, hipStream_t s) {
  const int bs = 256, nb = batch_blocks(op, bs);
  hipLaunchKernelGGL(k_cg_norm, dim3(nb), dim3(bs), 0, s, q, QCU_SEGS, st, partials);
  QCU_HIP(hipGetLastError());
  return finish(op, partials, bs, 1, FK_CG_ALPHA, st, nullptr, s);
}

int residual_parity(const BatchOp &op, SpinorPtr b, SpinorPtr x, SpinorPtr t, int parity, double *partials,
                    double *out, hipStream_t s) {
  const int bs = 256, nb = batch_blocks(op, bs);
  hipLaunchKernelGGL(k_residual, dim3(nb), dim3(bs), 0, s, b, x, t, parity, QCU_SEGS, partials);
  QCU_HIP(hipGetLastError());
  return finish(op, partials, bs, 2, FK_ACCUM, nullptr, out, s);
}

#undef QCU_SEGS

}  // namespace qcu
