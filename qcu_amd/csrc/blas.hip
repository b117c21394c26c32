// blas.hip -- fused BLAS-1 site kernels of the BiCGSTAB recurrence.
//
// Each kernel touches every vector once and reduces in FP64 with the fused
// grid reduction of qcu_reduce.h; the scalar epilogues of qcu_state.h keep
// rho, alpha, beta, omega on the device (the reference reads them back
// through cuBLAS and 1-thread kernels, src/bistabcg.cu:4-180,
// include/lattice_bistabcg.h:149-163).  Vectors may be FP64, FP32 or FP16 with
// a per-site scale; arithmetic is FP32 for the sloppy precisions and FP64
// otherwise, and every global sum is FP64.
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


QCU_D void fin_norm(SysState *st, double rr, int kind, double *out) {
  if (kind == FIN_NONE) return;
  if (kind == FIN_STORE) {
    *out = rr;
    return;
  }
  // FIN_INIT: r~ = r = b_hat, x = 0, p = v = 0
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

// r = b (converted), r~ = r, p = v = 0, x = 0; |r|^2 -> state
template <int Pb, int P, int Px>
__global__ void k_bicg_init(SpinorPtr r, SpinorPtr rt, SpinorPtr p, SpinorPtr v, SpinorPtr x,
                            SpinorPtr b, int vh, double *partials) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  double acc[1] = {0};
  if (i < vh) {
    cplx<RT<Pb>> bb[12];
    SpinorIO<Pb>::load(b, i, bb);
    cplx<RT<P>> rr[12], z[12];
    cvt<P, Pb>(bb, rr);
#pragma unroll
    for (int k = 0; k < 12; ++k) {
      z[k] = mk<RT<P>>(0, 0);
      acc[0] += (double)bb[k].re * bb[k].re + (double)bb[k].im * bb[k].im;
    }
    SpinorIO<P>::store(r, i, rr);
    SpinorIO<P>::store(rt, i, rr);
    SpinorIO<P>::store(p, i, z);
    SpinorIO<P>::store(v, i, z);
    cplx<RT<Px>> zx[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) zx[k] = mk<RT<Px>>(0, 0);
    SpinorIO<Px>::store(x, i, zx);
  }
  block_partial<1>(acc, partials, blockIdx.x);
}

// p = r + beta (p - omega v)
template <int P>
__global__ void k_p_update(SpinorPtr p, SpinorPtr r, SpinorPtr v, int vh, const SysState *st) {
  if (st->done) return;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= vh) return;
  typedef RT<P> T;
  const cplx<T> beta = cs<T>(st->beta), omega = cs<T>(st->omega);
  cplx<T> pp[12], rr[12], vv[12];
  SpinorIO<P>::load(p, i, pp);
  SpinorIO<P>::load(r, i, rr);
  SpinorIO<P>::load(v, i, vv);
#pragma unroll
  for (int k = 0; k < 12; ++k) pp[k] = rr[k] + beta * (pp[k] - omega * vv[k]);
  SpinorIO<P>::store(p, i, pp);
}

// s = r - alpha v
template <int P>
__global__ void k_s_update(SpinorPtr s, SpinorPtr r, SpinorPtr v, int vh, const SysState *st) {
  if (st->done) return;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= vh) return;
  typedef RT<P> T;
  const cplx<T> alpha = cs<T>(st->alpha);
  cplx<T> rr[12], vv[12];
  SpinorIO<P>::load(r, i, rr);
  SpinorIO<P>::load(v, i, vv);
#pragma unroll
  for (int k = 0; k < 12; ++k) rr[k] = rr[k] - alpha * vv[k];
  SpinorIO<P>::store(s, i, rr);
}

// x += alpha P + omega S ; r = s - omega t ; <r~, r>, |r|^2 -> rho, beta
// (P, S) = (p, s) for the plain method, (K p, K s) with a right preconditioner K
template <int Px, int P, int PC>
__global__ void k_xr_update(SpinorPtr x, SpinorPtr r, SpinorPtr rt, SpinorPtr p, SpinorPtr sx, SpinorPtr s,
                            SpinorPtr t, int vh, SysState *st, double *partials) {
  if (st->done) return;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  double acc[3] = {0, 0, 0};
  if (i < vh) {
    typedef RT<P> T;
    typedef RT<Px> TX;
    cplx<T> pp[12], ss[12], tt[12];
    SpinorIO<P>::load(p, i, pp);
    SpinorIO<P>::load(s, i, ss);
    SpinorIO<P>::load(t, i, tt);
    cplx<TX> xx[12];
    SpinorIO<Px>::load(x, i, xx);
    const cplx<TX> al = cs<TX>(st->alpha), omx = cs<TX>(st->omega);
    if (PC) {
      cplx<T> sh[12];
      SpinorIO<P>::load(sx, i, sh);
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
    SpinorIO<Px>::store(x, i, xx);
    const cplx<T> om = cs<T>(st->omega);
    cplx<T> rtt[12];
    SpinorIO<P>::load(rt, i, rtt);
#pragma unroll
    for (int k = 0; k < 12; ++k) {
      ss[k] = ss[k] - om * tt[k];
      acc[0] += (double)rtt[k].re * ss[k].re + (double)rtt[k].im * ss[k].im;
      acc[1] += (double)rtt[k].re * ss[k].im - (double)rtt[k].im * ss[k].re;
      acc[2] += (double)ss[k].re * ss[k].re + (double)ss[k].im * ss[k].im;
    }
    SpinorIO<P>::store(r, i, ss);
  }
  block_partial<3>(acc, partials, blockIdx.x);
}

// <t, s>, |t|^2 -> omega
template <int P>
__global__ void k_omega_dot(SpinorPtr t, SpinorPtr s, int vh, SysState *st, double *partials) {
  if (st->done) return;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  double acc[3] = {0, 0, 0};
  if (i < vh) {
    cplx<RT<P>> a[12], b[12];
    SpinorIO<P>::load(t, i, a);
    SpinorIO<P>::load(s, i, b);
#pragma unroll
    for (int k = 0; k < 12; ++k) {
      acc[0] += (double)a[k].re * b[k].re + (double)a[k].im * b[k].im;
      acc[1] += (double)a[k].re * b[k].im - (double)a[k].im * b[k].re;
      acc[2] += (double)a[k].re * a[k].re + (double)a[k].im * a[k].im;
    }
  }
  block_partial<3>(acc, partials, blockIdx.x);
}

// y = a x + b y
template <int Px, int Py>
__global__ void k_axpby(double a, SpinorPtr x, double b, SpinorPtr y, int vh) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= vh) return;
  typedef RT<Py> T;
  cplx<RT<Px>> xx[12];
  cplx<T> yy[12];
  if (a != 0.0) {
    SpinorIO<Px>::load(x, i, xx);
  } else {
#pragma unroll
    for (int k = 0; k < 12; ++k) xx[k] = mk<RT<Px>>(0, 0);
  }
  if (b != 0.0) {
    SpinorIO<Py>::load(y, i, yy);
  } else {
#pragma unroll
    for (int k = 0; k < 12; ++k) yy[k] = mk<T>(0, 0);
  }
#pragma unroll
  for (int k = 0; k < 12; ++k) {
    yy[k] = mk<T>((T)(a * xx[k].re + b * yy[k].re), (T)(a * xx[k].im + b * yy[k].im));
  }
  SpinorIO<Py>::store(y, i, yy);
}

// r = b - r, |r|^2 reduced; r in precision Pr, b in Pb
template <int Pb, int Pr>
__global__ void k_xmy_norm(SpinorPtr b, SpinorPtr r, int vh, double *partials) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  double acc[1] = {0};
  if (i < vh) {
    typedef RT<Pr> T;
    cplx<RT<Pb>> bb[12];
    cplx<T> rr[12];
    SpinorIO<Pb>::load(b, i, bb);
    SpinorIO<Pr>::load(r, i, rr);
#pragma unroll
    for (int k = 0; k < 12; ++k) {
      rr[k] = mk<T>((T)bb[k].re - rr[k].re, (T)bb[k].im - rr[k].im);
      acc[0] += (double)rr[k].re * rr[k].re + (double)rr[k].im * rr[k].im;
    }
    SpinorIO<Pr>::store(r, i, rr);
  }
  block_partial<1>(acc, partials, blockIdx.x);
}

template <int P>
__global__ void k_norm2(SpinorPtr x, int vh, double *partials) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  double acc[1] = {0};
  if (i < vh) {
    cplx<RT<P>> xx[12];
    SpinorIO<P>::load(x, i, xx);
#pragma unroll
    for (int k = 0; k < 12; ++k) acc[0] += (double)xx[k].re * xx[k].re + (double)xx[k].im * xx[k].im;
  }
  block_partial<1>(acc, partials, blockIdx.x);
}

// <r~, r>, |r|^2 after a reliable update
template <int P>
__global__ void k_reliable_dot(SpinorPtr rt, SpinorPtr r, int vh, double *partials) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  double acc[3] = {0, 0, 0};
  if (i < vh) {
    cplx<RT<P>> a[12], b[12];
    SpinorIO<P>::load(rt, i, a);
    SpinorIO<P>::load(r, i, b);
#pragma unroll
    for (int k = 0; k < 12; ++k) {
      acc[0] += (double)a[k].re * b[k].re + (double)a[k].im * b[k].im;
      acc[1] += (double)a[k].re * b[k].im - (double)a[k].im * b[k].re;
      acc[2] += (double)b[k].re * b[k].re + (double)b[k].im * b[k].im;
    }
  }
  block_partial<3>(acc, partials, blockIdx.x);
}

// out = C (x + a y), C the packed clover (inverse) of the parity or identity
__global__ void k_cl_axpy(SpinorPtr out, SpinorPtr x, double a, SpinorPtr y, CloverPtr C, int vh) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= vh) return;
  cplx<double> xx[12];
  SpinorIO<P64>::load(x, i, xx);
  if (a != 0.0) {
    cplx<double> yy[12];
    SpinorIO<P64>::load(y, i, yy);
#pragma unroll
    for (int k = 0; k < 12; ++k) xx[k] = xx[k] + a * yy[k];
  }
  if (C.p) {
    double c[72];
    CloverIO<P64>::load(C, i, c);
    clover_apply_site(c, xx);
  }
  SpinorIO<P64>::store(out, i, xx);
}

// ------------------------------------------------------------------ finishing kernel
// Sums the block partials in a fixed order and applies the epilogue.
template <int NR>
__global__ void __launch_bounds__(256) k_finish(const double *partials, int nblocks, int kind, SysState *st,
                                                double *out) {
  if (st && st->done && (kind == FK_ALPHA || kind == FK_OMEGA || kind == FK_RHO)) return;
  double acc4[4][NR];
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int k = 0; k < NR; ++k) acc4[u][k] = 0;
  const int bs = blockDim.x;
  int b = threadIdx.x;
  for (; b + 3 * bs < nblocks; b += 4 * bs) {
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int k = 0; k < NR; ++k) acc4[u][k] += partials[(size_t)(b + u * bs) * NR + k];
  }
  for (; b < nblocks; b += bs) {
#pragma unroll
    for (int k = 0; k < NR; ++k) acc4[0][k] += partials[(size_t)b * NR + k];
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
  switch (kind) {
    case FK_INIT: fin_norm(st, tot[0], FIN_INIT, nullptr); break;
    case FK_STORE:
      for (int k = 0; k < NR; ++k) out[k] = tot[k];
      break;
    case FK_ACCUM:
      for (int k = 0; k < NR; ++k) out[k] += tot[k];
      break;
    case FK_ALPHA: fin_alpha(st, tot[0], tot[1]); break;
    case FK_OMEGA: fin_omega(st, tot[0], tot[1], tot[2]); break;
    case FK_RHO: fin_rho(st, tot[0], tot[1], tot[2]); break;
    case FK_RELIABLE: fin_reliable(st, tot[0], tot[1], tot[2]); break;
    default: break;
  }
}

int finish(const double *partials, int nblocks, int nr, int kind, SysState *st, double *out, hipStream_t s) {
  if (kind == FK_NONE) return 0;
  switch (nr) {
    case 1: hipLaunchKernelGGL(k_finish<1>, dim3(1), dim3(256), 0, s, partials, nblocks, kind, st, out); break;
    case 2: hipLaunchKernelGGL(k_finish<2>, dim3(1), dim3(256), 0, s, partials, nblocks, kind, st, out); break;
    case 3: hipLaunchKernelGGL(k_finish<3>, dim3(1), dim3(256), 0, s, partials, nblocks, kind, st, out); break;
    default: set_error(QCU_ERR_ARG, "finish: bad reduction width"); return QCU_ERR_ARG;
  }
  QCU_HIP(hipGetLastError());
  return 0;
}

// ------------------------------------------------------------------ launchers
static inline int nblk(int vh, int bs) { return (vh + bs - 1) / bs; }

int bicg_init(SpinorPtr r, SpinorPtr rt, SpinorPtr p, SpinorPtr v, SpinorPtr x, int prec, SpinorPtr b, int pb, int vh,
              SysState *st, double *partials, const LaunchCfg &cfg, hipStream_t s) {
  const int bs = cfg.block, nb = nblk(vh, bs);
  // x precision: FP64 for the plain solver, FP32 for the sloppy accumulator
  if (pb == P64 && prec == P64)
    hipLaunchKernelGGL((k_bicg_init<P64, P64, P64>), dim3(nb), dim3(bs), 0, s, r, rt, p, v, x, b, vh, partials);
  else if (pb == P64 && prec == P32)
    hipLaunchKernelGGL((k_bicg_init<P64, P32, P32>), dim3(nb), dim3(bs), 0, s, r, rt, p, v, x, b, vh, partials);
  else if (pb == P64 && prec == P16)
    hipLaunchKernelGGL((k_bicg_init<P64, P16, P32>), dim3(nb), dim3(bs), 0, s, r, rt, p, v, x, b, vh, partials);
  else {
    set_error(QCU_ERR_ARG, "bicg_init: unsupported precision pair");
    return QCU_ERR_ARG;
  }
  QCU_HIP(hipGetLastError());
  return finish(partials, nb, 1, FK_INIT, st, nullptr, s);
}

int bicg_p_update(SpinorPtr p, SpinorPtr r, SpinorPtr v, int prec, int vh, SysState *st, const LaunchCfg &cfg,
                  hipStream_t s) {
  const int bs = cfg.block, nb = nblk(vh, bs);
  switch (prec) {
    case P64: hipLaunchKernelGGL(k_p_update<P64>, dim3(nb), dim3(bs), 0, s, p, r, v, vh, st); break;
    case P32: hipLaunchKernelGGL(k_p_update<P32>, dim3(nb), dim3(bs), 0, s, p, r, v, vh, st); break;
    default: hipLaunchKernelGGL(k_p_update<P16>, dim3(nb), dim3(bs), 0, s, p, r, v, vh, st); break;
  }
  QCU_HIP(hipGetLastError());
  return 0;
}

int bicg_s_update(SpinorPtr s_, SpinorPtr r, SpinorPtr v, int prec, int vh, SysState *st, const LaunchCfg &cfg,
                  hipStream_t s) {
  const int bs = cfg.block, nb = nblk(vh, bs);
  switch (prec) {
    case P64: hipLaunchKernelGGL(k_s_update<P64>, dim3(nb), dim3(bs), 0, s, s_, r, v, vh, st); break;
    case P32: hipLaunchKernelGGL(k_s_update<P32>, dim3(nb), dim3(bs), 0, s, s_, r, v, vh, st); break;
    default: hipLaunchKernelGGL(k_s_update<P16>, dim3(nb), dim3(bs), 0, s, s_, r, v, vh, st); break;
  }
  QCU_HIP(hipGetLastError());
  return 0;
}

template <int PC>
static int xr_launch(SpinorPtr x, int px, SpinorPtr r, SpinorPtr rt, SpinorPtr p, SpinorPtr sx, SpinorPtr s_,
                     SpinorPtr t, int prec, int vh, SysState *st, double *partials, const LaunchCfg &cfg,
                     hipStream_t s) {
  const int bs = cfg.block, nb = nblk(vh, bs);
  if (px == P64 && prec == P64)
    hipLaunchKernelGGL((k_xr_update<P64, P64, PC>), dim3(nb), dim3(bs), 0, s, x, r, rt, p, sx, s_, t, vh, st,
                       partials);
  else if (px == P32 && prec == P32)
    hipLaunchKernelGGL((k_xr_update<P32, P32, PC>), dim3(nb), dim3(bs), 0, s, x, r, rt, p, sx, s_, t, vh, st,
                       partials);
  else if (px == P32 && prec == P16)
    hipLaunchKernelGGL((k_xr_update<P32, P16, PC>), dim3(nb), dim3(bs), 0, s, x, r, rt, p, sx, s_, t, vh, st,
                       partials);
  else {
    set_error(QCU_ERR_ARG, "bicg_xr_update: unsupported precision pair");
    return QCU_ERR_ARG;
  }
  QCU_HIP(hipGetLastError());
  return finish(partials, nb, 3, FK_RHO, st, nullptr, s);
}

int bicg_xr_update(SpinorPtr x, int px, SpinorPtr r, SpinorPtr rt, SpinorPtr p, SpinorPtr s_, SpinorPtr t, int prec,
                   int vh, SysState *st, double *partials, const LaunchCfg &cfg, hipStream_t s) {
  return xr_launch<0>(x, px, r, rt, p, s_, s_, t, prec, vh, st, partials, cfg, s);
}

int bicg_xr_update_pc(SpinorPtr x, int px, SpinorPtr r, SpinorPtr rt, SpinorPtr ph, SpinorPtr sh, SpinorPtr s_,
                      SpinorPtr t, int prec, int vh, SysState *st, double *partials, const LaunchCfg &cfg,
                      hipStream_t s) {
  return xr_launch<1>(x, px, r, rt, ph, sh, s_, t, prec, vh, st, partials, cfg, s);
}

int omega_dot(SpinorPtr t, SpinorPtr s_, int prec, int vh, SysState *st, double *partials, hipStream_t s) {
  const int bs = 256, nb = nblk(vh, bs);
  switch (prec) {
    case P64: hipLaunchKernelGGL(k_omega_dot<P64>, dim3(nb), dim3(bs), 0, s, t, s_, vh, st, partials); break;
    case P32: hipLaunchKernelGGL(k_omega_dot<P32>, dim3(nb), dim3(bs), 0, s, t, s_, vh, st, partials); break;
    default: hipLaunchKernelGGL(k_omega_dot<P16>, dim3(nb), dim3(bs), 0, s, t, s_, vh, st, partials); break;
  }
  QCU_HIP(hipGetLastError());
  return finish(partials, nb, 3, FK_OMEGA, st, nullptr, s);
}

int axpby(double a, SpinorPtr x, int px, double b, SpinorPtr y, int py, int vh, hipStream_t s) {
  const int bs = 256, nb = nblk(vh, bs);
#define QCU_AXPBY(A, B)                                                                       \
  if (px == A && py == B) {                                                                 \
    hipLaunchKernelGGL((k_axpby<A, B>), dim3(nb), dim3(bs), 0, s, a, x, b, y, vh);          \
    QCU_HIP(hipGetLastError());                                                             \
    return 0;                                                                               \
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

int spinor_convert(SpinorPtr dst, int pdst, SpinorPtr src, int psrc, int vh, hipStream_t s) {
  return axpby(1.0, src, psrc, 0.0, dst, pdst, vh, s);
}

int zero_spinor(SpinorPtr x, int px, int vh, hipStream_t s) { return axpby(0.0, x, px, 0.0, x, px, vh, s); }

int xmy_norm(SpinorPtr b, int pb, SpinorPtr r, int pr, int vh, SysState *st, double *partials, int fin,
             hipStream_t s) {
  const int bs = 256, nb = nblk(vh, bs);
  QCU_CHECK(pb == P64 && pr == P64, QCU_ERR_ARG, "xmy_norm: unsupported precision pair");
  hipLaunchKernelGGL((k_xmy_norm<P64, P64>), dim3(nb), dim3(bs), 0, s, b, r, vh, partials);
  QCU_HIP(hipGetLastError());
  return finish(partials, nb, 1, fin == FIN_STORE ? FK_STORE : FK_NONE, st, st ? st->tmp : nullptr, s);
}

int reliable_dot(SpinorPtr rt, SpinorPtr r, int prec, int vh, SysState *st, double *partials, hipStream_t s) {
  const int bs = 256, nb = nblk(vh, bs);
  switch (prec) {
    case P64: hipLaunchKernelGGL(k_reliable_dot<P64>, dim3(nb), dim3(bs), 0, s, rt, r, vh, partials); break;
    case P32: hipLaunchKernelGGL(k_reliable_dot<P32>, dim3(nb), dim3(bs), 0, s, rt, r, vh, partials); break;
    default: hipLaunchKernelGGL(k_reliable_dot<P16>, dim3(nb), dim3(bs), 0, s, rt, r, vh, partials); break;
  }
  QCU_HIP(hipGetLastError());
  return finish(partials, nb, 3, FK_RELIABLE, st, nullptr, s);
}

int cl_axpy(SpinorPtr out, SpinorPtr x, double a, SpinorPtr y, CloverPtr C, int vh, hipStream_t s) {
  const int bs = 256, nb = nblk(vh, bs);
  hipLaunchKernelGGL(k_cl_axpy, dim3(nb), dim3(bs), 0, s, out, x, a, y, C, vh);
  QCU_HIP(hipGetLastError());
  return 0;
}

int norm2_dev(SpinorPtr x, int px, int vh, double *out_dev, double *partials, hipStream_t s) {
  const int bs = 256, nb = nblk(vh, bs);
  switch (px) {
    case P64: hipLaunchKernelGGL(k_norm2<P64>, dim3(nb), dim3(bs), 0, s, x, vh, partials); break;
    case P32: hipLaunchKernelGGL(k_norm2<P32>, dim3(nb), dim3(bs), 0, s, x, vh, partials); break;
    default: hipLaunchKernelGGL(k_norm2<P16>, dim3(nb), dim3(bs), 0, s, x, vh, partials); break;
  }
  QCU_HIP(hipGetLastError());
  return finish(partials, nb, 1, FK_STORE, nullptr, out_dev, s);
}

}  // namespace qcu
