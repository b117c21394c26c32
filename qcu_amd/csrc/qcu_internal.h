// qcu_internal.h -- host-side structures and kernel launchers shared by the
// translation units of libqcu.so.  Nothing here crosses the C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "qcu_common.h"

namespace qcu {

// --------------------------------------------------------------------------
// errors: recorded per thread, never thrown across the C ABI
// --------------------------------------------------------------------------
void set_error(int code, const std::string &msg);
void clear_error();
int last_error_code();

#define QCU_HIP(call)                                                                  \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess) {                                                            \
      ::qcu::set_error(100 + (int)e_, std::string(#call) + ": " + hipGetErrorString(e_)); \
      return 100 + (int)e_;                                                            \
    }                                                                                  \
  } while (0)

#define QCU_CHECK(cond, code, msg)       \
  do {                                   \
    if (!(cond)) {                       \
      ::qcu::set_error((code), (msg));   \
      return (code);                     \
    }                                    \
  } while (0)

#define QCU_TRY(expr)        \
  do {                       \
    int rc_ = (expr);        \
    if (rc_) return rc_;     \
  } while (0)

enum ErrCode {
  QCU_OK = 0,
  QCU_ERR_ARG = 1,
  QCU_ERR_DIMS = 2,
  QCU_ERR_GRID = 3,
  QCU_ERR_NOT_CONVERGED = 4,
  QCU_ERR_BREAKDOWN = 5,
  QCU_ERR_ALLOC = 6,
  QCU_ERR_STATE = 7,
};

// --------------------------------------------------------------------------
// device allocations
// --------------------------------------------------------------------------
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  int alloc(size_t n);
  void release();
  ~DevBuf() { release(); }
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
};

// Spinor storage for one parity in one precision (owning)
struct Spinor {
  int prec = P64;
  int vh = 0;
  DevBuf data, norm;
  int create(int prec_, int vh_);
  SpinorPtr ptr() const {
    SpinorPtr s;
    s.p = data.p;
    s.norm = (float *)norm.p;
    s.stride = vh;
    return s;
  }
};

size_t spinor_bytes(int prec, int vh);

// Per-system solver scalars, resident on the device.  Updated only by the
// single-thread epilogue of the reductions (see blas.hip).
struct SysState {
  double rho[2], rho_prev[2], alpha[2], beta[2], omega[2];
  double rr;        // |r|^2 of the iterated (sloppy) residual
  double rr_stop;   // stop when rr < rr_stop
  double rr_ref;    // |r|^2 at the last reliable update
  double delta2;    // reliable update when rr < delta2 * rr_ref
  double tmp[4];
  int done;         // 1: converged, 2: breakdown, 3: reliable update requested
  int iter;
  int pad[2];
};

// --------------------------------------------------------------------------
// kernel launchers (implemented in the .hip files)
// --------------------------------------------------------------------------
struct LaunchCfg {
  int block = 256;
  int remap = 1;
  int group = 0;  // hop scheduling groups (see dslash.hip)
};

// layout.hip
int gauge_ext_to_internal(const void *ext, const Geom &g, int prec, void *a, void *b, hipStream_t s);
int spinor_ext_to_internal(const void *ext, const Geom &g, SpinorPtr dst, int prec, hipStream_t s);
int spinor_internal_to_ext(SpinorPtr src, const Geom &g, void *ext, int prec, hipStream_t s);
int spinor_convert(SpinorPtr dst, int pdst, SpinorPtr src, int psrc, int vh, hipStream_t s);

// dslash.hip
struct HopOp {
  LinkPtr U;
  CloverPtr C[2];  // inverse clover per parity; C[p].p == nullptr for Wilson
  Geom g;
  int prec;
};
enum FinKind { FIN_NONE = 0, FIN_INIT = 1, FIN_STORE = 3 };
enum HopEpi { EPI_PLAIN = 0, EPI_XPAY = 1 };
enum HopRed { RED_NONE = 0, RED_DOT_Y_OUT = 1, RED_OUT_X = 2 };
// out = [A_p^-1] D_{p,1-p} in                                  (EPI_PLAIN)
// out = x - k2 [A_p^-1] D_{p,1-p} in                           (EPI_XPAY)
// RED_DOT_Y_OUT: <y,out> -> alpha update; RED_OUT_X: <out,x>,|out|^2 -> omega
int hop(const HopOp &op, SpinorPtr out, SpinorPtr in, int parity, int epi, SpinorPtr x, double k2,
        int red, SpinorPtr y, SysState *st, double *partials, const LaunchCfg &cfg, hipStream_t s,
        bool use_clover = true);

// clover.hip
// packed A (clov, optional) and A^-1 (cinv) of one parity from FP64 internal links
int clover_build(const LinkPtr &U64, const Geom &g, double coeff, int parity, void *clov, void *cinv,
                 int *bad_dev, hipStream_t s);
int clover_convert(const void *c64, const Geom &g, int prec, void *dst, float *norm, hipStream_t s);
// out[0] += |b - (A x - kappa t)|^2, out[1] += |b|^2 over one parity
int residual_parity(SpinorPtr b, SpinorPtr x, SpinorPtr t, CloverPtr A, double kappa, int vh,
                    double *partials, double *out, hipStream_t s);

// blas.hip
// one-block kernel: fixed-order sum of nblocks partials of width nr, then the
// epilogue `kind` (FinKindAll in qcu_reduce.h)
int finish(const double *partials, int nblocks, int nr, int kind, SysState *st, double *out, hipStream_t s);
int bicg_init(SpinorPtr r, SpinorPtr rt, SpinorPtr p, SpinorPtr v, SpinorPtr x, int prec,
              SpinorPtr b, int pb, int vh, SysState *st, double *partials, const LaunchCfg &cfg,
              hipStream_t s);
int bicg_p_update(SpinorPtr p, SpinorPtr r, SpinorPtr v, int prec, int vh, SysState *st,
                  const LaunchCfg &cfg, hipStream_t s);
int bicg_s_update(SpinorPtr s_, SpinorPtr r, SpinorPtr v, int prec, int vh, SysState *st,
                  const LaunchCfg &cfg, hipStream_t s);
int bicg_xr_update(SpinorPtr x, int px, SpinorPtr r, SpinorPtr rt, SpinorPtr p, SpinorPtr s_,
                   SpinorPtr t, int prec, int vh, SysState *st, double *partials,
                   const LaunchCfg &cfg, hipStream_t s);
// y = a*x + b*y  (real a,b, mixed precisions allowed)
int axpby(double a, SpinorPtr x, int px, double b, SpinorPtr y, int py, int vh, hipStream_t s);
// r = b - r; with FIN_STORE |r|^2 lands in st->tmp[0]
int xmy_norm(SpinorPtr b, int pb, SpinorPtr r, int pr, int vh, SysState *st, double *partials, int fin,
             hipStream_t s);
int reliable_dot(SpinorPtr rt, SpinorPtr r, int prec, int vh, SysState *st, double *partials,
                 hipStream_t s);
// out = C (x + a y) in FP64, C a packed clover matrix of one parity or identity (C.p == null)
int cl_axpy(SpinorPtr out, SpinorPtr x, double a, SpinorPtr y, CloverPtr C, int vh, hipStream_t s);
int norm2_dev(SpinorPtr x, int px, int vh, double *out_dev, double *partials, hipStream_t s);
int zero_spinor(SpinorPtr x, int px, int vh, hipStream_t s);

// gen.hip
int gen_gauge(void *ext, const Geom &g, double eps, uint64_t seed, hipStream_t s);
int gen_gauss(void *ext, long n_complex, uint64_t seed, hipStream_t s);

}  // namespace qcu
