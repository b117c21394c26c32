// qcu_internal.h -- host-side structures and kernel launchers shared by the
// translation units of libqcu.so.  Nothing here crosses the C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/qcu.h"
#include "qcu_batch.h"
#include "qcu_common.h"

namespace qcu {

// --------------------------------------------------------------------------
// errors: recorded per thread, never thrown across the C ABI
// --------------------------------------------------------------------------
void set_error(int code, const std::string &msg);
void clear_error();
int last_error_code();

#define QCU_HIP(call)                                                                      \
  do {                                                                                     \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess) {                                                                \
      ::qcu::set_error(QCU_ERR_HIP_BASE + (int)e_, std::string(#call) + ": " + hipGetErrorString(e_)); \
      return QCU_ERR_HIP_BASE + (int)e_;                                                   \
    }                                                                                      \
  } while (0)

#define QCU_CHECK(cond, code, msg)     \
  do {                                 \
    if (!(cond)) {                     \
      ::qcu::set_error((code), (msg)); \
      return (code);                   \
    }                                  \
  } while (0)

#define QCU_TRY(expr)    \
  do {                   \
    int rc_ = (expr);    \
    if (rc_) return rc_; \
  } while (0)

// error codes: the QCU_ERR_* enum of include/qcu.h; a failed HIP call returns
// QCU_ERR_HIP_BASE + its hipError_t

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

// Spinor storage of n (virtual) sites in one precision (owning)
struct Spinor {
  int prec = P64;
  int vh = 0;
  DevBuf data, norm;
  int create(int prec_, int n);
  SpinorPtr ptr() const {
    SpinorPtr s;
    s.p = data.p;
    s.norm = (float *)norm.p;
    s.stride = vh;
    return s;
  }
};

size_t spinor_bytes(int prec, long long n);

// Per-system solver scalars, resident on the device.  Updated only by the
// one-thread epilogues of the finishing kernel (blas.hip).
// done: 0 running, 1 iterated residual below the stop threshold, 2 breakdown,
//       3 reliable update requested, 4 finished (the host's decision)
struct SysState {
  double rho[2], rho_prev[2], alpha[2], beta[2], omega[2];
  double rr;        // |r|^2 of the iterated (sloppy) residual
  double rr_stop;   // stop when rr < rr_stop
  double rr_ref;    // |r|^2 at the last reliable update
  double delta2;    // reliable update when rr < delta2 * rr_ref
  double tmp[4];
  int done;
  int iter;
  int pad[2];
};

struct LaunchCfg {
  int block = 256;
  int remap = 1;
  int group = 0;  // hop scheduling groups (see dslash.hip)
};

// --------------------------------------------------------------------------
// batches: the segment table (qcu_batch.h) on the host and the device
// --------------------------------------------------------------------------
struct Lattice;
struct BatchOp {
  std::vector<SegDesc> segs;
  std::vector<const Lattice *> lats;
  DevBuf dseg;
  SegDesc *dev_segs = nullptr;
  int nsys = 0;
  long long nvsites = 0;
  bool has_clover = false;
  // one segment per lattice with nrhs[k] right-hand sides
  int build(const std::vector<const Lattice *> &lats, const std::vector<int> &nrhs, hipStream_t s);
  int max_blocks(int bsz) const;
};

// --------------------------------------------------------------------------
// kernel launchers (implemented in the .hip files)
// --------------------------------------------------------------------------
// layout.hip
int gauge_ext_to_internal(const void *ext, const Geom &g, int prec, void *a, void *b, hipStream_t s);
// 64-bit content hash of an external gauge field (scratch: gauge_hash_scratch_bytes() of device memory)
int gauge_hash(const void *ext, const Geom &g, void *scratch, unsigned long long *out, hipStream_t s);
size_t gauge_hash_scratch_bytes();
// external one-parity field <-> system (seg, rhs) of a batch vector
int spinor_ext_to_batch(const void *ext, const BatchOp &op, int seg, int rhs, SpinorPtr dst, int prec,
                        hipStream_t s);
int spinor_batch_to_ext(SpinorPtr src, const BatchOp &op, int seg, int rhs, void *ext, int prec, hipStream_t s);

// dslash.hip
enum HopEpi { EPI_PLAIN = 0, EPI_XPAY = 1 };
enum HopRed { RED_NONE = 0, RED_DOT_Y_OUT = 1, RED_OUT_X = 2 };
// out = [A_p^-1] D_{p,1-p} in                                  (EPI_PLAIN)
// out = x - kappa^2 [A_p^-1] D_{p,1-p} in                      (EPI_XPAY)
// RED_DOT_Y_OUT: <y,out> -> alpha update; RED_OUT_X: <out,x>,|out|^2 -> omega
int hop(const BatchOp &op, int prec, SpinorPtr out, SpinorPtr in, int parity, int epi, SpinorPtr x, int red,
        SpinorPtr y, SysState *st, double *partials, const LaunchCfg &cfg, hipStream_t s, bool use_clover = true);

// clover.hip
// packed A (clov, optional) and A^-1 (cinv) of one parity from FP64 internal links
int clover_build(const LinkPtr &U64, const Geom &g, double coeff, int parity, void *clov, void *cinv,
                 int *bad_dev, hipStream_t s);
int clover_convert(const void *c64, const Geom &g, int prec, void *dst, float *norm, hipStream_t s);

// blas.hip (every reduction is per system and finishes with finish())
int batch_blocks(const BatchOp &op, int bsz);
int finish(const BatchOp &op, const double *partials, int bsz, int nr, int kind, SysState *st, double *out,
           hipStream_t s);
int bicg_init(const BatchOp &op, SpinorPtr r, SpinorPtr rt, SpinorPtr p, SpinorPtr v, SpinorPtr x, int prec,
              SpinorPtr b, SysState *st, double *partials, hipStream_t s);
int bicg_p_update(const BatchOp &op, SpinorPtr p, SpinorPtr r, SpinorPtr v, int prec, SysState *st, hipStream_t s);
int bicg_s_update(const BatchOp &op, SpinorPtr s_, SpinorPtr r, SpinorPtr v, int prec, SysState *st, hipStream_t s);
int bicg_xr_update(const BatchOp &op, SpinorPtr x, int px, SpinorPtr r, SpinorPtr rt, SpinorPtr p, SpinorPtr s_,
                   SpinorPtr t, int prec, SysState *st, double *partials, hipStream_t s);
int bicg_xr_update_pc(const BatchOp &op, SpinorPtr x, int px, SpinorPtr r, SpinorPtr rt, SpinorPtr ph, SpinorPtr sh,
                      SpinorPtr s_, SpinorPtr t, int prec, SysState *st, double *partials, hipStream_t s);
int omega_dot(const BatchOp &op, SpinorPtr t, SpinorPtr s_, int prec, SysState *st, double *partials, hipStream_t s);
int reliable_dot(const BatchOp &op, SpinorPtr rt, SpinorPtr r, int prec, SysState *st, double *partials,
                 hipStream_t s);
// y = a*x + b*y  (real a, b; mixed precisions allowed)
int axpby(const BatchOp &op, double a, SpinorPtr x, int px, double b, SpinorPtr y, int py, hipStream_t s);
int spinor_convert(const BatchOp &op, SpinorPtr dst, int pdst, SpinorPtr src, int psrc, hipStream_t s);
int zero_spinor(const BatchOp &op, SpinorPtr x, int px, hipStream_t s);
// r = b - r (FP64); store: |r|^2 of system k into st[k].tmp[0]
int xmy_norm(const BatchOp &op, SpinorPtr b, SpinorPtr r, SysState *st, double *partials, int store, hipStream_t s);
// out_dev[k] = |x_k|^2
int norm2_dev(const BatchOp &op, SpinorPtr x, int px, double *out_dev, double *partials, hipStream_t s);
// out = C (x + a y) in FP64 (a = kappa of each segment when a_is_kappa);
// which: 0 identity, 1 inverse clover, 2 clover of the parity
int cl_axpy(const BatchOp &op, SpinorPtr out, SpinorPtr x, double a, bool a_is_kappa, SpinorPtr y, int which,
            int parity, hipStream_t s);
// out[2k] += |b - (A x - kappa t)|^2, out[2k+1] += |b|^2 of system k, one parity
int residual_parity(const BatchOp &op, SpinorPtr b, SpinorPtr x, SpinorPtr t, int parity, double *partials,
                    double *out, hipStream_t s);

// gen.hip
int gen_gauge(void *ext, const Geom &g, double eps, uint64_t seed, hipStream_t s);
int gen_gauss(void *ext, long n_complex, uint64_t seed, hipStream_t s);

// CG on the normal equations (FP64, per-system device scalars)
int cg_xr(const BatchOp &op, SpinorPtr x, SpinorPtr r, SpinorPtr p, SpinorPtr q, SysState *st, double *partials,
          hipStream_t s);
int cg_p(const BatchOp &op, SpinorPtr p, SpinorPtr z, SysState *st, hipStream_t s);
int g5_clinv(const BatchOp &op, SpinorPtr out, SpinorPtr in, SysState *st, hipStream_t s);
int cg_z(const BatchOp &op, SpinorPtr z, SpinorPtr r, SpinorPtr y, int kind, SysState *st, double *partials,
         hipStream_t s);
int cg_qnorm(const BatchOp &op, SpinorPtr q, SysState *st, double *partials, hipStream_t s);

}  // namespace qcu
