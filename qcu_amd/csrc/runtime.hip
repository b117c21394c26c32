// runtime.hip -- errors, device buffers, resident lattice state.
#include <cstring>
#include <string>

#include "qcu_lattice.h"
#include "qcu_solver.h"

namespace qcu {

// ---------------------------------------------------------------- errors
static thread_local int g_err_code = 0;
static thread_local std::string g_err_msg;

void set_error(int code, const std::string &msg) {
  g_err_code = code;
  g_err_msg = msg;
}
void clear_error() {
  g_err_code = 0;
  g_err_msg.clear();
}
int last_error_code() { return g_err_code; }
const char *last_error_msg() { return g_err_msg.c_str(); }

int current_device() {
  int d = 0;
  if (hipGetDevice(&d) != hipSuccess) return -1;
  return d;
}

// ---------------------------------------------------------------- buffers
int DevBuf::alloc(size_t n) {
  if (p && bytes >= n) return 0;
  release();
  if (n == 0) return 0;
  hipError_t e = hipMalloc(&p, n);
  if (e != hipSuccess) {
    p = nullptr;
    set_error(QCU_ERR_ALLOC, std::string("hipMalloc(") + std::to_string(n) + "): " + hipGetErrorString(e));
    return QCU_ERR_ALLOC;
  }
  bytes = n;
  return 0;
}

void DevBuf::release() {
  if (p) (void)hipFree(p);
  p = nullptr;
  bytes = 0;
}

size_t spinor_bytes(int prec, int vh) {
  switch (prec) {
    case P64: return (size_t)vh * 12 * 16;
    case P32: return (size_t)vh * 12 * 8;
    default: return (size_t)vh * 12 * 4;
  }
}

int Spinor::create(int prec_, int vh_) {
  prec = prec_;
  vh = vh_;
  QCU_TRY(data.alloc(spinor_bytes(prec, vh)));
  if (prec == P16) QCU_TRY(norm.alloc((size_t)vh * sizeof(float)));
  return 0;
}

// ---------------------------------------------------------------- lattice
int Lattice::init(const int L_[4], double kappa_, double csw_, int tbc) {
  for (int k = 0; k < 4; ++k) {
    QCU_CHECK(L_[k] > 0, QCU_ERR_DIMS, "lattice extents must be positive");
    L[k] = L_[k];
  }
  QCU_CHECK(L[0] % 2 == 0, QCU_ERR_DIMS, "Lx must be even (checkerboarded x)");
  QCU_CHECK(tbc == 1 || tbc == -1, QCU_ERR_ARG, "t_boundary must be +1 or -1");
  long vh = (long)L[0] * L[1] * L[2] * L[3] / 2;
  QCU_CHECK(vh < (1L << 30), QCU_ERR_DIMS, "lattice too large for 32-bit site indices");
  g.lxh = L[0] / 2;
  g.ly = L[1];
  g.lz = L[2];
  g.lt = L[3];
  g.vh = (int)vh;
  g.tbc = tbc;
  kappa = kappa_;
  csw = csw_;
  device = current_device();
  QCU_CHECK(device >= 0, QCU_ERR_STATE, "no HIP device");
  QCU_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  const size_t nb = max_blocks();
  const size_t bytes = nb * 4 * sizeof(double) + kNumStates * sizeof(SysState) + 64 * sizeof(double) + 64;
  QCU_TRY(scratch.alloc(bytes + 1024));
  char *base = (char *)scratch.p;
  partials = (double *)base;
  base += nb * 4 * sizeof(double);
  state = (SysState *)base;
  base += kNumStates * sizeof(SysState);
  dscal = (double *)base;
  base += 64 * sizeof(double);
  bad = (int *)base;
  QCU_HIP(hipMemsetAsync(scratch.p, 0, bytes, stream));
  QCU_HIP(hipHostMalloc((void **)&host_state, kNumStates * sizeof(SysState), hipHostMallocDefault));
  QCU_HIP(hipHostMalloc((void **)&host_scal, 64 * sizeof(double), hipHostMallocDefault));
  QCU_HIP(hipStreamSynchronize(stream));
  return 0;
}

size_t Lattice::max_blocks() const {
  // smallest block size used by any reducing kernel is 64 threads
  return (size_t)(g.vh + 63) / 64 + 1;
}

int Lattice::load_gauge(const void *ext, int mask) {
  QCU_CHECK(ext != nullptr, QCU_ERR_ARG, "gauge pointer is null");
  const size_t vh = g.vh;
  mask |= 1;  // FP64 links are always kept (clover build, reliable updates)
  prec_mask = mask;
  QCU_TRY(U64.alloc(8 * vh * 6 * 16));
  QCU_TRY(gauge_ext_to_internal(ext, g, P64, U64.p, nullptr, stream));
  if (mask & 2) {
    QCU_TRY(U32.alloc(8 * vh * 6 * 8));
    QCU_TRY(gauge_ext_to_internal(ext, g, P32, U32.p, nullptr, stream));
  }
  if (mask & 4) {
    QCU_TRY(U16a.alloc(8 * vh * 16));
    QCU_TRY(U16b.alloc(8 * vh * 8));
    QCU_TRY(gauge_ext_to_internal(ext, g, P16, U16a.p, U16b.p, stream));
  }
  if (has_clover()) {
    QCU_HIP(hipMemsetAsync(bad, 0, sizeof(int), stream));
    for (int p = 0; p < 2; ++p) {
      QCU_TRY(clov64[p].alloc(vh * 72 * 8));
      QCU_TRY(cinv64[p].alloc(vh * 72 * 8));
      QCU_TRY(clover_build(links(P64), g, kappa * csw, p, clov64[p].p, cinv64[p].p, bad, stream));
      if (mask & 2) {
        QCU_TRY(cinv32[p].alloc(vh * 72 * 4));
        QCU_TRY(clover_convert(cinv64[p].p, g, P32, cinv32[p].p, nullptr, stream));
      }
      if (mask & 4) {
        QCU_TRY(cinv16[p].alloc(vh * 72 * 2));
        QCU_TRY(cinv16n[p].alloc(vh * 4));
        QCU_TRY(clover_convert(cinv64[p].p, g, P16, cinv16[p].p, (float *)cinv16n[p].p, stream));
      }
    }
    int hbad = 0;
    QCU_HIP(hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, stream));
    QCU_HIP(hipStreamSynchronize(stream));
    QCU_CHECK(hbad == 0, QCU_ERR_STATE, "singular clover block");
  }
  QCU_HIP(hipStreamSynchronize(stream));
  gauge_loaded = true;
  return 0;
}

LinkPtr Lattice::links(int prec) const {
  LinkPtr l;
  l.vh = g.vh;
  l.b = nullptr;
  if (prec == P64) l.a = U64.p;
  else if (prec == P32) l.a = U32.p;
  else {
    l.a = U16a.p;
    l.b = U16b.p;
  }
  return l;
}

CloverPtr Lattice::clover_inv(int prec, int p) const {
  CloverPtr c;
  c.vh = g.vh;
  c.norm = nullptr;
  c.p = nullptr;
  if (!has_clover()) return c;
  if (prec == P64) c.p = cinv64[p].p;
  else if (prec == P32) c.p = cinv32[p].p;
  else {
    c.p = cinv16[p].p;
    c.norm = (float *)cinv16n[p].p;
  }
  return c;
}

CloverPtr Lattice::clover(int p) const {
  CloverPtr c;
  c.vh = g.vh;
  c.norm = nullptr;
  c.p = has_clover() ? clov64[p].p : nullptr;
  return c;
}

HopOp Lattice::op(int prec) const {
  HopOp o;
  o.U = links(prec);
  o.C[0] = clover_inv(prec, 0);
  o.C[1] = clover_inv(prec, 1);
  o.g = g;
  o.prec = prec;
  return o;
}

Lattice::~Lattice() {
  ws.reset();
  if (host_state) (void)hipHostFree(host_state);
  if (host_scal) (void)hipHostFree(host_scal);
  if (stream) (void)hipStreamDestroy(stream);
}

// |b - M x| / |b| on the full lattice, M = A - kappa D, x and b as internal FP64 parities
int lattice_true_residual(Lattice &lat, SpinorPtr xe, SpinorPtr xo, SpinorPtr be, SpinorPtr bo, double *res,
                          double *bnorm, SpinorPtr scratch) {
  Spinor own;
  SpinorPtr t = scratch;
  if (!t.p) {
    QCU_TRY(own.create(P64, lat.g.vh));
    t = own.ptr();
  }
  HopOp op = lat.op(P64);
  LaunchCfg cfg = tuned_cfg(lat, P64);
  QCU_HIP(hipMemsetAsync(lat.dscal, 0, 2 * sizeof(double), lat.stream));
  SpinorPtr none{};
  // even: t = D_eo x_o
  QCU_TRY(hop(op, t, xo, 0, EPI_PLAIN, none, 0, RED_NONE, none, nullptr, nullptr, cfg, lat.stream,
              false));
  QCU_TRY(residual_parity(be, xe, t, lat.clover(0), lat.kappa, lat.g.vh, lat.partials, lat.dscal,
                          lat.stream));
  QCU_TRY(hop(op, t, xe, 1, EPI_PLAIN, none, 0, RED_NONE, none, nullptr, nullptr, cfg, lat.stream,
              false));
  QCU_TRY(residual_parity(bo, xo, t, lat.clover(1), lat.kappa, lat.g.vh, lat.partials, lat.dscal,
                          lat.stream));
  QCU_HIP(hipMemcpyAsync(lat.host_scal, lat.dscal, 2 * sizeof(double), hipMemcpyDeviceToHost, lat.stream));
  QCU_HIP(hipStreamSynchronize(lat.stream));
  const double num = lat.host_scal[0], den = lat.host_scal[1];
  *res = den > 0 ? sqrt(num / den) : sqrt(num);
  if (bnorm) *bnorm = sqrt(den);
  return 0;
}

}  // namespace qcu
