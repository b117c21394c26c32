// api.hip -- the C ABI of libqcu.so (declared in include/qcu.h).
//
// The six reference entry points (include/qcu.h:21-30 of the reference) keep
// their signatures and semantics; instead of building an MPI/NCCL context,
// streams and cuBLAS handles on every call (include/lattice_set.h:73-181) they
// reuse one resident lattice per device and shape, and they never transpose the
// caller's buffers in place.
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>

#include "../../include/qcu.h"
#include "qcu_lattice.h"
#include "qcu_solver.h"

namespace qcu {
const char *last_error_msg();
int tune_lookup(const Lattice &lat, int prec, LaunchCfg *cfg);
}  // namespace qcu

using namespace qcu;

namespace {

std::mutex g_legacy_mu;
// resident lattices of the legacy entry points, one per (device, kind)
std::map<std::pair<int, int>, std::unique_ptr<Lattice>> g_legacy;

int check_dev_ptr(const void *p, const char *what) {
  if (!p) {
    set_error(QCU_ERR_ARG, std::string(what) + " is null");
    return QCU_ERR_ARG;
  }
  hipPointerAttribute_t attr;
  hipError_t e = hipPointerGetAttributes(&attr, p);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    set_error(QCU_ERR_ARG, std::string(what) + " is not a HIP device pointer");
    return QCU_ERR_ARG;
  }
  if (attr.type != hipMemoryTypeDevice && attr.type != hipMemoryTypeManaged) {
    set_error(QCU_ERR_ARG, std::string(what) + " is not device memory");
    return QCU_ERR_ARG;
  }
  return 0;
}

int check_param(const QcuParam *param) {
  QCU_CHECK(param != nullptr, QCU_ERR_ARG, "param is null");
  for (int k = 0; k < 4; ++k) QCU_CHECK(param->lattice_size[k] > 0, QCU_ERR_DIMS, "lattice_size must be positive");
  QCU_CHECK(param->lattice_size[0] % 2 == 0, QCU_ERR_DIMS, "lattice_size[0] must be even");
  return 0;
}

int check_grid(const QcuParam *grid) {
  QCU_CHECK(grid != nullptr, QCU_ERR_ARG, "grid is null");
  for (int k = 0; k < 4; ++k)
    QCU_CHECK(grid->lattice_size[k] == 1, QCU_ERR_GRID,
              "this build runs on one GPU: grid must be {1,1,1,1} (periodic wrap replaces the halo exchange)");
  return 0;
}

// kind 0: Wilson (kappa 0.125), kind 1: clover with kappa*csw = 1
Lattice *legacy_lattice(const QcuParam *param, int kind, int *rc) {
  int dev = current_device();
  auto key = std::make_pair(dev, kind);
  auto it = g_legacy.find(key);
  if (it != g_legacy.end()) {
    Lattice *l = it->second.get();
    if (std::memcmp(l->L, param->lattice_size, sizeof(l->L)) == 0) return l;
    g_legacy.erase(it);
  }
  std::unique_ptr<Lattice> l(new Lattice());
  *rc = kind == 0 ? l->init(param->lattice_size, 0.125, 0.0, 1) : l->init(param->lattice_size, 1.0, 1.0, 1);
  if (*rc) return nullptr;
  Lattice *raw = l.get();
  g_legacy[key] = std::move(l);
  return raw;
}

int legacy_dslash(void *out, void *in, void *gauge, QcuParam *param, int parity, int clover) {
  QCU_TRY(check_param(param));
  QCU_CHECK(parity == 0 || parity == 1, QCU_ERR_ARG, "parity must be 0 or 1");
  QCU_TRY(check_dev_ptr(out, "fermion_out"));
  QCU_TRY(check_dev_ptr(in, "fermion_in"));
  QCU_TRY(check_dev_ptr(gauge, "gauge"));
  std::lock_guard<std::mutex> lk(g_legacy_mu);
  int rc = 0;
  Lattice *lat = legacy_lattice(param, clover, &rc);
  if (!lat) return rc;
  QCU_TRY(lat->load_gauge(gauge, 1));
  Spinor a, b;
  QCU_TRY(a.create(P64, lat->g.vh));
  QCU_TRY(b.create(P64, lat->g.vh));
  QCU_TRY(spinor_ext_to_internal(in, lat->g, a.ptr(), P64, lat->stream));
  SpinorPtr none{};
  LaunchCfg cfg = tuned_cfg(*lat, P64);
  QCU_TRY(hop(lat->op(P64), b.ptr(), a.ptr(), parity, EPI_PLAIN, none, 0, RED_NONE, none, nullptr, nullptr,
              cfg, lat->stream, clover != 0));
  QCU_TRY(spinor_internal_to_ext(b.ptr(), lat->g, out, P64, lat->stream));
  QCU_HIP(hipStreamSynchronize(lat->stream));
  return 0;
}

int legacy_bicgstab(void *x, void *b, void *gauge, QcuParam *param) {
  QCU_TRY(check_param(param));
  QCU_TRY(check_dev_ptr(x, "fermion_out"));
  QCU_TRY(check_dev_ptr(b, "fermion_in"));
  QCU_TRY(check_dev_ptr(gauge, "gauge"));
  std::lock_guard<std::mutex> lk(g_legacy_mu);
  int rc = 0;
  Lattice *lat = legacy_lattice(param, 0, &rc);
  if (!lat) return rc;
  QCU_TRY(lat->load_gauge(gauge, 1));
  QcuSolveParam prm{};
  prm.tol = 1e-9;         // Re<r,r> < 1e-9 (include/define.h:68)
  prm.max_iter = 1000;    // include/define.h:67
  prm.precision = P64;
  prm.chunk = 8;
  prm.use_graph = 1;
  prm.legacy_stop = 1;
  QcuSolveInfo info{};
  QCU_TRY(full_solve(*lat, x, b, prm, &info));
  return 0;
}

}  // namespace

extern "C" {

// ------------------------------------------------------------------ errors
int qcuGetLastError(void) { return last_error_code(); }
const char *qcuGetLastErrorString(void) { return last_error_msg(); }
void qcuClearError(void) { clear_error(); }

// ------------------------------------------------------------------ reference entry points
void dslashQcu(void *fermion_out, void *fermion_in, void *gauge, QcuParam *param, int parity) {
  clear_error();
  (void)legacy_dslash(fermion_out, fermion_in, gauge, param, parity, 0);
}

void dslashCloverQcu(void *fermion_out, void *fermion_in, void *gauge, QcuParam *param, int parity) {
  clear_error();
  (void)legacy_dslash(fermion_out, fermion_in, gauge, param, parity, 1);
}

void ncclDslashQcu(void *fermion_out, void *fermion_in, void *gauge, QcuParam *param, int parity, QcuParam *grid) {
  clear_error();
  if (check_grid(grid)) return;
  (void)legacy_dslash(fermion_out, fermion_in, gauge, param, parity, 0);
}

void mpiDslashQcu(void *fermion_out, void *fermion_in, void *gauge, QcuParam *param, int parity, QcuParam *grid) {
  clear_error();
  if (check_grid(grid)) return;
  (void)legacy_dslash(fermion_out, fermion_in, gauge, param, parity, 0);
}

void mpiBistabCgQcu(void *gauge, QcuParam *param, QcuParam *grid) {
  // an empty function in the reference (src/draft.cu:30-31)
  clear_error();
  (void)gauge;
  (void)param;
  (void)grid;
}

void ncclBistabCgQcu(void *fermion_out, void *fermion_in, void *gauge, QcuParam *param, QcuParam *grid) {
  clear_error();
  if (check_grid(grid)) return;
  (void)legacy_bicgstab(fermion_out, fermion_in, gauge, param);
}

// ------------------------------------------------------------------ handle API
int qcuLatticeCreate(QcuLatticeHandle *h, const int L[4], double kappa, double csw, int t_boundary) {
  clear_error();
  QCU_CHECK(h && L, QCU_ERR_ARG, "null argument");
  std::unique_ptr<QcuLattice_s> o(new QcuLattice_s());
  QCU_TRY(o->lat.init(L, kappa, csw, t_boundary));
  *h = o.release();
  return 0;
}

int qcuLatticeDestroy(QcuLatticeHandle h) {
  clear_error();
  delete h;
  return 0;
}

int qcuLatticeLoadGauge(QcuLatticeHandle h, const void *gauge) {
  clear_error();
  QCU_CHECK(h, QCU_ERR_ARG, "null handle");
  QCU_TRY(check_dev_ptr(gauge, "gauge"));
  return h->lat.load_gauge(gauge, 7);
}

int qcuGaugeRandom(void *gauge, const int L[4], double eps, uint64_t seed) {
  clear_error();
  QCU_TRY(check_dev_ptr(gauge, "gauge"));
  QCU_CHECK(L && L[0] > 0 && L[1] > 0 && L[2] > 0 && L[3] > 0 && L[0] % 2 == 0, QCU_ERR_DIMS, "bad lattice size");
  Geom g{};
  g.lxh = L[0] / 2;
  g.ly = L[1];
  g.lz = L[2];
  g.lt = L[3];
  g.vh = L[0] / 2 * L[1] * L[2] * L[3];
  QCU_TRY(gen_gauge(gauge, g, eps, seed, nullptr));
  QCU_HIP(hipStreamSynchronize(nullptr));
  return 0;
}

int qcuGaussian(void *field, long n, uint64_t seed) {
  clear_error();
  QCU_TRY(check_dev_ptr(field, "field"));
  QCU_CHECK(n >= 0, QCU_ERR_ARG, "negative length");
  if (n == 0) return 0;
  QCU_TRY(gen_gauss(field, n, seed, nullptr));
  QCU_HIP(hipStreamSynchronize(nullptr));
  return 0;
}

int qcuDslash(QcuLatticeHandle h, void *out, const void *in, int parity, int clover, int prec) {
  clear_error();
  QCU_CHECK(h, QCU_ERR_ARG, "null handle");
  Lattice &lat = h->lat;
  QCU_CHECK(lat.gauge_loaded, QCU_ERR_STATE, "gauge field not loaded");
  QCU_CHECK(parity == 0 || parity == 1, QCU_ERR_ARG, "parity must be 0 or 1");
  QCU_CHECK(prec >= 0 && prec <= 2, QCU_ERR_ARG, "bad precision");
  QCU_CHECK(!clover || lat.has_clover(), QCU_ERR_STATE, "lattice has no clover term (csw = 0)");
  QCU_TRY(check_dev_ptr(out, "out"));
  QCU_TRY(check_dev_ptr(in, "in"));
  Spinor a, b;
  QCU_TRY(a.create(prec, lat.g.vh));
  QCU_TRY(b.create(prec, lat.g.vh));
  QCU_TRY(spinor_ext_to_internal(in, lat.g, a.ptr(), prec, lat.stream));
  SpinorPtr none{};
  QCU_TRY(hop(lat.op(prec), b.ptr(), a.ptr(), parity, EPI_PLAIN, none, 0, RED_NONE, none, nullptr, nullptr,
              tuned_cfg(lat, prec), lat.stream, clover != 0));
  QCU_TRY(spinor_internal_to_ext(b.ptr(), lat.g, out, prec, lat.stream));
  QCU_HIP(hipStreamSynchronize(lat.stream));
  return 0;
}

int qcuMatFull(QcuLatticeHandle h, void *out, const void *in) {
  clear_error();
  QCU_CHECK(h, QCU_ERR_ARG, "null handle");
  Lattice &lat = h->lat;
  QCU_CHECK(lat.gauge_loaded, QCU_ERR_STATE, "gauge field not loaded");
  QCU_TRY(check_dev_ptr(out, "out"));
  QCU_TRY(check_dev_ptr(in, "in"));
  const int vh = lat.g.vh;
  const size_t half = (size_t)vh * 12 * 16;
  Spinor x[2], t, o;
  for (int p = 0; p < 2; ++p) QCU_TRY(x[p].create(P64, vh));
  QCU_TRY(t.create(P64, vh));
  QCU_TRY(o.create(P64, vh));
  for (int p = 0; p < 2; ++p)
    QCU_TRY(spinor_ext_to_internal((const char *)in + p * half, lat.g, x[p].ptr(), P64, lat.stream));
  SpinorPtr none{};
  for (int p = 0; p < 2; ++p) {
    QCU_TRY(hop(lat.op(P64), t.ptr(), x[1 - p].ptr(), p, EPI_PLAIN, none, 0, RED_NONE, none, nullptr, nullptr,
                tuned_cfg(lat, P64), lat.stream, false));
    QCU_TRY(cl_axpy(o.ptr(), x[p].ptr(), 0.0, none, lat.clover(p), vh, lat.stream));
    QCU_TRY(axpby(-lat.kappa, t.ptr(), P64, 1.0, o.ptr(), P64, vh, lat.stream));
    QCU_TRY(spinor_internal_to_ext(o.ptr(), lat.g, (char *)out + p * half, P64, lat.stream));
  }
  QCU_HIP(hipStreamSynchronize(lat.stream));
  return 0;
}

int qcuTrueResidual(QcuLatticeHandle h, const void *x, const void *b, double *res) {
  clear_error();
  QCU_CHECK(h && res, QCU_ERR_ARG, "null argument");
  Lattice &lat = h->lat;
  QCU_CHECK(lat.gauge_loaded, QCU_ERR_STATE, "gauge field not loaded");
  QCU_TRY(check_dev_ptr(x, "x"));
  QCU_TRY(check_dev_ptr(b, "b"));
  const int vh = lat.g.vh;
  const size_t half = (size_t)vh * 12 * 16;
  Spinor xs[2], bs[2];
  for (int p = 0; p < 2; ++p) {
    QCU_TRY(xs[p].create(P64, vh));
    QCU_TRY(bs[p].create(P64, vh));
    QCU_TRY(spinor_ext_to_internal((const char *)x + p * half, lat.g, xs[p].ptr(), P64, lat.stream));
    QCU_TRY(spinor_ext_to_internal((const char *)b + p * half, lat.g, bs[p].ptr(), P64, lat.stream));
  }
  return lattice_true_residual(lat, xs[0].ptr(), xs[1].ptr(), bs[0].ptr(), bs[1].ptr(), res, nullptr, SpinorPtr{});
}

int qcuSolve(QcuLatticeHandle h, void *x, const void *b, const QcuSolveParam *p, QcuSolveInfo *info) {
  clear_error();
  QCU_CHECK(h && p, QCU_ERR_ARG, "null argument");
  QCU_CHECK(p->precision >= 0 && p->precision <= 2, QCU_ERR_ARG, "bad precision");
  QCU_CHECK(p->tol > 0, QCU_ERR_ARG, "tolerance must be positive");
  QCU_TRY(check_dev_ptr(x, "x"));
  QCU_TRY(check_dev_ptr(b, "b"));
  Lattice &lat = h->lat;
  QCU_CHECK(lat.gauge_loaded, QCU_ERR_STATE, "gauge field not loaded");
  QCU_CHECK(p->precision == P64 || (lat.prec_mask >> p->precision) & 1, QCU_ERR_STATE,
            "links of the requested precision are not resident");
  return full_solve(lat, x, b, *p, info);
}

int qcuBenchHop(QcuLatticeHandle h, int prec, int clover, int xpay, int iters, int block, int remap, double *ms) {
  clear_error();
  QCU_CHECK(h && ms, QCU_ERR_ARG, "null argument");
  QCU_CHECK(prec >= 0 && prec <= 2 && iters > 0, QCU_ERR_ARG, "bad precision or iteration count");
  Lattice &lat = h->lat;
  QCU_CHECK(lat.gauge_loaded, QCU_ERR_STATE, "gauge field not loaded");
  QCU_CHECK(!clover || lat.has_clover(), QCU_ERR_STATE, "lattice has no clover term");
  const int vh = lat.g.vh;
  Spinor a, b, c;
  QCU_TRY(a.create(prec, vh));
  QCU_TRY(b.create(prec, vh));
  QCU_TRY(c.create(prec, vh));
  Spinor g64;
  QCU_TRY(g64.create(P64, vh));
  QCU_TRY(gen_gauss(g64.data.p, (long)vh * 12, 77, lat.stream));
  QCU_TRY(spinor_convert(a.ptr(), prec, g64.ptr(), P64, vh, lat.stream));
  QCU_TRY(spinor_convert(c.ptr(), prec, g64.ptr(), P64, vh, lat.stream));
  LaunchCfg cfg = tuned_cfg(lat, prec);
  if (block > 0) cfg.block = block;
  if (remap >= 0) cfg.remap = remap;
  SpinorPtr none{};
  const double k2 = lat.kappa * lat.kappa;
  auto launch = [&](int k) {
    return hop(lat.op(prec), (k & 1) ? a.ptr() : b.ptr(), (k & 1) ? b.ptr() : a.ptr(), k & 1,
               xpay ? EPI_XPAY : EPI_PLAIN, c.ptr(), k2, RED_NONE, none, nullptr, nullptr, cfg, lat.stream,
               clover != 0);
  };
  for (int k = 0; k < 3; ++k) QCU_TRY(launch(k));
  hipEvent_t e0, e1;
  QCU_HIP(hipEventCreate(&e0));
  QCU_HIP(hipEventCreate(&e1));
  QCU_HIP(hipEventRecord(e0, lat.stream));
  for (int k = 0; k < iters; ++k) QCU_TRY(launch(k));
  QCU_HIP(hipEventRecord(e1, lat.stream));
  QCU_HIP(hipEventSynchronize(e1));
  float t = 0;
  QCU_HIP(hipEventElapsedTime(&t, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *ms = (double)t / iters;
  return 0;
}

}  // extern "C"
