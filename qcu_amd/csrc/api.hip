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
#include <string>

#include "../../include/qcu.h"
#include "qcu_lattice.h"
#include "qcu_solver.h"

namespace qcu {
const char *last_error_msg();
int bench_hop(const BatchOp &op, hipStream_t s, int prec, int clover, int xpay, int iters, const LaunchCfg &cfg,
              double *ms);
}  // namespace qcu

using namespace qcu;

struct QcuBatch_s {
  hipStream_t stream = nullptr;
  BatchOp op;
  Solver solver;
  ~QcuBatch_s() {
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace qcu {

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

}  // namespace qcu

namespace {

int check_method(const QcuSolveParam *p) {
  QCU_CHECK(p->method == QCU_METHOD_BICGSTAB || p->method == QCU_METHOD_CGNR, QCU_ERR_ARG, "unknown solver method");
  QCU_CHECK(p->method != QCU_METHOD_CGNR || (p->precision == QCU_PREC_FP64 && !p->precond), QCU_ERR_ARG,
            "the normal-equation CG runs in FP64 without a preconditioner");
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

// resident lattices of the legacy entry points, one per (device, kind):
// kind 0 Wilson with kappa 0.125, kind 1 clover with kappa*csw = 1
struct Legacy {
  Lattice lat;
  Spinor a, b;
  DevBuf hash_buf;               // scratch of the gauge hash
  const void *gauge = nullptr;   // caller's gauge of the resident links (null: none)
  unsigned long long hash = 0;   // its content hash
};

// The reference converts the links (and builds the clover term) on every
// call.  Here they stay resident while the caller passes the same gauge
// pointer with the same contents: a 64-bit hash of the field, one read of it,
// decides; any change of pointer or contents reloads.
int legacy_load_gauge(Legacy &L, const void *gauge) {
  if (!L.hash_buf.p) QCU_TRY(L.hash_buf.alloc(gauge_hash_scratch_bytes()));
  unsigned long long h = 0;
  QCU_TRY(gauge_hash(gauge, L.lat.g, L.hash_buf.p, &h, L.lat.stream));
  if (L.lat.gauge_loaded && L.gauge == gauge && L.hash == h) return 0;
  L.gauge = nullptr;  // invalid until the reload succeeds
  QCU_TRY(L.lat.load_gauge(gauge, 1));
  L.gauge = gauge;
  L.hash = h;
  return 0;
}
std::mutex g_legacy_mu;
std::map<std::pair<int, int>, std::unique_ptr<Legacy>> g_legacy;

Legacy *legacy_ctx(const QcuParam *param, int kind, int *rc) {
  const int dev = current_device();
  auto key = std::make_pair(dev, kind);
  auto it = g_legacy.find(key);
  if (it != g_legacy.end()) {
    if (std::memcmp(it->second->lat.L, param->lattice_size, sizeof(int) * 4) == 0) return it->second.get();
    g_legacy.erase(it);
  }
  std::unique_ptr<Legacy> l(new Legacy());
  *rc = kind == 0 ? l->lat.init(param->lattice_size, 0.125, 0.0, 1) : l->lat.init(param->lattice_size, 1.0, 1.0, 1);
  if (!*rc) *rc = l->a.create(P64, l->lat.g.vh);
  if (!*rc) *rc = l->b.create(P64, l->lat.g.vh);
  if (*rc) return nullptr;
  Legacy *raw = l.get();
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
  Legacy *L = legacy_ctx(param, clover, &rc);
  if (!L) return rc;
  Lattice &lat = L->lat;
  QCU_TRY(legacy_load_gauge(*L, gauge));
  QCU_TRY(spinor_ext_to_batch(in, lat.self, 0, 0, L->a.ptr(), P64, lat.stream));
  SpinorPtr none{};
  QCU_TRY(hop(lat.self, P64, L->b.ptr(), L->a.ptr(), parity, EPI_PLAIN, none, RED_NONE, none, nullptr, nullptr,
              tuned_cfg(lat.self, P64), lat.stream, clover != 0));
  QCU_TRY(spinor_batch_to_ext(L->b.ptr(), lat.self, 0, 0, out, P64, lat.stream));
  QCU_HIP(hipStreamSynchronize(lat.stream));
  return 0;
}

int lattice_solver(Lattice &lat, Solver **S) {
  if (!lat.solver) {
    std::unique_ptr<Solver> s(new Solver());
    QCU_TRY(s->init(&lat.self, lat.stream));
    lat.solver = std::move(s);
  }
  *S = lat.solver.get();
  return 0;
}

int legacy_bicgstab(void *x, void *b, void *gauge, QcuParam *param) {
  QCU_TRY(check_param(param));
  QCU_TRY(check_dev_ptr(x, "fermion_out"));
  QCU_TRY(check_dev_ptr(b, "fermion_in"));
  QCU_TRY(check_dev_ptr(gauge, "gauge"));
  std::lock_guard<std::mutex> lk(g_legacy_mu);
  int rc = 0;
  Legacy *L = legacy_ctx(param, 0, &rc);
  if (!L) return rc;
  QCU_TRY(legacy_load_gauge(*L, gauge));
  Solver *S = nullptr;
  QCU_TRY(lattice_solver(L->lat, &S));
  QcuSolveParam prm{};
  prm.tol = 1e-9;       // Re<r,r> < 1e-9 (include/define.h:68)
  prm.max_iter = 1000;  // include/define.h:67
  prm.precision = P64;
  prm.chunk = 8;
  prm.use_graph = 1;
  prm.legacy_stop = 1;
  QcuSolveInfo info{};
  void *xs[1] = {x};
  const void *bs[1] = {b};
  return full_solve(*S, xs, bs, prm, &info);
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
  QCU_CHECK(prec >= 0 && prec <= 2 && ((lat.prec_mask >> prec) & 1), QCU_ERR_ARG, "precision not resident");
  QCU_CHECK(!clover || lat.has_clover(), QCU_ERR_STATE, "lattice has no clover term (csw = 0)");
  QCU_TRY(check_dev_ptr(out, "out"));
  QCU_TRY(check_dev_ptr(in, "in"));
  Spinor a, b;
  QCU_TRY(a.create(prec, lat.g.vh));
  QCU_TRY(b.create(prec, lat.g.vh));
  QCU_TRY(spinor_ext_to_batch(in, lat.self, 0, 0, a.ptr(), prec, lat.stream));
  SpinorPtr none{};
  QCU_TRY(hop(lat.self, prec, b.ptr(), a.ptr(), parity, EPI_PLAIN, none, RED_NONE, none, nullptr, nullptr,
              tuned_cfg(lat.self, prec), lat.stream, clover != 0));
  QCU_TRY(spinor_batch_to_ext(b.ptr(), lat.self, 0, 0, out, prec, lat.stream));
  QCU_HIP(hipStreamSynchronize(lat.stream));
  return 0;
}

// ---- the hop without allocation ----------------------------------------------
// Staging of qcuDslashNoAlloc: the input and the output half spinor in the
// internal layout.  A file-scope device array: every GPU the library runs on
// holds its own copy until the process ends.  Host code never passes the
// array's name to a kernel (in host code it is a stand-in, not a device
// address); it uses the address hipGetSymbolAddress returns for the current
// device, fetched once per device.
constexpr size_t kNoAllocScratchBytes = (size_t)32 << 20;
static __device__ __align__(256) char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kNoAllocMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kNoAllocMaxDevices] = {};
}  // namespace

static int noalloc_scratch(int dev, void **ptr) {
  QCU_CHECK(dev >= 0 && dev < kNoAllocMaxDevices, QCU_ERR_STATE, "device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    QCU_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    QCU_CHECK(p != nullptr, QCU_ERR_STATE, "no device address for the scratch array");
    g_scratch_ptr[dev] = p;
  }
  *ptr = g_scratch_ptr[dev];
  return 0;
}

static size_t round256(size_t n) { return (n + 255) & ~(size_t)255; }

// bytes of one staged half spinor (the FP16 one carries its per-site scales)
static size_t noalloc_half_bytes(int prec, int vh) {
  return round256(spinor_bytes(prec, vh)) + (prec == P16 ? round256((size_t)vh * sizeof(float)) : 0);
}

int qcuDslashNoAlloc(QcuLatticeHandle h, void *out, const void *in, int parity, int clover, int prec) {
  clear_error();
  QCU_CHECK(h, QCU_ERR_ARG, "null handle");
  Lattice &lat = h->lat;
  QCU_CHECK(prec >= 0 && prec <= 2, QCU_ERR_ARG, "bad precision");
  const size_t half = noalloc_half_bytes(prec, lat.g.vh);
  QCU_CHECK(2 * half <= kNoAllocScratchBytes, QCU_ERR_DIMS,
            "qcuDslashNoAlloc: the two staged half spinors need " + std::to_string(2 * half) +
                " bytes, above the " + std::to_string(kNoAllocScratchBytes) + "-byte scratch array");
  QCU_CHECK(lat.gauge_loaded, QCU_ERR_STATE, "gauge field not loaded");
  QCU_CHECK(parity == 0 || parity == 1, QCU_ERR_ARG, "parity must be 0 or 1");
  QCU_CHECK((lat.prec_mask >> prec) & 1, QCU_ERR_ARG, "precision not resident");
  QCU_CHECK(!clover || lat.has_clover(), QCU_ERR_STATE, "lattice has no clover term (csw = 0)");
  QCU_TRY(check_dev_ptr(out, "out"));
  QCU_TRY(check_dev_ptr(in, "in"));
  int dev = -1;
  QCU_HIP(hipGetDevice(&dev));
  QCU_CHECK(dev == lat.device, QCU_ERR_STATE, "the current device is not the lattice's");
  void *base = nullptr;
  QCU_TRY(noalloc_scratch(dev, &base));
  SpinorPtr a, b;
  const size_t data = round256(spinor_bytes(prec, lat.g.vh));
  a.p = base;
  a.norm = prec == P16 ? (float *)((char *)base + data) : nullptr;
  a.stride = lat.g.vh;
  b.p = (char *)base + half;
  b.norm = prec == P16 ? (float *)((char *)base + half + data) : nullptr;
  b.stride = lat.g.vh;
  QCU_TRY(spinor_ext_to_batch(in, lat.self, 0, 0, a, prec, lat.stream));
  SpinorPtr none{};
  QCU_TRY(hop(lat.self, prec, b, a, parity, EPI_PLAIN, none, RED_NONE, none, nullptr, nullptr,
              tuned_cfg(lat.self, prec), lat.stream, clover != 0));
  QCU_TRY(spinor_batch_to_ext(b, lat.self, 0, 0, out, prec, lat.stream));
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
  const BatchOp &op = lat.self;
  const int vh = lat.g.vh;
  const size_t half = (size_t)vh * 12 * 16;
  Spinor x[2], t, o;
  for (int p = 0; p < 2; ++p) QCU_TRY(x[p].create(P64, vh));
  QCU_TRY(t.create(P64, vh));
  QCU_TRY(o.create(P64, vh));
  for (int p = 0; p < 2; ++p)
    QCU_TRY(spinor_ext_to_batch((const char *)in + p * half, op, 0, 0, x[p].ptr(), P64, lat.stream));
  SpinorPtr none{};
  for (int p = 0; p < 2; ++p) {
    QCU_TRY(hop(op, P64, t.ptr(), x[1 - p].ptr(), p, EPI_PLAIN, none, RED_NONE, none, nullptr, nullptr,
                tuned_cfg(op, P64), lat.stream, false));
    QCU_TRY(cl_axpy(op, o.ptr(), x[p].ptr(), 0.0, false, none, 2, p, lat.stream));
    QCU_TRY(axpby(op, -lat.kappa, t.ptr(), P64, 1.0, o.ptr(), P64, lat.stream));
    QCU_TRY(spinor_batch_to_ext(o.ptr(), op, 0, 0, (char *)out + p * half, P64, lat.stream));
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
  Solver *S = nullptr;
  QCU_TRY(lattice_solver(lat, &S));
  const void *xs[1] = {x}, *bs[1] = {b};
  return true_residuals(*S, xs, bs, res);
}

int qcuSolve(QcuLatticeHandle h, void *x, const void *b, const QcuSolveParam *p, QcuSolveInfo *info) {
  clear_error();
  QCU_CHECK(h && p, QCU_ERR_ARG, "null argument");
  QCU_CHECK(p->precision >= 0 && p->precision <= 2, QCU_ERR_ARG, "bad precision");
  QCU_CHECK(p->tol > 0, QCU_ERR_ARG, "tolerance must be positive");
  QCU_TRY(check_method(p));
  QCU_TRY(check_dev_ptr(x, "x"));
  QCU_TRY(check_dev_ptr(b, "b"));
  Lattice &lat = h->lat;
  QCU_CHECK(lat.gauge_loaded, QCU_ERR_STATE, "gauge field not loaded");
  QCU_CHECK((lat.prec_mask >> p->precision) & 1, QCU_ERR_STATE, "links of the requested precision are not resident");
  Solver *S = nullptr;
  QCU_TRY(lattice_solver(lat, &S));
  void *xs[1] = {x};
  const void *bs[1] = {b};
  QcuSolveInfo tmp{};
  QcuSolveParam q = *p;
  if (p->precond) {
    Precond *K = nullptr;
    QCU_TRY(net_precond(p->precond, &lat, &K));
    q.precond = K;
  }
  return full_solve(*S, xs, bs, q, info ? info : &tmp);
}

int qcuBenchHop(QcuLatticeHandle h, int prec, int clover, int xpay, int iters, int block, int remap, double *ms) {
  clear_error();
  QCU_CHECK(h && ms, QCU_ERR_ARG, "null argument");
  Lattice &lat = h->lat;
  QCU_CHECK(lat.gauge_loaded, QCU_ERR_STATE, "gauge field not loaded");
  QCU_CHECK(!clover || lat.has_clover(), QCU_ERR_STATE, "lattice has no clover term");
  LaunchCfg cfg = tuned_cfg(lat.self, prec);
  if (block > 0) cfg.block = block;
  if (remap >= 0) cfg.remap = remap;
  return bench_hop(lat.self, lat.stream, prec, clover, xpay, iters, cfg, ms);
}

// ------------------------------------------------------------------ batches
int qcuBatchCreate(QcuBatchHandle *h, const QcuLatticeHandle *lats, const int *nrhs, int nlat) {
  clear_error();
  QCU_CHECK(h && lats && nrhs && nlat > 0, QCU_ERR_ARG, "null argument");
  std::unique_ptr<QcuBatch_s> b(new QcuBatch_s());
  QCU_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
  std::vector<const Lattice *> ls;
  std::vector<int> nr;
  for (int k = 0; k < nlat; ++k) {
    QCU_CHECK(lats[k], QCU_ERR_ARG, "null lattice handle");
    ls.push_back(&lats[k]->lat);
    nr.push_back(nrhs[k]);
  }
  QCU_TRY(b->op.build(ls, nr, b->stream));
  QCU_TRY(b->solver.init(&b->op, b->stream));
  *h = b.release();
  return 0;
}

int qcuBatchDestroy(QcuBatchHandle h) {
  clear_error();
  delete h;
  return 0;
}

int qcuBatchSystems(QcuBatchHandle h, int *nsys) {
  clear_error();
  QCU_CHECK(h && nsys, QCU_ERR_ARG, "null argument");
  *nsys = h->op.nsys;
  return 0;
}

int qcuBatchSolve(QcuBatchHandle h, void *const *x, const void *const *b, const QcuSolveParam *p,
                  QcuSolveInfo *info) {
  clear_error();
  QCU_CHECK(h && x && b && p, QCU_ERR_ARG, "null argument");
  QCU_CHECK(p->precision >= 0 && p->precision <= 2, QCU_ERR_ARG, "bad precision");
  QCU_CHECK(p->tol > 0, QCU_ERR_ARG, "tolerance must be positive");
  QCU_CHECK(p->precond == nullptr, QCU_ERR_ARG, "batched solves take no preconditioner");
  QCU_TRY(check_method(p));
  for (const Lattice *l : h->op.lats)
    QCU_CHECK((l->prec_mask >> p->precision) & 1, QCU_ERR_STATE, "links of the requested precision are not resident");
  for (int k = 0; k < h->op.nsys; ++k) {
    QCU_TRY(check_dev_ptr(x[k], "x"));
    QCU_TRY(check_dev_ptr(b[k], "b"));
  }
  std::vector<QcuSolveInfo> tmp(h->op.nsys);
  return full_solve(h->solver, x, b, *p, info ? info : tmp.data());
}

int qcuBatchBenchHop(QcuBatchHandle h, int prec, int xpay, int iters, double *ms) {
  clear_error();
  QCU_CHECK(h && ms, QCU_ERR_ARG, "null argument");
  return bench_hop(h->op, h->stream, prec, h->op.has_clover, xpay, iters, tuned_cfg(h->op, prec), ms);
}

}  // extern "C"

namespace qcu {
const BatchOp &batch_op(const QcuBatch_s *b) { return b->op; }
hipStream_t batch_stream(const QcuBatch_s *b) { return b->stream; }
}  // namespace qcu
