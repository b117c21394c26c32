// runtime.hip -- errors, device buffers, resident lattice state, batches.
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

size_t spinor_bytes(int prec, long long n) { return (size_t)n * words_per_site(prec) * 16; }

int Spinor::create(int prec_, int n) {
  prec = prec_;
  vh = n;
  QCU_TRY(data.alloc(spinor_bytes(prec, n)));
  if (prec == P16) QCU_TRY(norm.alloc((size_t)n * sizeof(float)));
  return 0;
}

// ---------------------------------------------------------------- batches
int BatchOp::build(const std::vector<const Lattice *> &ls, const std::vector<int> &nrhs, hipStream_t s) {
  QCU_CHECK(!ls.empty() && ls.size() == nrhs.size(), QCU_ERR_ARG, "batch: one rhs count per lattice");
  segs.clear();
  lats = ls;
  nsys = 0;
  nvsites = 0;
  has_clover = false;
  for (size_t k = 0; k < ls.size(); ++k) {
    QCU_CHECK(ls[k] && ls[k]->gauge_loaded, QCU_ERR_STATE, "batch: gauge field not loaded");
    QCU_CHECK(nrhs[k] >= 1 && nrhs[k] <= kMaxRhs, QCU_ERR_ARG, "batch: 1 to 16 right-hand sides per lattice");
    // kernels address a segment's fields with 32-bit per-lane byte offsets of
    // 16 bytes per virtual site
    QCU_CHECK((long long)ls[k]->g.vh * nrhs[k] < kMaxVirtualSites, QCU_ERR_DIMS,
              "batch: segment too large (Vh * nrhs must be below 2^28)");
    SegDesc d = ls[k]->seg_desc(nrhs[k]);
    d.sys0 = nsys;
    d.voff = nvsites;
    nsys += nrhs[k];
    nvsites += d.nthr;
    has_clover = has_clover || ls[k]->has_clover();
    segs.push_back(d);
  }
  QCU_CHECK(!has_clover || [&] {
    for (const Lattice *l : ls)
      if (!l->has_clover()) return false;
    return true;
  }(), QCU_ERR_ARG, "batch: mixing Wilson and clover lattices is not supported");
  // block-partial slots per right-hand side: the block count at the smallest
  // block size a reducing kernel uses (64), the same stride for every segment
  const int pcap = max_blocks(64) + 1;
  for (SegDesc &d : segs) d.pcap = pcap;
  QCU_TRY(dseg.alloc(segs.size() * sizeof(SegDesc)));
  dev_segs = (SegDesc *)dseg.p;
  QCU_HIP(hipMemcpyAsync(dev_segs, segs.data(), segs.size() * sizeof(SegDesc), hipMemcpyHostToDevice, s));
  QCU_HIP(hipStreamSynchronize(s));
  return 0;
}

int BatchOp::max_blocks(int bsz) const {
  int nb = 0;
  for (const SegDesc &d : segs) nb += (d.nthr + bsz - 1) / bsz;
  return nb;
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
  // kernels address fields with 32-bit per-lane byte offsets, 16 bytes per site
  QCU_CHECK(vh < kMaxVirtualSites, QCU_ERR_DIMS, "lattice too large: Vh must be below 2^28");
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
  QCU_TRY(flag.alloc(256));
  return 0;
}

int Lattice::load_gauge(const void *ext, int mask) {
  QCU_CHECK(ext != nullptr, QCU_ERR_ARG, "gauge pointer is null");
  const size_t vh = g.vh;
  mask |= 1;  // FP64 links are always kept (clover build, reliable updates)
  // until every step below succeeds the lattice holds no usable gauge field:
  // a failed (re)load leaves it unloaded rather than half rebuilt
  gauge_loaded = false;
  prec_mask = 0;
  solver.reset();  // its batch table and graphs point at the previous state
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
    int *bad = (int *)flag.p;
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
  // the lattice's own one-system batch table is built from a loaded lattice
  prec_mask = mask;
  gauge_loaded = true;
  if (const int rc = self.build({this}, {1}, stream)) {
    gauge_loaded = false;
    prec_mask = 0;
    return rc;
  }
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

SegDesc Lattice::seg_desc(int nrhs) const {
  SegDesc d;
  std::memset(&d, 0, sizeof(d));
  d.g = g;
  d.nrhs = nrhs;
  d.nthr = g.vh * nrhs;
  d.kappa = kappa;
  for (int p = 0; p < 3; ++p) {
    d.U[p] = links(p);
    d.Cinv[p][0] = clover_inv(p, 0);
    d.Cinv[p][1] = clover_inv(p, 1);
  }
  d.A[0] = clover(0);
  d.A[1] = clover(1);
  return d;
}

Lattice::~Lattice() {
  solver.reset();
  if (stream) (void)hipStreamDestroy(stream);
}

}  // namespace qcu
