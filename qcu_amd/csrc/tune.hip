// tune.hip -- launch-shape autotuner and hop timing.
//
// For each (device, batch shape, precision) the tuner times the hop kernel
// with HIP events over a grid of block sizes, XCD-aware block orders and
// instruction-scheduling groupings, and keeps the fastest.  Results persist in
// a plain text cache file ("key block remap group" per line) so that later
// processes skip the search.
#include <atomic>
#include <cstdio>
#include <map>
#include <mutex>
#include <string>

#include "../../include/qcu.h"
#include "qcu_lattice.h"
#include "qcu_solver.h"

namespace qcu {

namespace {
std::mutex g_tune_mu;
std::map<std::string, LaunchCfg> g_tuned;
std::atomic<unsigned long long> g_tune_gen{1};
}  // namespace

// changes whenever a tuned launch shape is set, so that captured solver
// graphs (which hold the launch shapes of their capture) are re-captured
unsigned long long tune_generation() { return g_tune_gen.load(); }

static void set_tuned(const std::string &key, const LaunchCfg &c) {
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tuned[key] = c;
  }
  ++g_tune_gen;
}

static int valid_cfg(int block, int remap, int group) {
  QCU_CHECK((block == 64 || block == 128 || block == 256) && (remap == 0 || remap == 1) && group >= 0 && group <= 2,
            QCU_ERR_ARG, "launch shape: block 64, 128 or 256; remap 0 or 1; group 0, 1 or 2");
  return 0;
}

std::string tune_key(const BatchOp &op, int prec) {
  std::string k;
  hipDeviceProp_t prop;
  int dev = current_device();
  if (dev >= 0 && hipGetDeviceProperties(&prop, dev) == hipSuccess) k = prop.gcnArchName;
  char buf[96];
  for (const SegDesc &d : op.segs) {
    snprintf(buf, sizeof(buf), ":%dx%dx%dx%dr%d", 2 * d.g.lxh, d.g.ly, d.g.lz, d.g.lt, d.nrhs);
    k += buf;
  }
  snprintf(buf, sizeof(buf), ":p%d:c%d", prec, op.has_clover ? 1 : 0);
  return k + buf;
}

LaunchCfg tuned_cfg(const BatchOp &op, int prec) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto it = g_tuned.find(tune_key(op, prec));
  if (it != g_tuned.end()) return it->second;
  return LaunchCfg();
}

// average milliseconds per launch of the hop kernel on resident fields
int bench_hop(const BatchOp &op, hipStream_t s, int prec, int clover, int xpay, int iters, const LaunchCfg &cfg,
              double *ms) {
  QCU_CHECK(prec >= 0 && prec <= 2 && iters > 0, QCU_ERR_ARG, "bad precision or iteration count");
  for (const Lattice *l : op.lats)
    QCU_CHECK((l->prec_mask >> prec) & 1, QCU_ERR_STATE, "links of the requested precision are not resident");
  const int n = (int)op.nvsites;
  Spinor a, b, c, g64;
  QCU_TRY(a.create(prec, n));
  QCU_TRY(b.create(prec, n));
  QCU_TRY(c.create(prec, n));
  QCU_TRY(g64.create(P64, n));
  QCU_TRY(gen_gauss(g64.data.p, (long)n * 12, 77, s));
  QCU_TRY(spinor_convert(op, a.ptr(), prec, g64.ptr(), P64, s));
  QCU_TRY(spinor_convert(op, c.ptr(), prec, g64.ptr(), P64, s));
  SpinorPtr none{};
  auto launch = [&](int k) {
    return hop(op, prec, (k & 1) ? a.ptr() : b.ptr(), (k & 1) ? b.ptr() : a.ptr(), k & 1, xpay ? EPI_XPAY : EPI_PLAIN,
               c.ptr(), RED_NONE, none, nullptr, nullptr, cfg, s, clover != 0);
  };
  for (int k = 0; k < 3; ++k) QCU_TRY(launch(k));
  hipEvent_t e0, e1;
  QCU_HIP(hipEventCreate(&e0));
  QCU_HIP(hipEventCreate(&e1));
  QCU_HIP(hipEventRecord(e0, s));
  for (int k = 0; k < iters; ++k) QCU_TRY(launch(k));
  QCU_HIP(hipEventRecord(e1, s));
  QCU_HIP(hipEventSynchronize(e1));
  float t = 0;
  QCU_HIP(hipEventElapsedTime(&t, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *ms = (double)t / iters;
  return 0;
}

static int tune_op(const BatchOp &op, hipStream_t s, int prec, double *best_ms, LaunchCfg *out) {
  const int clover = op.has_clover ? 1 : 0;
  const int blocks[] = {64, 128, 256};
  LaunchCfg best;
  double bms = 1e30;
  for (int grp = 0; grp < 3; ++grp)
    for (int b : blocks)
      for (int r = 0; r < 2; ++r) {
        LaunchCfg c;
        c.block = b;
        c.remap = r;
        c.group = grp;
        double ms = 0;
        QCU_TRY(bench_hop(op, s, prec, clover, 1, 20, c, &ms));
        if (ms < bms) {
          bms = ms;
          best = c;
        }
      }
  set_tuned(tune_key(op, prec), best);
  if (best_ms) *best_ms = bms;
  if (out) *out = best;
  return 0;
}

}  // namespace qcu

using namespace qcu;

struct QcuBatch_s;
namespace qcu {
const BatchOp &batch_op(const QcuBatch_s *b);
hipStream_t batch_stream(const QcuBatch_s *b);
}  // namespace qcu

extern "C" {

int qcuTune(QcuLatticeHandle h, int prec, double *best_ms) {
  clear_error();
  QCU_CHECK(h, QCU_ERR_ARG, "null handle");
  QCU_CHECK(prec >= 0 && prec <= 2, QCU_ERR_ARG, "bad precision");
  QCU_CHECK(h->lat.gauge_loaded, QCU_ERR_STATE, "gauge field not loaded");
  return tune_op(h->lat.self, h->lat.stream, prec, best_ms, nullptr);
}

int qcuTuneSet(QcuLatticeHandle h, int prec, int block, int remap, int group) {
  clear_error();
  QCU_CHECK(h, QCU_ERR_ARG, "null handle");
  QCU_CHECK(h->lat.gauge_loaded, QCU_ERR_STATE, "gauge field not loaded");
  QCU_CHECK(prec >= 0 && prec <= 2, QCU_ERR_ARG, "bad precision");
  QCU_TRY(valid_cfg(block, remap, group));
  LaunchCfg c;
  c.block = block;
  c.remap = remap;
  c.group = group;
  set_tuned(tune_key(h->lat.self, prec), c);
  return 0;
}

int qcuBatchTuneSet(QcuBatchHandle h, int prec, int block, int remap, int group) {
  clear_error();
  QCU_CHECK(h, QCU_ERR_ARG, "null handle");
  QCU_CHECK(prec >= 0 && prec <= 2, QCU_ERR_ARG, "bad precision");
  QCU_TRY(valid_cfg(block, remap, group));
  LaunchCfg c;
  c.block = block;
  c.remap = remap;
  c.group = group;
  set_tuned(tune_key(batch_op(h), prec), c);
  return 0;
}

int qcuTuneGet(QcuLatticeHandle h, int prec, int *block, int *remap, int *group) {
  clear_error();
  QCU_CHECK(h && block && remap && group, QCU_ERR_ARG, "null argument");
  LaunchCfg c = tuned_cfg(h->lat.self, prec);
  *block = c.block;
  *remap = c.remap;
  *group = c.group;
  return 0;
}

int qcuBatchTune(QcuBatchHandle h, int prec, double *best_ms) {
  clear_error();
  QCU_CHECK(h, QCU_ERR_ARG, "null handle");
  QCU_CHECK(prec >= 0 && prec <= 2, QCU_ERR_ARG, "bad precision");
  return tune_op(batch_op(h), batch_stream(h), prec, best_ms, nullptr);
}

int qcuTuneSave(const char *path) {
  QCU_CHECK(path, QCU_ERR_ARG, "null path");
  FILE *f = fopen(path, "w");
  QCU_CHECK(f, QCU_ERR_ARG, std::string("cannot write ") + path);
  std::lock_guard<std::mutex> lk(g_tune_mu);
  for (auto &kv : g_tuned)
    fprintf(f, "%s %d %d %d\n", kv.first.c_str(), kv.second.block, kv.second.remap, kv.second.group);
  fclose(f);
  return 0;
}

int qcuTuneLoad(const char *path) {
  QCU_CHECK(path, QCU_ERR_ARG, "null path");
  FILE *f = fopen(path, "r");
  if (!f) return 0;  // no cache yet
  char key[512];
  int b, r, grp;
  std::lock_guard<std::mutex> lk(g_tune_mu);
  while (fscanf(f, "%511s %d %d %d", key, &b, &r, &grp) == 4) {
    if (b >= 64 && b <= 256 && b % 64 == 0 && (r == 0 || r == 1) && grp >= 0 && grp <= 2) {
      LaunchCfg c;
      c.block = b;
      c.remap = r;
      c.group = grp;
      g_tuned[key] = c;
    }
  }
  fclose(f);
  ++g_tune_gen;
  return 0;
}

}  // extern "C"
