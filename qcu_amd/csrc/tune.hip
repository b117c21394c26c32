// tune.hip -- launch-shape autotuner for the hop kernels.
//
// For each (device, lattice shape, precision, clover) the tuner times the hop
// kernel with HIP events over a small grid of block sizes and XCD remap
// settings and keeps the fastest.  Results persist in a plain text cache file
// (one "key block remap ms" line per entry) so that later processes skip the
// search.
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

std::string key_of(const Lattice &lat, int prec, int clover) {
  char buf[160];
  hipDeviceProp_t prop;
  const char *arch = "gfx";
  if (hipGetDeviceProperties(&prop, lat.device) == hipSuccess) arch = prop.gcnArchName;
  snprintf(buf, sizeof(buf), "%s:%dx%dx%dx%d:p%d:c%d", arch, lat.L[0], lat.L[1], lat.L[2], lat.L[3], prec,
           clover);
  return std::string(buf);
}
}  // namespace

LaunchCfg tuned_cfg(const Lattice &lat, int prec) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto it = g_tuned.find(key_of(lat, prec, lat.has_clover() ? 1 : 0));
  if (it != g_tuned.end()) return it->second;
  return LaunchCfg();
}

}  // namespace qcu

using namespace qcu;

extern "C" {

int qcuBenchHop(QcuLatticeHandle h, int prec, int clover, int xpay, int iters, int block, int remap, double *ms);

// Sweeps launch shapes of the hop kernel (the clover epilogue when the lattice
// has one) for one precision and records the fastest.  best_ms may be null.
int qcuTune(QcuLatticeHandle h, int prec, double *best_ms) {
  QCU_CHECK(h, QCU_ERR_ARG, "null handle");
  Lattice &lat = h->lat;
  const int clover = lat.has_clover() ? 1 : 0;
  const int blocks[] = {64, 128, 256};
  LaunchCfg best;
  double bms = 1e30;
  for (int grp = 0; grp < 3; ++grp)
    for (int b : blocks)
      for (int r = 0; r < 2; ++r) {
        {
          std::lock_guard<std::mutex> lk(g_tune_mu);
          LaunchCfg c;
          c.block = b;
          c.remap = r;
          c.group = grp;
          g_tuned[key_of(lat, prec, clover)] = c;
        }
        double ms = 0;
        QCU_TRY(qcuBenchHop(h, prec, clover, 1, 20, b, r, &ms));
        if (ms < bms) {
          bms = ms;
          best.block = b;
          best.remap = r;
          best.group = grp;
        }
      }
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tuned[key_of(lat, prec, clover)] = best;
  }
  if (best_ms) *best_ms = bms;
  return 0;
}

int qcuTuneGet(QcuLatticeHandle h, int prec, int *block, int *remap, int *group) {
  QCU_CHECK(h && block && remap && group, QCU_ERR_ARG, "null argument");
  LaunchCfg c = tuned_cfg(h->lat, prec);
  *block = c.block;
  *remap = c.remap;
  *group = c.group;
  return 0;
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
  char key[160];
  int b, r, grp;
  std::lock_guard<std::mutex> lk(g_tune_mu);
  while (fscanf(f, "%159s %d %d %d", key, &b, &r, &grp) == 4) {
    if (b >= 64 && b <= 256 && b % 64 == 0 && (r == 0 || r == 1) && grp >= 0 && grp <= 2) {
      LaunchCfg c;
      c.block = b;
      c.remap = r;
      c.group = grp;
      g_tuned[key] = c;
    }
  }
  fclose(f);
  return 0;
}

}  // extern "C"
