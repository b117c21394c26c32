// qcu_lattice.h -- resident lattice state: gauge links in every precision,
// packed clover (and inverse) per parity, solver work space.  One object per
// (gauge field, kappa, csw, boundary); nothing is allocated per operator call.
#pragma once
#include <map>
#include <memory>

#include "qcu_internal.h"

namespace qcu {

struct Workspace;  // solver vectors, see solver.hip

struct Lattice {
  int L[4] = {0, 0, 0, 0};
  Geom g{};
  double kappa = 0.125;
  double csw = 0.0;
  int device = 0;
  hipStream_t stream = nullptr;
  bool gauge_loaded = false;

  DevBuf U64, U32, U16a, U16b;
  DevBuf clov64[2];        // packed A (for the true residual)
  DevBuf cinv64[2];        // packed A^-1
  DevBuf cinv32[2], cinv16[2], cinv16n[2];
  DevBuf scratch;          // partials + state + small results + bad flag
  double *partials = nullptr;
  SysState *state = nullptr;     // kNumStates solver states
  double *dscal = nullptr;       // small FP64 results
  int *bad = nullptr;
  SysState *host_state = nullptr;  // pinned mirror
  double *host_scal = nullptr;

  std::unique_ptr<Workspace> ws;

  static constexpr int kNumStates = 64;

  int init(const int L_[4], double kappa_, double csw_, int tbc);
  // prec_mask bit k: keep links (and clover inverse) in precision k
  int load_gauge(const void *gauge_ext_dev, int prec_mask = 7);
  int prec_mask = 0;
  bool has_clover() const { return csw != 0.0; }
  HopOp op(int prec) const;
  LinkPtr links(int prec) const;
  CloverPtr clover_inv(int prec, int parity) const;
  CloverPtr clover(int parity) const;
  size_t max_blocks() const;
  ~Lattice();
};

// vectors in the external layout, helpers for the C entry points
// scratch: an FP64 half-field work vector, or {} to allocate one for the call
int lattice_true_residual(Lattice &lat, SpinorPtr xe, SpinorPtr xo, SpinorPtr be, SpinorPtr bo,
                          double *res, double *bnorm, SpinorPtr scratch);
LaunchCfg tuned_cfg(const Lattice &lat, int prec);

// device-local error string used by the legacy C entry points
int current_device();

}  // namespace qcu
