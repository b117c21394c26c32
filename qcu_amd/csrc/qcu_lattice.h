// qcu_lattice.h -- resident lattice state: gauge links in every precision,
// packed clover (and inverse) per parity.  One object per (gauge field, kappa,
// csw, boundary); nothing is allocated per operator call.
#pragma once
#include <memory>

#include "qcu_internal.h"

namespace qcu {

struct Solver;  // solver work space, see qcu_solver.h

struct Lattice {
  int L[4] = {0, 0, 0, 0};
  Geom g{};
  double kappa = 0.125;
  double csw = 0.0;
  int device = 0;
  hipStream_t stream = nullptr;
  bool gauge_loaded = false;
  int prec_mask = 0;

  DevBuf U64, U32, U16a, U16b;
  DevBuf clov64[2];  // packed A (for the true residual)
  DevBuf cinv64[2];  // packed A^-1
  DevBuf cinv32[2], cinv16[2], cinv16n[2];
  DevBuf flag;       // singular-clover counter

  BatchOp self;      // this lattice as a one-system batch
  std::unique_ptr<Solver> solver;

  int init(const int L_[4], double kappa_, double csw_, int tbc);
  // prec_mask bit k: keep links (and clover inverse) in precision k
  int load_gauge(const void *gauge_ext_dev, int prec_mask = 7);
  bool has_clover() const { return csw != 0.0; }
  LinkPtr links(int prec) const;
  CloverPtr clover_inv(int prec, int parity) const;
  CloverPtr clover(int parity) const;
  SegDesc seg_desc(int nrhs) const;
  ~Lattice();
};

int current_device();

}  // namespace qcu
