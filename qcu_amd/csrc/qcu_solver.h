// qcu_solver.h -- BiCGSTAB work space over a batch and the preconditioner
// interface.
#pragma once
#include <map>
#include <memory>
#include <vector>

#include "../../include/qcu.h"
#include "qcu_lattice.h"

namespace qcu {

struct Solver;

// A fixed right preconditioner K acting on odd-site batch vectors of one
// precision.
struct Precond {
  virtual ~Precond() {}
  // out = K in; must only enqueue work on S.stream (it is graph-captured);
  // st lets kernels skip systems that have stopped.
  virtual int apply(Solver &S, SpinorPtr out, SpinorPtr in, int prec, SysState *st) = 0;
  // unique over the process lifetime: a captured graph is valid only for the
  // preconditioner (and device buffers) it was captured with
  virtual unsigned long long uid() const = 0;
};

struct VecSet {
  int prec = P64, px = P64;
  Spinor r, rt, p, v, s, t, tmp, x, ph, sh;
  std::map<long, hipGraphExec_t> graphs;
  unsigned long long graph_net = 0;  // uid of the preconditioner the graphs were captured with
  unsigned long long graph_tune = 0; // tune_generation() when the graphs were captured
  int create(int prec, int px, int n, bool precond);
  ~VecSet();
};

struct Solver {
  const BatchOp *op = nullptr;
  hipStream_t stream = nullptr;
  int n = 0;      // virtual sites of the batch
  int nsys = 0;
  Spinor be, bo, bh, xe, xo, r64, t64;  // FP64 batch vectors
  std::map<int, std::unique_ptr<VecSet>> sets;
  DevBuf scratch;
  double *partials = nullptr;
  SysState *st = nullptr;    // nsys device states
  double *dscal = nullptr;   // 4 * nsys device scalars
  SysState *host_st = nullptr;
  double *host_scal = nullptr;
  Spinor cg_z, cg_p, cg_q;  // CG on the normal equations (FP64), made on first use
  int init(const BatchOp *op, hipStream_t s);
  VecSet *get_set(int prec, bool precond, int *rc);
  ~Solver();
};

struct SolveStats {
  int iters, reliable, restarts, converged;
};

LaunchCfg tuned_cfg(const BatchOp &op, int prec);
std::string tune_key(const BatchOp &op, int prec);
// changes whenever a tuned launch shape is set (captured graphs hold the old one)
unsigned long long tune_generation();

// 0 when p is a HIP device (or managed) pointer; else records the error
int check_dev_ptr(const void *p, const char *what);

// the preconditioner behind a QcuNetHandle; it must belong to lattice lat
int net_precond(void *handle, const Lattice *lat, Precond **out);

// b[k], x[k]: full-lattice external fields of system k (device pointers)
int full_solve(Solver &S, void *const *x_ext, const void *const *b_ext, const QcuSolveParam &prm,
               QcuSolveInfo *info);
// |b_k - M x_k| / |b_k| per system from external full fields
int true_residuals(Solver &S, const void *const *x_ext, const void *const *b_ext, double *res);

}  // namespace qcu

// the opaque handles of the C ABI
struct QcuLattice_s {
  qcu::Lattice lat;
};
