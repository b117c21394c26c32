// qcu_solver.h -- solver work space and the preconditioner interface.
#pragma once
#include <map>
#include <memory>

#include "../../include/qcu.h"
#include "qcu_lattice.h"

namespace qcu {

// A fixed right preconditioner K acting on odd-site spinors of one precision.
struct Precond {
  virtual ~Precond() {}
  // out = K in; must only enqueue work on lat.stream (it is graph-captured);
  // st may be used to skip work once the solve has stopped.
  virtual int apply(Lattice &lat, SpinorPtr out, SpinorPtr in, int prec, SysState *st) = 0;
};

struct VecSet {
  int prec = P64, px = P64;
  Spinor r, rt, p, v, s, t, tmp, x, ph, sh;
  std::map<long, hipGraphExec_t> graphs;
  Precond *graph_net = nullptr;
  int create(int prec, int px, int vh, bool precond);
  ~VecSet();
};

struct Workspace {
  int vh = 0;
  Spinor be, bo, bh, xe, xo, r64, t64;
  std::map<int, std::unique_ptr<VecSet>> sets;
  int ensure(int vh);
  VecSet *get_set(int prec, bool precond, int *rc);
};

struct SolveStats {
  int iters, reliable, restarts, converged;
  double schur_rr;
};

LaunchCfg tuned_cfg(const Lattice &lat, int prec);

int schur_solve(Lattice &lat, int prec, double target2, double delta, int max_iter, int chunk, bool use_graph,
                Precond *K, bool warm, SolveStats *stats);
int full_solve(Lattice &lat, void *x_ext, const void *b_ext, const QcuSolveParam &prm, QcuSolveInfo *info);

// blas.hip, preconditioned variants
int omega_dot(SpinorPtr t, SpinorPtr s, int prec, int vh, SysState *st, double *partials, hipStream_t strm);
int bicg_xr_update_pc(SpinorPtr x, int px, SpinorPtr r, SpinorPtr rt, SpinorPtr ph, SpinorPtr sh, SpinorPtr s,
                      SpinorPtr t, int prec, int vh, SysState *st, double *partials, const LaunchCfg &cfg,
                      hipStream_t strm);

}  // namespace qcu

// the opaque handle of the C ABI
struct QcuLattice_s {
  qcu::Lattice lat;
};
