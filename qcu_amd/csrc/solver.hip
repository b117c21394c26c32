// solver.hip -- even-odd BiCGSTAB on the odd-site Schur system, FP64 or
// mixed precision with FP64 reliable updates.
//
// Schur form (symmetric, reducing to include/lattice_bistabcg.h:99-132 and
// src/bistabcg.cu:73-96 for Wilson, A = 1):
//   Mhat x_o = x_o - kappa^2 Ao^-1 D_oe Ae^-1 D_eo x_o
//   bhat     = Ao^-1 (b_o + kappa D_oe Ae^-1 b_e)
//   x_e      = Ae^-1 (b_e + kappa D_eo x_o)
// The recurrence is the reference's (include/lattice_bistabcg.h:227-307); its
// scalars stay on the device and the host looks at them once per chunk of
// iterations.  A chunk is captured once as a hipGraph and replayed.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "../../include/qcu.h"
#include "qcu_lattice.h"
#include "qcu_solver.h"

namespace qcu {

// QCU_TRACE=1 prints the wall time of every solver phase to stderr
struct PhaseTrace {
  bool on;
  hipStream_t s;
  std::chrono::steady_clock::time_point t;
  explicit PhaseTrace(hipStream_t st) : on(getenv("QCU_TRACE") != nullptr), s(st), t(std::chrono::steady_clock::now()) {}
  void mark(const char *what) {
    if (!on) return;
    (void)hipStreamSynchronize(s);
    auto n = std::chrono::steady_clock::now();
    fprintf(stderr, "[qcu] %-28s %9.3f ms\n", what, std::chrono::duration<double, std::milli>(n - t).count());
    t = n;
  }
};

int VecSet::create(int prec_, int px_, int vh, bool precond) {
  prec = prec_;
  px = px_;
  QCU_TRY(r.create(prec, vh));
  QCU_TRY(rt.create(prec, vh));
  QCU_TRY(p.create(prec, vh));
  QCU_TRY(v.create(prec, vh));
  QCU_TRY(s.create(prec, vh));
  QCU_TRY(t.create(prec, vh));
  QCU_TRY(tmp.create(prec, vh));
  QCU_TRY(x.create(px, vh));
  if (precond) {
    QCU_TRY(ph.create(prec, vh));
    QCU_TRY(sh.create(prec, vh));
  }
  return 0;
}

VecSet::~VecSet() {
  for (auto &kv : graphs) {
    (void)hipGraphExecDestroy(kv.second);
  }
}

int Workspace::ensure(int vh_) {
  if (vh == vh_) return 0;
  vh = vh_;
  QCU_TRY(be.create(P64, vh));
  QCU_TRY(bo.create(P64, vh));
  QCU_TRY(bh.create(P64, vh));
  QCU_TRY(xe.create(P64, vh));
  QCU_TRY(xo.create(P64, vh));
  QCU_TRY(r64.create(P64, vh));
  QCU_TRY(t64.create(P64, vh));
  sets.clear();
  return 0;
}

VecSet *Workspace::get_set(int prec, bool precond, int *rc) {
  int key = prec * 2 + (precond ? 1 : 0);
  auto it = sets.find(key);
  if (it != sets.end()) return it->second.get();
  std::unique_ptr<VecSet> vs(new VecSet());
  int px = (prec == P64) ? P64 : P32;
  *rc = vs->create(prec, px, vh, precond);
  if (*rc) return nullptr;
  VecSet *raw = vs.get();
  sets[key] = std::move(vs);
  return raw;
}

// Mhat x (FP64) into out; t64 as scratch
static int apply_mhat64(Lattice &lat, SpinorPtr out, SpinorPtr x, SpinorPtr tmp) {
  HopOp op = lat.op(P64);
  LaunchCfg cfg = tuned_cfg(lat, P64);
  SpinorPtr none{};
  QCU_TRY(hop(op, tmp, x, 0, EPI_PLAIN, none, 0, RED_NONE, none, nullptr, nullptr, cfg, lat.stream));
  QCU_TRY(hop(op, out, tmp, 1, EPI_XPAY, x, lat.kappa * lat.kappa, RED_NONE, none, nullptr, nullptr,
              cfg, lat.stream));
  return 0;
}

// one BiCGSTAB iteration on the device, no host interaction
static int enqueue_iteration(Lattice &lat, VecSet &S, SysState *st, Precond *K) {
  const int prec = S.prec, vh = lat.g.vh;
  HopOp op = lat.op(prec);
  LaunchCfg cfg = tuned_cfg(lat, prec);
  LaunchCfg bcfg;
  bcfg.block = 256;
  SpinorPtr none{};
  const double k2 = lat.kappa * lat.kappa;
  QCU_TRY(bicg_p_update(S.p.ptr(), S.r.ptr(), S.v.ptr(), prec, vh, st, bcfg, lat.stream));
  SpinorPtr pin = S.p.ptr();
  if (K) {
    QCU_TRY(K->apply(lat, S.ph.ptr(), S.p.ptr(), prec, st));
    pin = S.ph.ptr();
  }
  QCU_TRY(hop(op, S.tmp.ptr(), pin, 0, EPI_PLAIN, none, 0, RED_NONE, none, st, nullptr, cfg, lat.stream));
  QCU_TRY(hop(op, S.v.ptr(), S.tmp.ptr(), 1, EPI_XPAY, pin, k2, RED_DOT_Y_OUT, S.rt.ptr(), st, lat.partials,
              cfg, lat.stream));
  QCU_TRY(bicg_s_update(S.s.ptr(), S.r.ptr(), S.v.ptr(), prec, vh, st, bcfg, lat.stream));
  SpinorPtr sin = S.s.ptr();
  if (K) {
    QCU_TRY(K->apply(lat, S.sh.ptr(), S.s.ptr(), prec, st));
    sin = S.sh.ptr();
  }
  QCU_TRY(hop(op, S.tmp.ptr(), sin, 0, EPI_PLAIN, none, 0, RED_NONE, none, st, nullptr, cfg, lat.stream));
  if (K) {
    // t = Mhat s_hat; <t, s> needs s, not s_hat: reduce in the update kernel below
    QCU_TRY(hop(op, S.t.ptr(), S.tmp.ptr(), 1, EPI_XPAY, sin, k2, RED_NONE, none, st, nullptr, cfg,
                lat.stream));
    QCU_TRY(omega_dot(S.t.ptr(), S.s.ptr(), prec, vh, st, lat.partials, lat.stream));
    QCU_TRY(bicg_xr_update_pc(S.x.ptr(), S.px, S.r.ptr(), S.rt.ptr(), S.ph.ptr(), S.sh.ptr(), S.s.ptr(),
                              S.t.ptr(), prec, vh, st, lat.partials, bcfg, lat.stream));
  } else {
    QCU_TRY(hop(op, S.t.ptr(), S.tmp.ptr(), 1, EPI_XPAY, sin, k2, RED_OUT_X, S.s.ptr(), st, lat.partials,
                cfg, lat.stream));
    QCU_TRY(bicg_xr_update(S.x.ptr(), S.px, S.r.ptr(), S.rt.ptr(), S.p.ptr(), S.s.ptr(), S.t.ptr(), prec, vh,
                           st, lat.partials, bcfg, lat.stream));
  }
  return 0;
}

static int run_chunk(Lattice &lat, VecSet &S, SysState *st, Precond *K, int chunk, bool use_graph) {
  if (!use_graph) {
    for (int k = 0; k < chunk; ++k) QCU_TRY(enqueue_iteration(lat, S, st, K));
    return 0;
  }
  const long key = ((long)chunk << 8) | (K ? 1 : 0);
  auto it = S.graphs.find(key);
  if (it == S.graphs.end() || S.graph_net != K) {
    if (it != S.graphs.end()) {
      (void)hipGraphExecDestroy(it->second);
      S.graphs.erase(it);
    }
    hipGraph_t graph;
    QCU_HIP(hipStreamBeginCapture(lat.stream, hipStreamCaptureModeThreadLocal));
    int rc = 0;
    for (int k = 0; k < chunk && !rc; ++k) rc = enqueue_iteration(lat, S, st, K);
    hipError_t e = hipStreamEndCapture(lat.stream, &graph);
    if (rc) return rc;
    QCU_HIP(e);
    hipGraphExec_t exec;
    QCU_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    (void)hipGraphDestroy(graph);
    S.graphs[key] = exec;
    S.graph_net = K;
    it = S.graphs.find(key);
  }
  QCU_HIP(hipGraphLaunch(it->second, lat.stream));
  return 0;
}

static int read_state(Lattice &lat, SysState *st, SysState *host) {
  QCU_HIP(hipMemcpyAsync(host, st, sizeof(SysState), hipMemcpyDeviceToHost, lat.stream));
  QCU_HIP(hipStreamSynchronize(lat.stream));
  return 0;
}

static int write_state(Lattice &lat, SysState *st, const SysState *host) {
  QCU_HIP(hipMemcpyAsync(st, host, sizeof(SysState), hipMemcpyHostToDevice, lat.stream));
  return 0;
}

// Solve Mhat x_o = bhat (bhat in ws.bh); the solution accumulates in ws.xo
// (FP64): the iterated correction S.x is folded into it at every reliable
// update.  warm: start from the current ws.xo instead of zero.
// target2: stop when |bhat - Mhat x_o|^2 < target2 (recomputed in FP64).
int schur_solve(Lattice &lat, int prec, double target2, double delta, int max_iter, int chunk, bool use_graph,
                Precond *K, bool warm, SolveStats *stats) {
  Workspace &ws = *lat.ws;
  const int vh = lat.g.vh;
  int rc = 0;
  VecSet *S = ws.get_set(prec, K != nullptr, &rc);
  if (!S) return rc;
  SysState *st = lat.state;
  SysState *hs = lat.host_state;
  std::memset(hs, 0, sizeof(SysState));
  hs->rr_stop = target2;
  hs->delta2 = (prec == P64) ? 0.0 : delta * delta;
  QCU_TRY(write_state(lat, st, hs));
  LaunchCfg bcfg;
  SpinorPtr src = ws.bh.ptr();
  if (warm) {
    QCU_TRY(apply_mhat64(lat, ws.r64.ptr(), ws.xo.ptr(), ws.t64.ptr()));
    QCU_TRY(xmy_norm(ws.bh.ptr(), P64, ws.r64.ptr(), P64, vh, st, lat.partials, FIN_NONE,
                     lat.stream));
    src = ws.r64.ptr();
  } else {
    QCU_TRY(zero_spinor(ws.xo.ptr(), P64, vh, lat.stream));
  }
  // S.x = 0, r = src, r~ = r, p = v = 0
  QCU_TRY(bicg_init(S->r.ptr(), S->rt.ptr(), S->p.ptr(), S->v.ptr(), S->x.ptr(), prec, src, P64, vh, st,
                    lat.partials, bcfg, lat.stream));
  QCU_TRY(read_state(lat, st, hs));
  PhaseTrace trace(lat.stream);
  trace.mark("  init");
  int iters = 0, reliable = 0, stuck = 0, restarts = 0;
  bool converged = false;
  double last_true = hs->rr;
  for (;;) {
    if (hs->done == 0) {
      const int n = std::min(chunk, max_iter - iters);
      if (n <= 0) break;
      QCU_TRY(run_chunk(lat, *S, st, K, n, use_graph && n == chunk));
      QCU_TRY(read_state(lat, st, hs));
      iters = hs->iter;
      trace.mark("  chunk");
      continue;
    }
    const int why = hs->done;
    if (why == 2 && ++restarts > 20) break;
    // reliable update: fold the correction into the FP64 solution and
    // recompute the residual in FP64 (also the convergence check)
    QCU_TRY(axpby(1.0, S->x.ptr(), S->px, 1.0, ws.xo.ptr(), P64, vh, lat.stream));
    QCU_TRY(zero_spinor(S->x.ptr(), S->px, vh, lat.stream));
    QCU_TRY(apply_mhat64(lat, ws.r64.ptr(), ws.xo.ptr(), ws.t64.ptr()));
    QCU_TRY(xmy_norm(ws.bh.ptr(), P64, ws.r64.ptr(), P64, vh, st, lat.partials, FIN_STORE,
                     lat.stream));
    QCU_TRY(read_state(lat, st, hs));
    const double rr_true = hs->tmp[0];
    ++reliable;
    if (rr_true < target2) {
      converged = true;
      last_true = rr_true;
      break;
    }
    if (iters >= max_iter) {
      last_true = rr_true;
      break;
    }
    if (why == 1) {
      // the iterated residual claimed convergence but the true one did not:
      // ask the recurrence for more (guards against a precision floor)
      if (++stuck > 8) {
        last_true = rr_true;
        break;
      }
      hs->rr_stop = std::min(hs->rr_stop, target2) * 0.01;
    }
    last_true = rr_true;
    QCU_TRY(spinor_convert(S->r.ptr(), prec, ws.r64.ptr(), P64, vh, lat.stream));
    if (why == 2) {
      // breakdown: restart the Krylov space from the current residual
      QCU_TRY(spinor_convert(S->rt.ptr(), prec, ws.r64.ptr(), P64, vh, lat.stream));
      QCU_TRY(zero_spinor(S->p.ptr(), prec, vh, lat.stream));
      QCU_TRY(zero_spinor(S->v.ptr(), prec, vh, lat.stream));
      hs->rho_prev[0] = 1;
      hs->rho_prev[1] = 0;
      hs->alpha[0] = hs->omega[0] = 1;
      hs->alpha[1] = hs->omega[1] = 0;
    }
    hs->done = 0;
    QCU_TRY(write_state(lat, st, hs));
    QCU_TRY(reliable_dot(S->rt.ptr(), S->r.ptr(), prec, vh, st, lat.partials, lat.stream));
    QCU_TRY(read_state(lat, st, hs));
    trace.mark("  reliable update");
  }
  QCU_TRY(axpby(1.0, S->x.ptr(), S->px, 1.0, ws.xo.ptr(), P64, vh, lat.stream));
  QCU_TRY(zero_spinor(S->x.ptr(), S->px, vh, lat.stream));
  stats->iters = iters;
  stats->reliable = reliable;
  stats->restarts = restarts;
  stats->converged = converged;
  stats->schur_rr = last_true;
  return 0;
}

// Full solve: b (external full, device) -> x (external full, device)
int full_solve(Lattice &lat, void *x_ext, const void *b_ext, const QcuSolveParam &prm, QcuSolveInfo *info) {
  QCU_CHECK(lat.gauge_loaded, QCU_ERR_STATE, "gauge field not loaded");
  if (!lat.ws) lat.ws.reset(new Workspace());
  Workspace &ws = *lat.ws;
  QCU_TRY(ws.ensure(lat.g.vh));
  const int vh = lat.g.vh;
  auto t0 = std::chrono::steady_clock::now();
  PhaseTrace trace(lat.stream);
  const char *b = (const char *)b_ext;
  char *x = (char *)x_ext;
  const size_t half = (size_t)vh * 12 * 16;
  QCU_TRY(spinor_ext_to_internal(b, lat.g, ws.be.ptr(), P64, lat.stream));
  QCU_TRY(spinor_ext_to_internal(b + half, lat.g, ws.bo.ptr(), P64, lat.stream));
  // bhat = Ao^-1 (b_o + kappa D_oe Ae^-1 b_e)
  HopOp op = lat.op(P64);
  LaunchCfg cfg = tuned_cfg(lat, P64);
  SpinorPtr none{};
  QCU_TRY(cl_axpy(ws.t64.ptr(), ws.be.ptr(), 0.0, none, lat.clover_inv(P64, 0), vh, lat.stream));
  QCU_TRY(hop(op, ws.r64.ptr(), ws.t64.ptr(), 1, EPI_PLAIN, none, 0, RED_NONE, none, nullptr, nullptr, cfg,
              lat.stream, false));
  QCU_TRY(cl_axpy(ws.bh.ptr(), ws.bo.ptr(), lat.kappa, ws.r64.ptr(), lat.clover_inv(P64, 1), vh, lat.stream));
  // |b|^2 for the stopping rule
  QCU_TRY(norm2_dev(ws.be.ptr(), P64, vh, lat.dscal, lat.partials, lat.stream));
  QCU_TRY(norm2_dev(ws.bo.ptr(), P64, vh, lat.dscal + 1, lat.partials, lat.stream));
  QCU_TRY(norm2_dev(ws.bh.ptr(), P64, vh, lat.dscal + 2, lat.partials, lat.stream));
  QCU_HIP(hipMemcpyAsync(lat.host_scal, lat.dscal, 3 * sizeof(double), hipMemcpyDeviceToHost, lat.stream));
  QCU_HIP(hipStreamSynchronize(lat.stream));
  const double b2 = lat.host_scal[0] + lat.host_scal[1];
  const double bh2 = lat.host_scal[2];
  trace.mark("b_hat + norms");
  double target2;
  if (prm.legacy_stop)
    target2 = prm.tol;  // Re<r,r> < tol on the Schur residual, as the reference
  else
    target2 = 0.25 * prm.tol * prm.tol * b2;
  const int chunk = prm.chunk > 0 ? prm.chunk : 8;
  const int max_iter = prm.max_iter > 0 ? prm.max_iter : 10000;
  const double delta = prm.reliable_delta > 0 ? prm.reliable_delta : 0.1;
  Precond *K = (Precond *)prm.precond;
  SolveStats st{};
  int total_iters = 0, total_rel = 0, fallback = 0, restarts = 0;
  double res = INFINITY;
  for (int attempt = 0; attempt < 6; ++attempt) {
    QCU_TRY(schur_solve(lat, prm.precision, target2, delta, max_iter - total_iters, chunk, prm.use_graph != 0, K,
                        attempt > 0, &st));
    trace.mark("schur solve");
    total_iters += st.iters;
    total_rel += st.reliable;
    restarts += st.restarts;
    if (!st.converged && K) {
      // stalled preconditioned solve: fall back to the plain method from the current x_o
      K = nullptr;
      fallback = 1;
      continue;
    }
    // x_e = Ae^-1 (b_e + kappa D_eo x_o)
    QCU_TRY(hop(op, ws.t64.ptr(), ws.xo.ptr(), 0, EPI_PLAIN, none, 0, RED_NONE, none, nullptr, nullptr, cfg,
                lat.stream, false));
    QCU_TRY(cl_axpy(ws.xe.ptr(), ws.be.ptr(), lat.kappa, ws.t64.ptr(), lat.clover_inv(P64, 0), vh, lat.stream));
    QCU_TRY(lattice_true_residual(lat, ws.xe.ptr(), ws.xo.ptr(), ws.be.ptr(), ws.bo.ptr(), &res, nullptr,
                                  ws.r64.ptr()));
    trace.mark("x_e + true residual");
    if (prm.legacy_stop || res <= prm.tol || total_iters >= max_iter || !st.converged) break;
    target2 *= 0.01;  // tighten the Schur target and continue from the current x_o
  }
  (void)bh2;
  QCU_TRY(spinor_internal_to_ext(ws.xe.ptr(), lat.g, x, P64, lat.stream));
  QCU_TRY(spinor_internal_to_ext(ws.xo.ptr(), lat.g, x + half, P64, lat.stream));
  QCU_HIP(hipStreamSynchronize(lat.stream));
  auto t1 = std::chrono::steady_clock::now();
  if (info) {
    info->iters = total_iters;
    info->reliable_updates = total_rel;
    info->converged = prm.legacy_stop ? st.converged : (res <= prm.tol);
    info->fallback = fallback;
    info->true_res = res;
    info->seconds = std::chrono::duration<double>(t1 - t0).count();
    info->restarts = restarts;
  }
  return 0;
}

}  // namespace qcu
