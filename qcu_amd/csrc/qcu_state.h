// qcu_state.h -- device-side scalar epilogues of the BiCGSTAB recurrence
// (van der Vorst; the order of updates follows include/lattice_bistabcg.h:227-307
// and src/bistabcg.cu:4-48, with all scalars kept on the device).
#pragma once
#include "qcu_internal.h"

namespace qcu {

QCU_D void cdiv(const double *a, const double *b, double *o) {
  double den = b[0] * b[0] + b[1] * b[1];
  double re = (a[0] * b[0] + a[1] * b[1]) / den;
  double im = (a[1] * b[0] - a[0] * b[1]) / den;
  o[0] = re;
  o[1] = im;
}
QCU_D void cmul(const double *a, const double *b, double *o) {
  double re = a[0] * b[0] - a[1] * b[1];
  double im = a[0] * b[1] + a[1] * b[0];
  o[0] = re;
  o[1] = im;
}
QCU_D bool finite2(double a, double b) { return isfinite(a) && isfinite(b); }

// alpha = rho / <r~, v>
QCU_D void fin_alpha(SysState *st, double dre, double dim) {
  double d[2] = {dre, dim};
  if (!(dre * dre + dim * dim > 0) || !finite2(dre, dim)) {
    st->done = 2;
    return;
  }
  cdiv(st->rho, d, st->alpha);
}

// omega = <t, s> / <t, t>
QCU_D void fin_omega(SysState *st, double tsre, double tsim, double tt) {
  if (!(tt > 0) || !isfinite(tt)) {
    // t = 0 means s = 0: the residual vanished; keep x update with omega = 0
    st->omega[0] = 0;
    st->omega[1] = 0;
    return;
  }
  st->omega[0] = tsre / tt;
  st->omega[1] = tsim / tt;
}

// rho' = <r~, r>, rr = |r|^2, beta = (rho'/rho) (alpha/omega)
QCU_D void fin_rho(SysState *st, double rre, double rim, double rr) {
  st->rho_prev[0] = st->rho[0];
  st->rho_prev[1] = st->rho[1];
  st->rho[0] = rre;
  st->rho[1] = rim;
  st->rr = rr;
  st->iter += 1;
  if (!isfinite(rr)) {
    st->done = 2;
    return;
  }
  if (rr < st->rr_stop) {
    st->done = 1;
    return;
  }
  double q1[2], q2[2];
  cdiv(st->rho, st->rho_prev, q1);
  double om2 = st->omega[0] * st->omega[0] + st->omega[1] * st->omega[1];
  double rp2 = st->rho_prev[0] * st->rho_prev[0] + st->rho_prev[1] * st->rho_prev[1];
  if (!(om2 > 0) || !(rp2 > 0)) {
    st->done = 2;
    return;
  }
  cdiv(st->alpha, st->omega, q2);
  cmul(q1, q2, st->beta);
  if (st->delta2 > 0 && rr < st->delta2 * st->rr_ref) st->done = 3;
}

// ---- CG on the normal equations Mhat^+ Mhat x = Mhat^+ b (CGNR): real scalars
// kept in the real parts of the state fields.  rr is the squared residual of
// the Schur system, |b - Mhat x|^2, carried by the recursion res -= alpha q.
QCU_D void fin_cg_init(SysState *st, double zz, double rr) {
  st->rho[0] = zz;
  st->beta[0] = 0;
  st->rr = rr;
  st->done = (rr < st->rr_stop) ? 1 : 0;
}

QCU_D void fin_cg_alpha(SysState *st, double qq) {
  if (!(qq > 0) || !isfinite(qq)) {
    st->done = 2;
    return;
  }
  st->alpha[0] = st->rho[0] / qq;
}

QCU_D void fin_cg_rho(SysState *st, double zz) {
  const double rr = st->tmp[0];
  st->iter += 1;
  st->rr = rr;
  if (!isfinite(rr) || !isfinite(zz)) {
    st->done = 2;
    return;
  }
  if (rr < st->rr_stop) {
    st->done = 1;
    return;
  }
  if (!(st->rho[0] > 0)) {
    st->done = 2;
    return;
  }
  st->beta[0] = zz / st->rho[0];
  st->rho[0] = zz;
}

}  // namespace qcu
