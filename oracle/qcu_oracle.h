/*
 * qcu_oracle.h -- FP64 host restatement of qcu's hot path (TEST INFRASTRUCTURE).
 *
 * This library is the checker for the HIP implementation in qcu_amd/: only the
 * test-suite, __graft_entry__.smoke() and the cpu_baseline leg of bench.py may
 * load it.  The product never links or calls it.
 *
 * Every routine works in the external (PyQuda) layouts that qcu's C ABI uses:
 *   gauge   complex128 [4 mu][2 parity][Lt][Lz][Ly][Lx/2][3 row][3 col]
 *   half    complex128 [Lt][Lz][Ly][Lx/2][4 spin][3 colour]       (one parity)
 *   full    complex128 [2 parity][Lt][Lz][Ly][Lx/2][4][3]         (even first)
 *   clover  complex128 [Vh][12][12]  (row sc0 = s0*3+c0, col sc1 = s1*3+c1)
 * with parity = (x+y+z+t) mod 2 and a site stored at x/2 of its parity array
 * (python/PyQuda/pyquda/field.py:22-63).
 *
 * t_boundary: +1 periodic, -1 antiperiodic in time.  The sign is applied as a
 * phase on every hop that crosses the t = Lt-1 <-> 0 boundary, after the
 * 12-real link reconstruction (the reference kernel wraps periodically).
 *
 * Pinning status of each routine is listed in qcu_oracle.c's header.
 */
#ifndef QCU_ORACLE_H
#define QCU_ORACLE_H
#include <complex.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef double _Complex ocplx;

/* ---- generators ---------------------------------------------------------- */
void or_gauge_random(ocplx *U, const int L[4], double eps, uint64_t seed);
void or_gauss_fill(ocplx *v, long n, uint64_t seed);
double or_uniform(uint64_t seed, uint64_t stream, uint64_t k);
void or_su3_exp(const double omega[8], double eps, ocplx out[9]);

/* ---- Wilson hop (src/wilson_dslash.cu:9-245) ----------------------------- */
void or_hop(ocplx *out, const ocplx *in, const ocplx *U, const int L[4],
            int parity, int t_boundary);
void or_hop_full(ocplx *out, const ocplx *in, const ocplx *U, const int L[4],
                 int t_boundary);

/* ---- clover (src/clover_dslash.cu:5-882) -------------------------------- */
void or_clover_build(ocplx *A, const ocplx *U, const int L[4], int parity,
                     double coeff);
int or_clover_invert(ocplx *Ainv, const ocplx *A, long nsites);
void or_clover_apply(ocplx *out, const ocplx *A, const ocplx *in, long nsites);
void or_clover_dslash(ocplx *out, const ocplx *in, const ocplx *U,
                      const int L[4], int parity, double coeff, int t_boundary);

/* ---- operators on the full lattice ------------------------------------- */
void or_mat_full(ocplx *out, const ocplx *in, const ocplx *U, const ocplx *Afull,
                 const int L[4], double kappa, int t_boundary);

/* ---- BLAS-1 -------------------------------------------------------------- */
ocplx or_dot(const ocplx *a, const ocplx *b, long n);
double or_norm2(const ocplx *a, long n);
/* out[0] = or_dot(a, b, n), for callers without C99 complex returns */
void or_dot_out(ocplx *out, const ocplx *a, const ocplx *b, long n);

/* ---- even-odd BiCGSTAB (include/lattice_bistabcg.h:221-340) ------------ */
typedef struct {
  double kappa;
  double csw;       /* coefficient kappa*csw enters the clover term; 0 => Wilson */
  int t_boundary;
  double tol;       /* see mode */
  int max_iter;
  int mode;         /* 0: stop on Re<r,r> < tol (reference), 1: relative |r|/|b_hat| < tol */
  int iters;        /* out */
  double schur_res; /* out: |b_hat - M_hat x_o| / |b_hat| recomputed */
  double true_res;  /* out: |b - M x| / |b| on the full lattice */
  double setup_s;   /* out: wall seconds of the clover build/inverse and b_hat */
  double iter_s;    /* out: wall seconds of the iterations */
} or_solve_param;
int or_bicgstab(ocplx *x, const ocplx *b, const ocplx *U, const int L[4],
                or_solve_param *p);
double or_true_residual(const ocplx *x, const ocplx *b, const ocplx *U,
                        const int L[4], double kappa, double csw, int t_boundary);

/* ---- learned gauge-covariant network ------------------------------------ */
/* weights: per layer l with (cin, cout): W[l] = complex [9 terms][cout][cin][4][4];
 * term 2*mu is the forward hop (A_mu), 2*mu+1 the backward hop (B_mu), 8 is C. */
long or_net_weight_count(int nlayers, const int *chan);
void or_net_forward(ocplx *out, const ocplx *in, const ocplx *W, int nlayers,
                    const int *chan, const ocplx *U, const int L[4], int t_boundary,
                    ocplx *acts);
void or_net_backward(ocplx *grad_in, ocplx *gradW, const ocplx *grad_out,
                     const ocplx *W, int nlayers, const int *chan,
                     const ocplx *U, const int L[4], int t_boundary,
                     const ocplx *acts);

/* ---- threads ------------------------------------------------------------ */
int or_num_threads(void);
void or_set_num_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
