/*
 * qcu_oracle.c -- FP64 OpenMP host restatement of qcu's hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  The HIP library in qcu_amd/ never links or calls
 * this file; the tests use it as the checker and bench.py times it as the
 * CPU baseline ("port": a restatement, not the reference build -- the
 * reference's .cu sources need nvcc, cuBLAS, NCCL and MPI).
 *
 * What each routine follows, and what pins it (tests/ holds the checks):
 *   or_hop          src/wilson_dslash.cu:9-245 (projectors, 12-real
 *                   reconstruction include/define.h:198-206, neighbour moves
 *                   include/define.h:127-134).  Pinned by an independent
 *                   lexicographic numpy dslash built from the DeGrand-Rossi
 *                   matrices of python/PyQuda/pyquda/utils/gamma.py:19-64
 *                   (2^4, 4^4, 4^3x8), the free-field plane-wave closed form,
 *                   gauge covariance and gamma5-hermiticity.
 *   or_clover_build src/clover_dslash.cu:5-837 (leaf products and the sign
 *                   table at :161-172, :288-300, :416-428, :548-560,
 *                   :680-692, :812-824; scale -1/8 and +1 at :826-835, here
 *                   generalised to -coeff/8 with coeff = kappa*csw).  Pinned by
 *                   the textbook 1 - coeff*sum sigma_munu F_munu with
 *                   sigma_munu = (i/2)[g_mu,g_nu] built from gamma.py data,
 *                   U=1 => A=1, hermiticity, gauge covariance.
 *   or_clover_invert include/define.h:334-365 (Gauss-Jordan without pivoting,
 *                   as src/clover_dslash.cu:839-855).  Pinned by A*Ainv = 1 and
 *                   numpy.linalg.inv.
 *   or_clover_apply src/clover_dslash.cu:857-882.  Pinned through
 *                   or_clover_dslash against numpy.
 *   or_bicgstab     include/lattice_bistabcg.h:221-340 with the even-odd
 *                   set-up of :99-132 and src/bistabcg.cu:49-180.  Pinned by a
 *                   dense numpy.linalg.solve of the full matrix on 2^4 / 4^4.
 *   or_gauge_random, or_gauss_fill  no reference counterpart (the reference
 *                   tests draw QUDA's gaussGauge and cupy randn).  Pinned by
 *                   unitarity, det U = 1, and Gaussian moments.
 *   or_net_forward / or_net_backward  no reference counterpart (new learned
 *                   preconditioner).  Pinned by an independent numpy
 *                   restatement (np.roll shifts, einsum mixing), linearity,
 *                   gauge covariance, the adjoint (dot-product) identity and
 *                   central finite differences of the loss.
 *   or_dot, or_norm2, or_true_residual  closed forms (sum conj(a) b); checked
 *                   against numpy.vdot and the dense operator.
 */
#include "qcu_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define I_ _Complex_I

int or_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
void or_set_num_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

/* ------------------------------------------------------------------------ */
/* lattice geometry                                                          */
/* ------------------------------------------------------------------------ */
typedef struct {
  int L[4];  /* full extents x,y,z,t */
  int lxh;   /* Lx/2 */
  long vh;   /* half volume */
} geom;

static geom mkgeom(const int L[4]) {
  geom g;
  for (int i = 0; i < 4; ++i) g.L[i] = L[i];
  g.lxh = L[0] / 2;
  g.vh = (long)g.lxh * L[1] * L[2] * L[3];
  return g;
}

/* half-site index of parity p -> full coordinates */
static void decode(const geom *g, int p, long i, int c[4]) {
  long r = i;
  int xh = (int)(r % g->lxh); r /= g->lxh;
  c[1] = (int)(r % g->L[1]); r /= g->L[1];
  c[2] = (int)(r % g->L[2]); r /= g->L[2];
  c[3] = (int)r;
  c[0] = 2 * xh + ((c[1] + c[2] + c[3] + p) & 1);
}

/* full coordinates -> (parity, half index) */
static long encode(const geom *g, const int c[4], int *p) {
  *p = (c[0] + c[1] + c[2] + c[3]) & 1;
  return (((long)c[3] * g->L[2] + c[2]) * g->L[1] + c[1]) * g->lxh + c[0] / 2;
}

/* neighbour of site c shifted by dir (+1/-1) in mu, periodic wrap */
static long neighbour(const geom *g, const int c[4], int mu, int dir, int *p) {
  int d[4] = {c[0], c[1], c[2], c[3]};
  d[mu] += dir;
  if (d[mu] < 0) d[mu] += g->L[mu];
  if (d[mu] >= g->L[mu]) d[mu] -= g->L[mu];
  return encode(g, d, p);
}

static const ocplx *link_ptr(const ocplx *U, const geom *g, int mu, int p, long i) {
  return U + ((((long)mu * 2 + p) * g->vh + i) * 9);
}

/* rows 0,1 read, row 2 = conj(row0 x row1): include/define.h:198-206 */
static void recon12(const ocplx *src, ocplx u[9]) {
  for (int k = 0; k < 6; ++k) u[k] = src[k];
  u[6] = conj(u[1] * u[5] - u[2] * u[4]);
  u[7] = conj(u[2] * u[3] - u[0] * u[5]);
  u[8] = conj(u[0] * u[4] - u[1] * u[3]);
}

/* ------------------------------------------------------------------------ */
/* counter-based generators                                                  */
/* ------------------------------------------------------------------------ */
static inline uint64_t mix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ULL;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

/* uniform in (0,1]: 53 random bits */
double or_uniform(uint64_t seed, uint64_t stream, uint64_t k) {
  uint64_t h = mix64(mix64(seed ^ mix64(stream)) + k);
  return ((double)(h >> 11) + 1.0) * (1.0 / 9007199254740992.0);
}

/* pair of N(0,1) deviates by Box-Muller from draws (2k, 2k+1) of a stream */
static void gauss_pair(uint64_t seed, uint64_t stream, uint64_t k, double *g0, double *g1) {
  double u1 = or_uniform(seed, stream, 2 * k);
  double u2 = or_uniform(seed, stream, 2 * k + 1);
  double r = sqrt(-2.0 * log(u1));
  double a = 6.283185307179586 * u2;
  *g0 = r * cos(a);
  *g1 = r * sin(a);
}

void or_gauss_fill(ocplx *v, long n, uint64_t seed) {
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; ++i) {
    double a, b;
    gauss_pair(seed, (uint64_t)i, 0, &a, &b);
    v[i] = a + I_ * b;
  }
}

static void mat3_mul(const ocplx *a, const ocplx *b, ocplx *c) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      ocplx s = 0;
      for (int k = 0; k < 3; ++k) s += a[i * 3 + k] * b[k * 3 + j];
      c[i * 3 + j] = s;
    }
}

/* U = exp(i eps H), H = sum_a omega_a lambda_a (Gell-Mann), by scaling and
 * squaring of a degree-16 Taylor series. */
void or_su3_exp(const double w[8], double eps, ocplx out[9]) {
  const double s3 = 0.57735026918962576451; /* 1/sqrt(3) */
  ocplx H[9];
  H[0] = w[2] + w[7] * s3;
  H[4] = -w[2] + w[7] * s3;
  H[8] = -2.0 * w[7] * s3;
  H[1] = w[0] - I_ * w[1];
  H[3] = w[0] + I_ * w[1];
  H[2] = w[3] - I_ * w[4];
  H[6] = w[3] + I_ * w[4];
  H[5] = w[5] - I_ * w[6];
  H[7] = w[5] + I_ * w[6];
  ocplx A[9];
  double nrm = 0;
  for (int k = 0; k < 9; ++k) {
    A[k] = I_ * eps * H[k];
    nrm += creal(A[k]) * creal(A[k]) + cimag(A[k]) * cimag(A[k]);
  }
  nrm = sqrt(nrm);
  int s = 0;
  while (nrm > 0.25) { nrm *= 0.5; ++s; }
  double sc = ldexp(1.0, -s);
  for (int k = 0; k < 9; ++k) A[k] *= sc;
  /* Horner: E = I + A(I + A/2(I + A/3(...))) */
  ocplx E[9], T[9];
  for (int k = 0; k < 9; ++k) E[k] = (k % 4 == 0) ? 1.0 : 0.0;
  for (int n = 16; n >= 1; --n) {
    mat3_mul(A, E, T);
    for (int k = 0; k < 9; ++k) E[k] = ((k % 4 == 0) ? 1.0 : 0.0) + T[k] / (double)n;
  }
  for (int q = 0; q < s; ++q) {
    mat3_mul(E, E, T);
    memcpy(E, T, sizeof(T));
  }
  memcpy(out, E, sizeof(E));
}

void or_gauge_random(ocplx *U, const int L[4], double eps, uint64_t seed) {
  geom g = mkgeom(L);
  long nlinks = 8 * g.vh;
#pragma omp parallel for schedule(static)
  for (long n = 0; n < nlinks; ++n) {
    double w[8];
    for (int k = 0; k < 4; ++k) gauss_pair(seed, (uint64_t)n, (uint64_t)k, &w[2 * k], &w[2 * k + 1]);
    or_su3_exp(w, eps, U + n * 9);
  }
}

/* ------------------------------------------------------------------------ */
/* Wilson hop: out_p(x) = sum_mu (1-g_mu) U_mu(x) psi(x+mu)                  */
/*                      + (1+g_mu) U_mu(x-mu)^dag psi(x-mu)                 */
/* ------------------------------------------------------------------------ */
static inline void mv(const ocplx u[9], const ocplx h[3], ocplx o[3]) {
  for (int c0 = 0; c0 < 3; ++c0) {
    ocplx s = 0;
    for (int c1 = 0; c1 < 3; ++c1) s += h[c1] * u[c0 * 3 + c1];
    o[c0] = s;
  }
}
static inline void mvd(const ocplx u[9], const ocplx h[3], ocplx o[3]) {
  for (int c0 = 0; c0 < 3; ++c0) {
    ocplx s = 0;
    for (int c1 = 0; c1 < 3; ++c1) s += h[c1] * conj(u[c1 * 3 + c0]);
    o[c0] = s;
  }
}

static void hop_site(ocplx *dst, const ocplx *in, const ocplx *U, const geom *g,
                     int p, long i, int tb) {
  int c[4];
  decode(g, p, i, c);
  ocplx d[12];
  for (int k = 0; k < 12; ++k) d[k] = 0;
  ocplx u[9], h0[3], h1[3], a0[3], a1[3];
  for (int mu = 0; mu < 4; ++mu) {
    /* backward: U_mu(x-mu)^dag (1+g_mu) psi(x-mu) */
    {
      int q;
      long j = neighbour(g, c, mu, -1, &q);
      recon12(link_ptr(U, g, mu, q, j), u);
      double ph = (mu == 3 && c[3] == 0) ? (double)tb : 1.0;
      const ocplx *s = in + j * 12;
      for (int cc = 0; cc < 3; ++cc) {
        const ocplx s0 = s[cc], s1 = s[3 + cc], s2 = s[6 + cc], s3 = s[9 + cc];
        switch (mu) {
          case 0: h0[cc] = s0 + I_ * s3; h1[cc] = s1 + I_ * s2; break;
          case 1: h0[cc] = s0 - s3;      h1[cc] = s1 + s2;      break;
          case 2: h0[cc] = s0 + I_ * s2; h1[cc] = s1 - I_ * s3; break;
          default: h0[cc] = s0 + s2;     h1[cc] = s1 + s3;      break;
        }
      }
      mvd(u, h0, a0);
      mvd(u, h1, a1);
      for (int cc = 0; cc < 3; ++cc) {
        ocplx x0 = ph * a0[cc], x1 = ph * a1[cc];
        d[cc] += x0;
        d[3 + cc] += x1;
        switch (mu) {
          case 0: d[6 + cc] -= I_ * x1; d[9 + cc] -= I_ * x0; break;
          case 1: d[6 + cc] += x1;      d[9 + cc] -= x0;      break;
          case 2: d[6 + cc] -= I_ * x0; d[9 + cc] += I_ * x1; break;
          default: d[6 + cc] += x0;     d[9 + cc] += x1;      break;
        }
      }
    }
    /* forward: U_mu(x) (1-g_mu) psi(x+mu) */
    {
      int q;
      long j = neighbour(g, c, mu, +1, &q);
      recon12(link_ptr(U, g, mu, p, i), u);
      double ph = (mu == 3 && c[3] == g->L[3] - 1) ? (double)tb : 1.0;
      const ocplx *s = in + j * 12;
      for (int cc = 0; cc < 3; ++cc) {
        const ocplx s0 = s[cc], s1 = s[3 + cc], s2 = s[6 + cc], s3 = s[9 + cc];
        switch (mu) {
          case 0: h0[cc] = s0 - I_ * s3; h1[cc] = s1 - I_ * s2; break;
          case 1: h0[cc] = s0 + s3;      h1[cc] = s1 - s2;      break;
          case 2: h0[cc] = s0 - I_ * s2; h1[cc] = s1 + I_ * s3; break;
          default: h0[cc] = s0 - s2;     h1[cc] = s1 - s3;      break;
        }
      }
      mv(u, h0, a0);
      mv(u, h1, a1);
      for (int cc = 0; cc < 3; ++cc) {
        ocplx x0 = ph * a0[cc], x1 = ph * a1[cc];
        d[cc] += x0;
        d[3 + cc] += x1;
        switch (mu) {
          case 0: d[6 + cc] += I_ * x1; d[9 + cc] += I_ * x0; break;
          case 1: d[6 + cc] -= x1;      d[9 + cc] += x0;      break;
          case 2: d[6 + cc] += I_ * x0; d[9 + cc] -= I_ * x1; break;
          default: d[6 + cc] -= x0;     d[9 + cc] -= x1;      break;
        }
      }
    }
  }
  for (int k = 0; k < 12; ++k) dst[k] = d[k];
}

void or_hop(ocplx *out, const ocplx *in, const ocplx *U, const int L[4], int parity,
            int tb) {
  geom g = mkgeom(L);
#pragma omp parallel for schedule(static)
  for (long i = 0; i < g.vh; ++i) hop_site(out + i * 12, in, U, &g, parity, i, tb);
}

void or_hop_full(ocplx *out, const ocplx *in, const ocplx *U, const int L[4], int tb) {
  geom g = mkgeom(L);
  long hs = g.vh * 12;
  or_hop(out, in + hs, U, L, 0, tb);
  or_hop(out + hs, in, U, L, 1, tb);
}

/* ------------------------------------------------------------------------ */
/* clover                                                                    */
/* ------------------------------------------------------------------------ */
static void mat3_dag(const ocplx *a, ocplx *b) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) b[i * 3 + j] = conj(a[j * 3 + i]);
}

/* full 3x3 link (12-real reconstruction, as every reference load) at site c */
static void get_link(const ocplx *U, const geom *g, int mu, const int c[4], ocplx u[9]) {
  int p;
  long i = encode(g, c, &p);
  recon12(link_ptr(U, g, mu, p, i), u);
}

static void shift(const geom *g, const int c[4], int mu, int dmu, int nu, int dnu, int out[4]) {
  for (int k = 0; k < 4; ++k) out[k] = c[k];
  if (dmu) { out[mu] = (out[mu] + dmu + g->L[mu]) % g->L[mu]; }
  if (dnu) { out[nu] = (out[nu] + dnu + g->L[nu]) % g->L[nu]; }
}

/* Q_munu(x): the four leaves in the order of src/clover_dslash.cu:46-160 */
static void clover_leaf_sum(const ocplx *U, const geom *g, const int c[4], int mu, int nu,
                            ocplx Q[9]) {
  ocplx a[9], b[9], t1[9], t2[9], t3[9];
  int s[4];
  for (int k = 0; k < 9; ++k) Q[k] = 0;
  /* U_mu(x) U_nu(x+mu) U_mu(x+nu)^dag U_nu(x)^dag */
  get_link(U, g, mu, c, a);
  shift(g, c, mu, 1, nu, 0, s); get_link(U, g, nu, s, b);
  mat3_mul(a, b, t1);
  shift(g, c, mu, 0, nu, 1, s); get_link(U, g, mu, s, a); mat3_dag(a, b);
  mat3_mul(t1, b, t2);
  get_link(U, g, nu, c, a); mat3_dag(a, b);
  mat3_mul(t2, b, t3);
  for (int k = 0; k < 9; ++k) Q[k] += t3[k];
  /* U_nu(x) U_mu(x-mu+nu)^dag U_nu(x-mu)^dag U_mu(x-mu) */
  get_link(U, g, nu, c, a);
  shift(g, c, mu, -1, nu, 1, s); get_link(U, g, mu, s, t1); mat3_dag(t1, b);
  mat3_mul(a, b, t1);
  shift(g, c, mu, -1, nu, 0, s); get_link(U, g, nu, s, a); mat3_dag(a, b);
  mat3_mul(t1, b, t2);
  shift(g, c, mu, -1, nu, 0, s); get_link(U, g, mu, s, a);
  mat3_mul(t2, a, t3);
  for (int k = 0; k < 9; ++k) Q[k] += t3[k];
  /* U_mu(x-mu)^dag U_nu(x-mu-nu)^dag U_mu(x-mu-nu) U_nu(x-nu) */
  shift(g, c, mu, -1, nu, 0, s); get_link(U, g, mu, s, a); mat3_dag(a, t1);
  shift(g, c, mu, -1, nu, -1, s); get_link(U, g, nu, s, a); mat3_dag(a, b);
  mat3_mul(t1, b, t2);
  shift(g, c, mu, -1, nu, -1, s); get_link(U, g, mu, s, a);
  mat3_mul(t2, a, t1);
  shift(g, c, mu, 0, nu, -1, s); get_link(U, g, nu, s, a);
  mat3_mul(t1, a, t3);
  for (int k = 0; k < 9; ++k) Q[k] += t3[k];
  /* U_nu(x-nu)^dag U_mu(x-nu) U_nu(x+mu-nu) U_mu(x)^dag */
  shift(g, c, mu, 0, nu, -1, s); get_link(U, g, nu, s, a); mat3_dag(a, t1);
  shift(g, c, mu, 0, nu, -1, s); get_link(U, g, mu, s, a);
  mat3_mul(t1, a, t2);
  shift(g, c, mu, 1, nu, -1, s); get_link(U, g, nu, s, a);
  mat3_mul(t2, a, t1);
  get_link(U, g, mu, c, a); mat3_dag(a, b);
  mat3_mul(t1, b, t3);
  for (int k = 0; k < 9; ++k) Q[k] += t3[k];
}

/* spin-block sign table: plane -> 4 entries (s0, s1, factor)
 * transcribed from src/clover_dslash.cu:161-172,288-300,416-428,548-560,
 * 680-692,812-824 (planes XY, XZ, XT, YZ, YT, ZT). */
typedef struct { int s0, s1; double re, im; } blk;
static const int planes[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
static const blk sign_table[6][4] = {
    {{0, 0, 0, -1}, {1, 1, 0, 1}, {2, 2, 0, -1}, {3, 3, 0, 1}},   /* XY */
    {{0, 1, -1, 0}, {1, 0, 1, 0}, {2, 3, -1, 0}, {3, 2, 1, 0}},   /* XZ */
    {{0, 1, 0, 1}, {1, 0, 0, 1}, {2, 3, 0, -1}, {3, 2, 0, -1}},   /* XT */
    {{0, 1, 0, -1}, {1, 0, 0, -1}, {2, 3, 0, -1}, {3, 2, 0, -1}}, /* YZ */
    {{0, 1, -1, 0}, {1, 0, 1, 0}, {2, 3, 1, 0}, {3, 2, -1, 0}},   /* YT */
    {{0, 0, 0, 1}, {1, 1, 0, -1}, {2, 2, 0, -1}, {3, 3, 0, 1}},   /* ZT */
};

void or_clover_build(ocplx *A, const ocplx *U, const int L[4], int parity, double coeff) {
  geom g = mkgeom(L);
#pragma omp parallel for schedule(static)
  for (long i = 0; i < g.vh; ++i) {
    int c[4];
    decode(&g, parity, i, c);
    ocplx cl[144];
    for (int k = 0; k < 144; ++k) cl[k] = 0;
    for (int pl = 0; pl < 6; ++pl) {
      ocplx Q[9];
      clover_leaf_sum(U, &g, c, planes[pl][0], planes[pl][1], Q);
      for (int e = 0; e < 4; ++e) {
        const blk *b = &sign_table[pl][e];
        ocplx f = b->re + I_ * b->im;
        for (int c0 = 0; c0 < 3; ++c0)
          for (int c1 = 0; c1 < 3; ++c1)
            cl[(b->s0 * 3 + c0) * 12 + b->s1 * 3 + c1] += (Q[c0 * 3 + c1] - conj(Q[c1 * 3 + c0])) * f;
      }
    }
    for (int k = 0; k < 144; ++k) cl[k] *= -0.125 * coeff;
    for (int k = 0; k < 12; ++k) cl[k * 13] += 1.0;
    memcpy(A + i * 144, cl, sizeof(cl));
  }
}

/* Gauss-Jordan without pivoting on [A | 1], include/define.h:334-365 */
int or_clover_invert(ocplx *Ainv, const ocplx *A, long nsites) {
  int bad = 0;
#pragma omp parallel for schedule(static) reduction(+ : bad)
  for (long i = 0; i < nsites; ++i) {
    ocplx aug[12 * 24];
    for (int r = 0; r < 12; ++r)
      for (int k = 0; k < 24; ++k)
        aug[r * 24 + k] = (k < 12) ? A[i * 144 + r * 12 + k] : ((k - 12 == r) ? 1.0 : 0.0);
    for (int r = 0; r < 12; ++r) {
      ocplx piv = aug[r * 24 + r];
      if (cabs(piv) == 0.0) { bad++; break; }
      for (int k = 0; k < 24; ++k) aug[r * 24 + k] /= piv;
      for (int j = 0; j < 12; ++j) {
        if (j == r) continue;
        ocplx f = aug[j * 24 + r];
        for (int k = 0; k < 24; ++k) aug[j * 24 + k] -= f * aug[r * 24 + k];
      }
    }
    for (int r = 0; r < 12; ++r)
      for (int k = 0; k < 12; ++k) Ainv[i * 144 + r * 12 + k] = aug[r * 24 + 12 + k];
  }
  return bad;
}

void or_clover_apply(ocplx *out, const ocplx *A, const ocplx *in, long nsites) {
#pragma omp parallel for schedule(static)
  for (long i = 0; i < nsites; ++i) {
    ocplx t[12];
    for (int r = 0; r < 12; ++r) {
      ocplx s = 0;
      for (int k = 0; k < 12; ++k) s += A[i * 144 + r * 12 + k] * in[i * 12 + k];
      t[r] = s;
    }
    memcpy(out + i * 12, t, sizeof(t));
  }
}

/* dslashCloverQcu: out = A_p^-1 D_{p,1-p} in  (src/nccl_clover_dslash.cu:3-79) */
void or_clover_dslash(ocplx *out, const ocplx *in, const ocplx *U, const int L[4], int parity,
                      double coeff, int tb) {
  geom g = mkgeom(L);
  ocplx *A = (ocplx *)malloc(sizeof(ocplx) * 144 * g.vh);
  ocplx *tmp = (ocplx *)malloc(sizeof(ocplx) * 12 * g.vh);
  or_hop(tmp, in, U, L, parity, tb);
  or_clover_build(A, U, L, parity, coeff);
  or_clover_invert(A, A, g.vh);
  or_clover_apply(out, A, tmp, g.vh);
  free(A);
  free(tmp);
}

/* ------------------------------------------------------------------------ */
/* BLAS-1                                                                    */
/* ------------------------------------------------------------------------ */
ocplx or_dot(const ocplx *a, const ocplx *b, long n) {
  double re = 0, im = 0;
#pragma omp parallel for schedule(static) reduction(+ : re, im)
  for (long i = 0; i < n; ++i) {
    ocplx t = conj(a[i]) * b[i];
    re += creal(t);
    im += cimag(t);
  }
  return re + I_ * im;
}
double or_norm2(const ocplx *a, long n) { return creal(or_dot(a, a, n)); }
void or_dot_out(ocplx *out, const ocplx *a, const ocplx *b, long n) { out[0] = or_dot(a, b, n); }

/* ------------------------------------------------------------------------ */
/* full-lattice operator M = A - kappa D  (A = 1 for Wilson)                 */
/* Afull: clover (not inverted) for both parities, [2][Vh][144], or NULL     */
/* ------------------------------------------------------------------------ */
void or_mat_full(ocplx *out, const ocplx *in, const ocplx *U, const ocplx *Afull, const int L[4],
                 double kappa, int tb) {
  geom g = mkgeom(L);
  long n = 2 * g.vh * 12;
  ocplx *d = (ocplx *)malloc(sizeof(ocplx) * n);
  or_hop_full(d, in, U, L, tb);
  if (Afull) {
    or_clover_apply(out, Afull, in, 2 * g.vh);
  } else {
    memcpy(out, in, sizeof(ocplx) * n);
  }
#pragma omp parallel for schedule(static)
  for (long k = 0; k < n; ++k) out[k] -= kappa * d[k];
  free(d);
}

/* ------------------------------------------------------------------------ */
/* even-odd BiCGSTAB                                                         */
/* Schur system on odd sites (the reference solves for x_o, src/bistabcg.cu) */
/*   Mhat x_o = x_o - kappa^2 Ao^-1 D_oe Ae^-1 D_eo x_o                      */
/*   bhat     = Ao^-1 (b_o + kappa D_oe Ae^-1 b_e)                           */
/*   x_e      = Ae^-1 (b_e + kappa D_eo x_o)                                 */
/* With csw = 0 (Ae = Ao = 1) these are include/lattice_bistabcg.h:99-132.  */
/* ------------------------------------------------------------------------ */
typedef struct {
  const ocplx *U;
  const int *L;
  geom g;
  double kappa;
  int tb;
  ocplx *Ainv[2]; /* NULL for Wilson */
  ocplx *tmp;
} eo_op;

static void eo_apply_cd(eo_op *op, ocplx *out, const ocplx *in, int parity) {
  or_hop(out, in, op->U, op->L, parity, op->tb);
  if (op->Ainv[parity]) or_clover_apply(out, op->Ainv[parity], out, op->g.vh);
}

static void eo_mhat(eo_op *op, ocplx *out, const ocplx *in) {
  long n = op->g.vh * 12;
  eo_apply_cd(op, op->tmp, in, 0);
  eo_apply_cd(op, out, op->tmp, 1);
  double k2 = op->kappa * op->kappa;
#pragma omp parallel for schedule(static)
  for (long k = 0; k < n; ++k) out[k] = in[k] - k2 * out[k];
}

static ocplx *alloc_c(long n) { return (ocplx *)calloc((size_t)n, sizeof(ocplx)); }

double or_true_residual(const ocplx *x, const ocplx *b, const ocplx *U, const int L[4],
                        double kappa, double csw, int tb) {
  geom g = mkgeom(L);
  long n = 2 * g.vh * 12;
  ocplx *Af = NULL;
  if (csw != 0.0) {
    Af = alloc_c(2 * g.vh * 144);
    or_clover_build(Af, U, L, 0, kappa * csw);
    or_clover_build(Af + g.vh * 144, U, L, 1, kappa * csw);
  }
  ocplx *r = alloc_c(n);
  or_mat_full(r, x, U, Af, L, kappa, tb);
  double num = 0, den = 0;
#pragma omp parallel for schedule(static) reduction(+ : num, den)
  for (long k = 0; k < n; ++k) {
    ocplx d = b[k] - r[k];
    num += creal(d) * creal(d) + cimag(d) * cimag(d);
    den += creal(b[k]) * creal(b[k]) + cimag(b[k]) * cimag(b[k]);
  }
  free(r);
  free(Af);
  return den > 0 ? sqrt(num / den) : sqrt(num);
}

static double wall(void) {
#ifdef _OPENMP
  return omp_get_wtime();
#else
  return 0.0;
#endif
}

int or_bicgstab(ocplx *x, const ocplx *b, const ocplx *U, const int L[4], or_solve_param *prm) {
  double t0 = wall();
  eo_op op;
  op.U = U;
  op.L = L;
  op.g = mkgeom(L);
  op.kappa = prm->kappa;
  op.tb = prm->t_boundary;
  long vh = op.g.vh, n = vh * 12;
  op.Ainv[0] = op.Ainv[1] = NULL;
  if (prm->csw != 0.0) {
    for (int p = 0; p < 2; ++p) {
      op.Ainv[p] = alloc_c(vh * 144);
      or_clover_build(op.Ainv[p], U, L, p, prm->kappa * prm->csw);
      or_clover_invert(op.Ainv[p], op.Ainv[p], vh);
    }
  }
  op.tmp = alloc_c(n);
  const ocplx *b_e = b, *b_o = b + n;
  ocplx *x_e = x, *x_o = x + n;
  ocplx *bh = alloc_c(n), *r = alloc_c(n), *rt = alloc_c(n), *p = alloc_c(n), *v = alloc_c(n),
        *s = alloc_c(n), *t = alloc_c(n), *w = alloc_c(n);
  /* bhat = Ao^-1 (b_o + kappa D_oe Ae^-1 b_e) */
  if (op.Ainv[0]) or_clover_apply(w, op.Ainv[0], b_e, vh); else memcpy(w, b_e, sizeof(ocplx) * n);
  or_hop(bh, w, U, L, 1, op.tb);
  for (long k = 0; k < n; ++k) bh[k] = b_o[k] + prm->kappa * bh[k];
  if (op.Ainv[1]) or_clover_apply(bh, op.Ainv[1], bh, vh);
  /* zero initial guess (the reference draws a random one, src/lattice_cuda.cu:3-14) */
  memset(x_o, 0, sizeof(ocplx) * n);
  memcpy(r, bh, sizeof(ocplx) * n);
  memcpy(rt, r, sizeof(ocplx) * n);
  ocplx rho_prev = 1, alpha = 1, omega = 1, rho, beta;
  double bnorm = sqrt(or_norm2(bh, n));
  double t1 = wall();
  prm->setup_s = t1 - t0;
  int it;
  for (it = 0; it < prm->max_iter; ++it) {
    rho = or_dot(rt, r, n);
    beta = (rho / rho_prev) * (alpha / omega);
    rho_prev = rho;
    for (long k = 0; k < n; ++k) p[k] = r[k] + (p[k] - v[k] * omega) * beta;
    double rr = or_norm2(r, n);
    int done = (prm->mode == 0) ? (rr < prm->tol) : (sqrt(rr) < prm->tol * bnorm);
    if (done) break;
    eo_mhat(&op, v, p);
    alpha = rho / or_dot(rt, v, n);
    for (long k = 0; k < n; ++k) s[k] = r[k] - v[k] * alpha;
    eo_mhat(&op, t, s);
    omega = or_dot(t, s, n) / or_dot(t, t, n);
    for (long k = 0; k < n; ++k) {
      x_o[k] = x_o[k] + p[k] * alpha + s[k] * omega;
      r[k] = s[k] - t[k] * omega;
    }
  }
  prm->iters = it;
  prm->iter_s = wall() - t1;
  /* Schur residual recomputed from x_o */
  eo_mhat(&op, w, x_o);
  double num = 0;
  for (long k = 0; k < n; ++k) {
    ocplx d = bh[k] - w[k];
    num += creal(d) * creal(d) + cimag(d) * cimag(d);
  }
  prm->schur_res = bnorm > 0 ? sqrt(num) / bnorm : sqrt(num);
  /* x_e = Ae^-1 (b_e + kappa D_eo x_o) */
  or_hop(w, x_o, U, L, 0, op.tb);
  for (long k = 0; k < n; ++k) x_e[k] = b_e[k] + prm->kappa * w[k];
  if (op.Ainv[0]) or_clover_apply(x_e, op.Ainv[0], x_e, vh);
  prm->true_res = or_true_residual(x, b, U, L, prm->kappa, prm->csw, prm->t_boundary);
  free(bh); free(r); free(rt); free(p); free(v); free(s); free(t); free(w);
  free(op.tmp);
  free(op.Ainv[0]);
  free(op.Ainv[1]);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* learned gauge-covariant network on the full lattice                       */
/* layer l maps chan[l] -> chan[l+1] channels:                               */
/*   out_o(x) = sum_i sum_mu [ A^{oi}_mu U_mu(x) in_i(x+mu)                  */
/*                          + B^{oi}_mu U_mu(x-mu)^dag in_i(x-mu) ]          */
/*            + sum_i C^{oi} in_i(x)                                         */
/* A, B, C: 4x4 complex spin matrices acting on each colour.                 */
/* transport T_{2mu} psi(x) = U_mu(x) psi(x+mu); T_{2mu+1} = adjoint of it.  */
/* ------------------------------------------------------------------------ */
long or_net_weight_count(int nl, const int *ch) {
  long n = 0;
  for (int l = 0; l < nl; ++l) n += 9L * ch[l] * ch[l + 1] * 16;
  return n;
}

static void transport(ocplx *dst, const ocplx *src, int term, const ocplx *U, const geom *g,
                      int tb) {
  long vh = g->vh;
  if (term == 8) {
    memcpy(dst, src, sizeof(ocplx) * 24 * vh);
    return;
  }
  int mu = term >> 1, fwd = !(term & 1);
#pragma omp parallel for schedule(static)
  for (long k = 0; k < 2 * vh; ++k) {
    int p = (int)(k / vh);
    long i = k % vh;
    int c[4], q;
    decode(g, p, i, c);
    long j = neighbour(g, c, mu, fwd ? 1 : -1, &q);
    ocplx u[9];
    double ph;
    if (fwd) {
      recon12(link_ptr(U, g, mu, p, i), u);
      ph = (mu == 3 && c[3] == g->L[3] - 1) ? (double)tb : 1.0;
    } else {
      recon12(link_ptr(U, g, mu, q, j), u);
      ph = (mu == 3 && c[3] == 0) ? (double)tb : 1.0;
    }
    const ocplx *s = src + ((long)q * vh + j) * 12;
    ocplx *d = dst + k * 12;
    for (int sp = 0; sp < 4; ++sp) {
      ocplx o[3];
      if (fwd) mv(u, s + sp * 3, o); else mvd(u, s + sp * 3, o);
      for (int cc = 0; cc < 3; ++cc) d[sp * 3 + cc] = ph * o[cc];
    }
  }
}

static long layer_woff(int l, const int *ch) {
  long n = 0;
  for (int k = 0; k < l; ++k) n += 9L * ch[k] * ch[k + 1] * 16;
  return n;
}

static void net_layer(ocplx *out, const ocplx *in, const ocplx *W, int cin, int cout,
                      const ocplx *U, const geom *g, int tb, ocplx *tmp) {
  long fs = 2 * g->vh * 12;
  memset(out, 0, sizeof(ocplx) * fs * cout);
  for (int t = 0; t < 9; ++t)
    for (int i = 0; i < cin; ++i) {
      transport(tmp, in + i * fs, t, U, g, tb);
      for (int o = 0; o < cout; ++o) {
        const ocplx *w = W + (((long)t * cout + o) * cin + i) * 16;
        ocplx *dst = out + o * fs;
#pragma omp parallel for schedule(static)
        for (long k = 0; k < 2 * g->vh; ++k)
          for (int s = 0; s < 4; ++s)
            for (int cc = 0; cc < 3; ++cc) {
              ocplx a = 0;
              for (int s2 = 0; s2 < 4; ++s2) a += w[s * 4 + s2] * tmp[k * 12 + s2 * 3 + cc];
              dst[k * 12 + s * 3 + cc] += a;
            }
      }
    }
}

void or_net_forward(ocplx *out, const ocplx *in, const ocplx *W, int nl, const int *ch,
                    const ocplx *U, const int L[4], int tb, ocplx *acts) {
  geom g = mkgeom(L);
  long fs = 2 * g.vh * 12;
  ocplx *tmp = alloc_c(fs);
  /* acts holds the input of every layer: layer l's block has ch[l] fields */
  long off = 0;
  memcpy(acts, in, sizeof(ocplx) * fs * ch[0]);
  for (int l = 0; l < nl; ++l) {
    ocplx *a_in = acts + off;
    ocplx *a_out = (l + 1 < nl) ? acts + off + fs * ch[l] : out;
    net_layer(a_out, a_in, W + layer_woff(l, ch), ch[l], ch[l + 1], U, &g, tb, tmp);
    off += fs * ch[l];
  }
  free(tmp);
}

/* Gradients as g_z = dL/dRe z + i dL/dIm z for a real loss L.              */
/* For out = W v:  g_W += g_out v^dag (summed over sites and colour),       */
/*                 g_v  = W^dag g_out;  transports: T_{2mu}^dag = T_{2mu+1}. */
void or_net_backward(ocplx *grad_in, ocplx *gradW, const ocplx *grad_out, const ocplx *W, int nl,
                     const int *ch, const ocplx *U, const int L[4], int tb, const ocplx *acts) {
  geom g = mkgeom(L);
  long fs = 2 * g.vh * 12;
  int cmax = 0;
  for (int l = 0; l <= nl; ++l) cmax = ch[l] > cmax ? ch[l] : cmax;
  ocplx *gcur = alloc_c(fs * cmax), *gnext = alloc_c(fs * cmax), *tmp = alloc_c(fs),
        *tmp2 = alloc_c(fs);
  memcpy(gcur, grad_out, sizeof(ocplx) * fs * ch[nl]);
  long aoff[64];
  aoff[0] = 0;
  for (int l = 1; l < nl; ++l) aoff[l] = aoff[l - 1] + fs * ch[l - 1];
  for (int l = nl - 1; l >= 0; --l) {
    int cin = ch[l], cout = ch[l + 1];
    const ocplx *a_in = acts + aoff[l];
    const ocplx *Wl = W + layer_woff(l, ch);
    ocplx *gW = gradW + layer_woff(l, ch);
    memset(gnext, 0, sizeof(ocplx) * fs * cin);
    for (int t = 0; t < 9; ++t) {
      int tadj = (t == 8) ? 8 : (t ^ 1);
      for (int i = 0; i < cin; ++i) {
        transport(tmp, a_in + i * fs, t, U, &g, tb);
        for (int o = 0; o < cout; ++o) {
          const ocplx *go = gcur + o * fs;
          ocplx acc[16];
          for (int q = 0; q < 16; ++q) {
            int s = q / 4, s2 = q % 4;
            double re = 0, im = 0;
#pragma omp parallel for schedule(static) reduction(+ : re, im)
            for (long k = 0; k < 2 * g.vh; ++k)
              for (int cc = 0; cc < 3; ++cc) {
                ocplx v = go[k * 12 + s * 3 + cc] * conj(tmp[k * 12 + s2 * 3 + cc]);
                re += creal(v);
                im += cimag(v);
              }
            acc[q] = re + I_ * im;
          }
          for (int q = 0; q < 16; ++q) gW[(((long)t * cout + o) * cin + i) * 16 + q] = acc[q];
        }
        /* g_in_i += T_t^dag ( sum_o W^{oi dag} g_o ) */
        memset(tmp2, 0, sizeof(ocplx) * fs);
        for (int o = 0; o < cout; ++o) {
          const ocplx *w = Wl + (((long)t * cout + o) * cin + i) * 16;
          const ocplx *go = gcur + o * fs;
#pragma omp parallel for schedule(static)
          for (long k = 0; k < 2 * g.vh; ++k)
            for (int s = 0; s < 4; ++s)
              for (int cc = 0; cc < 3; ++cc) {
                ocplx a = 0;
                for (int s2 = 0; s2 < 4; ++s2) a += conj(w[s2 * 4 + s]) * go[k * 12 + s2 * 3 + cc];
                tmp2[k * 12 + s * 3 + cc] += a;
              }
        }
        transport(tmp, tmp2, tadj, U, &g, tb);
        ocplx *gi = gnext + i * fs;
#pragma omp parallel for schedule(static)
        for (long k = 0; k < fs; ++k) gi[k] += tmp[k];
      }
    }
    ocplx *sw = gcur; gcur = gnext; gnext = sw;
  }
  if (grad_in) memcpy(grad_in, gcur, sizeof(ocplx) * fs * ch[0]);
  free(gcur); free(gnext); free(tmp); free(tmp2);
}
