// gen.hip -- seeded synthetic inputs generated on the device.
//
// Gauge links U = exp(i eps H), H = sum_a omega_a lambda_a (Gell-Mann,
// omega_a ~ N(0,1)), by scaling and squaring of a degree-16 Taylor series;
// Gaussian spinors with real and imaginary parts ~ N(0,1).  Random numbers
// come from a counter-based splitmix64 hash of (seed, stream, draw), so every
// value depends only on its index, never on the launch shape.
#include "qcu_internal.h"

namespace qcu {

QCU_D uint64_t mix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ULL;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

QCU_D double uniform01(uint64_t seed, uint64_t stream, uint64_t k) {
  uint64_t h = mix64(mix64(seed ^ mix64(stream)) + k);
  return ((double)(h >> 11) + 1.0) * (1.0 / 9007199254740992.0);
}

QCU_D void gauss_pair(uint64_t seed, uint64_t stream, uint64_t k, double &g0, double &g1) {
  double u1 = uniform01(seed, stream, 2 * k);
  double u2 = uniform01(seed, stream, 2 * k + 1);
  double r = sqrt(-2.0 * log(u1));
  double a = 6.283185307179586 * u2;
  g0 = r * cos(a);
  g1 = r * sin(a);
}

QCU_D void m3mul(const cplx<double> *a, const cplx<double> *b, cplx<double> *c) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      cplx<double> s = mk(0.0, 0.0);
#pragma unroll
      for (int k = 0; k < 3; ++k) s = s + a[i * 3 + k] * b[k * 3 + j];
      c[i * 3 + j] = s;
    }
}

__global__ void k_gen_gauge(double2 *U, long nlinks, double eps, uint64_t seed) {
  const long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nlinks) return;
  double w[8];
#pragma unroll
  for (int k = 0; k < 4; ++k) gauss_pair(seed, (uint64_t)n, (uint64_t)k, w[2 * k], w[2 * k + 1]);
  const double s3 = 0.57735026918962576451;
  cplx<double> H[9];
  H[0] = mk(w[2] + w[7] * s3, 0.0);
  H[4] = mk(-w[2] + w[7] * s3, 0.0);
  H[8] = mk(-2.0 * w[7] * s3, 0.0);
  H[1] = mk(w[0], -w[1]);
  H[3] = mk(w[0], w[1]);
  H[2] = mk(w[3], -w[4]);
  H[6] = mk(w[3], w[4]);
  H[5] = mk(w[5], -w[6]);
  H[7] = mk(w[5], w[6]);
  cplx<double> A[9];
  double nrm = 0;
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    A[k] = mk(-eps * H[k].im, eps * H[k].re);  // i eps H
    nrm += A[k].re * A[k].re + A[k].im * A[k].im;
  }
  nrm = sqrt(nrm);
  int s = 0;
  while (nrm > 0.25) {
    nrm *= 0.5;
    ++s;
  }
  const double sc = ldexp(1.0, -s);
#pragma unroll
  for (int k = 0; k < 9; ++k) A[k] = sc * A[k];
  cplx<double> E[9], T[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) E[k] = mk((k % 4 == 0) ? 1.0 : 0.0, 0.0);
  for (int m = 16; m >= 1; --m) {
    m3mul(A, E, T);
#pragma unroll
    for (int k = 0; k < 9; ++k) E[k] = mk(((k % 4 == 0) ? 1.0 : 0.0) + T[k].re / m, T[k].im / m);
  }
  for (int q = 0; q < s; ++q) {
    m3mul(E, E, T);
#pragma unroll
    for (int k = 0; k < 9; ++k) E[k] = T[k];
  }
  double2 *dst = U + n * 9;
#pragma unroll
  for (int k = 0; k < 9; ++k) dst[k] = make_double2(E[k].re, E[k].im);
}

__global__ void k_gen_gauss(double2 *v, long n, uint64_t seed) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double a, b;
  gauss_pair(seed, (uint64_t)i, 0, a, b);
  v[i] = make_double2(a, b);
}

int gen_gauge(void *ext, const Geom &g, double eps, uint64_t seed, hipStream_t s) {
  const long n = 8L * g.vh;
  const int bs = 256;
  hipLaunchKernelGGL(k_gen_gauge, dim3((unsigned)((n + bs - 1) / bs)), dim3(bs), 0, s, (double2 *)ext, n,
                     eps, seed);
  QCU_HIP(hipGetLastError());
  return 0;
}

int gen_gauss(void *ext, long n, uint64_t seed, hipStream_t s) {
  const int bs = 256;
  hipLaunchKernelGGL(k_gen_gauss, dim3((unsigned)((n + bs - 1) / bs)), dim3(bs), 0, s, (double2 *)ext, n,
                     seed);
  QCU_HIP(hipGetLastError());
  return 0;
}

}  // namespace qcu
