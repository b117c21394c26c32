// clover.hip -- clover term A = 1 - coeff/8 sum_{mu<nu} T_munu (Q_munu - Q_munu^dag)
// per site of one parity, and its inverse, stored as two packed Hermitian 6x6
// chiral blocks (spins {0,1} and {2,3}).
//
// Leaves and sign table follow src/clover_dslash.cu:5-837 (T_munu = -i sigma_munu
// in the DeGrand-Rossi basis; the reference hard-codes coeff = kappa*csw = 1 at
// :826-835).  The reference inverts the full 12x12 matrix by Gauss-Jordan
// without pivoting (src/clover_dslash.cu:839-855); A is block diagonal, so the
// two 6x6 blocks are inverted here, with partial pivoting, in FP64, once per
// gauge field instead of on every call.
#include "qcu_internal.h"
#include "qcu_reduce.h"

namespace qcu {

typedef cplx<double> cd;

QCU_D void link_at(const LinkPtr &U, const Geom &g, int mu, int x, int y, int z, int t, cd (&u)[9]) {
  const int lx = 2 * g.lxh;
  x = (x + lx) % lx;
  y = (y + g.ly) % g.ly;
  z = (z + g.lz) % g.lz;
  t = (t + g.lt) % g.lt;
  const int p = (x + y + z + t) & 1;
  const int i = ((t * g.lz + z) * g.ly + y) * g.lxh + (x >> 1);
  LinkIO<P64>::load(U, p * 4 + mu, i, u);
}

QCU_D void mm(const cd (&a)[9], const cd (&b)[9], cd (&c)[9]) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      cd s = mk(0.0, 0.0);
#pragma unroll
      for (int k = 0; k < 3; ++k) cmad(s, a[i * 3 + k], b[k * 3 + j]);
      c[i * 3 + j] = s;
    }
}
// c = a * b^dag
QCU_D void mmd(const cd (&a)[9], const cd (&b)[9], cd (&c)[9]) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      cd s = mk(0.0, 0.0);
#pragma unroll
      for (int k = 0; k < 3; ++k) s = s + a[i * 3 + k] * conj(b[j * 3 + k]);
      c[i * 3 + j] = s;
    }
}
// c = a^dag * b
QCU_D void mdm(const cd (&a)[9], const cd (&b)[9], cd (&c)[9]) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      cd s = mk(0.0, 0.0);
#pragma unroll
      for (int k = 0; k < 3; ++k) cmadj(s, a[k * 3 + i], b[k * 3 + j]);
      c[i * 3 + j] = s;
    }
}

// c = a^dag * b^dag
QCU_D void mdmd(const cd (&a)[9], const cd (&b)[9], cd (&c)[9]) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      cd s = mk(0.0, 0.0);
#pragma unroll
      for (int k = 0; k < 3; ++k) s = s + conj(a[k * 3 + i]) * conj(b[j * 3 + k]);
      c[i * 3 + j] = s;
    }
}

// Q_munu(x): four leaves, counter-clockwise in the (mu, nu) plane
QCU_D void leaf_sum(const LinkPtr &U, const Geom &g, const int (&c)[4], int mu, int nu, cd (&Q)[9]) {
  int e[4][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}};
  auto at = [&](int dm, int dn, int dir, cd (&u)[9]) {
    int x = c[0] + dm * e[mu][0] + dn * e[nu][0];
    int y = c[1] + dm * e[mu][1] + dn * e[nu][1];
    int z = c[2] + dm * e[mu][2] + dn * e[nu][2];
    int t = c[3] + dm * e[mu][3] + dn * e[nu][3];
    link_at(U, g, dir, x, y, z, t, u);
  };
  cd a[9], b[9], t1[9], t2[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) Q[k] = mk(0.0, 0.0);
  // U_mu(x) U_nu(x+mu) U_mu(x+nu)^dag U_nu(x)^dag
  at(0, 0, mu, a);
  at(1, 0, nu, b);
  mm(a, b, t1);
  at(0, 1, mu, a);
  mmd(t1, a, t2);
  at(0, 0, nu, a);
  mmd(t2, a, t1);
#pragma unroll
  for (int k = 0; k < 9; ++k) Q[k] = Q[k] + t1[k];
  // U_nu(x) U_mu(x-mu+nu)^dag U_nu(x-mu)^dag U_mu(x-mu)
  at(0, 0, nu, a);
  at(-1, 1, mu, b);
  mmd(a, b, t1);
  at(-1, 0, nu, a);
  mmd(t1, a, t2);
  at(-1, 0, mu, a);
  mm(t2, a, t1);
#pragma unroll
  for (int k = 0; k < 9; ++k) Q[k] = Q[k] + t1[k];
  // U_mu(x-mu)^dag U_nu(x-mu-nu)^dag U_mu(x-mu-nu) U_nu(x-nu)
  at(-1, 0, mu, a);
  at(-1, -1, nu, b);
  mdmd(a, b, t1);
  at(-1, -1, mu, a);
  mm(t1, a, t2);
  at(0, -1, nu, a);
  mm(t2, a, t1);
#pragma unroll
  for (int k = 0; k < 9; ++k) Q[k] = Q[k] + t1[k];
  // U_nu(x-nu)^dag U_mu(x-nu) U_nu(x+mu-nu) U_mu(x)^dag
  at(0, -1, nu, a);
  at(0, -1, mu, b);
  mdm(a, b, t1);
  at(1, -1, nu, a);
  mm(t1, a, t2);
  at(0, 0, mu, a);
  mmd(t2, a, t1);
#pragma unroll
  for (int k = 0; k < 9; ++k) Q[k] = Q[k] + t1[k];
}

// sign table entries (s0, s1, factor re, im) per plane, src/clover_dslash.cu
// :161-172 (XY), :288-300 (XZ), :416-428 (XT), :548-560 (YZ), :680-692 (YT), :812-824 (ZT)
__constant__ int kPlanes[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
__constant__ signed char kSign[6][4][4] = {
    {{0, 0, 0, -1}, {1, 1, 0, 1}, {2, 2, 0, -1}, {3, 3, 0, 1}},
    {{0, 1, -1, 0}, {1, 0, 1, 0}, {2, 3, -1, 0}, {3, 2, 1, 0}},
    {{0, 1, 0, 1}, {1, 0, 0, 1}, {2, 3, 0, -1}, {3, 2, 0, -1}},
    {{0, 1, 0, -1}, {1, 0, 0, -1}, {2, 3, 0, -1}, {3, 2, 0, -1}},
    {{0, 1, -1, 0}, {1, 0, 1, 0}, {2, 3, 1, 0}, {3, 2, -1, 0}},
    {{0, 0, 0, 1}, {1, 1, 0, -1}, {2, 2, 0, -1}, {3, 3, 0, 1}},
};

// Gauss-Jordan with partial pivoting on a 6x6 complex block, in place
QCU_D int invert6(cd (&m)[36]) {
  cd inv[36];
#pragma unroll
  for (int k = 0; k < 36; ++k) inv[k] = mk((k % 7 == 0) ? 1.0 : 0.0, 0.0);
  for (int col = 0; col < 6; ++col) {
    int piv = col;
    double best = norm2(m[col * 6 + col]);
    for (int r = col + 1; r < 6; ++r) {
      double v = norm2(m[r * 6 + col]);
      if (v > best) {
        best = v;
        piv = r;
      }
    }
    if (!(best > 0)) return 1;
    if (piv != col) {
      for (int k = 0; k < 6; ++k) {
        cd t = m[col * 6 + k];
        m[col * 6 + k] = m[piv * 6 + k];
        m[piv * 6 + k] = t;
        t = inv[col * 6 + k];
        inv[col * 6 + k] = inv[piv * 6 + k];
        inv[piv * 6 + k] = t;
      }
    }
    const cd pv = m[col * 6 + col];
    const double den = norm2(pv);
    const cd ip = mk(pv.re / den, -pv.im / den);
    for (int k = 0; k < 6; ++k) {
      m[col * 6 + k] = m[col * 6 + k] * ip;
      inv[col * 6 + k] = inv[col * 6 + k] * ip;
    }
    for (int r = 0; r < 6; ++r) {
      if (r == col) continue;
      const cd f = m[r * 6 + col];
      for (int k = 0; k < 6; ++k) {
        m[r * 6 + k] = m[r * 6 + k] - f * m[col * 6 + k];
        inv[r * 6 + k] = inv[r * 6 + k] - f * inv[col * 6 + k];
      }
    }
  }
#pragma unroll
  for (int k = 0; k < 36; ++k) m[k] = inv[k];
  return 0;
}

QCU_D void pack_store(double2 *dst, int vh, int i, const cd (&m)[36], int b) {
  // 6 diagonal reals, then the strict lower triangle row by row
  double buf[36];
  for (int r = 0; r < 6; ++r) buf[r] = m[r * 7].re;
  int k = 6;
  for (int r = 1; r < 6; ++r)
    for (int c = 0; c < r; ++c) {
      buf[k++] = m[r * 6 + c].re;
      buf[k++] = m[r * 6 + c].im;
    }
  for (int q = 0; q < 18; ++q) dst[(size_t)(18 * b + q) * vh + i] = make_double2(buf[2 * q], buf[2 * q + 1]);
}

// clov: packed A (may be null), cinv: packed A^-1, both double2 [36][Vh]
__global__ void __launch_bounds__(128) k_clover(LinkPtr U, Geom g, int parity, double coeff, double2 *clov, double2 *cinv,
                         int *bad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.vh) return;
  const Site s = decode(g, i, parity);
  const int c[4] = {2 * s.xh + s.xodd, s.y, s.z, s.t};
  cd blk[2][36];
  for (int b = 0; b < 2; ++b)
    for (int k = 0; k < 36; ++k) blk[b][k] = mk(0.0, 0.0);
  for (int pl = 0; pl < 6; ++pl) {
    cd Q[9];
    leaf_sum(U, g, c, kPlanes[pl][0], kPlanes[pl][1], Q);
    for (int e = 0; e < 4; ++e) {
      const int s0 = kSign[pl][e][0], s1 = kSign[pl][e][1];
      const cd f = mk((double)kSign[pl][e][2], (double)kSign[pl][e][3]);
      const int b = s0 >> 1;
      for (int c0 = 0; c0 < 3; ++c0)
        for (int c1 = 0; c1 < 3; ++c1) {
          cd fq = (Q[c0 * 3 + c1] - conj(Q[c1 * 3 + c0])) * f;
          cd &dst = blk[b][((s0 & 1) * 3 + c0) * 6 + (s1 & 1) * 3 + c1];
          dst = dst + fq;
        }
    }
  }
  const double sc = -0.125 * coeff;
  for (int b = 0; b < 2; ++b) {
    for (int k = 0; k < 36; ++k) blk[b][k] = sc * blk[b][k];
    for (int r = 0; r < 6; ++r) blk[b][r * 7].re += 1.0;
    if (clov) pack_store(clov, g.vh, i, blk[b], b);
    if (invert6(blk[b])) atomicAdd(bad, 1);
    pack_store(cinv, g.vh, i, blk[b], b);
  }
}

// FP64 packed clover -> FP32 / FP16 (+ per-site scale)
__global__ void __launch_bounds__(128) k_clover_cvt(const double2 *src, int vh, int prec, void *dst, float *norm) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= vh) return;
  double v[72];
  for (int k = 0; k < 36; ++k) {
    double2 a = src[(size_t)k * vh + i];
    v[2 * k] = a.x;
    v[2 * k + 1] = a.y;
  }
  if (prec == P32) {
    float4 *d = (float4 *)dst;
    for (int k = 0; k < 18; ++k)
      d[(size_t)k * vh + i] = make_float4((float)v[4 * k], (float)v[4 * k + 1], (float)v[4 * k + 2], (float)v[4 * k + 3]);
  } else {
    double m = 0;
    for (int k = 0; k < 72; ++k) m = fmax(m, fabs(v[k]));
    const double inv = m > 0 ? 1.0 / m : 0.0;
    uint4 *d = (uint4 *)dst;
    for (int k = 0; k < 9; ++k) {
      unsigned w[4];
      for (int q = 0; q < 4; ++q)
        w[q] = (unsigned)f2h((float)(v[8 * k + 2 * q] * inv)) | ((unsigned)f2h((float)(v[8 * k + 2 * q + 1] * inv)) << 16);
      d[(size_t)k * vh + i] = make_uint4(w[0], w[1], w[2], w[3]);
    }
    norm[i] = (float)m;
  }
}

int clover_build(const LinkPtr &U64, const Geom &g, double coeff, int parity, void *clov, void *cinv,
                 int *bad_dev, hipStream_t s) {
  const int bs = 128, nb = (g.vh + bs - 1) / bs;
  hipLaunchKernelGGL(k_clover, dim3(nb), dim3(bs), 0, s, U64, g, parity, coeff, (double2 *)clov,
                     (double2 *)cinv, bad_dev);
  QCU_HIP(hipGetLastError());
  return 0;
}

int clover_convert(const void *c64, const Geom &g, int prec, void *dst, float *norm, hipStream_t s) {
  const int bs = 128, nb = (g.vh + bs - 1) / bs;
  hipLaunchKernelGGL(k_clover_cvt, dim3(nb), dim3(bs), 0, s, (const double2 *)c64, g.vh, prec, dst, norm);
  QCU_HIP(hipGetLastError());
  return 0;
}

}  // namespace qcu
