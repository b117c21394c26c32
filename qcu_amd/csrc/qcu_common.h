// qcu_common.h -- shared device types for the MI355X (gfx950) hot path.
//
// Internal field layouts (all site-fastest so that a wavefront of 64 lanes
// reading 64 consecutive sites issues one contiguous 1 KiB (FP64) / 1 KiB
// (FP32 float4) / 1 KiB (FP16 uint4) transaction per load instruction):
//
//   FP64 spinor (one parity)  double2 [12 sc][Vh]
//   FP32 spinor               float4  [6 sc-pairs][Vh]
//   FP16 spinor               uint4   [3 x 8 halves][Vh] + float norm[Vh]
//   FP64 links (compressed)   double2 [2 p][4 mu][6 (rows 0,1)][Vh]
//   FP32 links                float4  [2][4][3][Vh]
//   FP16 links                uint4   [2][4][Vh] (halves 0..7) + uint2 [2][4][Vh]
//   clover inverse (packed)   36 reals per chiral block, 2 blocks per site:
//                             FP64 double2 [2 p][36][Vh], FP32 float4 [2][18][Vh],
//                             FP16 uint4 [2][9][Vh] + float norm[2][Vh]
//
// The half-site index is i = ((t*Lz + z)*Ly + y)*(Lx/2) + x/2 with the full
// x coordinate 2*(x/2) + ((y+z+t+parity)&1), as qcu's kernels decode it
// (src/wilson_dslash.cu:12-30).
#pragma once
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#define QCU_HD __host__ __device__ __forceinline__
#define QCU_D __device__ __forceinline__

namespace qcu {

template <class T>
struct cplx {
  T re, im;
};

template <class T>
QCU_HD cplx<T> mk(T r, T i) {
  cplx<T> c;
  c.re = r;
  c.im = i;
  return c;
}
template <class T>
QCU_HD cplx<T> operator+(cplx<T> a, cplx<T> b) { return mk(a.re + b.re, a.im + b.im); }
template <class T>
QCU_HD cplx<T> operator-(cplx<T> a, cplx<T> b) { return mk(a.re - b.re, a.im - b.im); }
template <class T>
QCU_HD cplx<T> operator*(cplx<T> a, cplx<T> b) {
  return mk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
template <class T>
QCU_HD cplx<T> operator*(T s, cplx<T> a) { return mk(s * a.re, s * a.im); }
template <class T>
QCU_HD cplx<T> conj(cplx<T> a) { return mk(a.re, -a.im); }
// i*a and -i*a
template <class T>
QCU_HD cplx<T> mul_i(cplx<T> a) { return mk(-a.im, a.re); }
template <class T>
QCU_HD cplx<T> mul_mi(cplx<T> a) { return mk(a.im, -a.re); }
// a += b*c ; a += conj(b)*c
template <class T>
QCU_HD void cmad(cplx<T> &a, cplx<T> b, cplx<T> c) {
  a.re += b.re * c.re - b.im * c.im;
  a.im += b.re * c.im + b.im * c.re;
}
template <class T>
QCU_HD void cmadj(cplx<T> &a, cplx<T> b, cplx<T> c) {
  a.re += b.re * c.re + b.im * c.im;
  a.im += b.re * c.im - b.im * c.re;
}
template <class T>
QCU_HD T norm2(cplx<T> a) { return a.re * a.re + a.im * a.im; }

// ---------------------------------------------------------------------------
// geometry
// ---------------------------------------------------------------------------
struct Geom {
  int lxh, ly, lz, lt;  // half x extent and full y, z, t
  int vh;               // half volume
  int tbc;              // +1 periodic, -1 antiperiodic time boundary
};

struct Site {
  int xh, y, z, t, xodd;  // xodd: full x coordinate is odd
};

QCU_D Site decode(const Geom &g, int i, int parity) {
  Site s;
  int r = i / g.lxh;
  s.xh = i - r * g.lxh;
  int r2 = r / g.ly;
  s.y = r - r2 * g.ly;
  s.t = r2 / g.lz;
  s.z = r2 - s.t * g.lz;
  s.xodd = (s.y + s.z + s.t + parity) & 1;
  return s;
}

// Loads and stores address a wavefront-uniform base (component k of a field,
// kept in scalar registers) plus a 32-bit per-lane byte offset, so that the
// compiler can use the scalar-base / 32-bit-offset form of global memory
// instructions instead of a 64-bit address per load and per component.
template <class V>
QCU_D V ld_at(const void *base, size_t comp_bytes, unsigned lane_off) {
  return *(const V *)((const char *)base + comp_bytes + lane_off);
}
template <class V>
QCU_D void st_at(void *base, size_t comp_bytes, unsigned lane_off, const V &v) {
  *(V *)((char *)base + comp_bytes + lane_off) = v;
}
// the same as a non-temporal store: for fields written once and read by a
// later kernel (the network's activations and term fields), which then need
// not displace the lines that the writing kernel still reads
QCU_D void st_nt(void *base, size_t comp_bytes, unsigned lane_off, const float4 &v) {
  typedef float f4 __attribute__((ext_vector_type(4)));
  const f4 w = {v.x, v.y, v.z, v.w};
  __builtin_nontemporal_store(w, (f4 *)((char *)base + comp_bytes + lane_off));
}
// a non-temporal load: for data the kernel reads once, which then need not
// displace the lines it reads again
QCU_D float4 ld_nt(const void *base, size_t comp_bytes, unsigned lane_off) {
  typedef float f4 __attribute__((ext_vector_type(4)));
  const f4 w = __builtin_nontemporal_load((const f4 *)((const char *)base + comp_bytes + lane_off));
  return make_float4(w.x, w.y, w.z, w.w);
}

// ---------------------------------------------------------------------------
// precision tags
// ---------------------------------------------------------------------------
enum Prec { P64 = 0, P32 = 1, P16 = 2 };

template <int P>
struct real_of {
  typedef float type;
};
template <>
struct real_of<P64> {
  typedef double type;
};

// Spinor field of one parity (or a slice of a batch) in internal storage.
struct SpinorPtr {
  void *p;      // payload
  float *norm;  // FP16 only
  int stride;   // sites between consecutive components (>= vh)
};

template <int P>
struct SpinorIO;

template <>
struct SpinorIO<P64> {
  typedef double real;
  static QCU_D void load(const SpinorPtr &f, int i, cplx<double> (&v)[12]) {
    const unsigned off = (unsigned)i * 16u;
#pragma unroll
    for (int k = 0; k < 12; ++k) {
      double2 a = ld_at<double2>(f.p, (size_t)k * f.stride * 16, off);
      v[k] = mk(a.x, a.y);
    }
  }
  static QCU_D void store(const SpinorPtr &f, int i, const cplx<double> (&v)[12]) {
    const unsigned off = (unsigned)i * 16u;
#pragma unroll
    for (int k = 0; k < 12; ++k) st_at<double2>(f.p, (size_t)k * f.stride * 16, off, make_double2(v[k].re, v[k].im));
  }
};

template <>
struct SpinorIO<P32> {
  typedef float real;
  static QCU_D void load(const SpinorPtr &f, int i, cplx<float> (&v)[12]) {
    const unsigned off = (unsigned)i * 16u;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      float4 a = ld_at<float4>(f.p, (size_t)k * f.stride * 16, off);
      v[2 * k] = mk(a.x, a.y);
      v[2 * k + 1] = mk(a.z, a.w);
    }
  }
  static QCU_D void store(const SpinorPtr &f, int i, const cplx<float> (&v)[12]) {
    const unsigned off = (unsigned)i * 16u;
#pragma unroll
    for (int k = 0; k < 6; ++k)
      st_at<float4>(f.p, (size_t)k * f.stride * 16, off,
                    make_float4(v[2 * k].re, v[2 * k].im, v[2 * k + 1].re, v[2 * k + 1].im));
  }
};

QCU_D float h2f(unsigned short h) { return __half2float(__ushort_as_half(h)); }
QCU_D unsigned short f2h(float f) { return __half_as_ushort(__float2half_rn(f)); }

template <>
struct SpinorIO<P16> {
  typedef float real;
  static QCU_D void load(const SpinorPtr &f, int i, cplx<float> (&v)[12]) {
    const unsigned off = (unsigned)i * 16u;
    const float nrm = ld_at<float>(f.norm, 0, (unsigned)i * 4u);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      uint4 a = ld_at<uint4>(f.p, (size_t)k * f.stride * 16, off);
      unsigned w[4] = {a.x, a.y, a.z, a.w};
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        v[4 * k + q] = mk(nrm * h2f((unsigned short)(w[q] & 0xffffu)),
                          nrm * h2f((unsigned short)(w[q] >> 16)));
      }
    }
  }
  static QCU_D void store(const SpinorPtr &f, int i, const cplx<float> (&v)[12]) {
    float m = 0.f;
#pragma unroll
    for (int k = 0; k < 12; ++k) m = fmaxf(m, fmaxf(fabsf(v[k].re), fabsf(v[k].im)));
    float inv = m > 0.f ? 1.0f / m : 0.f;
    const unsigned off = (unsigned)i * 16u;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      unsigned w[4];
#pragma unroll
      for (int q = 0; q < 4; ++q)
        w[q] = (unsigned)f2h(v[4 * k + q].re * inv) | ((unsigned)f2h(v[4 * k + q].im * inv) << 16);
      st_at<uint4>(f.p, (size_t)k * f.stride * 16, off, make_uint4(w[0], w[1], w[2], w[3]));
    }
    st_at<float>(f.norm, 0, (unsigned)i * 4u, m);
  }
};

// ---------------------------------------------------------------------------
// links: 6 complex (rows 0 and 1), row 2 reconstructed as conj(row0 x row1),
// the same 12-real compression qcu applies on every load (include/define.h:198-206)
// ---------------------------------------------------------------------------
struct LinkPtr {
  void *a;  // FP64 double2 / FP32 float4 / FP16 uint4 part
  void *b;  // FP16 uint2 part
  int vh;
};

template <class T>
QCU_D void recon_row2(cplx<T> (&u)[9]) {
  u[6] = conj(u[1] * u[5] - u[2] * u[4]);
  u[7] = conj(u[2] * u[3] - u[0] * u[5]);
  u[8] = conj(u[0] * u[4] - u[1] * u[3]);
}

template <int P>
struct LinkIO;

template <>
struct LinkIO<P64> {
  // (p, mu) slab index pm = p*4 + mu
  static QCU_D void load(const LinkPtr &L, int pm, int i, cplx<double> (&u)[9]) {
    const unsigned off = (unsigned)i * 16u;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      double2 a = ld_at<double2>(L.a, ((size_t)pm * 6 + k) * L.vh * 16, off);
      u[k] = mk(a.x, a.y);
    }
    recon_row2(u);
  }
};

template <>
struct LinkIO<P32> {
  static QCU_D void load(const LinkPtr &L, int pm, int i, cplx<float> (&u)[9]) {
    const unsigned off = (unsigned)i * 16u;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      float4 a = ld_at<float4>(L.a, ((size_t)pm * 3 + k) * L.vh * 16, off);
      u[2 * k] = mk(a.x, a.y);
      u[2 * k + 1] = mk(a.z, a.w);
    }
    recon_row2(u);
  }
};

template <>
struct LinkIO<P16> {
  static QCU_D void load(const LinkPtr &L, int pm, int i, cplx<float> (&u)[9]) {
    uint4 a = ld_at<uint4>(L.a, (size_t)pm * L.vh * 16, (unsigned)i * 16u);
    uint2 b = ld_at<uint2>(L.b, (size_t)pm * L.vh * 8, (unsigned)i * 8u);
    unsigned w[6] = {a.x, a.y, a.z, a.w, b.x, b.y};
#pragma unroll
    for (int k = 0; k < 6; ++k)
      u[k] = mk(h2f((unsigned short)(w[k] & 0xffffu)), h2f((unsigned short)(w[k] >> 16)));
    recon_row2(u);
  }
};

// ---------------------------------------------------------------------------
// packed clover inverse: per chiral block 6 real diagonal entries followed by
// the 15 complex entries of the strict lower triangle, row-major
// ((1,0),(2,0),(2,1),(3,0),...,(5,4)); 36 reals per block, 72 per site.
// ---------------------------------------------------------------------------
struct CloverPtr {
  void *p;
  float *norm;  // FP16 only
  int vh;
};

template <int P>
struct CloverIO;

template <>
struct CloverIO<P64> {
  static QCU_D void load(const CloverPtr &C, int i, double (&c)[72]) {
    const unsigned off = (unsigned)i * 16u;
#pragma unroll
    for (int k = 0; k < 36; ++k) {
      double2 a = ld_at<double2>(C.p, (size_t)k * C.vh * 16, off);
      c[2 * k] = a.x;
      c[2 * k + 1] = a.y;
    }
  }
};
template <>
struct CloverIO<P32> {
  static QCU_D void load(const CloverPtr &C, int i, float (&c)[72]) {
    const unsigned off = (unsigned)i * 16u;
#pragma unroll
    for (int k = 0; k < 18; ++k) {
      float4 a = ld_at<float4>(C.p, (size_t)k * C.vh * 16, off);
      c[4 * k] = a.x;
      c[4 * k + 1] = a.y;
      c[4 * k + 2] = a.z;
      c[4 * k + 3] = a.w;
    }
  }
};
template <>
struct CloverIO<P16> {
  static QCU_D void load(const CloverPtr &C, int i, float (&c)[72]) {
    const unsigned off = (unsigned)i * 16u;
    const float nrm = ld_at<float>(C.norm, 0, (unsigned)i * 4u);
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      uint4 a = ld_at<uint4>(C.p, (size_t)k * C.vh * 16, off);
      unsigned w[4] = {a.x, a.y, a.z, a.w};
#pragma unroll
      for (int q2 = 0; q2 < 4; ++q2) {
        c[8 * k + 2 * q2] = nrm * h2f((unsigned short)(w[q2] & 0xffffu));
        c[8 * k + 2 * q2 + 1] = nrm * h2f((unsigned short)(w[q2] >> 16));
      }
    }
  }
};

// out_b = C_b * v_b for chiral block b (spins {2b, 2b+1}) from its 36 packed reals
template <class T>
QCU_D void clover_apply_block(const T *cb, cplx<T> (&v)[12], int b) {
  cplx<T> o[6];
  const cplx<T> *x = v + 6 * b;
#pragma unroll
  for (int r = 0; r < 6; ++r) {
    cplx<T> acc = mk(cb[r] * x[r].re, cb[r] * x[r].im);
#pragma unroll
    for (int col = 0; col < 6; ++col) {
      if (col < r) {
        int k = 6 + 2 * (r * (r - 1) / 2 + col);
        cmad(acc, mk(cb[k], cb[k + 1]), x[col]);
      } else if (col > r) {
        int k = 6 + 2 * (col * (col - 1) / 2 + r);
        cmadj(acc, mk(cb[k], cb[k + 1]), x[col]);
      }
    }
    o[r] = acc;
  }
#pragma unroll
  for (int r = 0; r < 6; ++r) v[6 * b + r] = o[r];
}

// out = C * v for one site (both chiral blocks)
template <class T>
QCU_D void clover_apply_site(const T (&c)[72], cplx<T> (&v)[12]) {
  clover_apply_block(c, v, 0);
  clover_apply_block(c + 36, v, 1);
}

// load one chiral block (36 reals) of a packed FP64 / FP32 clover field
QCU_D void clover_load_block(const CloverPtr &C, int i, int b, double (&c)[36]) {
  const unsigned off = (unsigned)i * 16u;
#pragma unroll
  for (int k = 0; k < 18; ++k) {
    double2 a = ld_at<double2>(C.p, (size_t)(18 * b + k) * C.vh * 16, off);
    c[2 * k] = a.x;
    c[2 * k + 1] = a.y;
  }
}
QCU_D void clover_load_block(const CloverPtr &C, int i, int b, float (&c)[36]) {
  const unsigned off = (unsigned)i * 16u;
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    float4 a = ld_at<float4>(C.p, (size_t)(9 * b + k) * C.vh * 16, off);
    c[4 * k] = a.x;
    c[4 * k + 1] = a.y;
    c[4 * k + 2] = a.z;
    c[4 * k + 3] = a.w;
  }
}

// XCD-aware block remap: hardware places workgroup b on XCD b % 8; give each
// XCD a contiguous range of logical blocks so neighbouring t-slices share L2.
QCU_D int remap_block(int b, int n, int on) {
  if (!on) return b;
  int q = n >> 3, r = n & 7;
  int k = b & 7, j = b >> 3;
  return k * q + (k < r ? k : r) + j;
}

}  // namespace qcu
