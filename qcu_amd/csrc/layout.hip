// layout.hip -- external (PyQuda) <-> internal field layouts.
//
// External layouts (the C ABI, python/PyQuda/pyquda/field.py:72-78):
//   gauge  complex128 [4 mu][2 p][Vh][3][3]
//   spinor complex128 [Vh][4][3]  (one parity)
// The reference converts in place through a temporary and cuBLAS copies
// (src/lattice_cuda.cu:41-211); here the caller's buffers are only read (or
// only written), and the internal copies stay resident.
#include "qcu_internal.h"

namespace qcu {

// one thread per (p, mu, site): rows 0 and 1 of the 3x3 link
template <int P>
__global__ void __launch_bounds__(256) k_gauge_in(const double2 *ext, Geom g, void *a, void *b) {
  const int vh = g.vh;
  const long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= 8L * vh) return;
  const int pm = (int)(n / vh), i = (int)(n - (long)pm * vh);
  const int mu = pm >> 1, p = pm & 1;  // external order [mu][p]
  const double2 *src = ext + n * 9;
  double2 u[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) u[k] = src[k];
  const int slab = p * 4 + mu;  // internal order [p][mu]
  if (P == P64) {
    double2 *dst = (double2 *)a + (size_t)slab * 6 * vh + i;
#pragma unroll
    for (int k = 0; k < 6; ++k) dst[(size_t)k * vh] = u[k];
  } else if (P == P32) {
    float4 *dst = (float4 *)a + (size_t)slab * 3 * vh + i;
#pragma unroll
    for (int k = 0; k < 3; ++k)
      dst[(size_t)k * vh] = make_float4((float)u[2 * k].x, (float)u[2 * k].y, (float)u[2 * k + 1].x,
                                        (float)u[2 * k + 1].y);
  } else {
    unsigned w[6];
#pragma unroll
    for (int k = 0; k < 6; ++k)
      w[k] = (unsigned)f2h((float)u[k].x) | ((unsigned)f2h((float)u[k].y) << 16);
    ((uint4 *)a)[(size_t)slab * vh + i] = make_uint4(w[0], w[1], w[2], w[3]);
    ((uint2 *)b)[(size_t)slab * vh + i] = make_uint2(w[4], w[5]);
  }
}

int gauge_ext_to_internal(const void *ext, const Geom &g, int prec, void *a, void *b, hipStream_t s) {
  const long n = 8L * g.vh;
  const int bs = 256;
  const int nb = (int)((n + bs - 1) / bs);
  switch (prec) {
    case P64: hipLaunchKernelGGL(k_gauge_in<P64>, dim3(nb), dim3(bs), 0, s, (const double2 *)ext, g, a, b); break;
    case P32: hipLaunchKernelGGL(k_gauge_in<P32>, dim3(nb), dim3(bs), 0, s, (const double2 *)ext, g, a, b); break;
    case P16: hipLaunchKernelGGL(k_gauge_in<P16>, dim3(nb), dim3(bs), 0, s, (const double2 *)ext, g, a, b); break;
    default: set_error(QCU_ERR_ARG, "bad precision"); return QCU_ERR_ARG;
  }
  QCU_HIP(hipGetLastError());
  return 0;
}

// ---- content hash of an external gauge field -------------------------------
// A 64-bit summary of every bit of the field, position-dependent, used by the
// reference-named entry points to keep converted links and the clover term
// while the caller's gauge is unchanged.  Each complex entry k with bit
// patterns (a, b) is mixed with k and the mixed words are added modulo 2^64;
// addition commutes, so the result does not depend on the block order.
QCU_D unsigned long long hash_mix(unsigned long long a, unsigned long long b, unsigned long long k) {
  unsigned long long x = a * 0x9E3779B97F4A7C15ull + b * 0xC2B2AE3D27D4EB4Full + k * 0x165667B19E3779F9ull;
  x ^= x >> 29;
  x *= 0xBF58476D1CE4E5B9ull;
  x ^= x >> 32;
  x *= 0x94D049BB133111EBull;
  return x ^ (x >> 29);
}

constexpr int kHashBlocks = 1024;

__global__ void __launch_bounds__(256) k_gauge_hash(const double2 *ext, long n, unsigned long long *part) {
  unsigned long long h = 0;
  for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x) {
    const double2 v = ext[k];
    h += hash_mix((unsigned long long)__double_as_longlong(v.x), (unsigned long long)__double_as_longlong(v.y),
                  (unsigned long long)k);
  }
  __shared__ unsigned long long sh[256];
  sh[threadIdx.x] = h;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ void __launch_bounds__(256) k_hash_sum(unsigned long long *part, int n) {
  __shared__ unsigned long long sh[256];
  unsigned long long h = 0;
  for (int k = threadIdx.x; k < n; k += blockDim.x) h += part[k];
  sh[threadIdx.x] = h;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[n] = sh[0];
}

// scratch: (kHashBlocks + 1) 64-bit words of device memory
int gauge_hash(const void *ext, const Geom &g, void *scratch, unsigned long long *out, hipStream_t s) {
  const long n = 36L * 2 * g.vh;  // complex entries: [4][2][Vh][3][3]
  unsigned long long *part = (unsigned long long *)scratch;
  hipLaunchKernelGGL(k_gauge_hash, dim3(kHashBlocks), dim3(256), 0, s, (const double2 *)ext, n, part);
  hipLaunchKernelGGL(k_hash_sum, dim3(1), dim3(256), 0, s, part, kHashBlocks);
  QCU_HIP(hipGetLastError());
  QCU_HIP(hipMemcpyAsync(out, part + kHashBlocks, sizeof(*out), hipMemcpyDeviceToHost, s));
  QCU_HIP(hipStreamSynchronize(s));
  return 0;
}

size_t gauge_hash_scratch_bytes() { return (kHashBlocks + 1) * sizeof(unsigned long long); }

// external one-parity field [Vh][4][3] <-> system (segment seg, rhs r) of a batch vector
template <int P>
__global__ void k_spinor_in(const double2 *ext, int vh, int nrhs, int rhs, SpinorPtr dst) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= vh) return;
  typedef typename real_of<P>::type T;
  cplx<T> v[12];
  const double2 *src = ext + (size_t)i * 12;
#pragma unroll
  for (int k = 0; k < 12; ++k) {
    double2 a = src[k];
    v[k] = mk<T>((T)a.x, (T)a.y);
  }
  SpinorIO<P>::store(dst, i * nrhs + rhs, v);
}

template <int P>
__global__ void k_spinor_out(SpinorPtr src, int vh, int nrhs, int rhs, double2 *ext) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= vh) return;
  typedef typename real_of<P>::type T;
  cplx<T> v[12];
  SpinorIO<P>::load(src, i * nrhs + rhs, v);
  double2 *dst = ext + (size_t)i * 12;
#pragma unroll
  for (int k = 0; k < 12; ++k) dst[k] = make_double2((double)v[k].re, (double)v[k].im);
}

int spinor_ext_to_batch(const void *ext, const BatchOp &op, int seg, int rhs, SpinorPtr dst, int prec,
                        hipStream_t s) {
  QCU_CHECK(seg >= 0 && seg < (int)op.segs.size(), QCU_ERR_ARG, "bad segment");
  const SegDesc &d = op.segs[seg];
  QCU_CHECK(rhs >= 0 && rhs < d.nrhs, QCU_ERR_ARG, "bad right-hand side");
  const SpinorPtr v = seg_view(dst, prec, d.voff, d.nthr);
  const int vh = d.g.vh, bs = 256, nb = (vh + bs - 1) / bs;
  const double2 *e = (const double2 *)ext;
  switch (prec) {
    case P64: hipLaunchKernelGGL(k_spinor_in<P64>, dim3(nb), dim3(bs), 0, s, e, vh, d.nrhs, rhs, v); break;
    case P32: hipLaunchKernelGGL(k_spinor_in<P32>, dim3(nb), dim3(bs), 0, s, e, vh, d.nrhs, rhs, v); break;
    case P16: hipLaunchKernelGGL(k_spinor_in<P16>, dim3(nb), dim3(bs), 0, s, e, vh, d.nrhs, rhs, v); break;
    default: set_error(QCU_ERR_ARG, "bad precision"); return QCU_ERR_ARG;
  }
  QCU_HIP(hipGetLastError());
  return 0;
}

int spinor_batch_to_ext(SpinorPtr src, const BatchOp &op, int seg, int rhs, void *ext, int prec, hipStream_t s) {
  QCU_CHECK(seg >= 0 && seg < (int)op.segs.size(), QCU_ERR_ARG, "bad segment");
  const SegDesc &d = op.segs[seg];
  QCU_CHECK(rhs >= 0 && rhs < d.nrhs, QCU_ERR_ARG, "bad right-hand side");
  const SpinorPtr v = seg_view(src, prec, d.voff, d.nthr);
  const int vh = d.g.vh, bs = 256, nb = (vh + bs - 1) / bs;
  double2 *e = (double2 *)ext;
  switch (prec) {
    case P64: hipLaunchKernelGGL(k_spinor_out<P64>, dim3(nb), dim3(bs), 0, s, v, vh, d.nrhs, rhs, e); break;
    case P32: hipLaunchKernelGGL(k_spinor_out<P32>, dim3(nb), dim3(bs), 0, s, v, vh, d.nrhs, rhs, e); break;
    case P16: hipLaunchKernelGGL(k_spinor_out<P16>, dim3(nb), dim3(bs), 0, s, v, vh, d.nrhs, rhs, e); break;
    default: set_error(QCU_ERR_ARG, "bad precision"); return QCU_ERR_ARG;
  }
  QCU_HIP(hipGetLastError());
  return 0;
}

}  // namespace qcu
