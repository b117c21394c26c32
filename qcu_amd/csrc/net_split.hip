// net_split.hip -- the hidden layers of the learned preconditioner on the
// FP16 matrix cores at FP32 accuracy (new; no counterpart in the reference).
//
// A hidden layer with 8 input and 4 or 8 output channels is, per colour c,
// the complex GEMM
//   D[m][n] = sum_k W[m][k] V[k][n],  m = (output channel o, spin s),
//   k = (term t, input channel i, spin s2),  n = output site,
// where V holds the input channels transported from the term's neighbour
// (V_t(x) = M_t(x) in(neighbour_t(x)), M_t = +-U or +-U^dag, the identity for
// the site term).  It runs as a real GEMM on v_mfma_f32_32x32x16_f16 with each
// complex product embedded as a 2x2 real block,
//   [Re D]   [Wr  -Wi] [Re V]
//   [Im D] = [Wi   Wr] [Im V],
// and every FP32 operand split into two FP16 numbers, x = hi + lo with hi the
// FP16 rounding of x and lo that of the remainder x - hi, so that
//   W V = Whi Vhi + Whi Vlo + Wlo Vhi + O(2^-22 |W| |V|)
// in three FP16 products accumulated in FP32 (the split-operand method of
// Ootomo & Yokota, "Recovering single precision accuracy from Tensor Cores
// while surpassing the FP32 theoretical peak performance", IJHPCA 36 (2022);
// the lo x lo product is dropped).  FP16 holds 11 significant bits over 30
// binades, so both operands are scaled by exact powers of two first: the
// weights by one factor per layer (largest entry below 2^14, set when they
// are packed), the input field by one factor per application from its largest
// component, which the kernel that wrote the field records; the factors are
// divided out of D.  Values far below the largest lose their lo part to FP16
// underflow: their absolute error stays below 2^-25 of the scaled unit, 2^-39
// of the largest component.
//
// Lane layout of v_mfma_f32_32x32x16_f16 (64 lanes, g = lane / 32):
//   A: lane holds A[m = lane % 32][k = 8 g + q], q = 0..7;
//   B: lane holds B[k = 8 g + q][n = lane % 32];
//   D: 16 floats, D[m = 8 (r / 4) + 4 g + r % 4][n = lane % 32], r = 0..15.
// Index order chosen to match it:
//   M tile mt (32 rows): m = 8 ol + 2 s + ri, output channel o = 4 mt + ol,
//     spin s, ri = 0 (Re) / 1 (Im);
//   K step ks (16): k = 8 il + 2 s2 + ri, input channel i = 2 ks + il.
// So in K step ks lane group g transports input channel 2 ks + g of its site
// (all four spins, one colour per product), and after the last step lane
// (n, g) holds, for each output channel, spins 2g and 2g + 1 of site n: it
// stores float4 planes 3g .. 3g + 2 of the site-fastest FP32 layout.
//
// Work: a wave owns 32 consecutive half-sites of one output parity (a chunk)
// and runs the 9 terms x 4 K steps over them; the packed A operands of the
// layer (9 x MT x 4 x {hi, lo} x 1 KiB = 144 KiB for MT = 2) stay in LDS, so a
// block of 8 waves (two per SIMD) per CU loads them once and then walks a
// sequence of chunks.  The channel operand of the next step is loaded while
// the products of the current one run.
#include <cmath>
#include <cstdlib>

#include "qcu_internal.h"
#include "qcu_net.h"

namespace qcu {

typedef float f16v __attribute__((ext_vector_type(16)));
typedef float f4v __attribute__((ext_vector_type(4)));

// The two MFMA shapes.  NS sites per wave (the N of the product); the 64
// lanes form G = 64 / NS groups, lane group g supplying K rows 8 g .. 8 g + 7.
// NR accumulator registers per lane; row(r, g) the D row a lane's register r
// holds.
template <int NS>
struct SplitShape;
template <>
struct SplitShape<32> {  // v_mfma_f32_32x32x16_f16
  typedef f16v acc_t;
  static constexpr int NR = 16;
  static QCU_D acc_t mfma(h8v a, h8v b, acc_t c) { return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0); }
  static constexpr QCU_HD int row(int r, int g) { return 8 * (r / 4) + 4 * g + r % 4; }
};
template <>
struct SplitShape<16> {  // v_mfma_f32_16x16x32_f16
  typedef f4v acc_t;
  static constexpr int NR = 4;
  static QCU_D acc_t mfma(h8v a, h8v b, acc_t c) { return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0); }
  static constexpr QCU_HD int row(int r, int g) { return 4 * g + r; }
};

// Packs one layer's weights W[t][o][i][4x4] (complex, FP32) into the A
// operands A[t][mt][ks][hl][lane] (8 FP16 each; hl 0 = hi, 1 = lo) of shape
// NS, scaled by a power of two; winv[0] receives its inverse.  One block.
template <int NS, int MT>
__global__ void __launch_bounds__(256) k_net_pack_split(const float2 *W, int cin, int cout, uint4 *Apk, float *winv) {
  constexpr int G = 64 / NS, KS = 8 / G;
  __shared__ float red[256];
  const int tid = threadIdx.x;
  const int nw = 9 * cout * cin * 16;
  float mx = 0.f;
  for (int k = tid; k < nw; k += 256) mx = fmaxf(mx, fmaxf(fabsf(W[k].x), fabsf(W[k].y)));
  red[tid] = mx;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (tid < h) red[tid] = fmaxf(red[tid], red[tid + h]);
    __syncthreads();
  }
  const float sw = pow2_scale(red[0], 14);
  if (tid == 0) winv[0] = 1.f / sw;
  const int total = 9 * MT * KS * 2 * 64;
  for (int e = tid; e < total; e += 256) {
    const int lane = e & 63;
    int r = e >> 6;
    const int hl = r & 1;
    r >>= 1;
    const int ks = r % KS;
    r /= KS;
    const int mt = r % MT, t = r / MT;
    const int m = lane % NS;
    const int o = (NS / 8) * mt + (m >> 3), s = (m >> 1) & 3, ri = m & 1;
    const int i = G * ks + lane / NS;
    h8v out;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int s2 = q >> 1, rin = q & 1;
      float v = 0.f;
      if (o < cout && i < cin) {
        const float2 w = W[(((size_t)t * cout + o) * cin + i) * 16 + s * 4 + s2];
        v = ri == 0 ? (rin == 0 ? w.x : -w.y) : (rin == 0 ? w.y : w.x);
      }
      v *= sw;
      const _Float16 h = (_Float16)v;
      out[q] = hl == 0 ? h : (_Float16)(v - (float)h);
    }
    Apk[e] = __builtin_bit_cast(uint4, out);
  }
}

// Packs the single-output layer's weights W'[t][0][i][4x4] as the A operands
// of the fused epilogue (32-site shape): tile sel 0, 1 hold hop terms 4 sel ..
// 4 sel + 3 (rows m = 8 tl + 2 s + ri), tile 2 the site term (rows 0..7; the
// rest zero); K step ks covers input channels 2 ks, 2 ks + 1 in the order the
// epilogue holds them, k = 8 g + 4 cl + 2 sl + ri: channel 2 ks + cl, spin
// 2 g + sl.  A[sel][ks][hl][lane], 8 FP16 each; winv[0] = inverse scale.
__global__ void __launch_bounds__(256) k_net_pack_fuse(const float2 *W, uint4 *Apk, float *winv) {
  __shared__ float red[256];
  const int tid = threadIdx.x;
  const int nw = 9 * 8 * 16;
  float mx = 0.f;
  for (int k = tid; k < nw; k += 256) mx = fmaxf(mx, fmaxf(fabsf(W[k].x), fabsf(W[k].y)));
  red[tid] = mx;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (tid < h) red[tid] = fmaxf(red[tid], red[tid + h]);
    __syncthreads();
  }
  const float sw = pow2_scale(red[0], 14);
  if (tid == 0) winv[0] = 1.f / sw;
  for (int e = tid; e < 3 * 4 * 2 * 64; e += 256) {
    const int lane = e & 63;
    int r = e >> 6;
    const int hl = r & 1;
    r >>= 1;
    const int ks = r & 3, sel = r >> 2;
    const int m = lane & 31, g = lane >> 5;
    const int tl = m >> 3, s = (m >> 1) & 3, ri = m & 1;
    const int t = sel < 2 ? 4 * sel + tl : (tl == 0 ? 8 : -1);
    h8v out;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int cl = q >> 2, sl = (q >> 1) & 1, rin = q & 1;
      const int i = 2 * ks + cl, s2 = 2 * g + sl;
      float v = 0.f;
      if (t >= 0) {
        const float2 w = W[((size_t)t * 8 + i) * 16 + s * 4 + s2];
        v = ri == 0 ? (rin == 0 ? w.x : -w.y) : (rin == 0 ? w.y : w.x);
      }
      v *= sw;
      const _Float16 h = (_Float16)v;
      out[q] = hl == 0 ? h : (_Float16)(v - (float)h);
    }
    Apk[e] = __builtin_bit_cast(uint4, out);
  }
}

// v <- M v on the four spins of a colour-spin vector
QCU_D void transport4(const cf (&m)[9], cf (&v)[4][3]) {
#pragma unroll
  for (int s = 0; s < 4; ++s) transport1(m, v[s]);
}

// Epilogue of an 8 -> 8 layer fused with the next, single-output layer
// (32-site shape).  Lane (n, g) holds spins 2g, 2g + 1 of all 8 output
// channels of site n: exactly the K rows 8 g .. 8 g + 7 of each K step in the
// order k = 8 g + 4 cl + 2 sl + ri (channel 2 ks + cl, spin 2 g + sl), so
// the accumulators are the B operand of the mixing Y_t = W'_t out without
// any exchange between lanes.  The site's values get a power-of-two scale
// from their largest component (the two lanes of a site share it) before the
// FP16 split.  Even output sites produce the 8 hop-term fields (two M tiles),
// odd ones the site-term field (one tile, rows 0..7).
// Operands and output of the single-output mixing: selections 0 and 1 (the
// hop terms) at A01, selection 2 (the site term) at A2, each in LDS or global
// memory (A2: the packed operands from selection 2 on); winv2 their inverse
// scale; mix the term fields.
struct MixOut {
  const uint4 *A01, *A2;
  const float *winv2;
  char *mix;
};

// pre() is called once the accumulators are consumed: the caller issues its
// next loads there.
template <class Pre>
QCU_D void fused_epilogue(const MixOut &mo, const f16v (&acc)[3][2], float oscale, int i, int p, bool ok, int g,
                          int lane, size_t field, size_t plane, Pre pre) {
  float mx = 0.f;
#pragma unroll
  for (int cc = 0; cc < 3; ++cc)
#pragma unroll
    for (int mt = 0; mt < 2; ++mt)
#pragma unroll
      for (int r = 0; r < 16; ++r) mx = fmaxf(mx, fabsf(acc[cc][mt][r]));
  mx = fmaxf(mx, __shfl_xor(mx, 32));
  const float sn = pow2_scale(mx * oscale, 14);
  const float f = oscale * sn;                 // accumulator -> scaled B operand (a power of two)
  const float yscale = mo.winv2[0] / sn;       // mixed field -> its true value
  const unsigned off = (unsigned)i * 16u;
  // the B operands of the mixing, shared by its selections: per K step and
  // colour, the accumulators' rows of that K step, scaled and split
  h8v bh[4][3], bl[4][3];
#pragma unroll
  for (int ks = 0; ks < 4; ++ks)
#pragma unroll
    for (int cc = 0; cc < 3; ++cc) {
      const int mt = ks >> 1;
      unsigned uh[4], ul[4];
#pragma unroll
      for (int cl = 0; cl < 2; ++cl)
#pragma unroll
        for (int sl = 0; sl < 2; ++sl) {
          const int r = 4 * (2 * (ks & 1) + cl) + 2 * sl;
          split2(acc[cc][mt][r] * f, acc[cc][mt][r + 1] * f, uh[2 * cl + sl], ul[2 * cl + sl]);
        }
      bh[ks][cc] = as_h8(make_uint4(uh[0], uh[1], uh[2], uh[3]));
      bl[ks][cc] = as_h8(make_uint4(ul[0], ul[1], ul[2], ul[3]));
    }
  // the accumulators are consumed: the caller's next loads go out here
  __builtin_amdgcn_sched_barrier(0);
  pre();
  __builtin_amdgcn_sched_barrier(0);
  // selection sel: one M tile of the mixing (8 packed operands of 64 lanes:
  // K step ks, hi / lo)
  const int nsel = p == 0 ? 2 : 1;
#pragma unroll 1
  for (int k = 0; k < nsel; ++k) {
    const int sel = p == 0 ? k : 2;
    f16v dy[3];
#pragma unroll
    for (int cc = 0; cc < 3; ++cc) dy[cc] = (f16v){};
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) {
      h8v ah, al;
      if (p == 0) {
        ah = as_h8(mo.A01[((k * 4 + ks) * 2) * 64 + lane]);
        al = as_h8(mo.A01[((k * 4 + ks) * 2 + 1) * 64 + lane]);
      } else {
        ah = as_h8(mo.A2[(ks * 2) * 64 + lane]);
        al = as_h8(mo.A2[(ks * 2 + 1) * 64 + lane]);
      }
#pragma unroll
      for (int cc = 0; cc < 3; ++cc) {
        dy[cc] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bh[ks][cc], dy[cc], 0, 0, 0);
        dy[cc] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bl[ks][cc], dy[cc], 0, 0, 0);
        dy[cc] = __builtin_amdgcn_mfma_f32_32x32x16_f16(al, bh[ks][cc], dy[cc], 0, 0, 0);
      }
    }
    if (!ok) continue;
    // register r of dy: row 8 (r / 4) + 4 g + r % 4 = term 4 sel + r / 4 (the
    // site term: r / 4 == 0), spin 2 g + (r % 4) / 2
#pragma unroll
    for (int tl = 0; tl < 4; ++tl) {
      if (sel == 2 && tl > 0) break;
      const int t = sel == 2 ? 8 : 4 * sel + tl;
      char *ob = mo.mix + ((size_t)t * 2 + p) * field;
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        const int l0 = 2 * q, l1 = 2 * q + 1;
        const int e0 = 4 * tl + 2 * (l0 / 3), e1 = 4 * tl + 2 * (l1 / 3);
        float4 v;
        v.x = dy[l0 % 3][e0] * yscale;
        v.y = dy[l0 % 3][e0 + 1] * yscale;
        v.z = dy[l1 % 3][e1] * yscale;
        v.w = dy[l1 % 3][e1 + 1] * yscale;
        st_nt(ob, (size_t)(3 * g + q) * plane, off, v);
      }
    }
  }
}

// NS sites per wave, WAVES waves per block, MT M tiles (NS rows each: cout x 8 / NS)
template <int NS, int WAVES, int MT, bool FUSE = false, bool DIAG = false>
__global__ void __launch_bounds__(64 * WAVES) k_net_layer_split(NetSplitArgs a) {
  typedef SplitShape<NS> SH;
  static_assert(!FUSE || (NS == 32 && MT == 2), "the fused epilogue is written for the 32-site shape, 8 outputs");
  constexpr int G = 64 / NS, KS = 8 / G, NR = SH::NR;
  constexpr int NA = 9 * MT * KS * 2 * 64;
  // fused: the next layer's hop-term operands follow the weights (LDS total
  // 144 + 16 KiB = the CU's 160 KiB)
  constexpr int NA2 = FUSE ? 2 * 4 * 2 * 64 : 0;
  __shared__ uint4 lw[NA + NA2];
  if (a.st && a.st->done) return;  // uniform over the grid
  for (int k = threadIdx.x; k < NA; k += blockDim.x) lw[k] = a.A[k];
  if constexpr (FUSE)
    for (int k = threadIdx.x; k < NA2; k += blockDim.x) lw[NA + k] = a.A2[k];
  __syncthreads();
  const Geom g = a.g;
  const float sin_ = pow2_scale(__uint_as_float(*a.in_max), 14);
  const float oscale = a.winv[0] / sin_;
  const int lane = threadIdx.x & 63, gq = lane / NS, n = lane % NS;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform: chunk and parity in SGPRs
  const int per_par = (g.vh + NS - 1) / NS;
  const int nchunk = a.npar * per_par;
  // chunk sequence of this wave: plain interleave over the grid, or (order 1)
  // a contiguous eighth of the chunks per XCD (blocks are dealt to the 8 XCDs
  // round robin), interleaved over the XCD's waves
  // order 2: XCD x owns the z planes [x lz / 8, (x + 1) lz / 8) and its waves
  // walk that slab in site order with t slowest, so the y, z and t neighbours
  // a wave reads are read by waves of the same XCD (its L2) at about the same
  // time; it needs lz % 8 == 0 and whole chunks per (z, t) plane
  const int cpp = g.lxh * g.ly / NS;  // chunks per (z, t) plane
  // order 3 (diagonal): a chunk is one x row (NS == lxh), XCD x owns the
  // planes [x lz / 8, (x + 1) lz / 8) in pairs and walks each pair with t
  // slowest, both parities of a row side by side; every wave takes its terms
  // in an order rotated by phi = (y + 2 z + 3 t) mod 7 (see the term loop), so
  // that all the waves reading one input row read it in the same slot
  const bool diag = DIAG && a.order == 3 && NS == g.lxh && a.npar == 2 && (gridDim.x & 7) == 0 && (g.lz & 15) == 0;
  const bool slab = (a.order == 2 || diag) && (gridDim.x & 7) == 0 && (g.lz & 7) == 0 && cpp * NS == g.lxh * g.ly;
  int c0, cend, cstep;
  if (slab) {
    const int nb = gridDim.x >> 3;
    c0 = (int)(blockIdx.x >> 3) * WAVES + wave;
    cend = a.npar * cpp * (g.lz >> 3) * g.lt;  // chunks of one slab
    cstep = nb * WAVES;
  } else if (a.order == 1 && (gridDim.x & 7) == 0) {
    const int xcd = blockIdx.x & 7, nb = gridDim.x >> 3;
    c0 = (int)((long long)nchunk * xcd / 8) + (int)(blockIdx.x >> 3) * WAVES + wave;
    cend = (int)((long long)nchunk * (xcd + 1) / 8);
    cstep = nb * WAVES;
  } else {
    c0 = blockIdx.x * WAVES + wave;
    cend = nchunk;
    cstep = gridDim.x * WAVES;
  }
  const size_t field = (size_t)g.vh * 96, plane = (size_t)g.vh * 16;
  const unsigned chan_off = (unsigned)(2 * field);  // one channel (both parities)
  float omax = 0.f;
  // chunk c: output parity, first site of the lane (lanes past the last site
  // compute a copy of a valid site and store nothing), diagonal rotation
  struct Chunk {
    int p, i, phi;
    bool ok;
  };
  // The two parities of a site chunk are chunks c and c ^ 1.  Which is which
  // alternates every 4 chunks, so that the two waves of a SIMD (waves w and
  // w + 4 of a block) take different parities: the fused epilogue does twice
  // the work for even output sites.
  auto chunk_at = [&](int c) {
    Chunk k;
    k.p = a.parity0 + (a.npar == 2 ? ((c ^ (c >> 2)) & 1) : 0);
    int cs = a.npar == 2 ? c >> 1 : c;  // chunk in site order
    k.phi = 0;
    if (diag) {
      int r1 = cs;  // y fastest, then the plane within the pair, t, the pair
      const int y = r1 % g.ly;
      r1 /= g.ly;
      const int zz = r1 & 1;
      r1 >>= 1;
      const int tt = r1 % g.lt, pair = r1 / g.lt;
      const int z = (int)(blockIdx.x & 7) * (g.lz >> 3) + 2 * pair + zz;
      cs = (tt * g.lz + z) * g.ly + y;
      k.phi = (y + 2 * z + 3 * tt) % 7;
    } else if (slab) {
      const int dz = g.lz >> 3, within = cs % cpp, q = cs / cpp;
      cs = ((q / dz) * g.lz + (int)(blockIdx.x & 7) * dz + q % dz) * cpp + within;
    }
    const int base = cs * NS;
    k.i = base + n;
    k.ok = k.i < g.vh;
    if (!k.ok) k.i = base;
    return k;
  };
  typename SH::acc_t acc[3][MT];
  // the lane's channel operand in K step ks: input channel G ks + gq at the
  // term's neighbour; the base of channel G ks is uniform over the wave and
  // the lane group adds gq channels (a 32-bit offset, checked at launch)
  auto load_v = [&](float4 (&q)[6], const TermGeo &tg, int ks) {
    const char *cb = a.in + ((size_t)(2 * G * ks) + tg.jp) * field;
    const unsigned off = (unsigned)tg.j * 16u + (unsigned)gq * chan_off;
#pragma unroll
    for (int k = 0; k < 6; ++k) q[k] = ld_at<float4>(cb, k * plane, off);
  };
  auto load_link = [&](const TermGeo &tg, float4 (&l)[3]) {
#pragma unroll
    for (int k = 0; k < 3; ++k)
      l[k] = ld_at<float4>(a.U.a, ((size_t)tg.slab * 3 + k) * plane, (unsigned)tg.li * 16u);
  };
  // one K step: the channel operand in q is unpacked and transported, then q
  // is refilled with the next step's operand (next(q)) while the products run
  auto step = [&](float4 (&q)[6], const cf (&m)[9], bool hop, int t, int ks, auto next) {
    cf v[4][3];
#pragma unroll
    for (int s2 = 0; s2 < 4; ++s2)
#pragma unroll
      for (int cc = 0; cc < 3; ++cc) {
        const int ci = 3 * s2 + cc;
        const float4 f = q[ci >> 1];
        v[s2][cc] = (ci & 1) ? mk(f.z, f.w) : mk(f.x, f.y);
      }
    if (hop) {
      transport4(m, v);  // m carries the input scale
    } else {
#pragma unroll
      for (int s2 = 0; s2 < 4; ++s2)
#pragma unroll
        for (int cc = 0; cc < 3; ++cc) v[s2][cc] = sin_ * v[s2][cc];
    }
    // the refill is issued here, a whole step ahead of its use: the
    // scheduling barriers keep the compiler from sinking it below the
    // products (where it would reuse the operand registers and leave the
    // next step waiting on it)
    __builtin_amdgcn_sched_barrier(0);
    next(q);
    __builtin_amdgcn_sched_barrier(0);
    // the two FP16 halves of each colour's B operand
    unsigned uh[3][4], ul[3][4];
#pragma unroll
    for (int cc = 0; cc < 3; ++cc)
#pragma unroll
      for (int s2 = 0; s2 < 4; ++s2) split2(v[s2][cc].re, v[s2][cc].im, uh[cc][s2], ul[cc][s2]);
    // M tile by M tile: its two A halves, then three products per colour
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
      const int ab = (((t * MT + mt) * KS + ks) * 2) * 64 + lane;
      const h8v ah = as_h8(lw[ab]), al = as_h8(lw[ab + 64]);
#pragma unroll
      for (int cc = 0; cc < 3; ++cc) {
        const h8v bh = as_h8(make_uint4(uh[cc][0], uh[cc][1], uh[cc][2], uh[cc][3]));
        const h8v bl = as_h8(make_uint4(ul[cc][0], ul[cc][1], ul[cc][2], ul[cc][3]));
        acc[cc][mt] = SH::mfma(ah, bh, acc[cc][mt]);
        acc[cc][mt] = SH::mfma(ah, bl, acc[cc][mt]);
        acc[cc][mt] = SH::mfma(al, bh, acc[cc][mt]);
      }
    }
  };
  // the chunk's output: its 8 (or 4) channels, or with FUSE the next layer's
  // mixed term fields
  // pre() issues the wave's next loads: before the stores of a plain layer,
  // once the accumulators are split in the fused epilogue
  auto finish = [&](const Chunk &k, auto pre) {
    if constexpr (FUSE) {
      const MixOut mo{lw + NA, a.A2 + 2 * 4 * 2 * 64, a.winv2, a.mix};
      fused_epilogue(mo, acc, oscale, k.i, k.p, k.ok, gq, lane, field, plane, pre);
    } else {
      pre();
      if (!k.ok) return;
      // lane (n, gq) holds, per M tile and group of 4 registers j, output
      // channel (NS / 8) mt + row(4 j, gq) / 8, spins sb and sb + 1 with
      // sb = row(4 j, gq) % 8 / 2, all colours: float4 planes 3 sb / 2 .. + 2
      const unsigned off = (unsigned)k.i * 16u;
#pragma unroll
      for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int j = 0; j < NR / 4; ++j) {
          const int r0 = SH::row(4 * j, gq);
          const int o = (NS / 8) * mt + r0 / 8, sb = (r0 % 8) / 2;
          char *ob = a.out + ((size_t)o * 2 + k.p) * field;
#pragma unroll
          for (int q = 0; q < 3; ++q) {
            // complex li = 2q, 2q + 1 of the lane's six: spin sb + li / 3, colour li % 3
            const int l0 = 2 * q, l1 = 2 * q + 1;
            const int e0 = 4 * j + 2 * (l0 / 3), e1 = 4 * j + 2 * (l1 / 3);
            float4 f;
            f.x = acc[l0 % 3][mt][e0] * oscale;
            f.y = acc[l0 % 3][mt][e0 + 1] * oscale;
            f.z = acc[l1 % 3][mt][e1] * oscale;
            f.w = acc[l1 % 3][mt][e1 + 1] * oscale;
            omax = fmaxf(omax, fmaxf(fmaxf(fabsf(f.x), fabsf(f.y)), fmaxf(fabsf(f.z), fabsf(f.w))));
            st_nt(ob, (size_t)(3 * (sb / 2) + q) * plane, off, f);
          }
        }
    }
  };
  auto clear = [&]() {
#pragma unroll
    for (int cc = 0; cc < 3; ++cc)
#pragma unroll
      for (int mt = 0; mt < MT; ++mt) acc[cc][mt] = (typename SH::acc_t){};
  };
  if constexpr (DIAG) {
    // (diag holds: the launcher picks this instance only when it does)
#pragma unroll 1
    for (int c = c0; c < cend; c += cstep) {
      const Chunk ck = chunk_at(c);
      const int i = ck.i, p = ck.p;
      const Site site = decode(g, i, p);
      clear();
      // Diagonal schedule.  Slot sigma holds: 0 the local terms (x+, site,
      // x-), then y- 1, z- 2, t- 3, t+ 4, z+ 5, y+ 6, i.e. +-a_mu for the
      // backward / forward hop with a = (y 1, z 2, t 3).  At step k a wave
      // takes slot (k + phi) mod 7: a wave at R reads R + e_mu in the forward
      // slot -a_mu, so the reader of input row S through any term takes it in
      // step -phi(S) mod 7.  K step outer, so that one slot is one step.
      constexpr unsigned long long slot_terms[7] = {0x108, 0x3, 0x5, 0x7, 0x6, 0x4, 0x2};
      constexpr int slot_len[7] = {3, 1, 1, 1, 1, 1, 1};
      unsigned long long seq = 0;  // 9 terms, 4 bits each
      int nseq = 0;
#pragma unroll
      for (int k = 0; k < 7; ++k) {
        const int sg = (k + ck.phi) % 7;
        seq |= slot_terms[sg] << (4 * nseq);
        nseq += slot_len[sg];
      }
      auto term_at = [&](int j) { return (int)((seq >> (4 * j)) & 15u); };
      TermGeo tg = term_geo(g, site, i, p, term_at(0));
      float4 lraw[3];
      if (tg.hop) load_link(tg, lraw);
      float4 vb[6];
      load_v(vb, tg, 0);
#pragma unroll
      for (int ks = 0; ks < KS; ++ks) {
#pragma unroll 1
        for (int j = 0; j < 9; ++j) {
          const int t = term_at(j);
          cf m[9];
          term_matrix(lraw, tg.hop, tg.fwd, tg.ph * sin_, m);
          const bool last = ks + 1 == KS && j == 8;
          const TermGeo ng = term_geo(g, site, i, p, term_at(j == 8 ? 0 : j + 1));
          step(vb, m, tg.hop, t, ks, [&](float4(&q)[6]) {
            if (!last) {
              if (ng.hop) load_link(ng, lraw);
              load_v(q, ng, j == 8 ? ks + 1 : ks);
            }
          });
          tg = ng;
        }
      }
      finish(ck, []() {});
    }
  } else {
    // term outer, K step inner: each term's link is loaded and its third row
    // rebuilt once, the next term's link while this one runs.  The first
    // loads of a wave's next chunk are issued before the current chunk's
    // output is formed and stored, so that they overlap it.
    float4 lraw[3], vb[6];
    Chunk ck{};
    Site site{};
    TermGeo tg{};
    auto begin = [&](int c) {
      ck = chunk_at(c);
      site = decode(g, ck.i, ck.p);
      tg = term_geo(g, site, ck.i, ck.p, 0);
      load_link(tg, lraw);
      load_v(vb, tg, 0);
    };
    if (c0 < cend) begin(c0);
#pragma unroll 1
    for (int c = c0; c < cend; c += cstep) {
      const Chunk cur = ck;
      clear();
#pragma unroll 1
      for (int t = 0; t < 9; ++t) {
        cf m[9];
        term_matrix(lraw, tg.hop, tg.fwd, tg.ph * sin_, m);
        TermGeo ng = tg;
        if (t + 1 < 9) {
          ng = term_geo(g, site, cur.i, cur.p, t + 1);
          if (ng.hop) load_link(ng, lraw);
        }
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
          step(vb, m, tg.hop, t, ks, [&](float4(&q)[6]) {
            if (ks + 1 < KS) load_v(q, tg, ks + 1);
            else if (t + 1 < 9) load_v(q, ng, 0);
          });
        }
        tg = ng;
      }
      finish(cur, [&]() {
        if (c + cstep < cend) begin(c + cstep);
      });
    }
  }
  if (a.out_max) record_max(a.out_max, omax);
}

// ---------------------------------------------------------------------------
// The first layer (1 -> 8 channels) on an input that is zero on the even
// sites (the preconditioner's and training's input), split FP16 on
// v_mfma_f32_32x32x16_f16.  Even output sites take the 8 hop terms, odd ones
// the site term.  Per colour the even sites are the GEMM
//   D[(o, s, re/im)][site] = sum over (term t, spin s2, re/im) of W_t[o][s][s2] V_t[s2][site]
// with K step ks holding the forward (lane group 0) and backward (group 1)
// hop in direction ks: each lane transports the one input channel from its
// term's neighbour.  The odd sites are one more K step, the site term on
// lane group 0 (group 1 zero).  A wave owns 32 sites and produces both
// parities of their 8 output channels.
// Packs W[t][o][0][4x4] as A[ks][mt][hl][lane] (ks 0..3 the hop pairs, 4
// the site term; 8 FP16 each), scaled by a power of two; winv[0] receives
// its inverse.  One block.
__global__ void __launch_bounds__(256) k_net_pack_first(const float2 *W, uint4 *Apk, float *winv) {
  __shared__ float red[256];
  const int tid = threadIdx.x;
  const int nw = 9 * 8 * 16;
  float mx = 0.f;
  for (int k = tid; k < nw; k += 256) mx = fmaxf(mx, fmaxf(fabsf(W[k].x), fabsf(W[k].y)));
  red[tid] = mx;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (tid < h) red[tid] = fmaxf(red[tid], red[tid + h]);
    __syncthreads();
  }
  const float sw = pow2_scale(red[0], 14);
  if (tid == 0) winv[0] = 1.f / sw;
  for (int e = tid; e < kFirstPackOps; e += 256) {
    const int lane = e & 63;
    int r = e >> 6;
    const int hl = r & 1;
    r >>= 1;
    const int mt = r & 1, ks = r >> 1;
    const int m = lane & 31, g = lane >> 5;
    const int o = 4 * mt + (m >> 3), s = (m >> 1) & 3, ri = m & 1;
    const int t = ks < 4 ? 2 * ks + g : (g == 0 ? 8 : -1);
    h8v out;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int s2 = q >> 1, rin = q & 1;
      float v = 0.f;
      if (t >= 0) {
        const float2 w = W[((size_t)t * 8 + o) * 16 + s * 4 + s2];
        v = ri == 0 ? (rin == 0 ? w.x : -w.y) : (rin == 0 ? w.y : w.x);
      }
      v *= sw;
      const _Float16 h = (_Float16)v;
      out[q] = hl == 0 ? h : (_Float16)(v - (float)h);
    }
    Apk[e] = __builtin_bit_cast(uint4, out);
  }
}

// Chunk order of the kernels that take 32 sites of one parity per chunk
// (k_net_first, k_net_second): where the lattice splits into whole z slabs,
// one per XCD (the grid a multiple of 8 blocks, block b on XCD b mod 8), each
// XCD sweeps its slab plane by plane in t, so that the rows its chunks read
// through the z and t hops are read by its own waves while they are still in
// its L2; otherwise chunks are interleaved over the grid.
struct SlabOrder {
  bool slab;
  int cpp, dz;           // chunks per z plane, z planes per slab
  int c0, cend, cstep;   // the wave's chunks: c0, c0 + cstep, ... < cend
};
QCU_D SlabOrder slab_order(const Geom &g, int waves, int wave) {
  SlabOrder o;
  o.cpp = (g.lxh * g.ly) / 32;
  o.dz = g.lz >> 3;
  o.slab = (gridDim.x & 7) == 0 && (g.lz & 7) == 0 && o.cpp * 32 == g.lxh * g.ly;
  if (o.slab) {
    o.c0 = (int)(blockIdx.x >> 3) * waves + wave;
    o.cend = o.cpp * o.dz * g.lt;
    o.cstep = (int)(gridDim.x >> 3) * waves;
  } else {
    o.c0 = blockIdx.x * waves + wave;
    o.cend = (g.vh + 31) / 32;
    o.cstep = gridDim.x * waves;
  }
  return o;
}
// chunk c's index in site order (its first site is 32 times it)
QCU_D int slab_chunk(const SlabOrder &o, const Geom &g, int c) {
  if (!o.slab) return c;
  const int within = c % o.cpp, q = c / o.cpp;
  return ((q / o.dz) * g.lz + (int)(blockIdx.x & 7) * o.dz + q % o.dz) * o.cpp + within;
}

// EVEN_ONLY: the even output sites only (the odd ones are then never read:
// the next layer takes their site term from the input itself, k_net_second)
template <int WAVES, bool EVEN_ONLY>
__global__ void __launch_bounds__(64 * WAVES) k_net_first(NetFirstArgs a) {
  typedef SplitShape<32> SH;
  __shared__ uint4 lw[kFirstPackOps];
  if (a.st && a.st->done) return;  // uniform over the grid
  for (int k = threadIdx.x; k < kFirstPackOps; k += blockDim.x) lw[k] = a.A[k];
  __syncthreads();
  const Geom g = a.g;
  const float sin_ = pow2_scale(__uint_as_float(*a.in_max), 14);
  const float oscale = a.winv[0] / sin_;
  const int lane = threadIdx.x & 63, gq = lane >> 5, n = lane & 31;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const size_t field = (size_t)g.vh * 96, plane = (size_t)g.vh * 16;
  const char *in1 = a.in + field;  // the input's odd sites
  float omax = 0.f;
  f16v acc[3][2];
  auto clear = [&]() {
#pragma unroll
    for (int cc = 0; cc < 3; ++cc) acc[cc][0] = acc[cc][1] = (f16v){};
  };
  // products of one K step: the lane's transported (and scaled) 4 spins x 3
  // colours, split, against the two M tiles of the step
  auto products = [&](const cf (&v)[4][3], int ks) {
    unsigned uh[3][4], ul[3][4];
#pragma unroll
    for (int cc = 0; cc < 3; ++cc)
#pragma unroll
      for (int s2 = 0; s2 < 4; ++s2) split2(v[s2][cc].re, v[s2][cc].im, uh[cc][s2], ul[cc][s2]);
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) {
      const int ab = ((ks * 2 + mt) * 2) * 64 + lane;
      const h8v ah = as_h8(lw[ab]), al = as_h8(lw[ab + 64]);
#pragma unroll
      for (int cc = 0; cc < 3; ++cc) {
        const h8v bh = as_h8(make_uint4(uh[cc][0], uh[cc][1], uh[cc][2], uh[cc][3]));
        const h8v bl = as_h8(make_uint4(ul[cc][0], ul[cc][1], ul[cc][2], ul[cc][3]));
        acc[cc][mt] = SH::mfma(ah, bh, acc[cc][mt]);
        acc[cc][mt] = SH::mfma(ah, bl, acc[cc][mt]);
        acc[cc][mt] = SH::mfma(al, bh, acc[cc][mt]);
      }
    }
  };
  auto unpack = [](const float4 (&q)[6], cf (&v)[4][3]) {
#pragma unroll
    for (int s2 = 0; s2 < 4; ++s2)
#pragma unroll
      for (int cc = 0; cc < 3; ++cc) {
        const int ci = 3 * s2 + cc;
        const float4 f = q[ci >> 1];
        v[s2][cc] = (ci & 1) ? mk(f.z, f.w) : mk(f.x, f.y);
      }
  };
  // lane (n, gq) holds spins 2 gq, 2 gq + 1 of the 8 output channels of its
  // site: float4 planes 3 gq .. 3 gq + 2 of each channel
  auto store = [&](int i, int p) {
    const unsigned off = (unsigned)i * 16u;
#pragma unroll
    for (int mt = 0; mt < 2; ++mt)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int o = 4 * mt + j;
        char *ob = a.out + ((size_t)o * 2 + p) * field;
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          const int l0 = 2 * q, l1 = 2 * q + 1;
          const int e0 = 4 * j + 2 * (l0 / 3), e1 = 4 * j + 2 * (l1 / 3);
          float4 f;
          f.x = acc[l0 % 3][mt][e0] * oscale;
          f.y = acc[l0 % 3][mt][e0 + 1] * oscale;
          f.z = acc[l1 % 3][mt][e1] * oscale;
          f.w = acc[l1 % 3][mt][e1 + 1] * oscale;
          omax = fmaxf(omax, fmaxf(fmaxf(fabsf(f.x), fabsf(f.y)), fmaxf(fabsf(f.z), fabsf(f.w))));
          st_nt(ob, (size_t)(3 * gq + q) * plane, off, f);
        }
      }
  };
  auto load_v = [&](float4 (&q)[6], int j) {
#pragma unroll
    for (int k = 0; k < 6; ++k) q[k] = ld_at<float4>(in1, k * plane, (unsigned)j * 16u);
  };
  // each link is read once (by the even sites' hops)
  auto load_link = [&](const TermGeo &tg, float4 (&l)[3]) {
#pragma unroll
    for (int k = 0; k < 3; ++k) l[k] = ld_nt(a.U.a, ((size_t)tg.slab * 3 + k) * plane, (unsigned)tg.li * 16u);
  };
  int i = 0;
  bool ok = false;
  Site se{};
  TermGeo tg{};
  float4 lraw[3], vb[6];
  const SlabOrder sl = slab_order(g, WAVES, wave);
  // chunk c: the lane's site (lanes past the last site compute a copy of a
  // valid site and store nothing) and the first hop pair's loads
  auto begin = [&](int c) {
    const int base = slab_chunk(sl, g, c) * 32;
    i = base + n;
    ok = i < g.vh;
    if (!ok) i = base;
    se = decode(g, i, 0);
    tg = term_geo_u(g, se, i, 0, term_const(g, gq));
    load_link(tg, lraw);
    load_v(vb, tg.j);
  };
  const int cstep = sl.cstep, cend = sl.cend;
  int c = sl.c0;
  if (c < cend) begin(c);
#pragma unroll 1
  for (; c < cend; c += cstep) {
    // even output sites: K step ks = the hops in direction ks, forward on
    // lane group 0, backward on group 1; the next step's link and
    // neighbour are loaded while this step's products run
    const int ie = i;
    const bool oke = ok;
    clear();
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) {
      cf m[9], v[4][3];
      term_matrix(lraw, 1, tg.fwd, tg.ph * sin_, m);
      unpack(vb, v);
      transport4(m, v);
      __builtin_amdgcn_sched_barrier(0);
      if (ks < 3) {
        tg = term_geo_u(g, se, i, 0, term_const(g, 2 * (ks + 1) + gq));
        load_link(tg, lraw);
        load_v(vb, tg.j);
      } else if constexpr (EVEN_ONLY) {
        if (c + cstep < cend) begin(c + cstep);  // the next chunk's first loads
      } else {
        load_v(vb, i);  // the odd pass: the site's own input
      }
      __builtin_amdgcn_sched_barrier(0);
      products(v, ks);
    }
    if (oke) store(ie, 0);
    if constexpr (EVEN_ONLY) continue;
    // odd output sites: the site term on lane group 0; the next chunk's first
    // loads go out first
    clear();
    {
      cf v[4][3];
      unpack(vb, v);
      const int io = i;
      const bool oko = ok;
      __builtin_amdgcn_sched_barrier(0);
      if (c + cstep < cend) begin(c + cstep);
      __builtin_amdgcn_sched_barrier(0);
      const float f = gq == 0 ? sin_ : 0.f;
#pragma unroll
      for (int s2 = 0; s2 < 4; ++s2)
#pragma unroll
        for (int cc = 0; cc < 3; ++cc) v[s2][cc] = f * v[s2][cc];
      products(v, 4);
      if (oko) store(io, 1);
    }
  }
  if (a.out_max) record_max(a.out_max, omax);
}

int net_pack_first(const float2 *W, void *Apk, float *winv, hipStream_t s) {
  hipLaunchKernelGGL(k_net_pack_first, dim3(1), dim3(256), 0, s, W, (uint4 *)Apk, winv);
  QCU_HIP(hipGetLastError());
  return 0;
}

int net_layer_first(const NetFirstArgs &a, hipStream_t s, bool even_only) {
  QCU_CHECK((size_t)a.g.vh * 96 < ((size_t)1 << 32), QCU_ERR_DIMS, "net: lattice too large for the first layer");
  static int ncu = 0;
  if (!ncu) {
    int dev = 0;
    QCU_HIP(hipGetDevice(&dev));
    QCU_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    if (ncu <= 0) ncu = 256;
  }
  // 8 waves per block, one block per CU (213 registers: two waves per SIMD)
  constexpr int waves = 8;
  const int nchunk = (a.g.vh + 31) / 32;
  int nblk = ncu;
  if (nchunk < nblk * waves) nblk = (nchunk + waves - 1) / waves;
  if (even_only) hipLaunchKernelGGL((k_net_first<waves, true>), dim3(nblk), dim3(64 * waves), 0, s, a);
  else hipLaunchKernelGGL((k_net_first<waves, false>), dim3(nblk), dim3(64 * waves), 0, s, a);
  QCU_HIP(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------------------
// The second layer (8 -> 8) fused with what the first one leaves implicit.
// With the network input zero on the even sites, the first layer's output is
// its site term W1_8 in(y) at odd y, and its 8 hop terms at even y.  So, by
// linearity (the spin matrices commute with the colour transports), the
// second layer's output is
//   even x: sum_t (W2_t W1_8 + W2_8 W1_t) T_t in(x + t)      (t = the 8 hops)
//   odd x:  sum_t W2_t T_t L1(x + t) + W2_8 W1_8 in(x)        (L1 at even sites)
// The even sites read only the one-channel input (like k_net_first), and the
// first layer writes only its even sites.  Products are split FP16 as in
// k_net_layer_split; the combined matrices E_t = W2_t W1_8 + W2_8 W1_t and
// S = W2_8 W1_8 are formed here (FP64 sums, rounded to FP32) and packed like
// k_net_first's operands: A[ks][mt][hl][lane], ks 0..3 the hop pairs of E, 4
// the site term S; winv[0], winv[1] the inverse scales of E and S.
__global__ void __launch_bounds__(256) k_net_pack_second(const float2 *W1, const float2 *W2, uint4 *Apk,
                                                         float *winv) {
  __shared__ float2 cm[9 * 8 * 16];  // E_0..7, then S: [t][o][s][s2]
  __shared__ float red[2][256];
  const int tid = threadIdx.x;
  float mxe = 0.f, mxs = 0.f;
  for (int e = tid; e < 9 * 8 * 16; e += 256) {
    const int t = e >> 7, o = (e >> 4) & 7, s = (e >> 2) & 3, s2 = e & 3;
    double re = 0.0, im = 0.0;
    // sum_i sum_s' A[o][i][s][s'] B[i][s'][s2], A of the second layer (cin 8),
    // B of the first (cin 1); W[t][o][i][s][s2] = W[((t cout + o) cin + i) 16 + 4 s + s2]
    auto acc = [&](int ta, int tb) {
      for (int i = 0; i < 8; ++i)
        for (int sp = 0; sp < 4; ++sp) {
          const float2 a = W2[(((size_t)ta * 8 + o) * 8 + i) * 16 + s * 4 + sp];
          const float2 b = W1[((size_t)tb * 8 + i) * 16 + sp * 4 + s2];
          re += (double)a.x * b.x - (double)a.y * b.y;
          im += (double)a.x * b.y + (double)a.y * b.x;
        }
    };
    if (t < 8) {
      acc(t, 8);
      acc(8, t);
    } else {
      acc(8, 8);
    }
    const float2 v = make_float2((float)re, (float)im);
    cm[e] = v;
    const float m = fmaxf(fabsf(v.x), fabsf(v.y));
    if (t < 8) mxe = fmaxf(mxe, m);
    else mxs = fmaxf(mxs, m);
  }
  red[0][tid] = mxe;
  red[1][tid] = mxs;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (tid < h) {
      red[0][tid] = fmaxf(red[0][tid], red[0][tid + h]);
      red[1][tid] = fmaxf(red[1][tid], red[1][tid + h]);
    }
    __syncthreads();
  }
  const float swe = pow2_scale(red[0][0], 14), sws = pow2_scale(red[1][0], 14);
  if (tid == 0) {
    winv[0] = 1.f / swe;
    winv[1] = 1.f / sws;
  }
  for (int e = tid; e < kFirstPackOps; e += 256) {
    const int lane = e & 63;
    int r = e >> 6;
    const int hl = r & 1;
    r >>= 1;
    const int mt = r & 1, ks = r >> 1;
    const int m = lane & 31, g = lane >> 5;
    const int o = 4 * mt + (m >> 3), s = (m >> 1) & 3, ri = m & 1;
    const int t = ks < 4 ? 2 * ks + g : (g == 0 ? 8 : -1);
    const float sw = ks < 4 ? swe : sws;
    h8v out;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int s2 = q >> 1, rin = q & 1;
      float v = 0.f;
      if (t >= 0) {
        const float2 w = cm[(t * 8 + o) * 16 + s * 4 + s2];
        v = ri == 0 ? (rin == 0 ? w.x : -w.y) : (rin == 0 ? w.y : w.x);
      }
      v *= sw;
      const _Float16 h = (_Float16)v;
      out[q] = hl == 0 ? h : (_Float16)(v - (float)h);
    }
    Apk[e] = __builtin_bit_cast(uint4, out);
  }
}

// MIX: the epilogue also writes the next layer's single-output term fields
// (fused_epilogue with the hop terms' operands in global memory, the site
// term's in LDS)
template <int WAVES, bool MIX>
__global__ void __launch_bounds__(64 * WAVES) k_net_second(NetSecondArgs a) {
  typedef SplitShape<32> SH;
  constexpr int NH = 8 * 2 * 4 * 2 * 64;  // the second layer's hop-term operands (split packing, terms 0..7)
  constexpr int N2 = MIX ? 4 * 2 * 64 : 0;  // the mixing's site-term operands (8 KiB: 156 KiB in all)
  __shared__ uint4 lw[NH + kFirstPackOps + N2];
  if (a.st && a.st->done) return;  // uniform over the grid
  for (int k = threadIdx.x; k < NH; k += blockDim.x) lw[k] = a.Ah[k];
  for (int k = threadIdx.x; k < kFirstPackOps; k += blockDim.x) lw[NH + k] = a.A[k];
  for (int k = threadIdx.x; k < N2; k += blockDim.x) lw[NH + kFirstPackOps + k] = a.A2[2 * 4 * 2 * 64 + k];
  __syncthreads();
  const uint4 *lh = lw, *le = lw + NH, *ls2 = lw + NH + kFirstPackOps;
  const Geom g = a.g;
  // operand scales: s0 the input's, s1 the first layer's even sites'
  const float s0 = pow2_scale(__uint_as_float(*a.in_max), 14);
  const float s1 = pow2_scale(__uint_as_float(*a.l1_max), 14);
  const float oe = a.winv[0] / s0;    // even sites: E against the input
  const float oo = a.winvh[0] / s1;   // odd sites: the hop terms against L1
  const float rs = (a.winv[1] * s1) / (a.winvh[0] * s0);  // the site term's part -> the hop terms' scale
  const int lane = threadIdx.x & 63, gq = lane >> 5, n = lane & 31;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const size_t field = (size_t)g.vh * 96, plane = (size_t)g.vh * 16;
  const unsigned chan_off = (unsigned)(2 * field);  // one channel (both parities)
  const char *in1 = a.in + field;  // the input's odd sites
  float omax = 0.f;
  f16v acc[3][2];
  auto clear = [&]() {
#pragma unroll
    for (int cc = 0; cc < 3; ++cc) acc[cc][0] = acc[cc][1] = (f16v){};
  };
  // products of one K step: the lane's (scaled) 4 spins x 3 colours, split,
  // against the two M tiles at A (tile mt at A + mt * mstride, lo at + 64)
  auto products = [&](const cf (&v)[4][3], const uint4 *A, int mstride) {
    unsigned uh[3][4], ul[3][4];
#pragma unroll
    for (int cc = 0; cc < 3; ++cc)
#pragma unroll
      for (int s2 = 0; s2 < 4; ++s2) split2(v[s2][cc].re, v[s2][cc].im, uh[cc][s2], ul[cc][s2]);
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) {
      const h8v ah = as_h8(A[mt * mstride + lane]), al = as_h8(A[mt * mstride + 64 + lane]);
#pragma unroll
      for (int cc = 0; cc < 3; ++cc) {
        const h8v bh = as_h8(make_uint4(uh[cc][0], uh[cc][1], uh[cc][2], uh[cc][3]));
        const h8v bl = as_h8(make_uint4(ul[cc][0], ul[cc][1], ul[cc][2], ul[cc][3]));
        acc[cc][mt] = SH::mfma(ah, bh, acc[cc][mt]);
        acc[cc][mt] = SH::mfma(ah, bl, acc[cc][mt]);
        acc[cc][mt] = SH::mfma(al, bh, acc[cc][mt]);
      }
    }
  };
  auto unpack = [](const float4 (&q)[6], cf (&v)[4][3]) {
#pragma unroll
    for (int s2 = 0; s2 < 4; ++s2)
#pragma unroll
      for (int cc = 0; cc < 3; ++cc) {
        const int ci = 3 * s2 + cc;
        const float4 f = q[ci >> 1];
        v[s2][cc] = (ci & 1) ? mk(f.z, f.w) : mk(f.x, f.y);
      }
  };
  // lane (n, gq) holds spins 2 gq, 2 gq + 1 of the 8 output channels of its
  // site: float4 planes 3 gq .. 3 gq + 2 of each channel
  auto store = [&](int i, int p, float os) {
    const unsigned off = (unsigned)i * 16u;
#pragma unroll
    for (int mt = 0; mt < 2; ++mt)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int o = 4 * mt + j;
        char *ob = a.out + ((size_t)o * 2 + p) * field;
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          const int l0 = 2 * q, l1 = 2 * q + 1;
          const int e0 = 4 * j + 2 * (l0 / 3), e1 = 4 * j + 2 * (l1 / 3);
          float4 f;
          f.x = acc[l0 % 3][mt][e0] * os;
          f.y = acc[l0 % 3][mt][e0 + 1] * os;
          f.z = acc[l1 % 3][mt][e1] * os;
          f.w = acc[l1 % 3][mt][e1 + 1] * os;
          omax = fmaxf(omax, fmaxf(fmaxf(fabsf(f.x), fabsf(f.y)), fmaxf(fabsf(f.z), fabsf(f.w))));
          st_nt(ob, (size_t)(3 * gq + q) * plane, off, f);
        }
      }
  };
  // the one-channel input at odd site j
  auto load_in = [&](float4 (&q)[6], int j) {
#pragma unroll
    for (int k = 0; k < 6; ++k) q[k] = ld_at<float4>(in1, k * plane, (unsigned)j * 16u);
  };
  // the first layer's channels 2 ks + gq at the term's (even) neighbour
  auto load_l1 = [&](float4 (&q)[6], const TermGeo &t, int ks) {
    const char *cb = a.l1 + ((size_t)(4 * ks) + t.jp) * field;
    const unsigned off = (unsigned)t.j * 16u + (unsigned)gq * chan_off;
#pragma unroll
    for (int k = 0; k < 6; ++k) q[k] = ld_at<float4>(cb, k * plane, off);
  };
  auto load_link = [&](const TermGeo &t, float4 (&l)[3]) {
#pragma unroll
    for (int k = 0; k < 3; ++k)
      l[k] = ld_at<float4>(a.U.a, ((size_t)t.slab * 3 + k) * plane, (unsigned)t.li * 16u);
  };
  int i = 0;
  bool ok = false;
  Site se{};
  TermGeo tg{};
  float4 lraw[3], vb[6];
  const SlabOrder sl = slab_order(g, WAVES, wave);
  // chunk c: the lane's site (lanes past the last site compute a copy of a
  // valid site and store nothing) and the even pass's first loads
  auto begin = [&](int c) {
    const int base = slab_chunk(sl, g, c) * 32;
    i = base + n;
    ok = i < g.vh;
    if (!ok) i = base;
    se = decode(g, i, 0);
    tg = term_geo_u(g, se, i, 0, term_const(g, gq));
    load_link(tg, lraw);
    load_in(vb, tg.j);
  };
  const int cstep = sl.cstep, cend = sl.cend;
  int c = sl.c0;
  if (c < cend) begin(c);
#pragma unroll 1
  for (; c < cend; c += cstep) {
    const int ic = i;
    const bool okc = ok;
    // even output sites: K step ks = the hops in direction ks of the input
    // (forward on lane group 0, backward on group 1) against E
    clear();
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) {
      cf m[9], v[4][3];
      term_matrix(lraw, 1, tg.fwd, tg.ph * s0, m);
      unpack(vb, v);
      transport4(m, v);
      __builtin_amdgcn_sched_barrier(0);
      if (ks < 3) {
        tg = term_geo_u(g, se, ic, 0, term_const(g, 2 * (ks + 1) + gq));
        load_link(tg, lraw);
        load_in(vb, tg.j);
      } else if constexpr (!MIX) {
        load_in(vb, ic);  // the odd pass's site term: the site's own input
      }
      __builtin_amdgcn_sched_barrier(0);
      products(v, le + (ks * 2) * 2 * 64, 2 * 64);
    }
    if (okc) store(ic, 0, oe);
    if constexpr (MIX) {
      // the hop-term fields of the single-output mixing; the odd pass's site
      // operand is loaded once the accumulators are split
      const MixOut mo{a.A2, ls2, a.winv2, a.mix};
      fused_epilogue(mo, acc, oe, ic, 0, okc, gq, lane, field, plane, [&]() { load_in(vb, ic); });
    }
    // odd output sites: the site term S on lane group 0 first, then the
    // accumulators take the hop terms' scale and the 8 hop terms follow
    // against the first layer's even sites (the term's link and the next K
    // step's channels loaded one step ahead, as in k_net_layer_split)
    const Site so = decode(g, ic, 1);
    clear();
    {
      cf v[4][3];
      unpack(vb, v);
      __builtin_amdgcn_sched_barrier(0);
      tg = term_geo(g, so, ic, 1, 0);
      load_link(tg, lraw);
      load_l1(vb, tg, 0);
      __builtin_amdgcn_sched_barrier(0);
      const float f = gq == 0 ? s0 : 0.f;
#pragma unroll
      for (int s2 = 0; s2 < 4; ++s2)
#pragma unroll
        for (int cc = 0; cc < 3; ++cc) v[s2][cc] = f * v[s2][cc];
      products(v, le + (4 * 2) * 2 * 64, 2 * 64);
#pragma unroll
      for (int cc = 0; cc < 3; ++cc) {
        acc[cc][0] *= rs;
        acc[cc][1] *= rs;
      }
    }
#pragma unroll 1
    for (int t = 0; t < 8; ++t) {
      cf m[9];
      term_matrix(lraw, 1, tg.fwd, tg.ph * s1, m);
      TermGeo ng = tg;
      if (t + 1 < 8) {
        ng = term_geo(g, so, ic, 1, t + 1);
        load_link(ng, lraw);
      }
#pragma unroll
      for (int ks = 0; ks < 4; ++ks) {
        cf v[4][3];
        unpack(vb, v);
        transport4(m, v);
        __builtin_amdgcn_sched_barrier(0);
        if (ks + 1 < 4) load_l1(vb, tg, ks + 1);
        else if (t + 1 < 8) load_l1(vb, ng, 0);
        __builtin_amdgcn_sched_barrier(0);
        products(v, lh + ((t * 2 * 4 + ks) * 2) * 64, 4 * 2 * 64);
      }
      tg = ng;
    }
    if constexpr (MIX) {
      if (okc) store(ic, 1, oo);
      // the site-term field; the next chunk's first loads once the accumulators are split
      const MixOut mo{a.A2, ls2, a.winv2, a.mix};
      fused_epilogue(mo, acc, oo, ic, 1, okc, gq, lane, field, plane, [&]() {
        if (c + cstep < cend) begin(c + cstep);
      });
    } else {
      if (c + cstep < cend) begin(c + cstep);  // the next chunk's first loads, ahead of the stores
      if (okc) store(ic, 1, oo);
    }
  }
  if (a.out_max) record_max(a.out_max, omax);
}

// The last layer's site term composed with the layer before it:
// Wc[u][0][i] = W4_8[0][j] W3_u[j][i] (sum over j and the inner spin), the
// weights of a single-output layer over that layer's input (u = 0..8), laid
// out as a (cin 8, cout 1) layer for net_pack_fuse.  FP64 sums, FP32 result.
__global__ void __launch_bounds__(256) k_net_compose_last(const float2 *W3, const float2 *W4, float2 *Wc) {
  for (int e = threadIdx.x; e < 9 * 8 * 16; e += blockDim.x) {
    const int u = e >> 7, i = (e >> 4) & 7, s = (e >> 2) & 3, s2 = e & 3;
    double re = 0.0, im = 0.0;
    for (int j = 0; j < 8; ++j)
      for (int sp = 0; sp < 4; ++sp) {
        const float2 a = W4[((size_t)8 * 8 + j) * 16 + s * 4 + sp];
        const float2 b = W3[(((size_t)u * 8 + j) * 8 + i) * 16 + sp * 4 + s2];
        re += (double)a.x * b.x - (double)a.y * b.y;
        im += (double)a.x * b.y + (double)a.y * b.x;
      }
    Wc[e] = make_float2((float)re, (float)im);
  }
}

int net_compose_last(const float2 *W3, const float2 *W4, float2 *Wc, hipStream_t s) {
  hipLaunchKernelGGL(k_net_compose_last, dim3(1), dim3(256), 0, s, W3, W4, Wc);
  QCU_HIP(hipGetLastError());
  return 0;
}

int net_pack_second(const float2 *W1, const float2 *W2, void *Apk, float *winv, hipStream_t s) {
  hipLaunchKernelGGL(k_net_pack_second, dim3(1), dim3(256), 0, s, W1, W2, (uint4 *)Apk, winv);
  QCU_HIP(hipGetLastError());
  return 0;
}

int net_layer_second(const NetSecondArgs &a, hipStream_t s) {
  QCU_CHECK((size_t)a.g.vh * 96 * 2 + (size_t)a.g.vh * 16 < ((size_t)1 << 32), QCU_ERR_DIMS,
            "net: lattice too large for the second layer");
  static int ncu = 0;
  if (!ncu) {
    int dev = 0;
    QCU_HIP(hipGetDevice(&dev));
    QCU_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    if (ncu <= 0) ncu = 256;
  }
  // 8 waves per block, one block per CU (the operands fill 148 KiB of its LDS, 156 KiB with the mixing)
  constexpr int waves = 8;
  const int nchunk = (a.g.vh + 31) / 32;
  int nblk = ncu;
  if (nchunk < nblk * waves) nblk = (nchunk + waves - 1) / waves;
  if (a.A2) hipLaunchKernelGGL((k_net_second<waves, true>), dim3(nblk), dim3(64 * waves), 0, s, a);
  else hipLaunchKernelGGL((k_net_second<waves, false>), dim3(nblk), dim3(64 * waves), 0, s, a);
  QCU_HIP(hipGetLastError());
  return 0;
}

// MFMA shape of the split kernel: 32 sites per wave (v_mfma_f32_32x32x16_f16)
// or 16 (v_mfma_f32_16x16x32_f16); QCU_NET_SPLIT_SHAPE=16 selects the second
// (read when a network is created: its packed weights follow the shape)
int net_split_shape_default() {
  const char *e = getenv("QCU_NET_SPLIT_SHAPE");
  return (e && atoi(e) == 16) ? 16 : 32;
}

int net_pack_split(const float2 *W, int cin, int cout, void *Apk, float *winv, int ns, hipStream_t s) {
  QCU_CHECK(cin == 8 && (cout == 4 || cout == 8), QCU_ERR_ARG, "net: split-FP16 layer needs 8 -> 4 or 8 channels");
  QCU_CHECK(ns == 16 || ns == 32, QCU_ERR_ARG, "net: split shape 16 or 32");
  uint4 *A = (uint4 *)Apk;
  if (ns == 32) {
    if (cout == 8) hipLaunchKernelGGL((k_net_pack_split<32, 2>), dim3(1), dim3(256), 0, s, W, cin, cout, A, winv);
    else hipLaunchKernelGGL((k_net_pack_split<32, 1>), dim3(1), dim3(256), 0, s, W, cin, cout, A, winv);
  } else {
    if (cout == 8) hipLaunchKernelGGL((k_net_pack_split<16, 4>), dim3(1), dim3(256), 0, s, W, cin, cout, A, winv);
    else hipLaunchKernelGGL((k_net_pack_split<16, 2>), dim3(1), dim3(256), 0, s, W, cin, cout, A, winv);
  }
  QCU_HIP(hipGetLastError());
  return 0;
}

bool net_split_can_fuse(int ns) { return ns == 32; }

int net_pack_fuse(const float2 *W, void *Apk, float *winv, hipStream_t s) {
  hipLaunchKernelGGL(k_net_pack_fuse, dim3(1), dim3(256), 0, s, W, (uint4 *)Apk, winv);
  QCU_HIP(hipGetLastError());
  return 0;
}

int net_layer_split(const NetSplitArgs &a, int cout, hipStream_t s) {
  QCU_CHECK(a.npar == 1 || a.npar == 2, QCU_ERR_ARG, "net: one or two output parities");
  const int ns = a.ns;
  QCU_CHECK(ns == 16 || ns == 32, QCU_ERR_ARG, "net: split shape 16 or 32");
  static int wenv = -1;
  if (wenv < 0) {
    const char *e = getenv("QCU_NET_SPLIT_WAVES");
    wenv = e ? atoi(e) : 0;
  }
  const int waves = (wenv == 4 || wenv == 8 || (wenv == 12 && ns == 16)) ? wenv : (ns == 32 ? 8 : 12);
  // per-lane byte offsets within a K step's channels: (64 / NS - 1) x 192 + 6 x 16 bytes per site
  QCU_CHECK((size_t)a.g.vh * ((64 / ns - 1) * 192 + 96) < ((size_t)1 << 32), QCU_ERR_DIMS,
            "net: lattice too large for the split-FP16 layer");
  static int ncu = 0;
  if (!ncu) {
    int dev = 0;
    QCU_HIP(hipGetDevice(&dev));
    QCU_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    if (ncu <= 0) ncu = 256;
  }
  const int nchunk = a.npar * ((a.g.vh + ns - 1) / ns);
  // one block per CU (the weights fill its LDS); small lattices: no more blocks than chunks need
  int nblk = nchunk < ncu * waves ? (nchunk + waves - 1) / waves : ncu;
  if (a.order >= 1 && a.order <= 3) nblk = (nblk + 7) & ~7;
  const dim3 grid(nblk), block(64 * waves);
#define QCU_SPLIT_LAUNCH(NS_, W_)                                                          \
  if (ns == NS_ && waves == W_) {                                                          \
    if (cout == 8) hipLaunchKernelGGL((k_net_layer_split<NS_, W_, 64 / NS_>), grid, block, 0, s, a); \
    else hipLaunchKernelGGL((k_net_layer_split<NS_, W_, 32 / NS_>), grid, block, 0, s, a);           \
  }
  const bool diag = a.order == 3 && ns == 16 && a.g.lxh == 16 && a.npar == 2 && (a.g.lz & 15) == 0 && cout == 8 &&
                    !a.A2;
  if (diag) {
    const int nb8 = (a.g.vh * 2 / 16 + 8 * 8 - 1) / (8 * 8);  // blocks of 8 waves, at most one per CU
    const dim3 g8((unsigned)(((nb8 < ncu ? nb8 : ncu) + 7) & ~7)), b8(64 * 8);
    hipLaunchKernelGGL((k_net_layer_split<16, 8, 4, false, true>), g8, b8, 0, s, a);
  } else if (a.A2) {
    QCU_CHECK(ns == 32 && cout == 8 && a.mix && a.winv2, QCU_ERR_ARG, "net: fused layer needs the 32-site shape");
    if (waves == 4) hipLaunchKernelGGL((k_net_layer_split<32, 4, 2, true>), grid, block, 0, s, a);
    else hipLaunchKernelGGL((k_net_layer_split<32, 8, 2, true>), grid, block, 0, s, a);
  } else {
    QCU_SPLIT_LAUNCH(32, 4) QCU_SPLIT_LAUNCH(32, 8) QCU_SPLIT_LAUNCH(16, 4) QCU_SPLIT_LAUNCH(16, 8)
    QCU_SPLIT_LAUNCH(16, 12)
  }
#undef QCU_SPLIT_LAUNCH
  QCU_HIP(hipGetLastError());
  return 0;
}

}  // namespace qcu
