// qcu_net.h -- lattice geometry of the learned preconditioner's layers,
// shared by its kernels: channel fields, the neighbour and link of each term,
// and the colour transport.
#pragma once
#include <cmath>

#include "qcu_common.h"

namespace qcu {

typedef cplx<float> cf;

// channel c, parity p of a channel-major activation block: one FP32 spinor
// field of vh sites (96 bytes per site) per (channel, parity)
QCU_D SpinorPtr chan_field(const SpinorPtr &base, int c, int p, int vh) {
  SpinorPtr f;
  f.p = (char *)base.p + ((size_t)c * 2 + p) * (size_t)vh * 96;
  f.norm = nullptr;
  f.stride = vh;
  return f;
}

// neighbour site (half index) and link for term t at output site i of parity p
struct TermGeo {
  int j;      // input site (half index)
  int jp;     // input parity
  int hop;    // 0: site term
  int fwd;
  int slab;   // link slab (parity * 4 + mu)
  int li;     // link site
  float ph;
};

QCU_D TermGeo term_geo(const Geom &g, const Site &s, int i, int p, int t) {
  TermGeo r;
  r.j = i;
  r.jp = p;
  r.hop = t < 8;
  r.fwd = !(t & 1);
  r.slab = 0;
  r.li = i;
  r.ph = 1.f;
  if (!r.hop) return r;
  const int mu = t >> 1;
  const int q = 1 - p;
  const int sy = g.lxh, sz = g.lxh * g.ly, st_ = sz * g.lz;
  int j = i;
  if (mu == 0) {
    const int xe = 1 - s.xodd;
    j = r.fwd ? i + s.xodd - s.xodd * (int)(s.xh == g.lxh - 1) * g.lxh : i - xe + xe * (int)(s.xh == 0) * g.lxh;
  } else if (mu == 1) {
    j = r.fwd ? i + sy - (int)(s.y == g.ly - 1) * g.ly * sy : i - sy + (int)(s.y == 0) * g.ly * sy;
  } else if (mu == 2) {
    j = r.fwd ? i + sz - (int)(s.z == g.lz - 1) * g.lz * sz : i - sz + (int)(s.z == 0) * g.lz * sz;
  } else {
    j = r.fwd ? i + st_ - (int)(s.t == g.lt - 1) * g.lt * st_ : i - st_ + (int)(s.t == 0) * g.lt * st_;
    const bool cross = r.fwd ? (s.t == g.lt - 1) : (s.t == 0);
    r.ph = (cross && g.tbc < 0) ? -1.f : 1.f;
  }
  r.j = j;
  r.jp = q;
  r.slab = r.fwd ? p * 4 + mu : q * 4 + mu;
  r.li = r.fwd ? i : j;
  return r;
}

// The same geometry for a term t that is uniform over the wave, without
// branches: the direction-dependent quantities are selected, not branched on,
// so that a loop over sites with a fixed term has one straight-line body.  The
// term's constants are formed once (TermConst, scalar registers) as sums of
// products and masks rather than selections among struct fields, which the
// compiler turns into indexed accesses to a stack copy of the struct.
struct TermConst {
  int t, mu, fwd, hop;
  int len, str;  // extent and half-site stride of direction mu (mu > 0)
  int aper;      // hops across the t boundary carry -1
  int sel[3];    // all-ones for the coordinate (y, z, t) along mu, else 0
};
QCU_D TermConst term_const(const Geom &g, int t) {
  TermConst c;
  c.t = t;
  c.mu = t >> 1;
  c.fwd = !(t & 1);
  c.hop = t < 8;
  const int m1 = c.mu == 1, m2 = c.mu == 2, m3 = c.mu == 3;
  c.len = m1 * g.ly + m2 * g.lz + m3 * g.lt;
  c.str = m1 * g.lxh + m2 * g.lxh * g.ly + m3 * g.lxh * g.ly * g.lz;
  c.aper = m3 * (g.tbc < 0);
  c.sel[0] = -m1;
  c.sel[1] = -m2;
  c.sel[2] = -m3;
  return c;
}
QCU_D TermGeo term_geo_u(const Geom &g, const Site &s, int i, int p, const TermConst &k) {
  // x: the even-odd neighbour in the half-site index
  const int xe = 1 - s.xodd;
  const int jx = k.fwd ? i + s.xodd - s.xodd * (int)(s.xh == g.lxh - 1) * g.lxh
                       : i - xe + xe * (int)(s.xh == 0) * g.lxh;
  // y, z, t: a stride with periodic wrap
  // the coordinate along mu by masks (a selection among the Site fields
  // becomes an indexed access to a stack copy of the Site)
  const int c = (s.y & k.sel[0]) | (s.z & k.sel[1]) | (s.t & k.sel[2]);
  const int cross = k.fwd ? (int)(c == k.len - 1) : (int)(c == 0);
  const int jo = k.fwd ? i + k.str - cross * k.len * k.str : i - k.str + cross * k.len * k.str;
  TermGeo r;
  r.j = !k.hop ? i : k.mu == 0 ? jx : jo;
  r.jp = k.hop ? 1 - p : p;
  r.hop = k.hop;
  r.fwd = k.fwd;
  // the site term gets a valid link (slab 0 at its own site), never applied
  r.slab = !k.hop ? 0 : k.fwd ? p * 4 + k.mu : (1 - p) * 4 + k.mu;
  r.li = k.fwd ? i : r.j;
  r.ph = (k.hop && k.aper && cross) ? -1.f : 1.f;
  return r;
}

// the colour matrix of a term, applied as v <- M v: M = ph U for a forward
// hop and ph U^dag for a backward one, from the two stored rows of U; the
// identity for the site term
QCU_D void term_matrix(const float4 (&raw)[3], int hop, int fwd, float ph, cf (&m)[9]) {
  if (!hop) {
#pragma unroll
    for (int k = 0; k < 9; ++k) m[k] = mk(k % 4 == 0 ? 1.f : 0.f, 0.f);
    return;
  }
  cf u[9];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    u[2 * k] = mk(raw[k].x, raw[k].y);
    u[2 * k + 1] = mk(raw[k].z, raw[k].w);
  }
  recon_row2(u);
  if (fwd) {
#pragma unroll
    for (int k = 0; k < 9; ++k) m[k] = mk(ph * u[k].re, ph * u[k].im);
  } else {
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) m[r * 3 + c] = mk(ph * u[c * 3 + r].re, -ph * u[c * 3 + r].im);
  }
}

// v <- M v as a chain of fused multiply-adds
QCU_D void transport1(const cf (&m)[9], cf (&v)[3]) {
  cf a[3];
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    float re = m[r * 3].re * v[0].re, im = m[r * 3].re * v[0].im;
    re = fmaf(-m[r * 3].im, v[0].im, re);
    im = fmaf(m[r * 3].im, v[0].re, im);
#pragma unroll
    for (int c = 1; c < 3; ++c) {
      re = fmaf(m[r * 3 + c].re, v[c].re, re);
      re = fmaf(-m[r * 3 + c].im, v[c].im, re);
      im = fmaf(m[r * 3 + c].re, v[c].im, im);
      im = fmaf(m[r * 3 + c].im, v[c].re, im);
    }
    a[r] = mk(re, im);
  }
#pragma unroll
  for (int r = 0; r < 3; ++r) v[r] = a[r];
}

struct SysState;

// Hidden layer on the FP16 matrix cores (net_split.hip): 8 input channels,
// 4 or 8 output channels, both in the channel-major FP32 layout.
struct NetSplitArgs {
  const char *in;           // input channel fields: (c * 2 + parity) * vh * 96 bytes
  char *out;                // output channel fields, same layout
  const uint4 *A;           // packed split weights (net_pack_split)
  const float *winv;        // inverse of the weights' power-of-two scale
  const unsigned *in_max;   // largest |component| of the input (float bits)
  unsigned *out_max;        // receives the largest |component| of the output (or null)
  LinkPtr U;                // FP32 links
  Geom g;
  int parity0, npar;        // output parities parity0 .. parity0 + npar - 1
  int ns;                   // sites per wave: 32 (v_mfma_f32_32x32x16_f16) or 16 (16x16x32), as packed
  int order;                // chunk order: 0 interleaved over the grid, 1 one eighth per XCD, 2 a z slab per
                            // XCD, 3 z-plane pairs per XCD with the diagonal term schedule (16 sites, Lx = 32)
  const SysState *st;       // early-out flag of a captured solver iteration (or null)
  // fused single-output layer (8 -> 8 layers only, when the next layer maps 8
  // channels to 1 and only its mixed term fields are needed): instead of its
  // output channels the kernel writes, per output site, the next layer's
  // mixed term fields Y_t = W'_t out (hop terms at even sites, the site term
  // at odd sites) into mix, in the channel layout (term t, parity q at slot
  // (t, q)); null A2: no fusion
  const uint4 *A2;          // packed W' (net_pack_fuse)
  const float *winv2;       // its inverse scale
  char *mix;
};
int net_pack_split(const float2 *W, int cin, int cout, void *Apk, float *winv, int ns, hipStream_t s);
int net_split_shape_default();
int net_layer_split(const NetSplitArgs &a, int cout, hipStream_t s);
// the 8 -> 1 weights W'[t][0][i][4x4] as the fused epilogue's A operands
int net_pack_fuse(const float2 *W, void *Apk, float *winv, hipStream_t s);
// whether net_layer_split can fuse the next single-output layer (32-site shape)
bool net_split_can_fuse(int ns);
constexpr size_t kSplitPackBytes = (size_t)9 * 2 * 4 * 2 * 64 * 16;  // A operands of an 8 -> 8 layer
constexpr size_t kFusePackBytes = (size_t)3 * 4 * 2 * 64 * 16;       // those of the fused 8 -> 1 mixing
constexpr int kFirstPackOps = 5 * 2 * 2 * 64;                         // A operands of the first (1 -> 8) layer

// the first layer (1 -> 8 channels) on an input that is zero on the even
// sites: act[0] -> act[1], split FP16 (net_layer_first)
struct NetFirstArgs {
  const char *in;           // the input channel field (both parities; only the odd sites are read)
  char *out;                // the 8 output channel fields: (c * 2 + parity) * vh * 96 bytes
  const uint4 *A;           // packed operands (net_pack_first)
  const float *winv;        // inverse of the weights' power-of-two scale
  const unsigned *in_max;   // largest |component| of the input (float bits)
  unsigned *out_max;        // receives that of the output (or null)
  LinkPtr U;                // FP32 links
  Geom g;
  const SysState *st;       // early-out flag of a captured solver iteration (or null)
};
int net_pack_first(const float2 *W, void *Apk, float *winv, hipStream_t s);
int net_layer_first(const NetFirstArgs &a, hipStream_t s, bool even_only = false);

// the second layer (8 -> 8) after a first layer on an input zero on the even
// sites, with the first layer's odd sites left implicit (net_layer_second):
// even outputs from the input alone, odd outputs from the first layer's even
// sites and the input
struct NetSecondArgs {
  const char *in;           // the network input (channel 0 of both parities; the odd sites are read)
  const char *l1;           // the first layer's 8 channels (both parities; the even sites are read)
  char *out;                // the 8 output channels
  const uint4 *A;           // combined operands E_t, S (net_pack_second)
  const float *winv;        // their inverse scales: [0] E, [1] S
  const uint4 *Ah;          // the second layer's own split operands (net_pack_split; terms 0..7 used)
  const float *winvh;       // their inverse scale
  const unsigned *in_max;   // largest |component| of the input (float bits)
  const unsigned *l1_max;   // that of the first layer's even sites
  unsigned *out_max;        // receives that of the output (or null)
  LinkPtr U;                // FP32 links
  Geom g;
  const SysState *st;       // early-out flag of a captured solver iteration (or null)
  // optional: the last layer's site-term mixing composed with the next layer,
  // written as the single-output layer's term fields (hop terms at even
  // sites, the site term at odd ones; see NetSplitArgs); null A2: none
  const uint4 *A2 = nullptr;
  const float *winv2 = nullptr;
  char *mix = nullptr;
};
int net_pack_second(const float2 *W1, const float2 *W2, void *Apk, float *winv, hipStream_t s);
// Wc[u][0][i][4x4] = W4_8[0][j] W3_u[j][i] (8 -> 8 layer W3, 8 -> 1 layer W4; u = 0..8)
int net_compose_last(const float2 *W3, const float2 *W4, float2 *Wc, hipStream_t s);
int net_layer_second(const NetSecondArgs &a, hipStream_t s);

typedef _Float16 h8v __attribute__((ext_vector_type(8)));

// exact power of two that brings a largest magnitude mx below 2^top (1 for a
// zero or non-finite mx; the exponent is clamped so that the factor is finite)
QCU_HD float pow2_scale(float mx, int top) {
  if (!(mx > 0.f) || !(mx <= 3.4e38f)) return 1.f;
  int e;
  (void)frexpf(mx, &e);  // mx = f 2^e with 0.5 <= f < 1, so mx < 2^e
  e = e < -110 ? -110 : e;
  return ldexpf(1.f, top - e);
}

QCU_D h8v as_h8(uint4 u) { return __builtin_bit_cast(h8v, u); }

// the two FP16 halves of a complex FP32 value: hi = (RN(re), RN(im)) packed,
// lo = the rounded remainders.  The remainder x - hi is exact in FP32; one
// v_fma_mix_f32 per component forms it from the packed FP16 operand directly.
QCU_D void split2(float re, float im, unsigned &hi, unsigned &lo) {
  typedef _Float16 h2v __attribute__((ext_vector_type(2)));
  typedef float f2v __attribute__((ext_vector_type(2)));
  const h2v h = __builtin_convertvector((f2v){re, im}, h2v);
  hi = __builtin_bit_cast(unsigned, h);
  float rr, ri;
  asm("v_fma_mix_f32 %0, -%1, 1.0, %2 op_sel_hi:[1,0,0]" : "=v"(rr) : "v"(hi), "v"(re));
  asm("v_fma_mix_f32 %0, -%1, 1.0, %2 op_sel:[1,0,0] op_sel_hi:[1,0,0]" : "=v"(ri) : "v"(hi), "v"(im));
  const h2v l = __builtin_convertvector((f2v){rr, ri}, h2v);
  lo = __builtin_bit_cast(unsigned, l);
}

// Records the largest |component| a wave wrote into *slot: the wave maximum,
// then one atomic maximum on the float bits (non-negative floats order like
// unsigned integers).  Every lane of the wave must call it.
QCU_D void record_max(unsigned *slot, float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off));
  if ((threadIdx.x & 63) == 0) {
    const unsigned b = __float_as_uint(v);
    if (b > *(volatile unsigned *)slot) atomicMax(slot, b);
  }
}

}  // namespace qcu
