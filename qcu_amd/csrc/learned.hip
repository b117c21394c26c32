// learned.hip -- learned gauge-covariant linear preconditioner (new; no
// counterpart in the reference).
//
// Network: L layers on the full lattice, layer l maps chan[l] -> chan[l+1]
// channels of spinor fields:
//   out_o(x) = sum_i [ sum_mu ( A^{oi}_mu U_mu(x) in_i(x+mu) + B^{oi}_mu U_mu(x-mu)^dag in_i(x-mu) )
//                      + C^{oi} in_i(x) ]
// with learned 4x4 complex spin matrices A, B, C (acting on each colour) and a
// -1 phase on hops across the time boundary for antiperiodic lattices.  Terms
// t = 2 mu (forward hop), 2 mu + 1 (backward hop), 8 (site).  Everything is
// FP32.  Weights are W[l][t][o][i][4x4] complex.
//
// As a right preconditioner of the odd-site Schur operator the input (an odd
// vector) is embedded in the full lattice with zero even sites, and the odd
// part of the single output channel is kept.
//
// Kernels:
// * layer forward on the matrix cores (k_net_layer_mfma: the channel and spin
//   mixing as real FP32 MFMA products, one colour transport per lane and step)
//   for layers with several output channels; a single-output layer mixes its
//   input channels per term on the matrix cores (k_net_mix) and transports
//   one mixed field per term (k_net_gather); the vector-unit kernel
//   (k_net_layer, one thread per site and group of output channels) is the
//   cross-check;
// * the data backward is the forward kernel with conjugate-transposed weights
//   and swapped hop directions;
// * the weight gradient is a deterministic two-stage reduction of
//   g_o(x) (T_t a_i)(x)^dag over sites and colours (matrix cores, with a
//   vector cross-check);
// * Adam runs on the device.
// Training minimises |Mhat K b - b|^2 / |b|^2.
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/qcu.h"
#include "qcu_batch.h"
#include "qcu_lattice.h"
#include "qcu_reduce.h"
#include "qcu_solver.h"
#include "qcu_net.h"

namespace qcu {


constexpr int kNetMaxLayers = 8;
constexpr int kNetMaxChan = 8;
constexpr int kWgradBlocks = 120;  // site chunks per weight-gradient combination

// Order in which the waves of the matrix-core layer visit the lattice.  A wave
// takes a chunk of 16 consecutive half-sites.  In the tiled order (an x row
// is a whole number of chunks) the two output parities of a chunk are
// neighbours, t runs fastest over the rows of a ty x tz cross-section tile,
// and the tiles come last; with the XCD-aware block remap each XCD walks whole
// tiles, so the t, y and z neighbours a wave reads are read by waves running
// at about the same time on the same XCD and come from its L2.  The input
// parity a site term reads is the one the hop terms of the other output
// parity read, so running both parities together also shares those reads.
struct WaveOrder {
  int tiled;   // 0: chunks in site order, parity from the grid's y index
  int nx;      // chunks per x row
  int ty, tz;  // rows per tile in y and z
  int nyt;     // tiles along y
  int npar;    // output parities (1 or 2)
  int nwaves;  // waves in the tiled grid
};

struct NetLayerArgs {
  SpinorPtr in;   // channel c, parity p at in.p + (c * 2 + p) * vh * 96 bytes, stride vh
  SpinorPtr out;  // same layout
  const float2 *W;
  LinkPtr U;      // FP32 links
  Geom g;
  int cin, cout;
  int parity0;    // first output parity (the grid's y index adds to it)
  int zero_par;   // parity on which the input is known to be zero (-1: none)
  int ngroups;    // output-channel groups
  const SysState *st;
  WaveOrder ord;  // matrix-core kernel only
  unsigned *out_max;  // matrix-core kernel: receives the largest |component| written (or null)
};

// the colour matrix of term tg at this site (FP32 links)
// NT: non-temporal loads (links the kernel reads once)
template <bool NT = false>
QCU_D void term_link_matrix(const LinkPtr &U, const TermGeo &tg, cf (&m)[9]) {
  float4 raw[3];
  if (tg.hop) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const size_t cb = ((size_t)tg.slab * 3 + k) * U.vh * 16;
      raw[k] = NT ? ld_nt(U.a, cb, (unsigned)tg.li * 16u) : ld_at<float4>(U.a, cb, (unsigned)tg.li * 16u);
    }
  }
  term_matrix(raw, tg.hop, tg.fwd, tg.ph, m);
}

// v <- M v on each of the four spins of a spinor
QCU_D void transport_spins(const cf (&m)[9], cf (&v)[12]) {
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    cf w[3] = {v[s * 3], v[s * 3 + 1], v[s * 3 + 2]};
    transport1(m, w);
#pragma unroll
    for (int c = 0; c < 3; ++c) v[s * 3 + c] = w[c];
  }
}

// grid: x = site blocks x output-channel groups (group fastest, so that the
// groups of one site block run together and share its loads in L2; the
// XCD-aware remap keeps neighbouring site blocks on one L2), y = output parity.
// CO output channels per thread; CI: input channel count known at compile time
// (0: a.cin at run time).  The term is the outer loop.  The colour transport and
// the spin matrices act on different indices, so they commute: with fewer
// output than input channels the input channels are first mixed at the
// neighbour and each output channel is transported once per term
// (sum_i W_i (M v_i) = M sum_i W_i v_i), which saves (cin - CO) / cin of the
// transports of the 8 -> 1 layer.
template <int CO, int CI>
__global__ void __launch_bounds__(256) k_net_layer(NetLayerArgs a) {
  if (a.st && a.st->done) return;
  const Geom g = a.g;
  const int blk = remap_block(blockIdx.x, gridDim.x, 1);
  const int sb = blk / a.ngroups;
  const int og = (blk - sb * a.ngroups) * CO;
  const int i = sb * blockDim.x + threadIdx.x;
  const int p = a.parity0 + blockIdx.y;
  if (i >= g.vh) return;
  const Site site = decode(g, i, p);
  cf acc[CO][12];
#pragma unroll
  for (int o = 0; o < CO; ++o)
#pragma unroll
    for (int k = 0; k < 12; ++k) acc[o][k] = mk(0.f, 0.f);
  const int cin = CI > 0 ? CI : a.cin, cout = a.cout;
  const int t_lo = (a.zero_par == 1 - p) ? 8 : 0, t_hi = (a.zero_par == p) ? 8 : 9;
  // one output channel per thread: mixing first costs no extra live accumulators
  const bool mix_first = CO == 1 && cin > 1;
  // acc_o[s][c] += sum_s2 W[t][o][ci][s][s2] v[s2][c]
  auto mix = [&](cf (&dst)[CO][12], const cf (&v)[12], int t, int ci) {
#pragma unroll
    for (int o = 0; o < CO; ++o) {
      if (og + o >= cout) break;
      const float2 *w = a.W + (((size_t)t * cout + og + o) * cin + ci) * 16;
#pragma unroll
      for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int s2 = 0; s2 < 4; ++s2) {
          const float2 ww = w[s * 4 + s2];
          const cf wc = mk(ww.x, ww.y);
#pragma unroll
          for (int c = 0; c < 3; ++c) cmad(dst[o][s * 3 + c], wc, v[s2 * 3 + c]);
        }
    }
  };
#pragma unroll 1
  for (int t = t_lo; t < t_hi; ++t) {
    const TermGeo tg = term_geo(g, site, i, p, t);
    cf m[9];
    term_link_matrix(a.U, tg, m);
    if constexpr (CO == 1) {
      if (mix_first) {
        cf tmp[1][12];
#pragma unroll
        for (int k = 0; k < 12; ++k) tmp[0][k] = mk(0.f, 0.f);
#pragma unroll 1
        for (int ci = 0; ci < cin; ++ci) {
          cf v[12];
          SpinorIO<P32>::load(chan_field(a.in, ci, tg.jp, g.vh), tg.j, v);
          mix(tmp, v, t, ci);
        }
        if (tg.hop) transport_spins(m, tmp[0]);
#pragma unroll
        for (int k = 0; k < 12; ++k) acc[0][k] = acc[0][k] + tmp[0][k];
        continue;
      }
    }
#pragma unroll 1
    for (int ci = 0; ci < cin; ++ci) {
      cf v[12];
      SpinorIO<P32>::load(chan_field(a.in, ci, tg.jp, g.vh), tg.j, v);
      if (tg.hop) transport_spins(m, v);
      mix(acc, v, t, ci);
    }
  }
#pragma unroll
  for (int o = 0; o < CO; ++o) {
    if (og + o >= cout) break;
    SpinorIO<P32>::store(chan_field(a.out, og + o, p, g.vh), i, acc[o]);
  }
}

// ---- matrix-core layer -----------------------------------------------------
typedef float f4v __attribute__((ext_vector_type(4)));

// A operands of one layer: Apk[((t * cin + i) * MT + mt) * 3 + part][lane],
// part 0 = Re W, 1 = Im W, 2 = Re W + Im W, row m = lane % 16 -> (o = 4 mt + m / 4,
// s = m % 4), column k = lane / 16 -> s2
constexpr int kApkParts = 3;
__global__ void k_net_pack(const float2 *W, float *Apk, int cin, int cout, int mtiles) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  const int total = 9 * cin * mtiles * kApkParts * 64;
  if (idx >= total) return;
  const int lane = idx & 63;
  int r = idx >> 6;
  const int part = r % kApkParts;
  r /= kApkParts;
  const int mt = r % mtiles;
  r /= mtiles;
  const int i = r % cin, t = r / cin;
  const int ml = lane & 15, s2 = lane >> 4;
  const int o = 4 * mt + (ml >> 2), sp = ml & 3;
  float v = 0.f;
  if (o < cout) {
    const float2 w = W[(((size_t)t * cout + o) * cin + i) * 16 + sp * 4 + s2];
    v = part == 0 ? w.x : part == 1 ? w.y : w.x + w.y;
  }
  Apk[idx] = v;
}

// The channel and spin mixing of one layer is a complex GEMM per colour,
//   D[m][n] = sum_k W[m][k] V[k][n],  m = (output channel o, spin s),
//   k = (term t, input channel i, spin s2),  n = site,
// done as three real products on v_mfma_f32_16x16x4_f32 in FP32 (the
// three-multiplication complex product):
//   P1 += Wr Vr,  P2 += Wi Vi,  P3 += (Wr + Wi)(Vr + Vi),
//   Re D = P1 - P2,  Im D = P3 - P1 - P2,
// with Wr + Wi packed beside the weights and Vr + Vi formed per step; three
// products instead of four cut the matrix-core work by a quarter.
// Lane l of an MFMA step holds A[m = l % 16][k = l / 16], B[k = l / 16][n = l % 16]
// and returns D[m = 4 (l / 16) + j][n = l % 16], j = 0..3.  A K step is the four
// spins s2 of one (t, i): lane group g = l / 16 transports spin g of its site,
// so no transport is repeated.  An M tile is four output channels times four
// spins, so lane group g ends with all 12 colour-spin components of channel
// 4 mt + g at its site.  Each wave owns 16 sites.
//
// The term is the outer loop: the colour matrix of each term is built once and
// the input channels stream through it; the link of the next term is loaded
// while the current one runs.  The spinor and weight operands of each step are
// loaded one step ahead into the other of two operand sets, which swap roles
// every step; no load sits behind a branch (the last step loads a valid dummy),
// so each step waits only for its own operands.
struct MfmaOps {
  float2 s[3];        // the lane's spin: colours 0..2
  float w[2][3];      // [mt][Re W, Im W, Re W + Im W]
};

template <int MT, int CI>
QCU_D void net_layer_mfma_body(const NetLayerArgs &a, const float *Apk) {
  if (a.st && a.st->done) return;
  const Geom g = a.g;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int blk = remap_block(blockIdx.x, gridDim.x, 1);
  const int w = blk * (int)(blockDim.x >> 6) + wave;
  int base, p;
  if (a.ord.tiled) {
    if (w >= a.ord.nwaves) return;  // uniform over the wave
    const WaveOrder &o = a.ord;
    int r = w;
    p = a.parity0;
    if (o.npar == 2) {
      p += r & 1;
      r >>= 1;
    }
    const int xc = r % o.nx;
    r /= o.nx;
    const int yy = r % o.ty;
    r /= o.ty;
    const int zz = r % o.tz;
    r /= o.tz;
    const int t = r % g.lt;
    r /= g.lt;
    const int y = (r % o.nyt) * o.ty + yy, z = (r / o.nyt) * o.tz + zz;
    base = ((t * g.lz + z) * g.ly + y) * g.lxh + xc * 16;
  } else {
    base = w * 16;
    p = a.parity0 + blockIdx.y;
    if (base >= g.vh) return;  // uniform over the wave
  }
  const int gq = lane >> 4;
  int i = base + (lane & 15);
  const bool ok = i < g.vh;
  if (!ok) i = base;
  const Site site = decode(g, i, p);
  // spin gq of a site is floats 6 gq .. 6 gq + 5 of its 24; each float2 of it
  // lies in one float4 plane of the site-fastest layout
  const unsigned vh16 = (unsigned)g.vh * 16u;
  unsigned poff[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const int f = 6 * gq + 2 * k;
    poff[k] = (unsigned)(f >> 2) * vh16 + (unsigned)(f & 3) * 4u;
  }
  f4v p1[3][MT], p2[3][MT], p3[3][MT];
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
      p1[c][mt] = (f4v){0.f, 0.f, 0.f, 0.f};
      p2[c][mt] = (f4v){0.f, 0.f, 0.f, 0.f};
      p3[c][mt] = (f4v){0.f, 0.f, 0.f, 0.f};
    }
  const int cin = CI > 0 ? CI : a.cin;
  // terms that read the zero parity are skipped: hop terms read parity 1 - p,
  // the site term parity p
  int t_lo = 0, t_hi = 9;
  if (a.zero_par == 1 - p) t_lo = 8;
  if (a.zero_par == p) t_hi = 8;
  const int q0 = t_lo * cin, nq = t_hi * cin;
  const size_t field = (size_t)g.vh * 96, chan_bytes = 2 * field;
  // per term: the input field's parity base and the lane's three offsets
  struct TermIn {
    const char *fb;
    unsigned o[3];
  };
  auto term_in = [&](const TermGeo &tg) {
    TermIn r;
    r.fb = (const char *)a.in.p + (size_t)tg.jp * field;
    const unsigned jo = (unsigned)tg.j * 16u;
#pragma unroll
    for (int k = 0; k < 3; ++k) r.o[k] = jo + poff[k];
    return r;
  };
  auto link_raw = [&](const TermGeo &tg, float4 (&l)[3]) {
#pragma unroll
    for (int k = 0; k < 3; ++k) l[k] = ld_at<float4>(a.U.a, ((size_t)tg.slab * 3 + k) * a.U.vh * 16, (unsigned)tg.li * 16u);
  };
  auto load_ops = [&](MfmaOps &x, const TermIn &ti, int ci, int q) {
    const char *cb = ti.fb + (size_t)ci * chan_bytes;
#pragma unroll
    for (int k = 0; k < 3; ++k) x.s[k] = ld_at<float2>(cb, 0, ti.o[k]);
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
      for (int part = 0; part < 3; ++part)
        x.w[mt][part] = ld_at<float>(Apk, (((size_t)q * MT + mt) * 3 + part) * 256, (unsigned)lane * 4u);
  };
  auto compute = [&](const MfmaOps &x, const cf (&m)[9]) {
    cf v[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) v[k] = mk(x.s[k].x, x.s[k].y);
    transport1(m, v);  // the site term's matrix is the identity
    // lanes past the last site hold a copy of a valid site: their columns of D
    // are computed and never stored
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const float br = v[c].re, bi = v[c].im, bs = br + bi;
#pragma unroll
      for (int mt = 0; mt < MT; ++mt) {
        p1[c][mt] = __builtin_amdgcn_mfma_f32_16x16x4f32(x.w[mt][0], br, p1[c][mt], 0, 0, 0);
        p2[c][mt] = __builtin_amdgcn_mfma_f32_16x16x4f32(x.w[mt][1], bi, p2[c][mt], 0, 0, 0);
        p3[c][mt] = __builtin_amdgcn_mfma_f32_16x16x4f32(x.w[mt][2], bs, p3[c][mt], 0, 0, 0);
      }
    }
  };
  // prologue: first term's link and first step's operands
  TermGeo tg = term_geo(g, site, i, p, t_lo);
  TermIn cur = term_in(tg);
  float4 nl[3];
  if (tg.hop) link_raw(tg, nl);
  MfmaOps A, B;
  load_ops(A, cur, 0, q0);
#pragma unroll 1
  for (int t = t_lo; t < t_hi; ++t) {
    cf m[9];
    term_matrix(nl, tg.hop, tg.fwd, tg.ph, m);
    // the next term: geometry and link (a dummy copy of this term after the last)
    TermGeo ng = tg;
    TermIn nxt = cur;
    if (t + 1 < t_hi) {
      ng = term_geo(g, site, i, p, t + 1);
      nxt = term_in(ng);
      if (ng.hop) link_raw(ng, nl);
    }
    const int qt = t * cin;
    const int qn = qt + cin < nq ? qt + cin : nq - 1;  // first step of the next term
    // the scheduling barriers keep each prefetch ahead of the step it overlaps
    if constexpr (CI > 0 && CI % 2 == 0) {
      // a compile-time even channel count: the term is straight-line code, so
      // each step waits for exactly its own operands
#pragma unroll
      for (int ci = 0; ci < CI; ci += 2) {
        load_ops(B, cur, ci + 1, qt + ci + 1);
        __builtin_amdgcn_sched_barrier(0);
        compute(A, m);
        if (ci + 2 < CI) load_ops(A, cur, ci + 2, qt + ci + 2);
        else load_ops(A, nxt, 0, qn);
        __builtin_amdgcn_sched_barrier(0);
        compute(B, m);
      }
    } else {
      int ci = 0;
#pragma unroll 1
      for (; ci + 1 < cin; ci += 2) {
        load_ops(B, cur, ci + 1, qt + ci + 1);
        __builtin_amdgcn_sched_barrier(0);
        compute(A, m);
        const bool here = ci + 2 < cin;
        load_ops(A, here ? cur : nxt, here ? ci + 2 : 0, here ? qt + ci + 2 : qn);
        __builtin_amdgcn_sched_barrier(0);
        compute(B, m);
      }
      if (ci < cin) {  // odd channel count: the last channel of the term
        load_ops(B, nxt, 0, qn);
        __builtin_amdgcn_sched_barrier(0);
        compute(A, m);
        A = B;
      }
    }
    tg = ng;
    cur = nxt;
  }
  float omax = 0.f;
  if (ok) {
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
      const int o = 4 * mt + gq;
      if (o >= a.cout) continue;
      cf out[12];
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const float a = p1[c][mt][j], b = p2[c][mt][j];
          out[j * 3 + c] = mk(a - b, p3[c][mt][j] - a - b);
          omax = fmaxf(omax, fmaxf(fabsf(out[j * 3 + c].re), fabsf(out[j * 3 + c].im)));
        }
      SpinorIO<P32>::store(chan_field(a.out, o, p, g.vh), i, out);
    }
  }
  // the next layer's split-FP16 operands are scaled by this field's largest component
  if (a.out_max) record_max(a.out_max, omax);
}

// CI: input channel count known at compile time (0: a.cin at run time)
template <int MT, int CI>
__global__ void __launch_bounds__(256) k_net_layer_mfma(NetLayerArgs a, const float *Apk) {
  net_layer_mfma_body<MT, CI>(a, Apk);
}

// ---- single-output layer in two passes --------------------------------------
// With one output channel the matrix-core layer fills a quarter of its tiles
// and still transports every input channel.  Mixing and transport commute, so
// the layer is split:
//   pass 1 (k_net_mix, matrix cores) mixes the input channels once per term at
//     every input site, Y_t(j) = sum_i W[t][0][i] in_i(j): per colour a complex
//     GEMM with m = (term, spin), k = (spin s2) per input channel, n = site, as
//     three real products; two M tiles hold the 8 hop terms, a third (rows 0-3
//     used) the site term;
//   pass 2 (k_net_gather, vector units) transports and sums them at the output
//     sites, out(x) = sum_{t<8} M_t(x) Y_t(neighbour_t(x)) + Y_8(x),
// one transport per term instead of one per term and input channel.  The term
// fields Y_t use the channel layout: term t, parity q at slot (t, q).
struct NetMixArgs {
  SpinorPtr in;     // cin channel fields
  SpinorPtr mix;    // 9 term fields
  const float2 *W;  // the layer's weights W[t][0][i][4x4]
  int vh, cin;
  int out_par;      // the output parity of a one-parity layer (-1: both parities)
  const SysState *st;
};

// tiles T0 .. T0 + NT - 1 of input parity q for the 16 sites from base
template <int CI, int T0, int NT>
QCU_D void net_mix_body(const NetMixArgs &a, int q, int base) {
  const int lane = threadIdx.x & 63, gq = lane >> 4, ml = lane & 15;
  const int vh = a.vh;
  int i = base + ml;
  const bool ok = i < vh;
  if (!ok) i = base;
  // the lane's spin gq of site i: floats 6 gq .. 6 gq + 5 of the 24
  const unsigned vh16 = (unsigned)vh * 16u;
  unsigned off[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const int f = 6 * gq + 2 * k;
    off[k] = (unsigned)(f >> 2) * vh16 + (unsigned)(f & 3) * 4u + (unsigned)i * 16u;
  }
  f4v p1[3][NT], p2[3][NT], p3[3][NT];
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int e = 0; e < NT; ++e) {
      p1[c][e] = (f4v){0.f, 0.f, 0.f, 0.f};
      p2[c][e] = (f4v){0.f, 0.f, 0.f, 0.f};
      p3[c][e] = (f4v){0.f, 0.f, 0.f, 0.f};
    }
  const int cin = CI > 0 ? CI : a.cin;
  const size_t field = (size_t)vh * 96;
  const char *fb = (const char *)a.in.p + (size_t)q * field;
  // A[m = ml][k = gq] of tile et: term 4 et + ml / 4 (tile 2: the site term in
  // rows 0-3, zero below), spin ml % 4, input spin gq
  auto step = [&](int ci) {
    const char *cb = fb + (size_t)ci * 2 * field;
    float2 s[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) s[k] = ld_at<float2>(cb, 0, off[k]);
    float w[NT][3];
#pragma unroll
    for (int e = 0; e < NT; ++e) {
      const int et = T0 + e;
      const int t = et < 2 ? 4 * et + (ml >> 2) : 8;
      float2 wv = a.W[((size_t)t * cin + ci) * 16 + (ml & 3) * 4 + gq];
      if (et == 2 && ml >= 4) wv = make_float2(0.f, 0.f);
      w[e][0] = wv.x;
      w[e][1] = wv.y;
      w[e][2] = wv.x + wv.y;
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const float br = s[c].x, bi = s[c].y, bs = br + bi;
#pragma unroll
      for (int e = 0; e < NT; ++e) {
        p1[c][e] = __builtin_amdgcn_mfma_f32_16x16x4f32(w[e][0], br, p1[c][e], 0, 0, 0);
        p2[c][e] = __builtin_amdgcn_mfma_f32_16x16x4f32(w[e][1], bi, p2[c][e], 0, 0, 0);
        p3[c][e] = __builtin_amdgcn_mfma_f32_16x16x4f32(w[e][2], bs, p3[c][e], 0, 0, 0);
      }
    }
  };
  if constexpr (CI > 0) {
#pragma unroll
    for (int ci = 0; ci < CI; ++ci) step(ci);
  } else {
#pragma unroll 1
    for (int ci = 0; ci < cin; ++ci) step(ci);
  }
  if (!ok) return;
  // lane group gq holds rows 4 gq .. 4 gq + 3: spins 0..3 of term 4 et + gq
#pragma unroll
  for (int e = 0; e < NT; ++e) {
    const int et = T0 + e;
    if (et == 2 && gq != 0) continue;
    const int t = et < 2 ? 4 * et + gq : 8;
    cf out[12];
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const float x = p1[c][e][j], y = p2[c][e][j];
        out[j * 3 + c] = mk(x - y, p3[c][e][j] - x - y);
      }
    SpinorIO<P32>::store(chan_field(a.mix, t, q, vh), i, out);
  }
}

// grid: x = 16-site chunks (one per wave), y = input parity q.  BOTH: every
// term on both parities (both output parities); otherwise the hop terms on
// the parity the output parity's hops read and the site term on the output
// parity.
template <int CI, bool BOTH>
__global__ void __launch_bounds__(256) k_net_mix(NetMixArgs a) {
  if (a.st && a.st->done) return;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int base = ((int)blockIdx.x * (int)(blockDim.x >> 6) + wave) * 16;
  if (base >= a.vh) return;  // uniform over the wave
  const int q = blockIdx.y;
  if constexpr (BOTH) {
    net_mix_body<CI, 0, 3>(a, q, base);
  } else {
    if (q == a.out_par) net_mix_body<CI, 2, 1>(a, q, base);
    else net_mix_body<CI, 0, 2>(a, q, base);
  }
}

struct NetGatherArgs {
  SpinorPtr mix;  // the 9 term fields of k_net_mix
  SpinorPtr mix2; // TWO: a second set of term fields, added to the first; its site term replaces the first's
  SpinorPtr out;  // the output channel (FP32 channel layout), or with DIRECT an odd vector of precision P
  LinkPtr U;      // FP32 links
  Geom g;
  int parity0;    // first output parity (the grid's y index adds to it)
  const SysState *st;
};

// one thread per output site: the site term plus the 8 transported hop terms.
// DIRECT: the odd output goes straight into a vector of precision P (the
// solver's), instead of channel 0 of the FP32 output block.  TWO: the hop
// terms are the sums of two sets of term fields, the site term is the
// second set's.  Every term field and, with one output parity, every link is
// read once: non-temporal loads.
template <int P = P32, bool DIRECT = false, bool TWO = false>
__global__ void __launch_bounds__(256) k_net_gather(NetGatherArgs a) {
  if (a.st && a.st->done) return;
  const Geom g = a.g;
  const int i = remap_block(blockIdx.x, gridDim.x, 1) * blockDim.x + threadIdx.x;
  const int p = a.parity0 + blockIdx.y;
  if (i >= g.vh) return;
  const Site site = decode(g, i, p);
  auto load = [&](const SpinorPtr &f, int j, cf(&v)[12]) {
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      const float4 q = ld_nt(f.p, (size_t)k * f.stride * 16, (unsigned)j * 16u);
      v[2 * k] = mk(q.x, q.y);
      v[2 * k + 1] = mk(q.z, q.w);
    }
  };
  cf acc[12];
  load(chan_field(TWO ? a.mix2 : a.mix, 8, p, g.vh), i, acc);
#pragma unroll 2
  for (int t = 0; t < 8; ++t) {
    const TermGeo tg = term_geo(g, site, i, p, t);
    cf m[9];
    if (gridDim.y == 1) term_link_matrix<true>(a.U, tg, m);
    else term_link_matrix(a.U, tg, m);
    cf v[12];
    load(chan_field(a.mix, t, tg.jp, g.vh), tg.j, v);
    if constexpr (TWO) {
      cf w[12];
      load(chan_field(a.mix2, t, tg.jp, g.vh), tg.j, w);
#pragma unroll
      for (int k = 0; k < 12; ++k) v[k] = v[k] + w[k];
    }
    transport_spins(m, v);
#pragma unroll
    for (int k = 0; k < 12; ++k) acc[k] = acc[k] + v[k];
  }
  if constexpr (DIRECT) {
    typedef typename real_of<P>::type T;
    cplx<T> v[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) v[k] = mk<T>((T)acc[k].re, (T)acc[k].im);
    SpinorIO<P>::store(a.out, i, v);
  } else {
    SpinorIO<P32>::store(chan_field(a.out, 0, p, g.vh), i, acc);
  }
}

// weight gradient partials: for term t, input channel ci, outputs og..og+CO-1,
//   sum over sites of both parities and colours of g_o(x)[s][c] conj((T_t a_ci)(x)[s2][c])
struct NetWgradArgs {
  SpinorPtr act;   // layer input channels
  SpinorPtr gout;  // gradient w.r.t. the layer output channels
  LinkPtr U;
  Geom g;
  int cin, cout;
  float *partials;  // [combo][block][CO * 32]
  int act_zero;     // parity on which the input is known to be zero (-1: none)
  int gout_zero;    // parity on which the output gradient is known to be zero (-1: none)
  // largest |component| of the input and of the output gradient (float bits),
  // recorded by the kernels that wrote them: the split-FP16 products scale
  // their operands by the powers of two these give
  const unsigned *amax = nullptr, *gmax = nullptr;
};

template <int CO>
__global__ void __launch_bounds__(256) k_net_wgrad(NetWgradArgs a) {
  const Geom g = a.g;
  const int t = blockIdx.y / a.cin, ci = blockIdx.y % a.cin;
  const int og = blockIdx.z * CO;
  float acc[CO][32];
#pragma unroll
  for (int o = 0; o < CO; ++o)
#pragma unroll
    for (int k = 0; k < 32; ++k) acc[o][k] = 0.f;
  const int nsite = 2 * g.vh;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nsite; k += gridDim.x * blockDim.x) {
    const int p = k >= g.vh ? 1 : 0;
    const int i = k - p * g.vh;
    const TermGeo tg = term_geo(g, decode(g, i, p), i, p, t);
    cf v[12];
    SpinorIO<P32>::load(chan_field(a.act, ci, tg.jp, g.vh), tg.j, v);
    if (tg.hop) {
      cf m[9];
      term_link_matrix(a.U, tg, m);
      transport_spins(m, v);
    }
#pragma unroll
    for (int o = 0; o < CO; ++o) {
      if (og + o >= a.cout) break;
      cf gg[12];
      SpinorIO<P32>::load(chan_field(a.gout, og + o, p, g.vh), i, gg);
#pragma unroll
      for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int s2 = 0; s2 < 4; ++s2) {
          cf e = mk(0.f, 0.f);
#pragma unroll
          for (int c = 0; c < 3; ++c) cmadj(e, v[s2 * 3 + c], gg[s * 3 + c]);  // g * conj(v)
          acc[o][2 * (s * 4 + s2)] += e.re;
          acc[o][2 * (s * 4 + s2) + 1] += e.im;
        }
    }
  }
  // block reduction in a fixed order: waves by shuffles, then LDS across waves
  __shared__ float sh[4][CO * 32];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int o = 0; o < CO; ++o)
#pragma unroll
    for (int k = 0; k < 32; ++k) {
      float v = acc[o][k];
#pragma unroll
      for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
      if (lane == 0) sh[wave][o * 32 + k] = v;
    }
  __syncthreads();
  const int combo = blockIdx.z * gridDim.y + blockIdx.y;
  float *dst = a.partials + ((size_t)combo * gridDim.x + blockIdx.x) * (CO * 32);
  for (int k = threadIdx.x; k < CO * 32; k += blockDim.x) {
    float s = 0.f;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += sh[w][k];
    dst[k] = s;
  }
}

// Weight gradient on the matrix cores.  For term t,
//   gW[t][o][i][s][s2] = sum over sites x of both parities and colours c of
//                        g_o(x)[s][c] conj(v_i(x)[s2][c]),  v_i = T_t a_i,
// i.e. D[m = (o, s)][n = (i, s2)] = sum_k G[m][k] conj(V)[k][n] over k = (x, c).
// Two forms of the products:
// * FP32 (SPLIT false): three real products per colour on
//   v_mfma_f32_16x16x4_f32 (the three-multiplication complex product, with
//   G conj(V)):
//     P1 += Gr Vr,  P2 += Gi Vi,  P3 += (Gr + Gi)(Vr - Vi),
//     Re D = P1 + P2,  Im D = P3 - P1 + P2
//   (accumulators dr = P1, dn = P2, di = P3).  A K step is 4 sites of one
//   colour: lane l (group q = l / 16, index n = l % 16) feeds A[m = n][k = q] =
//   g of row (o = 4 mt + n / 4, s = n % 4) at site q of the group, and
//   B[k = q][n] = v of column (i = 4 nt + n / 4, s2 = n % 4) at the same site.
// * split FP16 (SPLIT true; the 8 -> 8 layers, whose two operand fields have
//   recorded maxima): v_mfma_f32_16x16x32_f16 with K step = the 3 colours of 4
//   sites as real pairs, lane group q supplying k = 8 q .. 8 q + 7 = (Re, Im)
//   of colours 0, 1, 2 of site q and two zeros.  With A = (Gr, Gi) pairs,
//     Re D = A (Vr, Vi),  Im D = A (-Vi, Vr)
//   (G conj V = Gr Vr + Gi Vi + i (Gi Vr - Gr Vi)), each product of FP32
//   operands split as x = hi + lo in three FP16 products (Ootomo & Yokota,
//   IJHPCA 36 (2022); lo x lo dropped), after scaling G and V by powers of two
//   from their fields' largest components (G below 2^14; V, a transported
//   colour vector, below sqrt(3) 2^14).  Accumulators dr = Re D, di = Im D in
//   the scaled units; k_net_wgrad_fin divides the scales out in FP64.
// Either way lane (q, n) ends holding D rows (o = 4 mt + q, s = j), column n.
// Each wave sweeps groups of 4 sites; the 4 waves of a block are summed
// through LDS and the block sums are added in a fixed order by
// k_net_wgrad_fin.
//
// The grid is one row of 9 x kWgBlocks blocks.  The 9 blocks (one per term)
// of a site block sb are 8 block indices apart inside a window of 72: blocks
// go to the XCDs in turn, so all 9 run at about the same time on one XCD and
// read the output gradient and the input field of their sites from its L2,
// instead of from HBM once per term.  XCD k takes the site blocks
// k nsb / 8 .. (k + 1) nsb / 8 - 1, so at each step of the waves' sweep it
// works on one contiguous stretch of sites (their y neighbours included).
constexpr int kWgBlocks = 1024;  // blocks per term (a multiple of 8)

template <int MT, int NT, bool SPLIT, bool LDS>
__global__ void __launch_bounds__(256) k_net_wgrad_mfma(NetWgradArgs a, float *part) {
  const Geom g = a.g;
  const int b = blockIdx.x;
  const int nsb = gridDim.x / 9;
  const int t = (b >> 3) % 9, sb = (b & 7) * (nsb >> 3) + b / 72;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int q = lane >> 4, n = lane & 15;
  // output parities this term needs: a nonzero output gradient there, and an
  // input (parity 1 - p for a hop, p for the site term) not known to be zero;
  // the sites of the others contribute exact zeros and are skipped
  int pmask = 3;
#pragma unroll
  for (int pp = 0; pp < 2; ++pp) {
    const int pin = t < 8 ? 1 - pp : pp;
    if (pp == a.gout_zero || pin == a.act_zero) pmask &= ~(1 << pp);
  }
  const int site0 = (pmask & 1) ? 0 : g.vh;
  const int nsite = ((pmask & 1) + (pmask >> 1)) * g.vh, ngroups = (nsite + 3) / 4;
  const int nw = nsb * 4;
  const size_t vh = (size_t)g.vh;
  const int rs = n & 3, rslot = (3 * rs) >> 1, rodd = (3 * rs) & 1;  // spin of this lane's row / column
  f4v dr[MT][NT], di[MT][NT], dn[SPLIT ? 1 : MT][SPLIT ? 1 : NT];
#pragma unroll
  for (int mt = 0; mt < MT; ++mt)
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
      dr[mt][nt] = (f4v){0.f, 0.f, 0.f, 0.f};
      di[mt][nt] = (f4v){0.f, 0.f, 0.f, 0.f};
      if (!SPLIT) dn[mt][nt] = (f4v){0.f, 0.f, 0.f, 0.f};
    }
  // operand scales of the split products (powers of two, exact)
  const float sg = SPLIT ? pow2_scale(__uint_as_float(*a.gmax), 14) : 1.f;
  const float sv = SPLIT ? pow2_scale(__uint_as_float(*a.amax), 14) : 1.f;
  // the lane's spin (3 colours) out of its two raw float4 slots
  auto spin3 = [&](const float4 (&r)[2], float keep, cf (&v)[3]) {
    const float4 x = r[0], y = r[1];
    if (rodd) {
      v[0] = mk(keep * x.z, keep * x.w);
      v[1] = mk(keep * y.x, keep * y.y);
      v[2] = mk(keep * y.z, keep * y.w);
    } else {
      v[0] = mk(keep * x.x, keep * x.y);
      v[1] = mk(keep * x.z, keep * x.w);
      v[2] = mk(keep * y.x, keep * y.y);
    }
  };
  // one group of 4 sites: geometry and raw operands (loaded one group ahead)
  struct Geo {
    int p, i;
    float keep;
    TermGeo tg;
  };
  struct Ops {
    float4 g[MT][2], v[NT][2], l[3];
  };
  // lanes whose row or column channel does not exist read channel 0 and are
  // scaled by zero, so every lane issues the same loads
  const int o_n = n >> 2;
  float grow[MT], gcol[NT];
#pragma unroll
  for (int mt = 0; mt < MT; ++mt) grow[mt] = 4 * mt + o_n < a.cout ? sg : 0.f;
#pragma unroll
  for (int nt = 0; nt < NT; ++nt) gcol[nt] = 4 * nt + o_n < a.cin ? sv : 0.f;
  const TermConst tcon = term_const(g, t);
  auto geo = [&](int grp) {
    Geo r;
    int k = grp * 4 + q;
    const bool ok = k < nsite;
    if (!ok) k = 0;  // a valid site; its contribution is zeroed
    k += site0;
    r.p = k >= g.vh ? 1 : 0;
    r.i = k - r.p * g.vh;
    r.keep = ok ? 1.f : 0.f;
    r.tg = term_geo_u(g, decode(g, r.i, r.p), r.i, r.p, tcon);
    return r;
  };
  // every term loads a link (the site term a valid one it does not apply), so
  // that the loads of the sweep below sit behind no branch
  auto load = [&](Ops &x, const Geo &ge) {
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
      const int o = grow[mt] != 0.f ? 4 * mt + o_n : 0;
      const float4 *b = (const float4 *)chan_field(a.gout, o, ge.p, g.vh).p + ge.i + (size_t)rslot * vh;
      x.g[mt][0] = b[0];
      x.g[mt][1] = b[vh];
    }
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
      const int ch = gcol[nt] != 0.f ? 4 * nt + o_n : 0;
      const float4 *b = (const float4 *)chan_field(a.act, ch, ge.tg.jp, g.vh).p + ge.tg.j + (size_t)rslot * vh;
      x.v[nt][0] = b[0];
      x.v[nt][1] = b[vh];
    }
#pragma unroll
    for (int k = 0; k < 3; ++k)
      x.l[k] = ld_at<float4>(a.U.a, ((size_t)ge.tg.slab * 3 + k) * a.U.vh * 16, (unsigned)ge.tg.li * 16u);
  };
  // the products of one group: rows x.g, columns x.v transported by u (a
  // hop term; the site term leaves them as they are), sites weighted by keep
  auto products = [&](const Ops &x, float keep, const cf (&u)[9]) {
    const bool hop = tcon.hop;  // uniform
    // rows: output gradient of channels 4 mt + n / 4, spin n % 4
    cf gv[MT][3];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) spin3(x.g[mt], keep * grow[mt], gv[mt]);
    // columns: transported inputs of channels 4 nt + n / 4, spin n % 4
    cf vv[NT][3];
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
      spin3(x.v[nt], gcol[nt], vv[nt]);
      if (hop) transport1(u, vv[nt]);
    }
    if (SPLIT) {
      uint4 ah[MT], al[MT];
#pragma unroll
      for (int mt = 0; mt < MT; ++mt) {
        split2(gv[mt][0].re, gv[mt][0].im, ah[mt].x, al[mt].x);
        split2(gv[mt][1].re, gv[mt][1].im, ah[mt].y, al[mt].y);
        split2(gv[mt][2].re, gv[mt][2].im, ah[mt].z, al[mt].z);
        ah[mt].w = 0u;
        al[mt].w = 0u;
      }
      // (Re, Im) -> (-Im, Re) on a packed FP16 pair
      auto rot = [](unsigned u) { return __builtin_amdgcn_alignbit(u, u, 16) ^ 0x8000u; };
#pragma unroll
      for (int nt = 0; nt < NT; ++nt) {
        uint4 bh, bl;
        split2(vv[nt][0].re, vv[nt][0].im, bh.x, bl.x);
        split2(vv[nt][1].re, vv[nt][1].im, bh.y, bl.y);
        split2(vv[nt][2].re, vv[nt][2].im, bh.z, bl.z);
        bh.w = 0u;
        bl.w = 0u;
        const uint4 ch = make_uint4(rot(bh.x), rot(bh.y), rot(bh.z), 0u);
        const uint4 cl = make_uint4(rot(bl.x), rot(bl.y), rot(bl.z), 0u);
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
          f4v r = dr[mt][nt], i = di[mt][nt];
          r = __builtin_amdgcn_mfma_f32_16x16x32_f16(as_h8(ah[mt]), as_h8(bh), r, 0, 0, 0);
          i = __builtin_amdgcn_mfma_f32_16x16x32_f16(as_h8(ah[mt]), as_h8(ch), i, 0, 0, 0);
          r = __builtin_amdgcn_mfma_f32_16x16x32_f16(as_h8(ah[mt]), as_h8(bl), r, 0, 0, 0);
          i = __builtin_amdgcn_mfma_f32_16x16x32_f16(as_h8(ah[mt]), as_h8(cl), i, 0, 0, 0);
          r = __builtin_amdgcn_mfma_f32_16x16x32_f16(as_h8(al[mt]), as_h8(bh), r, 0, 0, 0);
          i = __builtin_amdgcn_mfma_f32_16x16x32_f16(as_h8(al[mt]), as_h8(ch), i, 0, 0, 0);
          dr[mt][nt] = r;
          di[mt][nt] = i;
        }
      }
      return;
    }
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
          const float ar = gv[mt][c].re, ai = gv[mt][c].im, br = vv[nt][c].re, bi = vv[nt][c].im;
          dr[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar, br, dr[mt][nt], 0, 0, 0);
          dn[SPLIT ? 0 : mt][SPLIT ? 0 : nt] =
              __builtin_amdgcn_mfma_f32_16x16x4f32(ai, bi, dn[SPLIT ? 0 : mt][SPLIT ? 0 : nt], 0, 0, 0);
          di[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar + ai, br - bi, di[mt][nt], 0, 0, 0);
        }
  };
  auto compute = [&](const Ops &x, const Geo &ge) {
    cf u[9];
    if (tcon.hop) term_matrix(x.l, 1, ge.tg.fwd, ge.tg.ph, u);
    products(x, ge.keep, u);
  };
  // shared memory: the staged operands of a step (LDS path), then the block
  // sum of the epilogue (after a barrier)
  constexpr int NREG = MT * NT * 2 * 4;
  constexpr int kStage = LDS ? 2 * 48 * 32 + 32 * 9 / 2 : 0;  // float4s: sG, sV, sM
  constexpr int kSum = 4 * NREG * 64 / 4;
  __shared__ float4 smem[kStage > kSum ? kStage : kSum];
  if constexpr (LDS) {
    // The block's 4 waves take 8 consecutive groups, 32 consecutive sites, per
    // step: wave w the groups w and w + 4 of the step.  Their output gradient
    // and input channels are loaded by the whole block with whole lines:
    // thread (plane, site) reads float4 slot plane % 6 of channel plane / 6 at
    // one of the 32 sites, so a wave's load touches 2 planes x 512 contiguous
    // bytes.  They are staged in LDS plane by plane ([plane][site]), and each
    // wave reads its rows and columns from there.  (A wave that read its
    // matrix-core operands straight from memory would touch 16 planes x 64
    // bytes per load, 16 MiB apart, all in one vector-cache set.)  The next
    // step's loads are issued before the current step's products.
    float4 *sG = smem, *sV = smem + 48 * 32;
    float2 *sM = (float2 *)(smem + 2 * 48 * 32);  // the term's colour matrix at each site of the step
    const int tid = threadIdx.x;
    const int fs = tid & 31;  // the thread's site in a step; its planes: (tid >> 5) + 8 r
    // the thread's six planes (separate variables: arrays of them end up in
    // scratch memory in some instances)
    auto plane_base = [&](int r, const SpinorPtr &f, int nch) {
      const int plane = (tid >> 5) + 8 * r, ch = plane / 6, slot = plane - 6 * ch;
      return (const float4 *)f.p + ((size_t)(ch < nch ? ch : 0) * 12 + slot) * vh;
    };
    const float4 *gb0 = plane_base(0, a.gout, a.cout), *gb1 = plane_base(1, a.gout, a.cout),
                 *gb2 = plane_base(2, a.gout, a.cout), *gb3 = plane_base(3, a.gout, a.cout),
                 *gb4 = plane_base(4, a.gout, a.cout), *gb5 = plane_base(5, a.gout, a.cout);
    const float4 *vb0 = plane_base(0, a.act, a.cin), *vb1 = plane_base(1, a.act, a.cin),
                 *vb2 = plane_base(2, a.act, a.cin), *vb3 = plane_base(3, a.act, a.cin),
                 *vb4 = plane_base(4, a.act, a.cin), *vb5 = plane_base(5, a.act, a.cin);
    // threads 0..31 (one per site of the step) also load the site's link and
    // build the term's colour matrix there, once per site
    const bool linker = tid < 32;
    const int nw8 = nsb * 8;  // groups between a block's steps
    // the step's loads go straight into LDS (global_load_lds_dwordx4: lane l's
    // 16 bytes land at the wave's LDS base + 16 l, so a wave's load fills
    // planes 2 w + 8 r and 2 w + 8 r + 1 of the step); the links into registers
    auto issue = [&](int g0, float4 (&l)[3], float &ph) {
      int k = g0 * 4 + fs;
      if (k >= nsite) k = 0;
      k += site0;
      const int p = k >= g.vh ? 1 : 0, i = k - p * g.vh;
      const TermGeo tg = term_geo_u(g, decode(g, i, p), i, p, tcon);
      const unsigned og = (unsigned)p * 6u * (unsigned)g.vh + (unsigned)i;
      const unsigned ov = (unsigned)tg.jp * 6u * (unsigned)g.vh + (unsigned)tg.j;
      typedef void __attribute__((address_space(3))) *lds_t;
      const int wb = 2 * wave * 32;
#define QCU_WG_LDS(R_)                                                                                   \
  __builtin_amdgcn_global_load_lds((const void *)(gb##R_ + og), (lds_t)(sG + wb + 8 * (R_) * 32), 16, 0, 0); \
  __builtin_amdgcn_global_load_lds((const void *)(vb##R_ + ov), (lds_t)(sV + wb + 8 * (R_) * 32), 16, 0, 0);
      QCU_WG_LDS(0) QCU_WG_LDS(1) QCU_WG_LDS(2) QCU_WG_LDS(3) QCU_WG_LDS(4) QCU_WG_LDS(5)
#undef QCU_WG_LDS
      if (linker && tcon.hop) {
        const size_t ps = (size_t)a.U.vh * 16;
        l[0] = ld_at<float4>(a.U.a, ((size_t)tg.slab * 3 + 0) * ps, (unsigned)tg.li * 16u);
        l[1] = ld_at<float4>(a.U.a, ((size_t)tg.slab * 3 + 1) * ps, (unsigned)tg.li * 16u);
        l[2] = ld_at<float4>(a.U.a, ((size_t)tg.slab * 3 + 2) * ps, (unsigned)tg.li * 16u);
        ph = tg.ph;
      }
    };
    // the products of the wave's group h (0, 1) of the step
    auto group = [&](int g0, int h) {
      const int s = 16 * h + 4 * wave + q;
      Ops x;
#pragma unroll
      for (int mt = 0; mt < MT; ++mt) {
        const int o = grow[mt] != 0.f ? 4 * mt + o_n : 0;
        x.g[mt][0] = sG[(o * 6 + rslot) * 32 + s];
        x.g[mt][1] = sG[(o * 6 + rslot + 1) * 32 + s];
      }
#pragma unroll
      for (int nt = 0; nt < NT; ++nt) {
        const int ch = gcol[nt] != 0.f ? 4 * nt + o_n : 0;
        x.v[nt][0] = sV[(ch * 6 + rslot) * 32 + s];
        x.v[nt][1] = sV[(ch * 6 + rslot + 1) * 32 + s];
      }
      cf u[9];
      if (tcon.hop) {
#pragma unroll
        for (int k = 0; k < 9; ++k) {
          const float2 e = sM[s * 9 + k];
          u[k] = mk(e.x, e.y);
        }
      }
      products(x, (g0 + 4 * h + wave) * 4 + q < nsite ? 1.f : 0.f, u);
    };
    int g0 = sb * 8;  // the block's first group; block-uniform, as every branch below
#pragma unroll 1
    for (; g0 < ngroups; g0 += nw8) {
      __syncthreads();  // the previous step's LDS reads are done
      float4 raw[3];
      float ph = 1.f;
      issue(g0, raw, ph);
      __builtin_amdgcn_s_waitcnt(0);  // this wave's loads, those into LDS included, have landed
      if (linker && tcon.hop) {
        cf m[9];
        term_matrix(raw, 1, tcon.fwd, ph, m);
#pragma unroll
        for (int k = 0; k < 9; ++k) sM[tid * 9 + k] = make_float2(m[k].re, m[k].im);
      }
      __syncthreads();
      group(g0, 0);
      group(g0, 1);
    }
    __syncthreads();  // the block sum below reuses the staging memory
  } else {
  // groups grp, grp + nw, ...: the operands of the next group load into the
  // other operand set while this one multiplies (the last group loads a valid
  // dummy: itself), so no load sits behind a branch
  int grp = sb * 4 + wave;
  if (grp < ngroups) {
    Geo gA = geo(grp), gB;
    Ops A, B;
    load(A, gA);
#pragma unroll 1
    while (true) {
      const bool m1 = grp + nw < ngroups;
      gB = geo(m1 ? grp + nw : grp);
      load(B, gB);
      __builtin_amdgcn_sched_barrier(0);
      compute(A, gA);
      if (!m1) break;
      grp += nw;
      const bool m2 = grp + nw < ngroups;
      gA = geo(m2 ? grp + nw : grp);
      load(A, gA);
      __builtin_amdgcn_sched_barrier(0);
      compute(B, gB);
      if (!m2) break;
      grp += nw;
    }
  }
  }
  // block sum in a fixed order: registers (mt, nt, part, j) of every lane
  float(*sh)[NREG][64] = (float(*)[NREG][64])smem;
#pragma unroll
  for (int mt = 0; mt < MT; ++mt)
#pragma unroll
    for (int nt = 0; nt < NT; ++nt)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int r = ((mt * NT + nt) * 2) * 4 + j;
        if (SPLIT) {
          sh[wave][r][lane] = dr[mt][nt][j];
          sh[wave][r + 4][lane] = di[mt][nt][j];
        } else {
          const float p2 = dn[SPLIT ? 0 : mt][SPLIT ? 0 : nt][j];
          sh[wave][r][lane] = dr[mt][nt][j] + p2;
          sh[wave][r + 4][lane] = di[mt][nt][j] - dr[mt][nt][j] + p2;
        }
      }
  __syncthreads();
  for (int e = threadIdx.x; e < NREG * 64; e += blockDim.x) {
    const int r = e >> 6, l = e & 63;
    const float v = (sh[0][r][l] + sh[1][r][l]) + (sh[2][r][l] + sh[3][r][l]);
    part[((size_t)(t * nsb + sb) * NREG + r) * 64 + l] = v;
  }
}

// gW[t][o][i][s][s2] = sum over the blocks of term t, FP64 accumulation in block
// order, times the inverse operand scales of the split products (null maxima:
// unscaled FP32 products)
template <int MT, int NT>
__global__ void k_net_wgrad_fin(const float *part, int nblk, int cin, int cout, float2 *gW, const unsigned *gmax,
                                const unsigned *amax) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;  // over 9 * cout * cin * 16
  const int total = 9 * cout * cin * 16;
  if (idx >= total) return;
  const int e = idx % 16, s = e >> 2, s2 = e & 3;
  const int ci = (idx / 16) % cin;
  const int o = (idx / 16 / cin) % cout;
  const int t = idx / 16 / cin / cout;
  constexpr int NREG = MT * NT * 2 * 4;
  const int mt = o >> 2, nt = ci >> 2;
  const int lane = (o & 3) * 16 + (ci & 3) * 4 + s2;
  const int r = ((mt * NT + nt) * 2) * 4 + s;
  // 8 interleaved FP64 sums (block b into sum b % 8) keep 16 loads in flight;
  // they are combined in a fixed order, so the result is the same every run
  constexpr int kAcc = 8;
  double re[kAcc] = {}, im[kAcc] = {};
  const float *src = part + ((size_t)t * nblk * NREG + r) * 64 + lane;
  const size_t bstride = (size_t)NREG * 64;
  int b = 0;
  for (; b + kAcc <= nblk; b += kAcc)
#pragma unroll
    for (int k = 0; k < kAcc; ++k) {
      re[k] += src[(b + k) * bstride];
      im[k] += src[(b + k) * bstride + 4 * 64];
    }
  for (; b < nblk; ++b) {
    re[b % kAcc] += src[b * bstride];
    im[b % kAcc] += src[b * bstride + 4 * 64];
  }
  double sr = 0, si = 0;
#pragma unroll
  for (int k = 0; k < kAcc; ++k) {
    sr += re[k];
    si += im[k];
  }
  if (gmax) {
    const double inv = 1.0 / ((double)pow2_scale(__uint_as_float(*gmax), 14) *
                              (double)pow2_scale(__uint_as_float(*amax), 14));
    sr *= inv;
    si *= inv;
  }
  gW[idx] = make_float2((float)sr, (float)si);
}

// second stage: gW[t][o][ci][16] = sum over blocks (FP64 accumulation)
__global__ void k_net_wgrad_sum(const float *partials, int nblk, int cin, int cout, int co, int ngroups,
                                float2 *gW, int accumulate) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;  // over 9 * cout * cin * 16
  const int total = 9 * cout * cin * 16;
  if (idx >= total) return;
  const int e = idx % 16;
  const int ci = (idx / 16) % cin;
  const int o = (idx / 16 / cin) % cout;
  const int t = idx / 16 / cin / cout;
  const int gz = o / co, ol = o % co;
  const int combo = gz * (9 * cin) + t * cin + ci;
  double re = 0, im = 0;
  for (int b = 0; b < nblk; ++b) {
    const float *src = partials + ((size_t)combo * nblk + b) * (co * 32) + ol * 32 + 2 * e;
    re += src[0];
    im += src[1];
  }
  float2 v = make_float2((float)re, (float)im);
  if (accumulate) {
    v.x += gW[idx].x;
    v.y += gW[idx].y;
  }
  gW[idx] = v;
}

// W'[t'][i][o] = (W[t][o][i])^dag with t' the adjoint term (hops swap direction)
__global__ void k_net_transpose(const float2 *W, float2 *Wt, int cin, int cout) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  const int total = 9 * cout * cin * 16;
  if (idx >= total) return;
  const int e = idx % 16, s = e / 4, s2 = e % 4;
  const int ci = (idx / 16) % cin;
  const int o = (idx / 16 / cin) % cout;
  const int t = idx / 16 / cin / cout;
  const int ta = t == 8 ? 8 : (t ^ 1);
  const float2 w = W[idx];
  // Wt has cin' = cout, cout' = cin: index [ta][ci][o][s2][s]
  Wt[(((size_t)ta * cin + ci) * cout + o) * 16 + s2 * 4 + s] = make_float2(w.x, -w.y);
}

// embed an odd-site vector (precision P) as channel 0 of a full FP32 field
// (rec: receives the largest |component|, or null; then 4 sites per thread
// and one atomic per block of 256 threads); zero_even: the even sites are
// written as zeros (else left as they are, for passes that never read them)
template <int P>
__global__ void __launch_bounds__(256) k_net_embed(SpinorPtr full, SpinorPtr odd, int vh, const SysState *st,
                                                  unsigned *rec = nullptr, int zero_even = 1) {
  if (st && st->done) return;
  typedef typename real_of<P>::type T;
  const int per = rec ? 4 : 1;
  float mx = 0.f;
  for (int k = 0; k < per; ++k) {
    const int i = (blockIdx.x * per + k) * blockDim.x + threadIdx.x;
    if (i >= vh) break;
    cplx<T> v[12];
    SpinorIO<P>::load(odd, i, v);
    cf w[12], z[12];
#pragma unroll
    for (int c = 0; c < 12; ++c) {
      w[c] = mk((float)v[c].re, (float)v[c].im);
      z[c] = mk(0.f, 0.f);
      mx = fmaxf(mx, fmaxf(fabsf(w[c].re), fabsf(w[c].im)));
    }
    SpinorIO<P32>::store(chan_field(full, 0, 1, vh), i, w);
    if (zero_even) SpinorIO<P32>::store(chan_field(full, 0, 0, vh), i, z);
  }
  if (!rec) return;  // uniform
  __shared__ float red[256];
  red[threadIdx.x] = mx;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + h]);
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicMax(rec, __float_as_uint(red[0]));
}

template <int P>
__global__ void k_net_extract(SpinorPtr odd, SpinorPtr full, int vh, const SysState *st) {
  if (st && st->done) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= vh) return;
  typedef typename real_of<P>::type T;
  cf w[12];
  SpinorIO<P32>::load(chan_field(full, 0, 1, vh), i, w);
  cplx<T> v[12];
#pragma unroll
  for (int k = 0; k < 12; ++k) v[k] = mk<T>((T)w[k].re, (T)w[k].im);
  SpinorIO<P>::store(odd, i, v);
}

// z = gamma5 Ao^-1 r (FP32, odd sites); gamma5 = diag(1, 1, -1, -1) in this basis
__global__ void k_g5_clover(SpinorPtr z, SpinorPtr r, CloverPtr C, int vh) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= vh) return;
  cf v[12];
  SpinorIO<P32>::load(r, i, v);
  if (C.p) {
    float cb[36];
    clover_load_block(C, i, 0, cb);
    clover_apply_block(cb, v, 0);
    clover_load_block(C, i, 1, cb);
    clover_apply_block(cb, v, 1);
  }
#pragma unroll
  for (int k = 6; k < 12; ++k) v[k] = mk(-v[k].re, -v[k].im);
  SpinorIO<P32>::store(z, i, v);
}

// out = a * (r - k2 gamma5 u)
__global__ void k_g5_xpay(SpinorPtr out, SpinorPtr r, SpinorPtr u, float k2, float a, int vh) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= vh) return;
  cf rv[12], uv[12];
  SpinorIO<P32>::load(r, i, rv);
  SpinorIO<P32>::load(u, i, uv);
#pragma unroll
  for (int k = 0; k < 12; ++k) {
    const float sg = k < 6 ? 1.f : -1.f;
    rv[k] = mk(a * (rv[k].re - k2 * sg * uv[k].re), a * (rv[k].im - k2 * sg * uv[k].im));
  }
  SpinorIO<P32>::store(out, i, rv);
}

// res = y - b (FP32 odd), |res|^2 into partials per block (FP64)
__global__ void k_sub_norm(SpinorPtr res, SpinorPtr y, SpinorPtr b, float bscale, int vh, double *partials) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double acc[1] = {0};
  if (i < vh) {
    cf yv[12], bv[12];
    SpinorIO<P32>::load(y, i, yv);
    SpinorIO<P32>::load(b, i, bv);
#pragma unroll
    for (int k = 0; k < 12; ++k) {
      yv[k] = mk(yv[k].re - bscale * bv[k].re, yv[k].im - bscale * bv[k].im);
      acc[0] += (double)yv[k].re * yv[k].re + (double)yv[k].im * yv[k].im;
    }
    SpinorIO<P32>::store(res, i, yv);
  }
  block_partial<1>(acc, partials, blockIdx.x);
}

// FP32 odd vector <- external one-parity complex128 field, scaled
__global__ void k_load_odd(SpinorPtr dst, const double2 *ext, int vh) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= vh) return;
  cf v[12];
  for (int k = 0; k < 12; ++k) {
    const double2 a = ext[(size_t)i * 12 + k];
    v[k] = mk((float)a.x, (float)a.y);
  }
  SpinorIO<P32>::store(dst, i, v);
}

__global__ void k_store_odd(double2 *ext, SpinorPtr src, int vh) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= vh) return;
  cf v[12];
  SpinorIO<P32>::load(src, i, v);
  for (int k = 0; k < 12; ++k) ext[(size_t)i * 12 + k] = make_double2(v[k].re, v[k].im);
}

// external [C][2][Vh][12] complex128 <-> internal channel fields (both parities)
__global__ void k_full_in(SpinorPtr base, const double2 *ext, int nch, int vh) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)nch * 2 * vh) return;
  const int i = (int)(idx % vh), p = (int)((idx / vh) % 2), c = (int)(idx / vh / 2);
  cf v[12];
  for (int k = 0; k < 12; ++k) {
    const double2 a = ext[((size_t)(c * 2 + p) * vh + i) * 12 + k];
    v[k] = mk((float)a.x, (float)a.y);
  }
  SpinorIO<P32>::store(chan_field(base, c, p, vh), i, v);
}

__global__ void k_full_out(double2 *ext, SpinorPtr base, int nch, int vh) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)nch * 2 * vh) return;
  const int i = (int)(idx % vh), p = (int)((idx / vh) % 2), c = (int)(idx / vh / 2);
  cf v[12];
  SpinorIO<P32>::load(chan_field(base, c, p, vh), i, v);
  for (int k = 0; k < 12; ++k) ext[((size_t)(c * 2 + p) * vh + i) * 12 + k] = make_double2(v[k].re, v[k].im);
}

__global__ void k_add_f32(float *acc, const float *x, int n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) acc[k] += x[k];
}

// Adam on interleaved (re, im) FP32 parameters
__global__ void k_adam(float *w, float *m, float *v, const float *g, int n, float lr, float b1, float b2, float c1,
                       float c2, float eps, float gscale) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const float gg = g[k] * gscale;
  const float mm = b1 * m[k] + (1.f - b1) * gg;
  const float vv = b2 * v[k] + (1.f - b2) * gg * gg;
  m[k] = mm;
  v[k] = vv;
  w[k] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}

// ---------------------------------------------------------------- host side
// operands of a layer on the split-FP16 kernel
struct SplitIO {
  const void *A;            // packed weights
  const float *winv;        // inverse weight scale
  const unsigned *in_max;   // largest |component| of the input field
  unsigned *out_max;        // receives that of the output field
  // fusion with the next, single-output layer (null A2: none)
  const void *A2 = nullptr;
  const float *winv2 = nullptr;
  void *mix = nullptr;
};

struct Net : Precond {
  Lattice *lat = nullptr;
  int nl = 0;
  int chan[kNetMaxLayers + 1] = {0};
  std::vector<size_t> woff;  // float2 offset of each layer's weights
  size_t nw = 0;             // complex weights
  DevBuf W, Wt, gW, m, v, wpart;
  DevBuf act;                // layer inputs: layer l has chan[l] channel fields
  std::vector<size_t> aoff;  // byte offset of layer l's input block
  DevBuf gbuf[2];            // gradient ping-pong (max channels)
  DevBuf tmp;                // odd FP32 vectors for training
  DevBuf red;                // reduction partials
  int adam_t = 0;
  int co_max = 4;   // output channels per thread of the vector layer kernel (1, 2, 3, 4 or 8)
  bool use_mfma = true;  // matrix-core layer kernel (QCU_NET_KERNEL=valu selects the vector one)
  DevBuf Apk;       // A operands of the matrix-core kernel for one layer
  DevBuf mixb;      // term fields of the two-pass single-output layer (9 fields)
  int block = 128;  // threads per block of the layer kernel
  int tile = 8;     // rows per tile side of the matrix-core wave order (0: site order)
  // split-FP16 hidden layers (8 -> 4 or 8 channels; QCU_NET_SPLIT=0 keeps them
  // on the FP32 matrix-core kernel)
  bool use_split = true;
  int split_order = -1;         // chunk order of the split kernel (QCU_NET_SPLIT_ORDER; -1: z slabs per XCD
                                // for one output parity, else interleaved)
  int split_ns = 32;            // sites per wave of the split kernel (QCU_NET_SPLIT_SHAPE)
  bool split_wgrad = true;      // 8 -> 8 weight gradients on split FP16 (QCU_NET_SPLIT_WGRAD=0: FP32)
  bool wg_lds = true;           // weight-gradient operands staged through LDS (QCU_NET_WGRAD_LDS=0: direct loads)
  DevBuf spk;                   // packed operands: [direction][layer] x kSplitPackBytes
  DevBuf smax;                  // [direction][layer + 1] largest |component| of each field (float bits)
  DevBuf swinv;                 // [direction][layer] inverse weight scale, then the fused layer's
  DevBuf fpk;                   // packed operands of the fused single-output layer
  DevBuf p2pk;                  // combined operands of the second layer (net_pack_second), then 2 inverse scales
  DevBuf mixb2;                 // the second layer's share of the last layer's term fields (third_split)
  DevBuf wcomp;                 // W4_8 composed with W3 (net_compose_last)
  DevBuf fpk2;                  // its packed operands (net_pack_fuse), then its inverse scale
  unsigned long long wver = 1;  // weight version (bumped on every change of W)
  unsigned long long packed_ver = 0;
  size_t field_bytes() const { return (size_t)lat->g.vh * 2 * 96; }  // one channel, both parities
  // layer k (or its transpose, dir 1: cin <-> cout) runs on the split kernel
  bool split_layer(int k, int dir) const {
    const int ci = dir ? chan[k + 1] : chan[k], co = dir ? chan[k] : chan[k + 1];
    return use_mfma && use_split && ci == 8 && (co == 4 || co == 8);
  }
  SplitIO split_io(int k, int dir) {
    SplitIO io;
    io.A = (const char *)spk.p + ((size_t)dir * kNetMaxLayers + k) * kSplitPackBytes;
    io.winv = (const float *)swinv.p + dir * kNetMaxLayers + k;
    // forward: input act[k], output act[k + 1]; transpose: input g[k + 1], output g[k]
    unsigned *mx = (unsigned *)smax.p + dir * (kNetMaxLayers + 1);
    io.in_max = dir ? mx + k + 1 : mx + k;
    io.out_max = dir ? mx + k : mx + k + 1;
    return io;
  }
  int pack_split(int dir, hipStream_t s);
  int reset_max(int dir, hipStream_t s);
  // the first layer (1 -> 8) on an input that is zero on the even sites runs
  // on the split kernel net_layer_first when its input's largest component
  // was recorded (load_input)
  bool first_split() const { return use_mfma && use_split && nl >= 2 && chan[0] == 1 && chan[1] == 8; }
  // ... and then the second layer (8 -> 8) on net_layer_second, the first
  // layer's odd sites left implicit, when only the network output is needed
  bool second_split() const { return first_split() && nl >= 3 && chan[2] == 8 && split_layer(1, 0) && split_ns == 32; }
  float *p2winv() const { return (float *)((char *)p2pk.p + (size_t)kFirstPackOps * 16); }
  // ... and, in a 1-8-8-8-1 network, the third layer on its even output sites
  // only: its odd outputs feed only the last layer's site term, whose
  // composition with the third layer (W4_8 W3_u) the second layer applies in
  // its epilogue, writing mixb2
  bool third_split() const { return second_split() && nl == 4 && fusable() && mixb2.p != nullptr; }
  float *fpk2winv() const { return (float *)((char *)fpk2.p + kFusePackBytes); }
  bool input_fresh = false;  // load_input ran since the last forward pass: maxima cleared, input's recorded
  // act[0] <- the odd vector in (precision prec) as channel 0, zero on the
  // even sites (zero_even false: the even sites are left unwritten, for a
  // forward pass that never reads them); clears the field maxima of the next
  // forward pass first
  int load_input(SpinorPtr in, int prec, const SysState *st, hipStream_t s, bool zero_even = true);
  // the preconditioner's forward pass (apply) reads the input's odd sites only:
  // its first two layers run as net_layer_second
  bool apply_reads_odd_input_only() const { return second_split() && !(fusable() && nl == 3); }
  // the last layer (8 -> 1 channels) can be folded into the epilogue of the
  // 8 -> 8 layer before it when only the network output is needed
  bool fusable() const {
    return nl >= 2 && chan[nl] == 1 && chan[nl - 1] == 8 && chan[nl - 2] == 8 && split_layer(nl - 2, 0) &&
           net_split_can_fuse(split_ns) && mixb.p != nullptr;
  }
  // the second pass of the single-output layer: transports and sums the term fields in mixb
  // direct_prec >= 0: out is an odd vector of that precision (the odd output parity only);
  // two: the term fields are the sums of mixb and mixb2 (site term: mixb2's)
  int gather(SpinorPtr out, int parity0, int nparity, const SysState *st, hipStream_t s, int direct_prec = -1,
             bool two = false);

  int init(Lattice *l, int nlayers, const int *ch, uint64_t seed, double init_scale);
  SpinorPtr act_ptr(int layer) const {
    SpinorPtr s;
    s.p = (char *)act.p + aoff[layer];
    s.norm = nullptr;
    s.stride = lat->g.vh;
    return s;
  }
  // io: the split kernel's operands and field maxima (null: FP32 kernels only,
  // no maximum recorded)
  int layer(SpinorPtr in, SpinorPtr out, const float2 *Wl, int cin, int cout, int parity0, int nparity,
            const SysState *st, hipStream_t s, int zero_par = -1, const SplitIO *io = nullptr);
  // zero_par0: parity on which the network input is zero (-1: none);
  // keep_acts false: only act[nl] is needed afterwards (the last layer may be fused)
  // direct (with odd_only_last, keep_acts false): the odd output may go straight into this
  // vector of precision direct_prec; *wrote tells whether it did (fused last layer)
  // only >= 0: run layer `only` alone, on the activations and field maxima a
  // full pass of the same kind left (timing)
  int forward(const SysState *st, hipStream_t s, bool odd_only_last, int zero_par0, bool keep_acts = true,
              const SpinorPtr *direct = nullptr, int direct_prec = P32, bool *wrote = nullptr, int only = -1);
  // gout_zero_par: parity on which gout_last is zero; act0_zero_par: parity on
  // which the network input act[0] is zero (-1: none)
  int backward(SpinorPtr gout_last, hipStream_t s, bool want_grad_in, SpinorPtr grad_in, int gout_zero_par = -1,
               int act0_zero_par = -1);
  int apply(Solver &S, SpinorPtr out, SpinorPtr in, int prec, SysState *st) override;
  unsigned long long id = 0;
  // changes with the weights too: solver graphs captured with other weights
  // (or their packed operands) are re-captured
  unsigned long long uid() const override { return (id << 32) ^ wver; }
};

static std::atomic<unsigned long long> g_net_ids{0};

static size_t ndiv(size_t a, size_t b) { return (a + b - 1) / b; }

int Net::init(Lattice *l, int nlayers, const int *ch, uint64_t seed, double init_scale) {
  QCU_CHECK(l && l->gauge_loaded, QCU_ERR_STATE, "net: lattice has no gauge field");
  QCU_CHECK(nlayers >= 1 && nlayers <= kNetMaxLayers, QCU_ERR_ARG, "net: 1 to 8 layers");
  QCU_CHECK((l->prec_mask >> P32) & 1, QCU_ERR_STATE, "net: FP32 links not resident");
  // the layer kernels address one channel field of one parity (96 bytes per
  // site) with 32-bit per-lane byte offsets
  QCU_CHECK((size_t)l->g.vh * 96 < ((size_t)1 << 32), QCU_ERR_DIMS, "net: lattice too large (Vh * 96 bytes >= 4 GiB)");
  lat = l;
  nl = nlayers;
  id = ++g_net_ids;
  for (int k = 0; k <= nl; ++k) {
    QCU_CHECK(ch[k] >= 1 && ch[k] <= kNetMaxChan, QCU_ERR_ARG, "net: 1 to 8 channels per layer");
    chan[k] = ch[k];
  }
  QCU_CHECK(chan[0] == 1 && chan[nl] == 1, QCU_ERR_ARG, "net: one input and one output channel");
  woff.assign(nl + 1, 0);
  for (int k = 0; k < nl; ++k) woff[k + 1] = woff[k] + (size_t)9 * chan[k] * chan[k + 1] * 16;
  nw = woff[nl];
  QCU_TRY(W.alloc(nw * 8));
  QCU_TRY(Wt.alloc(nw * 8));
  QCU_TRY(gW.alloc(nw * 8));
  QCU_TRY(m.alloc(nw * 8));
  QCU_TRY(v.alloc(nw * 8));
  aoff.assign(nl + 1, 0);
  size_t off = 0;
  for (int k = 0; k <= nl; ++k) {
    aoff[k] = off;
    off += chan[k] * field_bytes();
  }
  QCU_TRY(act.alloc(off));
  int cmax = 1;
  for (int k = 0; k <= nl; ++k) cmax = std::max(cmax, chan[k]);
  QCU_TRY(gbuf[0].alloc(cmax * field_bytes()));
  QCU_TRY(gbuf[1].alloc(cmax * field_bytes()));
  QCU_TRY(tmp.alloc((size_t)6 * lat->g.vh * 96));
  QCU_TRY(red.alloc(((size_t)lat->g.vh / 64 + 16) * 8));
  // weight-gradient partials: combos (9 cin * groups) * blocks * CO*32 floats
  size_t wp = 0;
  for (int k = 0; k < nl; ++k) {
    const int co = chan[k + 1] >= 2 ? 2 : 1;
    wp = std::max(wp, (size_t)9 * chan[k] * ndiv(chan[k + 1], co) * kWgradBlocks * co * 32);
  }
  // matrix-core weight gradient: 9 terms x blocks x (up to 2 x 2 tiles x 2 parts x 4) registers x 64 lanes
  wp = std::max(wp, (size_t)9 * kWgBlocks * 32 * 64);
  QCU_TRY(wpart.alloc(wp * 4));
  // initialisation: the identity through channel 0 plus small seeded noise
  std::vector<float> h(nw * 2, 0.f);
  uint64_t z = seed * 0x9e3779b97f4a7c15ULL + 12345;
  auto rnd = [&]() {
    z ^= z << 13;
    z ^= z >> 7;
    z ^= z << 17;
    return ((double)(z >> 11) / 9007199254740992.0) * 2.0 - 1.0;
  };
  for (int k = 0; k < nl; ++k) {
    const int ci = chan[k], co = chan[k + 1];
    const double sc = init_scale / sqrt(9.0 * ci * 4.0);
    for (size_t e = woff[k]; e < woff[k + 1]; ++e) {
      h[2 * e] = (float)(sc * rnd());
      h[2 * e + 1] = (float)(sc * rnd());
    }
    // site term, output 0 <- input 0: unit spin matrix
    for (int s = 0; s < 4; ++s) h[2 * (woff[k] + (((size_t)8 * co + 0) * ci + 0) * 16 + s * 5)] += 1.f;
  }
  QCU_HIP(hipMemcpy(W.p, h.data(), nw * 8, hipMemcpyHostToDevice));
  ++wver;
  QCU_HIP(hipMemset(m.p, 0, nw * 8));
  QCU_HIP(hipMemset(v.p, 0, nw * 8));
  adam_t = 0;
  QCU_TRY(Apk.alloc((size_t)9 * kNetMaxChan * 2 * kApkParts * 64 * 4));
  QCU_TRY(spk.alloc((size_t)2 * kNetMaxLayers * kSplitPackBytes));
  QCU_TRY(smax.alloc((size_t)2 * (kNetMaxLayers + 1) * 4));
  QCU_TRY(swinv.alloc((size_t)(2 * kNetMaxLayers + 1) * 4));
  QCU_TRY(fpk.alloc(kFusePackBytes));
  QCU_TRY(p2pk.alloc((size_t)kFirstPackOps * 16 + 16));
  QCU_TRY(wcomp.alloc((size_t)9 * 8 * 16 * sizeof(float2)));
  QCU_TRY(fpk2.alloc(kFusePackBytes + 16));
  QCU_HIP(hipMemset(smax.p, 0, (size_t)2 * (kNetMaxLayers + 1) * 4));
  if (const char *e = getenv("QCU_NET_SPLIT")) use_split = atoi(e) != 0;
  if (const char *e = getenv("QCU_NET_SPLIT_ORDER")) split_order = atoi(e);
  if (const char *e = getenv("QCU_NET_SPLIT_WGRAD")) split_wgrad = atoi(e) != 0;
  if (const char *e = getenv("QCU_NET_WGRAD_LDS")) wg_lds = atoi(e) != 0;
  split_ns = net_split_shape_default();
  if (const char *e = getenv("QCU_NET_KERNEL")) use_mfma = strcmp(e, "valu") != 0;
  if (const char *e = getenv("QCU_NET_TILE")) tile = atoi(e);  // 0: site order
  // a layer with one output and several input channels (the last layer, and
  // the input gradient of the first) runs in two passes through mixb
  int cbig = 1;
  for (int k = 0; k <= nl; ++k) cbig = std::max(cbig, chan[k]);
  if (use_mfma && cbig > 1) QCU_TRY(mixb.alloc(9 * field_bytes()));
  // the second set of term fields of a 1-8-8-8-1 preconditioner
  if (use_mfma && use_split && nl == 4 && chan[0] == 1 && chan[1] == 8 && chan[2] == 8 && chan[3] == 8 &&
      chan[4] == 1)
    QCU_TRY(mixb2.alloc(9 * field_bytes()));
  // launch shape of the layer kernel: QCU_NET_SHAPE="co,block" overrides it
  if (const char *e = getenv("QCU_NET_SHAPE")) {
    int c = 0, b = 0;
    if (sscanf(e, "%d,%d", &c, &b) == 2 && (c == 1 || c == 2 || c == 3 || c == 4 || c == 8) &&
        (b == 64 || b == 128 || b == 256)) {
      co_max = c;
      block = b;
    }
  }
  return 0;
}

int Net::layer(SpinorPtr in, SpinorPtr out, const float2 *Wl, int cin, int cout, int parity0, int nparity,
               const SysState *st, hipStream_t s, int zero_par, const SplitIO *io) {
  NetLayerArgs a;
  a.in = in;
  a.out = out;
  a.W = Wl;
  a.U = lat->links(P32);
  a.g = lat->g;
  a.cin = cin;
  a.cout = cout;
  a.parity0 = parity0;
  a.zero_par = zero_par;
  a.st = st;
  a.out_max = io ? io->out_max : nullptr;
  // a hidden layer with packed split operands: FP16 matrix cores
  if (io && io->A && use_mfma && use_split && cin == 8 && (cout == 4 || cout == 8) && zero_par < 0) {
    NetSplitArgs b;
    b.in = (const char *)in.p;
    b.out = (char *)out.p;
    b.A = (const uint4 *)io->A;
    b.winv = io->winv;
    b.in_max = io->in_max;
    b.out_max = io->out_max;
    b.U = a.U;
    b.g = lat->g;
    b.parity0 = parity0;
    b.npar = nparity;
    b.ns = split_ns;
    b.order = split_order >= 0 ? split_order : (nparity == 1 ? 2 : 0);
    b.st = st;
    b.A2 = (const uint4 *)io->A2;
    b.winv2 = io->winv2;
    b.mix = (char *)io->mix;
    return net_layer_split(b, cout, s);
  }
  // one output channel: the input channels are mixed per term on the matrix
  // cores at the input sites, then transported once per term
  if (use_mfma && cout == 1 && cin > 1 && zero_par < 0) {
    QCU_CHECK(mixb.p && (nparity == 1 || parity0 == 0), QCU_ERR_STATE, "net: two-pass layer set-up");
    const int vh = lat->g.vh;
    NetMixArgs m;
    m.in = in;
    m.mix.p = mixb.p;
    m.mix.norm = nullptr;
    m.mix.stride = vh;
    m.W = Wl;
    m.vh = vh;
    m.cin = cin;
    m.out_par = nparity == 2 ? -1 : parity0;
    m.st = st;
    const dim3 g1((unsigned)ndiv(vh, 64), 2);  // 4 waves of 16 sites per block
    const bool both = nparity == 2;
    if (cin == 8) {
      if (both) hipLaunchKernelGGL((k_net_mix<8, true>), g1, dim3(256), 0, s, m);
      else hipLaunchKernelGGL((k_net_mix<8, false>), g1, dim3(256), 0, s, m);
    } else {
      if (both) hipLaunchKernelGGL((k_net_mix<0, true>), g1, dim3(256), 0, s, m);
      else hipLaunchKernelGGL((k_net_mix<0, false>), g1, dim3(256), 0, s, m);
    }
    QCU_HIP(hipGetLastError());
    return gather(out, parity0, nparity, st, s);
  }
  if (use_mfma && cout > 1) {
    const int mtiles = (cout + 3) / 4;
    const int total = 9 * cin * mtiles * kApkParts * 64;
    hipLaunchKernelGGL(k_net_pack, dim3((unsigned)ndiv(total, 256)), dim3(256), 0, s, Wl, (float *)Apk.p, cin, cout,
                       mtiles);
    const int wpb = block / 64;  // waves per block
    const Geom &g = lat->g;
    a.ord = WaveOrder{};
    dim3 grid((unsigned)ndiv(g.vh, 16 * wpb), nparity, 1);
    // a layer whose input is zero on one parity (the preconditioner's first
    // layer) keeps the site order, one output parity per grid row: its two
    // parities do different work (8 hop terms against the site term), and at
    // 32^3x64 it runs in 0.68 ms that way against 0.75 ms in the tiled order
    if (tile > 0 && g.lxh % 16 == 0 && zero_par < 0) {
      auto divisor = [&](int n) {  // largest divisor of n not above the tile size
        int d = tile < n ? tile : n;
        while (n % d) --d;
        return d;
      };
      a.ord.tiled = 1;
      a.ord.nx = g.lxh / 16;
      a.ord.ty = divisor(g.ly);
      a.ord.tz = divisor(g.lz);
      a.ord.nyt = g.ly / a.ord.ty;
      a.ord.npar = nparity;
      a.ord.nwaves = nparity * (g.vh / 16);
      grid = dim3((unsigned)ndiv(a.ord.nwaves, wpb), 1, 1);
    }
    const float *ap = (const float *)Apk.p;
    const int ci = cin == 8 ? 8 : cin == 1 ? 1 : 0;
#define QCU_NET_MFMA_CASE(MT_, CI_) \
  if (mtiles == MT_ && ci == CI_) hipLaunchKernelGGL((k_net_layer_mfma<MT_, CI_>), grid, dim3(block), 0, s, a, ap);
    QCU_NET_MFMA_CASE(1, 0) QCU_NET_MFMA_CASE(1, 1) QCU_NET_MFMA_CASE(1, 8)
    QCU_NET_MFMA_CASE(2, 0) QCU_NET_MFMA_CASE(2, 1) QCU_NET_MFMA_CASE(2, 8)
#undef QCU_NET_MFMA_CASE
    QCU_HIP(hipGetLastError());
    return 0;
  }
  const int bs = block;
  int co = cout < co_max ? cout : co_max;
  if (co > 4 && co < 8) co = 4;
  a.ngroups = (int)ndiv(cout, co);
  dim3 grid((unsigned)(ndiv(lat->g.vh, bs) * a.ngroups), nparity, 1);
  const int ci = cin == 1 ? 1 : cin == 8 ? 8 : 0;
#define QCU_NET_CASE(CO_, CI_) \
  if (co == CO_ && ci == CI_) hipLaunchKernelGGL((k_net_layer<CO_, CI_>), grid, dim3(bs), 0, s, a);
  QCU_NET_CASE(1, 0) QCU_NET_CASE(1, 1) QCU_NET_CASE(1, 8)
  QCU_NET_CASE(2, 0) QCU_NET_CASE(2, 1) QCU_NET_CASE(2, 8)
  QCU_NET_CASE(3, 0) QCU_NET_CASE(3, 1) QCU_NET_CASE(3, 8)
  QCU_NET_CASE(4, 0) QCU_NET_CASE(4, 1) QCU_NET_CASE(4, 8)
  QCU_NET_CASE(8, 0) QCU_NET_CASE(8, 1) QCU_NET_CASE(8, 8)
#undef QCU_NET_CASE
  QCU_HIP(hipGetLastError());
  return 0;
}

// act[0] holds the input; fills act[1..nl] (act[nl] = output)
int Net::gather(SpinorPtr out, int parity0, int nparity, const SysState *st, hipStream_t s, int direct_prec,
                bool two) {
  NetGatherArgs ga;
  ga.mix.p = mixb.p;
  ga.mix.norm = nullptr;
  ga.mix.stride = lat->g.vh;
  ga.mix2 = ga.mix;
  ga.mix2.p = mixb2.p;
  QCU_CHECK(!two || mixb2.p, QCU_ERR_STATE, "net: second term-field set not allocated");
  ga.out = out;
  ga.U = lat->links(P32);
  ga.g = lat->g;
  ga.parity0 = parity0;
  ga.st = st;
  const dim3 grid((unsigned)ndiv(lat->g.vh, 256), nparity);
  QCU_CHECK(direct_prec < 0 || (parity0 == 1 && nparity == 1), QCU_ERR_ARG, "net: direct output is the odd parity");
  if (two) {
    switch (direct_prec) {
      case P64: hipLaunchKernelGGL((k_net_gather<P64, true, true>), grid, dim3(256), 0, s, ga); break;
      case P32: hipLaunchKernelGGL((k_net_gather<P32, true, true>), grid, dim3(256), 0, s, ga); break;
      case P16: hipLaunchKernelGGL((k_net_gather<P16, true, true>), grid, dim3(256), 0, s, ga); break;
      default: hipLaunchKernelGGL((k_net_gather<P32, false, true>), grid, dim3(256), 0, s, ga); break;
    }
  } else {
    switch (direct_prec) {
      case P64: hipLaunchKernelGGL((k_net_gather<P64, true>), grid, dim3(256), 0, s, ga); break;
      case P32: hipLaunchKernelGGL((k_net_gather<P32, true>), grid, dim3(256), 0, s, ga); break;
      case P16: hipLaunchKernelGGL((k_net_gather<P16, true>), grid, dim3(256), 0, s, ga); break;
      default: hipLaunchKernelGGL((k_net_gather<P32, false>), grid, dim3(256), 0, s, ga); break;
    }
  }
  QCU_HIP(hipGetLastError());
  return 0;
}

// Packs the split operands of every split layer in direction dir (0: the
// weights, 1: their transposes in Wt) and records the weight version packed.
int Net::pack_split(int dir, hipStream_t s) {
  if (dir == 0 && first_split()) {
    const SplitIO io = split_io(0, 0);
    QCU_TRY(net_pack_first((const float2 *)W.p + woff[0], (void *)io.A, (float *)io.winv, s));
  }
  if (dir == 0 && second_split())
    QCU_TRY(net_pack_second((const float2 *)W.p + woff[0], (const float2 *)W.p + woff[1], p2pk.p, p2winv(), s));
  if (dir == 0 && third_split()) {
    QCU_TRY(net_compose_last((const float2 *)W.p + woff[2], (const float2 *)W.p + woff[3], (float2 *)wcomp.p, s));
    QCU_TRY(net_pack_fuse((const float2 *)wcomp.p, fpk2.p, fpk2winv(), s));
  }
  for (int k = 0; k < nl; ++k) {
    if (!split_layer(k, dir)) continue;
    const SplitIO io = split_io(k, dir);
    const float2 *w = (const float2 *)(dir ? Wt.p : W.p) + woff[k];
    QCU_TRY(net_pack_split(w, dir ? chan[k + 1] : chan[k], dir ? chan[k] : chan[k + 1], (void *)io.A,
                           (float *)io.winv, split_ns, s));
  }
  if (dir == 0 && fusable())
    QCU_TRY(net_pack_fuse((const float2 *)W.p + woff[nl - 1], fpk.p, (float *)swinv.p + 2 * kNetMaxLayers, s));
  if (dir == 0) packed_ver = wver;
  return 0;
}

__global__ void k_fill_u32(unsigned *p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

// clears the field maxima of direction dir before a pass writes them
int Net::reset_max(int dir, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, (unsigned *)smax.p + dir * (kNetMaxLayers + 1),
                     kNetMaxLayers + 1);
  QCU_HIP(hipGetLastError());
  return 0;
}

int Net::forward(const SysState *st, hipStream_t s, bool odd_only_last, int zero_par0, bool keep_acts,
                 const SpinorPtr *direct, int direct_prec, bool *wrote, int only) {
  if (wrote) *wrote = false;
  bool any = first_split();
  for (int k = 0; k < nl; ++k) any = any || split_layer(k, 0);
  // the first layer on the split kernel needs the input's maximum, recorded by load_input
  const bool first = first_split() && zero_par0 == 0 && input_fresh;
  if (any) {
    if (packed_ver != wver) QCU_TRY(pack_split(0, s));
    if (!input_fresh && only < 0) QCU_TRY(reset_max(0, s));
  }
  input_fresh = false;
  // fused: the 8 -> 8 layer before the last writes the last layer's mixed
  // term fields, so act[nl - 1] is never written and the last layer is only
  // its gather pass
  const bool fuse = !keep_acts && odd_only_last && fusable();
  // the first two layers as net_layer_first (even sites) + net_layer_second
  const bool second = first && !keep_acts && second_split() && !(fuse && nl - 2 == 1);
  const bool third = second && fuse && third_split();
  for (int k = 0; k < nl; ++k) {
    const bool last = k == nl - 1;
    const int p0 = (last && odd_only_last) ? 1 : 0;
    // third: the fused layer on its even output sites only
    const int np = (last && odd_only_last) || (third && k == nl - 2) ? 1 : 2;
    if (only >= 0 && k != only) continue;
    if (fuse && last) {
      if (direct) {
        QCU_TRY(gather(*direct, p0, np, st, s, direct_prec, third));
        if (wrote) *wrote = true;
      } else {
        QCU_TRY(gather(act_ptr(k + 1), p0, np, st, s, -1, third));
      }
      break;
    }
    // the kernel before a split layer records the largest component it writes
    SplitIO io = split_io(k, 0);
    if (!split_layer(k, 0) && !(k == 0 && first)) io.A = nullptr;
    if (!(k + 1 < nl && split_layer(k + 1, 0))) io.out_max = nullptr;
    if (fuse && k == nl - 2) {
      io.A2 = fpk.p;
      io.winv2 = (const float *)swinv.p + 2 * kNetMaxLayers;
      io.mix = mixb.p;
    }
    if (k == 0 && first) {
      NetFirstArgs f;
      f.in = (const char *)act_ptr(0).p;
      f.out = (char *)act_ptr(1).p;
      f.A = (const uint4 *)io.A;
      f.winv = io.winv;
      f.in_max = io.in_max;
      f.out_max = io.out_max;
      f.U = lat->links(P32);
      f.g = lat->g;
      f.st = st;
      QCU_TRY(net_layer_first(f, s, second));
      continue;
    }
    if (k == 1 && second) {
      NetSecondArgs q;
      q.in = (const char *)act_ptr(0).p;
      q.l1 = (const char *)act_ptr(1).p;
      q.out = (char *)act_ptr(2).p;
      q.A = (const uint4 *)p2pk.p;
      q.winv = p2winv();
      q.Ah = (const uint4 *)io.A;
      q.winvh = io.winv;
      q.in_max = split_io(0, 0).in_max;
      q.l1_max = io.in_max;
      q.out_max = io.out_max;
      q.U = lat->links(P32);
      q.g = lat->g;
      q.st = st;
      if (third) {
        q.A2 = (const uint4 *)fpk2.p;
        q.winv2 = fpk2winv();
        q.mix = (char *)mixb2.p;
      }
      QCU_TRY(net_layer_second(q, s));
      continue;
    }
    QCU_TRY(layer(act_ptr(k), act_ptr(k + 1), (const float2 *)W.p + woff[k], chan[k], chan[k + 1], p0, np, st, s,
                  k == 0 ? zero_par0 : -1, any ? &io : nullptr));
  }
  return 0;
}

// gout_last: gradient w.r.t. act[nl]; accumulates nothing, writes gW (this sample)
int Net::backward(SpinorPtr gout, hipStream_t s, bool want_grad_in, SpinorPtr grad_in, int gout_zero_par,
                  int act0_zero_par) {
  // conjugate-transposed weights for the data gradient
  for (int k = 0; k < nl; ++k) {
    const int total = 9 * chan[k] * chan[k + 1] * 16;
    hipLaunchKernelGGL(k_net_transpose, dim3((unsigned)ndiv(total, 256)), dim3(256), 0, s,
                       (const float2 *)W.p + woff[k], (float2 *)Wt.p + woff[k], chan[k], chan[k + 1]);
  }
  QCU_HIP(hipGetLastError());
  bool any = false;
  for (int k = 0; k < nl; ++k) any = any || split_layer(k, 1);
  if (any) {
    QCU_TRY(pack_split(1, s));
    QCU_TRY(reset_max(1, s));
  }
  SpinorPtr gcur = gout;
  int ping = 0;
  for (int k = nl - 1; k >= 0; --k) {
    const int cin = chan[k], cout = chan[k + 1];
    // weight gradient of layer k
    NetWgradArgs a;
    a.act = act_ptr(k);
    a.gout = gcur;
    a.U = lat->links(P32);
    a.g = lat->g;
    a.cin = cin;
    a.cout = cout;
    a.partials = (float *)wpart.p;
    a.act_zero = k == 0 ? act0_zero_par : -1;
    a.gout_zero = k == nl - 1 ? gout_zero_par : -1;
    // a layer that is split in both directions has both field maxima: its
    // input's from the forward pass, its output gradient's from this one
    const bool wsplit = split_wgrad && split_layer(k, 0) && split_layer(k, 1) && cin == 8 && cout == 8;
    if (wsplit) {
      a.amax = split_io(k, 0).in_max;
      a.gmax = split_io(k, 1).in_max;
    }
    const int total = 9 * cout * cin * 16;
    if (use_mfma) {
      const int mt = (cout + 3) / 4, nt = (cin + 3) / 4;
      static_assert(kWgBlocks % 8 == 0, "term blocks of a site block share an XCD");
      const dim3 grid(9 * kWgBlocks);
      float *pp = (float *)wpart.p;
      float2 *dst = (float2 *)gW.p + woff[k];
      const dim3 fg((unsigned)ndiv(total, 256));
      auto run = [&](auto mtc, auto ntc, auto splitc) {
        constexpr int MT_ = decltype(mtc)::value, NT_ = decltype(ntc)::value;
        if (wg_lds)
          hipLaunchKernelGGL((k_net_wgrad_mfma<MT_, NT_, decltype(splitc)::value, true>), grid, dim3(256), 0, s, a, pp);
        else
          hipLaunchKernelGGL((k_net_wgrad_mfma<MT_, NT_, decltype(splitc)::value, false>), grid, dim3(256), 0, s, a, pp);
        hipLaunchKernelGGL((k_net_wgrad_fin<MT_, NT_>), fg, dim3(256), 0, s, (const float *)pp, kWgBlocks, cin, cout,
                           dst, a.gmax, a.amax);
      };
      using I1 = std::integral_constant<int, 1>;
      using I2 = std::integral_constant<int, 2>;
      using F = std::false_type;
      if (mt == 1 && nt == 1) run(I1{}, I1{}, F{});
      if (mt == 1 && nt == 2) run(I1{}, I2{}, F{});
      if (mt == 2 && nt == 1) run(I2{}, I1{}, F{});
      // the split form exists for the 8 -> 8 shape only
      if (mt == 2 && nt == 2) {
        if (wsplit) run(I2{}, I2{}, std::true_type{});
        else run(I2{}, I2{}, F{});
      }
      QCU_HIP(hipGetLastError());
    } else {
      const int co = cout >= 2 ? 2 : 1;
      const int ng = (int)ndiv(cout, co);
      dim3 grid(kWgradBlocks, 9 * cin, ng);
      if (co == 2) hipLaunchKernelGGL(k_net_wgrad<2>, grid, dim3(256), 0, s, a);
      else hipLaunchKernelGGL(k_net_wgrad<1>, grid, dim3(256), 0, s, a);
      QCU_HIP(hipGetLastError());
      hipLaunchKernelGGL(k_net_wgrad_sum, dim3((unsigned)ndiv(total, 256)), dim3(256), 0, s,
                         (const float *)wpart.p, kWgradBlocks, cin, cout, co, ng, (float2 *)gW.p + woff[k], 0);
      QCU_HIP(hipGetLastError());
    }
    if (k == 0 && !want_grad_in) break;
    // data gradient: the forward layer with the transposed weights (cin <-> cout)
    SpinorPtr gnext;
    if (k == 0) {
      gnext = grad_in;
    } else {
      gnext.p = gbuf[ping].p;
      gnext.norm = nullptr;
      gnext.stride = lat->g.vh;
    }
    SplitIO io = split_io(k, 1);
    if (!split_layer(k, 1)) io.A = nullptr;
    if (!(k > 0 && split_layer(k - 1, 1))) io.out_max = nullptr;
    QCU_TRY(layer(gcur, gnext, (const float2 *)Wt.p + woff[k], cout, cin, 0, 2, nullptr, s,
                  k == nl - 1 ? gout_zero_par : -1, any ? &io : nullptr));
    gcur = gnext;
    ping ^= 1;
  }
  return 0;
}

int Net::load_input(SpinorPtr in, int prec, const SysState *st, hipStream_t s, bool zero_even) {
  const int vh = lat->g.vh, bs = 256;
  const unsigned nb = (unsigned)ndiv(vh, bs);
  const SpinorPtr a0 = act_ptr(0);
  bool any = first_split();
  for (int k = 0; k < nl; ++k) any = any || split_layer(k, 0);
  if (any) QCU_TRY(reset_max(0, s));
  unsigned *rec = first_split() ? (unsigned *)split_io(0, 0).in_max : nullptr;
  const unsigned nbr = rec ? (unsigned)ndiv(vh, 4 * bs) : nb;
  const int ze = zero_even ? 1 : 0;
  switch (prec) {
    case P64: hipLaunchKernelGGL(k_net_embed<P64>, dim3(nbr), dim3(bs), 0, s, a0, in, vh, st, rec, ze); break;
    case P32: hipLaunchKernelGGL(k_net_embed<P32>, dim3(nbr), dim3(bs), 0, s, a0, in, vh, st, rec, ze); break;
    default: hipLaunchKernelGGL(k_net_embed<P16>, dim3(nbr), dim3(bs), 0, s, a0, in, vh, st, rec, ze); break;
  }
  QCU_HIP(hipGetLastError());
  input_fresh = any;
  return 0;
}

int Net::apply(Solver &S, SpinorPtr out, SpinorPtr in, int prec, SysState *st) {
  const int vh = lat->g.vh, bs = 256;
  const unsigned nb = (unsigned)ndiv(vh, bs);
  QCU_TRY(load_input(in, prec, st, S.stream, !apply_reads_odd_input_only()));
  bool wrote = false;
  QCU_TRY(forward(st, S.stream, true, 0, false, &out, prec, &wrote));
  if (wrote) return 0;  // the fused last layer wrote out
  const SpinorPtr an = act_ptr(nl);
  switch (prec) {
    case P64: hipLaunchKernelGGL(k_net_extract<P64>, dim3(nb), dim3(bs), 0, S.stream, out, an, vh, st); break;
    case P32: hipLaunchKernelGGL(k_net_extract<P32>, dim3(nb), dim3(bs), 0, S.stream, out, an, vh, st); break;
    default: hipLaunchKernelGGL(k_net_extract<P16>, dim3(nb), dim3(bs), 0, S.stream, out, an, vh, st); break;
  }
  QCU_HIP(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------- training
// Mhat (FP32) on odd fields through the lattice's one-system batch
static int mhat32(Lattice &lat, SpinorPtr out, SpinorPtr x, SpinorPtr tmp, hipStream_t s) {
  SpinorPtr none{};
  const LaunchCfg cfg = tuned_cfg(lat.self, P32);
  QCU_TRY(hop(lat.self, P32, tmp, x, 0, EPI_PLAIN, none, RED_NONE, none, nullptr, nullptr, cfg, s));
  QCU_TRY(hop(lat.self, P32, out, tmp, 1, EPI_XPAY, x, RED_NONE, none, nullptr, nullptr, cfg, s));
  return 0;
}

// out = a * Mhat^dag r = a (r - kappa^2 g5 D_oe Ae^-1 D_eo g5 Ao^-1 r)
static int mhat32_dag(Lattice &lat, SpinorPtr out, SpinorPtr r, SpinorPtr t1, SpinorPtr t2, float a, hipStream_t s) {
  const int vh = lat.g.vh, bs = 256;
  const unsigned nb = (unsigned)ndiv(vh, bs);
  SpinorPtr none{};
  const LaunchCfg cfg = tuned_cfg(lat.self, P32);
  hipLaunchKernelGGL(k_g5_clover, dim3(nb), dim3(bs), 0, s, t1, r, lat.clover_inv(P32, 1), vh);
  QCU_HIP(hipGetLastError());
  QCU_TRY(hop(lat.self, P32, t2, t1, 0, EPI_PLAIN, none, RED_NONE, none, nullptr, nullptr, cfg, s, true));
  QCU_TRY(hop(lat.self, P32, t1, t2, 1, EPI_PLAIN, none, RED_NONE, none, nullptr, nullptr, cfg, s, false));
  hipLaunchKernelGGL(k_g5_xpay, dim3(nb), dim3(bs), 0, s, out, r, t1, (float)(lat.kappa * lat.kappa), a, vh);
  QCU_HIP(hipGetLastError());
  return 0;
}

struct NetTrainer {
  Net &net;
  hipStream_t s;
  SpinorPtr odd(int k) const {
    SpinorPtr p;
    p.p = (char *)net.tmp.p + (size_t)k * net.lat->g.vh * 96;
    p.norm = nullptr;
    p.stride = net.lat->g.vh;
    return p;
  }
  // loss |Mhat K b - b|^2 / |b|^2 of one odd source already in odd(0); optionally
  // accumulates the weight gradient (gW += sample gradient) via gsum
  int sample(double *loss, bool grad, float *gsum) {
    Lattice &lat = *net.lat;
    const int vh = lat.g.vh, bs = 256;
    const unsigned nb = (unsigned)ndiv(vh, bs);
    // |b|^2
    double *red = (double *)net.red.p;
    hipLaunchKernelGGL(k_sub_norm, dim3(nb), dim3(bs), 0, s, odd(5), odd(0), odd(0), 0.f, vh, red);
    QCU_HIP(hipGetLastError());
    std::vector<double> h(nb);
    QCU_HIP(hipMemcpyAsync(h.data(), red, nb * 8, hipMemcpyDeviceToHost, s));
    QCU_HIP(hipStreamSynchronize(s));
    double b2 = 0;
    for (double x : h) b2 += x;
    if (!(b2 > 0)) {
      *loss = 0;
      return 0;
    }
    // y = K b / |b|
    QCU_TRY(net.load_input(odd(0), P32, nullptr, s));
    // only the odd part of the output enters the loss: the last layer runs on
    // the odd sites only, as in the preconditioner
    QCU_TRY(net.forward(nullptr, s, true, 0));
    hipLaunchKernelGGL(k_net_extract<P32>, dim3(nb), dim3(bs), 0, s, odd(1), net.act_ptr(net.nl), vh, nullptr);
    QCU_HIP(hipGetLastError());
    QCU_TRY(mhat32(lat, odd(2), odd(1), odd(3), s));
    // res = Mhat K b - b, unnormalised, then scaled
    hipLaunchKernelGGL(k_sub_norm, dim3(nb), dim3(bs), 0, s, odd(4), odd(2), odd(0), 1.f, vh, red);
    QCU_HIP(hipGetLastError());
    QCU_HIP(hipMemcpyAsync(h.data(), red, nb * 8, hipMemcpyDeviceToHost, s));
    QCU_HIP(hipStreamSynchronize(s));
    double r2 = 0;
    for (double x : h) r2 += x;
    *loss = r2 / b2;
    if (!grad) return 0;
    // dL/d(K b) on odd sites = 2 Mhat^dag res / |b|^2 ; zero on even sites
    QCU_TRY(mhat32_dag(lat, odd(2), odd(4), odd(3), odd(5), (float)(2.0 / b2), s));
    SpinorPtr gl;
    gl.p = net.gbuf[1].p;
    gl.norm = nullptr;
    gl.stride = vh;
    hipLaunchKernelGGL(k_net_embed<P32>, dim3(nb), dim3(bs), 0, s, gl, odd(2), vh, nullptr);
    QCU_HIP(hipGetLastError());
    // the backward pass ping-pongs through gbuf[0] and gbuf[1]; the last layer's
    // output gradient lives in gbuf[1] and is consumed first.  It is zero on the
    // even sites, and so is the network input.
    SpinorPtr none{};
    QCU_TRY(net.backward(gl, s, false, none, 0, 0));
    (void)gsum;
    return 0;
  }
};

}  // namespace qcu

using namespace qcu;

struct QcuNet_s {
  Net net;
};

namespace qcu {
int net_precond(void *handle, const Lattice *lat, Precond **out) {
  QcuNet_s *h = (QcuNet_s *)handle;
  QCU_CHECK(h->net.lat == lat, QCU_ERR_ARG, "the preconditioner was built for another lattice");
  // pack the split operands now, on the lattice stream, rather than inside a
  // captured solver iteration
  Net &net = h->net;
  bool any = false;
  for (int k = 0; k < net.nl; ++k) any = any || net.split_layer(k, 0);
  if (any && net.packed_ver != net.wver) QCU_TRY(net.pack_split(0, net.lat->stream));
  *out = &net;
  return 0;
}

// uniform values in [-1, 1) from a hash of the index
__global__ void k_fill_hash(float *p, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  unsigned h = (unsigned)i * 2654435761u;
  h ^= h >> 15;
  h *= 2246822519u;
  h ^= h >> 13;
  p[i] = (float)(h >> 8) * (2.f / 16777216.f) - 1.f;
}

// the network runs on its lattice's resident FP32 links: a lattice whose gauge
// (re)load failed has none
static int net_ready(const Net &net) {
  QCU_CHECK(net.lat->gauge_loaded && ((net.lat->prec_mask >> P32) & 1), QCU_ERR_STATE,
            "net: the lattice has no resident gauge field");
  return 0;
}
}  // namespace qcu

extern "C" {

int qcuNetCreate(QcuNetHandle *h, QcuLatticeHandle lat, int nlayers, const int *chan, uint64_t seed,
                 double init_scale) {
  clear_error();
  QCU_CHECK(h && lat && chan, QCU_ERR_ARG, "null argument");
  std::unique_ptr<QcuNet_s> n(new QcuNet_s());
  QCU_TRY(n->net.init(&lat->lat, nlayers, chan, seed, init_scale));
  *h = n.release();
  return 0;
}

int qcuNetDestroy(QcuNetHandle h) {
  clear_error();
  delete h;
  return 0;
}

long qcuNetWeightCount(QcuNetHandle h) { return h ? (long)h->net.nw : 0; }

// weights as complex128 [n] on the host
int qcuNetGetWeights(QcuNetHandle h, double *host, long n) {
  clear_error();
  QCU_CHECK(h && host && n == (long)h->net.nw, QCU_ERR_ARG, "weight count mismatch");
  std::vector<float> f(2 * n);
  QCU_HIP(hipMemcpy(f.data(), h->net.W.p, n * 8, hipMemcpyDeviceToHost));
  for (long k = 0; k < 2 * n; ++k) host[k] = f[k];
  return 0;
}

int qcuNetSetWeights(QcuNetHandle h, const double *host, long n) {
  clear_error();
  QCU_CHECK(h && host && n == (long)h->net.nw, QCU_ERR_ARG, "weight count mismatch");
  std::vector<float> f(2 * n);
  for (long k = 0; k < 2 * n; ++k) f[k] = (float)host[k];
  QCU_HIP(hipMemcpy(h->net.W.p, f.data(), n * 8, hipMemcpyHostToDevice));
  ++h->net.wver;
  return 0;
}

int qcuNetGetGradient(QcuNetHandle h, double *host, long n) {
  clear_error();
  QCU_CHECK(h && host && n == (long)h->net.nw, QCU_ERR_ARG, "weight count mismatch");
  std::vector<float> f(2 * n);
  QCU_HIP(hipMemcpy(f.data(), h->net.gW.p, n * 8, hipMemcpyDeviceToHost));
  for (long k = 0; k < 2 * n; ++k) host[k] = f[k];
  return 0;
}

// full-lattice forward: in [1][2][Vh][12], out [1][2][Vh][12] complex128 (device)
int qcuNetForward(QcuNetHandle h, void *out, const void *in) {
  clear_error();
  QCU_CHECK(h, QCU_ERR_ARG, "null handle");
  QCU_TRY(check_dev_ptr(out, "out"));
  QCU_TRY(check_dev_ptr(in, "in"));
  Net &net = h->net;
  QCU_TRY(net_ready(net));
  hipStream_t s = net.lat->stream;
  const int vh = net.lat->g.vh;
  const long n0 = (long)net.chan[0] * 2 * vh;
  hipLaunchKernelGGL(k_full_in, dim3((unsigned)ndiv(n0, 256)), dim3(256), 0, s, net.act_ptr(0),
                     (const double2 *)in, net.chan[0], vh);
  QCU_HIP(hipGetLastError());
  QCU_TRY(net.forward(nullptr, s, false, -1));
  const long nn = (long)net.chan[net.nl] * 2 * vh;
  hipLaunchKernelGGL(k_full_out, dim3((unsigned)ndiv(nn, 256)), dim3(256), 0, s, (double2 *)out,
                     net.act_ptr(net.nl), net.chan[net.nl], vh);
  QCU_HIP(hipGetLastError());
  QCU_HIP(hipStreamSynchronize(s));
  return 0;
}

// full-lattice backward for a given input and output gradient: fills the
// weight gradient (read with qcuNetGetGradient) and grad_in (may be null)
int qcuNetBackward(QcuNetHandle h, void *grad_in, const void *grad_out, const void *in) {
  clear_error();
  QCU_CHECK(h, QCU_ERR_ARG, "null handle");
  QCU_TRY(check_dev_ptr(grad_out, "grad_out"));
  QCU_TRY(check_dev_ptr(in, "in"));
  if (grad_in) QCU_TRY(check_dev_ptr(grad_in, "grad_in"));
  Net &net = h->net;
  QCU_TRY(net_ready(net));
  hipStream_t s = net.lat->stream;
  const int vh = net.lat->g.vh;
  const long n0 = (long)net.chan[0] * 2 * vh;
  hipLaunchKernelGGL(k_full_in, dim3((unsigned)ndiv(n0, 256)), dim3(256), 0, s, net.act_ptr(0),
                     (const double2 *)in, net.chan[0], vh);
  QCU_TRY(net.forward(nullptr, s, false, -1));
  SpinorPtr go;
  go.p = net.gbuf[1].p;
  go.norm = nullptr;
  go.stride = vh;
  const long nn = (long)net.chan[net.nl] * 2 * vh;
  hipLaunchKernelGGL(k_full_in, dim3((unsigned)ndiv(nn, 256)), dim3(256), 0, s, go, (const double2 *)grad_out,
                     net.chan[net.nl], vh);
  QCU_HIP(hipGetLastError());
  SpinorPtr gi;
  gi.p = net.tmp.p;  // room for one channel on both parities (6 odd vectors)
  gi.norm = nullptr;
  gi.stride = vh;
  QCU_TRY(net.backward(go, s, grad_in != nullptr, gi));
  if (grad_in) {
    hipLaunchKernelGGL(k_full_out, dim3((unsigned)ndiv(n0, 256)), dim3(256), 0, s, (double2 *)grad_in, gi,
                       net.chan[0], vh);
    QCU_HIP(hipGetLastError());
  }
  QCU_HIP(hipStreamSynchronize(s));
  return 0;
}

// Trains on odd-site sources srcs[k] (external one-parity complex128 fields on
// the device, e.g. bhat of the Schur system): `epochs` passes, Adam, mini-batches
// of `batch` sources.  losses[e] = mean loss of epoch e (before the updates of
// that epoch's batches are applied in turn).  seconds: wall time.
int qcuNetTrain(QcuNetHandle h, const void *const *srcs, int n, int epochs, int batch, double lr, double *losses,
                double *seconds) {
  clear_error();
  QCU_CHECK(h && srcs && n > 0 && epochs > 0 && batch > 0, QCU_ERR_ARG, "bad training arguments");
  for (int k = 0; k < n; ++k) QCU_TRY(check_dev_ptr(srcs[k], "srcs[k]"));
  Net &net = h->net;
  QCU_TRY(net_ready(net));
  Lattice &lat = *net.lat;
  hipStream_t s = lat.stream;
  const int vh = lat.g.vh, bs = 256;
  const unsigned nb = (unsigned)ndiv(vh, bs);
  NetTrainer tr{net, s};
  DevBuf gsum;
  QCU_TRY(gsum.alloc(net.nw * 8));
  auto t0 = std::chrono::steady_clock::now();
  const int nfl = (int)(2 * net.nw);
  for (int e = 0; e < epochs; ++e) {
    double esum = 0;
    for (int b0 = 0; b0 < n; b0 += batch) {
      const int b1 = std::min(n, b0 + batch);
      QCU_HIP(hipMemsetAsync(gsum.p, 0, net.nw * 8, s));
      for (int k = b0; k < b1; ++k) {
        hipLaunchKernelGGL(k_load_odd, dim3(nb), dim3(bs), 0, s, tr.odd(0), (const double2 *)srcs[k], vh);
        QCU_HIP(hipGetLastError());
        double loss = 0;
        QCU_TRY(tr.sample(&loss, true, nullptr));
        esum += loss;
        // gsum += gW
        hipLaunchKernelGGL(k_add_f32, dim3((unsigned)ndiv(nfl, 256)), dim3(256), 0, s, (float *)gsum.p,
                           (const float *)net.gW.p, nfl);
        QCU_HIP(hipGetLastError());
      }
      net.adam_t += 1;
      const float b1f = 0.9f, b2f = 0.999f;
      const float c1 = 1.f - powf(b1f, (float)net.adam_t), c2 = 1.f - powf(b2f, (float)net.adam_t);
      hipLaunchKernelGGL(k_adam, dim3((unsigned)ndiv(nfl, 256)), dim3(256), 0, s, (float *)net.W.p, (float *)net.m.p,
                         (float *)net.v.p, (const float *)gsum.p, nfl, (float)lr, b1f, b2f, c1, c2, 1e-8f,
                         1.f / (float)(b1 - b0));
      QCU_HIP(hipGetLastError());
      ++net.wver;
    }
    if (losses) losses[e] = esum / n;
  }
  QCU_HIP(hipStreamSynchronize(s));
  if (seconds) *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return 0;
}

int qcuNetApplyOdd(QcuNetHandle h, void *out, const void *in) {
  clear_error();
  QCU_CHECK(h, QCU_ERR_ARG, "null handle");
  QCU_TRY(check_dev_ptr(out, "out"));
  QCU_TRY(check_dev_ptr(in, "in"));
  Net &net = h->net;
  QCU_TRY(net_ready(net));
  hipStream_t s = net.lat->stream;
  const int vh = net.lat->g.vh, bs = 256;
  const unsigned nb = (unsigned)ndiv(vh, bs);
  NetTrainer tr{net, s};
  hipLaunchKernelGGL(k_load_odd, dim3(nb), dim3(bs), 0, s, tr.odd(0), (const double2 *)in, vh);
  QCU_HIP(hipGetLastError());
  QCU_TRY(net.load_input(tr.odd(0), P32, nullptr, s, !net.apply_reads_odd_input_only()));
  QCU_TRY(net.forward(nullptr, s, true, 0, false));
  SpinorPtr o1 = net.act_ptr(net.nl);
  o1.p = (char *)o1.p + (size_t)vh * 96;  // channel 0, parity 1
  hipLaunchKernelGGL(k_store_odd, dim3(nb), dim3(bs), 0, s, (double2 *)out, o1, vh);
  QCU_HIP(hipGetLastError());
  QCU_HIP(hipStreamSynchronize(s));
  return 0;
}

int qcuNetLossGrad(QcuNetHandle h, const void *src, double *loss) {
  clear_error();
  QCU_CHECK(h && loss, QCU_ERR_ARG, "null argument");
  QCU_TRY(check_dev_ptr(src, "src"));
  Net &net = h->net;
  QCU_TRY(net_ready(net));
  hipStream_t s = net.lat->stream;
  const int vh = net.lat->g.vh, bs = 256;
  NetTrainer tr{net, s};
  hipLaunchKernelGGL(k_load_odd, dim3((unsigned)ndiv(vh, bs)), dim3(bs), 0, s, tr.odd(0), (const double2 *)src, vh);
  QCU_HIP(hipGetLastError());
  QCU_TRY(tr.sample(loss, true, nullptr));
  QCU_HIP(hipStreamSynchronize(s));
  return 0;
}

// average ms per launch of layer `layer`'s forward kernel over both parities
// (on the current activations), timed with HIP events on the lattice stream;
// flops: algorithmic flops of one launch
int qcuNetBenchLayer(QcuNetHandle h, int layer, int iters, double *ms, double *flops) {
  clear_error();
  QCU_CHECK(h && ms && iters > 0, QCU_ERR_ARG, "bad arguments");
  Net &net = h->net;
  QCU_CHECK(layer >= 0 && layer < net.nl, QCU_ERR_ARG, "layer out of range");
  QCU_TRY(net_ready(net));
  hipStream_t s = net.lat->stream;
  const int ci = net.chan[layer], co = net.chan[layer + 1];
  // a full forward pass over both parities first (the preconditioner's pass
  // leaves some activations implicit): the layer then runs on complete
  // activations, the operands packed from the current weights and the input
  // maximum that pass recorded
  QCU_TRY(net.forward(nullptr, s, false, -1));
  const bool split = net.split_layer(layer, 0);
  if (split && net.packed_ver != net.wver) QCU_TRY(net.pack_split(0, s));
  SplitIO io = net.split_io(layer, 0);
  io.out_max = nullptr;
  auto launch = [&]() {
    return net.layer(net.act_ptr(layer), net.act_ptr(layer + 1), (const float2 *)net.W.p + net.woff[layer], ci, co, 0,
                     2, nullptr, s, -1, split ? &io : nullptr);
  };
  QCU_TRY(launch());
  hipEvent_t e0, e1;
  QCU_HIP(hipEventCreate(&e0));
  QCU_HIP(hipEventCreate(&e1));
  QCU_HIP(hipEventRecord(e0, s));
  for (int k = 0; k < iters; ++k) QCU_TRY(launch());
  QCU_HIP(hipEventRecord(e1, s));
  QCU_HIP(hipEventSynchronize(e1));
  float t = 0;
  QCU_HIP(hipEventElapsedTime(&t, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *ms = (double)t / iters;
  if (flops) {
    // per output site: 9 terms x cout x cin spin-matrix products on 3 colours
    // (16 complex multiply-adds each) + 8 hops x cin transports of 4 spins
    // (9 complex multiply-adds each).  A complex multiply-add is 8 real flops,
    // except in the matrix-core mixing, which uses three real products per
    // complex product (6 flops); the split-FP16 kernel is counted at 8 flops,
    // the FP32 arithmetic it reproduces (it executes three FP16 products of
    // 8 flops each).  The two-pass single-output layer transports one mixed
    // field per hop term.
    const bool two_pass = net.use_mfma && co == 1 && ci > 1;
    const bool mfma = net.use_mfma && (co > 1 || two_pass) && !split;
    const double mix = (mfma ? 6.0 : 8.0) * 9.0 * co * ci * 16 * 3;
    const double per_site = mix + 8.0 * (8.0 * (two_pass ? 1 : ci) * 4 * 9);
    *flops = per_site * 2.0 * net.lat->g.vh;
  }
  return 0;
}

// average ms of layer `layer`'s kernels as the preconditioner runs them: one
// application first (the same deterministic input as qcuNetBenchApply), then
// that layer alone on the activations and field maxima the application left
// (the last layer of a fused network: its gather pass), timed with HIP events
// on the lattice stream
int qcuNetBenchPrecondLayer(QcuNetHandle h, int layer, int iters, double *ms) {
  clear_error();
  QCU_CHECK(h && ms && iters > 0, QCU_ERR_ARG, "bad arguments");
  Net &net = h->net;
  QCU_TRY(net_ready(net));
  QCU_CHECK(layer >= 0 && layer < net.nl, QCU_ERR_ARG, "layer out of range");
  hipStream_t s = net.lat->stream;
  const int vh = net.lat->g.vh;
  NetTrainer tr{net, s};
  const size_t nf = (size_t)vh * 24;
  hipLaunchKernelGGL(k_fill_hash, dim3((unsigned)ndiv(nf, 256)), dim3(256), 0, s, (float *)tr.odd(0).p, nf);
  QCU_HIP(hipGetLastError());
  QCU_TRY(net.load_input(tr.odd(0), P32, nullptr, s, !net.apply_reads_odd_input_only()));
  bool wrote = false;
  const SpinorPtr o1 = tr.odd(1);
  QCU_TRY(net.forward(nullptr, s, true, 0, false, &o1, P32, &wrote));
  const bool fresh = net.first_split();  // the pass ran the path that needs a recorded input maximum
  auto launch = [&]() {
    net.input_fresh = fresh;
    return net.forward(nullptr, s, true, 0, false, &o1, P32, &wrote, layer);
  };
  QCU_TRY(launch());
  hipEvent_t e0, e1;
  QCU_HIP(hipEventCreate(&e0));
  QCU_HIP(hipEventCreate(&e1));
  QCU_HIP(hipEventRecord(e0, s));
  for (int r = 0; r < iters; ++r) QCU_TRY(launch());
  QCU_HIP(hipEventRecord(e1, s));
  QCU_HIP(hipEventSynchronize(e1));
  float t = 0;
  QCU_HIP(hipEventElapsedTime(&t, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *ms = (double)t / iters;
  return 0;
}

// average ms of one preconditioner application as the solver runs it (odd
// FP32 field embedded, the forward pass with only the output kept, odd part
// extracted), on internal buffers, timed with HIP events on the lattice stream
int qcuNetBenchApply(QcuNetHandle h, int iters, double *ms) {
  clear_error();
  QCU_CHECK(h && ms && iters > 0, QCU_ERR_ARG, "bad arguments");
  Net &net = h->net;
  QCU_TRY(net_ready(net));
  hipStream_t s = net.lat->stream;
  const int vh = net.lat->g.vh, bs = 256;
  const unsigned nb = (unsigned)ndiv(vh, bs);
  NetTrainer tr{net, s};
  // as in the solver: the fused last layer writes the odd output vector itself
  auto once = [&]() {
    QCU_TRY(net.load_input(tr.odd(0), P32, nullptr, s, !net.apply_reads_odd_input_only()));
    bool wrote = false;
    const SpinorPtr o1 = tr.odd(1);
    QCU_TRY(net.forward(nullptr, s, true, 0, false, &o1, P32, &wrote));
    if (!wrote) {
      hipLaunchKernelGGL(k_net_extract<P32>, dim3(nb), dim3(bs), 0, s, tr.odd(1), net.act_ptr(net.nl), vh, nullptr);
      QCU_HIP(hipGetLastError());
    }
    return 0;
  };
  // a deterministic input of unit size (all-zero operands would run the
  // matrix cores at an unrepresentative clock)
  const size_t nf = (size_t)vh * 24;
  hipLaunchKernelGGL(k_fill_hash, dim3((unsigned)ndiv(nf, 256)), dim3(256), 0, s, (float *)tr.odd(0).p, nf);
  QCU_HIP(hipGetLastError());
  QCU_TRY(once());
  hipEvent_t e0, e1;
  QCU_HIP(hipEventCreate(&e0));
  QCU_HIP(hipEventCreate(&e1));
  QCU_HIP(hipEventRecord(e0, s));
  for (int k = 0; k < iters; ++k) QCU_TRY(once());
  QCU_HIP(hipEventRecord(e1, s));
  QCU_HIP(hipEventSynchronize(e1));
  float t = 0;
  QCU_HIP(hipEventElapsedTime(&t, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *ms = (double)t / iters;
  return 0;
}

// Weights file: one text line of JSON metadata, then the complex weights as
// little-endian FP32 (re, im) pairs in the order of qcuNetGetWeights.
//   QCUNET1 {"format": "qcu-net", "version": 1, "layers": 4, "channels": [1, 8, 8, 8, 1],
//            "weights": 25344, "dtype": "complex64-le", ...}
// The lattice extents, kappa, c_sw and boundary are recorded for reference;
// the weights act site by site, so they load onto any lattice.
int qcuNetSave(QcuNetHandle h, const char *path) {
  clear_error();
  QCU_CHECK(h && path, QCU_ERR_ARG, "null argument");
  Net &net = h->net;
  std::vector<float> f(2 * net.nw);
  QCU_HIP(hipMemcpy(f.data(), net.W.p, net.nw * 8, hipMemcpyDeviceToHost));
  std::string ch;
  for (int k = 0; k <= net.nl; ++k) ch += (k ? ", " : "") + std::to_string(net.chan[k]);
  const Lattice &l = *net.lat;
  char meta[1024];
  snprintf(meta, sizeof(meta),
           "QCUNET1 {\"format\": \"qcu-net\", \"version\": 1, \"layers\": %d, \"channels\": [%s], "
           "\"weights\": %zu, \"dtype\": \"complex64-le\", "
           "\"weight_order\": \"layer, term (2 mu forward, 2 mu + 1 backward, 8 site), output channel, "
           "input channel, spin row, spin column\", \"lattice\": [%d, %d, %d, %d], \"kappa\": %.17g, "
           "\"csw\": %.17g, \"t_boundary\": %d, \"adam_steps\": %d}\n",
           net.nl, ch.c_str(), net.nw, l.L[0], l.L[1], l.L[2], l.L[3], l.kappa, l.csw, l.g.tbc, net.adam_t);
  FILE *fp = fopen(path, "wb");
  QCU_CHECK(fp, QCU_ERR_ARG, std::string("cannot write ") + path);
  const size_t n = strlen(meta);
  const bool ok = fwrite(meta, 1, n, fp) == n && fwrite(f.data(), 4, f.size(), fp) == f.size();
  const bool closed = fclose(fp) == 0;
  QCU_CHECK(ok && closed, QCU_ERR_ARG, std::string("write failed: ") + path);
  return 0;
}

int qcuNetLoad(QcuNetHandle h, const char *path) {
  clear_error();
  QCU_CHECK(h && path, QCU_ERR_ARG, "null argument");
  Net &net = h->net;
  FILE *fp = fopen(path, "rb");
  QCU_CHECK(fp, QCU_ERR_ARG, std::string("cannot read ") + path);
  std::string meta;
  int c;
  while ((c = fgetc(fp)) != EOF && c != '\n' && meta.size() < 4096) meta.push_back((char)c);
  auto fail = [&](const std::string &why) {
    fclose(fp);
    set_error(QCU_ERR_ARG, std::string(path) + ": " + why);
    return QCU_ERR_ARG;
  };
  if (c != '\n' || meta.compare(0, 8, "QCUNET1 ") != 0) return fail("not a qcu network file");
  // channels and weight count must match this network
  const size_t pc = meta.find("\"channels\": [");
  if (pc == std::string::npos) return fail("no channel list");
  std::vector<int> ch;
  const char *q = meta.c_str() + pc + 13;
  while (*q && *q != ']') {
    char *end = nullptr;
    const long v = strtol(q, &end, 10);
    if (end == q) return fail("bad channel list");
    ch.push_back((int)v);
    q = end;
    while (*q == ',' || *q == ' ') ++q;
  }
  bool same = (int)ch.size() == net.nl + 1;
  for (int k = 0; same && k <= net.nl; ++k) same = ch[k] == net.chan[k];
  if (!same) return fail("channel list differs from this network's");
  const size_t pw = meta.find("\"weights\": ");
  if (pw == std::string::npos || strtoull(meta.c_str() + pw + 11, nullptr, 10) != net.nw)
    return fail("weight count differs from this network's");
  std::vector<float> f(2 * net.nw);
  if (fread(f.data(), 4, f.size(), fp) != f.size()) return fail("file ends before the last weight");
  if (fgetc(fp) != EOF) return fail("trailing data after the weights");
  fclose(fp);
  for (float x : f)
    QCU_CHECK(std::isfinite(x), QCU_ERR_ARG, std::string(path) + ": non-finite weight");
  QCU_HIP(hipMemcpy(net.W.p, f.data(), net.nw * 8, hipMemcpyHostToDevice));
  ++net.wver;
  // a loaded network starts a fresh optimiser state
  QCU_HIP(hipMemset(net.m.p, 0, net.nw * 8));
  QCU_HIP(hipMemset(net.v.p, 0, net.nw * 8));
  net.adam_t = 0;
  return 0;
}

// mean loss over sources without training
int qcuNetLoss(QcuNetHandle h, const void *const *srcs, int n, double *loss) {
  clear_error();
  QCU_CHECK(h && srcs && n > 0 && loss, QCU_ERR_ARG, "bad arguments");
  for (int k = 0; k < n; ++k) QCU_TRY(check_dev_ptr(srcs[k], "srcs[k]"));
  Net &net = h->net;
  QCU_TRY(net_ready(net));
  hipStream_t s = net.lat->stream;
  const int vh = net.lat->g.vh, bs = 256;
  NetTrainer tr{net, s};
  double sum = 0;
  for (int k = 0; k < n; ++k) {
    hipLaunchKernelGGL(k_load_odd, dim3((unsigned)ndiv(vh, bs)), dim3(bs), 0, s, tr.odd(0), (const double2 *)srcs[k],
                       vh);
    QCU_HIP(hipGetLastError());
    double l = 0;
    QCU_TRY(tr.sample(&l, false, nullptr));
    sum += l;
  }
  *loss = sum / n;
  return 0;
}

}  // extern "C"
