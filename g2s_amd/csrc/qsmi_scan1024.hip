// Scan kernels on a 1024 x 1024 FFT grid for 2-D TIs of extent 513..1024
// (the 1000^2 benchmark workload), gfx950.
//
// Grid. The map of a node is needed only at valid positions p, where every
// p + d_j lies inside the TI. There the circular correlation on any grid of
// at least the TI's extent, with the TI zero-padded, equals the linear one;
// so these kernels correlate on 1024 x 1024 (a power of two) whatever the
// TI's own FFT size (1000 for 1000^2). Devices with a circular TI use the
// generic kernels on the TI's FFT grid.
//
// One wavefront = one length-1024 transform: lane l holds the 16 elements
// l + 64 m (m = 0..15). Stockham radix 16, 16, 4: the first stage reads the
// lane's own elements, the last leaves element l + 64 m in register m
// again, and the two exchanges in between go through a wavefront-private
// 8.5 KiB LDS buffer (one pad slot per 16 elements). No block barrier is
// needed after the twiddle table is loaded: every wavefront runs its own
// work items.
//
// Layouts:
//   YT [A][k0][k1]            TI spectra, k0 < 513, k1 < 1024 (k1 fastest)
//   CT [node][c1/2][k0][c1%2] the column-inverse intermediate in row pairs:
//                             the rows kernel reads one pair as contiguous
//                             16-byte loads
// Work items are handed out wavefront by wavefront, XCD-aware in the pencil
// kernel: the waves of XCD x work on column slice x (65 columns), so each
// XCD's L2 holds 1/8 of the TI spectra, and consecutive waves take
// consecutive columns of one node (their 16-byte CT writes fill common lines).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <algorithm>

#include "qsmi_kernels.hpp"
#include "qsmi_select.hpp"
#include "qsmi_taps.hpp"

namespace qsmi {
namespace w1k {

constexpr int F = 1024;
constexpr int H0 = 513;             // half spectrum along dim0
constexpr int NXCD = 8;
constexpr int BUF = 1088;           // padded exchange buffer (float2)
constexpr int GSLOT = 2 * 16 * kTapPass;  // tap sums of two sub-variables per slot (k, c)
// LDS float2 per pencil wavefront; = 2 mod 16 so that the wavefronts'
// buffers start 4 banks apart (the block's transposed store reads the same
// offset of every buffer)
constexpr int WSTRIDE = BUF + GSLOT + 2;
constexpr int NW = 16;              // rows wavefronts per block (one block per CU, 4 per SIMD)
constexpr int NWP = 12;             // pencil wavefronts per block = columns per group (3 per SIMD)
constexpr int NGRP = (H0 + NWP - 1) / NWP;  // 43 column groups
// Column groups of a node on XCD x: x + 8 i (i < GX_BASE), so the XCD's L2
// holds a fixed eighth of the TI spectra, and the last GX_EXTRA groups
// 8 GX_BASE + e on XCD (node + node / 8 + e) mod 8: over every 8 nodes
// n = 8 c .. 8 c + 7 each XCD takes GX_EXTRA of them, so all XCDs get the
// same number of groups.
constexpr int GX_BASE = NGRP / NXCD;   // 5
constexpr int GX_EXTRA = NGRP % NXCD;  // 3
constexpr int CT_PAIR = 520 * 2;    // float2 per row pair: 513 columns padded to 64-byte multiples
constexpr long long CT_STRIDE = (long long)(F / 2) * CT_PAIR;  // float2 per node
static_assert(WSTRIDE % 16 == 2, "buffer bank offset");
static_assert((NWP * 16) % 64 == 0 && (CT_PAIR * 8) % 64 == 0, "64-byte aligned group stores");

__device__ __forceinline__ int pidx(int e) { return e + (e >> 4); }

// orders this wavefront's LDS accesses across lanes (no hardware barrier)
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename V2>
struct Cx;
template <>
struct Cx<float2> {
  using S = float;
  static __device__ __forceinline__ float2 mk(float x, float y) { return make_float2(x, y); }
};
template <>
struct Cx<double2> {
  using S = double;
  static __device__ __forceinline__ double2 mk(double x, double y) { return make_double2(x, y); }
};

template <typename V2>
__device__ __forceinline__ V2 add(V2 a, V2 b) {
  return Cx<V2>::mk(a.x + b.x, a.y + b.y);
}
template <typename V2>
__device__ __forceinline__ V2 sub(V2 a, V2 b) {
  return Cx<V2>::mk(a.x - b.x, a.y - b.y);
}
template <typename V2>
__device__ __forceinline__ V2 mul(V2 a, V2 b) {
  return Cx<V2>::mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * (DIR i)
template <int DIR, typename V2>
__device__ __forceinline__ V2 mul_i(V2 a) {
  return (DIR < 0) ? Cx<V2>::mk(a.y, -a.x) : Cx<V2>::mk(-a.y, a.x);
}
// table entry w = exp(-2 pi i t / F) applied in direction DIR
template <int DIR, typename V2>
__device__ __forceinline__ V2 tw_dir(V2 w) {
  return (DIR < 0) ? w : Cx<V2>::mk(w.x, -w.y);
}

template <int DIR, typename V2>
__device__ __forceinline__ void dft4(V2& a, V2& b, V2& c, V2& d) {
  const V2 s02 = add(a, c), d02 = sub(a, c);
  const V2 s13 = add(b, d), d13 = mul_i<DIR>(sub(b, d));
  a = add(s02, s13);
  c = sub(s02, s13);
  b = add(d02, d13);
  d = sub(d02, d13);
}

// 16-point DFT in direction DIR (exp(DIR 2 pi i n k / 16)), 4 x 4:
// n = 4 n1 + n2, k = k1 + 4 k2
template <int DIR, typename V2>
__device__ __forceinline__ void dft16(V2 (&u)[16]) {
  using S = typename Cx<V2>::S;
  const S c1 = (S)0.92387953251128675613, s1 = (S)0.38268343236508977173;  // cos, sin(2 pi / 16)
  const S r2 = (S)0.70710678118654752440;
#pragma unroll
  for (int n2 = 0; n2 < 4; ++n2) dft4<DIR>(u[n2], u[4 + n2], u[8 + n2], u[12 + n2]);
  // u[4 k1 + n2] now holds a_{n2}(k1); twiddle by W16^(n2 k1)
  auto rot = [](V2 a, S c, S s) { return Cx<V2>::mk(a.x * c - a.y * s, a.x * s + a.y * c); };
  const S d = (S)DIR;
  u[4 + 1] = rot(u[4 + 1], c1, d * s1);     // n2 1, k1 1: W^1
  u[8 + 1] = rot(u[8 + 1], r2, d * r2);     // n2 1, k1 2: W^2
  u[12 + 1] = rot(u[12 + 1], s1, d * c1);   // n2 1, k1 3: W^3
  u[4 + 2] = rot(u[4 + 2], r2, d * r2);     // n2 2, k1 1: W^2
  u[8 + 2] = mul_i<DIR>(u[8 + 2]);          // n2 2, k1 2: W^4
  u[12 + 2] = rot(u[12 + 2], -r2, d * r2);  // n2 2, k1 3: W^6
  u[4 + 3] = rot(u[4 + 3], s1, d * c1);     // n2 3, k1 1: W^3
  u[8 + 3] = rot(u[8 + 3], -r2, d * r2);    // n2 3, k1 2: W^6
  u[12 + 3] = rot(u[12 + 3], -c1, -d * s1); // n2 3, k1 3: W^9
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) dft4<DIR>(u[4 * k1 + 0], u[4 * k1 + 1], u[4 * k1 + 2], u[4 * k1 + 3]);
  // u[4 k1 + k2] holds X(k1 + 4 k2): transpose to u[k1 + 4 k2]
  V2 t[16];
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1)
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) t[k1 + 4 * k2] = u[4 * k1 + k2];
#pragma unroll
  for (int i = 0; i < 16; ++i) u[i] = t[i];
}

// Length-1024 transform of one wavefront in direction DIR. On entry and exit
// u[m] = element l + 64 m. buf: this wavefront's BUF-element LDS buffer; tw:
// exp(-2 pi i t / 1024), t < 1024.
// The padded slot of element e is pidx(e) = e + (e >> 4); the per-stage
// bases below are pidx's value with the lane's part factored out, so every
// access is a base plus a constant.
template <int DIR, typename V2>
__device__ __forceinline__ void fft1024(V2 (&u)[16], V2* buf, const V2* tw, int l) {
  const int hi = l >> 4, k = l & 15;
  // stage 1: radix 16, Ns = 1 (butterfly l reads l + 64 q), no twiddles;
  // writes 16 l + q: slot 17 l + q
  dft16<DIR>(u);
  wave_sync();  // earlier readers of buf are done
  V2* b1 = buf + 17 * l;
#pragma unroll
  for (int q = 0; q < 16; ++q) b1[q] = u[q];
  wave_sync();
  // stage 2: radix 16, Ns = 16: reads l + 64 q (slot l + hi + 68 q), writes
  // 256 hi + k + 16 q (slot 272 hi + k + 17 q)
  const V2* b2 = buf + l + hi;
#pragma unroll
  for (int q = 0; q < 16; ++q) u[q] = b2[68 * q];
  const V2* t2 = tw + 4 * k;
#pragma unroll
  for (int q = 1; q < 16; ++q) u[q] = mul(u[q], tw_dir<DIR>(t2[(q - 1) * 4 * k]));
  dft16<DIR>(u);
  wave_sync();
  V2* b3 = buf + 272 * hi + k;
#pragma unroll
  for (int q = 0; q < 16; ++q) b3[17 * q] = u[q];
  wave_sync();
  // stage 3: radix 4, Ns = 256: butterflies j = l + 64 b read j + 256 q
  // (slot l + hi + 68 b + 272 q) and write element j + 256 q = l + 64 (b + 4 q),
  // kept in register b + 4 q
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    const int j = l + 64 * b;
    const V2* b4 = buf + l + hi + 68 * b;
    V2 v0 = b4[0], v1 = b4[272], v2 = b4[544], v3 = b4[816];
    v1 = mul(v1, tw_dir<DIR>(tw[j]));
    v2 = mul(v2, tw_dir<DIR>(tw[2 * j]));
    v3 = mul(v3, tw_dir<DIR>(tw[3 * j]));
    dft4<DIR>(v0, v1, v2, v3);
    u[b] = v0;
    u[b + 4] = v1;
    u[b + 8] = v2;
    u[b + 12] = v3;
  }
}

// block-shared twiddle table, then every wave proceeds on its own
template <typename V2>
__device__ __forceinline__ void load_twiddles(V2* twl, const V2* tw) {
  for (int i = threadIdx.x; i < F; i += blockDim.x) twl[i] = tw[i];
  __syncthreads();
}

}  // namespace w1k

using namespace w1k;

// ---------------------------------------------------------------------------
// Tap lists of each node (build_taps1024), one wavefront per node.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(64) k_taps1024(const NodeInfo* __restrict__ info, const int* __restrict__ nbOff,
                                                const float* __restrict__ nbVal, int V, TapInfo1024* __restrict__ taps,
                                                uint16_t* __restrict__ tapD0, float* __restrict__ tapVal,
                                                const int* __restrict__ nodeCount) {
  const int node = blockIdx.x;
  if (nodeCount && node >= *nodeCount) return;
  const NodeInfo ni = info[node];
  build_taps1024(node, ni.nb, ni.nbStart, nbOff, nbVal, V, taps, tapD0, tapVal);
}

// ---------------------------------------------------------------------------
// Pencil along dim1: one wavefront per (node, column k0 < 513). The forward
// transform of the node's sparse kernel column, X(k1) = sum_j v_j
// exp(2 pi i (k0 d0_j + k1 d1_j) / 1024), in one of three ways:
//  * T <= kTapMax distinct rows r_t: with k1 = l + 64 m,
//      X(l + 64 m) = sum_c exp(-2 pi i m c / 16) S_l(c),
//      S_l(c) = sum_{t: r_t mod 16 = c} G(t) exp(-2 pi i l r_t / 1024),
//    G(t) = sum over the row's neighbours of v exp(+2 pi i k0 d0 / 1024) in
//    list order: T twiddled MACs per lane and one 16-point DFT, and X arrives
//    in the registers (elements l + 64 m) the MAC and the inverse use;
//  * more rows (T <= kTapNbMax): the row sums into a zeroed LDS column and a
//    full forward transform;
//  * more than kTapNbMax neighbours (no lists): lane l owns rows l + 64 i
//    and adds their neighbours in list order.
// Then M = sum_a Y_a X_a and the inverse along dim1 into CT (rows lo1..hi1).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(64 * NWP) k_pencil1024(
    ScanGeom g, const float2* __restrict__ YT, const float2* __restrict__ tw, const NodeInfo* __restrict__ info,
    const int* __restrict__ nbOff, const float* __restrict__ nbVal, const TapInfo1024* __restrict__ taps,
    const uint16_t* __restrict__ tapD0, const float* __restrict__ tapVal, float2* __restrict__ CT, int nNodes,
    const int* __restrict__ count) {
  __shared__ float2 twl[F];
  __shared__ float2 wbuf[NWP][WSTRIDE];
  // the node's tap twiddles W(t, l), shared by the block's columns, at slot
  // (k, c) of the k-th tap of class c: row (16 k + c) of 64 lanes
  __shared__ float2 wtab[kTapPass * 16 * 64];
  // the first pass reads every class's slot: a class without taps multiplies
  // its zero tap sum by whatever its slot holds, which must be finite (LDS
  // is not cleared between kernels; later items leave earlier twiddles)
  for (int i = threadIdx.x; i < 16 * 64; i += blockDim.x) wtab[i] = make_float2(0.f, 0.f);
  load_twiddles(twl, tw);  // ends with a barrier
  const int l = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform (scalar) from here on
  float2* buf = wbuf[w];
  float2* gbuf = wbuf[w] + BUF;
  // a block takes whole nodes (node bX, bX + blocksX, ... of its XCD): the
  // node's lists and tap twiddles are loaded and built once for all of the
  // XCD's column groups of the node
  const int xcd = blockIdx.x % NXCD;
  const int blocksX = gridDim.x / NXCD, bX = blockIdx.x / NXCD;
  const int nAct = count ? min(*count, nNodes) : nNodes;
  const int V = g.V;
  for (int node = bX; node < nAct; node += blocksX) {  // block-uniform
    // the lane index, opaque to the compiler per node and per group:
    // lane-dependent addresses are recomputed instead of held (spilled)
    // across the loops
    int lx = l;
    asm volatile("" : "+v"(lx));
    lx &= 63;  // the range is known again
    const TapInfo1024* tp = taps + node;
    const uint16_t* td = tapD0 + (long long)node * kTapNbMax;
    const float* tv = tapVal + (long long)node * kTapNbMax * V;
    // one round of loads for the node: every address below is known without
    // the others' values (lists are read in full up to 128 slots and 64
    // taps; unused entries are ignored)
    const NodeInfo ni = info[node];
    const int T = tp->T;        // block-uniform
    const int kmax = tp->passes;
    const int myRow = tp->row[lx];
    const int mySlot = tp->slot[lx < kTapMax ? lx : 0];
    const int myS0 = tp->start[lx], myS1 = tp->start[lx + 1];
    const uint32_t sdA = td[lx], sdB = td[lx + 64];
    const float svA0 = tv[lx * V], svA1 = tv[lx * V + (V > 1 ? 1 : 0)];
    const float svB0 = tv[(lx + 64) * V], svB1 = tv[(lx + 64) * V + (V > 1 ? 1 : 0)];
    if (ni.lo[1] > ni.hi[1] || ni.lo[0] > ni.hi[0]) continue;  // block-uniform
    const bool tapPath = kmax > 0;  // block-uniform
    if (tapPath) {
      // W(t, l) = exp(-2 pi i l r_t / 1024), the same for every column of the
      // node: wave w builds the slots of taps w, w + NWP, ... (the previous
      // node's last store phase ended with a barrier)
      for (int t = w; t < T; t += NWP) {
        const int r = __builtin_amdgcn_readlane(myRow, t), sl = __builtin_amdgcn_readlane(mySlot, t);
        wtab[sl * 64 + lx] = twl[__umul24((uint32_t)lx, (uint32_t)r) & (F - 1)];
      }
      __syncthreads();
    }
    const int lxn = lx;
    for (int gi = 0; gi < GX_BASE + GX_EXTRA; ++gi) {  // block-uniform
    int cg = xcd + NXCD * gi;
    if (gi >= GX_BASE) {  // one of the node's last groups: on this XCD when its turn
      const int e = gi - GX_BASE;
      if ((node + node / NXCD + e) % NXCD != xcd) continue;
      cg = NXCD * GX_BASE + e;
    }
    int lx = lxn;
    asm volatile("" : "+v"(lx));
    lx &= 63;
    const int k0 = cg * NWP + w;
    const bool colOk = k0 < H0;  // wave-uniform; idle waves still take the barriers
    if (colOk) {
      float2 M[16];
#pragma unroll
      for (int m = 0; m < 16; ++m) M[m] = make_float2(0.f, 0.f);
      const int nb = ni.nb;
      if (tapPath) {
        const float2* wl = wtab + lx;
        for (int a0 = 0; a0 < V; a0 += 2) {
          const bool two = a0 + 1 < V;
          // every slot's contribution v exp(+2 pi i k0 d0 / 1024) into buf ...
          wave_sync();
          for (int s = lx; s < nb; s += 64) {
            uint32_t d0m;
            float v0, v1;
            if (a0 == 0 && s < 128) {  // the prefetched slots
              const bool A = s < 64;
              d0m = A ? sdA : sdB;
              v0 = A ? svA0 : svB0;
              v1 = A ? svA1 : svB1;
            } else {
              d0m = td[s];
              v0 = tv[(long long)s * V + a0];
              v1 = tv[(long long)s * V + a0 + (two ? 1 : 0)];
            }
            if (!two) v1 = 0.f;
            const float2 wv = twl[__umul24((uint32_t)k0, d0m) & (F - 1)];
            buf[2 * s] = make_float2(v0 * wv.x, -v0 * wv.y);
            buf[2 * s + 1] = make_float2(v1 * wv.x, -v1 * wv.y);
          }
          // the first pass's slots of classes without taps add zero
          if (lx < 32) gbuf[lx] = make_float2(0.f, 0.f);
          wave_sync();
          // ... then lane t sums tap t's slots in list order
          if (lx < T) {
            float2 g0 = make_float2(0.f, 0.f), g1 = make_float2(0.f, 0.f);
  #pragma unroll 4
            for (int s = myS0; s < myS1; ++s) {
              const float2 c0 = buf[2 * s], c1 = buf[2 * s + 1];
              g0.x += c0.x;
              g0.y += c0.y;
              g1.x += c1.x;
              g1.y += c1.y;
            }
            gbuf[2 * mySlot] = g0;
            gbuf[2 * mySlot + 1] = g1;
          }
          wave_sync();
          // both sub-variables' class sums share each twiddle read; the
          // first TI spectrum column is loaded before the tap MACs
          const float2* Y0 = YT + ((long long)g.A[a0] * H0 + k0) * F + lx;
          float2 y[16];
  #pragma unroll
          for (int m = 0; m < 16; ++m) y[m] = Y0[64 * m];
          float2 S0[16], S1[16];
  #pragma unroll
          for (int c = 0; c < 16; ++c) {
            S0[c] = make_float2(0.f, 0.f);
            S1[c] = make_float2(0.f, 0.f);
          }
          const float4* G4 = reinterpret_cast<const float4*>(gbuf);  // slot (k, c): (G0, G1)
          int cn[16];  // taps per class (uniform)
  #pragma unroll
          for (int c = 0; c < 16; ++c) cn[c] = tp->cls[c + 1] - tp->cls[c];
          // pass k adds the k-th tap of every class: classes are independent
          // accumulators, so a pass's twiddle and tap-sum reads are issued
          // together (constant offsets from one base); each class still adds
          // its taps in order
          for (int k = 0; k < kmax; ++k) {
            const float2* wk = wl + k * 16 * 64;
            const float4* gk = G4 + k * 16;
  #pragma unroll
            for (int c = 0; c < 16; ++c) {
              // pass 0 takes every class (branch-free, its reads issued
              // together; empty slots hold G = 0 and an unused finite W)
              if (k == 0 || k < cn[c]) {  // uniform
                const float2 wv = wk[c * 64];
                const float4 gg = gk[c];
                S0[c].x = fmaf(gg.x, wv.x, S0[c].x);
                S0[c].x = fmaf(-gg.y, wv.y, S0[c].x);
                S0[c].y = fmaf(gg.x, wv.y, S0[c].y);
                S0[c].y = fmaf(gg.y, wv.x, S0[c].y);
                S1[c].x = fmaf(gg.z, wv.x, S1[c].x);
                S1[c].x = fmaf(-gg.w, wv.y, S1[c].x);
                S1[c].y = fmaf(gg.z, wv.y, S1[c].y);
                S1[c].y = fmaf(gg.w, wv.x, S1[c].y);
              }
            }
          }
          dft16<-1>(S0);
  #pragma unroll
          for (int m = 0; m < 16; ++m) {
            M[m].x = fmaf(y[m].x, S0[m].x, M[m].x);
            M[m].x = fmaf(-y[m].y, S0[m].y, M[m].x);
            M[m].y = fmaf(y[m].x, S0[m].y, M[m].y);
            M[m].y = fmaf(y[m].y, S0[m].x, M[m].y);
          }
          if (two) {
            const float2* Y1 = YT + ((long long)g.A[a0 + 1] * H0 + k0) * F + lx;
  #pragma unroll
            for (int m = 0; m < 16; ++m) y[m] = Y1[64 * m];
            dft16<-1>(S1);
  #pragma unroll
            for (int m = 0; m < 16; ++m) {
              M[m].x = fmaf(y[m].x, S1[m].x, M[m].x);
              M[m].x = fmaf(-y[m].y, S1[m].y, M[m].x);
              M[m].y = fmaf(y[m].x, S1[m].y, M[m].y);
              M[m].y = fmaf(y[m].y, S1[m].x, M[m].y);
            }
          }
        }
      } else if (T > 0) {
        // more rows (or more than kTapPass in a class): per sub-variable, the
        // row sums scattered into a zeroed column, then a full forward transform
        for (int a = 0; a < V; ++a) {
          wave_sync();
          for (int s = lx; s < nb; s += 64) {
            const float2 wv = twl[__umul24((uint32_t)k0, (uint32_t)td[s]) & (F - 1)];
            const float v = tv[(long long)s * V + a];
            buf[s] = make_float2(v * wv.x, -v * wv.y);
          }
          wave_sync();
          float2 gs[8];  // taps lx + 64 i (T <= kTapNbMax < 512)
  #pragma unroll
          for (int i = 0; i < 8; ++i) {
            gs[i] = make_float2(0.f, 0.f);
            const int t = lx + 64 * i;
            if (t < T) {
              const int e1 = tp->start[t + 1];
              for (int s = tp->start[t]; s < e1; ++s) {
                const float2 c = buf[s];
                gs[i].x += c.x;
                gs[i].y += c.y;
              }
            }
          }
          wave_sync();
  #pragma unroll
          for (int m = 0; m < 16; ++m) buf[pidx(lx + 64 * m)] = make_float2(0.f, 0.f);
          wave_sync();
  #pragma unroll
          for (int i = 0; i < 8; ++i) {
            const int t = lx + 64 * i;
            if (t < T) buf[pidx(tp->row[t])] = gs[i];
          }
          wave_sync();
          float2 u[16];
  #pragma unroll
          for (int m = 0; m < 16; ++m) u[m] = buf[pidx(lx + 64 * m)];
          fft1024<-1>(u, buf, twl, lx);
          const float2* Y = YT + ((long long)g.A[a] * H0 + k0) * F + lx;
  #pragma unroll
          for (int m = 0; m < 16; ++m) {
            const float2 y = Y[64 * m];
            M[m].x = fmaf(y.x, u[m].x, M[m].x);
            M[m].x = fmaf(-y.y, u[m].y, M[m].x);
            M[m].y = fmaf(y.x, u[m].y, M[m].y);
            M[m].y = fmaf(y.y, u[m].x, M[m].y);
          }
        }
      } else if (nb > 0) {
        // no lists: lane lx owns rows lx + 64 i and adds their neighbours in
        // list order (the wave walks the list; loads are wave-uniform)
        for (int a = 0; a < V; ++a) {
          float2 u[16];
  #pragma unroll
          for (int m = 0; m < 16; ++m) u[m] = make_float2(0.f, 0.f);
          for (int s = 0; s < nb; ++s) {
            const long long jj = ni.nbStart + s;
            const int row = (-nbOff[jj * 2 + 1]) & (F - 1);
            if ((row & 63) == lx) {
              const float2 wv = twl[(k0 * (nbOff[jj * 2 + 0] & (F - 1))) & (F - 1)];
              const float v = nbVal[jj * V + a];
              const int i = row >> 6;
  #pragma unroll
              for (int m = 0; m < 16; ++m)
                if (m == i) {
                  u[m].x = fmaf(v, wv.x, u[m].x);
                  u[m].y = fmaf(-v, wv.y, u[m].y);
                }
            }
          }
          fft1024<-1>(u, buf, twl, lx);
          const float2* Y = YT + ((long long)g.A[a] * H0 + k0) * F + lx;
  #pragma unroll
          for (int m = 0; m < 16; ++m) {
            const float2 y = Y[64 * m];
            M[m].x = fmaf(y.x, u[m].x, M[m].x);
            M[m].x = fmaf(-y.y, u[m].y, M[m].x);
            M[m].y = fmaf(y.x, u[m].y, M[m].y);
            M[m].y = fmaf(y.y, u[m].x, M[m].y);
          }
        }
      }
      // inverse along dim1, input already in registers; the column into
      // the first 1024 slots of this wavefront's buffer
      fft1024<+1>(M, buf, twl, lx);
      wave_sync();
#pragma unroll
      for (int m = 0; m < 16; ++m) buf[lx + 64 * m] = M[m];
    }
    // transposed store: per row pair, the group's columns as one contiguous
    // 16 x (columns) byte run of CT (lane -> pair offset lx / NWP, column
    // lx % NWP: five pairs per store instruction)
    __syncthreads();
    {
      const int ncol = min(NWP, H0 - cg * NWP);
      const int pLo = ni.lo[1] >> 1, pHi = ni.hi[1] >> 1;
      const int po = lx / NWP, c = lx - NWP * po;
      float4* Cn = reinterpret_cast<float4*>(CT + (long long)node * CT_STRIDE) + cg * NWP + c;
      const float4* src = reinterpret_cast<const float4*>(wbuf[c < ncol ? c : 0]);
      for (int p0 = pLo + 5 * w; p0 <= pHi; p0 += 5 * NWP) {
        const int p = p0 + po;
        if (po < 5 && c < ncol && p <= pHi) Cn[(long long)p * (CT_PAIR / 2)] = src[p];
      }
    }
    __syncthreads();
    }  // column groups
  }
}

namespace w1k {

// Row pair p of one node's CT: Hermitian completion of the half spectrum
// (columns k > 512 are conj(Z[1024 - k]); the imaginary parts of columns 0
// and 512 do not enter a c2r), then one two-for-one c2r along dim0. On exit
// u[m] holds position c0 = l + 64 m of row 2p (x) and row 2p+1 (y).
// column k = l + 64 m: m < 8 direct; m = 8: the Nyquist column (l = 0)
// or conj(Z[512 - l]); m > 8: conj(Z[1024 - k])
template <bool BOTH>
__device__ __forceinline__ void rows_pair_load(const float4* __restrict__ rowp, bool va, int l, float2 (&u)[16]) {
  float4 ab[16];
#pragma unroll
  for (int m = 0; m < 8; ++m) ab[m] = rowp[l + 64 * m];
  ab[8] = rowp[l == 0 ? 512 : 512 - l];
#pragma unroll
  for (int m = 9; m < 16; ++m) ab[m] = rowp[F - 64 * m - l];
#pragma unroll
  for (int m = 0; m < 16; ++m) {
    float2 A = make_float2(ab[m].x, ab[m].y), B = make_float2(ab[m].z, ab[m].w);
    if (!BOTH) {  // one valid row: the other (never written for this node) is zero
      if (va)
        B = make_float2(0.f, 0.f);
      else
        A = make_float2(0.f, 0.f);
    }
    if ((m == 0 || m == 8) && l == 0) {  // columns 0 and 512: real
      A.y = 0.f;
      B.y = 0.f;
    }
    const bool cj = m > 8 || (m == 8 && l > 0);
    // z = A + i B with A, B conjugated where cj
    u[m] = cj ? make_float2(A.x + B.y, B.x - A.y) : make_float2(A.x - B.y, A.y + B.x);
  }
}

__device__ __forceinline__ void rows_pair(const float2* __restrict__ Cn, int p, bool va, bool vb, float2* buf,
                                          const float2* twl, int l, float2 (&u)[16]) {
  const float4* rowp = reinterpret_cast<const float4*>(Cn + (long long)p * CT_PAIR);
  if (va && vb)  // wave-uniform
    rows_pair_load<true>(rowp, va, l, u);
  else
    rows_pair_load<false>(rowp, va, l, u);
  fft1024<+1>(u, buf, twl, l);
}

}  // namespace w1k

// ---------------------------------------------------------------------------
// Rows along dim0: one wavefront per (node, row pair p < 512): the pair's
// two-for-one c2r, then scale + delta0, the valid-position mask and the
// wavefront's K best by the device rule -> cand[node][p][K] (-inf
// candidates for pairs without a valid row), and with the exact selection
// the number of its values within 2 eps of its K-th best -> gcnt[node][p].
// ---------------------------------------------------------------------------
template <int KT, bool QUANT>
__global__ void __launch_bounds__(64 * NW) k_rows1024(ScanGeom g, const float2* __restrict__ tw,
                                                     const NodeInfo* __restrict__ info, const float2* __restrict__ CT,
                                                     int nNodes, int K, Cand* __restrict__ cand,
                                                     float* __restrict__ map, const int* __restrict__ count,
                                                     int* __restrict__ gcnt) {
  __shared__ float2 twl[F];
  __shared__ float2 wbuf[NW][BUF];
  load_twiddles(twl, tw);
  const int l = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  float2* buf = wbuf[w];
  const int waves = gridDim.x * NW;
  const int nAct = count ? min(*count, nNodes) : nNodes;
  const long long items = (long long)nAct * (F / 2);
  const int s0 = g.s[0];
  for (long long it = (long long)blockIdx.x * NW + w; it < items; it += waves) {
    const int node = (int)(it / (F / 2));
    const int p = (int)(it - (long long)node * (F / 2));
    const NodeInfo ni = info[node];
    Cand* out = cand ? cand + ((long long)node * (F / 2) + p) * K : nullptr;
    const int ca = 2 * p;
    const bool va = ca >= ni.lo[1] && ca <= ni.hi[1];
    const bool vb = ca + 1 >= ni.lo[1] && ca + 1 <= ni.hi[1];
    if (!(va || vb) || ni.lo[0] > ni.hi[0]) {  // wave-uniform
      if (out && l < K) out[l] = worst_cand();
      if (gcnt && l == 0) gcnt[(long long)node * (F / 2) + p] = 0;
      continue;
    }
    int lx = l;  // opaque per item (see k_pencil1024)
    asm volatile("" : "+v"(lx));
    lx &= 63;
    float2 u[16];
    rows_pair(CT + (long long)node * CT_STRIDE, p, va, vb, buf, twl, lx, u);
    const float scale = ni.scale, d0 = ni.delta0, quantum = ni.quantum;
    const int lo0 = ni.lo[0], hi0 = ni.hi[0];
    if (map) {
      float* mp = map + (long long)node * g.s[0] * g.s[1];
#pragma unroll
      for (int m = 0; m < 16; ++m) {
        const int c0 = lx + 64 * m;
        if (c0 < lo0 || c0 > hi0) continue;
        if (va) mp[(uint32_t)(c0 + s0 * ca)] = u[m].x * scale + d0;
        if (vb) mp[(uint32_t)(c0 + s0 * (ca + 1))] = u[m].y * scale + d0;
      }
    }
    if (!out) continue;
    // keys of this lane's 32 positions, -inf where invalid (QUANT: the
    // device's nodes carry a tie quantum)
    float key[32];
    const float invq = QUANT ? 1.f / quantum : 0.f;
#pragma unroll
    for (int m = 0; m < 16; ++m) {
      const int c0 = lx + 64 * m;
      const bool okc = c0 >= lo0 && c0 <= hi0;
      float ka = u[m].x * scale + d0, kb = u[m].y * scale + d0;
      if (QUANT) {
        ka = rintf(ka * invq);
        kb = rintf(kb * invq);
      }
      key[2 * m] = (okc && va) ? ka : -INFINITY;
      key[2 * m + 1] = (okc && vb) ? kb : -INFINITY;
    }
    if (ni.exOn) {  // node-uniform: verbatim exclusion disk
#pragma unroll
      for (int m = 0; m < 16; ++m) {
        const int c0 = lx + 64 * m;
        if (excluded(ni, s0, g.s[1], c0, ca, 0)) key[2 * m] = -INFINITY;
        if (excluded(ni, s0, g.s[1], c0, ca + 1, 0)) key[2 * m + 1] = -INFINITY;
      }
    }
    // pass 1: tau = the K-th largest of the lanes' maxima. The K lanes
    // holding them give K keys >= tau, so tau <= the wavefront's K-th best
    // key and every key of the K best is >= tau (also excludes the -inf of
    // invalid positions).
    float mk[1] = {key[0]};
#pragma unroll
    for (int e = 1; e < 31; e += 2) mk[0] = fmaxf(fmaxf(mk[0], key[e]), key[e + 1]);
    mk[0] = fmaxf(mk[0], key[31]);
    const float tau2 = fmaxf(wave_kth_key<1>(mk, K), -3.402823466e38f);
    if (gcnt && ni.eps > 0.f && !QUANT) {
      // exact selection: values within 2 eps of tau (at least those within
      // 2 eps of the wavefront's K-th best)
      const float th = (tau2 <= -3.402823466e38f) ? -INFINITY : tau2 - 2.f * ni.eps * 1.0009765625f;
      int c = 0;
#pragma unroll
      for (int e = 0; e < 32; ++e) c += __popcll(__ballot(key[e] >= th));
      if (lx == 0) gcnt[(long long)node * (F / 2) + p] = c;
    }
    // pass 2: exact device rule among the keys >= tau (K and a few more)
    RankList<KT> rl;
    rl.init();
#pragma unroll
    for (int e = 0; e < 32; ++e) {
      if (key[e] >= tau2) {
        const int m = e >> 1, wr = e & 1;
        const uint32_t pos = (uint32_t)(lx + 64 * m + s0 * (ca + wr));
        rl.insert(key[e], tie_hash(ni.seedbits, pos), pos, (wr ? u[m].y : u[m].x) * scale + d0);
      }
    }
    wave_extract<KT>(rl, K, out);
  }
}

// ---------------------------------------------------------------------------
// Exact selection, second pass: for every queued (node, row pair) whose
// candidates may hide more positions within the node's band, the pair again
// (same transform and mask as k_rows1024), appending every valid position
// whose FP32 value is >= theta[node] to the node's band list.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(64 * NW) k_rows1024_collect(
    ScanGeom g, const float2* __restrict__ tw, const NodeInfo* __restrict__ info, const float2* __restrict__ CT,
    const int2* __restrict__ queue, const int* __restrict__ qcnt, const float* __restrict__ theta,
    uint32_t* __restrict__ bandPos, int* __restrict__ bandCnt) {
  __shared__ float2 twl[F];
  __shared__ float2 wbuf[NW][BUF];
  const int items = *qcnt;
  if ((int)blockIdx.x * NW >= items) return;  // block-uniform
  load_twiddles(twl, tw);
  const int l = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  float2* buf = wbuf[w];
  const int s0 = g.s[0];
  for (int it = blockIdx.x * NW + w; it < items; it += gridDim.x * NW) {
    const int2 qi = queue[it];
    const int node = qi.x, p = qi.y;
    const NodeInfo ni = info[node];
    const int ca = 2 * p;
    const bool va = ca >= ni.lo[1] && ca <= ni.hi[1];
    const bool vb = ca + 1 >= ni.lo[1] && ca + 1 <= ni.hi[1];
    if (!(va || vb)) continue;
    int lx = l;  // opaque per item (see k_pencil1024)
    asm volatile("" : "+v"(lx));
    lx &= 63;
    float2 u[16];
    rows_pair(CT + (long long)node * CT_STRIDE, p, va, vb, buf, twl, lx, u);
    const float th = theta[node];
    const float scale = ni.scale, d0 = ni.delta0;
    const int lo0 = ni.lo[0], hi0 = ni.hi[0];
    uint32_t* outp = bandPos + (long long)node * kBandCap;
#pragma unroll
    for (int e = 0; e < 32; ++e) {
      const int m = e >> 1, wr = e & 1;
      const int c0 = lx + 64 * m;
      const float val = (wr ? u[m].y : u[m].x) * scale + d0;
      bool ok = (wr ? vb : va) && c0 >= lo0 && c0 <= hi0 && val >= th;
      if (ok && ni.exOn && excluded(ni, s0, g.s[1], c0, ca + wr, 0)) ok = false;
      const unsigned long long msk = __ballot(ok);
      if (msk == 0ull) continue;
      const int leader = __ffsll((long long)msk) - 1;
      int b0 = 0;
      if (lx == leader) b0 = atomicAdd(&bandCnt[node], __popcll(msk));
      b0 = __shfl(b0, leader);
      if (ok) {
        const int idx = b0 + __popcll(msk & ((1ull << lx) - 1ull));
        if (idx < kBandCap) outp[idx] = (uint32_t)(c0 + s0 * (ca + wr));
      }
    }
  }
}

// ---------------------------------------------------------------------------
// TI spectra on the 1024 x 1024 grid, in FP64 (rounded to FP32 once):
// rows (A, y < s1): the zero-padded real row's forward transform, k0 < 513
// -> tmp[A][y][k0]; columns (A, k0): the forward transform along y (zero
// above s1) -> YT[A][k0][k1].
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_spec1024_rows(const float* __restrict__ sub, int s0, int s1, long long S,
                                                      int nSub, const double2* __restrict__ tw,
                                                      double2* __restrict__ tmp) {
  __shared__ double2 twl[F];
  __shared__ double2 wbuf[4][BUF];
  load_twiddles(twl, tw);
  const int l = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const long long it = (long long)blockIdx.x * 4 + w;
  if (it >= (long long)nSub * s1) return;  // after the block's only barrier
  const int A = (int)(it / s1), y = (int)(it - (long long)A * s1);
  const float* row = sub + A * S + (long long)y * s0;
  double2 u[16];
#pragma unroll
  for (int m = 0; m < 16; ++m) {
    const int x = l + 64 * m;
    u[m] = make_double2(x < s0 ? (double)row[x] : 0.0, 0.0);
  }
  fft1024<-1>(u, wbuf[w], twl, l);
  double2* o = tmp + ((long long)A * F + y) * H0;
#pragma unroll
  for (int m = 0; m < 16; ++m) {
    const int k0 = l + 64 * m;
    if (k0 < H0) o[k0] = u[m];
  }
}

__global__ void __launch_bounds__(256) k_spec1024_cols(const double2* __restrict__ tmp, int s1, int nSub,
                                                      const double2* __restrict__ tw, float2* __restrict__ YT) {
  __shared__ double2 twl[F];
  __shared__ double2 wbuf[4][BUF];
  load_twiddles(twl, tw);
  const int l = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const long long it = (long long)blockIdx.x * 4 + w;
  if (it >= (long long)nSub * H0) return;
  const int A = (int)(it / H0), k0 = (int)(it - (long long)A * H0);
  double2 u[16];
#pragma unroll
  for (int m = 0; m < 16; ++m) {
    const int y = l + 64 * m;
    u[m] = y < s1 ? tmp[((long long)A * F + y) * H0 + k0] : make_double2(0.0, 0.0);
  }
  fft1024<-1>(u, wbuf[w], twl, l);
  float2* o = YT + ((long long)A * H0 + k0) * F + l;
#pragma unroll
  for (int m = 0; m < 16; ++m) o[64 * m] = make_float2((float)u[m].x, (float)u[m].y);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
long long ct1024_stride() { return CT_STRIDE; }
int rows1024_groups() { return F / 2; }

namespace {
int cu_count() {
  static int n = 0;
  if (n == 0) {
    int dev = 0, v = 0;
    (void)hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || v <= 0) v = 256;
    n = v;
  }
  return n;
}
}  // namespace

// every block holds NW wavefronts and one twiddle table (one block per CU);
// a multiple of 8 blocks so the XCD of block b is b % 8
int w1k_blocks() { return std::max(NXCD, (cu_count() / NXCD) * NXCD); }

hipError_t spectra1024(const float* dSub, int s0, int s1, long long S, int nSub, const double2* twd, float2* YT,
                       hipStream_t st) {
  double2* tmp = nullptr;
  hipError_t e = hipMalloc(&tmp, (size_t)nSub * F * H0 * sizeof(double2));
  if (e != hipSuccess) return e;
  e = hipMemsetAsync(tmp, 0, (size_t)nSub * F * H0 * sizeof(double2), st);
  if (e == hipSuccess) {
    const long long rows = (long long)nSub * s1, cols = (long long)nSub * H0;
    hipLaunchKernelGGL(k_spec1024_rows, dim3((unsigned)((rows + 3) / 4)), dim3(256), 0, st, dSub, s0, s1, S, nSub,
                       twd, tmp);
    hipLaunchKernelGGL(k_spec1024_cols, dim3((unsigned)((cols + 3) / 4)), dim3(256), 0, st, tmp, s1, nSub, twd, YT);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(st);
  }
  (void)hipFree(tmp);
  return e;
}

void launch_pencil1024(const ScanGeom& g, const float2* YT, const float2* tw, int nNodes, const NodeInfo* info,
                       const int* nbOff, const float* nbVal, TapInfo1024* taps, uint16_t* tapD0, float* tapVal,
                       float2* CT, const int* count, bool tapsBuilt, hipStream_t st) {
  if (!tapsBuilt)
    hipLaunchKernelGGL(k_taps1024, dim3((unsigned)nNodes), dim3(64), 0, st, info, nbOff, nbVal, g.V, taps, tapD0,
                       tapVal, count);
  // at most one block per CU and per node of the XCD
  const int perX = std::min(w1k_blocks() / NXCD, nNodes);
  hipLaunchKernelGGL(k_pencil1024, dim3((unsigned)(std::max(perX, 1) * NXCD)), dim3(64 * NWP), 0, st, g, YT, tw, info,
                     nbOff, nbVal, taps, tapD0, tapVal, CT, nNodes, count);
}

void launch_rows1024(const ScanGeom& g, const float2* tw, const NodeInfo* info, const float2* CT, int nNodes, int K,
                     bool quant, Cand* cand, float* map, const int* count, int* gcnt, hipStream_t st) {
  const long long items = (long long)nNodes * (F / 2);
  const int blocks = (int)std::max<long long>(1, std::min<long long>(w1k_blocks(), (items + NW - 1) / NW));
  const dim3 grid((unsigned)blocks), block(64 * NW);
#define QSMI_ROWS1024(KT_, Q_) \
  hipLaunchKernelGGL((k_rows1024<KT_, Q_>), grid, block, 0, st, g, tw, info, CT, nNodes, K, cand, map, count, gcnt)
  if (quant) {
    if (K <= 2) QSMI_ROWS1024(2, true);
    else if (K <= 4) QSMI_ROWS1024(4, true);
    else if (K <= 8) QSMI_ROWS1024(8, true);
    else QSMI_ROWS1024(16, true);
  } else {
    if (K <= 2) QSMI_ROWS1024(2, false);
    else if (K <= 4) QSMI_ROWS1024(4, false);
    else if (K <= 8) QSMI_ROWS1024(8, false);
    else QSMI_ROWS1024(16, false);
  }
#undef QSMI_ROWS1024
}

void launch_rows1024_collect(const ScanGeom& g, const float2* tw, const NodeInfo* info, const float2* CT,
                             const int2* queue, const int* qcnt, const float* theta, uint32_t* bandPos, int* bandCnt,
                             int maxItems, hipStream_t st) {
  const int blocks = std::max(1, std::min(w1k_blocks(), (maxItems + NW - 1) / NW));
  hipLaunchKernelGGL(k_rows1024_collect, dim3((unsigned)blocks), dim3(64 * NW), 0, st, g, tw, info, CT, queue, qcnt,
                     theta, bandPos, bandCnt);
}

}  // namespace qsmi
