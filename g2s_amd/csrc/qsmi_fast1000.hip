// Specialised scan kernels for 2-D FFT grids of 1000 x 1000 (the 1000^2
// benchmark workload), gfx950.
//
// One length-1000 transform per 128-thread block (two wavefronts):
// 1000 = 10*10*10, Stockham radix 10; thread j < 100 owns butterfly j of every
// stage (10 points in registers), threads 100..127 idle. Stage outputs are
// exchanged through one 8 KiB LDS buffer. After the last stage thread j
// holds elements {j + 100 q}, exactly the elements the first stage of the
// next transform reads, so forward FFT -> MAC -> inverse FFT runs without an
// LDS round trip in between.
//
// Layouts:
//   YT   [A][k0][k1]            TI spectra, column-contiguous (k1 fastest)
//   CT   [node][c1/2][k0][c1%2]  the column-inverse intermediate in row
//                               pairs: the rows kernel reads one pair as
//                               contiguous 16-byte loads; the pencil kernel's
//                               16-byte writes of neighbouring columns land in
//                               the same lines on one XCD (L2 merges them)
// Block-to-work mapping is XCD-aware: consecutive block ids go to different
// XCDs, so block id b works on column slice (b % 8); each XCD's L2 then holds
// only 1/8 of the TI spectra.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <algorithm>

#include "qsmi_kernels.hpp"
#include "qsmi_select.hpp"

namespace qsmi {
namespace f1k {

constexpr int L = 1000;
constexpr int TR = 2;  // rows per CT tile (a row pair: one two-for-one c2r)
constexpr int NXCD = 8;

// butterfly-local twiddle + radix-10 for one lane's butterfly j at stage Ns
template <int DIR>
__device__ __forceinline__ void stage_compute(float2* u, int j, int Ns, const float2* __restrict__ tw) {
  if (Ns > 1) {
    const int k = j % Ns;
    const int st = (L / 10) / Ns;  // twiddle stride: exponent q*k*st
#pragma unroll
    for (int q = 1; q < 10; ++q) u[q] = cmul(u[q], twid<DIR>(tw[q * k * st]));
  }
  dft10<DIR>(u);
}

__device__ __forceinline__ void load_x(const float2* buf, float2 (&u)[10], int j) {
#pragma unroll
  for (int q = 0; q < 10; ++q) u[q] = buf[j + 100 * q];
}

__device__ __forceinline__ void store_y(float2* buf, const float2 (&u)[10], int j, int Ns) {
  const int k = j % Ns;
  const int base = (j - k) * 10 + k;
#pragma unroll
  for (int q = 0; q < 10; ++q) buf[base + q * Ns] = u[q];
}

// Full transform. On entry: `fromLds` -> input in buf (natural order), else
// input already in u (elements j + 100 q). On exit u holds elements j + 100 q.
// Every thread of the block must call it (barriers inside); only threads
// with `act` (butterfly index j < 100 of a busy group) compute.
template <int DIR>
__device__ __forceinline__ void fft1000(float2* buf, float2 (&u)[10], const float2* tw, int j, bool act,
                                        bool fromLds) {
  if (fromLds) {
    if (act) load_x(buf, u, j);
    __syncthreads();
  }
  if (act) {  // stage 1 (Ns = 1)
    stage_compute<DIR>(u, j, 1, tw);
    store_y(buf, u, j, 1);
  }
  __syncthreads();
  if (act) load_x(buf, u, j);
  __syncthreads();
  if (act) {  // stage 2 (Ns = 10)
    stage_compute<DIR>(u, j, 10, tw);
    store_y(buf, u, j, 10);
  }
  __syncthreads();
  if (act) load_x(buf, u, j);
  __syncthreads();
  if (act) stage_compute<DIR>(u, j, 100, tw);  // stage 3: outputs j + 100 q stay in registers
}

template <int DIR>
__device__ __forceinline__ void fft1000(float2* buf, float2 (&u)[10], const float2* tw, int j, bool fromLds) {
  fft1000<DIR>(buf, u, tw, j, j < 100, fromLds);
}

// Row pair p of one node's CT: Hermitian completion of the half spectrum
// (h0 = 501: columns k = j + 100 q >= 501 are conj(Z[1000 - k]); the
// imaginary parts of columns 0 and 500 do not enter a c2r), then one
// two-for-one c2r along dim0. On exit u[q] holds position c0 = j + 100 q of
// row 2p (x) and row 2p+1 (y). Every thread of the block calls it (barriers
// inside); va / vb: the rows are valid, act = va || vb.
template <int G>
__device__ __forceinline__ void rows1k_pair(const float2* __restrict__ Cn, int p, int h0, bool va, bool vb, bool act,
                                            bool gin, int grp, int j, float2 (&buf)[G][L], const float2* twl,
                                            float2 (&u)[10]) {
  if (act) {
    const float4* rowp = reinterpret_cast<const float4*>(Cn + (long long)p * h0 * TR);
#pragma unroll
    for (int q = 0; q < 10; ++q) {
      const int k = j + 100 * q;
      const bool cj = (q >= 6) || (q == 5 && j > 0);
      const int kk = cj ? (L - k) : k;
      const float4 ab = rowp[kk];  // (row 2p, row 2p+1) at column kk
      float2 A = va ? make_float2(ab.x, ab.y) : make_float2(0.f, 0.f);
      float2 B = vb ? make_float2(ab.z, ab.w) : make_float2(0.f, 0.f);
      if ((q == 0 || q == 5) && j == 0) {
        A.y = 0.f;
        B.y = 0.f;
      }
      if (cj) {
        A.y = -A.y;
        B.y = -B.y;
      }
      u[q] = make_float2(A.x - B.y, A.y + B.x);
    }
  }
  fft1000<+1>(buf[gin ? grp : 0], u, twl, j, act, false);
}

}  // namespace f1k

// ---------------------------------------------------------------------------
// Distinct neighbour rows of each node (one 64-thread block per node):
// TapInfo with the rows ordered by (row mod 10, row), and the neighbours'
// slots in tap order: tapPos[node][j] = start[tap of j] + (number of earlier
// neighbours in the same tap), so the pencil can sum every tap's
// contributions in neighbour-list order (a fixed summation order, hence
// reproducible FP32 maps). T = 0 (full transform) above kTapMax rows or
// kTapNbMax neighbours.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(64) k_taps1k(const NodeInfo* __restrict__ info, const int* __restrict__ nbOff,
                                              TapInfo* __restrict__ taps, uint16_t* __restrict__ tapPos,
                                              const int* __restrict__ nodeCount) {
  using namespace f1k;
  __shared__ int flag[L];
  __shared__ short tix[L];           // tap index of a present row
  __shared__ short tapJ[kTapNbMax];  // tap of neighbour j
  __shared__ int cntT[kTapMax];
  __shared__ int startL[kTapMax + 1];
  const int node = blockIdx.x, lane = threadIdx.x;
  if (nodeCount && node >= *nodeCount) return;
  const NodeInfo ni = info[node];
  for (int i = lane; i < L; i += 64) flag[i] = 0;
  __syncthreads();
  for (int t = lane; t < ni.nb; t += 64) {
    int r = (-nbOff[(ni.nbStart + t) * 2 + 1]) % L;
    if (r < 0) r += L;
    flag[r] = 1;
  }
  __syncthreads();
  TapInfo* tp = taps + node;
  int count = 0;
  if (lane == 0) tp->cls[0] = 0;
  for (int c = 0; c < 10; ++c) {
    for (int h = 0; h < 2; ++h) {
      const int m = lane + 64 * h;
      const int row = c + 10 * m;
      const bool present = m < 100 && flag[row] != 0;
      const unsigned long long mask = __ballot(present);
      const int before = __popcll(mask & ((1ull << lane) - 1ull));
      if (present && count + before < kTapMax) {
        tp->row[count + before] = row;
        tix[row] = (short)(count + before);
      }
      count += __popcll(mask);
    }
    if (lane == 0) tp->cls[c + 1] = count;
  }
  const int T = (count <= kTapMax && ni.nb <= kTapNbMax) ? count : 0;
  if (lane == 0) tp->T = T;
  if (T == 0) return;  // block-uniform
  for (int t = lane; t < T; t += 64) cntT[t] = 0;
  __syncthreads();
  for (int t = lane; t < ni.nb; t += 64) {
    int r = (-nbOff[(ni.nbStart + t) * 2 + 1]) % L;
    if (r < 0) r += L;
    const short tt = tix[r];
    tapJ[t] = tt;
    atomicAdd(&cntT[tt], 1);
  }
  __syncthreads();
  if (lane == 0) {
    int acc = 0;
    for (int t = 0; t < T; ++t) {
      startL[t] = acc;
      tp->start[t] = acc;
      acc += cntT[t];
    }
    startL[T] = acc;
    tp->start[T] = acc;
  }
  __syncthreads();
  uint16_t* out = tapPos + (long long)node * kTapNbMax;
  for (int t = lane; t < ni.nb; t += 64) {
    const short tt = tapJ[t];
    int r = 0;
    for (int q = 0; q < t; ++q) r += (tapJ[q] == tt) ? 1 : 0;
    out[t] = (uint16_t)(startL[tt] + r);
  }
}

// ---------------------------------------------------------------------------
// Pencil along dim1 for 1000 x 1000 grids. 128 threads per block; block b
// works on column k0 = slice(b % 8) and a chunk of `npb` nodes.
//
// Forward transform of a node's sparse kernel column, two ways:
//  * T <= kTapMax distinct neighbour rows r_t: with k1 = j + 100 q,
//      X(k1) = sum_c exp(-2 pi i q c / 10) S_j(c),
//      S_j(c) = sum_{t: r_t mod 10 = c} G(t) exp(-2 pi i j r_t / 1000),
//    where G(t) sums value * exp(+2 pi i k0 d0 / 1000) over the row's
//    neighbours: T twiddled MACs and one radix-10 butterfly per lane, no
//    data exchange, and X arrives in exactly the registers (elements
//    j + 100 q) the MAC and the inverse transform use. Two sub-variables
//    share each twiddle.
//  * otherwise: scatter into an 8 KiB column and a full Stockham transform.
// Twiddles are read from an LDS copy of the table.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(3))) k_pencil1k(ScanGeom g, const float2* __restrict__ YT,
                                                 const float2* __restrict__ tw1k, int nNodes, int npb,
                                                 const NodeInfo* __restrict__ info, const int* __restrict__ nbOff,
                                                 const float* __restrict__ nbVal, const TapInfo* __restrict__ taps,
                                                 const uint16_t* __restrict__ tapPos, float2* __restrict__ CT,
                                                 long long ctStride, const int* __restrict__ count, int chunks) {
  using namespace f1k;
  constexpr int kChunk = 512;         // neighbours per pass of the full-transform column build
  __shared__ float2 sb[L];            // contributions in tap order / column / exchange buffer
  __shared__ float2 twl[L];           // exp(-2 pi i t / 1000)
  __shared__ float2 G[2 * kTapMax];   // [tap][sub-variable of the pair]
  __shared__ int rowL[kTapMax];
  __shared__ int startL[kTapMax + 1];
  __shared__ float2 cl[kChunk];       // full-transform path: contributions in list order
  __shared__ short rl[kChunk];        // and their rows
  const int lane = threadIdx.x;  // butterfly index j for lane < 100
  const int h0 = g.h0;  // 501
  const int slice = (h0 + NXCD - 1) / NXCD;
  const int bid = blockIdx.x;
  const int xcd = bid % NXCD;
  const int r = bid / NXCD;
  const int k0 = xcd * slice + (r % slice);
  const int chunk = r / slice;
  if (k0 >= h0) return;
  // with a device count only nodes [0, count) are scanned, spread over all
  // node chunks of the grid
  int nAct = nNodes, per = npb;
  if (count) {
    nAct = min(*count, nNodes);
    per = max(1, (nAct + chunks - 1) / chunks);
  }
  const int nodeBeg = chunk * per;
  const int nodeEnd = min(nAct, nodeBeg + per);
  if (nodeBeg >= nodeEnd) return;
  const int V = g.V;
  const bool act = lane < 100;
  for (int i = lane; i < L; i += 128) twl[i] = tw1k[i];
  __syncthreads();
  for (int node = nodeBeg; node < nodeEnd; ++node) {
    const NodeInfo ni = info[node];
    if (ni.lo[1] > ni.hi[1] || ni.lo[0] > ni.hi[0]) continue;
    float2 M[10];
#pragma unroll
    for (int q = 0; q < 10; ++q) M[q] = make_float2(0.f, 0.f);
    const TapInfo* tp = taps + node;
    const int T = tp->T;
    if (T > 0) {
      const uint16_t* tpos = tapPos + (long long)node * kTapNbMax;
      if (lane < T) rowL[lane] = tp->row[lane];
      if (lane <= T) startL[lane] = tp->start[lane];
      for (int a0 = 0; a0 < V; a0 += 2) {
        const bool two = a0 + 1 < V;
        // every neighbour's value * exp(+2 pi i k0 d0 / 1000) into its slot in
        // tap order (sb as [slot][sub-variable of the pair]) ...
        for (int t = lane; t < ni.nb; t += 128) {
          const int jj = ni.nbStart + t;
          const float v0 = nbVal[(long long)jj * V + a0];
          const float v1 = two ? nbVal[(long long)jj * V + a0 + 1] : 0.f;
          const int d0 = nbOff[jj * 2 + 0];
          int e = (k0 * d0) % L;
          if (e < 0) e += L;
          const float2 w = twl[e];
          const int slot = tpos[t];
          sb[2 * slot] = make_float2(v0 * w.x, -v0 * w.y);
          sb[2 * slot + 1] = make_float2(v1 * w.x, -v1 * w.y);
        }
        __syncthreads();
        // ... then one lane per tap sums its row's neighbours in list order
        if (lane < T) {
          float2 g0 = make_float2(0.f, 0.f), g1 = make_float2(0.f, 0.f);
          const int e1 = startL[lane + 1];
          for (int i = startL[lane]; i < e1; ++i) {
            const float2 c0 = sb[2 * i], c1 = sb[2 * i + 1];
            g0.x += c0.x;
            g0.y += c0.y;
            g1.x += c1.x;
            g1.y += c1.y;
          }
          G[2 * lane] = g0;
          G[2 * lane + 1] = g1;
        }
        __syncthreads();
        if (act) {
          float2 S0[10], S1[10];
#pragma unroll
          for (int c = 0; c < 10; ++c) {
            float2 s0 = make_float2(0.f, 0.f), s1 = make_float2(0.f, 0.f);
            const int te = tp->cls[c + 1];
            for (int t = tp->cls[c]; t < te; ++t) {
              // (j * r) mod 1000 in float: exact for every j < 100, r < 1000
              const uint32_t p = (uint32_t)lane * (uint32_t)rowL[t];
              const uint32_t qd = (uint32_t)((float)p * 0.001f);
              const float2 w = twl[p - qd * 1000u];
              const float2 g0 = G[2 * t], g1 = G[2 * t + 1];
              s0.x = fmaf(g0.x, w.x, s0.x);
              s0.x = fmaf(-g0.y, w.y, s0.x);
              s0.y = fmaf(g0.x, w.y, s0.y);
              s0.y = fmaf(g0.y, w.x, s0.y);
              s1.x = fmaf(g1.x, w.x, s1.x);
              s1.x = fmaf(-g1.y, w.y, s1.x);
              s1.y = fmaf(g1.x, w.y, s1.y);
              s1.y = fmaf(g1.y, w.x, s1.y);
            }
            S0[c] = s0;
            S1[c] = s1;
          }
          dft10<-1>(S0);
          const float2* Y0 = YT + ((long long)g.A[a0] * h0 + k0) * L;
#pragma unroll
          for (int q = 0; q < 10; ++q) {
            const float2 y = Y0[lane + 100 * q];
            M[q].x = fmaf(y.x, S0[q].x, M[q].x);
            M[q].x = fmaf(-y.y, S0[q].y, M[q].x);
            M[q].y = fmaf(y.x, S0[q].y, M[q].y);
            M[q].y = fmaf(y.y, S0[q].x, M[q].y);
          }
          if (two) {
            dft10<-1>(S1);
            const float2* Y1 = YT + ((long long)g.A[a0 + 1] * h0 + k0) * L;
#pragma unroll
            for (int q = 0; q < 10; ++q) {
              const float2 y = Y1[lane + 100 * q];
              M[q].x = fmaf(y.x, S1[q].x, M[q].x);
              M[q].x = fmaf(-y.y, S1[q].y, M[q].x);
              M[q].y = fmaf(y.x, S1[q].y, M[q].y);
              M[q].y = fmaf(y.y, S1[q].x, M[q].y);
            }
          }
        }
        __syncthreads();
      }
    } else {
      for (int a = 0; a < V; ++a) {
        // the sparse neighbour column: value * exp(+2 pi i k0 d0 / 1000) at
        // row (-d1) mod 1000; lane l owns rows l + 128 i and sums their
        // neighbours in list order
        float2 rs[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) rs[i] = make_float2(0.f, 0.f);
        for (int c0 = 0; c0 < ni.nb; c0 += kChunk) {
          const int cn = min(kChunk, ni.nb - c0);
          for (int t = lane; t < cn; t += 128) {
            const int jj = ni.nbStart + c0 + t;
            const float v = nbVal[(long long)jj * V + a];
            const int d0 = nbOff[jj * 2 + 0];
            const int d1 = nbOff[jj * 2 + 1];
            int e = (k0 * d0) % L;
            if (e < 0) e += L;
            const float2 w = twl[e];
            int q = (-d1) % L;
            if (q < 0) q += L;
            cl[t] = make_float2(v * w.x, -v * w.y);
            rl[t] = (short)q;
          }
          __syncthreads();
          for (int t = 0; t < cn; ++t) {
            const int q = rl[t];
            if ((q & 127) == lane) {
              const float2 c = cl[t];
              const int i = q >> 7;
#pragma unroll
              for (int k = 0; k < 8; ++k)
                if (k == i) {
                  rs[k].x += c.x;
                  rs[k].y += c.y;
                }
            }
          }
          __syncthreads();
        }
#pragma unroll
        for (int i = 0; i < 8; ++i)
          if (lane + 128 * i < L) sb[lane + 128 * i] = rs[i];
        __syncthreads();
        float2 u[10];
        fft1000<-1>(sb, u, twl, lane, true);
        if (act) {
          const float2* Y = YT + ((long long)g.A[a] * h0 + k0) * L;
#pragma unroll
          for (int q = 0; q < 10; ++q) {
            const float2 y = Y[lane + 100 * q];
            const float2 x = u[q];
            M[q].x = fmaf(y.x, x.x, M[q].x);
            M[q].x = fmaf(-y.y, x.y, M[q].x);
            M[q].y = fmaf(y.x, x.y, M[q].y);
            M[q].y = fmaf(y.y, x.x, M[q].y);
          }
        }
        __syncthreads();
      }
    }
    // inverse along dim1, input already in registers
    fft1000<+1>(sb, M, twl, lane, false);
    if (act) {
      float2* Cn = CT + (long long)node * ctStride;
      const int lo = ni.lo[1], hi = ni.hi[1];
#pragma unroll
      for (int q = 0; q < 10; ++q) {
        const int c1 = lane + 100 * q;
        if (c1 >= lo && c1 <= hi) Cn[((long long)(c1 / f1k::TR) * h0 + k0) * f1k::TR + (c1 % f1k::TR)] = M[q];
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Rows for 1000 x 1000: G row pairs (2p, 2p+1) of one node per block, one
// group of 100 threads per pair (block of 64*ceil(100G/64) threads); a
// pair is one two-for-one c2r along dim0: thread j of a group loads
// Z[j + 100 q] of both rows as one 16-byte load. Epilogue: scale + delta0,
// the valid-position mask, and the block's K best of all G pairs by the
// device rule -> cand[node][block][K]. Blocks without a valid row write -inf
// candidates.
// ---------------------------------------------------------------------------
template <int KT, int G>
__global__ void __launch_bounds__(64 * ((100 * G + 63) / 64)) k_rows1k(ScanGeom g, const float2* __restrict__ tw1k,
                                                                     const NodeInfo* __restrict__ info,
                                                                     const float2* __restrict__ CT,
                                                                     long long ctStride, int K,
                                                                     Cand* __restrict__ cand, int candBlocks,
                                                                     float* __restrict__ map,
                                                                     const int* __restrict__ count,
                                                                     int* __restrict__ gcnt) {
  using namespace f1k;
  __shared__ float2 buf[G][L];
  __shared__ float2 twl[L];
  const int tid = threadIdx.x;
  const int grp = tid / 100;
  const int j = tid - 100 * grp;
  const int node = blockIdx.y;
  if (count && node >= *count) return;  // beyond the compacted nodes
  const NodeInfo ni = info[node];
  const int h0 = g.h0;
  // one candidate group of K per wavefront of the block
  const int wavesPerBlock = blockDim.x >> 6;
  Cand* out = cand ? cand + ((long long)node * candBlocks + (long long)blockIdx.x * wavesPerBlock) * K : nullptr;
  const int p = blockIdx.x * G + grp;  // this group's pair
  const int ca = 2 * p;
  const bool gin = grp < G && p < L / 2;
  const bool va = gin && (ca >= ni.lo[1] && ca <= ni.hi[1]);
  const bool vb = gin && (ca + 1 >= ni.lo[1] && ca + 1 <= ni.hi[1]);
  // block-uniform exit when none of the block's rows is valid
  const int rlo = max(ni.lo[1], 2 * blockIdx.x * G), rhi = min(ni.hi[1], 2 * (blockIdx.x * G + G) - 1);
  if (rlo > rhi || ni.lo[0] > ni.hi[0]) {
    if (out && tid < K * wavesPerBlock) {
      Cand w;
      w.key = -INFINITY;
      w.h = 0u;
      w.pos = 0xFFFFFFFFu;
      w.val = -INFINITY;
      out[tid] = w;
    }
    return;
  }
  for (int i = tid; i < L; i += blockDim.x) twl[i] = tw1k[i];  // first read after the FFT's first barrier
  const bool act = va || vb;  // this thread's group has a valid row (j < 100 by construction)
  float2 u[10];
  f1k::rows1k_pair<G>(CT + (long long)node * ctStride, p, h0, va, vb, act, gin, grp, j, buf, twl, u);
  const float scale = ni.scale, d0 = ni.delta0, quantum = ni.quantum;
  const int s0 = g.s[0];
  const int lo0 = ni.lo[0], hi0 = ni.hi[0];
  if (map && act) {
#pragma unroll
    for (int q = 0; q < 10; ++q) {
      const int c0 = j + 100 * q;
      if (c0 < lo0 || c0 > hi0) continue;
      if (va) map[(long long)node * g.s[0] * g.s[1] + (uint32_t)(c0 + s0 * ca)] = u[q].x * scale + d0;
      if (vb) map[(long long)node * g.s[0] * g.s[1] + (uint32_t)(c0 + s0 * (ca + 1))] = u[q].y * scale + d0;
    }
  }
  if (!out) return;
  // keys of this thread's 20 positions, -inf where invalid
  float key[20];
  if (quantum > 0.f) {
    const float invq = 1.f / quantum;
#pragma unroll
    for (int q = 0; q < 10; ++q) {
      key[2 * q] = rintf((u[q].x * scale + d0) * invq);
      key[2 * q + 1] = rintf((u[q].y * scale + d0) * invq);
    }
  } else {
#pragma unroll
    for (int q = 0; q < 10; ++q) {
      key[2 * q] = u[q].x * scale + d0;
      key[2 * q + 1] = u[q].y * scale + d0;
    }
  }
#pragma unroll
  for (int q = 0; q < 10; ++q) {
    const int c0 = j + 100 * q;
    const bool okc = act && c0 >= lo0 && c0 <= hi0;
    key[2 * q] = (okc && va) ? key[2 * q] : -INFINITY;
    key[2 * q + 1] = (okc && vb) ? key[2 * q + 1] : -INFINITY;
  }
  if (ni.exOn) {  // node-uniform: verbatim exclusion disk
#pragma unroll
    for (int q = 0; q < 10; ++q) {
      const int c0 = j + 100 * q;
      if (excluded(ni, s0, g.s[1], c0, ca, 0)) key[2 * q] = -INFINITY;
      if (excluded(ni, s0, g.s[1], c0, ca + 1, 0)) key[2 * q + 1] = -INFINITY;
    }
  }
  // pass 1: per-thread KT largest keys (min/max network)
  float m[KT];
#pragma unroll
  for (int i = 0; i < KT; ++i) m[i] = -INFINITY;
#pragma unroll
  for (int e = 0; e < 20; ++e) {
    float x = key[e];
#pragma unroll
    for (int i = 0; i < KT; ++i) {
      const float hi = fmaxf(m[i], x);
      x = fminf(m[i], x);
      m[i] = hi;
    }
  }
  // tau2 also excludes the -inf of invalid positions
  const float tau2 = fmaxf(wave_kth_key<KT>(m, K), -3.402823466e38f);
  if (gcnt && ni.eps > 0.f && quantum == 0.f) {
    // exact selection: how many of the wavefront's values lie within 2 eps
    // of its K-th best (more than K: the node's band may reach beyond the
    // wavefront's candidates)
    const float th = (tau2 <= -3.402823466e38f) ? -INFINITY : tau2 - 2.f * ni.eps * 1.0009765625f;
    int c = 0;
#pragma unroll
    for (int e = 0; e < 20; ++e) c += (key[e] >= th) ? 1 : 0;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off);
    if ((tid & 63) == 0)
      gcnt[(long long)node * candBlocks + (long long)blockIdx.x * wavesPerBlock + (tid >> 6)] = c;
  }
  // pass 2: exact device rule among the keys >= tau (normally K per block)
  RankList<KT> rl;
  rl.init();
#pragma unroll
  for (int e = 0; e < 20; ++e) {
    if (key[e] >= tau2) {
      const int q = e >> 1, w = e & 1;
      const uint32_t pos = (uint32_t)(j + 100 * q + s0 * (ca + w));
      rl.insert(key[e], tie_hash(ni.seedbits, pos), pos, (w ? u[q].y : u[q].x) * scale + d0);
    }
  }
  wave_extract<KT>(rl, K, out + (threadIdx.x >> 6) * K);
}

// ---------------------------------------------------------------------------
// Exact selection, second pass on the 1000 x 1000 path: for every queued
// (node, rows block) whose candidates may hide more positions within the
// node's band, the block's rows again (same FFT and epilogue as k_rows1k),
// appending every valid position whose FP32 value is >= theta[node] to the
// node's band list. Persistent grid over the device queue count.
// ---------------------------------------------------------------------------
template <int G>
__global__ void __launch_bounds__(64 * ((100 * G + 63) / 64)) k_rows1k_collect(
    ScanGeom g, const float2* __restrict__ tw1k, const NodeInfo* __restrict__ info, const float2* __restrict__ CT,
    long long ctStride, const int2* __restrict__ queue, const int* __restrict__ qcnt,
    const float* __restrict__ theta, uint32_t* __restrict__ bandPos, int* __restrict__ bandCnt) {
  using namespace f1k;
  __shared__ float2 buf[G][L];
  __shared__ float2 twl[L];
  const int tid = threadIdx.x;
  const int grp = tid / 100;
  const int j = tid - 100 * grp;
  const int lane = tid & 63;
  const int h0 = g.h0, s0 = g.s[0];
  const int items = *qcnt;
  if ((int)blockIdx.x >= items) return;
  for (int i = tid; i < L; i += blockDim.x) twl[i] = tw1k[i];  // first read after the FFT's first barrier
  for (int it = blockIdx.x; it < items; it += gridDim.x) {
    const int2 qi = queue[it];
    const int node = qi.x, bx = qi.y;
    const NodeInfo ni = info[node];
    const int p = bx * G + grp;
    const int ca = 2 * p;
    const bool gin = grp < G && p < L / 2;
    const bool va = gin && (ca >= ni.lo[1] && ca <= ni.hi[1]);
    const bool vb = gin && (ca + 1 >= ni.lo[1] && ca + 1 <= ni.hi[1]);
    const bool act = va || vb;
    float2 u[10];
    rows1k_pair<G>(CT + (long long)node * ctStride, p, h0, va, vb, act, gin, grp, j, buf, twl, u);
    const float th = theta[node];
    const float scale = ni.scale, d0 = ni.delta0;
    const int lo0 = ni.lo[0], hi0 = ni.hi[0];
    uint32_t* out = bandPos + (long long)node * kBandCap;
#pragma unroll
    for (int e = 0; e < 20; ++e) {
      const int q = e >> 1, w = e & 1;
      const int c0 = j + 100 * q;
      const float val = (w ? u[q].y : u[q].x) * scale + d0;
      bool ok = act && (w ? vb : va) && c0 >= lo0 && c0 <= hi0 && val >= th;
      if (ok && ni.exOn && excluded(ni, s0, g.s[1], c0, ca + w, 0)) ok = false;
      const unsigned long long m = __ballot(ok);
      if (m == 0ull) continue;
      const int leader = __ffsll((long long)m) - 1;
      int base = 0;
      if (lane == leader) base = atomicAdd(&bandCnt[node], __popcll(m));
      base = __shfl(base, leader);
      if (ok) {
        const int idx = base + __popcll(m & ((1ull << lane) - 1ull));
        if (idx < kBandCap) out[idx] = (uint32_t)(c0 + s0 * (ca + w));
      }
    }
  }
}

// YT[A][k0][k1] <- Yhat[A][k1][k0] (once per TI)
__global__ void k_transpose_spectra(const float2* __restrict__ Y, float2* __restrict__ YT, int h0, int f1, int nA) {
  __shared__ float2 tile[32][33];
  const int A = blockIdx.z;
  const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
  const float2* src = Y + (long long)A * h0 * f1;
  float2* dst = YT + (long long)A * h0 * f1;
  for (int i = threadIdx.y; i < 32; i += blockDim.y) {
    const int k1 = by + i, k0 = bx + threadIdx.x;
    if (k1 < f1 && k0 < h0) tile[i][threadIdx.x] = src[(long long)k1 * h0 + k0];
  }
  __syncthreads();
  for (int i = threadIdx.y; i < 32; i += blockDim.y) {
    const int k0 = bx + i, k1 = by + threadIdx.x;
    if (k1 < f1 && k0 < h0) dst[(long long)k0 * f1 + k1] = tile[threadIdx.x][i];
  }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
long long ct1k_stride(int h0) { return (long long)((1000 + f1k::TR - 1) / f1k::TR) * h0 * f1k::TR; }

void launch_transpose_spectra(const float2* Y, float2* YT, int h0, int f1, int nA, hipStream_t st) {
  dim3 grid((h0 + 31) / 32, (f1 + 31) / 32, nA);
  hipLaunchKernelGGL(k_transpose_spectra, grid, dim3(32, 8), 0, st, Y, YT, h0, f1, nA);
}

void launch_pencil1k(const ScanGeom& g, const float2* YT, const float2* tw1k, int nNodes, int npb,
                     const NodeInfo* info, const int* nbOff, const float* nbVal, TapInfo* taps, uint16_t* tapPos,
                     float2* CT, const int* count, hipStream_t st) {
  hipLaunchKernelGGL(k_taps1k, dim3((unsigned)nNodes), dim3(64), 0, st, info, nbOff, taps, tapPos, count);
  const int slice = (g.h0 + f1k::NXCD - 1) / f1k::NXCD;
  const int chunks = (nNodes + npb - 1) / npb;
  const long long blocks = (long long)f1k::NXCD * slice * chunks;
  hipLaunchKernelGGL(k_pencil1k, dim3((unsigned)blocks), dim3(128), 0, st, g, YT, tw1k, nNodes, npb, info, nbOff, nbVal,
                     taps, tapPos, CT, ct1k_stride(g.h0), count, chunks);
}

// row pairs per rows block: 5 pairs -> 512 threads, 98 % of the lanes busy
constexpr int kRowsG = 5;

constexpr int kRowsBlocks = (1000 / 2 + kRowsG - 1) / kRowsG;
constexpr int kRowsWaves = (100 * kRowsG + 63) / 64;

// candidate groups per node: one per wavefront of every rows block
int rows1k_tiles() { return kRowsBlocks * kRowsWaves; }
int rows1k_blocks() { return kRowsBlocks; }
int rows1k_waves_per_block() { return kRowsWaves; }

void launch_rows1k_collect(const ScanGeom& g, const float2* tw1k, const NodeInfo* info, const float2* CT,
                           const int2* queue, const int* qcnt, const float* theta, uint32_t* bandPos, int* bandCnt,
                           int maxItems, hipStream_t st) {
  // persistent grid: at most one block per possible item, at most 2048
  const int blocks = std::max(1, std::min(maxItems, 2048));
  hipLaunchKernelGGL((k_rows1k_collect<kRowsG>), dim3((unsigned)blocks), dim3(64 * kRowsWaves), 0, st, g, tw1k, info,
                     CT, ct1k_stride(g.h0), queue, qcnt, theta, bandPos, bandCnt);
}

hipError_t set_rows1k_lds() { return hipSuccess; }

void launch_rows1k(const ScanGeom& g, const float2* tw1k, const NodeInfo* info, const float2* CT, int nNodes, int K,
                   Cand* cand, float* map, const int* count, int* gcnt, hipStream_t st) {
  const int blocks = kRowsBlocks, groups = rows1k_tiles();
  dim3 grid((unsigned)blocks, (unsigned)nNodes);
  dim3 block(64 * ((100 * kRowsG + 63) / 64));
  const long long cs = ct1k_stride(g.h0);
  if (K <= 2)
    hipLaunchKernelGGL((k_rows1k<2, kRowsG>), grid, block, 0, st, g, tw1k, info, CT, cs, K, cand, groups, map, count,
                       gcnt);
  else if (K <= 4)
    hipLaunchKernelGGL((k_rows1k<4, kRowsG>), grid, block, 0, st, g, tw1k, info, CT, cs, K, cand, groups, map, count,
                       gcnt);
  else if (K <= 8)
    hipLaunchKernelGGL((k_rows1k<8, kRowsG>), grid, block, 0, st, g, tw1k, info, CT, cs, K, cand, groups, map, count,
                       gcnt);
  else
    hipLaunchKernelGGL((k_rows1k<16, kRowsG>), grid, block, 0, st, g, tw1k, info, CT, cs, K, cand, groups, map,
                       count, gcnt);
}

}  // namespace qsmi
