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

#include "qsmi_kernels.hpp"

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
// Every thread of the block must call it (barriers inside); threads >= 100
// idle.
template <int DIR>
__device__ __forceinline__ void fft1000(float2* buf, float2 (&u)[10], const float2* tw, int j, bool fromLds) {
  const bool act = j < 100;
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

}  // namespace f1k

// ---------------------------------------------------------------------------
// Pencil along dim1 for 1000 x 1000 grids. 128 threads per block; block b
// works on column k0 = slice(b % 8) and a chunk of `npb` nodes, one active
// sub-variable at a time through one 8 KiB scatter/exchange buffer.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(128) k_pencil1k(ScanGeom g, const float2* __restrict__ YT,
                                                 const float2* __restrict__ tw1k, int nNodes, int npb,
                                                 const NodeInfo* __restrict__ info, const int* __restrict__ nbOff,
                                                 const float* __restrict__ nbVal, float2* __restrict__ CT,
                                                 long long ctStride) {
  using namespace f1k;
  __shared__ float2 sb[L];
  const int lane = threadIdx.x;  // butterfly index j for lane < 100
  const int h0 = g.h0;  // 501
  const int slice = (h0 + NXCD - 1) / NXCD;
  const int bid = blockIdx.x;
  const int xcd = bid % NXCD;
  const int r = bid / NXCD;
  const int k0 = xcd * slice + (r % slice);
  const int chunk = r / slice;
  if (k0 >= h0) return;
  const int nodeBeg = chunk * npb;
  const int nodeEnd = min(nNodes, nodeBeg + npb);
  const int V = g.V;
  const bool act = lane < 100;
  for (int node = nodeBeg; node < nodeEnd; ++node) {
    const NodeInfo ni = info[node];
    if (ni.lo[1] > ni.hi[1] || ni.lo[0] > ni.hi[0]) continue;
    float2 M[10];
#pragma unroll
    for (int q = 0; q < 10; ++q) M[q] = make_float2(0.f, 0.f);
    for (int a = 0; a < V; ++a) {
      // zero the scatter buffer, then the sparse neighbour column:
      // value * exp(+2 pi i k0 d0 / 1000) at row (-d1) mod 1000
      for (int i = lane; i < L; i += 128) sb[i] = make_float2(0.f, 0.f);
      __syncthreads();
      for (int t = lane; t < ni.nb; t += 128) {
        const int jj = ni.nbStart + t;
        const float v = nbVal[(long long)jj * V + a];
        if (v == 0.f) continue;
        const int d0 = nbOff[jj * 2 + 0];
        const int d1 = nbOff[jj * 2 + 1];
        int e = (k0 * d0) % L;
        if (e < 0) e += L;
        const float2 w = tw1k[e];
        int q = (-d1) % L;
        if (q < 0) q += L;
        atomicAdd(&sb[q].x, v * w.x);
        atomicAdd(&sb[q].y, -v * w.y);
      }
      __syncthreads();
      float2 u[10];
      fft1000<-1>(sb, u, tw1k, lane, true);
      if (act) {
        const float2* Y = YT + ((long long)g.A[a] * h0 + k0) * L;
#pragma unroll
        for (int q = 0; q < 10; ++q) {
          const float2 y = Y[lane + 100 * q];
          const float2 x = u[q];
          M[q].x += y.x * x.x - y.y * x.y;
          M[q].y += y.x * x.y + y.y * x.x;
        }
      }
      __syncthreads();
    }
    // inverse along dim1, input already in registers
    fft1000<+1>(sb, M, tw1k, lane, false);
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

// Block-wide K best of per-thread sorted lists (blockDim multiple of 64),
// best first into out[0..K). Ends with a barrier.
template <int KT>
__device__ void block_topk(Cand* c, int K, Cand* out, Cand* red) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nwaves = blockDim.x >> 6;
  for (int it = 0; it < K; ++it) {
    Cand best = c[0];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      Cand o;
      o.key = __shfl_xor(best.key, off);
      o.h = __shfl_xor(best.h, off);
      o.pos = __shfl_xor(best.pos, off);
      o.val = __shfl_xor(best.val, off);
      if (cand_better(o, best)) best = o;
    }
    if (lane == 0) red[wave] = best;
    __syncthreads();
    Cand w = red[0];
    for (int i = 1; i < nwaves; ++i)
      if (cand_better(red[i], w)) w = red[i];
    if (tid == 0) out[it] = w;
    if (c[0].pos == w.pos && c[0].pos != 0xFFFFFFFFu) {
#pragma unroll
      for (int i = 0; i < KT - 1; ++i) c[i] = c[i + 1];
      c[KT - 1].key = -INFINITY;
      c[KT - 1].h = 0u;
      c[KT - 1].pos = 0xFFFFFFFFu;
      c[KT - 1].val = -INFINITY;
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Rows for 1000 x 1000: one 128-thread block per row pair (2p, 2p+1) of one
// node (two-for-one c2r along dim0); thread j loads Z[j + 100 q] of both rows
// as one 16-byte load; epilogue, mask, block top-K -> cand[node][p][K].
// Pairs without a valid row write -inf candidates.
// ---------------------------------------------------------------------------
template <int KT>
__global__ void __launch_bounds__(128) k_rows1k_pair(ScanGeom g, const float2* __restrict__ tw1k,
                                                     const NodeInfo* __restrict__ info,
                                                     const float2* __restrict__ CT, long long ctStride, int K,
                                                     Cand* __restrict__ cand, int candPairs,
                                                     float* __restrict__ map) {
  using namespace f1k;
  __shared__ float2 buf[L];
  __shared__ Cand red[2];
  const int j = threadIdx.x;
  const int node = blockIdx.y;
  const int p = blockIdx.x;
  const NodeInfo ni = info[node];
  const int h0 = g.h0;
  Cand* out = cand ? cand + ((long long)node * candPairs + p) * K : nullptr;
  const int ca = 2 * p;
  const bool va = (ca >= ni.lo[1] && ca <= ni.hi[1]), vb = (ca + 1 >= ni.lo[1] && ca + 1 <= ni.hi[1]);
  if ((!va && !vb) || ni.lo[0] > ni.hi[0]) {
    if (out && j < K) {
      Cand w;
      w.key = -INFINITY;
      w.h = 0u;
      w.pos = 0xFFFFFFFFu;
      w.val = -INFINITY;
      out[j] = w;
    }
    return;
  }
  const float4* rowp = reinterpret_cast<const float4*>(CT + (long long)node * ctStride + (long long)p * h0 * TR);
  const bool act = j < 100;
  float2 u[10];
  if (act) {
#pragma unroll
    for (int q = 0; q < 10; ++q) {
      const int k = j + 100 * q;
      const bool cj = (k >= h0);
      const int kk = cj ? (L - k) : k;
      const float4 ab = rowp[kk];  // (row 2p, row 2p+1) at column kk
      float2 A = va ? make_float2(ab.x, ab.y) : make_float2(0.f, 0.f);
      float2 B = vb ? make_float2(ab.z, ab.w) : make_float2(0.f, 0.f);
      if (kk == 0 || 2 * kk == L) {  // a Hermitian c2r ignores these imaginary parts
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
  fft1000<+1>(buf, u, tw1k, j, false);
  Cand c[KT];
#pragma unroll
  for (int i = 0; i < KT; ++i) {
    c[i].key = -INFINITY;
    c[i].h = 0u;
    c[i].pos = 0xFFFFFFFFu;
    c[i].val = -INFINITY;
  }
  const float scale = ni.scale, d0 = ni.delta0, quantum = ni.quantum;
  const int s0 = g.s[0];
  if (act) {
#pragma unroll
    for (int q = 0; q < 10; ++q) {
      const int c0 = j + 100 * q;
      if (c0 < ni.lo[0] || c0 > ni.hi[0]) continue;
#pragma unroll
      for (int w = 0; w < 2; ++w) {
        if (w == 0 ? !va : !vb) continue;
        const float val = (w == 0 ? u[q].x : u[q].y) * scale + d0;
        const uint32_t pos = (uint32_t)(c0 + s0 * (ca + w));
        if (map) map[(long long)node * g.s[0] * g.s[1] + pos] = val;
        if (out) {
          const float key = (quantum > 0.f) ? rintf(val * (1.f / quantum)) : val;
          if (!(key >= c[KT - 1].key)) continue;
          Cand cd;
          cd.key = key;
          cd.val = val;
          cd.pos = pos;
          cd.h = tie_hash(ni.seedbits, pos);
          if (!cand_better(cd, c[KT - 1])) continue;
#pragma unroll
          for (int i = 0; i < KT; ++i) {
            if (cand_better(cd, c[i])) {
              Cand tt = c[i];
              c[i] = cd;
              cd = tt;
            }
          }
        }
      }
    }
  }
  if (out) block_topk<KT>(c, K, out, red);
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
                     const NodeInfo* info, const int* nbOff, const float* nbVal, float2* CT, hipStream_t st) {
  const int slice = (g.h0 + f1k::NXCD - 1) / f1k::NXCD;
  const int chunks = (nNodes + npb - 1) / npb;
  const long long blocks = (long long)f1k::NXCD * slice * chunks;
  hipLaunchKernelGGL(k_pencil1k, dim3((unsigned)blocks), dim3(128), 0, st, g, YT, tw1k, nNodes, npb, info, nbOff, nbVal,
                     CT, ct1k_stride(g.h0));
}

int rows1k_tiles() { return (1000 + f1k::TR - 1) / f1k::TR; }

hipError_t set_rows1k_lds() { return hipSuccess; }

void launch_rows1k(const ScanGeom& g, const float2* tw1k, const NodeInfo* info, const float2* CT, int nNodes, int K,
                   Cand* cand, float* map, hipStream_t st) {
  const int pairs = rows1k_tiles();
  dim3 grid((unsigned)pairs, (unsigned)nNodes);
  const long long cs = ct1k_stride(g.h0);
  if (K <= 2)
    hipLaunchKernelGGL(k_rows1k_pair<2>, grid, dim3(128), 0, st, g, tw1k, info, CT, cs, K, cand, pairs, map);
  else if (K <= 4)
    hipLaunchKernelGGL(k_rows1k_pair<4>, grid, dim3(128), 0, st, g, tw1k, info, CT, cs, K, cand, pairs, map);
  else if (K <= 8)
    hipLaunchKernelGGL(k_rows1k_pair<8>, grid, dim3(128), 0, st, g, tw1k, info, CT, cs, K, cand, pairs, map);
  else
    hipLaunchKernelGGL(k_rows1k_pair<16>, grid, dim3(128), 0, st, g, tw1k, info, CT, cs, K, cand, pairs, map);
}

}  // namespace qsmi
