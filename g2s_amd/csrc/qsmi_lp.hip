// Learned proposal for QuickSampling on gfx950 (the "learned-proposal QS"
// form): a pattern encoder proposes a shortlist of TI positions for each path
// node, the shortlist is rescored exactly, and the full FFT scan stays the
// fallback and the judge.
//
//  * Two MLP towers, 240 -> 256 -> 256 -> 256 -> 128 (ReLU), fp32 master
//    weights; every product runs on MFMA (v_mfma_f32_32x32x16_bf16: operands
//    rounded to bf16, fp32 accumulation):
//      key tower:   a TI position p: 80 slots, slot i = canonical offset d_i
//                   (the 80 nearest non-centre offsets of the sorted offset
//                   list): (d_i/16, (y(p + d_i) - mu)/sigma);
//      query tower: a node's neighbours in the same slots: the neighbour at
//                   offset d_i fills slot i with (d_i/16, (x - mu)/sigma);
//                   neighbours off the stencil are left to the exact rescoring,
//                   empty slots are zero.
//    In this layout the exact mismatch over the stencil is a bilinear form of
//    the two inputs, which the towers approximate.
//  * score(node, p) = <q(node), k(p)>; training minimises the softmax
//    cross-entropy of the exact best position against in-batch and random
//    negatives (InfoNCE), Adam on the fp32 weights.
//  * Proposal per node: a per-node score threshold from a fixed sample of
//    key positions, one scoring GEMM whose epilogue appends the positions
//    above the threshold, then the exact mismatch of every proposed position
//    (direct sum over the neighbours) and the device selection rule.
//
// Lane layout of v_mfma_f32_32x32x16_bf16 (64 lanes, l = lane):
//   A operand: A[l % 32][8 (l / 32) + e], e = 0..7
//   B operand: B[8 (l / 32) + e][l % 32]
//   result r:  C[8 (r / 4) + 4 (l / 32) + r % 4][l % 32], r = 0..15
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <algorithm>
#include <cstring>
#include <random>
#include <vector>

#include "qsmi_internal.hpp"
#include "qsmi_kernels.hpp"
#include "qsmi_select.hpp"

namespace qsmi {
namespace lp {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int NB = 80;    // neighbour slots of the encoders
constexpr int IN = 3 * NB;  // 240 input features
constexpr int HID = 256;
constexpr int EMB = 128;
constexpr float OFFSET_SCALE = 1.f / 16.f;
// bias offsets of the four layers inside one tower's parameters
__device__ __forceinline__ long long boff_dev(int l) {
  return l == 0 ? 240LL * 256 : l == 1 ? 240LL * 256 + 256 + 256LL * 256
                              : l == 2 ? 240LL * 256 + 256 + 2 * (256LL * 256 + 256) - 256
                                       : 240LL * 256 + 256 + 2 * (256LL * 256 + 256) + 256LL * 128;
}
constexpr int SLOT_R = 8;                  // slot map covers |d0|, |d1| <= 8
constexpr int SLOT_W = 2 * SLOT_R + 1;     // 17 x 17 map: canonical slot or -1

constexpr int NSEG = 16;  // shortlist segments per node (one counter each)

constexpr int GT = 64;   // block tile (M and N)
constexpr int GK = 32;   // K step
constexpr int LDT = GK + 8;  // LDS row stride (bf16), 16-byte aligned fragments

// ---------------------------------------------------------------------------
// GEMM on MFMA: C[M,N] = op(A)[M,K] op(B)[K,N] (fp32 storage).
//   TA: A stored [K][M] (row stride lda), else [M][K]
//   TB: B stored [N][K] (row stride ldb), else [K][N]
// Epilogue EPI_STORE: C = (acc ? C : 0) + AB (+ bias[n]) (ReLU when relu);
//          EPI_SELECT: positions n with AB[m][n] >= thr[m] are appended to
//          list[m][...] (counter cnt[m], capacity cap; overflow only counts).
// 256 threads: four wavefronts, each a 32 x 32 quarter of the 64 x 64 tile.
// ---------------------------------------------------------------------------
enum { EPI_STORE = 0, EPI_SELECT = 1 };

struct GemmArgs {
  int M, N, K;
  const float* A;
  int lda;
  const float* B;
  int ldb;
  float* C;
  int ldc;
  int accumulate;
  const float* bias;
  int relu;
  // select epilogue
  const float* thr;
  uint32_t* list;
  int* cnt;
  int cap;
};

template <bool TA, bool TB, int EPI>
__global__ void __launch_bounds__(256) k_gemm(GemmArgs g) {
  __shared__ __attribute__((aligned(16))) __bf16 As[GT * LDT];
  __shared__ __attribute__((aligned(16))) __bf16 Bs[GT * LDT];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int bm = blockIdx.y * GT, bn = blockIdx.x * GT;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  f32x16 acc;
#pragma unroll
  for (int r = 0; r < 16; ++r) acc[r] = 0.f;
  for (int k0 = 0; k0 < g.K; k0 += GK) {
    // A tile -> As[m][k]
    if (!TA) {
      const int m = tid >> 2, kb = (tid & 3) * 8;
      const int gm = bm + m;
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int gk = k0 + kb + e;
        const float v = (gm < g.M && gk < g.K) ? g.A[(long long)gm * g.lda + gk] : 0.f;
        As[m * LDT + kb + e] = (__bf16)v;
      }
    } else {
      const int k = tid >> 3, mb = (tid & 7) * 8;
      const int gk = k0 + k;
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int gm = bm + mb + e;
        const float v = (gm < g.M && gk < g.K) ? g.A[(long long)gk * g.lda + gm] : 0.f;
        As[(mb + e) * LDT + k] = (__bf16)v;
      }
    }
    // B tile -> Bs[n][k]
    if (TB) {
      const int n = tid >> 2, kb = (tid & 3) * 8;
      const int gn = bn + n;
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int gk = k0 + kb + e;
        const float v = (gn < g.N && gk < g.K) ? g.B[(long long)gn * g.ldb + gk] : 0.f;
        Bs[n * LDT + kb + e] = (__bf16)v;
      }
    } else {
      const int k = tid >> 3, nb = (tid & 7) * 8;
      const int gk = k0 + k;
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int gn = bn + nb + e;
        const float v = (gn < g.N && gk < g.K) ? g.B[(long long)gk * g.ldb + gn] : 0.f;
        Bs[(nb + e) * LDT + k] = (__bf16)v;
      }
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < GK / 16; ++s) {
      const bf16x8 a = *reinterpret_cast<const bf16x8*>(&As[(wm + (lane & 31)) * LDT + 16 * s + 8 * (lane >> 5)]);
      const bf16x8 b = *reinterpret_cast<const bf16x8*>(&Bs[(wn + (lane & 31)) * LDT + 16 * s + 8 * (lane >> 5)]);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
    }
    __syncthreads();
  }
  const int gn = bn + wn + (lane & 31);
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int gm = bm + wm + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
    if (gm >= g.M || gn >= g.N) continue;
    float v = acc[r];
    if (EPI == EPI_STORE) {
      float* c = g.C + (long long)gm * g.ldc + gn;
      if (g.accumulate) v += *c;
      if (g.bias) v += g.bias[gn];
      if (g.relu) v = fmaxf(v, 0.f);
      *c = v;
    } else {
      if (v >= g.thr[gm]) {
        const int slot = atomicAdd(&g.cnt[gm], 1);
        if (slot < g.cap) g.list[(long long)gm * g.cap + slot] = (uint32_t)gn;
      }
    }
  }
}

void gemm(bool ta, bool tb, const GemmArgs& a, hipStream_t st) {
  dim3 grid((a.N + GT - 1) / GT, (a.M + GT - 1) / GT);
  if (!ta && !tb) hipLaunchKernelGGL((k_gemm<false, false, EPI_STORE>), grid, dim3(256), 0, st, a);
  else if (!ta && tb) hipLaunchKernelGGL((k_gemm<false, true, EPI_STORE>), grid, dim3(256), 0, st, a);
  else if (ta && !tb) hipLaunchKernelGGL((k_gemm<true, false, EPI_STORE>), grid, dim3(256), 0, st, a);
  else hipLaunchKernelGGL((k_gemm<true, true, EPI_STORE>), grid, dim3(256), 0, st, a);
}

void gemm_select(const GemmArgs& a, hipStream_t st) {
  dim3 grid((a.N + GT - 1) / GT, (a.M + GT - 1) / GT);
  hipLaunchKernelGGL((k_gemm<false, true, EPI_SELECT>), grid, dim3(256), 0, st, a);
}

GemmArgs gargs(int M, int N, int K, const float* A, int lda, const float* B, int ldb, float* C, int ldc) {
  GemmArgs a;
  memset(&a, 0, sizeof(a));
  a.M = M;
  a.N = N;
  a.K = K;
  a.A = A;
  a.lda = lda;
  a.B = B;
  a.ldb = ldb;
  a.C = C;
  a.ldc = ldc;
  return a;
}

// ---------------------------------------------------------------------------
// element-wise kernels
// ---------------------------------------------------------------------------
// dY *= (Y > 0) (ReLU backward, Y = the layer's output after ReLU)
__global__ void k_relu_bwd(float* __restrict__ dy, const float* __restrict__ y, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && !(y[i] > 0.f)) dy[i] = 0.f;
}

// db[n] (+)= sum_r dY[r][n]
__global__ void k_colsum(const float* __restrict__ dy, int rows, int cols, float* __restrict__ db) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= cols) return;
  float s = 0.f;
  for (int r = 0; r < rows; ++r) s += dy[(long long)r * cols + c];
  db[c] = s;
}

// Softmax cross-entropy of row b of G * invTau against column b (one block
// per row); dG = invTau * (softmax - onehot) / B; loss[b] = lse - logit_b.
__global__ void __launch_bounds__(256) k_softmax_ce(const float* __restrict__ Gm, int B, int C, float invTau,
                                                    float* __restrict__ dG, float* __restrict__ loss) {
  __shared__ float red[256];
  const int b = blockIdx.x, t = threadIdx.x;
  const float* row = Gm + (long long)b * C;
  float mx = -INFINITY;
  for (int c = t; c < C; c += 256) mx = fmaxf(mx, row[c] * invTau);
  red[t] = mx;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) red[t] = fmaxf(red[t], red[t + s]);
    __syncthreads();
  }
  mx = red[0];
  __syncthreads();
  float sum = 0.f;
  for (int c = t; c < C; c += 256) sum += expf(row[c] * invTau - mx);
  red[t] = sum;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) red[t] += red[t + s];
    __syncthreads();
  }
  sum = red[0];
  const float lse = mx + logf(sum);
  for (int c = t; c < C; c += 256) {
    const float p = expf(row[c] * invTau - lse);
    dG[(long long)b * C + c] = invTau * (p - (c == b ? 1.f : 0.f)) / (float)B;
  }
  if (t == 0) loss[b] = lse - row[b] * invTau;
}

__global__ void k_adam(float* __restrict__ w, const float* __restrict__ gr, float* __restrict__ m,
                       float* __restrict__ v, long long n, float lr, float b1, float b2, float eps, float c1,
                       float c2) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float g = gr[i];
  const float mi = b1 * m[i] + (1.f - b1) * g;
  const float vi = b2 * v[i] + (1.f - b2) * g * g;
  m[i] = mi;
  v[i] = vi;
  w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// ---------------------------------------------------------------------------
// features
// ---------------------------------------------------------------------------
// Key features of TI positions (pos == nullptr: positions first..first+n-1):
// slot i = canonical offset d_i: (d0/16, d1/16, (y(p + d) - mu)/sigma),
// all zero when p + d leaves the TI.
__global__ void k_key_features(int n, const uint32_t* __restrict__ pos, long long first, const float* __restrict__ ti,
                               int nVar, int s0, int s1, const int* __restrict__ canon, float mu, float invSigma,
                               float* __restrict__ out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)n * NB) return;
  const long long r = i / NB;
  const int slot = (int)(i - r * NB);
  const long long p = pos ? (long long)pos[r] : first + r;
  const int x0 = (int)(p % s0), x1 = (int)(p / s0);
  const int d0 = canon[2 * slot], d1 = canon[2 * slot + 1];
  const int t0 = x0 + d0, t1 = x1 + d1;
  float* o = out + r * IN + 3 * slot;
  if (t0 >= 0 && t0 < s0 && t1 >= 0 && t1 < s1) {
    o[0] = d0 * OFFSET_SCALE;
    o[1] = d1 * OFFSET_SCALE;
    o[2] = (ti[((long long)t1 * s0 + t0) * nVar] - mu) * invSigma;
  } else {
    o[0] = o[1] = o[2] = 0.f;
  }
}

__device__ __forceinline__ int slot_of(const int* __restrict__ slotMap, int d0, int d1) {
  if (d0 < -SLOT_R || d0 > SLOT_R || d1 < -SLOT_R || d1 > SLOT_R) return -1;
  return slotMap[(d1 + SLOT_R) * SLOT_W + (d0 + SLOT_R)];
}

// Query features of recorded pairs (one block per row): offsets
// [pair][NB][2], values [pair][NB] (NaN = empty), placed in canonical slots.
__global__ void __launch_bounds__(128) k_pair_features(const int* __restrict__ idx, const int* __restrict__ offs,
                                                      const float* __restrict__ vals,
                                                      const int* __restrict__ slotMap, float mu, float invSigma,
                                                      float* __restrict__ out) {
  const long long r = blockIdx.x;
  const long long q = idx[r];
  float* o = out + r * IN;
  for (int i = threadIdx.x; i < IN; i += blockDim.x) o[i] = 0.f;
  __syncthreads();
  for (int j = threadIdx.x; j < NB; j += blockDim.x) {
    const float v = vals[q * NB + j];
    if (isnan(v)) continue;
    const int d0 = offs[(q * NB + j) * 2], d1 = offs[(q * NB + j) * 2 + 1];
    const int s = slot_of(slotMap, d0, d1);
    if (s < 0) continue;
    o[3 * s] = d0 * OFFSET_SCALE;
    o[3 * s + 1] = d1 * OFFSET_SCALE;
    o[3 * s + 2] = (v - mu) * invSigma;
  }
}

// Query features of a path wave (one block per node): neighbours 1..nb-1
// (entry 0 is the simulated cell itself), values from the simulation grid.
__global__ void __launch_bounds__(128) k_wave_features(const NodeInfo* __restrict__ info, const int* __restrict__ nbOff,
                                                      const float* __restrict__ di, int nVar, int sd0,
                                                      const int* __restrict__ slotMap, float mu, float invSigma,
                                                      float* __restrict__ out, float* __restrict__ cover,
                                                      int* __restrict__ exCnt) {
  __shared__ int covered;
  const int node = blockIdx.x;
  if (node == 0 && threadIdx.x == 0) *exCnt = 0;  // k_lp_flags of this wave compacts into it
  const NodeInfo ni = info[node];
  float* o = out + (long long)node * IN;
  for (int i = threadIdx.x; i < IN; i += blockDim.x) o[i] = 0.f;
  if (threadIdx.x == 0) covered = 0;
  __syncthreads();
  const int c0 = (int)(ni.cell % (unsigned)sd0), c1 = (int)(ni.cell / (unsigned)sd0);
  for (int j = 1 + threadIdx.x; j < ni.nb; j += blockDim.x) {
    const int d0 = nbOff[(ni.nbStart + j) * 2], d1 = nbOff[(ni.nbStart + j) * 2 + 1];
    const int s = slot_of(slotMap, d0, d1);
    if (s < 0) continue;
    const float v = di[((long long)(c1 + d1) * sd0 + (c0 + d0)) * nVar];
    o[3 * s] = d0 * OFFSET_SCALE;
    o[3 * s + 1] = d1 * OFFSET_SCALE;
    o[3 * s + 2] = (v - mu) * invSigma;
    atomicAdd(&covered, 1);
  }
  __syncthreads();
  if (threadIdx.x == 0) cover[node] = ni.nb > 1 ? (float)covered / (float)(ni.nb - 1) : 0.f;
}

// Stencil coverage of every path node before the loop (one wavefront per
// node): ok[i] = 1 when at least minCover of the node's neighbours 1..nb-1
// fall on the key stencil (the rule k_wave_features + k_lp_flags apply), so
// waves without such a node skip the proposal entirely.
__global__ void __launch_bounds__(64) k_cover_all(int nPath, int n, const int* __restrict__ nbIdx,
                                                 const int* __restrict__ nbCnt, const int* __restrict__ offs,
                                                 const int* __restrict__ slotMap, float minCover,
                                                 unsigned char* __restrict__ ok) {
  const int i = blockIdx.x;
  if (i >= nPath) return;
  const int nb = nbCnt[i];
  int c = 0;
  for (int j = 1 + (int)threadIdx.x; j < nb; j += 64) {
    const int o = nbIdx[(long long)i * n + j];
    c += slot_of(slotMap, offs[2 * o], offs[2 * o + 1]) >= 0 ? 1 : 0;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off);
  if (threadIdx.x == 0) ok[i] = (nb > 1 && (float)c / (float)(nb - 1) >= minCover) ? 1 : 0;
}

// Per-node score threshold from the sampled scores (one block per row of
// n = 8192 floats): every thread keeps the 4 largest of its 32 values and
// thr = the r-th largest of the 1024 survivors (r <= 1024), found by a radix
// select over their order-preserving bit patterns (four 8-bit digits, one
// LDS histogram per digit). When a thread held more than 4 values above it,
// thr is a lower bound of the r-th largest, which only lengthens the
// shortlist.
__global__ void __launch_bounds__(256) k_row_kth(const float* __restrict__ Gm, int n, int r,
                                                const float* __restrict__ cover, float minCover,
                                                float* __restrict__ thr) {
  __shared__ int hist[256];
  __shared__ uint32_t sel[2];  // [0] prefix of the selected digits, [1] rank left
  if (!(cover[blockIdx.x] >= minCover)) {  // routed to the exact scan: propose nothing
    if (threadIdx.x == 0) thr[blockIdx.x] = INFINITY;
    return;
  }
  const int tid = threadIdx.x, lane = tid & 63;
  const float* row = Gm + (long long)blockIdx.x * n;
  float m0 = -INFINITY, m1 = -INFINITY, m2 = -INFINITY, m3 = -INFINITY;
  for (int i = tid; i < n; i += 256) {
    float x = row[i];
    float h = fmaxf(m0, x); x = fminf(m0, x); m0 = h;
    h = fmaxf(m1, x); x = fminf(m1, x); m1 = h;
    h = fmaxf(m2, x); x = fminf(m2, x); m2 = h;
    m3 = fmaxf(m3, x);
  }
  const uint32_t k4[4] = {key_bits(m0), key_bits(m1), key_bits(m2), key_bits(m3)};
  if (tid == 0) {
    sel[0] = 0u;
    sel[1] = (uint32_t)min(max(r, 1), 1024);
  }
  for (int d = 3; d >= 0; --d) {
    const int sh = 8 * d;
    const uint32_t hiMask = d == 3 ? 0u : (0xFFFFFFFFu << (sh + 8));
    hist[tid] = 0;
    __syncthreads();
    const uint32_t pre = sel[0];
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if ((k4[q] & hiMask) == pre) atomicAdd(&hist[(k4[q] >> sh) & 255], 1);
    __syncthreads();
    if (tid < 64) {
      // digits from the top: lane l holds digits 255-4l .. 252-4l
      const int c0 = hist[255 - 4 * lane], c1 = hist[254 - 4 * lane], c2 = hist[253 - 4 * lane],
                c3 = hist[252 - 4 * lane];
      const int sum = c0 + c1 + c2 + c3;
      int incl = sum;
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        const int v = __shfl_up(incl, off);
        if (lane >= off) incl += v;
      }
      const int rem = (int)sel[1];
      const int excl = incl - sum;
      const bool mine = excl < rem && incl >= rem;
      if (mine) {  // exactly one lane
        int acc = excl, digit = 255 - 4 * lane;
        const int cs[4] = {c0, c1, c2, c3};
        for (int q = 0; q < 4; ++q) {
          if (acc + cs[q] >= rem) {
            digit = 255 - 4 * lane - q;
            break;
          }
          acc += cs[q];
        }
        sel[0] = pre | ((uint32_t)digit << sh);
        sel[1] = (uint32_t)(rem - acc);
      }
    }
    __syncthreads();
  }
  if (tid == 0) thr[blockIdx.x] = key_of_bits(sel[0]);
}

// fp32 -> bf16 copy (the keys of the scoring product)
__global__ void k_to_bf16(const float* __restrict__ src, __bf16* __restrict__ dst, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = (__bf16)src[i];
}

// ---------------------------------------------------------------------------
// Scoring of every TI position for the nodes of a wave with the threshold
// epilogue: S[m][n] = <Q[m], keys[n]> (bf16 operands, fp32 accumulation);
// positions with S >= thr[m] are appended to segment (block % NSEG) of
// list[m] (counter cnt[m][seg], segment capacity cap / NSEG).
// Q (M <= 256 rows) sits in LDS as bf16; each wavefront takes 64 keys at a
// time, holds their B fragments for all of K = 128 in registers and sweeps
// the M tiles.
// Appends are staged in LDS as (row << 24 | position) (positions < 2^24),
// each wavefront in its own region of the stage with a wave-uniform fill
// count: a ballot per accumulator register finds the hits, the rare
// registers with a hit write their positions at the count plus the lane's
// rank in the ballot. After a block iteration (four tiles) in which a region
// passed half full, and after the last one, the block counts its staged hits
// per row, reserves each row's room in the global list with one atomic per
// row, and scatters. A full region spills single hits straight to the list.
// ---------------------------------------------------------------------------
constexpr int kStageW = 256;  // staged hits per wavefront region

__device__ __forceinline__ void score_emit(uint32_t e, uint32_t* __restrict__ list, int* __restrict__ cnt, int cap,
                                           int seg, int segCap) {
  const int row = (int)(e >> 24);
  const int slot = atomicAdd(&cnt[row * NSEG + seg], 1);
  if (slot < segCap) list[(long long)row * cap + seg * segCap + slot] = e & 0xFFFFFFu;
}

__global__ void __launch_bounds__(256) k_score(int M, long long N, const float* __restrict__ Q,
                                              const __bf16* __restrict__ keys, const float* __restrict__ thr,
                                              uint32_t* __restrict__ list, int* __restrict__ cnt, int cap) {
  constexpr int QLD = EMB + 8;  // bf16 row stride of Q in LDS (16-byte aligned fragments)
  __shared__ __attribute__((aligned(16))) __bf16 qs[256 * QLD];
  __shared__ __attribute__((aligned(16))) float th[256];
  __shared__ int rowCnt[256];
  __shared__ int rowBase[256];
  __shared__ uint32_t stg[4 * kStageW];
  __shared__ int stgW[4];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int mt = (M + 31) / 32;
  for (int i = tid; i < mt * 32 * EMB; i += 256) {
    const int m = i / EMB, k = i - m * EMB;
    qs[m * QLD + k] = (__bf16)(m < M ? Q[(long long)m * EMB + k] : 0.f);
  }
  for (int i = tid; i < 256; i += 256) {
    th[i] = i < M ? thr[i] : INFINITY;
    rowCnt[i] = 0;
  }
  __syncthreads();
  const long long tiles = (N + 63) / 64;
  const int seg = blockIdx.x % NSEG, segCap = cap / NSEG;
  const uint64_t below = (1ull << lane) - 1ull;  // lanes ranked before this one
  uint32_t* myStg = stg + wave * kStageW;
  int nW = 0;  // wave-uniform fill of this wavefront's region
  // block-uniform loop: every wavefront runs the same iterations (barriers)
  const long long step = (long long)gridDim.x * 4;
  const bf16x8 z = {};
  for (long long t0 = (long long)blockIdx.x * 4; t0 < tiles; t0 += step) {
    const long long tile = t0 + wave;
    if (tile < tiles) {
      const long long n0 = tile * 64;
      const long long na = n0 + (lane & 31), nb2 = na + 32;
      bf16x8 b0[8], b1[8];
#pragma unroll
      for (int s = 0; s < 8; ++s) {
        b0[s] = na < N ? *reinterpret_cast<const bf16x8*>(keys + na * EMB + 16 * s + 8 * (lane >> 5)) : z;
        b1[s] = nb2 < N ? *reinterpret_cast<const bf16x8*>(keys + nb2 * EMB + 16 * s + 8 * (lane >> 5)) : z;
      }
      // lanes whose key exists (wave masks: the compares below stay scalar)
      const uint64_t inA = __ballot(na < N), inB = __ballot(nb2 < N);
      // M tiles software-pipelined: the MFMAs of tile m + 1 are issued before
      // the threshold epilogue of tile m, which then overlaps them
      auto mfma_tile = [&](int m, f32x16& c0, f32x16& c1) {
#pragma unroll
        for (int r = 0; r < 16; ++r) c0[r] = c1[r] = 0.f;
#pragma unroll
        for (int s = 0; s < 8; ++s) {
          const bf16x8 a = *reinterpret_cast<const bf16x8*>(&qs[(m * 32 + (lane & 31)) * QLD + 16 * s + 8 * (lane >> 5)]);
          c0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b0[s], c0, 0, 0, 0);
          c1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b1[s], c1, 0, 0, 0);
        }
      };
      f32x16 c0, c1, d0, d1;
      mfma_tile(0, c0, c1);
      for (int m = 0; m < mt; ++m) {
        if (m + 1 < mt) mfma_tile(m + 1, d0, d1);
        // thresholds of this lane's 16 rows: four runs of four consecutive rows
        float4 tq[4];
#pragma unroll
        for (int q = 0; q < 4; ++q)
          tq[q] = *reinterpret_cast<const float4*>(&th[m * 32 + 8 * q + 4 * (lane >> 5)]);
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const float4 t4 = tq[r >> 2];
          const float t = (r & 3) == 0 ? t4.x : (r & 3) == 1 ? t4.y : (r & 3) == 2 ? t4.z : t4.w;
          const uint64_t ma = __ballot(c0[r] >= t) & inA, mb = __ballot(c1[r] >= t) & inB;
          if (ma | mb) {  // wave-uniform: rare
            const int row = m * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
            const int ca = __popcll(ma);
            const int ia = nW + __popcll(ma & below), ib = nW + ca + __popcll(mb & below);
            const bool ha = (ma >> lane) & 1ull, hb = (mb >> lane) & 1ull;
            if (ha) {
              const uint32_t e = ((uint32_t)row << 24) | (uint32_t)na;
              if (ia < kStageW) myStg[ia] = e;
              else score_emit(e, list, cnt, cap, seg, segCap);
            }
            if (hb) {
              const uint32_t e = ((uint32_t)row << 24) | (uint32_t)nb2;
              if (ib < kStageW) myStg[ib] = e;
              else score_emit(e, list, cnt, cap, seg, segCap);
            }
            nW += ca + __popcll(mb);
          }
        }
        c0 = d0;
        c1 = d1;
      }
    }
    // flush when a region passed half full, and after the last iteration:
    // per-row counts, one global reservation per row, scatter
    if (lane == 0) stgW[wave] = min(nW, kStageW);
    __syncthreads();
    const int n0w = stgW[0], n1w = stgW[1], n2w = stgW[2], n3w = stgW[3];
    const bool last = t0 + step >= tiles;
    const int nMax = max(max(n0w, n1w), max(n2w, n3w));
    if (nMax >= kStageW / 2 || (last && nMax > 0)) {
      const int nAll[4] = {n0w, n1w, n2w, n3w};
      for (int w = 0; w < 4; ++w)
        for (int i = tid; i < nAll[w]; i += 256) atomicAdd(&rowCnt[stg[w * kStageW + i] >> 24], 1);
      __syncthreads();
      {
        const int c = rowCnt[tid];
        rowBase[tid] = c ? atomicAdd(&cnt[tid * NSEG + seg], c) : 0;
        rowCnt[tid] = 0;
      }
      __syncthreads();
      for (int w = 0; w < 4; ++w)
        for (int i = tid; i < nAll[w]; i += 256) {
          const uint32_t e = stg[w * kStageW + i];
          const int row = (int)(e >> 24);
          const int slot = rowBase[row] + atomicAdd(&rowCnt[row], 1);
          if (slot < segCap) list[(long long)row * cap + seg * segCap + slot] = e & 0xFFFFFFu;
        }
      __syncthreads();
      rowCnt[tid] = 0;
      nW = 0;
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Fused query-tower forward for inference: 32 rows per block (four
// wavefronts), activations in LDS as bf16 (exactly the operands the MFMA
// GEMM path rounds to), weights read as transposed bf16 copies
// Wt[n][k] (k padded to 256) so every B fragment is one 16-byte load.
// Same products, same K order and same rounding points as the layer-by-layer
// GEMMs.
// ---------------------------------------------------------------------------
constexpr int KP = 256;        // padded K of every layer
constexpr int ALD = KP + 8;    // LDS row stride (bf16)

__global__ void k_wt_bf16(const float* __restrict__ W, int din, int dout, __bf16* __restrict__ Wt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= dout * KP) return;
  const int n = i / KP, k = i - n * KP;
  Wt[i] = (__bf16)(k < din ? W[(long long)k * dout + n] : 0.f);
}

__global__ void __launch_bounds__(256) k_mlp_fwd(int M, const float* __restrict__ X, const __bf16* __restrict__ Wt,
                                                const float* __restrict__ W, float* __restrict__ E) {
  __shared__ __attribute__((aligned(16))) __bf16 act[2][32 * ALD];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int r0 = blockIdx.x * 32;
  for (int i = tid; i < 32 * KP; i += 256) {
    const int r = i / KP, k = i - r * KP;
    act[0][r * ALD + k] = (__bf16)((r0 + r < M && k < IN) ? X[(long long)(r0 + r) * IN + k] : 0.f);
  }
  __syncthreads();
  int cur = 0;
  long long wtOff = 0;
  for (int l = 0; l < 4; ++l) {
    const int dout = l < 3 ? HID : EMB;
    const float* bias = W + boff_dev(l);
    const __bf16* wt = Wt + wtOff;
    const int subPerWave = dout / 32 / 4;  // 2 for 256 outputs, 1 for 128
    for (int sub = 0; sub < subPerWave; ++sub) {
      const int n0 = (wave * subPerWave + sub) * 32;
      f32x16 c;
#pragma unroll
      for (int r = 0; r < 16; ++r) c[r] = 0.f;
      for (int s = 0; s < KP / 16; ++s) {
        const bf16x8 a = *reinterpret_cast<const bf16x8*>(&act[cur][(lane & 31) * ALD + 16 * s + 8 * (lane >> 5)]);
        const bf16x8 b = *reinterpret_cast<const bf16x8*>(wt + (long long)(n0 + (lane & 31)) * KP + 16 * s + 8 * (lane >> 5));
        c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
      }
      const int n = n0 + (lane & 31);
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int row = 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        float v = c[r] + bias[n];
        if (l < 3) {
          act[cur ^ 1][row * ALD + n] = (__bf16)fmaxf(v, 0.f);
        } else if (r0 + row < M) {
          E[(long long)(r0 + row) * EMB + n] = v;
        }
      }
    }
    wtOff += (long long)dout * KP;
    cur ^= 1;
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Exact rescoring of the proposed positions. Stage 1 (block = 256 proposals
// of one node, grid (chunks, nodes)): each wavefront evaluates one position
// at a time with its lanes over the node's neighbours (a position's
// neighbours share a few cache lines) -> ebuf. Stage 2 (block per node): the
// K best by the device rule -> cand[node][K], best[node] = the largest E
// (-inf when no proposed position is valid).
// ---------------------------------------------------------------------------
constexpr int kRescoreChunk = 256;

__global__ void __launch_bounds__(256) k_rescore(const NodeInfo* __restrict__ info, const int* __restrict__ nbOff,
                                                const float* __restrict__ nbVal, int V, int yy, const int* __restrict__ act,
                                                const float* __restrict__ sub, long long S, int s0,
                                                const uint32_t* __restrict__ list, const int* __restrict__ cnt, int cap,
                                                float* __restrict__ ebuf) {
  __shared__ int doff[128];
  __shared__ float dval[128 * 4];
  const int node = blockIdx.y, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int c0 = blockIdx.x * kRescoreChunk;
  const int m = cap;  // every entry of the segmented list; empty ones score -inf
  const int segCap = cap / NSEG;
  const NodeInfo ni = info[node];
  int any = 0;
  for (int g = 0; g < NSEG; ++g) any |= cnt[node * NSEG + g];
  if (!any) {  // nothing proposed (e.g. routed to the exact scan)
    float* eb0 = ebuf + (long long)node * cap;
    for (int c = c0 + tid; c < min(m, c0 + kRescoreChunk); c += 256) eb0[c] = -INFINITY;
    return;
  }
  const int nb = min(ni.nb, 128);
  for (int j = tid; j < nb; j += 256) {
    doff[j] = nbOff[(ni.nbStart + j) * 2] + s0 * nbOff[(ni.nbStart + j) * 2 + 1];
    for (int a = 0; a < V && a < 4; ++a) dval[j * 4 + a] = nbVal[(long long)(ni.nbStart + j) * V + a];
  }
  __syncthreads();
  const bool usable = ni.nb > 0 && ni.nb <= 128 && V <= 4 && ni.lo[0] <= ni.hi[0] && ni.lo[1] <= ni.hi[1];
  float* eb = ebuf + (long long)node * cap;
  const int cEnd = min(m, c0 + kRescoreChunk);
  // 8 lanes per position, eight positions per wavefront at a time
  const int g = lane >> 3, gl = lane & 7;
  const int* segCnt = cnt + node * NSEG;
  for (int c = c0 + wave * 8 + g; c - g < cEnd; c += 32) {
    const bool inRange = c < cEnd;
    const bool have = inRange && (c % segCap) < segCnt[c / segCap];
    const uint32_t p = have ? list[(long long)node * cap + c] : 0u;
    const int p0 = (int)(p % (uint32_t)s0), p1 = (int)(p / (uint32_t)s0);
    const bool valid = have && usable && p0 >= ni.lo[0] && p0 <= ni.hi[0] && p1 >= ni.lo[1] && p1 <= ni.hi[1];
    float e = 0.f;
    if (valid) {
      if (yy) {  // one continuous variable: images (y^2, y), one load per neighbour
        const float* ys = sub + S + p;
        float y[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) {
          const int j = gl + 8 * u;
          y[u] = j < nb ? ys[doff[j]] : 0.f;
        }
#pragma unroll
        for (int u = 0; u < 16; ++u) {
          const int j = gl + 8 * u;
          if (j < nb) e += dval[j * 4] * (y[u] * y[u]) + dval[j * 4 + 1] * y[u];
        }
      } else {
        for (int j = gl; j < nb; j += 8) {
          const long long t = (long long)p + doff[j];
          for (int a = 0; a < V; ++a) e += dval[j * 4 + a] * sub[(long long)act[a] * S + t];
        }
      }
    }
#pragma unroll
    for (int off = 4; off > 0; off >>= 1) e += __shfl_xor(e, off);
    if (gl == 0 && inRange) eb[c] = valid ? e + ni.delta0 : -INFINITY;
  }
}

template <int KT>
__global__ void __launch_bounds__(256) k_rescore_select(const NodeInfo* __restrict__ info,
                                                       const uint32_t* __restrict__ list, const int* __restrict__ cnt,
                                                       int cap, int K, const float* __restrict__ ebuf,
                                                       Cand* __restrict__ cand, float* __restrict__ best) {
  __shared__ TopkScratch red;
  __shared__ float bestW[4];
  const int node = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const NodeInfo ni = info[node];
  const int m = cap;  // segmented list: empty entries hold -inf
  const float* eb = ebuf + (long long)node * cap;
  float mk[KT];
#pragma unroll
  for (int i = 0; i < KT; ++i) mk[i] = -INFINITY;
  for (int c = tid; c < m; c += 256) {
    float x = eb[c];
#pragma unroll
    for (int i = 0; i < KT; ++i) {
      const float hi = fmaxf(mk[i], x);
      x = fminf(mk[i], x);
      mk[i] = hi;
    }
  }
  const float tau2 = fmaxf(block_kth_key<KT>(mk, K, &red), -3.402823466e38f);
  RankList<KT> rl;
  rl.init();
  float bestE = -INFINITY;
  for (int c = tid; c < m; c += 256) {
    const float e = eb[c];
    bestE = fmaxf(bestE, e);
    if (e >= tau2) {
      const uint32_t p = list[(long long)node * cap + c];
      rl.insert(e, tie_hash(ni.seedbits, p), p, e);
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) bestE = fmaxf(bestE, __shfl_xor(bestE, off));
  if (lane == 0) bestW[wave] = bestE;
  rl.block_extract(K, cand + (long long)node * K, &red);  // its barriers order bestW
  if (tid == 0) {
    float b = bestW[0];
    for (int w = 1; w < 4; ++w) b = fmaxf(b, bestW[w]);
    best[node] = b;
  }
}

}  // namespace lp
}  // namespace qsmi

// ===========================================================================
// host side
// ===========================================================================
namespace qsmi {
namespace lp {

constexpr int kDin[4] = {IN, HID, HID, HID};
constexpr int kDout[4] = {HID, HID, HID, EMB};
constexpr long long kWoff[4] = {0, 240LL * 256 + 256, 240LL * 256 + 256 + 256LL * 256 + 256,
                                240LL * 256 + 256 + 2 * (256LL * 256 + 256)};
constexpr long long kTower = 240LL * 256 + 256 + 2 * (256LL * 256 + 256) + 256LL * 128 + 128;  // 226,176
constexpr int kSample = 8192;   // key positions of the per-node threshold sample
constexpr int kKeyChunk = 32768;

__host__ __device__ constexpr long long boff(int l) { return kWoff[l] + (long long)kDin[l] * kDout[l]; }

// gather rows of a [n][128] array
__global__ void k_gather_rows(const float* __restrict__ src, const long long* __restrict__ rows, int n,
                              float* __restrict__ dst) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)n * EMB) return;
  const long long r = i / EMB;
  dst[i] = src[rows[r] * EMB + (i - r * EMB)];
}

// block-mean of loss[0..B) added to acc[0]
__global__ void k_loss_accum(const float* __restrict__ loss, int B, float w, float* __restrict__ acc) {
  __shared__ float red[256];
  float s = 0.f;
  for (int i = threadIdx.x; i < B; i += 256) s += loss[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) acc[0] += red[0] * w;
}

}  // namespace lp
}  // namespace qsmi

using namespace qsmi;
using namespace qsmi::lp;

struct qsmi_lp {
  int gpu = 0;
  hipStream_t stream = nullptr;
  float* dW = nullptr;   // parameters (query tower, key tower)
  float* dG = nullptr;   // gradients
  float* dM = nullptr;   // Adam moments
  float* dV = nullptr;
  int adamT = 0;
  // scratch, grown on demand
  char* ws = nullptr;
  size_t wsCap = 0;
  // keys of one TI
  float* dKeys = nullptr;
  float* dKsample = nullptr;
  __bf16* dKeysB = nullptr;  // the keys as bf16 (scoring operand)
  __bf16* dWqT = nullptr;    // query tower, transposed bf16 (fused forward)
  int* dSlot = nullptr;  // canonical slot map of the keys' stencil
  long long keyS = 0;
  int keyDims[2] = {0, 0};
  float mu = 0.f, sigma = 1.f;
  // proposal state of one wave
  int pCap = 0, pMaxNodes = 0, pK = 0, pShort = 0;
  int pYY = 0;  // the TI's images are (y^2, y) of one continuous variable
  char* pws = nullptr;
  float *pXq = nullptr, *pH = nullptr, *pQ = nullptr, *pGs = nullptr, *pThr = nullptr, *pEbuf = nullptr,
        *pBest = nullptr, *pCover = nullptr;
  uint32_t* pList = nullptr;
  int* pCnt = nullptr;
  int* pFlag = nullptr;
  Cand* pCand = nullptr;
  NodeInfo* pInfoEx = nullptr;
  NodeInfo* pInfoC = nullptr;  // the exact-scan nodes of the wave, compacted
  int* pPos = nullptr;         // node -> index in pInfoC (-1: not scanned)
  int* pExCnt = nullptr;       // number of compacted nodes
  int* pAct = nullptr;
  // proposed, fallbacks, overflows, shortlist sum, audit nodes, audit hits, audit best hits, audit positions
  unsigned long long* pCounters = nullptr;
};

namespace {

struct Acts {
  float* X;
  float* H[3];
  float* E;
};

// bump allocation in the handle's scratch buffer (device, 256-byte aligned)
struct Bump {
  qsmi_lp* lp;
  size_t off = 0;
  std::vector<std::pair<void**, size_t>> reqs;
  template <class T>
  void add(T** p, size_t n) {
    reqs.push_back({(void**)p, ((n * sizeof(T) + 255) / 256) * 256});
  }
  int commit() {
    size_t total = 0;
    for (auto& r : reqs) total += r.second;
    if (total > lp->wsCap) {
      if (lp->ws) (void)hipFree(lp->ws);
      lp->ws = nullptr;
      lp->wsCap = 0;
      if (hipMalloc(&lp->ws, total) != hipSuccess) return fail(QSMI_ERR_NOMEM, "learned proposal scratch");
      lp->wsCap = total;
    }
    size_t o = 0;
    for (auto& r : reqs) {
      *r.first = lp->ws + o;
      o += r.second;
    }
    return QSMI_OK;
  }
};

void add_acts(Bump& b, Acts& a, int n) {
  b.add(&a.X, (size_t)n * IN);
  for (int l = 0; l < 3; ++l) b.add(&a.H[l], (size_t)n * HID);
  b.add(&a.E, (size_t)n * EMB);
}

void forward(qsmi_lp* lp, int t, const Acts& a, int n, hipStream_t st) {
  const float* W = lp->dW + t * kTower;
  const float* in = a.X;
  for (int l = 0; l < 4; ++l) {
    float* out = l < 3 ? a.H[l] : a.E;
    GemmArgs g = gargs(n, kDout[l], kDin[l], in, kDin[l], W + kWoff[l], kDout[l], out, kDout[l]);
    g.bias = W + boff(l);
    g.relu = l < 3;
    gemm(false, false, g, st);
    in = out;
  }
}

// gradients of tower t's parameters from dE (n x 128); dHa, dHb: n x 256 scratch
void backward(qsmi_lp* lp, int t, const Acts& a, float* dE, int n, float* dHa, float* dHb) {
  const float* W = lp->dW + t * kTower;
  float* G = lp->dG + t * kTower;
  float* dOut = dE;
  for (int l = 3; l >= 0; --l) {
    const float* in = l > 0 ? a.H[l - 1] : a.X;
    gemm(true, false, gargs(kDin[l], kDout[l], n, in, kDin[l], dOut, kDout[l], G + kWoff[l], kDout[l]), lp->stream);
    hipLaunchKernelGGL(k_colsum, dim3((kDout[l] + 255) / 256), dim3(256), 0, lp->stream, dOut, n, kDout[l],
                       G + boff(l));
    if (l > 0) {
      float* dIn = (l & 1) ? dHa : dHb;
      gemm(false, true, gargs(n, kDin[l], kDout[l], dOut, kDout[l], W + kWoff[l], kDout[l], dIn, kDin[l]),
           lp->stream);
      const long long cnt = (long long)n * kDin[l];
      hipLaunchKernelGGL(k_relu_bwd, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, lp->stream, dIn, a.H[l - 1],
                         cnt);
      dOut = dIn;
    }
  }
}

struct LossBufs {
  float *Gm, *dG, *loss, *dQ, *dK, *dHa, *dHb;
};

void add_loss(Bump& b, LossBufs& L, int B, int C) {
  b.add(&L.Gm, (size_t)B * C);
  b.add(&L.dG, (size_t)B * C);
  b.add(&L.loss, (size_t)B);
  b.add(&L.dQ, (size_t)B * EMB);
  b.add(&L.dK, (size_t)C * EMB);
  b.add(&L.dHa, (size_t)std::max(B, C) * HID);
  b.add(&L.dHb, (size_t)std::max(B, C) * HID);
}

// loss over (q, k) activations already forwarded; all gradients into lp->dG
void loss_backward(qsmi_lp* lp, const Acts& q, const Acts& k, const LossBufs& L, int B, int C, float invTau) {
  hipStream_t st = lp->stream;
  gemm(false, true, gargs(B, C, EMB, q.E, EMB, k.E, EMB, L.Gm, C), st);
  hipLaunchKernelGGL(k_softmax_ce, dim3(B), dim3(256), 0, st, L.Gm, B, C, invTau, L.dG, L.loss);
  gemm(false, false, gargs(B, EMB, C, L.dG, C, k.E, EMB, L.dQ, EMB), st);
  gemm(true, false, gargs(C, EMB, B, L.dG, C, q.E, EMB, L.dK, EMB), st);
  backward(lp, 0, q, L.dQ, B, L.dHa, L.dHb);
  backward(lp, 1, k, L.dK, C, L.dHa, L.dHb);
}

void adam_step(qsmi_lp* lp, float lr) {
  const float b1 = 0.9f, b2 = 0.999f;
  lp->adamT++;
  const float c1 = 1.f - powf(b1, (float)lp->adamT), c2 = 1.f - powf(b2, (float)lp->adamT);
  const long long n = 2 * kTower;
  hipLaunchKernelGGL(k_adam, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, lp->stream, lp->dW, lp->dG, lp->dM,
                     lp->dV, n, lr, b1, b2, 1e-8f, c1, c2);
}

int hchk(hipError_t e, const char* what) {
  if (e == hipSuccess) return QSMI_OK;
  return fail(e == hipErrorOutOfMemory ? QSMI_ERR_NOMEM : QSMI_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

template <class T>
int upload(const T* src, size_t n, T** dst, std::vector<void*>& owned, hipStream_t st) {
  int rc = hchk(hipMalloc((void**)dst, std::max<size_t>(n, 1) * sizeof(T)), "learned proposal upload");
  if (rc) return rc;
  owned.push_back(*dst);
  if (n && src) return hchk(hipMemcpyAsync(*dst, src, n * sizeof(T), hipMemcpyHostToDevice, st), "learned proposal upload");
  return QSMI_OK;
}

void free_all(std::vector<void*>& v) {
  for (void* p : v) (void)hipFree(p);
  v.clear();
}

// canonical slot of every offset of the 17 x 17 window (-1: off the stencil)
std::vector<int> slot_map(const int* canon) {
  std::vector<int> m(SLOT_W * SLOT_W, -1);
  for (int i = 0; i < NB; ++i) {
    const int d0 = canon[2 * i], d1 = canon[2 * i + 1];
    if (d0 >= -SLOT_R && d0 <= SLOT_R && d1 >= -SLOT_R && d1 <= SLOT_R) m[(d1 + SLOT_R) * SLOT_W + d0 + SLOT_R] = i;
  }
  return m;
}

void mean_sigma(const float* ti, size_t S, float* mu, float* sigma) {
  double s = 0, s2 = 0;
  for (size_t i = 0; i < S; ++i) {
    s += ti[i];
    s2 += (double)ti[i] * ti[i];
  }
  const double m = s / (double)S;
  *mu = (float)m;
  *sigma = (float)std::sqrt(std::max(1e-30, s2 / (double)S - m * m));
}

}  // namespace

extern "C" {

int qsmi_lp_param_count(long long* count) {
  if (!count) return fail(QSMI_ERR_ARG, "null argument");
  *count = 2 * kTower;
  return QSMI_OK;
}

int qsmi_lp_create(int gpu, unsigned seed, qsmi_lp** out) {
  if (!out) return fail(QSMI_ERR_ARG, "null argument");
  int rc = hchk(hipSetDevice(gpu), "hipSetDevice");
  if (rc) return rc;
  qsmi_lp* lp = new qsmi_lp();
  lp->gpu = gpu;
  const long long n = 2 * kTower;
  if ((rc = hchk(hipStreamCreateWithFlags(&lp->stream, hipStreamNonBlocking), "stream")) ||
      (rc = hchk(hipMalloc(&lp->dW, n * sizeof(float)), "weights")) ||
      (rc = hchk(hipMalloc(&lp->dG, n * sizeof(float)), "gradients")) ||
      (rc = hchk(hipMalloc(&lp->dM, n * sizeof(float)), "moments")) ||
      (rc = hchk(hipMalloc(&lp->dV, n * sizeof(float)), "moments"))) {
    qsmi_lp_destroy(lp);
    return rc;
  }
  // He-uniform weights, zero biases (mt19937 + uniform_real_distribution)
  std::vector<float> w(n, 0.f);
  std::mt19937 gen(seed);
  for (int t = 0; t < 2; ++t)
    for (int l = 0; l < 4; ++l) {
      const float lim = std::sqrt(6.f / (float)kDin[l]) * (l == 3 ? 0.5f : 1.f);
      std::uniform_real_distribution<float> U(-lim, lim);
      float* W = w.data() + t * kTower + kWoff[l];
      for (long long i = 0; i < (long long)kDin[l] * kDout[l]; ++i) W[i] = U(gen);
    }
  if ((rc = hchk(hipMemcpy(lp->dW, w.data(), n * sizeof(float), hipMemcpyHostToDevice), "weights")) ||
      (rc = hchk(hipMemset(lp->dM, 0, n * sizeof(float)), "moments")) ||
      (rc = hchk(hipMemset(lp->dV, 0, n * sizeof(float)), "moments"))) {
    qsmi_lp_destroy(lp);
    return rc;
  }
  *out = lp;
  return QSMI_OK;
}

int qsmi_lp_destroy(qsmi_lp* lp) {
  if (!lp) return QSMI_OK;
  (void)hipSetDevice(lp->gpu);
  if (lp->stream) (void)hipStreamSynchronize(lp->stream);
  for (void* p : {(void*)lp->dW, (void*)lp->dG, (void*)lp->dM, (void*)lp->dV, (void*)lp->ws, (void*)lp->dKeys,
                  (void*)lp->dKsample, (void*)lp->dKeysB, (void*)lp->dSlot, (void*)lp->dWqT, (void*)lp->pws})
    if (p) (void)hipFree(p);
  if (lp->stream) (void)hipStreamDestroy(lp->stream);
  delete lp;
  return QSMI_OK;
}

int qsmi_lp_get_params(qsmi_lp* lp, float* out) {
  if (!lp || !out) return fail(QSMI_ERR_ARG, "null argument");
  (void)hipSetDevice(lp->gpu);
  int rc = hchk(hipMemcpyAsync(out, lp->dW, 2 * kTower * sizeof(float), hipMemcpyDeviceToHost, lp->stream), "copy");
  if (!rc) rc = hchk(hipStreamSynchronize(lp->stream), "sync");
  return rc;
}

int qsmi_lp_set_params(qsmi_lp* lp, const float* in) {
  if (!lp || !in) return fail(QSMI_ERR_ARG, "null argument");
  (void)hipSetDevice(lp->gpu);
  int rc = hchk(hipMemcpyAsync(lp->dW, in, 2 * kTower * sizeof(float), hipMemcpyHostToDevice, lp->stream), "copy");
  if (!rc) rc = hchk(hipStreamSynchronize(lp->stream), "sync");
  lp->adamT = 0;
  if (!rc) rc = hchk(hipMemsetAsync(lp->dM, 0, 2 * kTower * sizeof(float), lp->stream), "moments");
  if (!rc) rc = hchk(hipMemsetAsync(lp->dV, 0, 2 * kTower * sizeof(float), lp->stream), "moments");
  if (!rc) rc = hchk(hipStreamSynchronize(lp->stream), "sync");
  return rc;
}

int qsmi_lp_gemm(qsmi_lp* lp, int transA, int transB, int M, int N, int K, const float* A, const float* B,
                 float* C) {
  if (!lp || !A || !B || !C) return fail(QSMI_ERR_ARG, "null argument");
  if (M < 1 || N < 1 || K < 1) return fail(QSMI_ERR_ARG, "empty product");
  (void)hipSetDevice(lp->gpu);
  std::vector<void*> own;
  float *dA, *dB, *dC;
  int rc = upload(A, (size_t)M * K, &dA, own, lp->stream);
  if (!rc) rc = upload(B, (size_t)K * N, &dB, own, lp->stream);
  if (!rc) rc = upload<float>(nullptr, (size_t)M * N, &dC, own, lp->stream);
  if (!rc) {
    GemmArgs g = gargs(M, N, K, dA, transA ? M : K, dB, transB ? K : N, dC, N);
    gemm(transA != 0, transB != 0, g, lp->stream);
    rc = hchk(hipGetLastError(), "gemm");
    if (!rc) rc = hchk(hipMemcpyAsync(C, dC, (size_t)M * N * sizeof(float), hipMemcpyDeviceToHost, lp->stream), "copy");
    if (!rc) rc = hchk(hipStreamSynchronize(lp->stream), "sync");
  }
  (void)hipStreamSynchronize(lp->stream);
  free_all(own);
  return rc;
}

int qsmi_lp_encode(qsmi_lp* lp, int tower, int n, const float* features, float* embeddings) {
  if (!lp || !features || !embeddings) return fail(QSMI_ERR_ARG, "null argument");
  if (tower < 0 || tower > 1 || n < 1) return fail(QSMI_ERR_ARG, "tower must be 0 or 1, n >= 1");
  (void)hipSetDevice(lp->gpu);
  Bump b{lp};
  Acts a;
  add_acts(b, a, n);
  int rc = b.commit();
  if (rc) return rc;
  rc = hchk(hipMemcpyAsync(a.X, features, (size_t)n * IN * sizeof(float), hipMemcpyHostToDevice, lp->stream), "copy");
  if (rc) return rc;
  forward(lp, tower, a, n, lp->stream);
  rc = hchk(hipGetLastError(), "encode");
  if (!rc) rc = hchk(hipMemcpyAsync(embeddings, a.E, (size_t)n * EMB * sizeof(float), hipMemcpyDeviceToHost, lp->stream), "copy");
  if (!rc) rc = hchk(hipStreamSynchronize(lp->stream), "sync");
  return rc;
}

int qsmi_lp_loss_grad(qsmi_lp* lp, int B, int C, const float* qfeat, const float* kfeat, float invTau, float* loss,
                      float* grad) {
  if (!lp || !qfeat || !kfeat || !loss) return fail(QSMI_ERR_ARG, "null argument");
  if (B < 1 || C < B || C > 65536) return fail(QSMI_ERR_ARG, "need 1 <= B <= C <= 65536");
  (void)hipSetDevice(lp->gpu);
  Bump b{lp};
  Acts q, k;
  LossBufs L;
  add_acts(b, q, B);
  add_acts(b, k, C);
  add_loss(b, L, B, C);
  int rc = b.commit();
  if (rc) return rc;
  hipStream_t st = lp->stream;
  rc = hchk(hipMemcpyAsync(q.X, qfeat, (size_t)B * IN * sizeof(float), hipMemcpyHostToDevice, st), "copy");
  if (!rc) rc = hchk(hipMemcpyAsync(k.X, kfeat, (size_t)C * IN * sizeof(float), hipMemcpyHostToDevice, st), "copy");
  if (rc) return rc;
  forward(lp, 0, q, B, st);
  forward(lp, 1, k, C, st);
  loss_backward(lp, q, k, L, B, C, invTau);
  rc = hchk(hipGetLastError(), "loss");
  std::vector<float> lv(B);
  if (!rc) rc = hchk(hipMemcpyAsync(lv.data(), L.loss, B * sizeof(float), hipMemcpyDeviceToHost, st), "copy");
  if (!rc && grad) rc = hchk(hipMemcpyAsync(grad, lp->dG, 2 * kTower * sizeof(float), hipMemcpyDeviceToHost, st), "copy");
  if (!rc) rc = hchk(hipStreamSynchronize(st), "sync");
  if (!rc) {
    double s = 0;
    for (float x : lv) s += x;
    *loss = (float)(s / B);
  }
  return rc;
}

int qsmi_lp_train(qsmi_lp* lp, const qsmi_lp_train_set* set, const qsmi_lp_train_opts* opts, float* lossOut) {
  if (!lp || !set || !opts) return fail(QSMI_ERR_ARG, "null argument");
  if (!set->ti || !set->canon || !set->offsets || !set->values || !set->best)
    return fail(QSMI_ERR_ARG, "null training set array");
  const int B = opts->batch > 0 ? opts->batch : 256;
  const int R = opts->negatives > 0 ? opts->negatives : 768;
  const int C = B + R;
  const float lr = opts->lr > 0.f ? opts->lr : 1e-3f;
  const float invTau = opts->invTau > 0.f ? opts->invTau : 10.f;
  const int steps = opts->steps;
  if (steps < 0 || B > 4096 || C > 16384) return fail(QSMI_ERR_ARG, "batch/negatives out of range");
  const size_t S = (size_t)set->tiSize[0] * set->tiSize[1];
  if (S < 2) return fail(QSMI_ERR_ARG, "empty TI");
  std::vector<int> usable;
  for (int i = 0; i < set->nPairs; ++i)
    if (set->best[i] != 0xFFFFFFFFu && set->best[i] < S) usable.push_back(i);
  if (usable.empty()) return fail(QSMI_ERR_ARG, "no usable training pair");
  (void)hipSetDevice(lp->gpu);
  hipStream_t st = lp->stream;
  float mu, sigma;
  mean_sigma(set->ti, S, &mu, &sigma);
  // every step's query indices and candidate positions, drawn up front
  std::mt19937 gen(opts->seed);
  std::uniform_int_distribution<int> pickPair(0, (int)usable.size() - 1);
  std::uniform_int_distribution<unsigned> pickPos(0, (unsigned)(S - 1));
  std::vector<int> qidx((size_t)steps * B);
  std::vector<uint32_t> cpos((size_t)steps * C);
  for (int s = 0; s < steps; ++s) {
    for (int b = 0; b < B; ++b) {
      const int pi = usable[pickPair(gen)];
      qidx[(size_t)s * B + b] = pi;
      cpos[(size_t)s * C + b] = set->best[pi];
    }
    for (int r = 0; r < R; ++r) cpos[(size_t)s * C + B + r] = pickPos(gen);
  }
  std::vector<void*> own;
  float *dTi, *dVal, *dAcc;
  int *dCanon, *dOff, *dQidx;
  uint32_t* dCpos;
  const int nBlk = (steps + 49) / 50;
  int rc = upload(set->ti, S, &dTi, own, st);
  if (!rc) rc = upload(set->canon, (size_t)2 * NB, &dCanon, own, st);
  const std::vector<int> smap = slot_map(set->canon);
  int* dSlot;
  if (!rc) rc = upload(smap.data(), smap.size(), &dSlot, own, st);
  if (!rc) rc = upload(set->offsets, (size_t)set->nPairs * NB * 2, &dOff, own, st);
  if (!rc) rc = upload(set->values, (size_t)set->nPairs * NB, &dVal, own, st);
  if (!rc) rc = upload(qidx.data(), qidx.size(), &dQidx, own, st);
  if (!rc) rc = upload(cpos.data(), cpos.size(), &dCpos, own, st);
  if (!rc) rc = upload<float>(nullptr, (size_t)std::max(nBlk, 1), &dAcc, own, st);
  if (!rc) rc = hchk(hipMemsetAsync(dAcc, 0, std::max(nBlk, 1) * sizeof(float), st), "memset");
  Bump b{lp};
  Acts q, k;
  LossBufs L;
  add_acts(b, q, B);
  add_acts(b, k, C);
  add_loss(b, L, B, C);
  if (!rc) rc = b.commit();
  if (rc) {
    (void)hipStreamSynchronize(st);
    free_all(own);
    return rc;
  }
  const float invS = 1.f / sigma;
  for (int s = 0; s < steps; ++s) {
    hipLaunchKernelGGL(k_pair_features, dim3(B), dim3(128), 0, st, dQidx + (size_t)s * B, dOff, dVal, dSlot, mu, invS,
                       q.X);
    hipLaunchKernelGGL(k_key_features, dim3((C * NB + 255) / 256), dim3(256), 0, st, C, dCpos + (size_t)s * C, 0LL,
                       dTi, 1, (int)set->tiSize[0], (int)set->tiSize[1], dCanon, mu, invS, k.X);
    forward(lp, 0, q, B, st);
    forward(lp, 1, k, C, st);
    loss_backward(lp, q, k, L, B, C, invTau);
    adam_step(lp, lr);
    hipLaunchKernelGGL(k_loss_accum, dim3(1), dim3(256), 0, st, L.loss, B,
                       1.f / (float)B / (float)std::min(50, steps - (s / 50) * 50), dAcc + s / 50);
  }
  rc = hchk(hipGetLastError(), "training step");
  std::vector<float> acc(std::max(nBlk, 1));
  if (!rc) rc = hchk(hipMemcpyAsync(acc.data(), dAcc, acc.size() * sizeof(float), hipMemcpyDeviceToHost, st), "copy");
  if (!rc) rc = hchk(hipStreamSynchronize(st), "sync");
  (void)hipStreamSynchronize(st);
  free_all(own);
  if (!rc && lossOut)
    for (int i = 0; i < nBlk; ++i) lossOut[i] = acc[i];
  return rc;
}

int qsmi_lp_build_keys(qsmi_lp* lp, const unsigned* tiSize, const float* ti, const int* canon) {
  if (!lp || !tiSize || !ti || !canon) return fail(QSMI_ERR_ARG, "null argument");
  const long long S = (long long)tiSize[0] * tiSize[1];
  if (S < kSample || S > 0x7FFFFFFFLL) return fail(QSMI_ERR_ARG, "TI must hold between 8192 and 2^31 cells");
  (void)hipSetDevice(lp->gpu);
  hipStream_t st = lp->stream;
  (void)hipStreamSynchronize(st);
  if (lp->dKeys) (void)hipFree(lp->dKeys);
  if (lp->dKsample) (void)hipFree(lp->dKsample);
  if (lp->dKeysB) (void)hipFree(lp->dKeysB);
  lp->dKeys = lp->dKsample = nullptr;
  lp->dKeysB = nullptr;
  lp->keyS = 0;
  int rc = hchk(hipMalloc(&lp->dKeys, (size_t)S * EMB * sizeof(float)), "keys");
  if (!rc) rc = hchk(hipMalloc(&lp->dKsample, (size_t)kSample * EMB * sizeof(float)), "keys");
  if (!rc) rc = hchk(hipMalloc(&lp->dKeysB, (size_t)S * EMB * sizeof(__bf16)), "keys");
  if (!rc && !lp->dSlot) rc = hchk(hipMalloc(&lp->dSlot, SLOT_W * SLOT_W * sizeof(int)), "slots");
  if (rc) return rc;
  mean_sigma(ti, (size_t)S, &lp->mu, &lp->sigma);
  std::vector<void*> own;
  float* dTi;
  int* dCanon;
  long long* dRows;
  std::vector<long long> rows(kSample);
  for (int i = 0; i < kSample; ++i) rows[i] = ((long long)i * S + S / 2) / kSample;
  rc = upload(ti, (size_t)S, &dTi, own, st);
  if (!rc) rc = upload(canon, (size_t)2 * NB, &dCanon, own, st);
  if (!rc) rc = upload(rows.data(), rows.size(), &dRows, own, st);
  if (!rc) {
    const std::vector<int> smap = slot_map(canon);
    rc = hchk(hipMemcpyAsync(lp->dSlot, smap.data(), smap.size() * sizeof(int), hipMemcpyHostToDevice, st), "slots");
  }
  Bump b{lp};
  Acts a;
  add_acts(b, a, kKeyChunk);
  if (!rc) rc = b.commit();
  if (!rc) {
    for (long long first = 0; first < S; first += kKeyChunk) {
      const int n = (int)std::min<long long>(kKeyChunk, S - first);
      hipLaunchKernelGGL(k_key_features, dim3((unsigned)(((long long)n * NB + 255) / 256)), dim3(256), 0, st, n,
                         (const uint32_t*)nullptr, first, dTi, 1, (int)tiSize[0], (int)tiSize[1], dCanon, lp->mu,
                         1.f / lp->sigma, a.X);
      Acts o = a;
      o.E = lp->dKeys + first * EMB;
      forward(lp, 1, o, n, st);
    }
    hipLaunchKernelGGL(k_gather_rows, dim3((kSample * EMB + 255) / 256), dim3(256), 0, st, lp->dKeys, dRows, kSample,
                       lp->dKsample);
    hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)((S * EMB + 255) / 256)), dim3(256), 0, st, lp->dKeys, lp->dKeysB,
                       S * EMB);
    rc = hchk(hipGetLastError(), "keys");
  }
  if (!rc) rc = hchk(hipStreamSynchronize(st), "sync");
  (void)hipStreamSynchronize(st);
  free_all(own);
  if (!rc) {
    lp->keyS = S;
    lp->keyDims[0] = (int)tiSize[0];
    lp->keyDims[1] = (int)tiSize[1];
  }
  return rc;
}

}  // extern "C"

// ===========================================================================
// learned proposal inside the path loop (one wave at a time, device stream)
// ===========================================================================
namespace qsmi {
namespace lp {

// Route each node: 0 = drew from the marginal, 1 = exact scan (fallback),
// 2 = proposal. infoEx: the wave's NodeInfo with the valid box emptied for
// the nodes the exact scan must skip.
__global__ void k_lp_flags(int nN, int w0, const int* __restrict__ order, const NodeInfo* __restrict__ info,
                           const int* __restrict__ cnt, int cap, const float* __restrict__ best,
                           const Cand* __restrict__ cand, int K, int every, float mseLimit,
                           const float* __restrict__ cover, float minCover, int audit,
                           int* __restrict__ flag, NodeInfo* __restrict__ infoEx, NodeInfo* __restrict__ infoC,
                           int* __restrict__ pos, int* __restrict__ exCnt,
                           unsigned long long* __restrict__ counters) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nN) return;
  NodeInfo ni = info[q];
  int f = 0;
  if (ni.nb > 0) {
    const int i = order[w0 + q];
    bool overflow = false;
    int total = 0;
    for (int g = 0; g < NSEG; ++g) {
      const int c = cnt[q * NSEG + g];
      overflow |= c > cap / NSEG;
      total += min(c, cap / NSEG);
    }
    const bool few = !(cand[(long long)q * K + K - 1].key > -INFINITY);
    const bool periodic = every > 0 && (i % every) == 0;
    const float mse = -best[q] / (float)max(1, ni.nb - 1);
    const bool poor = mseLimit > 0.f && !(mse <= mseLimit);
    const bool unseen = !(cover[q] >= minCover);
    f = (overflow || few || periodic || poor || unseen) ? 1 : 2;
    if (f == 1) {
      atomicAdd(&counters[1], 1ull);
      if (overflow) atomicAdd(&counters[2], 1ull);
    } else {
      atomicAdd(&counters[0], 1ull);
      atomicAdd(&counters[3], (unsigned long long)total);
    }
  }
  flag[q] = f;
  const bool scan = f == 1 || (f == 2 && audit);
  // compacted list of the nodes to scan (slot order is irrelevant: every
  // node's scan is independent), and the full list with empty boxes
  int k = -1;
  if (scan) {
    k = atomicAdd(exCnt, 1);
    infoC[k] = ni;
  }
  pos[q] = k;
  if (!scan)
    for (int d = 0; d < 3; ++d) {  // empty valid box: every scan kernel skips the node
      ni.lo[d] = 1;
      ni.hi[d] = 0;
    }
  infoEx[q] = ni;
}

// Fallback nodes: cand[node] = the K best of the exact scan's block
// candidates. Audit of proposal nodes: how many of the exact K best the
// shortlist held.
template <int KT>
__global__ void __launch_bounds__(256) k_lp_combine(const int* __restrict__ flag, const NodeInfo* __restrict__ info,
                                                   const Cand* __restrict__ candEx, const int* __restrict__ pos,
                                                   int blocks, int rowsPerBlock, int K, int audit,
                                                   const uint32_t* __restrict__ list, const int* __restrict__ cnt,
                                                   int cap, Cand* __restrict__ cand,
                                                   unsigned long long* __restrict__ counters) {
  __shared__ TopkScratch red;
  __shared__ Cand ex[kMaxK];
  __shared__ int found[kMaxK];
  const int node = blockIdx.x, tid = threadIdx.x;
  const int f = flag[node];
  if (f == 0 || (f == 2 && !audit)) return;
  // blocks holding candidates: all of them, or (generic rows kernel) only
  // those covering the node's valid rows, as the merge stage counts them
  int nblk = blocks;
  if (rowsPerBlock > 0) {
    const NodeInfo ni = info[node];
    const int n1v = ni.hi[1] - ni.lo[1] + 1, n2v = ni.hi[2] - ni.lo[2] + 1, n0v = ni.hi[0] - ni.lo[0] + 1;
    const long long nrv = (n1v > 0 && n2v > 0 && n0v > 0) ? (long long)n1v * n2v : 0;
    nblk = (int)((nrv + rowsPerBlock - 1) / rowsPerBlock);
  }
  RankList<KT> rl;
  rl.init();
  const Cand* ce = candEx + (long long)(pos ? pos[node] : node) * blocks * K;
  for (int i = tid; i < nblk * K; i += 256) {
    const Cand c = ce[i];
    if (c.key > -INFINITY) rl.insert(c.key, c.h, c.pos, c.val);
  }
  rl.block_extract(K, f == 1 ? cand + (long long)node * K : ex, &red);
  if (f == 1) return;
  // audit
  if (tid < K) found[tid] = 0;
  __syncthreads();
  const int segCap = cap / NSEG;
  for (int i = tid; i < cap; i += 256) {
    if ((i % segCap) >= cnt[node * NSEG + i / segCap]) continue;
    const uint32_t p = list[(long long)node * cap + i];
    for (int r = 0; r < K; ++r)
      if (ex[r].pos == p) found[r] = 1;
  }
  __syncthreads();
  if (tid == 0) {
    int hits = 0, valid = 0;
    for (int r = 0; r < K; ++r)
      if (ex[r].pos != 0xFFFFFFFFu) {
        valid++;
        hits += found[r];
      }
    atomicAdd(&counters[4], 1ull);
    atomicAdd(&counters[5], (unsigned long long)hits);
    atomicAdd(&counters[6], (unsigned long long)((valid > 0 && found[0]) ? 1 : 0));
    atomicAdd(&counters[7], (unsigned long long)valid);
  }
}

// training-pair record of one wave (block per node)
__global__ void __launch_bounds__(128) k_lp_record(int w0, const int* __restrict__ order,
                                                  const NodeInfo* __restrict__ info, const int* __restrict__ nbOff,
                                                  const float* __restrict__ di, int nVar, int sd0,
                                                  const uint32_t* __restrict__ outPos, int K, int cap,
                                                  int* __restrict__ recOff, float* __restrict__ recVal,
                                                  uint32_t* __restrict__ recBest) {
  const int q = blockIdx.x;
  const int i = order[w0 + q];
  if (i >= cap) return;
  const NodeInfo ni = info[q];
  const int c0 = (int)(ni.cell % (unsigned)sd0), c1 = (int)(ni.cell / (unsigned)sd0);
  for (int slot = threadIdx.x; slot < NB; slot += blockDim.x) {
    const int j = slot + 1;
    int d0 = 0, d1 = 0;
    float v = NAN;
    if (ni.nb > 0 && j < ni.nb) {
      d0 = nbOff[(ni.nbStart + j) * 2];
      d1 = nbOff[(ni.nbStart + j) * 2 + 1];
      v = di[((long long)(c1 + d1) * sd0 + (c0 + d0)) * nVar];
    }
    recOff[((long long)i * NB + slot) * 2] = d0;
    recOff[((long long)i * NB + slot) * 2 + 1] = d1;
    recVal[(long long)i * NB + slot] = v;
  }
  if (threadIdx.x == 0) recBest[i] = ni.nb > 0 ? outPos[(long long)q * K] : 0xFFFFFFFFu;
}

}  // namespace lp

int lp_prepare(qsmi_lp* lp, const qsmi_ti* t, int maxNodes, int K, int shortlist, const int* active, int V,
               int circular) {
  using namespace lp;
  if (circular) return fail(QSMI_ERR_UNSUPPORTED, "learned proposal: circular TI unsupported");
  if (!lp->dKeys || lp->keyS != t->S || lp->keyDims[0] != t->s[0] || lp->keyDims[1] != t->s[1])
    return fail(QSMI_ERR_ARG, "learned proposal: keys were not built for this TI");
  if (t->D != 2) return fail(QSMI_ERR_UNSUPPORTED, "learned proposal: 2-D only");
  if (V > 4) return fail(QSMI_ERR_UNSUPPORTED, "learned proposal: at most 4 active sub-variables");
  if (maxNodes > 256) return fail(QSMI_ERR_UNSUPPORTED, "learned proposal: waves of at most 256 nodes");
  if (t->S > (1LL << 24)) return fail(QSMI_ERR_UNSUPPORTED, "learned proposal: TI of at most 2^24 cells");
  const int L = shortlist > 0 ? shortlist : 4096;
  const int cap = NSEG * ((2 * L + NSEG - 1) / NSEG);  // room for twice the target, in NSEG equal segments
  if (!(lp->pws && lp->pMaxNodes >= maxNodes && lp->pK == K && lp->pShort == L)) {
  if (lp->pws) (void)hipFree(lp->pws);
  lp->pws = nullptr;
  const size_t n = (size_t)std::max(maxNodes, 1);
  std::vector<std::pair<void**, size_t>> req = {
      {(void**)&lp->pXq, n * IN * 4},          {(void**)&lp->pH, 3 * n * HID * 4},
      {(void**)&lp->pQ, n * EMB * 4},          {(void**)&lp->pGs, n * kSample * 4},
      {(void**)&lp->pThr, n * 4},              {(void**)&lp->pList, n * cap * 4},
      {(void**)&lp->pCnt, n * NSEG * 4},       {(void**)&lp->pEbuf, n * cap * 4},
      {(void**)&lp->pBest, n * 4},             {(void**)&lp->pFlag, n * 4},
      {(void**)&lp->pCover, n * 4},
      {(void**)&lp->pCand, n * K * sizeof(Cand)}, {(void**)&lp->pInfoEx, n * sizeof(NodeInfo)},
      {(void**)&lp->pInfoC, n * sizeof(NodeInfo)}, {(void**)&lp->pPos, n * 4}, {(void**)&lp->pExCnt, 4},
      {(void**)&lp->pCounters, 8 * 8},         {(void**)&lp->pAct, 4 * 4}};
  size_t total = 0;
  for (auto& r : req) total += (r.second + 255) / 256 * 256;
  if (hipMalloc(&lp->pws, total) != hipSuccess) return fail(QSMI_ERR_NOMEM, "learned proposal: wave buffers");
  size_t o = 0;
  for (auto& r : req) {
    *r.first = lp->pws + o;
    o += (r.second + 255) / 256 * 256;
  }
  lp->pMaxNodes = (int)n;
  lp->pK = K;
  lp->pShort = L;
  lp->pCap = cap;
  }
  if (!lp->dWqT && hipMalloc(&lp->dWqT, (size_t)(3 * HID + EMB) * KP * sizeof(__bf16)) != hipSuccess)
    return fail(QSMI_ERR_NOMEM, "learned proposal: transposed weights");
  {
    long long o = 0;
    for (int l = 0; l < 4; ++l) {
      hipLaunchKernelGGL(k_wt_bf16, dim3((kDout[l] * KP + 255) / 256), dim3(256), 0, lp->stream, lp->dW + kWoff[l],
                         kDin[l], kDout[l], lp->dWqT + o);
      o += (long long)kDout[l] * KP;
    }
    if (hipStreamSynchronize(lp->stream) != hipSuccess) return fail(QSMI_ERR_HIP, "learned proposal: weights");
  }
  int act[4] = {0, 0, 0, 0};
  for (int a = 0; a < V; ++a) act[a] = active[a];
  lp->pYY = (t->nSub == 2 && V == 2 && active[0] == 0 && active[1] == 1) ? 1 : 0;
  if (hipMemcpy(lp->pAct, act, sizeof(act), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(lp->pCounters, 0, 8 * 8) != hipSuccess)
    return fail(QSMI_ERR_HIP, "learned proposal: set-up copies");
  return QSMI_OK;
}

void lp_propose(qsmi_lp* lp, const qsmi_ti* t, int V, int nN, const NodeInfo* info, const int* nbOff,
                const float* nbVal, const float* di, int nVar, int sd0, int K, float minCover, hipStream_t st) {
  using namespace lp;
  Acts a;
  a.X = lp->pXq;
  for (int l = 0; l < 3; ++l) a.H[l] = lp->pH + (size_t)l * lp->pMaxNodes * HID;
  a.E = lp->pQ;
  hipLaunchKernelGGL(k_wave_features, dim3(nN), dim3(128), 0, st, info, nbOff, di, nVar, sd0, lp->dSlot, lp->mu,
                     1.f / lp->sigma, a.X, lp->pCover, lp->pExCnt);
  hipLaunchKernelGGL(k_mlp_fwd, dim3((nN + 31) / 32), dim3(256), 0, st, nN, a.X, lp->dWqT, lp->dW, a.E);
  // per-node threshold: the r-th best sampled score, r = L * sample / S
  gemm(false, true, gargs(nN, kSample, EMB, a.E, EMB, lp->dKsample, EMB, lp->pGs, kSample), st);
  const int r = (int)std::max<long long>(1, ((long long)lp->pShort * kSample + lp->keyS - 1) / lp->keyS);
  hipLaunchKernelGGL(k_row_kth, dim3(nN), dim3(256), 0, st, lp->pGs, kSample, r, lp->pCover, minCover, lp->pThr);
  (void)hipMemsetAsync(lp->pCnt, 0, (size_t)nN * NSEG * sizeof(int), st);
  hipLaunchKernelGGL(k_score, dim3(512), dim3(256), 0, st, nN, lp->keyS, a.E, lp->dKeysB, lp->pThr, lp->pList,
                     lp->pCnt, lp->pCap);
  hipLaunchKernelGGL(k_rescore, dim3((lp->pCap + kRescoreChunk - 1) / kRescoreChunk, nN), dim3(256), 0, st, info,
                     nbOff, nbVal, V, lp->pYY, lp->pAct, t->dSub, t->S, t->s[0], lp->pList, lp->pCnt, lp->pCap,
                     lp->pEbuf);
#define QSMI_LP_SELECT(KT)                                                                                     \
  hipLaunchKernelGGL(k_rescore_select<KT>, dim3(nN), dim3(256), 0, st, info, lp->pList, lp->pCnt, lp->pCap, K, \
                     lp->pEbuf, lp->pCand, lp->pBest)
  if (K <= 2) QSMI_LP_SELECT(2);
  else if (K <= 4) QSMI_LP_SELECT(4);
  else if (K <= 8) QSMI_LP_SELECT(8);
  else QSMI_LP_SELECT(16);
#undef QSMI_LP_SELECT
}

void lp_flags(qsmi_lp* lp, int nN, int w0, const int* order, const NodeInfo* info, int K, int every, float mseSigma2,
              float minCover, int audit, hipStream_t st) {
  using namespace lp;
  hipLaunchKernelGGL(k_lp_flags, dim3((nN + 127) / 128), dim3(128), 0, st, nN, w0, order, info, lp->pCnt, lp->pCap,
                     lp->pBest, lp->pCand, K, every, mseSigma2 > 0.f ? mseSigma2 * lp->sigma * lp->sigma : 0.f,
                     lp->pCover, minCover, audit, lp->pFlag, lp->pInfoEx, lp->pInfoC, lp->pPos, lp->pExCnt,
                     lp->pCounters);
}

int lp_cover_all(qsmi_lp* lp, int nPath, int n, const int* dNbIdx, const int* dNbCnt, const int* dOffs, float minCover,
                 unsigned char* dOk, hipStream_t st) {
  using namespace lp;
  if (!lp->dSlot) return fail(QSMI_ERR_ARG, "learned proposal: keys were not built");
  if (nPath <= 0) return QSMI_OK;
  hipLaunchKernelGGL(k_cover_all, dim3(nPath), dim3(64), 0, st, nPath, n, dNbIdx, dNbCnt, dOffs, lp->dSlot,
                     minCover > 0.f ? minCover : 0.5f, dOk);
  return QSMI_OK;
}

void lp_combine(qsmi_lp* lp, int nN, const NodeInfo* info, const Cand* candEx, bool compacted, int blocks,
                int rowsPerBlock, int K, int audit, hipStream_t st) {
  using namespace lp;
  const int* pos = compacted ? lp->pPos : nullptr;
#define QSMI_LP_COMBINE(KT)                                                                                         \
  hipLaunchKernelGGL(k_lp_combine<KT>, dim3(nN), dim3(256), 0, st, lp->pFlag, info, candEx, pos, blocks,            \
                     rowsPerBlock, K, audit,                                                                        \
                     lp->pList, lp->pCnt, lp->pCap, lp->pCand, lp->pCounters)
  if (K <= 2) QSMI_LP_COMBINE(2);
  else if (K <= 4) QSMI_LP_COMBINE(4);
  else if (K <= 8) QSMI_LP_COMBINE(8);
  else QSMI_LP_COMBINE(16);
#undef QSMI_LP_COMBINE
}

NodeInfo* lp_info_ex(qsmi_lp* lp) { return lp->pInfoEx; }
NodeInfo* lp_info_compact(qsmi_lp* lp) { return lp->pInfoC; }
const int* lp_compact_count(qsmi_lp* lp) { return lp->pExCnt; }
Cand* lp_cand(qsmi_lp* lp) { return lp->pCand; }

int lp_read_counters(qsmi_lp* lp, long long out[8]) {
  unsigned long long c[8];
  if (hipMemcpy(c, lp->pCounters, sizeof(c), hipMemcpyDeviceToHost) != hipSuccess)
    return fail(QSMI_ERR_HIP, "learned proposal: counters");
  for (int i = 0; i < 8; ++i) out[i] = (long long)c[i];
  return QSMI_OK;
}

void lp_record(int nN, int w0, const int* order, const NodeInfo* info, const int* nbOff, const float* di, int nVar,
               int sd0, const uint32_t* outPos, int K, int cap, int* recOff, float* recVal, uint32_t* recBest,
               hipStream_t st) {
  hipLaunchKernelGGL(lp::k_lp_record, dim3(nN), dim3(128), 0, st, w0, order, info, nbOff, di, nVar, sd0, outPos, K,
                     cap, recOff, recVal, recBest);
}

}  // namespace qsmi
