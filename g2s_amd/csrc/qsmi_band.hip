// Exact selection of the K best positions of a node (gfx950).
//
// The FP32 maps of the scan carry an absolute rounding error that is set by
// the norms of the TI sub-images and of the neighbour values, not by the
// local magnitude of the mismatch: on a heavy-tailed TI (the 1000^2
// log-normal workload) the error is ~1e-2 everywhere, larger than the gap
// between the best positions of a low-valued neighbourhood. The selection
// therefore never trusts FP32 values near the K-th best:
//
//   1. merge (k_merge_band): per node, the K best FP32 candidates of the
//      rows stage give T, the K-th best value; the band is every position
//      whose FP32 value is >= theta = T - 2 eps (eps: the node's error
//      bound, NodeInfo::eps). Every position whose exact value reaches the
//      exact K-th best lies in the band (its FP32 value is at least its
//      exact value - eps, and the exact K-th best is at least T - eps).
//      Each rows group (wavefront or block) emitted its K best and the
//      number of its values within 2 eps of its own K-th best; a group whose
//      K-th candidate is in the band and which holds more than K such values
//      may hold band positions it did not emit, so its rows block is queued
//      for a second pass; the candidates of the other groups in the band go
//      to the node's list.
//   2. second pass (k_rows1024_collect / k_rows_collect): the queued rows
//      blocks again, appending every position >= theta.
//   3. rescoring (k_band_rescore): the exact value of every listed
//      position in FP64, E = delta0 + sum_j sum_a v_ja Y_a(p + d_j) summed in
//      that order with separately rounded products and sums (no fused
//      multiply-add). The lists of all nodes of the call are cut into tiles
//      of 64 positions spread over one grid, so a crowded node's list is
//      scored by as many wavefronts as it has tiles. A node whose band
//      overflows the list capacity (full fallback) is rescored over its whole
//      valid box instead, in kFullTiles tiles that keep their K best.
//   4. k_band_final: per node, the K best by the device rule on the FP64
//      values, then the rank draw and the simulation write.
// Nodes with eps == 0 (categorical-only TIs with a tie quantum, TIs with
// NaN) keep the FP32 selection and finish in step 1.
//
// Replaces the selection of the reference's scan (fKst over the FP32 map,
// src/CPUThreadAcceleratorDevice.cpp:427-468, include/fKb.hpp:907-954) with
// an exact one; the map itself is src/CPUThreadAcceleratorDevice.cpp:251-357.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <algorithm>

#include "qsmi_internal.hpp"
#include "qsmi_select.hpp"

namespace qsmi {

int rows1024_groups();
void launch_rows1024_collect(const ScanGeom& g, const float2* tw, const NodeInfo* info, const float2* CT,
                           const int2* queue, const int* qcnt, const float* theta, uint32_t* bandPos, int* bandCnt,
                           int maxItems, hipStream_t st);
void launch_rows_collect(const ScanGeom& g, const FftPlan& plan0, int lognv, const float2* tw0, const NodeInfo* info,
                         const float2* C, const int2* queue, const int* qcnt, const float* theta, uint32_t* bandPos,
                         int* bandCnt, int maxItems, hipStream_t st);

namespace {

constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr int kFullTiles = 128;  // rescoring tiles of a full-fallback node (slices of its valid box)
constexpr int kSegNodes = 2048;  // nodes per rescoring launch (their tile prefix sums are held in LDS)

// Device view of the band buffers.
struct BandDev {
  uint32_t* pos;   // [node][kBandCap]
  int* cnt;        // [node]: listed positions (-1: node finished by the merge)
  float* theta;    // [node]
  int2* queue;     // (node, rows block) items of the second pass
  int* qcnt;       // [0] queue items (zero between calls)
  double* val;     // [node][kBandCap] FP64 values of the listed positions
  Cand64* part;    // [node][kFullTiles][K] full-fallback tile results
  unsigned long long* stats;  // exact nodes, listed positions, second-pass blocks, full-fallback nodes
};

BandDev dev_view(const BandBufs& b) {
  BandDev d;
  d.pos = b.pos;
  d.cnt = b.cnt;
  d.theta = b.theta;
  d.queue = b.queue;
  d.qcnt = b.qcnt;
  d.val = b.val;
  d.part = b.part;
  d.stats = b.stats;
  return d;
}

// Exact value of position p (FP64, the oracle's summation order).
__device__ __forceinline__ double exact_value(const ScanGeom& g, const NodeInfo& ni, const int* __restrict__ off,
                                              const double* __restrict__ v, const float* __restrict__ Y, long long S,
                                              uint32_t p, double nscale) {
  const int D = g.D, V = g.V;
  int c[3] = {0, 0, 0};
  {
    uint32_t rem = p;
    for (int d = 0; d < D; ++d) {
      c[d] = (int)(rem % (uint32_t)g.s[d]);
      rem /= (uint32_t)g.s[d];
    }
  }
  double acc = (double)ni.delta0;
  for (int j = 0; j < ni.nb; ++j) {
    long long t = 0, mul = 1;
    for (int d = 0; d < D; ++d) {
      int x = c[d] + off[j * D + d];
      if (g.circular) x = ((x % g.s[d]) + g.s[d]) % g.s[d];
      t += (long long)x * mul;
      mul *= g.s[d];
    }
    for (int a = 0; a < V; ++a)
      acc = __dadd_rn(acc, __dmul_rn(v[j * V + a], (double)Y[(long long)g.A[a] * S + t]));
  }
  return nscale == 1.0 ? acc : acc * nscale;
}

// The same value for 2-D, non-circular scans with V = VV <= 2 active
// sub-variables: the same sum in the same order, with the TI loads of the
// next kExactBatch neighbours in flight while the products of the current
// ones are summed (the position is valid, so every p + d_j lies inside the
// TI; loads past the last neighbour re-read neighbour nb - 1 and are not
// summed).
constexpr int kExactBatch = 16;

template <int VV>
__device__ __forceinline__ void exact_load(const int* __restrict__ off, const float* __restrict__ Y0,
                                           const float* __restrict__ Y1, int c0, int c1, int s0, int nb, int j0,
                                           float (&y0)[kExactBatch], float (&y1)[kExactBatch]) {
#pragma unroll
  for (int u = 0; u < kExactBatch; ++u) {
    const int j = min(j0 + u, nb - 1);
    const long long t = (long long)(c0 + off[2 * j]) + (long long)(c1 + off[2 * j + 1]) * s0;
    y0[u] = Y0[t];
    if (VV == 2) y1[u] = Y1[t];
  }
}

template <int VV>
__device__ __forceinline__ void exact_sum(const double* __restrict__ v, int nb, int j0,
                                          const float (&y0)[kExactBatch], const float (&y1)[kExactBatch],
                                          double& acc) {
#pragma unroll
  for (int u = 0; u < kExactBatch; ++u) {
    if (j0 + u < nb) {
      acc = __dadd_rn(acc, __dmul_rn(v[(j0 + u) * VV], (double)y0[u]));
      if (VV == 2) acc = __dadd_rn(acc, __dmul_rn(v[(j0 + u) * VV + 1], (double)y1[u]));
    }
  }
}

template <int VV>
__device__ __forceinline__ double exact_value_2d(const ScanGeom& g, const NodeInfo& ni, const int* __restrict__ off,
                                                 const double* __restrict__ v, const float* __restrict__ Y,
                                                 long long S, uint32_t p, double nscale) {
  const int s0 = g.s[0];
  const int c0 = (int)(p % (uint32_t)s0), c1 = (int)(p / (uint32_t)s0);
  const float* Y0 = Y + (long long)g.A[0] * S;
  const float* Y1 = Y + (long long)g.A[VV - 1] * S;
  const int nb = ni.nb;
  double acc = (double)ni.delta0;
  if (nb > 0) {
    float a0[kExactBatch], a1[kExactBatch], b0[kExactBatch], b1[kExactBatch];
    exact_load<VV>(off, Y0, Y1, c0, c1, s0, nb, 0, a0, a1);
    for (int j0 = 0; j0 < nb; j0 += 2 * kExactBatch) {
      exact_load<VV>(off, Y0, Y1, c0, c1, s0, nb, j0 + kExactBatch, b0, b1);
      exact_sum<VV>(v, nb, j0, a0, a1, acc);
      exact_load<VV>(off, Y0, Y1, c0, c1, s0, nb, j0 + 2 * kExactBatch, a0, a1);
      exact_sum<VV>(v, nb, j0 + kExactBatch, b0, b1, acc);
    }
  }
  return nscale == 1.0 ? acc : acc * nscale;
}

__device__ __forceinline__ double exact_value_any(const ScanGeom& g, const NodeInfo& ni, const int* __restrict__ off,
                                                  const double* __restrict__ v, const float* __restrict__ Y,
                                                  long long S, uint32_t p, double nscale) {
  if (g.D == 2 && !g.circular && g.V == 2) return exact_value_2d<2>(g, ni, off, v, Y, S, p, nscale);
  if (g.D == 2 && !g.circular && g.V == 1) return exact_value_2d<1>(g, ni, off, v, Y, S, p, nscale);
  return exact_value(g, ni, off, v, Y, S, p, nscale);
}

// The node's neighbour offsets and alpha-combined values in LDS:
// [nb * D] int, then [nb * V] double (8-byte aligned).
__device__ __forceinline__ void stage_neighbours(const ScanGeom& g, const NodeInfo& ni, const int* __restrict__ nbOff,
                                                 const float* __restrict__ nbVal, char* sm, int** offOut,
                                                 double** valOut) {
  const int D = g.D, V = g.V, nb = ni.nb;
  int* so = reinterpret_cast<int*>(sm);
  double* sv = reinterpret_cast<double*>(sm + (((size_t)nb * D * sizeof(int) + 7) & ~(size_t)7));
  for (int t = threadIdx.x; t < nb * D; t += blockDim.x) so[t] = nbOff[(long long)ni.nbStart * D + t];
  for (int t = threadIdx.x; t < nb * V; t += blockDim.x) sv[t] = (double)nbVal[(long long)ni.nbStart * V + t];
  *offOut = so;
  *valOut = sv;
}

__global__ void k_band_reset(int* __restrict__ qcnt, int* __restrict__ cnt0) {
  if (threadIdx.x < 2) qcnt[threadIdx.x] = 0;
  if (threadIdx.x == 0 && cnt0) *cnt0 = 0;
}

// Step 1. One 256-thread block per node. cand: [node][candGroups][K], gcnt:
// [node][candGroups];
// groups of one rows block are consecutive (wpb per block); rowsPerBlock > 0
// (generic rows kernel): only the blocks covering the node's valid rows
// wrote candidates (one group each).
template <int KT>
__global__ void __launch_bounds__(256) k_merge_band(const NodeInfo* __restrict__ info, const Cand* __restrict__ cand,
                                                    const int* __restrict__ gcnt, int candGroups, int wpb,
                                                    int rowsPerBlock, int K, float kf,
                                                    float* __restrict__ outVal, uint32_t* __restrict__ outPos,
                                                    Cand* __restrict__ outCand, uint32_t* __restrict__ chosen,
                                                    float* __restrict__ simVals, const float* __restrict__ tiVals,
                                                    int nVar, uint32_t* __restrict__ indexOut, BandDev bb,
                                                    int exactOn, const int* __restrict__ count) {
  __shared__ Cand red[16];
  __shared__ Cand best[kMaxK];
  __shared__ uint32_t bpos[kMaxK];
  __shared__ uint32_t flags[kMaxBandBlocks / 32];
  __shared__ int lcnt;
  const int node = blockIdx.x, tid = threadIdx.x;
  if (count && node >= *count) return;
  const NodeInfo ni = info[node];
  if (simVals && ni.nb == 0) {  // drawn from the marginal by the gather stage
    if (tid == 0) bb.cnt[node] = -1;
    return;
  }
  int nGroups = candGroups, nBlocks = candGroups / wpb;
  if (rowsPerBlock > 0) {
    const int n1v = ni.hi[1] - ni.lo[1] + 1;
    const int n2v = ni.hi[2] - ni.lo[2] + 1;
    const int n0v = ni.hi[0] - ni.lo[0] + 1;
    const long long nrv = (n1v > 0 && n2v > 0 && n0v > 0) ? (long long)n1v * n2v : 0;
    nGroups = (int)((nrv + rowsPerBlock - 1) / rowsPerBlock);
    nBlocks = nGroups;
  }
  TopK<KT> tk;
  tk.init();
  const Cand* cn = cand + (long long)node * candGroups * K;
  const int* gc = gcnt + (long long)node * candGroups;
  for (int i = tid; i < nGroups * K; i += blockDim.x) tk.push(cn[i]);
  block_topk<KT>(tk, K, best, red);
  if (!exactOn || !(ni.eps > 0.f)) {  // FP32 selection
    if (tid < K) {
      outVal[(long long)node * K + tid] = best[tid].val;
      outPos[(long long)node * K + tid] = best[tid].pos;
      if (outCand) outCand[(long long)node * K + tid] = best[tid];
      bpos[tid] = best[tid].pos;
    }
    __syncthreads();
    if (tid == 0) {
      bb.cnt[node] = -1;
      if (chosen || simVals) draw_write(ni, bpos, K, kf, chosen, node, simVals, tiVals, nVar, indexOut);
    }
    return;
  }
  // the band: FP32 values >= theta (every position when fewer than K are valid)
  const float th = (best[K - 1].pos == kNone) ? -INFINITY : best[K - 1].val - 2.f * ni.eps * 1.0009765625f;
  for (int i = tid; i < kMaxBandBlocks / 32; i += blockDim.x) flags[i] = 0u;
  if (tid == 0) lcnt = 0;
  __syncthreads();
  // a group may hide band positions when its K-th candidate is in the band
  // and more than K of its values lie within 2 eps of that K-th value (the
  // band lies inside that range: theta >= its K-th value - 2 eps)
  for (int gi = tid; gi < nGroups; gi += blockDim.x) {
    const Cand last = cn[(long long)gi * K + K - 1];
    if (last.pos != kNone && last.val >= th && gc[gi] > K) {
      const int rb = gi / wpb;
      atomicOr(&flags[rb >> 5], 1u << (rb & 31));
    }
  }
  __syncthreads();
  uint32_t* out = bb.pos + (long long)node * kBandCap;
  for (int i = tid; i < nGroups * K; i += blockDim.x) {
    const int rb = (i / K) / wpb;
    if ((flags[rb >> 5] >> (rb & 31)) & 1u) continue;
    const Cand c = cn[i];
    if (c.pos != kNone && c.val >= th) {
      const int k = atomicAdd(&lcnt, 1);
      if (k < kBandCap) out[k] = c.pos;
    }
  }
  int nq = 0;
  for (int rb = tid; rb < nBlocks; rb += blockDim.x) {
    if ((flags[rb >> 5] >> (rb & 31)) & 1u) {
      const int q = atomicAdd(&bb.qcnt[0], 1);  // capacity: every (node, block) of the wave
      bb.queue[q] = make_int2(node, rb);
      ++nq;
    }
  }
  if (nq) atomicAdd(&bb.stats[2], (unsigned long long)nq);
  __syncthreads();
  if (tid == 0) {
    bb.cnt[node] = lcnt;
    bb.theta[node] = th;
    atomicAdd(&bb.stats[0], 1ull);
  }
}

__device__ __forceinline__ int band_tiles(int cnt) {
  return cnt < 0 ? 0 : cnt > kBandCap ? kFullTiles : (cnt + 63) >> 6;
}

// Step 3. Nodes [node0, node0 + nSeg) (nSeg <= kSegNodes), one wavefront
// per block. Every block sums the nodes' tile counts in LDS, then takes tiles
// blockIdx.x, blockIdx.x + gridDim.x, ...: a list tile scores 64 listed
// positions into val, a full-fallback tile scores a slice of the node's valid
// box and writes its K best to part. Dynamic LDS: the tile node's neighbours.
template <int KT>
__global__ void __launch_bounds__(64) k_band_rescore(ScanGeom g, const NodeInfo* __restrict__ info,
                                                      const int* __restrict__ nbOff, const float* __restrict__ nbVal,
                                                      const float* __restrict__ Y, long long S, double nscale, int K,
                                                      int node0, int nSeg, BandDev bb, const int* __restrict__ count) {
  extern __shared__ char sm[];
  __shared__ int pre[kSegNodes + 1];
  __shared__ Cand64 red[1];
  const int tid = threadIdx.x;
  int nLive = nSeg;
  if (count) nLive = min(nSeg, max(0, *count - node0));
  if (nLive <= 0) return;
  // exclusive prefix of the tile counts: thread tid owns a run of nodes
  const int per = (nLive + 63) >> 6;
  const int r0 = min(nLive, tid * per), r1 = min(nLive, r0 + per);
  int run = 0;
  for (int i = r0; i < r1; ++i) run += band_tiles(bb.cnt[node0 + i]);
  int incl = run;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(incl, o);
    if (tid >= o) incl += y;
  }
  int acc = incl - run;
  for (int i = r0; i < r1; ++i) {
    pre[i] = acc;
    acc += band_tiles(bb.cnt[node0 + i]);
  }
  if (tid == 63) pre[nLive] = incl;
  __syncthreads();
  const int total = pre[nLive];
  int cur = -1;
  int* so = nullptr;
  double* sv = nullptr;
  for (int tile = blockIdx.x; tile < total; tile += gridDim.x) {
    // the node holding the tile: the last with pre[n] <= tile
    int lo = 0, hi = nLive - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (pre[mid] <= tile) lo = mid;
      else hi = mid - 1;
    }
    const int node = node0 + lo, lt = tile - pre[lo];
    const NodeInfo ni = info[node];
    if (node != cur) {
      __syncthreads();  // the previous node's neighbours are read no more
      stage_neighbours(g, ni, nbOff, nbVal, sm, &so, &sv);
      __syncthreads();
      cur = node;
    }
    const int cnt = bb.cnt[node];
    if (cnt <= kBandCap) {
      const int i = lt * 64 + tid;
      if (i < cnt) {
        const uint32_t p = bb.pos[(long long)node * kBandCap + i];
        bb.val[(long long)node * kBandCap + i] = exact_value_any(g, ni, so, sv, Y, S, p, nscale);
      }
    } else {  // the list overflowed: slice lt of the whole valid box
      const long long n0 = ni.hi[0] - ni.lo[0] + 1, n1 = ni.hi[1] - ni.lo[1] + 1, n2 = ni.hi[2] - ni.lo[2] + 1;
      const long long box = (n0 > 0 && n1 > 0 && n2 > 0) ? n0 * n1 * n2 : 0;
      const long long len = (box + kFullTiles - 1) / kFullTiles;
      const long long beg = (long long)lt * len, end = min(box, beg + len);
      TopK64<KT> tk;
      tk.init();
      for (long long i = beg + tid; i < end; i += 64) {
        const int c0 = ni.lo[0] + (int)(i % n0);
        const int c1 = ni.lo[1] + (int)((i / n0) % n1);
        const int c2 = ni.lo[2] + (int)(i / (n0 * n1));
        if (ni.exOn && excluded(ni, g.s[0], g.s[1], c0, c1, c2)) continue;
        const uint32_t p = (uint32_t)(c0 + (long long)g.s[0] * (c1 + (long long)g.s[1] * c2));
        Cand64 c;
        c.val = exact_value_any(g, ni, so, sv, Y, S, p, nscale);
        c.h = tie_hash(ni.seedbits, p);
        c.pos = p;
        tk.push(c);
      }
      block_topk64<KT>(tk, K, bb.part + ((long long)node * kFullTiles + lt) * K, red);
    }
  }
}

// Step 4. One block per node: the K best of its listed positions' FP64
// values (of its tiles' K best after a full fallback) by the device rule,
// outputs, the draw and the simulation write. Block 0 clears the second-pass
// queue count for the next call (the collect kernels read it before this
// launch).
template <int KT>
__global__ void __launch_bounds__(256) k_band_final(const NodeInfo* __restrict__ info, int K, float kf,
                                                    float* __restrict__ outVal, uint32_t* __restrict__ outPos,
                                                    Cand* __restrict__ outCand, uint32_t* __restrict__ chosen,
                                                    float* __restrict__ simVals, const float* __restrict__ tiVals,
                                                    int nVar, uint32_t* __restrict__ indexOut, BandDev bb,
                                                    const int* __restrict__ count) {
  __shared__ Cand64 red[4];
  __shared__ Cand64 best[kMaxK];
  __shared__ uint32_t bpos[kMaxK];
  const int node = blockIdx.x, tid = threadIdx.x;
  if (node == 0 && tid == 0) bb.qcnt[0] = 0;
  if (count && node >= *count) return;
  const int cnt = bb.cnt[node];
  if (cnt < 0) return;  // finished by the merge
  const NodeInfo ni = info[node];
  TopK64<KT> tk;
  tk.init();
  if (cnt <= kBandCap) {
    const uint32_t* lst = bb.pos + (long long)node * kBandCap;
    const double* vl = bb.val + (long long)node * kBandCap;
    for (int i = tid; i < cnt; i += blockDim.x) {
      Cand64 c;
      c.pos = lst[i];
      c.val = vl[i];
      c.h = tie_hash(ni.seedbits, c.pos);
      tk.push(c);
    }
    if (tid == 0) atomicAdd(&bb.stats[1], (unsigned long long)cnt);
  } else {
    const Cand64* pp = bb.part + (long long)node * kFullTiles * K;
    for (int i = tid; i < kFullTiles * K; i += blockDim.x) tk.push(pp[i]);
    if (tid == 0) atomicAdd(&bb.stats[3], 1ull);
  }
  block_topk64<KT>(tk, K, best, red);
  if (tid < K) {
    const Cand64 c = best[tid];
    outVal[(long long)node * K + tid] = (float)c.val;
    outPos[(long long)node * K + tid] = c.pos;
    if (outCand) {
      Cand w;
      w.key = (float)c.val;
      w.h = c.h;
      w.pos = c.pos;
      w.val = (float)c.val;
      outCand[(long long)node * K + tid] = w;
    }
    bpos[tid] = c.pos;
  }
  __syncthreads();
  if (tid == 0 && (chosen || simVals)) draw_write(ni, bpos, K, kf, chosen, node, simVals, tiVals, nVar, indexOut);
}

// Per-node API: the band of a materialised map (position space, -inf
// outside the valid box and at excluded positions) from the FP32 K-th best
// outVal[K-1]; node 0 of the band buffers.
__global__ void __launch_bounds__(256) k_map_collect(const float* __restrict__ map, long long S,
                                                     const float* __restrict__ outVal,
                                                     const uint32_t* __restrict__ outPos, int K, float eps,
                                                     BandDev bb) {
  const int lane = threadIdx.x & 63;
  const float th = (outPos[K - 1] == kNone) ? -INFINITY : outVal[K - 1] - 2.f * eps * 1.0009765625f;
  const long long stride = (long long)gridDim.x * blockDim.x;
  const long long n = ((S + 63) / 64) * 64;  // every lane of a wavefront runs the same rounds
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i - lane < n; i += stride) {
    const bool ok = i < S && map[i] >= th && map[i] > -INFINITY;
    const unsigned long long m = __ballot(ok);
    if (m == 0ull) continue;
    const int leader = __ffsll((long long)m) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(&bb.cnt[0], __popcll(m));
    base = __shfl(base, leader);
    if (ok) {
      const int idx = base + __popcll(m & ((1ull << lane) - 1ull));
      if (idx < kBandCap) bb.pos[idx] = (uint32_t)i;
    }
  }
}

int kt_of(int K) { return K <= 2 ? 2 : K <= 4 ? 4 : K <= 8 ? 8 : 16; }

size_t neighbour_lds(const ScanGeom& g, int nbMax) {
  return (((size_t)nbMax * g.D * sizeof(int) + 7) & ~(size_t)7) + (size_t)nbMax * g.V * sizeof(double);
}

}  // namespace

bool exact_selection_on(const qsmi_device* dev) {
  if (dev->quantum != 0.f || dev->cross) return false;
  const qsmi_ti* t = dev->ti;
  const int blocks = row_blocks(dev);
  return blocks <= kMaxBandBlocks;
}

void band_free(BandBufs* b) {
  (void)hipFree(b->pos);
  (void)hipFree(b->cnt);
  (void)hipFree(b->theta);
  (void)hipFree(b->queue);
  (void)hipFree(b->qcnt);
  (void)hipFree(b->val);
  (void)hipFree(b->part);
  (void)hipFree(b->stats);
  *b = BandBufs();
}

int band_grow(BandBufs* b, int nNodes, int blocksPerNode, int K) {
  const size_t n = (size_t)std::max(nNodes, 1);
  const size_t q = n * (size_t)std::max(blocksPerNode, 1);
  if (n <= b->capNodes && q <= b->capQueue && K <= b->K) return QSMI_OK;
  const size_t nn = std::max(n, b->capNodes);
  const size_t qq = std::max(q, b->capQueue);
  const int kk = std::max(K, b->K);
  // grown between launches of one stream: work queued on the old buffers
  // finishes first (the callers' streams do not synchronise with the null
  // stream, so wait for the device before the statistics are read back)
  unsigned long long keep[4] = {0, 0, 0, 0};
  const int zero[2] = {0, 0};  // the queue count starts at zero; k_band_final clears it after each call
  if (b->stats) {
    (void)hipDeviceSynchronize();
    (void)hipMemcpy(keep, b->stats, sizeof(keep), hipMemcpyDeviceToHost);
  }
  band_free(b);
  if (hipMalloc(&b->pos, nn * kBandCap * sizeof(uint32_t)) != hipSuccess ||
      hipMalloc(&b->cnt, nn * sizeof(int)) != hipSuccess || hipMalloc(&b->theta, nn * sizeof(float)) != hipSuccess ||
      hipMalloc(&b->queue, qq * sizeof(int2)) != hipSuccess || hipMalloc(&b->qcnt, 2 * sizeof(int)) != hipSuccess ||
      hipMalloc(&b->val, nn * kBandCap * sizeof(double)) != hipSuccess ||
      hipMalloc(&b->part, nn * kFullTiles * (size_t)kk * sizeof(Cand64)) != hipSuccess ||
      hipMalloc(&b->stats, sizeof(keep)) != hipSuccess ||
      hipMemcpy(b->stats, keep, sizeof(keep), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b->qcnt, zero, sizeof(zero), hipMemcpyHostToDevice) != hipSuccess) {
    band_free(b);
    return fail(QSMI_ERR_NOMEM, "exact selection buffers: out of device memory");
  }
  b->capNodes = nn;
  b->capQueue = qq;
  b->K = kk;
  return QSMI_OK;
}

// Buffers for one node carved from memory the caller owns (base aligned to
// 256 bytes): the layout band_scratch_bytes(K) sizes. band_grow then finds
// them large enough and allocates nothing.
size_t band_scratch_bytes(int K) {
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  return al(kBandCap * sizeof(uint32_t)) + 2 * al(sizeof(int)) + al(sizeof(float)) + al(sizeof(int2)) +
         al(kBandCap * sizeof(double)) + al((size_t)kFullTiles * K * sizeof(Cand64)) +
         al(4 * sizeof(unsigned long long));
}

void band_view(BandBufs* b, void* base, int K) {
  auto al = [](size_t v) { return (v + 255) & ~(size_t)255; };
  char* p = static_cast<char*>(base);
  b->pos = reinterpret_cast<uint32_t*>(p);
  p += al(kBandCap * sizeof(uint32_t));
  b->cnt = reinterpret_cast<int*>(p);
  p += al(sizeof(int));
  b->theta = reinterpret_cast<float*>(p);
  p += al(sizeof(float));
  b->queue = reinterpret_cast<int2*>(p);
  p += al(sizeof(int2));
  b->qcnt = reinterpret_cast<int*>(p);
  p += al(sizeof(int));
  b->val = reinterpret_cast<double*>(p);
  p += al(kBandCap * sizeof(double));
  b->part = reinterpret_cast<Cand64*>(p);
  p += al((size_t)kFullTiles * K * sizeof(Cand64));
  b->stats = reinterpret_cast<unsigned long long*>(p);
  b->capNodes = 1;
  b->capQueue = 1;
  b->K = K;
}

int band_reserve(qsmi_device* dev, BandBufs* bb, int nNodes, int K) {
  const int blocks = use_w1k(dev) ? rows1024_groups() : row_blocks(dev);
  return band_grow(bb, nNodes, exact_selection_on(dev) ? blocks : 1, K);
}

int band_read_stats(const BandBufs* b, long long out[4]) {
  unsigned long long v[4] = {0, 0, 0, 0};
  if (b->stats && hipMemcpy(v, b->stats, sizeof(v), hipMemcpyDeviceToHost) != hipSuccess)
    return fail(QSMI_ERR_HIP, "exact selection statistics");
  for (int i = 0; i < 4; ++i) out[i] = (long long)v[i];
  return QSMI_OK;
}

namespace {

// steps 3 and 4 for nNodes nodes whose lists are in bb
template <int KT>
void launch_rescore(const ScanGeom& g, const NodeInfo* info, const int* nbOff, const float* nbVal, const qsmi_ti* t,
                    double nscale, int nNodes, int K, float kf, float* outVal, uint32_t* outPos, Cand* outCand,
                    uint32_t* chosen, float* simVals, const float* tiVals, int nVar, uint32_t* indexOut,
                    const BandDev& bd, const int* count, size_t lds, hipStream_t st) {
  for (int n0 = 0; n0 < nNodes; n0 += kSegNodes) {
    const int nSeg = std::min(kSegNodes, nNodes - n0);
    // about two tiles per node (most lists hold under 64 positions); blocks
    // beyond the tile total return after the prefix
    const int grid = std::min(8192, 2 * nSeg + 512);
    hipLaunchKernelGGL(k_band_rescore<KT>, dim3((unsigned)grid), dim3(64), lds, st, g, info, nbOff, nbVal, t->dSub,
                       t->S, nscale, K, n0, nSeg, bd, count);
  }
  hipLaunchKernelGGL(k_band_final<KT>, dim3((unsigned)nNodes), dim3(256), 0, st, info, K, kf, outVal, outPos, outCand,
                     chosen, simVals, tiVals, nVar, indexOut, bd, count);
}

}  // namespace

int band_select(qsmi_device* dev, const ScanGeom& g, const NodeInfo* info, const Cand* cand, const int* gcnt,
                int nNodes, int K, float kf, const float2* C, const int* nbOff, const float* nbVal, int nbMax,
                float* outVal, uint32_t* outPos, Cand* outCand, uint32_t* chosen, float* simVals,
                const float* tiVals, int nVar, uint32_t* indexOut, BandBufs* bb, const int* count) {
  qsmi_ti* t = dev->ti;
  hipStream_t st = dev->stream;
  int candGroups, wpb, rpb, blocks;
  if (use_w1k(dev)) {  // one candidate group per row-pair wavefront
    candGroups = rows1024_groups();
    wpb = 1;
    rpb = 0;
    blocks = candGroups;
  } else {
    candGroups = row_blocks(dev);
    wpb = 1;
    rpb = rows_per_block(dev);
    blocks = candGroups;
  }
  if (dev->cross) {  // one row per transform in the cross-measurement form (no exact selection)
    rpb = 1 << t->lognv[0];
    candGroups = (int)(((long long)t->s[1] * t->s[2] + rpb - 1) / rpb);
    blocks = candGroups;
  }
  const bool exact = exact_selection_on(dev);
  int rc = band_grow(bb, nNodes, exact ? blocks : 1, K);
  if (rc) return rc;
  const BandDev bd = dev_view(*bb);
  const int KT = kt_of(K);
#define QSMI_BAND_MERGE(KT_)                                                                                      \
  hipLaunchKernelGGL(k_merge_band<KT_>, dim3((unsigned)nNodes), dim3(256), 0, st, info, cand, gcnt, candGroups, \
                     wpb, rpb, K, kf, outVal, outPos, outCand, chosen, simVals, tiVals, nVar, indexOut, bd,       \
                     exact ? 1 : 0, count)
  switch (KT) {
    case 2: QSMI_BAND_MERGE(2); break;
    case 4: QSMI_BAND_MERGE(4); break;
    case 8: QSMI_BAND_MERGE(8); break;
    default: QSMI_BAND_MERGE(16); break;
  }
#undef QSMI_BAND_MERGE
  if (!exact) return QSMI_OK;
  const int maxItems = nNodes * blocks;
  if (use_w1k(dev))
    launch_rows1024_collect(g, t->dTw1024, info, C, bb->queue, bb->qcnt, bb->theta, bb->pos, bb->cnt, maxItems, st);
  else
    launch_rows_collect(g, t->plan[0], t->lognv[0], t->dTw[0], info, C, bb->queue, bb->qcnt, bb->theta, bb->pos,
                        bb->cnt, maxItems, st);
  const double nscale = dev->trueMismatch ? 1.0 : (double)t->N;
  const size_t lds = neighbour_lds(g, nbMax);
  if (lds > 48 * 1024) return fail(QSMI_ERR_UNSUPPORTED, "exact selection: n x (variables) too large for LDS");
  switch (KT) {
    case 2: launch_rescore<2>(g, info, nbOff, nbVal, t, nscale, nNodes, K, kf, outVal, outPos, outCand, chosen, simVals,
                              tiVals, nVar, indexOut, bd, count, lds, st); break;
    case 4: launch_rescore<4>(g, info, nbOff, nbVal, t, nscale, nNodes, K, kf, outVal, outPos, outCand, chosen, simVals,
                              tiVals, nVar, indexOut, bd, count, lds, st); break;
    case 8: launch_rescore<8>(g, info, nbOff, nbVal, t, nscale, nNodes, K, kf, outVal, outPos, outCand, chosen, simVals,
                              tiVals, nVar, indexOut, bd, count, lds, st); break;
    default: launch_rescore<16>(g, info, nbOff, nbVal, t, nscale, nNodes, K, kf, outVal, outPos, outCand, chosen,
                                simVals, tiVals, nVar, indexOut, bd, count, lds, st); break;
  }
  return QSMI_OK;
}

int band_select_map(qsmi_device* dev, const NodeInfo* info, int nb, int K, float eps, float* outVal,
                    uint32_t* outPos, BandBufs* bb) {
  qsmi_ti* t = dev->ti;
  hipStream_t st = dev->stream;
  int rc = band_grow(bb, 1, 1, K);
  if (rc) return rc;
  const BandDev bd = dev_view(*bb);
  const ScanGeom g = geom_of(dev);
  hipLaunchKernelGGL(k_band_reset, dim3(1), dim3(64), 0, st, bb->qcnt, bb->cnt);
  const long long S = t->S;
  const int blocks = (int)std::min<long long>(1024, (S + 255) / 256);
  hipLaunchKernelGGL(k_map_collect, dim3((unsigned)std::max(blocks, 1)), dim3(256), 0, st, dev->dMap, S, outVal,
                     outPos, K, eps, bd);
  const double nscale = dev->trueMismatch ? 1.0 : (double)t->N;
  const size_t lds = neighbour_lds(g, nb);
  if (lds > 48 * 1024) return fail(QSMI_ERR_UNSUPPORTED, "exact selection: n x (variables) too large for LDS");
  const float kf = (float)K;
  switch (kt_of(K)) {
    case 2: launch_rescore<2>(g, info, dev->dOff, dev->dVal, t, nscale, 1, K, kf, outVal, outPos, nullptr, nullptr,
                              nullptr, nullptr, 0, nullptr, bd, nullptr, lds, st); break;
    case 4: launch_rescore<4>(g, info, dev->dOff, dev->dVal, t, nscale, 1, K, kf, outVal, outPos, nullptr, nullptr,
                              nullptr, nullptr, 0, nullptr, bd, nullptr, lds, st); break;
    case 8: launch_rescore<8>(g, info, dev->dOff, dev->dVal, t, nscale, 1, K, kf, outVal, outPos, nullptr, nullptr,
                              nullptr, nullptr, 0, nullptr, bd, nullptr, lds, st); break;
    default: launch_rescore<16>(g, info, dev->dOff, dev->dVal, t, nscale, 1, K, kf, outVal, outPos, nullptr, nullptr,
                                nullptr, nullptr, 0, nullptr, bd, nullptr, lds, st); break;
  }
  return QSMI_OK;
}

}  // namespace qsmi
