// Tap lists of the 1024-grid spectrum stage, built by one wavefront per
// node (k_taps1024, or the path loop's gather kernel right after it writes
// the node's neighbours). Included by the translation units of both kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "qsmi_kernels.hpp"

namespace qsmi {

constexpr int kTapF = 1024;  // rows of the scan grid

// Distinct neighbour rows ("taps", row = (-d1) mod 1024) of one node:
// TapInfo1024 with the rows ordered by (row mod 16, row) and every tap's
// slot range; the slots hold the neighbours in tap order (a tap's neighbours
// in neighbour-list order, so every tap sum has a fixed summation order),
// each with d0 mod 1024 (tapD0) and its V values (tapVal). T = 0 (no lists)
// above kTapNbMax neighbours. Called by all 64 lanes of one wavefront (a
// 64-thread block) with the node's neighbours [nbStart, nbStart + nb) of
// nbOff / nbVal already visible to it.
__device__ __forceinline__ void build_taps1024(int node, int nb, int nbStart, const int* __restrict__ nbOff,
                                               const float* __restrict__ nbVal, int V,
                                               TapInfo1024* __restrict__ taps, uint16_t* __restrict__ tapD0,
                                               float* __restrict__ tapVal) {
  __shared__ uint8_t flag[kTapF];
  __shared__ short tix[kTapF];       // tap index of a present row
  __shared__ short tapJ[kTapNbMax];  // tap of neighbour j
  __shared__ int cntT[kTapNbMax];
  __shared__ int startL[kTapNbMax + 1];
  const int lane = threadIdx.x & 63;
  TapInfo1024* tp = taps + node;
  if (nb > kTapNbMax || nb <= 0) {  // block-uniform
    if (lane == 0) {
      tp->T = 0;
      tp->passes = 0;
    }
    return;
  }
  for (int i = lane; i < kTapF; i += 64) flag[i] = 0;
  __syncthreads();
  for (int t = lane; t < nb; t += 64) flag[(-nbOff[(nbStart + t) * 2 + 1]) & (kTapF - 1)] = 1;
  __syncthreads();
  int count = 0, kmax = 0;
  if (lane == 0) tp->cls[0] = 0;
  for (int c = 0; c < 16; ++c) {
    const int row = c + 16 * lane;  // the 64 rows of class c
    const bool present = flag[row] != 0;
    const unsigned long long mask = __ballot(present);
    const int before = __popcll(mask & ((1ull << lane) - 1ull));
    if (present) {  // count + before < nb <= kTapNbMax
      tp->row[count + before] = (uint16_t)row;
      tix[row] = (short)(count + before);
      // tap path slot: the before-th tap of class c
      if (count + before < kTapMax && before < kTapPass) tp->slot[count + before] = (uint8_t)(16 * before + c);
    }
    count += __popcll(mask);
    kmax = max(kmax, __popcll(mask));
    if (lane == 0) tp->cls[c + 1] = count;
  }
  const int T = count;
  if (lane == 0) {
    tp->T = T;
    tp->passes = (T <= kTapMax && kmax <= kTapPass) ? kmax : 0;
  }
  for (int t = lane; t < T; t += 64) cntT[t] = 0;
  __syncthreads();
  for (int t = lane; t < nb; t += 64) {
    const short tt = tix[(-nbOff[(nbStart + t) * 2 + 1]) & (kTapF - 1)];
    tapJ[t] = tt;
    atomicAdd(&cntT[tt], 1);
  }
  __syncthreads();
  if (lane == 0) {
    int acc = 0;
    for (int t = 0; t < T; ++t) {
      startL[t] = acc;
      tp->start[t] = (uint16_t)acc;
      acc += cntT[t];
    }
    startL[T] = acc;
    tp->start[T] = (uint16_t)acc;
  }
  __syncthreads();
  uint16_t* od = tapD0 + (long long)node * kTapNbMax;
  float* ov = tapVal + (long long)node * kTapNbMax * V;
  // neighbour t's slot: its tap's start plus the number of earlier
  // neighbours on the same tap, by one ballot per tap over 64 neighbours at
  // a time (startL[u] advances past each round's neighbours of tap u)
  for (int t0 = 0; t0 < nb; t0 += 64) {  // wave-uniform
    const int t = t0 + lane;
    const int tt = t < nb ? (int)tapJ[t] : -1;
    int slot = 0;
    for (int u = 0; u < T; ++u) {  // wave-uniform
      const unsigned long long m = __ballot(tt == u);
      if (m == 0ull) continue;
      const int base = startL[u];
      if (tt == u) slot = base + __popcll(m & ((1ull << lane) - 1ull));
      if (lane == 0) startL[u] = base + __popcll(m);  // after every lane's read (in order)
    }
    if (t < nb) {
      const long long jj = nbStart + t;
      od[slot] = (uint16_t)(nbOff[jj * 2 + 0] & (kTapF - 1));
      for (int a = 0; a < V; ++a) ov[(long long)slot * V + a] = nbVal[jj * V + a];
    }
  }
}

}  // namespace qsmi
