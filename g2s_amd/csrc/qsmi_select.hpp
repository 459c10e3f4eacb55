// Block top-K of the device selection rule (shared by the scan's rows
// kernels and the learned proposal's rescoring kernel).
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "qsmi_kernels.hpp"

namespace qsmi {

// ---------------------------------------------------------------------------
// top-K of the device rule
// ---------------------------------------------------------------------------
__device__ __forceinline__ Cand worst_cand() {
  Cand c;
  c.key = -INFINITY;
  c.h = 0u;
  c.pos = 0xFFFFFFFFu;
  c.val = -INFINITY;
  return c;
}

template <int KT>
struct TopK {
  Cand c[KT];
  __device__ __forceinline__ void init() {
#pragma unroll
    for (int i = 0; i < KT; ++i) c[i] = worst_cand();
  }
  __device__ __forceinline__ void push(Cand cur) {
    if (!cand_better(cur, c[KT - 1])) return;
#pragma unroll
    for (int i = 0; i < KT; ++i) {
      if (cand_better(cur, c[i])) {
        Cand t = c[i];
        c[i] = cur;
        cur = t;
      }
    }
  }
  // value fast path: the hash is only computed when the key can enter
  __device__ __forceinline__ void push_value(float val, float quantum, uint32_t pos, uint32_t seedbits) {
    float key = (quantum > 0.f) ? rintf(val * (1.f / quantum)) : val;
    if (!(key >= c[KT - 1].key)) return;  // also rejects NaN
    Cand cd;
    cd.key = key;
    cd.val = val;
    cd.pos = pos;
    cd.h = tie_hash(seedbits, pos);
    push(cd);
  }
  __device__ __forceinline__ void pop() {
#pragma unroll
    for (int i = 0; i < KT - 1; ++i) c[i] = c[i + 1];
    c[KT - 1] = worst_cand();
  }
};

__device__ __forceinline__ Cand shfl_xor_cand(const Cand& a, int off) {
  Cand o;
  o.key = __shfl_xor(a.key, off);
  o.h = __shfl_xor(a.h, off);
  o.pos = __shfl_xor(a.pos, off);
  o.val = __shfl_xor(a.val, off);
  return o;
}

// Block-wide K best of the per-thread lists, best first, into out[0..K).
// blockDim.x must be a multiple of 64. Ends with a barrier. Returns the K-th
// best (the worst candidate when fewer than K exist) in every thread.
template <int KT>
__device__ Cand block_topk(TopK<KT>& t, int K, Cand* out, Cand* red) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nwaves = blockDim.x >> 6;
  Cand last = worst_cand();
  for (int it = 0; it < K; ++it) {
    Cand best = t.c[0];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      Cand o = shfl_xor_cand(best, off);
      if (cand_better(o, best)) best = o;
    }
    if (lane == 0) red[wave] = best;
    __syncthreads();
    Cand w = red[0];
    for (int i = 1; i < nwaves; ++i)
      if (cand_better(red[i], w)) w = red[i];
    if (tid == 0) out[it] = w;
    if (t.c[0].pos == w.pos && t.c[0].pos != 0xFFFFFFFFu) t.pop();
    last = w;
    __syncthreads();
  }
  return last;
}

// ---------------------------------------------------------------------------
// Block top-K of the device rule in two passes (blockDim a multiple of 64,
// at most 8 wavefronts):
//  1. every thread keeps its KT largest keys (floats only); the block finds
//     tau, a lower bound of its K-th largest key counted with multiplicity;
//  2. only values with key >= tau (normally exactly K per block) get the tie
//     hash and enter a per-thread list ordered by the 64-bit rank
//     (orderable key bits << 32 | hash), from which K rounds of block maxima
//     extract the result. The hash is a bijection of the position for a
//     fixed seed, so ranks are distinct and (key desc, hash desc) is the
//     whole order of the rule (the position tie-break never triggers).
// ---------------------------------------------------------------------------
struct TopkScratch {
  float f[2][8];
  int n[2][8];
  unsigned long long r[2][8];
  uint32_t pos;
  float val;
};

// orderable image of a key: larger key -> larger integer; -0 folds into +0
__device__ __forceinline__ uint32_t key_bits(float k) {
  uint32_t u = __float_as_uint(k);
  if (u == 0x80000000u) u = 0u;
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__device__ __forceinline__ float key_of_bits(uint32_t b) {
  return __uint_as_float((b & 0x80000000u) ? (b & 0x7FFFFFFFu) : ~b);
}

// tau <= the block's K-th largest key (multiplicity counted); -inf when the
// block holds fewer than K keys. m: this thread's KT largest keys, descending
// (consumed). Block-uniform result; ends after a barrier.
template <int KT>
__device__ float block_kth_key(float (&m)[KT], int K, TopkScratch* s) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nw = blockDim.x >> 6;
  int need = K;
  float tau = -INFINITY;
  for (int it = 0; it < K; ++it) {
    const int b = it & 1;
    float M = m[0];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) M = fmaxf(M, __shfl_xor(M, off));
    if (lane == 0) s->f[b][wave] = M;
    __syncthreads();
    M = s->f[b][0];
    for (int i = 1; i < nw; ++i) M = fmaxf(M, s->f[b][i]);
    tau = M;
    if (M == -INFINITY) break;
    int c = 0;
#pragma unroll
    for (int i = 0; i < KT; ++i) c += (m[i] == M) ? 1 : 0;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off);
    if (lane == 0) s->n[b][wave] = c;
    __syncthreads();
    int tot = 0;
    for (int i = 0; i < nw; ++i) tot += s->n[b][i];
    if (tot >= need) break;
    need -= tot;
    // the entries equal to M sit at the head of the list: pop them
#pragma unroll
    for (int r = 0; r < KT; ++r) {
      if (m[0] == M) {
#pragma unroll
        for (int i = 0; i < KT - 1; ++i) m[i] = m[i + 1];
        m[KT - 1] = -INFINITY;
      }
    }
  }
  return tau;
}

template <int KT>
struct RankList {
  unsigned long long r[KT];  // 0 = empty
  uint32_t pos[KT];
  float val[KT];
  __device__ __forceinline__ void init() {
#pragma unroll
    for (int i = 0; i < KT; ++i) {
      r[i] = 0ull;
      pos[i] = 0xFFFFFFFFu;
      val[i] = -INFINITY;
    }
  }
  __device__ __forceinline__ void insert(float key, uint32_t h, uint32_t p, float v) {
    unsigned long long c = ((unsigned long long)key_bits(key) << 32) | h;
#pragma unroll
    for (int i = 0; i < KT; ++i) {
      if (c > r[i]) {
        const unsigned long long tc = r[i];
        const uint32_t tp = pos[i];
        const float tv = val[i];
        r[i] = c;
        pos[i] = p;
        val[i] = v;
        c = tc;
        p = tp;
        v = tv;
      }
    }
  }
  // K rounds of block maxima -> out[0..K), best first; ends after a barrier
  __device__ void block_extract(int K, Cand* out, TopkScratch* s) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nw = blockDim.x >> 6;
    for (int it = 0; it < K; ++it) {
      const int b = it & 1;
      unsigned long long M = r[0];
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(M, off);
        M = o > M ? o : M;
      }
      if (lane == 0) s->r[b][wave] = M;
      __syncthreads();
      M = s->r[b][0];
      for (int i = 1; i < nw; ++i) M = s->r[b][i] > M ? s->r[b][i] : M;
      if (M != 0ull && r[0] == M) {  // exactly one thread holds it
        s->pos = pos[0];
        s->val = val[0];
#pragma unroll
        for (int i = 0; i < KT - 1; ++i) {
          r[i] = r[i + 1];
          pos[i] = pos[i + 1];
          val[i] = val[i + 1];
        }
        r[KT - 1] = 0ull;
        pos[KT - 1] = 0xFFFFFFFFu;
        val[KT - 1] = -INFINITY;
      }
      __syncthreads();
      if (tid == 0) {
        Cand w;
        if (M != 0ull) {
          w.key = key_of_bits((uint32_t)(M >> 32));
          w.h = (uint32_t)M;
          w.pos = s->pos;
          w.val = s->val;
        } else {
          w.key = -INFINITY;
          w.h = 0u;
          w.pos = 0xFFFFFFFFu;
          w.val = -INFINITY;
        }
        out[it] = w;
      }
    }
  }
};



// ---------------------------------------------------------------------------
// The same two passes inside one wavefront (shuffles only, no barriers).
// ---------------------------------------------------------------------------
template <int KT>
__device__ float wave_kth_key(float (&m)[KT], int K) {
  int need = K;
  float tau = -INFINITY;
  for (int it = 0; it < K; ++it) {
    float M = m[0];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) M = fmaxf(M, __shfl_xor(M, off));
    tau = M;
    if (M == -INFINITY) break;
    int c = 0;
#pragma unroll
    for (int i = 0; i < KT; ++i) c += (m[i] == M) ? 1 : 0;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off);
    if (c >= need) break;
    need -= c;
#pragma unroll
    for (int r = 0; r < KT; ++r) {
      if (m[0] == M) {
#pragma unroll
        for (int i = 0; i < KT - 1; ++i) m[i] = m[i + 1];
        m[KT - 1] = -INFINITY;
      }
    }
  }
  return tau;
}

// K rounds of wavefront maxima of a RankList -> out[0..K) (lane 0 writes)
template <int KT>
__device__ void wave_extract(RankList<KT>& rl, int K, Cand* out) {
  const int lane = threadIdx.x & 63;
  for (int it = 0; it < K; ++it) {
    unsigned long long M = rl.r[0];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const unsigned long long o = __shfl_xor(M, off);
      M = o > M ? o : M;
    }
    const bool mine = M != 0ull && rl.r[0] == M;  // at most one lane
    const unsigned long long ball = __ballot(mine);
    const int src = ball ? __ffsll((long long)ball) - 1 : 0;
    const uint32_t pos = __shfl(rl.pos[0], src);
    const float val = __shfl(rl.val[0], src);
    if (mine) {
#pragma unroll
      for (int i = 0; i < KT - 1; ++i) {
        rl.r[i] = rl.r[i + 1];
        rl.pos[i] = rl.pos[i + 1];
        rl.val[i] = rl.val[i + 1];
      }
      rl.r[KT - 1] = 0ull;
      rl.pos[KT - 1] = 0xFFFFFFFFu;
      rl.val[KT - 1] = -INFINITY;
    }
    if (lane == 0) {
      Cand w;
      if (M != 0ull) {
        w.key = key_of_bits((uint32_t)(M >> 32));
        w.h = (uint32_t)M;
        w.pos = pos;
        w.val = val;
      } else {
        w.key = -INFINITY;
        w.h = 0u;
        w.pos = 0xFFFFFFFFu;
        w.val = -INFINITY;
      }
      out[it] = w;
    }
  }
}

// ---------------------------------------------------------------------------
// Top-K of the exact selection (FP64 values, device rule).
// ---------------------------------------------------------------------------
__device__ __forceinline__ Cand64 worst_cand64() {
  Cand64 c;
  c.val = -INFINITY;
  c.h = 0u;
  c.pos = 0xFFFFFFFFu;
  return c;
}

template <int KT>
struct TopK64 {
  Cand64 c[KT];
  __device__ __forceinline__ void init() {
#pragma unroll
    for (int i = 0; i < KT; ++i) c[i] = worst_cand64();
  }
  __device__ __forceinline__ void push(Cand64 cur) {
    if (!cand64_better(cur, c[KT - 1])) return;
#pragma unroll
    for (int i = 0; i < KT; ++i) {
      if (cand64_better(cur, c[i])) {
        Cand64 t = c[i];
        c[i] = cur;
        cur = t;
      }
    }
  }
  __device__ __forceinline__ void pop() {
#pragma unroll
    for (int i = 0; i < KT - 1; ++i) c[i] = c[i + 1];
    c[KT - 1] = worst_cand64();
  }
};

__device__ __forceinline__ Cand64 shfl_xor_cand64(const Cand64& a, int off) {
  Cand64 o;
  o.val = __shfl_xor(a.val, off);
  o.h = __shfl_xor(a.h, off);
  o.pos = __shfl_xor(a.pos, off);
  return o;
}

// Block-wide K best, best first, into out[0..K) (LDS or global); red holds
// one entry per wavefront. blockDim.x a multiple of 64. Ends with a barrier.
template <int KT>
__device__ void block_topk64(TopK64<KT>& t, int K, Cand64* out, Cand64* red) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nwaves = blockDim.x >> 6;
  for (int it = 0; it < K; ++it) {
    Cand64 best = t.c[0];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      Cand64 o = shfl_xor_cand64(best, off);
      if (cand64_better(o, best)) best = o;
    }
    if (lane == 0) red[wave] = best;
    __syncthreads();
    Cand64 w = red[0];
    for (int i = 1; i < nwaves; ++i)
      if (cand64_better(red[i], w)) w = red[i];
    if (tid == 0) out[it] = w;
    if (t.c[0].pos == w.pos && t.c[0].pos != 0xFFFFFFFFu) t.pop();
    __syncthreads();
  }
}

// QuantileSamplingModule::sample's rank for one TI (include/
// quantileSamplingModule.hpp:101-113): floor(seed * k * (ceil(1/k)/1)), k
// from the node's k image when given (include/simulation.hpp:355-357),
// clamped to the node's list; then the simulation write (include/
// simulation.hpp:658-666): the TI value at the chosen position for every
// variable the cell does not hold yet (a partially informed cell keeps its
// values) and importDataIndex = index * nTI + TI (one TI). pos: the node's
// K best, best first (0xFFFFFFFF: none).
__device__ __forceinline__ void draw_write(const NodeInfo& ni, const uint32_t* pos, int K, float kf,
                                           uint32_t* chosen, int node, float* simVals, const float* tiVals,
                                           int nVar, uint32_t* indexOut) {
  const float kk = ni.kLocal > 0.f ? ni.kLocal : kf;
  int r = (int)floorf(ni.seed * kk * (ceilf(1.f / kk) / 1.f));
  if (r > (int)ceilf(kk) - 1) r = (int)ceilf(kk) - 1;  // the node's list holds ceil(k)
  if (r > K - 1) r = K - 1;
  if (r < 0) r = 0;
  const uint32_t p = pos[r];
  if (chosen) chosen[node] = p;
  if (simVals && p != 0xFFFFFFFFu) {
    for (int v = 0; v < nVar; ++v) {
      float* dst = simVals + (long long)ni.cell * nVar + v;
      if (isnan(*dst)) *dst = tiVals[(long long)p * nVar + v];
    }
    indexOut[ni.cell] = p;
  }
}

}  // namespace qsmi
