// Device-resident QuickSampling path loop: the simulation grid, the index
// image, the TI values and every node's neighbour list live in HBM; per path
// wave a gather kernel builds the scan inputs of each node from the current
// grid (include/simulation.hpp:447-537 and the QuantileSamplingModule
// conversion, include/quantileSamplingModule.hpp:368-457), the scan kernels
// run, and the merge kernel writes the drawn TI values back into the grid
// (simulation.hpp:658-666). The host only enqueues launches; waves follow each
// other in stream order, so a wave reads what earlier waves wrote.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "qsmi_internal.hpp"
#include "qsmi_taps.hpp"

namespace qsmi {

struct GatherParams {
  int D;
  int sd[3];       // simulation grid
  int nVar, nExp, nSub, V, n;
  int kd[3];       // expanded kernel grid
  int kctr[3];     // QSM kernel centre dims/2
  int types[8];    // per original variable: 0 continuous, 1 categorical
  int expVar[16];  // expanded index -> original variable
  int expCat[16];  // expanded index -> offset of its class value in catvals, -1 continuous
  int conv[16];    // per neighbour sub-variable: 0 = P0 (w), 1 = P1 (w*x), 2 = P2 (w*x^2)
  int subOf[16];   // per neighbour sub-variable: expanded index
  int act[16];     // active TI sub-variables (alpha row non-zero)
  float alpha[256];  // [A + B*nSub]
  float alphaX[256]; // cross-measurement form: the count layer's coefficients
  int cross;         // 1: TI with NaN (two layers per node, no delta0)
  int s[3];        // TI grid (valid box)
  int circular;
  float scale;     // map scale (1/N true mismatch, 1 otherwise)
  int trueMismatch;
  float quantum;
  unsigned tiCells;
  int noVerbatim;  // 1: exclude the position that continues the nearest neighbour's copy (-nV)
  unsigned f[3];   // TI FFT grid (the reference's index space)
  int nTI;         // training images (tiVals holds them one after the other)
  unsigned tiCum[9];  // first cell of each TI in tiVals, then the total
  int circSim;          // 1: the simulation grid wraps around (-csim)
  int kdK[8][3];        // per kernel of a kernel-index image: its grid (kernel 0: kd)
  long long kerOff[8];  // its offset in the kernel array (floats)
  float rms[16];        // per active sub-variable: rms of its TI sub-image (FFT grid)
  float epsScale;       // kMapErrC * scale * N; 0: exact selection off (tie quantum, NaN TI)
};

__device__ __forceinline__ uint32_t seed_bits_dev(float seed) {
  if (!isfinite(seed) || seed <= 0.f) return 0u;
  if (seed >= 1.f) return 0xFFFFFFFFu;
  return (uint32_t)floor((double)seed * 4294967295.0);
}

// One wavefront per node of the wave (order[w0 + q]). In the cross-
// measurement form node q's count layer is written as node nN + q (same
// neighbours, the count-layer values, an offset copy of its own). With taps
// (1024-grid scans) the wavefront then builds the node's tap lists
// (build_taps1024) from the neighbours it wrote. wave (a replayed graph):
// {w0, nN} read on the device, blocks from nN on return.
__global__ void __launch_bounds__(64) k_gather(GatherParams P, int w0, int nN, const int* __restrict__ order,
                                               const unsigned* __restrict__ pathCells,
                                               const float* __restrict__ pathSeeds, const int* __restrict__ nbIdx,
                                               const int* __restrict__ nbCnt, const int* __restrict__ offs,
                                               const float* __restrict__ kern, const float* __restrict__ catvals,
                                               float* __restrict__ di, uint32_t* __restrict__ index,
                                               const float* __restrict__ tiVals, NodeInfo* __restrict__ info,
                                               int* __restrict__ nbOff, float* __restrict__ nbVal,
                                               const uint8_t* __restrict__ centreOk,
                                               const uint32_t* __restrict__ complete, uint32_t nComplete,
                                               const float* __restrict__ pathK, const int* __restrict__ nodeKer,
                                               const int* __restrict__ nodeTi, TapInfo1024* __restrict__ taps,
                                               uint16_t* __restrict__ tapD0, float* __restrict__ tapVal,
                                               const int* __restrict__ wave) {
  const int q = blockIdx.x;
  const int lane = threadIdx.x;
  if (wave) {
    w0 = wave[0];
    nN = wave[1];
    if (q >= nN) return;
  }
  const int i = order[w0 + q];
  const unsigned cell = pathCells[i];
  const float seed = pathSeeds[i];
  const int cnt = nbCnt[i];
  const int D = P.D;
  // the node's kernel (include/simulation.hpp:348-351: a kernel-index image
  // picks one per cell), centre dims/2 (quantileSamplingModule.hpp:362-366)
  const int kk = nodeKer ? nodeKer[i] : 0;
  const int kdv[3] = {P.kdK[kk][0], P.kdK[kk][1], P.kdK[kk][2]};
  const int kctr[3] = {kdv[0] / 2, kdv[1] / 2, kdv[2] / 2};
  const float* __restrict__ kn = kern + P.kerOff[kk];
  NodeInfo ni;
  ni.cell = cell;
  ni.seed = seed;
  ni.seedbits = seed_bits_dev(seed);
  ni.quantum = P.quantum;
  ni.scale = P.scale;
  ni.nbStart = q * P.n;
  ni.cmin = 0.f;
  ni.exOn = 0;
  ni.exPos = 0;
  ni.exR2 = 0.f;
  ni.kLocal = pathK ? pathK[i] : 0.f;
  ni.tiReq = nodeTi ? nodeTi[i] : -1;
  ni.eps = 0.f;
  // only the cell itself, holding no value: draw from the marginal
  // (simulation.hpp:586-606); a partially informed cell with only itself as
  // neighbour is scanned with its own values (simulation.hpp:581-584)
  bool centreData = false;
  if (cnt == 1)
    for (int v = 0; v < P.nVar; ++v) centreData |= !isnan(di[(long long)cell * P.nVar + v]);
  if (cnt <= 1 && !centreData) {
    if (lane == 0) {
      // over all TIs in order (include/simulation.hpp:586-606): the TI whose
      // cells hold the position
      unsigned position = (unsigned)(int)floorf(seed * (float)P.tiCum[P.nTI]);
      if (P.cross && !centreOk[position]) {
        // incomplete cell: rank among the complete cells (simulation.hpp:
        // 617-654; rank 0 keeps cell 0 as the reference's walk does)
        const unsigned r = (unsigned)(int)floorf(seed * (float)(nComplete / (unsigned)P.nVar)) * (unsigned)P.nVar;
        position = r == 0 ? 0u : complete[r - 1];
      }
      for (int v = 0; v < P.nVar; ++v)  // the variables the cell does not hold
        if (isnan(di[(long long)cell * P.nVar + v])) di[(long long)cell * P.nVar + v] = tiVals[(long long)position * P.nVar + v];
      int tt = 0;
      while (tt + 1 < P.nTI && position >= P.tiCum[tt + 1]) ++tt;
      index[cell] = (position - P.tiCum[tt]) * (unsigned)P.nTI + (unsigned)tt;
      ni.nb = 0;
      ni.delta0 = 0.f;
      for (int d = 0; d < 3; ++d) {
        ni.lo[d] = 1;
        ni.hi[d] = 0;
      }
      info[q] = ni;
      if (P.cross) info[nN + q] = ni;
    }
    if (taps) build_taps1024(q, 0, q * P.n, nbOff, nbVal, P.V, taps, tapD0, tapVal);  // no lists
    return;
  }
  int cc[3] = {0, 0, 0};
  {
    unsigned rem = cell;
    for (int d = 0; d < D; ++d) {
      cc[d] = (int)(rem % (unsigned)P.sd[d]);
      rem /= (unsigned)P.sd[d];
    }
  }
  const int* idx = nbIdx + (long long)i * P.n;
  int mn[3] = {0, 0, 0}, mx[3] = {0, 0, 0};
  float wmin = INFINITY;  // smallest weight of a neighbour value (cross form)
  float eAcc = 0.f;       // sum_a rms_a |v_ja| over this lane's neighbours (map error bound)
  for (int j = lane; j < cnt; j += 64) {
    const int* dd = offs + (long long)idx[j] * D;
    long long t = 0, mul = 1;
    bool inK = true;
    long long kc = 0;
    for (int d = 0; d < D; ++d) {
      int x = cc[d] + dd[d];
      if (P.circSim) x = ((x % P.sd[d]) + P.sd[d]) % P.sd[d];
      t += (long long)x * mul;
      mul *= P.sd[d];
      mn[d] = min(mn[d], dd[d]);
      mx[d] = max(mx[d], dd[d]);
    }
    for (int d = D - 1; d >= 0; --d) {  // DataImage::indexWithDelta from the kernel centre
      int c = kctr[d] + dd[d];
      inK &= (c >= 0 && c < kdv[d]);
      c = ((c % kdv[d]) + kdv[d]) % kdv[d];
      kc = kc * kdv[d] + c;
    }
    const float* x = di + t * P.nVar;
    for (int d = 0; d < D; ++d) nbOff[((long long)q * P.n + j) * D + d] = dd[d];
    if (P.cross) {
      for (int d = 0; d < D; ++d) nbOff[((long long)(nN + q) * P.n + j) * D + d] = dd[d];
      for (int v = 0; v < P.nVar; ++v) {
        const float w = kn[kc * P.nExp + v];
        if (inK && !isnan(x[v]) && w > 0.f) wmin = fminf(wmin, w);
      }
    }
    // v_a = sum_B alpha(A_a, B) u_B with u_B the QSM conversion of sub-variable B
    for (int layer = 0; layer < (P.cross ? 2 : 1); ++layer) {
      const float* al_ = layer == 0 ? P.alpha : P.alphaX;
      const long long row = (long long)(layer == 0 ? q : nN + q) * P.n + j;
      for (int a = 0; a < P.V; ++a) {
        const int A = P.act[a];
        float acc = 0.f;
        for (int B = 0; B < P.nSub; ++B) {
          const float al = al_[A + B * P.nSub];
          if (al == 0.f) continue;
          const int e = P.subOf[B];
          const float xv = x[P.expVar[e]];
          const float xe = (P.expCat[e] < 0) ? xv : ((xv == catvals[P.expCat[e]]) ? 1.f : 0.f);
          const float w = kn[kc * P.nExp + e];
          const float u = (inK && !isnan(xe)) ? (P.conv[B] == 0 ? w * 1.f : P.conv[B] == 2 ? w * xe * xe : w * xe) : 0.f;
          acc += al * u;
        }
        nbVal[row * P.V + a] = acc;
        if (layer == 0) eAcc += P.rms[a] * fabsf(acc);
      }
    }
  }
  // bounding box (min/max start at 0, src/AcceleratorDevice.cpp:95-108)
  for (int d = 0; d < D; ++d) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      mn[d] = min(mn[d], __shfl_xor(mn[d], off));
      mx[d] = max(mx[d], __shfl_xor(mx[d], off));
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) wmin = fminf(wmin, __shfl_xor(wmin, off));
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) eAcc += __shfl_xor(eAcc, off);
  // delta0 in the reference's order: variable by variable, neighbour by
  // neighbour, float accumulation of -w x^2 read at expanded index v. The
  // terms are formed in parallel (one lane per neighbour) and summed by lane
  // 0 in that order, so the float result is the sequential one.
  extern __shared__ float term[];  // [nVar][n] terms, then [nVar][n] presence flags
  unsigned char* has = reinterpret_cast<unsigned char*>(term + P.nVar * P.n);
  const int nT = cnt;
  for (int j = lane; j < nT; j += 64) {
    const int* dd = offs + (long long)idx[j] * D;
    long long t = 0, mul = 1, kc = 0;
    bool inK = true;
    for (int d = 0; d < D; ++d) {
      int x = cc[d] + dd[d];
      if (P.circSim) x = ((x % P.sd[d]) + P.sd[d]) % P.sd[d];
      t += (long long)x * mul;
      mul *= P.sd[d];
    }
    for (int d = D - 1; d >= 0; --d) {
      int c = kctr[d] + dd[d];
      inK &= (c >= 0 && c < kdv[d]);
      c = ((c % kdv[d]) + kdv[d]) % kdv[d];
      kc = kc * kdv[d] + c;
    }
    for (int v = 0; v < P.nVar; ++v) {
      float tv = 0.f;
      bool h = false;
      if (P.types[v] == 0) {
        const int e = v;  // the reference reads the expanded value at the original index
        const float xv = di[t * P.nVar + P.expVar[e]];
        const float xe = (P.expCat[e] < 0) ? xv : ((xv == catvals[P.expCat[e]]) ? 1.f : 0.f);
        if (inK && !isnan(xe)) {
          tv = -1.f * (kn[kc * P.nExp + (v % P.nExp)] * xe * xe);
          h = true;
        }
      }
      term[v * P.n + j] = tv;
      has[v * P.n + j] = h ? 1 : 0;
    }
  }
  __syncthreads();
  // the sequential sum, in that order, on every lane: each round of 64 terms
  // is read from LDS by all lanes at once and added lane by lane from
  // registers (a term without a value adds +0, which leaves the sum as it is)
  float delta0 = 0.f;
  for (int v = 0; v < P.nVar; ++v) {
    if (P.types[v] != 0) continue;
    for (int j0 = 0; j0 < nT; j0 += 64) {
      const int j = j0 + lane;
      const int tb = __float_as_int((j < nT && has[v * P.n + j]) ? term[v * P.n + j] : 0.f);
      const int cntR = min(64, nT - j0);
      for (int q = 0; q < cntR; ++q) delta0 += __int_as_float(__builtin_amdgcn_readlane(tb, q));
    }
  }
  if (lane == 0) {
    for (int d = 0; d < 3; ++d) {
      if (d >= D) {
        ni.lo[d] = ni.hi[d] = 0;
      } else if (P.circular) {
        ni.lo[d] = 0;
        ni.hi[d] = P.s[d] - 1;
      } else {
        ni.lo[d] = -mn[d];
        ni.hi[d] = P.s[d] - 1 - mx[d];
      }
    }
    ni.nb = cnt;
    ni.delta0 = (P.trueMismatch && !P.cross) ? delta0 : 0.f;
    ni.eps = P.epsScale * eAcc;
    if (P.noVerbatim) {
      // include/simulation.hpp:567-580: the TI position the nearest
      // neighbour was copied from, moved back by its offset (wrapped, its
      // validity ignored), then QuantileSamplingModule's
      // setValueInErrorArray(cvtPositionToIndex(.)) (include/
      // quantileSamplingModule.hpp:566-571): the map entry at the index of
      // that position, i.e. position cvtIndexToPosition(index), both maps
      // with the node's bounding-box minimum in unsigned arithmetic
      // (src/CPUThreadDevice.cpp:257-284)
      const int* d1 = offs + (long long)idx[1] * D;
      long long t1 = 0, mul = 1;
      for (int d = 0; d < D; ++d) {
        int x = cc[d] + d1[d];
        if (P.circSim) x = ((x % P.sd[d]) + P.sd[d]) % P.sd[d];
        t1 += (long long)x * mul;
        mul *= P.sd[d];
      }
      unsigned rem = index[t1];
      unsigned v[3] = {0, 0, 0};
      for (int d = 0; d < D; ++d) {
        int c = (int)(rem % (unsigned)P.s[d]) - d1[d];
        c = ((c % P.s[d]) + P.s[d]) % P.s[d];
        v[d] = (unsigned)c;
        rem /= (unsigned)P.s[d];
      }
      unsigned vpos = 0;
      for (int d = D - 1; d >= 0; --d) vpos = vpos * (unsigned)P.s[d] + v[d];
      unsigned S = 1, N = 1;
      for (int d = 0; d < D; ++d) {
        S *= (unsigned)P.s[d];
        N *= P.f[d];
      }
      unsigned idxv = 0, div = S;
      for (int d = D - 1; d >= 0; --d) {
        const unsigned f = P.f[d], sd = (unsigned)P.s[d];
        div /= sd;
        idxv = idxv * f + (f - (vpos / div + (unsigned)mn[d]) % sd - 1);
      }
      unsigned pos = 0;
      div = N;
      for (int d = D - 1; d >= 0; --d) {
        const unsigned f = P.f[d], sd = (unsigned)P.s[d];
        div /= f;
        pos = pos * sd + ((f - (idxv / div) % f - (unsigned)mn[d] - 1) + sd) % sd;
      }
      ni.exOn = 1;
      ni.exPos = pos;
      ni.exR2 = 0.f;
    }
    // the count layer is a sum of neighbour weights (0 or >= the smallest
    // one): below half of it the FFT value is noise around zero
    ni.cmin = (P.cross && wmin < INFINITY) ? 0.5f * wmin : 0.f;
    info[q] = ni;
    if (P.cross) {
      ni.nbStart = (nN + q) * P.n;
      ni.delta0 = 0.f;
      info[nN + q] = ni;
    }
  }
  if (taps) {
    __syncthreads();  // this wavefront's neighbour offsets and values are written
    build_taps1024(q, cnt, q * P.n, nbOff, nbVal, P.V, taps, tapD0, tapVal);
  }
}

// ---------------------------------------------------------------------------
// Neighbour sets of every path node on the device (include/simulation.hpp:
// 447-511: walk the distance-sorted offset list and take the first n cells
// informed before the node, the cell itself first). One wavefront per node
// tests 64 offsets at a time; a ballot keeps the hits in offset order, so the
// result is the serial walk's. postS[cell] = -1 for hard data, the path index
// otherwise. Outputs: nbIdx [node][n] offset indices, nbPj [node][n] the
// neighbours' path indices (-1: hard data), nbCnt [node].
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(64) k_nbsearch(int D, int sd0, int sd1, int sd2, const unsigned* __restrict__ cells,
                                                const int* __restrict__ postS, const int* __restrict__ offs, int nOff,
                                                int n, int* __restrict__ nbIdx, int* __restrict__ nbPj,
                                                int* __restrict__ nbCnt, const int* __restrict__ nNode,
                                                const int* __restrict__ nodeKer, const int* __restrict__ offStart,
                                                const int* __restrict__ offCount, int circSim) {
  const long long i = blockIdx.x;
  const int lane = threadIdx.x;
  const int sd[3] = {sd0, sd1, sd2};
  int cc[3] = {0, 0, 0};
  {
    unsigned rem = cells[i];
    for (int d = 0; d < D; ++d) {
      cc[d] = (int)(rem % (unsigned)sd[d]);
      rem /= (unsigned)sd[d];
    }
  }
  int* out = nbIdx + i * n;
  int* outP = nbPj + i * n;
  // this node's count from an n image (include/simulation.hpp:337-343), at most n
  const int nI = nNode ? min(max(nNode[i], 0), n) : n;
  // the node's kernel's offset list (kernel-index image), indices global
  const int o0 = nodeKer ? offStart[nodeKer[i]] : 0;
  const int nO = nodeKer ? offCount[nodeKer[i]] : nOff;
  int cnt = 0;
  for (int base = 0; base < nO && cnt < nI; base += 64) {
    const int o = o0 + base + lane;
    bool hit = false;
    int pj = -1;
    if (base + lane < nO) {
      bool inside = true;
      long long t = 0, mul = 1;
      for (int d = 0; d < D; ++d) {
        int x = cc[d] + offs[(long long)o * D + d];
        // include/simulation.hpp:394: `|| circularSim` takes the wrapped cell
        if (circSim) x = ((x % sd[d]) + sd[d]) % sd[d];
        inside &= (x >= 0 && x < sd[d]);
        t += (long long)x * mul;
        mul *= sd[d];
      }
      if (inside) {
        pj = postS[t];
        hit = pj <= (int)i;
      }
    }
    const unsigned long long m = __ballot(hit);
    const int before = __popcll(m & ((1ull << lane) - 1ull));
    if (hit && cnt + before < nI) {
      out[cnt + before] = o;
      outP[cnt + before] = pj;
    }
    cnt += __popcll(m);
  }
  if (lane == 0) nbCnt[i] = min(cnt, nI);
}

int nb_search_device(int D, const unsigned* sd, size_t nPath, const unsigned* dCells, const int* dPostS,
                     const int* dOffs, int nOff, int n, int* dNbIdx, int* dNbPj, int* dNbCnt, hipStream_t st,
                     const int* dNNode, const int* dNodeKer, const int* dOffStart, const int* dOffCount,
                     int circSim) {
  if (nPath == 0) return QSMI_OK;
  hipLaunchKernelGGL(k_nbsearch, dim3((unsigned)nPath), dim3(64), 0, st, D, (int)sd[0], (int)sd[1], (int)sd[2], dCells,
                     dPostS, dOffs, nOff, n, dNbIdx, dNbPj, dNbCnt, dNNode, dNodeKer, dOffStart, dOffCount, circSim);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(QSMI_ERR_HIP, std::string("neighbour search: ") + hipGetErrorString(e));
  return QSMI_OK;
}

// ---------------------------------------------------------------------------
// Dependency levels of the path nodes: 1 + the deepest level among a node's
// path neighbours (hard data and the cell itself do not count), one
// wavefront per node. A node's neighbours precede it on the path, so node i
// waits only for levels of nodes < i; workgroups start in index order on
// every XCD, so the lowest unfinished node is always resident with its
// neighbours' levels written. Levels start at -1 (hipMemsetAsync before the
// launch) and are read and written by device-scope atomics (coherent across
// the XCDs' L2s), relaxed: nothing else is published with them. The wait is
// bounded: a neighbour level still unwritten after
// kLevelPolls polls sets *err and counts as 0 (the caller then recomputes the
// levels on the host), so every wavefront ends.
// ---------------------------------------------------------------------------
constexpr int kLevelPolls = 1 << 21;  // x s_sleep(4) (256 cycles): ~0.2 s, far above any real wait

__global__ void __launch_bounds__(64) k_levels(int n, const int* __restrict__ nbPj, const int* __restrict__ nbCnt,
                                               int* level, int* err) {
  const int i = blockIdx.x, lane = threadIdx.x;
  const int cnt = nbCnt[i];
  int lv = 0;
  for (int j = lane; j < cnt; j += 64) {
    const int pj = nbPj[(long long)i * n + j];
    if (pj >= 0 && pj < i) {
      int v = __hip_atomic_load(&level[pj], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      for (int polls = 0; v < 0; ++polls) {
        if (polls >= kLevelPolls) {
          __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          v = 0;
          break;
        }
        __builtin_amdgcn_s_sleep(4);
        v = __hip_atomic_load(&level[pj], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      lv = max(lv, v + 1);
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) lv = max(lv, __shfl_xor(lv, off));
  if (lane == 0) __hip_atomic_store(&level[i], lv, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

int levels_device(size_t nPath, int n, const int* dNbPj, const int* dNbCnt, int* dLevel, int* dErr, hipStream_t st) {
  if (nPath == 0) return QSMI_OK;
  if (nPath > 0x7fffffffull) return fail(QSMI_ERR_UNSUPPORTED, "path longer than 2^31 nodes");
  hipError_t e = hipMemsetAsync(dLevel, 0xFF, nPath * sizeof(int), st);
  if (e == hipSuccess) e = hipMemsetAsync(dErr, 0, sizeof(int), st);
  if (e != hipSuccess) return fail(QSMI_ERR_HIP, std::string("path levels: ") + hipGetErrorString(e));
  hipLaunchKernelGGL(k_levels, dim3((unsigned)nPath), dim3(64), 0, st, n, dNbPj, dNbCnt, dLevel, dErr);
  e = hipGetLastError();
  if (e != hipSuccess) return fail(QSMI_ERR_HIP, std::string("path levels: ") + hipGetErrorString(e));
  return QSMI_OK;
}

namespace {

// Swaps a device's stream, wave buffers and tap tables with its side lane's
// for the lifetime of the object (when on).
struct SideSwap {
  qsmi_device* d;
  bool on;
  SideSwap(qsmi_device* dev, bool enable) : d(dev), on(enable) {
    if (on) flip();
  }
  ~SideSwap() {
    if (on) flip();
  }
  void flip() {
    std::swap(d->stream, d->side.stream);
    std::swap(d->wb, d->side.wb);
    std::swap(d->dTap, d->side.dTap);
    std::swap(d->dTapD0, d->side.dTapD0);
    std::swap(d->dTapVal, d->side.dTapVal);
    std::swap(d->capTap, d->side.capTap);
    std::swap(d->capTapV, d->side.capTapV);
  }
};

struct LaneEvents {
  hipEvent_t e[2] = {nullptr, nullptr};
  ~LaneEvents() {
    for (hipEvent_t x : e)
      if (x) (void)hipEventDestroy(x);
  }
};

// Progress of the path loop (include/simulation.hpp:685-692 reports the
// share of simulated cells): every kStep waves an event on the device's
// stream; before one is reused the host waits for it, so the host runs at
// most kRing * kStep waves ahead of the GPU and a line on stderr, printed
// about every `seconds`, tells how far the GPU is.
struct Progress {
  static constexpr size_t kStep = 64, kRing = 8;
  int seconds = 0;
  hipEvent_t ev[kRing] = {};
  bool used[kRing] = {};
  std::chrono::steady_clock::time_point t0, last;
  explicit Progress(int s) : seconds(s) {
    if (seconds <= 0) return;
    for (auto& e : ev)
      if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) e = nullptr;
    t0 = last = std::chrono::steady_clock::now();
  }
  ~Progress() {
    for (hipEvent_t e : ev)
      if (e) (void)hipEventDestroy(e);
  }
  // before wave w is enqueued (waves w0 of the waves' first node)
  void tick(size_t w, size_t nWaves, long long nodesBefore, long long nodesAll, hipStream_t st) {
    if (seconds <= 0 || w == 0 || w % kStep) return;
    const size_t slot = (w / kStep) % kRing;
    if (!ev[slot]) return;
    if (used[slot]) (void)hipEventSynchronize(ev[slot]);  // the GPU has passed wave w - kRing * kStep
    (void)hipEventRecord(ev[slot], st);
    used[slot] = true;
    const auto now = std::chrono::steady_clock::now();
    if (std::chrono::duration<double>(now - last).count() < seconds) return;
    last = now;
    const size_t done = w > kRing * kStep ? w - kRing * kStep : 0;
    std::fprintf(stderr, "[qsmi] path loop: %zu of %zu waves done, %lld of %lld nodes enqueued, %.0f s\n", done,
                 nWaves, nodesBefore, nodesAll, std::chrono::duration<double>(now - t0).count());
    std::fflush(stderr);
  }
};

// Captured wave steps (tuning key loop.graph): the kernels of one plain
// 1024-grid wave are captured once per lane at the largest wave size and
// the graph is replayed for each of the lane's waves. Its first node,
// k_wave_next, copies the lane's next {w0, nN} from the lane's list to
// cur[0..1] and advances the cursor cur[2]; the gather reads {w0, nN} there
// and every later kernel reads nN (cur + 1) as its device node count, so a
// replay needs no parameter update.
__global__ void k_wave_next(const int2* __restrict__ desc, int* __restrict__ cur) {
  if (threadIdx.x == 0) {
    const int j = cur[2];
    const int2 d = desc[j];
    cur[0] = d.x;
    cur[1] = d.y;
    cur[2] = j + 1;
  }
}

struct WaveGraphs {
  hipGraph_t graph[2] = {nullptr, nullptr};
  hipGraphExec_t exec[2] = {nullptr, nullptr};
  int2* dDesc = nullptr;  // [2][nWaves]: each lane's waves {w0, nN} in launch order
  int* dCur = nullptr;    // [2][4]: {w0, nN, cursor, unused} of each lane
  ~WaveGraphs() {
    if (exec[0] || exec[1]) (void)hipDeviceSynchronize();  // replays still in flight on an error return
    for (int q = 0; q < 2; ++q) {
      if (exec[q]) (void)hipGraphExecDestroy(exec[q]);
      if (graph[q]) (void)hipGraphDestroy(graph[q]);
    }
    if (dDesc) (void)hipFree(dDesc);
    if (dCur) (void)hipFree(dCur);
  }
};

}  // namespace

int run_resident_waves(qsmi_device* dev, const ResidentLoop& R, double* stageMsOut, long long* scannedOut,
                       int* wavesOut, long long* bandStats) {
  qsmi_ti* t = dev->ti;
  hipStream_t st = dev->stream;
  const int D = t->D;
  if (R.nVar > 8 || R.nExp > 16) return fail(QSMI_ERR_UNSUPPORTED, "too many variables for the device gather");
  const size_t gatherLds = (size_t)R.nVar * R.n * (sizeof(float) + 1);
  if (gatherLds > 60 * 1024) return fail(QSMI_ERR_UNSUPPORTED, "n x nVar too large for the device gather (<= 12288)");
  GatherParams P;
  memset(&P, 0, sizeof(P));
  P.D = D;
  for (int d = 0; d < 3; ++d) {
    P.sd[d] = d < D ? R.sd[d] : 1;
    P.kd[d] = d < D ? R.kd[d] : 1;
    P.kctr[d] = P.kd[d] / 2;
    P.kdK[0][d] = P.kd[d];
    P.s[d] = t->s[d];
  }
  P.nVar = R.nVar;
  P.nExp = R.nExp;
  P.nSub = t->nSub;
  P.V = (int)dev->active.size();
  P.n = R.n;
  for (int v = 0; v < R.nVar; ++v) P.types[v] = R.types[v];
  for (int e = 0; e < R.nExp; ++e) {
    P.expVar[e] = R.expVar[e];
    P.expCat[e] = R.expCat[e];
  }
  for (int B = 0; B < t->nSub; ++B) {
    P.conv[B] = R.conv[B];
    P.subOf[B] = R.subOf[B];
  }
  for (int a = 0; a < P.V; ++a) P.act[a] = dev->active[a];
  for (int q = 0; q < t->nSub * t->nSub; ++q) P.alpha[q] = dev->alpha[q];
  P.cross = dev->cross ? 1 : 0;
  P.noVerbatim = R.noVerbatim ? 1 : 0;
  P.circSim = R.circSim ? 1 : 0;
  for (int d = 0; d < 3; ++d) P.f[d] = d < D ? (unsigned)t->f[d] : 1u;
  if (R.noVerbatim && R.lp) return fail(QSMI_ERR_UNSUPPORTED, "learned proposal with the verbatim exclusion");
  if (dev->cross) {
    if (R.lp) return fail(QSMI_ERR_UNSUPPORTED, "learned proposal: TI with NaN unsupported");
    for (int q = 0; q < t->nSub * t->nSub; ++q) P.alphaX[q] = dev->alphaX[q];
  }
  P.circular = dev->circular;
  P.trueMismatch = dev->trueMismatch;
  P.scale = map_scale(dev);
  for (int a = 0; a < P.V; ++a) P.rms[a] = t->rms[dev->active[a]];
  P.epsScale = (dev->quantum == 0.f && !dev->cross)
                   ? (float)((double)kMapErrC * (double)P.scale * scan_cells(dev))
                   : 0.f;
  P.quantum = dev->quantum;
  P.tiCells = (unsigned)t->S;
  const int nTI = 1 + (int)R.extraDevs.size();
  P.nTI = nTI;
  if (nTI > 8) return fail(QSMI_ERR_UNSUPPORTED, "at most 8 training images");
  for (int q = 0; q <= 8; ++q) P.tiCum[q] = q == 0 ? 0u : (q <= nTI ? R.tiCum[q] : R.tiCum[nTI]);
  if (nTI == 1) P.tiCum[1] = (unsigned)t->S;
  TiCum cum;
  for (int q = 0; q < 9; ++q) cum.c[q] = P.tiCum[q];
  if (R.nKer > 8) return fail(QSMI_ERR_UNSUPPORTED, "at most 8 kernels");
  for (int q = 1; q < R.nKer; ++q) {
    for (int d = 0; d < 3; ++d) P.kdK[q][d] = d < D ? R.kdK[q][d] : 1;
    P.kerOff[q] = R.kerOff[q];
  }
  if (R.dNodeKer && (R.lp || R.recCap > 0))
    return fail(QSMI_ERR_UNSUPPORTED, "kernel-index image: not with the learned proposal or the record");
  if (R.dPathK && (R.lp || R.recCap > 0))
    return fail(QSMI_ERR_UNSUPPORTED, "k image: not with the learned proposal or the record");
  if (nTI > 1 && (R.lp || dev->cross || R.noVerbatim || R.recCap > 0))
    return fail(QSMI_ERR_UNSUPPORTED, "several TIs: not with the learned proposal, NaN TIs, -nV or the record");

  const int K = (int)std::ceil(std::max(R.k, R.kMax));
  int maxWave = 0;
  for (size_t w = 0; w + 1 < R.waveStart.size(); ++w) maxWave = std::max(maxWave, R.waveStart[w + 1] - R.waveStart[w]);
  if (!dev->wb) dev->wb = new WaveBufs();
  WaveBufs* b = dev->wb;
  // the cross-measurement form scans two layers per node (nodes q, nN + q)
  const int layers = dev->cross ? 2 : 1;
  int rc = grow_wave_bufs(dev, b, layers * std::max(maxWave, 1), layers * std::max(maxWave, 1) * R.n, K);
  if (rc) return rc;
  if (use_w1k(dev) && (rc = ensure_tap_tables(dev, std::max(maxWave, 1), geom_of(dev).V))) return rc;
  if (R.lp && (rc = lp_prepare(R.lp, t, std::max(maxWave, 1), K, R.lpShortlist, dev->active.data(),
                               (int)dev->active.size(), dev->circular)))
    return rc;
  // several TIs: every extra device works on this device's stream and
  // wave inputs; per-TI K best lists [t][node][K] for the cross-TI merge
  std::vector<hipStream_t> keepStreams;
  float* dValAll = nullptr;
  uint32_t* dPosAll = nullptr;
  const long long tiStride = (long long)std::max(maxWave, 1) * K;
  if (nTI > 1) {
    for (qsmi_device* x : R.extraDevs) {
      if (!x->wb) x->wb = new WaveBufs();
      if ((rc = grow_wave_bufs(x, x->wb, std::max(maxWave, 1), std::max(maxWave, 1) * R.n, K))) return rc;
      if (use_w1k(x) && (rc = ensure_tap_tables(x, std::max(maxWave, 1), geom_of(x).V))) return rc;
    }
    if (hipMalloc(&dValAll, (size_t)nTI * tiStride * sizeof(float)) != hipSuccess ||
        hipMalloc(&dPosAll, (size_t)nTI * tiStride * sizeof(uint32_t)) != hipSuccess) {
      (void)hipFree(dValAll);
      return fail(QSMI_ERR_NOMEM, "several TIs: candidate lists");
    }
    for (qsmi_device* x : R.extraDevs) {
      keepStreams.push_back(x->stream);
      x->stream = st;
    }
  }
  auto restore = [&]() {
    for (size_t i = 0; i < keepStreams.size(); ++i) R.extraDevs[i]->stream = keepStreams[i];
    keepStreams.clear();
    if (dValAll) (void)hipFree(dValAll);
    if (dPosAll) (void)hipFree(dPosAll);
    dValAll = nullptr;
    dPosAll = nullptr;
  };
  ScanGeom g = geom_of(dev);
  // rows stage: the cross form puts one row (both layers) in each FFT
  const int rpb = dev->cross ? (1 << t->lognv[0]) : rows_per_block(dev);
  const int rb = dev->cross ? (int)(((long long)t->s[1] * t->s[2] + rpb - 1) / rpb) : row_blocks(dev);
  const bool savedTiming = dev->timing;
  dev->timing = false;  // the per-call events of launch_spectrum are not used here
  const size_t nWaves = R.waveStart.size() > 0 ? R.waveStart.size() - 1 : 0;
  const bool timing = R.stageTiming > 0;
  const size_t every = (size_t)std::max(1, R.stageTiming);  // timed waves: w % every == 0
  if (timing) {
    const size_t need = ((nWaves + every - 1) / every) * 5;
    while (dev->evPool.size() < need) {
      hipEvent_t ev;
      hipError_t e = hipEventCreate(&ev);
      if (e != hipSuccess) {
        dev->timing = savedTiming;
        return fail(QSMI_ERR_HIP, std::string("event pool: ") + hipGetErrorString(e));
      }
      dev->evPool.push_back(ev);
    }
  }
  long long scanned = 0;
  int waves = 0;
  // Two streams: no node of a level has a neighbour in the same level, so
  // consecutive waves of one level are independent; they alternate between
  // the device's stream (lane 0) and its side stream (lane 1), each with its
  // own wave buffers. At a level change each lane waits for the other's
  // earlier work.
  const bool twoLanes = !R.lp && nTI == 1 && R.recCap <= 0 && R.waveLevel.size() == nWaves && nWaves > 1;
  LaneEvents lev;
  if (twoLanes) {
    if (!dev->side.stream && hipStreamCreateWithFlags(&dev->side.stream, hipStreamNonBlocking) != hipSuccess) {
      dev->timing = savedTiming;
      return fail(QSMI_ERR_HIP, "resident waves: side stream");
    }
    if (!dev->side.wb) dev->side.wb = new WaveBufs();
    {
      SideSwap sw(dev, true);
      rc = grow_wave_bufs(dev, dev->wb, layers * std::max(maxWave, 1), layers * std::max(maxWave, 1) * R.n, K);
      if (!rc && use_w1k(dev)) rc = ensure_tap_tables(dev, std::max(maxWave, 1), geom_of(dev).V);
    }
    for (int q = 0; q < 2 && !rc; ++q)
      if (hipEventCreateWithFlags(&lev.e[q], hipEventDisableTiming) != hipSuccess)
        rc = fail(QSMI_ERR_HIP, "resident waves: lane events");
    if (rc) {
      dev->timing = savedTiming;
      return rc;
    }
  }
  // captured wave steps: the plain 1024-grid loop on two lanes (no sampled
  // stage events then)
  const bool graphOn = t->tune.loopGraph && twoLanes && use_w1k(dev) && !dev->cross;
  WaveGraphs G;
  if (graphOn) {
    std::vector<int2> desc[2];
    int prev = -1, ln = 0;  // the lanes the loop below gives the waves
    for (size_t w = 0; w < nWaves; ++w) {
      const int w0 = R.waveStart[w], nN = R.waveStart[w + 1] - w0;
      if (nN <= 0) continue;
      if (R.waveLevel[w] != prev) {
        prev = R.waveLevel[w];
        ln = 0;
      } else {
        ln ^= 1;
      }
      desc[ln].push_back(make_int2(w0, nN));
    }
    const int cur0[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (hipMalloc(&G.dDesc, 2 * nWaves * sizeof(int2)) != hipSuccess || hipMalloc(&G.dCur, sizeof(cur0)) != hipSuccess)
      rc = fail(QSMI_ERR_NOMEM, "wave graphs: out of device memory");
    for (int q = 0; q < 2 && !rc; ++q)
      if (!desc[q].empty() && hipMemcpy(G.dDesc + (size_t)q * nWaves, desc[q].data(), desc[q].size() * sizeof(int2),
                                        hipMemcpyHostToDevice) != hipSuccess)
        rc = fail(QSMI_ERR_HIP, "wave graphs: wave lists");
    if (!rc && hipMemcpy(G.dCur, cur0, sizeof(cur0), hipMemcpyHostToDevice) != hipSuccess)
      rc = fail(QSMI_ERR_HIP, "wave graphs: cursors");
    // every buffer at the largest wave: the captured calls allocate nothing
    for (int q = 0; q < 2 && !rc; ++q) {
      SideSwap sw(dev, q == 1);
      rc = band_reserve(dev, &dev->wb->band, maxWave, K);
    }
    if (rc) {
      dev->timing = savedTiming;
      return rc;
    }
  }
  // captures the wave step of the current lane (the device's stream and
  // buffers, swapped in by SideSwap)
  auto captureLane = [&](int ln) -> int {
    hipStream_t cs = dev->stream;
    WaveBufs* cb = dev->wb;
    int* cur = G.dCur + 4 * ln;
    hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return fail(QSMI_ERR_HIP, std::string("wave graph: capture: ") + hipGetErrorString(e));
    hipLaunchKernelGGL(k_wave_next, dim3(1), dim3(64), 0, cs, G.dDesc + (size_t)ln * nWaves, cur);
    hipLaunchKernelGGL(k_gather, dim3(maxWave), dim3(64), gatherLds, cs, P, 0, maxWave, R.dOrder, R.dPathCells,
                       R.dPathSeeds, R.dNbIdx, R.dNbCnt, R.dOffs, R.dKernel, R.dCatvals, R.dDi, R.dIndex, R.dTi,
                       cb->dInfo, cb->dOff, cb->dVal, dev->dCentreOk, dev->dComplete, dev->nComplete, R.dPathK,
                       R.dNodeKer, R.dNodeTi, dev->dTap, dev->dTapD0, dev->dTapVal, cur);
    int lrc = launch_spectrum(dev, g, maxWave, cb->dInfo, cb->dOff, cb->dVal, cb->dOrd, cb->dC, cur + 1, true);
    if (!lrc) {
      launch_rows(dev, g, cb->dInfo, cb->dC, maxWave, rb, K, cb->dCand, cb->dGcnt, nullptr, cur + 1, 0);
      lrc = band_select(dev, g, cb->dInfo, cb->dCand, cb->dGcnt, maxWave, K, R.k, cb->dC, cb->dOff, cb->dVal, R.n,
                        cb->dOutVal, cb->dOutPos, nullptr, nullptr, R.dDi, R.dTi, R.nVar, R.dIndex, &cb->band,
                        cur + 1);
    }
    hipGraph_t gr = nullptr;
    e = hipStreamEndCapture(cs, &gr);
    if (lrc || e != hipSuccess || !gr) {
      if (gr) (void)hipGraphDestroy(gr);
      return lrc ? lrc : fail(QSMI_ERR_HIP, std::string("wave graph: end capture: ") + hipGetErrorString(e));
    }
    G.graph[ln] = gr;
    e = hipGraphInstantiate(&G.exec[ln], gr, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      G.exec[ln] = nullptr;
      return fail(QSMI_ERR_HIP, std::string("wave graph: instantiate: ") + hipGetErrorString(e));
    }
    return QSMI_OK;
  };
  int prevLevel = -1, lane = 0;
  bool used1 = false, need1 = false;
  Progress progress(R.progressSec);
  const long long nodesAll = nWaves ? (long long)R.waveStart[nWaves] : 0;
  for (size_t w = 0; w < nWaves; ++w) {
    const int w0 = R.waveStart[w], nN = R.waveStart[w + 1] - w0;
    if (nN <= 0) continue;
    progress.tick(w, nWaves, w0, nodesAll, dev->stream);
    if (twoLanes) {
      if (R.waveLevel[w] != prevLevel) {  // a new level starts on lane 0
        if (prevLevel >= 0) {
          (void)hipEventRecord(lev.e[0], dev->stream);
          if (used1) {
            (void)hipEventRecord(lev.e[1], dev->side.stream);
            (void)hipStreamWaitEvent(dev->stream, lev.e[1], 0);
          }
          need1 = true;  // lane 1 waits for lane 0's earlier levels before its next wave
        }
        prevLevel = R.waveLevel[w];
        lane = 0;
        used1 = false;
      } else {
        lane ^= 1;
      }
      if (lane == 1) {
        if (need1) (void)hipStreamWaitEvent(dev->side.stream, lev.e[0], 0);
        need1 = false;
        used1 = true;
      }
    }
    SideSwap sideSwap(dev, twoLanes && lane == 1);
    hipStream_t st = dev->stream;  // this wave's lane
    WaveBufs* b = dev->wb;
    if (graphOn) {
      if (!G.exec[lane] && (rc = captureLane(lane))) {
        dev->timing = savedTiming;
        return rc;
      }
      const hipError_t e = hipGraphLaunch(G.exec[lane], st);
      if (e != hipSuccess) {
        dev->timing = savedTiming;
        return fail(QSMI_ERR_HIP, std::string("wave graph: launch: ") + hipGetErrorString(e));
      }
      scanned += nN;
      waves++;
      continue;
    }
    hipEvent_t* ev = (timing && w % every == 0) ? &dev->evPool[(w / every) * 5] : nullptr;
    if (ev) (void)hipEventRecord(ev[0], st);
    // plain 1024-grid waves: the gather also builds the tap lists
    const bool fuseTaps = use_w1k(dev) && !R.lp && nTI == 1 && !dev->cross;
    hipLaunchKernelGGL(k_gather, dim3(nN), dim3(64), gatherLds, st, P, w0, nN, R.dOrder, R.dPathCells, R.dPathSeeds,
                       R.dNbIdx, R.dNbCnt, R.dOffs, R.dKernel, R.dCatvals, R.dDi, R.dIndex, R.dTi, b->dInfo, b->dOff,
                       b->dVal, dev->dCentreOk, dev->dComplete, dev->nComplete, R.dPathK, R.dNodeKer, R.dNodeTi,
                       fuseTaps ? dev->dTap : nullptr, fuseTaps ? dev->dTapD0 : nullptr,
                       fuseTaps ? dev->dTapVal : nullptr, nullptr);
    if (ev) (void)hipEventRecord(ev[1], st);
    const bool useLp = R.lp && (R.waveLp.empty() || R.waveLp[w]);
    if (R.lp && !useLp) R.lpSkipped += nN;
    if (useLp) {
      // learned proposal for every node, the exact scan for the routed ones
      lp_propose(R.lp, t, P.V, nN, b->dInfo, b->dOff, b->dVal, R.dDi, R.nVar, R.sd[0], K, R.lpMinCover, st);
      lp_flags(R.lp, nN, w0, R.dOrder, b->dInfo, K, R.lpEvery, R.lpMse, R.lpMinCover, R.lpAudit, st);
      // the nodes to scan: compacted to the front on the 1024-grid path (the
      // kernels read the count on the device), else the whole wave with the
      // other nodes' boxes emptied
      const bool compact = use_w1k(dev);
      const NodeInfo* infoEx = compact ? lp_info_compact(R.lp) : lp_info_ex(R.lp);
      const int* exCount = compact ? lp_compact_count(R.lp) : nullptr;
      if ((rc = launch_spectrum(dev, g, nN, infoEx, b->dOff, b->dVal, b->dOrd, b->dC, exCount))) return rc;
      if (ev) (void)hipEventRecord(ev[2], st);
      launch_rows(dev, g, infoEx, b->dC, nN, rb, K, b->dCand, b->dGcnt, nullptr, exCount);
      // the scanned nodes' K best by the exact selection, as candidate lists
      if ((rc = band_select(dev, g, infoEx, b->dCand, b->dGcnt, nN, K, R.k, b->dC, b->dOff, b->dVal, R.n, b->dOutVal,
                            b->dOutPos, b->dCandEx, nullptr, nullptr, nullptr, 0, nullptr, &b->band, exCount)))
        return rc;
      lp_combine(R.lp, nN, compact ? b->dInfo : infoEx, b->dCandEx, compact, 1, 0, K, R.lpAudit, st);
      if (ev) (void)hipEventRecord(ev[3], st);
      launch_merge_draw(g, b->dInfo, lp_cand(R.lp), nN, 1, K, 0, K, R.k, b->dOutVal, b->dOutPos, nullptr, R.dDi,
                        R.dTi, R.nVar, R.dIndex, st);
    } else if (nTI > 1) {
      // every TI's scan and K best, then the merge across TIs and the draw
      for (int tt = 0; tt < nTI; ++tt) {
        qsmi_device* dv = tt == 0 ? dev : R.extraDevs[tt - 1];
        WaveBufs* xb = tt == 0 ? b : dv->wb;
        const ScanGeom gt = geom_of(dv);
        // the node boxes on this TI's extent
        const NodeInfo* it = b->dInfo;
        if (tt > 0) {
          // this TI's boxes, map scale and error bound: the bound scales
          // with the TI's sub-image norms (largest ratio over the active
          // sub-variables) and with scale * N
          const float sc = map_scale(dv);
          double ratio = 0.0;
          for (int A : dev->active)
            ratio = std::max(ratio, t->rms[A] > 0.f ? (double)dv->ti->rms[A] / (double)t->rms[A] : 1e30);
          if (!dv->trueMismatch) ratio *= (double)dv->ti->N / (double)t->N;
          launch_info_shift(b->dInfo, xb->dInfo, nN, dv->ti->s[0] - t->s[0], dv->ti->s[1] - t->s[1],
                            dv->ti->s[2] - t->s[2], sc, (float)std::min(ratio, 1e30), st);
          it = xb->dInfo;
        }
        const int rbt = tt == 0 ? rb : row_blocks(dv);
        if ((rc = launch_spectrum(dv, gt, nN, it, b->dOff, b->dVal, xb->dOrd, xb->dC))) {
          restore();
          return rc;
        }
        launch_rows(dv, gt, it, xb->dC, nN, rbt, K, xb->dCand, xb->dGcnt, nullptr, nullptr, 0);
        if ((rc = band_select(dv, gt, it, xb->dCand, xb->dGcnt, nN, K, R.k, xb->dC, b->dOff, b->dVal, R.n,
                              dValAll + tt * tiStride, dPosAll + tt * tiStride, nullptr, nullptr, nullptr, nullptr,
                              R.nVar, nullptr, &xb->band))) {
          restore();
          return rc;
        }
      }
      if (ev) (void)hipEventRecord(ev[2], st);
      if (ev) (void)hipEventRecord(ev[3], st);
      launch_merge_ti(b->dInfo, nN, nTI, K, R.k, dValAll, dPosAll, tiStride, dev->quantum, R.dDi, R.dTi, R.nVar,
                      cum, R.dIndex, st);
    } else {
      if ((rc = launch_spectrum(dev, g, layers * nN, b->dInfo, b->dOff, b->dVal, b->dOrd, b->dC, nullptr, fuseTaps)))
        return rc;
      if (ev) (void)hipEventRecord(ev[2], st);
      launch_rows(dev, g, b->dInfo, b->dC, nN, rb, K, b->dCand, b->dGcnt, nullptr, nullptr, dev->cross ? nN : 0);
      if (ev) (void)hipEventRecord(ev[3], st);
      if ((rc = band_select(dev, g, b->dInfo, b->dCand, b->dGcnt, nN, K, R.k, b->dC, b->dOff, b->dVal, R.n, b->dOutVal,
                            b->dOutPos, nullptr, nullptr, R.dDi, R.dTi, R.nVar, R.dIndex, &b->band)))
        return rc;
    }
    if (R.recCap > 0)
      lp_record(nN, w0, R.dOrder, b->dInfo, b->dOff, R.dDi, R.nVar, R.sd[0], b->dOutPos, K, R.recCap, R.dRecOff,
                R.dRecVal, R.dRecBest, st);
    if (ev) (void)hipEventRecord(ev[4], st);
    scanned += nN;
    waves++;
  }
  dev->timing = savedTiming;
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    (void)hipStreamSynchronize(st);
    if (twoLanes) (void)hipStreamSynchronize(dev->side.stream);
    restore();
    return fail(QSMI_ERR_HIP, std::string("resident waves: ") + hipGetErrorString(e));
  }
  e = hipStreamSynchronize(st);
  if (e == hipSuccess && twoLanes) e = hipStreamSynchronize(dev->side.stream);
  if (e == hipSuccess && bandStats) {
    for (int q = 0; q < 4; ++q) bandStats[q] = 0;
    long long one[4];
    if ((rc = band_read_stats(&b->band, one))) {
      restore();
      return rc;
    }
    for (int q = 0; q < 4; ++q) bandStats[q] += one[q];
    if (twoLanes) {
      if ((rc = band_read_stats(&dev->side.wb->band, one))) {
        restore();
        return rc;
      }
      for (int q = 0; q < 4; ++q) bandStats[q] += one[q];
    }
    for (qsmi_device* x : R.extraDevs) {
      if ((rc = band_read_stats(&x->wb->band, one))) {
        restore();
        return rc;
      }
      for (int q = 0; q < 4; ++q) bandStats[q] += one[q];
    }
  }
  restore();
  if (e != hipSuccess) return fail(QSMI_ERR_HIP, std::string("resident waves: ") + hipGetErrorString(e));
  double stage[4] = {0, 0, 0, 0};
  if (timing) {
    size_t timed = 0, all = 0;
    for (size_t w = 0; w < nWaves; ++w) {
      if (R.waveStart[w + 1] - R.waveStart[w] <= 0) continue;
      ++all;
      if (w % every) continue;
      ++timed;
      hipEvent_t* ev = &dev->evPool[(w / every) * 5];
      float ms[4] = {0, 0, 0, 0};
      for (int q = 0; q < 4; ++q) (void)hipEventElapsedTime(&ms[q], ev[q], ev[q + 1]);
      stage[1] += ms[0];
      stage[0] += ms[1];
      stage[2] += ms[2];
      stage[3] += ms[3];
    }
    // the timed waves stand for all of them
    if (timed) for (int q = 0; q < 4; ++q) stage[q] *= (double)all / (double)timed;
  }
  for (int q = 0; q < 4; ++q) stageMsOut[q] = stage[q];
  *scannedOut = scanned;
  *wavesOut = waves;
  return QSMI_OK;
}

}  // namespace qsmi
