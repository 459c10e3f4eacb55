// Shared device-side types of the MI355X QuickSampling scan.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "qsmi_fft.hpp"

namespace qsmi {

constexpr int kMaxSub = 16;     // TI sub-variable spectra per device
constexpr int kMaxK = 16;       // ceil(k) supported by the select
constexpr int kMaxVecElems = 4096;  // L * nvec of one LDS column block

// Geometry of one scan (TI grid s, FFT grid f, half spectrum along dim0).
struct ScanGeom {
  int D;
  int s[3];
  int f[3];
  int h0;       // f0/2 + 1
  int P;        // pencil dimension = D-1
  int V;        // active TI sub-variable spectra
  int A[kMaxSub];  // their indices in the TI spectra array
  int circular;
  float invN;   // 1 / prod(f)
  long long H;  // half-spectrum size h0*f1*f2
};

// Distinct neighbour rows ("taps") of one node on the 1024-row scan grid
// (row = (-d1) mod 1024), ordered by (row mod 16, row), and the node's
// neighbours in tap order, for the sparse column transform. T = 0 when the
// node has more than kTapNbMax neighbours (no lists); up to kTapMax taps the
// column transform is the tap sum, beyond it a full FFT of the column the
// taps fill.
constexpr int kTapMax = 48;
constexpr int kTapPass = 3;      // most taps of one class on the tap path
constexpr int kTapNbMax = 500;
struct TapInfo1024 {
  int T;
  int passes;                      // tap path: the most taps of one class (<= kTapPass); 0: full transform
  int cls[17];                     // taps of class c: [cls[c], cls[c+1])
  uint16_t row[kTapNbMax];
  uint16_t start[kTapNbMax + 1];   // tap t's neighbours: slots [start[t], start[t+1]) in tap order
  uint8_t slot[kTapMax];           // tap path: tap t is the k-th of class c: slot 16 k + c
};

// Per path node of a wave. Valid TI centres: lo[i] <= c_i <= hi[i].
struct NodeInfo {
  int lo[3];
  int hi[3];
  int nbStart;
  int nb;
  float delta0;   // added to the map (true mismatch) — see scale
  float scale;    // map = raw * scale + delta0
  float seed;
  uint32_t seedbits;  // seedFromUnitFloat(seed)
  float quantum;      // tie quantum (0: none)
  uint32_t cell;      // simulation cell (simulation driver only)
  float cmin;         // cross-measurement form: a count layer value (a sum of
                      // neighbour weights) below it is zero: half the
                      // smallest neighbour weight (0 otherwise)
  int exOn;           // 1: no candidate within the disk below (verbatim exclusion)
  uint32_t exPos;     // disk centre, TI position
  float exR2;         // squared radius: positions p with sum_d (p_d - o_d)^2 <= exR2
  float kLocal;       // this node's k from a k image (0: the run's k)
  int tiReq;          // several TIs: the only TI this node samples (-1: all; a TI-index image)
  float eps;          // bound of the FP32 map error of this node, in map units; > 0 switches the
                      // exact selection on (every position within 2 eps of the K-th best FP32
                      // value is rescored in FP64 before the device rule picks the K best)
};

// Several TIs: the first cell of each TI in their concatenation, then the total.
struct TiCum {
  unsigned c[9];
};

// Verbatim exclusion test of TI position (c0, c1, c2) for a node.
__device__ __forceinline__ bool excluded(const NodeInfo& ni, int s0, int s1, int c0, int c1, int c2) {
  const int o0 = (int)(ni.exPos % (uint32_t)s0);
  const int o1 = (int)((ni.exPos / (uint32_t)s0) % (uint32_t)s1);
  const int o2 = (int)(ni.exPos / ((uint32_t)s0 * (uint32_t)s1));
  const float d0 = (float)(c0 - o0), d1 = (float)(c1 - o1), d2 = (float)(c2 - o2);
  return d0 * d0 + d1 * d1 + d2 * d2 <= ni.exR2;
}

// Candidate of the device selection rule.
struct Cand {
  float key;    // value (or rint(value/quantum))
  uint32_t h;   // tie hash
  uint32_t pos; // TI position (flattened, dim0 fastest)
  float val;    // map value
};

__device__ __forceinline__ bool cand_better(const Cand& a, const Cand& b) {
  if (a.key != b.key) return a.key > b.key;
  if (a.h != b.h) return a.h > b.h;
  return a.pos < b.pos;
}

__host__ __device__ __forceinline__ uint32_t fmix32(uint32_t h) {
  h ^= h >> 16;
  h *= 0x85ebca6bu;
  h ^= h >> 13;
  h *= 0xc2b2ae35u;
  h ^= h >> 16;
  return h;
}

__host__ __device__ __forceinline__ uint32_t tie_hash(uint32_t seedbits, uint32_t pos) {
  return fmix32(pos * 0x9E3779B1u + seedbits);
}

// Candidate of the exact selection: FP64 value, the same device rule
// (value desc, hash desc, position asc).
struct Cand64 {
  double val;
  uint32_t h;
  uint32_t pos;
};

__device__ __forceinline__ bool cand64_better(const Cand64& a, const Cand64& b) {
  if (a.val != b.val) return a.val > b.val;
  if (a.h != b.h) return a.h > b.h;
  return a.pos < b.pos;
}

// FP32 map error bound per node: eps = kMapErrC * (scale * N) *
// sum_a rms(Y_a) * sum_j |v_ja|, with rms(Y_a) the root mean square of TI
// sub-image a over the FFT grid (zero padding included). The FFT rounding
// error of a correlation is proportional to the norms of its two operands;
// the constant holds the measured worst case with a margin.
constexpr float kMapErrC = 4e-5f;

// Exact selection work space per wave node: positions to rescore.
constexpr int kBandCap = 8192;
// rows blocks per node the exact selection tracks (larger grids keep the
// FP32 selection)
constexpr int kMaxBandBlocks = 8192;

}  // namespace qsmi
