// QuickSampling CPU oracle — TEST INFRASTRUCTURE ONLY.
//
// A plain CPU restatement of the reference QuickSampling scan and its callers,
// used by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg
// as the checker. Nothing in the product (g2s_amd/) links, loads or calls it.
//
// Every function cites the reference file:line it restates. How each one is
// pinned (golden vectors from oracle/_ref, closed forms, invariants, numpy):
//   size rule ............ tests/test_oracle_ref.py (vs compiled reference)
//   TI sub-images, alpha . tests/test_oracle_ref.py (vs compiled reference)
//   kernel gen/expand .... tests/test_oracle_ref.py (vs compiled reference)
//   fKst select .......... tests/test_oracle_ref.py (vs compiled reference,
//                          tie-heavy maps, leading -inf runs)
//   offset list order .... tests/test_oracle_ref.py (golden and live: the
//                          comparator's weights from the compiled DataImage
//                          distance/indexWithDelta, the generation loop and
//                          std::sort restated)
//   FFT .................. tests/test_oracle_math.py (vs numpy.fft)
//   maps ................. tests/test_oracle_math.py (direct FP64 vs FFT FP64
//                          vs numpy brute force; self-retrieval; valid count;
//                          linearity in alpha; translation)
//   index maps / draw .... tests/test_oracle_math.py (round trip, closed form)
//   path / seeds ......... tests/test_oracle_sim.py (libstdc++ routines;
//                          permutation and range checks)
//   neighbour gather ..... tests/test_oracle_sim.py (brute-force nearest
//                          informed search)
#include <algorithm>
#include <climits>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <functional>
#include <limits>
#include <numeric>
#include <random>
#include <vector>

#include "fft_cpu.hpp"

#ifdef _OPENMP
#include <omp.h>
#endif

using namespace orc;

extern "C" {

// ---------------------------------------------------------------------------
// a1. FFT size rule — src/sharedMemoryManager.cpp:26-58.
// Smallest size >= src, up to the next power of two, of the form
// 2^a 3^b 5^c 7^d times at most one factor 11 or 13.
// ---------------------------------------------------------------------------
unsigned orc_fft_size_1d(unsigned src) {
  if (src <= 1) return src;
  unsigned limit = 1;
  while (limit < src) limit <<= 1;
  for (unsigned cand = src; cand <= limit; ++cand) {
    unsigned v = cand;
    for (unsigned p = 2; p < 8; ++p)
      while (v % p == 0) v /= p;
    if (v % 11 == 0)
      v /= 11;
    else if (v % 13 == 0)
      v /= 13;
    if (v == 1) return cand;
  }
  return src;
}

void orc_fft_size(int D, const unsigned* src, unsigned* out) {
  for (int i = 0; i < D; ++i) out[i] = orc_fft_size_1d(src[i]);
}

// ---------------------------------------------------------------------------
// Category discovery — src/qs.cpp:1500-1522: values of a categorical
// variable in order of first appearance over the TI (NaN skipped).
// Returns the number of categories written (<= maxOut).
// ---------------------------------------------------------------------------
int orc_categories(size_t cells, int nVar, const float* data, int var, float* out, int maxOut) {
  int n = 0;
  for (size_t c = 0; c < cells; ++c) {
    float v = data[c * nVar + var];
    if (std::isnan(v)) continue;
    bool present = false;
    for (int k = 0; k < n; ++k) present |= (out[k] == v);
    if (!present) {
      if (n >= maxOut) return -1;
      out[n++] = v;
    }
  }
  return n;
}

// ---------------------------------------------------------------------------
// a2. TI sub-variable images — include/DataImage.hpp:538-599.
// Continuous variable -> [y^2, y]; categorical -> [1(y==c_k)] for each class;
// NaN -> 0. The reference stores them point-reflected on the FFT grid
// (index N-1-idx); `reflected`=1 reproduces that layout on the f-grid,
// `reflected`=0 writes them unreflected on the f-grid (zero padding beyond
// the TI extent). types: 0 continuous, 1 categorical.
// ncat[v] and catvals (concatenated per categorical variable) describe the
// classes. Output: nSub images of f0*f1*f2 floats.
// ---------------------------------------------------------------------------
int orc_ti_subimages(int D, const unsigned* s, const unsigned* f, int nVar, const int* types,
                     const float* data, const int* ncat, const float* catvals, int reflected,
                     float* out) {
  size_t F = 1, S = 1;
  for (int i = 0; i < D; ++i) {
    F *= f[i];
    S *= s[i];
  }
  int sub = 0;
  int catIdx = 0;
  const float* cv = catvals;
  for (int v = 0; v < nVar; ++v) {
    int nsv = (types[v] == 0) ? 2 : ncat[catIdx];
    for (int k = 0; k < nsv; ++k) std::fill(out + (size_t)(sub + k) * F, out + (size_t)(sub + k + 1) * F, 0.f);
    for (size_t c = 0; c < S; ++c) {
      // TI coordinate -> f-grid flat index
      size_t rem = c, fidx = 0, mul = 1;
      for (int i = 0; i < D; ++i) {
        size_t ci = rem % s[i];
        rem /= s[i];
        fidx += ci * mul;
        mul *= f[i];
      }
      size_t dst = reflected ? (F - 1 - fidx) : fidx;
      float y = data[c * nVar + v];
      if (types[v] == 0) {
        out[(size_t)sub * F + dst] = std::isnan(y) ? 0.f : y * y;
        out[(size_t)(sub + 1) * F + dst] = std::isnan(y) ? 0.f : y;
      } else {
        for (int k = 0; k < nsv; ++k) out[(size_t)(sub + k) * F + dst] = (y == cv[k]) ? 1.f : 0.f;
      }
    }
    if (types[v] != 0) {
      cv += ncat[catIdx];
      catIdx++;
    }
    sub += nsv;
  }
  return sub;
}

// ---------------------------------------------------------------------------
// Cross-measurement sub-images (TIs with NaN; continuous variables only):
// include/DataImage.hpp:538-599 with needCrossMeasurement — per variable
// [y^2, y, m] with NaN -> 0 in the first two and m = 1 where y is a number,
// 0 where it is NaN; same f-grid / reflection layout as orc_ti_subimages.
// Returns nSub (3 per variable), or -1 for a categorical variable.
// ---------------------------------------------------------------------------
int orc_ti_subimages_x(int D, const unsigned* s, const unsigned* f, int nVar, const int* types,
                       const float* data, int reflected, float* out) {
  size_t F = 1, S = 1;
  for (int i = 0; i < D; ++i) {
    F *= f[i];
    S *= s[i];
  }
  for (int v = 0; v < nVar; ++v)
    if (types[v] != 0) return -1;
  std::fill(out, out + (size_t)3 * nVar * F, 0.f);
  for (int v = 0; v < nVar; ++v) {
    for (size_t c = 0; c < S; ++c) {
      size_t rem = c, fidx = 0, mul = 1;
      for (int i = 0; i < D; ++i) {
        size_t ci = rem % s[i];
        rem /= s[i];
        fidx += ci * mul;
        mul *= f[i];
      }
      size_t dst = reflected ? (F - 1 - fidx) : fidx;
      float y = data[c * nVar + v];
      out[(size_t)(3 * v) * F + dst] = std::isnan(y) ? 0.f : y * y;
      out[(size_t)(3 * v + 1) * F + dst] = std::isnan(y) ? 0.f : y;
      out[(size_t)(3 * v + 2) * F + dst] = std::isnan(y) ? 0.f : 1.f;
    }
  }
  return 3 * nVar;
}

// ---------------------------------------------------------------------------
// Cross-measurement coefficient layers — include/DataImage.hpp:626-710 with
// forXMeasurement (continuous variables only): regular layer diag(-1, +2, -1)
// over [y^2 x P0, y x P1, m x P2] (no constant term: delta0 = 0), and the
// cross layer alphaX(m, P0) = 1, i.e. the weighted count of TI cells with a
// value under the neighbours. convTypes: 0 = P0 (w), 1 = P1 (w x),
// 2 = P2 (w x^2). Returns nSub, or -1 for a categorical variable.
// ---------------------------------------------------------------------------
int orc_coef_matrix_x(int nVar, const int* types, float* alpha, float* alphaX, int* convTypes, int* subVarOf) {
  for (int v = 0; v < nVar; ++v)
    if (types[v] != 0) return -1;
  const int nSub = 3 * nVar;
  std::fill(alpha, alpha + nSub * nSub, 0.f);
  std::fill(alphaX, alphaX + nSub * nSub, 0.f);
  for (int v = 0; v < nVar; ++v) {
    const int b = 3 * v;
    alpha[b + b * nSub] = -1.f;
    alpha[(b + 1) + (b + 1) * nSub] = 2.f;
    alpha[(b + 2) + (b + 2) * nSub] = -1.f;
    alphaX[(b + 2) + b * nSub] = 1.f;
    for (int q = 0; q < 3; ++q) {
      convTypes[b + q] = q;
      subVarOf[b + q] = v;
    }
  }
  return nSub;
}

// ---------------------------------------------------------------------------
// a3. Coefficient matrix — include/DataImage.hpp:626-710 with
// include/OperationMatrix.hpp:9-73 (layout alpha[A + B*nSub]).
// No cross-measurement layer (TIs without NaN). Continuous: (P0, -1) for the
// y^2 image against w, (P1, +2) for y against w*x, constant P2 with -1.
// Categorical: (P1, +1) per class.
// convTypes[sub] receives the conversion of each neighbour sub-variable
// (0=P0, 1=P1). Returns nSub.
// ---------------------------------------------------------------------------
int orc_coef_matrix(int nVar, const int* types, const int* ncat, float* alpha, int* convTypes,
                    int* subVarOf) {
  int nSub = 0, catIdx = 0;
  for (int v = 0; v < nVar; ++v) nSub += (types[v] == 0) ? 2 : ncat[catIdx++];
  std::fill(alpha, alpha + nSub * nSub, 0.f);
  int pos = 0;
  catIdx = 0;
  int expanded = 0;  // index of the one-hot expanded variable
  for (int v = 0; v < nVar; ++v) {
    if (types[v] == 0) {
      alpha[pos + pos * nSub] = -1.f;
      convTypes[pos] = 0;
      subVarOf[pos] = expanded;
      pos++;
      alpha[pos + pos * nSub] = 2.f;
      convTypes[pos] = 1;
      subVarOf[pos] = expanded;
      pos++;
      expanded++;
    } else {
      for (int k = 0; k < ncat[catIdx]; ++k) {
        alpha[pos + pos * nSub] = 1.f;
        convTypes[pos] = 1;
        subVarOf[pos] = expanded;
        pos++;
        expanded++;
      }
      catIdx++;
    }
  }
  return nSub;
}

// ---------------------------------------------------------------------------
// Kernel generation — include/DataImage.hpp:433-480 (UNIFORM=0,
// EXPONENTIAL=2, GAUSSIAN=7) and distance2ToCenter :518-534.
// Output: cells(kdims) x nVar, variable fastest.
// ---------------------------------------------------------------------------
static float dist2_to_center(int D, const unsigned* dims, size_t index) {
  float value = 0.f;
  size_t val = index;
  for (int i = 0; i < D; ++i) {
    int d = std::abs(int(val % dims[i]) - (int(dims[i]) - 1) / 2);
    val /= dims[i];
    value += (float)(d * d);
  }
  return value;
}

void orc_kernel_generate(int D, const unsigned* dims, int nVar, const int* ktype, const float* weight,
                         const float* alphas, float* out) {
  size_t cells = 1;
  for (int i = 0; i < D; ++i) cells *= dims[i];
  for (int j = 0; j < nVar; ++j) {
    for (size_t i = 0; i < cells; ++i) {
      float v = 0.f;
      if (ktype[j] == 0) v = weight[j];
      // the reference calls the global (double) exp/sqrt on float arguments
      if (ktype[j] == 2) {
        float dist = (float)::sqrt((double)dist2_to_center(D, dims, i));
        v = (float)::exp((double)(-alphas[j] * dist));
      }
      if (ktype[j] == 7) {
        float d2 = dist2_to_center(D, dims, i) * alphas[j] * alphas[j];
        v = (float)(1 / ::sqrt(2 * M_PI) * ::exp((double)(-d2 / 2.f)));
      }
      out[i * nVar + j] = v;
    }
  }
}

// offsetKernel4categories — include/DataImage.hpp:482-508 (no X-measurement):
// channel i of the input kernel is repeated factor[i] times.
int orc_kernel_expand(size_t cells, int nVarIn, const float* kin, const int* factor, float* out) {
  int cum = 0;
  for (int i = 0; i < nVarIn; ++i) cum += factor[i];
  int pos = 0;
  for (int i = 0; i < nVarIn; ++i)
    for (int j = 0; j < factor[i]; ++j) {
      for (size_t k = 0; k < cells; ++k) out[k * cum + pos] = kin[k * nVarIn + i];
      pos++;
    }
  return cum;
}

// ---------------------------------------------------------------------------
// Neighbour offset list — src/qs.cpp:1055-1137: every offset of the kernel
// box (half-width (dim+1)/2-1 per dim) generated in the reference's order,
// NaN-kernel offsets removed, then std::sort by decreasing -|d|^2 (libstdc++
// introsort, unstable, so the generation order matters for ties).
// out: count x D ints. Returns the count.
// ---------------------------------------------------------------------------
int orc_offsets(int D, const unsigned* kdims, int kNVar, const float* kernel, int* out) {
  std::vector<std::vector<int>> pos;
  pos.push_back(std::vector<int>());
  for (int i = 0; i < D; ++i) {
    size_t orig = pos.size();
    int half = (int(kdims[i]) + 1) / 2;
    pos.resize(orig * (2 * half - 1));
    for (size_t k = 0; k < orig; ++k) pos[k].push_back(0);
    for (int j = 1; j < half; ++j) {
      std::copy(pos.begin(), pos.begin() + orig, pos.begin() + orig * (2 * j - 1));
      std::copy(pos.begin(), pos.begin() + orig, pos.begin() + orig * (2 * j));
      for (size_t k = orig * (2 * j - 1); k < orig * (2 * j); ++k) pos[k][i] = j;
      for (size_t k = orig * (2 * j); k < orig * (2 * j + 1); ++k) pos[k][i] = -j;
    }
  }
  // centre used by qs.cpp for the NaN removal and for the sort key
  auto kindex = [&](const std::vector<int>& v) {
    size_t idx = 0;
    for (int i = D - 1; i >= 0; --i) {
      int c = (int(kdims[i]) - 1) / 2 + v[i];
      c = ((c % int(kdims[i])) + int(kdims[i])) % int(kdims[i]);
      idx = idx * kdims[i] + (size_t)c;
    }
    return idx;
  };
  if (kernel) {
    pos.erase(std::remove_if(pos.begin(), pos.end(),
                             [&](std::vector<int>& v) { return std::isnan(kernel[kindex(v) * kNVar]); }),
              pos.end());
  }
  auto d2 = [&](const std::vector<int>& v) {
    return dist2_to_center(D, kdims, kindex(v));
  };
  std::sort(pos.begin(), pos.end(), [&](std::vector<int>& a, std::vector<int>& b) { return -d2(a) > -d2(b); });
  for (size_t k = 0; k < pos.size(); ++k)
    for (int i = 0; i < D; ++i) out[k * D + i] = pos[k][i];
  return (int)pos.size();
}

// ---------------------------------------------------------------------------
// a14. Random path and per-node seeds — src/qs.cpp:975, 1159-1201, 1250-1258.
// Cells fully informed in the DI go first (swap scan), the rest is shuffled
// with std::shuffle(mt19937(seed)); then one U[0,1) float per path entry
// from the same generator (redrawn once when it equals 1).
// informed: 1 per cell when every variable is set. Returns beginPath.
// ---------------------------------------------------------------------------
unsigned orc_path(size_t cells, const unsigned char* informed, unsigned seed, unsigned* path, float* seeds) {
  std::mt19937 gen(seed);
  for (size_t i = 0; i < cells; ++i) path[i] = (unsigned)i;
  size_t begin = 0;
  for (size_t i = 0; i < cells; ++i)
    if (informed[i]) {
      std::swap(path[begin], path[i]);
      begin++;
    }
  std::shuffle(path + begin, path + cells, gen);
  std::uniform_real_distribution<float> u(0.f, 1.f);
  for (size_t i = 0; i < cells; ++i) {
    seeds[i] = u(gen);
    if (seeds[i] == 1.f) seeds[i] = u(gen);
  }
  return (unsigned)begin;
}

// ---------------------------------------------------------------------------
// a9/a10. fKst::findKBiggest with the RNG tie rule — include/fKb.hpp:58-93,
// :121-132, :907-954 — seeded as QuantileSamplingModule does for a plain
// device (include/quantileSamplingModule.hpp:45-49, :533-539):
// mt19937.seed(seedFromUnitFloat(seed)); uniform_real_distribution<float>.
// The scalar path is the one the reference build uses (no -march flags).
// ---------------------------------------------------------------------------
static unsigned seed_from_unit_float(float seed) {
  if (!std::isfinite(seed) || seed <= 0.f) return 0;
  if (seed >= 1.f) return std::numeric_limits<unsigned>::max();
  return (unsigned)std::floor((double)seed * (double)std::numeric_limits<unsigned>::max());
}

unsigned orc_seed_from_unit_float(float seed) { return seed_from_unit_float(seed); }

}  // extern "C"

template <typename RNG>
static void fkst_add(const float* data, unsigned short k, float* out, unsigned* pos, RNG& rng, unsigned& cpt,
                     unsigned i) {
  short position = k - 2;
  short likeLast = k - 2;
  while (likeLast > -1 && out[k - 1] == out[likeLast]) likeLast--;
  int xes = k - likeLast - 1;
  int x = k - xes + int(std::floor(rng() * xes));
  std::swap(pos[x], pos[k - 1]);
  if (data[i] > out[k - 1]) {
    if (xes == 1)
      cpt = 0;
    else
      cpt++;
    while (position > -1 && data[i] > out[position]) {
      out[position + 1] = out[position];
      pos[position + 1] = pos[position];
      position--;
    }
    out[position + 1] = data[i];
    pos[position + 1] = i;
  } else {
    cpt++;
    if ((xes + cpt) * rng() < xes) pos[position + 1] = i;
  }
}

extern "C" {

void orc_fkst(const float* data, unsigned N, unsigned short k, float seed, float* out, unsigned* pos) {
  std::mt19937 gen;
  gen.seed(seed_from_unit_float(seed));
  std::uniform_real_distribution<float> dist(0.f, 1.f);
  auto rng = [&]() { return dist(gen); };
  std::fill(out, out + k, -INFINITY);
  std::fill(pos, pos + k, UINT_MAX);
  unsigned cpt = 0;
  for (unsigned i = 0; i < N; ++i)
    if (data[i] >= out[k - 1]) fkst_add(data, k, out, pos, rng, cpt, i);
}

// ---------------------------------------------------------------------------
// Device selection rule (the rule the MI355X select kernel implements).
// The reference's rule is sequential (reservoir draws in scan order); a
// parallel select cannot reproduce its RNG stream, so the device orders
// candidates by the strict total order
//   (value_q desc, hash(seed,position) desc, position asc)
// where value_q = value when quantum == 0, else rint(value/quantum) (exact
// ties made exact for integer-valued categorical maps), and
// hash = fmix32(position * 0x9E3779B1 + seedFromUnitFloat(seed)).
// Uniform among exact ties, deterministic given the node seed.
// ---------------------------------------------------------------------------
static inline uint32_t fmix32(uint32_t h) {
  h ^= h >> 16;
  h *= 0x85ebca6bu;
  h ^= h >> 13;
  h *= 0xc2b2ae35u;
  h ^= h >> 16;
  return h;
}

uint32_t orc_tie_hash(uint32_t seedbits, uint32_t position) { return fmix32(position * 0x9E3779B1u + seedbits); }

struct Cand {
  float key;
  uint32_t h;
  uint32_t pos;
  float val;
};
static inline bool cand_better(const Cand& a, const Cand& b) {
  if (a.key != b.key) return a.key > b.key;
  if (a.h != b.h) return a.h > b.h;
  return a.pos < b.pos;
}

// map: position-space map over the TI grid s (dim0 fastest); lo/hi: valid box
// (inclusive) per dim. Positions outside the box or with non-finite values
// are excluded. Returns the number of candidates found (<= K); the rest of
// out/pos is filled with -inf / UINT_MAX.
int orc_select_device(int D, const unsigned* s, const float* map, const int* lo, const int* hi, int K, float seed,
                      float quantum, float* outVal, unsigned* outPos) {
  uint32_t sb = seed_from_unit_float(seed);
  std::vector<Cand> best;
  size_t S = 1;
  for (int i = 0; i < D; ++i) S *= s[i];
  for (size_t c = 0; c < S; ++c) {
    size_t rem = c;
    bool ok = true;
    for (int i = 0; i < D; ++i) {
      int ci = int(rem % s[i]);
      rem /= s[i];
      ok &= (ci >= lo[i] && ci <= hi[i]);
    }
    float v = map[c];
    if (!ok || !std::isfinite(v)) continue;
    Cand cd;
    cd.val = v;
    cd.key = (quantum > 0.f) ? std::rint(v / quantum) : v;
    cd.pos = (uint32_t)c;
    cd.h = orc_tie_hash(sb, (uint32_t)c);
    if ((int)best.size() < K) {
      best.push_back(cd);
      std::sort(best.begin(), best.end(), cand_better);
    } else if (cand_better(cd, best.back())) {
      best.back() = cd;
      std::sort(best.begin(), best.end(), cand_better);
    }
  }
  for (int i = 0; i < K; ++i) {
    if (i < (int)best.size()) {
      outVal[i] = best[i].val;
      outPos[i] = best[i].pos;
    } else {
      outVal[i] = -INFINITY;
      outPos[i] = UINT_MAX;
    }
  }
  return (int)best.size();
}

// The device rule on an FP64 map (keys = the FP64 values, or rint(value /
// quantum) with a tie quantum): the selection the library's exact selection
// makes. Same conventions as orc_select_device.
int orc_select_device_f64(int D, const unsigned* s, const double* map, const int* lo, const int* hi, int K,
                          float seed, float quantum, double* outVal, unsigned* outPos) {
  uint32_t sb = seed_from_unit_float(seed);
  struct Cd {
    double key;
    uint32_t h, pos;
    double val;
  };
  auto better = [](const Cd& a, const Cd& b) {
    if (a.key != b.key) return a.key > b.key;
    if (a.h != b.h) return a.h > b.h;
    return a.pos < b.pos;
  };
  std::vector<Cd> best;
  size_t S = 1;
  for (int i = 0; i < D; ++i) S *= s[i];
  for (size_t c = 0; c < S; ++c) {
    size_t rem = c;
    bool ok = true;
    for (int i = 0; i < D; ++i) {
      int ci = int(rem % s[i]);
      rem /= s[i];
      ok &= (ci >= lo[i] && ci <= hi[i]);
    }
    const double v = map[c];
    if (!ok || !std::isfinite(v)) continue;
    Cd cd{quantum > 0.f ? std::rint(v / (double)quantum) : v, orc_tie_hash(sb, (uint32_t)c), (uint32_t)c, v};
    if ((int)best.size() < K) {
      best.push_back(cd);
      std::sort(best.begin(), best.end(), better);
    } else if (better(cd, best.back())) {
      best.back() = cd;
      std::sort(best.begin(), best.end(), better);
    }
  }
  for (int i = 0; i < K; ++i) {
    if (i < (int)best.size()) {
      outVal[i] = best[i].val;
      outPos[i] = best[i].pos;
    } else {
      outVal[i] = -INFINITY;
      outPos[i] = UINT_MAX;
    }
  }
  return (int)best.size();
}

// ---------------------------------------------------------------------------
// a11. Candidate draw — include/quantileSamplingModule.hpp:101-113:
// rank = floor(seed * k * (ceil(nTI/k)/nTI)), computed in float like the
// reference (seed*localk*... with float localk).
// ---------------------------------------------------------------------------
unsigned orc_draw_rank(float seed, float k, unsigned nTI) {
  return (unsigned)int(std::floor(seed * k * (std::ceil(nTI / k) / nTI)));
}

// ---------------------------------------------------------------------------
// a12. Index <-> position maps — src/CPUThreadDevice.cpp:257-284.
// ---------------------------------------------------------------------------
unsigned orc_index_to_position(int D, const unsigned* f, const unsigned* s, const int* mn, unsigned index) {
  unsigned position = 0;
  unsigned div = 1;
  for (int i = 0; i < D; ++i) div *= f[i];
  for (int i = D - 1; i >= 0; --i) {
    div /= f[i];
    position = position * s[i] + ((f[i] - (index / div) % f[i] - mn[i] - 1) + s[i]) % s[i];
  }
  return position;
}

unsigned orc_position_to_index(int D, const unsigned* f, const unsigned* s, const int* mn, unsigned position) {
  unsigned index = 0;
  unsigned div = 1;
  for (int i = D - 1; i >= 0; --i) div *= s[i];
  for (int i = D - 1; i >= 0; --i) {
    div /= s[i];
    index = index * f[i] + (f[i] - (position / div + mn[i]) % s[i] - 1);
  }
  return index;
}

// ---------------------------------------------------------------------------
// Mismatch maps.
//
// Notation: nb neighbours with offsets off[j*D+i] (relative to the simulated
// cell, the cell itself included), converted values u[j*nSub+B] (QSM
// conversion, quantileSamplingModule.hpp:368-410), alpha[A+B*nSub], TI
// sub-images Y_A unreflected on the TI grid s.
//
// (1) orc_map_direct_f64 — the closed form in position space, FP64:
//     E(c) = delta0 + sum_j sum_{A,B} alpha_AB u_jB Y_A(c + d_j)
//     for every TI cell c whose whole neighbourhood lies inside the TI
//     (c_i in [-min_i, s_i-1-max_i], min/max start at 0 as in
//     src/AcceleratorDevice.cpp:95-108); -inf elsewhere (or, circular,
//     c + d_j wrapped and every c valid).
// (2) orc_map_refflow — the reference dataflow in index space
//     (src/CPUThreadAcceleratorDevice.cpp:251-357 via
//     src/AcceleratorDevice.cpp:92-136): reflected TI spectra, neighbour
//     images shifted by -min, partial r2c, MAC alpha*C*D
//     (include/complexMulti.hpp:58-67), unnormalised c2r, wrap-around mask to
//     -inf per dim, then /N + delta0 when trueMismatch. Templated on the
//     arithmetic precision (float = the reference's FFTW single precision).
// ---------------------------------------------------------------------------
void orc_bbox(int D, int nb, const int* off, int* mn, int* mx) {
  for (int i = 0; i < D; ++i) mn[i] = mx[i] = 0;
  for (int j = 0; j < nb; ++j)
    for (int i = 0; i < D; ++i) {
      mn[i] = std::min(mn[i], off[j * D + i]);
      mx[i] = std::max(mx[i], off[j * D + i]);
    }
}

void orc_map_direct_f64(int D, const unsigned* s, int nSub, const float* Y /* nSub x S */, const float* alpha,
                        int nb, const int* off, const float* u, double delta0, int circular, double* out) {
  size_t S = 1;
  for (int i = 0; i < D; ++i) S *= s[i];
  int mn[3], mx[3];
  orc_bbox(D, nb, off, mn, mx);
  // v_jA = sum_B alpha_AB u_jB
  std::vector<double> v((size_t)nb * nSub, 0.0);
  for (int j = 0; j < nb; ++j)
    for (int A = 0; A < nSub; ++A) {
      double acc = 0;
      for (int B = 0; B < nSub; ++B) acc += (double)alpha[A + B * nSub] * (double)u[j * nSub + B];
      v[(size_t)j * nSub + A] = acc;
    }
#pragma omp parallel for schedule(static)
  for (long long cc = 0; cc < (long long)S; ++cc) {
    size_t c = (size_t)cc;
    int ci[3] = {0, 0, 0};
    size_t rem = c;
    bool ok = true;
    for (int i = 0; i < D; ++i) {
      ci[i] = int(rem % s[i]);
      rem /= s[i];
      if (!circular) ok &= (ci[i] >= -mn[i] && ci[i] <= int(s[i]) - 1 - mx[i]);
    }
    if (!ok) {
      out[c] = -INFINITY;
      continue;
    }
    double acc = delta0;
    for (int j = 0; j < nb; ++j) {
      size_t t = 0, mul = 1;
      for (int i = 0; i < D; ++i) {
        int ti = ci[i] + off[j * D + i];
        if (circular) ti = ((ti % int(s[i])) + int(s[i])) % int(s[i]);
        t += (size_t)ti * mul;
        mul *= s[i];
      }
      for (int A = 0; A < nSub; ++A) acc += v[(size_t)j * nSub + A] * (double)Y[(size_t)A * S + t];
    }
    out[c] = acc;
  }
}

}  // extern "C"

template <typename T>
static void map_refflow(int D, const unsigned* s_, const unsigned* f_, int nSub,
                        const std::complex<T>* Yhat /* nSub x half, spectra of reflected images */,
                        const float* alpha, int nb, const int* off, const float* u, float delta0, int circular,
                        int trueMismatch, float* out /* f-grid, index space */) {
  using C = std::complex<T>;
  FftND<T> fft(D, f_);
  size_t N = fft.real_size(), H = fft.half_size();
  // extents padded to 3 dimensions
  unsigned f[3] = {1, 1, 1}, s[3] = {1, 1, 1};
  for (int i = 0; i < D; ++i) {
    f[i] = f_[i];
    s[i] = s_[i];
  }
  int mn[3] = {0, 0, 0}, mx[3] = {0, 0, 0};
  orc_bbox(D, nb, off, mn, mx);
  std::vector<C> acc(H, C(0, 0)), in(H);
  std::vector<T> real(N);
  for (int B = 0; B < nSub; ++B) {
    bool need = false;
    for (int A = 0; A < nSub; ++A) need |= alpha[A + B * nSub] != 0.f;
    if (!need) continue;
    std::fill(real.begin(), real.end(), (T)0);
    std::vector<char> rows(f[1] * f[2], 0);
    for (int j = 0; j < nb; ++j) {
      size_t idx = 0, mul = 1;
      for (int i = 0; i < D; ++i) {
        idx += (size_t)(off[j * D + i] - mn[i]) * mul;
        mul *= f[i];
      }
      real[idx] = (T)u[j * nSub + B];
      rows[idx / f[0]] = 1;
    }
    fft.r2c(real.data(), in.data(), &rows);
    for (int A = 0; A < nSub; ++A) {
      float a = alpha[A + B * nSub];
      if (a == 0.f) continue;
      const C* y = Yhat + (size_t)A * H;
      for (size_t h = 0; h < H; ++h) acc[h] += (T)a * (y[h] * in[h]);
    }
  }
  fft.c2r(acc.data(), real.data());
  for (size_t i = 0; i < N; ++i) out[i] = (float)real[i];
  if (!circular) {
    for (int i = 0; i < D; ++i) {
      size_t block = 1;
      for (int j = 0; j < i; ++j) block *= f[j];
      block *= f[i] - (s[i] - (mx[i] - mn[i]));
      size_t delta = 1;
      for (int j = 0; j <= i; ++j) delta *= f[j];
      for (size_t j = 0; j < N; j += delta) std::fill(out + j, out + j + block, -INFINITY);
    }
  }
  if (trueMismatch)
    for (size_t i = 0; i < N; ++i) out[i] = out[i] / (float)N + delta0;
}

extern "C" {

// Spectra of the (reflected, f-grid) TI sub-images; double or float.
void orc_ti_spectra_f64(int D, const unsigned* f, int nSub, const float* imgs, double* outInterleaved) {
  FftND<double> fft(D, f);
  size_t N = fft.real_size(), H = fft.half_size();
  std::vector<double> r(N);
  auto* o = reinterpret_cast<std::complex<double>*>(outInterleaved);
  for (int A = 0; A < nSub; ++A) {
    for (size_t i = 0; i < N; ++i) r[i] = imgs[(size_t)A * N + i];
    fft.r2c(r.data(), o + (size_t)A * H);
  }
}

void orc_ti_spectra_f32(int D, const unsigned* f, int nSub, const float* imgs, float* outInterleaved) {
  FftND<float> fft(D, f);
  size_t N = fft.real_size(), H = fft.half_size();
  auto* o = reinterpret_cast<std::complex<float>*>(outInterleaved);
  for (int A = 0; A < nSub; ++A) fft.r2c(imgs + (size_t)A * N, o + (size_t)A * H);
}

void orc_map_refflow_f64(int D, const unsigned* s, const unsigned* f, int nSub, const double* Yhat,
                         const float* alpha, int nb, const int* off, const float* u, float delta0, int circular,
                         int trueMismatch, float* out) {
  map_refflow<double>(D, s, f, nSub, reinterpret_cast<const std::complex<double>*>(Yhat), alpha, nb, off, u, delta0,
                      circular, trueMismatch, out);
}

void orc_map_refflow_f32(int D, const unsigned* s, const unsigned* f, int nSub, const float* Yhat,
                         const float* alpha, int nb, const int* off, const float* u, float delta0, int circular,
                         int trueMismatch, float* out) {
  map_refflow<float>(D, s, f, nSub, reinterpret_cast<const std::complex<float>*>(Yhat), alpha, nb, off, u, delta0,
                     circular, trueMismatch, out);
}

// 1-D transform exported for the FFT pin against numpy.
void orc_fft1d_f64(size_t n, double* inout, int sign) {
  FftPlan1D<double> p(n);
  std::vector<std::complex<double>> w;
  p.exec(reinterpret_cast<std::complex<double>*>(inout), sign, w);
}

void orc_rfftn_f64(int D, const unsigned* f, const double* in, double* outInterleaved) {
  FftND<double> fft(D, f);
  fft.r2c(in, reinterpret_cast<std::complex<double>*>(outInterleaved));
}

void orc_irfftn_f64(int D, const unsigned* f, double* specInterleaved, double* out) {
  FftND<double> fft(D, f);
  fft.c2r(reinterpret_cast<std::complex<double>*>(specInterleaved), out);
}

// ---------------------------------------------------------------------------
// a5. Neighbour conversion and delta0 — include/quantileSamplingModule.hpp:
// 362-457. Input: nb neighbours (offsets off, one-hot expanded values
// vals[j*nExp + e], NaN for "no value"), the expanded kernel (kdims,
// nExp channels). Each neighbour sub-variable B belongs to expanded variable
// subVarOf[B] with conversion convTypes[B] (0:P0=w, 1:P1=w*x). w is the
// kernel value at the offset (kernel centre dims/2, DataImage::indexWithDelta
// include/DataImage.hpp:343-376); offsets outside the kernel or NaN values
// give 0. delta0 = sum over ORIGINAL continuous variables i of
// -w x^2, where the reference reads the expanded value and kernel channel at
// the original index i (quantileSamplingModule.hpp:415-455; identical to the
// expanded index whenever continuous variables come first), accumulated in
// float in the reference's loop order.
// ---------------------------------------------------------------------------
// DataImage::indexWithDelta — include/DataImage.hpp:343-376: the cell at
// `position + delta` with each coordinate wrapped into the grid; returns
// whether no wrap was needed.
int orc_index_with_delta(int D, const unsigned* dims, unsigned position, const int* delta, unsigned* location) {
  int val[3];
  bool ok = true;
  for (int i = 0; i < D; ++i) {
    val[i] = int(position % dims[i]) + delta[i];
    ok &= (val[i] >= 0) && (val[i] < int(dims[i]));
    const int dim = int(dims[i]);
    val[i] = ((val[i] % dim) + dim) % dim;
    position /= dims[i];
  }
  unsigned fin = 0;
  for (int i = D - 1; i >= 0; --i) fin = fin * dims[i] + (unsigned)val[i];
  *location = fin;
  return ok ? 1 : 0;
}

static bool kernel_index(int D, const unsigned* kdims, const int* d, size_t* kidx) {
  // QSM's kernel centre: dims/2 per dimension (quantileSamplingModule.hpp:362-366)
  unsigned centre = 0;
  for (int i = D - 1; i >= 0; --i) centre = centre * kdims[i] + kdims[i] / 2;
  unsigned loc = 0;
  int ok = orc_index_with_delta(D, kdims, centre, d, &loc);
  *kidx = loc;
  return ok != 0;
}

float orc_convert_neighbors(int D, int nb, const int* off, int nExp, const float* vals, int nSub,
                            const int* convTypes, const int* subVarOf, int nVarOrig, const int* typesOrig,
                            const unsigned* kdims, const float* kernel, float* u) {
  for (int j = 0; j < nb; ++j) {
    size_t kidx;
    bool inside = kernel_index(D, kdims, off + j * D, &kidx);
    for (int B = 0; B < nSub; ++B) {
      int e = subVarOf[B];
      float x = vals[j * nExp + e];
      float w = kernel[kidx * nExp + (e % nExp)];
      float r = 0.f;
      if (inside && !std::isnan(x)) r = convTypes[B] == 0 ? w * 1.f : convTypes[B] == 2 ? w * x * x : w * x;
      u[j * nSub + B] = r;
    }
  }
  float delta0 = 0.f;
  for (int i = 0; i < nVarOrig; ++i) {
    if (typesOrig[i] != 0) continue;
    for (int j = 0; j < nb; ++j) {
      size_t kidx;
      bool inside = kernel_index(D, kdims, off + j * D, &kidx);
      float x = vals[j * nExp + i];
      if (inside && !std::isnan(x)) delta0 += -1.f * (kernel[kidx * nExp + (i % nExp)] * x * x);
    }
  }
  return delta0;
}

}  // extern "C"
