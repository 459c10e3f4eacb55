// QuickSampling host layer above the qsmi C-ABI: the steps of qs.cpp,
// QuantileSamplingModule and simulation() around the scan, restated in C++
// for the MI355X device (include/qsmi.h, "host layer").
//
// The path loop differs from the reference's in scheduling only: the
// neighbour set of every path node depends on the path and the geometry, not
// on simulated values (include/simulation.hpp:447-511 takes the first n cells
// along the offset list whose path order is <= the node's), so the host
// computes all neighbour sets first, gives each node a dependency level
// (1 + the highest level among the path nodes it reads), and scans each level
// in waves of independent nodes. Every node therefore reads exactly the
// values the serial loop (and the reference's spin-waiting OpenMP loop,
// simulation.hpp:463-473) would give it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "../../include/qsmi.h"
#include "qsmi_internal.hpp"

namespace {

}  // namespace

extern "C" int qsmi_internal_fail(int code, const char* msg);

namespace {

int hfail(int code, const std::string& m) { return qsmi_internal_fail(code, m.c_str()); }

struct KernelImg {
  int D = 0;
  unsigned dims[3] = {1, 1, 1};
  int nch = 0;
  std::vector<float> data;  // cells x nch
};

float dist2_center(int D, const unsigned* dims, const int* d) {
  float v = 0.f;
  for (int i = 0; i < D; ++i) v += (float)(d[i] * d[i]);
  (void)dims;
  return v;
}

// kernel cell of an offset relative to the kernel centre `ctr`; false when
// the offset leaves the kernel (wrapped index still returned), as
// DataImage::indexWithDelta (include/DataImage.hpp:343-376)
bool kernel_cell(const KernelImg& k, const int* ctr, const int* d, size_t* cell) {
  bool ok = true;
  size_t idx = 0;
  for (int i = k.D - 1; i >= 0; --i) {
    int c = ctr[i] + d[i];
    ok &= (c >= 0 && c < (int)k.dims[i]);
    c = ((c % (int)k.dims[i]) + (int)k.dims[i]) % (int)k.dims[i];
    idx = idx * k.dims[i] + (size_t)c;
  }
  *cell = idx;
  return ok;
}

}  // namespace

extern "C" {

int qsmi_categories(size_t cells, int nVar, const float* data, int var, float* out, int maxOut, int* count) {
  if (!data || !out || !count || var < 0 || var >= nVar) return hfail(QSMI_ERR_ARG, "invalid argument");
  int n = 0;
  for (size_t c = 0; c < cells; ++c) {
    const float v = data[c * nVar + var];
    if (std::isnan(v)) continue;
    bool seen = false;
    for (int q = 0; q < n && !seen; ++q) seen = (out[q] == v);
    if (seen) continue;
    if (n >= maxOut) return hfail(QSMI_ERR_UNSUPPORTED, "too many categories");
    out[n++] = v;
  }
  *count = n;
  return QSMI_OK;
}

int qsmi_coef_matrix(int nVar, const int* types, const int* ncat, float* alpha, int* convTypes, int* subVarOf,
                     int* nSubOut) {
  if (!types || !alpha || !convTypes || !subVarOf || !nSubOut) return hfail(QSMI_ERR_ARG, "null argument");
  int nSub = 0, ci = 0;
  for (int v = 0; v < nVar; ++v) nSub += types[v] == 0 ? 2 : ncat[ci++];
  if (nSub > 16) return hfail(QSMI_ERR_UNSUPPORTED, "more than 16 sub-variables");
  std::fill(alpha, alpha + nSub * nSub, 0.f);
  int b = 0, e = 0;
  ci = 0;
  for (int v = 0; v < nVar; ++v) {
    if (types[v] == 0) {
      // (y^2, w) with -1 and (y, w*x) with +2; the P2 term goes to delta0
      alpha[b + b * nSub] = -1.f;
      convTypes[b] = 0;
      subVarOf[b++] = e;
      alpha[b + b * nSub] = 2.f;
      convTypes[b] = 1;
      subVarOf[b++] = e++;
    } else {
      for (int q = 0; q < ncat[ci]; ++q) {
        alpha[b + b * nSub] = 1.f;
        convTypes[b] = 1;
        subVarOf[b++] = e++;
      }
      ci++;
    }
  }
  *nSubOut = nSub;
  return QSMI_OK;
}

int qsmi_ti_subimages(int ndim, const unsigned* s, int nVar, const int* types, const float* ti, const int* ncat,
                      const float* catvals, float* out) {
  if (!s || !types || !ti || !out || ndim < 1 || ndim > 3) return hfail(QSMI_ERR_ARG, "invalid argument");
  size_t S = 1;
  for (int i = 0; i < ndim; ++i) S *= s[i];
  size_t b = 0;
  int ci = 0;
  const float* cv = catvals;
  for (int v = 0; v < nVar; ++v) {
    if (types[v] == 0) {
      float* y2 = out + b * S;
      float* y1 = out + (b + 1) * S;
      for (size_t c = 0; c < S; ++c) {
        const float y = ti[c * nVar + v];
        y2[c] = std::isnan(y) ? 0.f : y * y;
        y1[c] = std::isnan(y) ? 0.f : y;
      }
      b += 2;
    } else {
      for (int q = 0; q < ncat[ci]; ++q) {
        float* o = out + (b + q) * S;
        for (size_t c = 0; c < S; ++c) o[c] = (ti[c * nVar + v] == cv[q]) ? 1.f : 0.f;
      }
      b += ncat[ci];
      cv += ncat[ci];
      ci++;
    }
  }
  return QSMI_OK;
}

// The point reflection of DataImage::convertInput4Xcorr (include/
// DataImage.hpp:538-599) undone: TI cell c of sub-image A sits at
// N - 1 - flat_fft(c) of the reflected image, N = prod(fftSize).
int qsmi_subimages_from_reflected(int ndim, const unsigned* srcSize, const unsigned* fftSize, int nSub,
                                  const float* const* reflected, float* out) {
  if (!srcSize || !fftSize || !reflected || !out || ndim < 1 || ndim > 3 || nSub < 1)
    return hfail(QSMI_ERR_ARG, "invalid argument");
  unsigned s[3] = {1, 1, 1}, f[3] = {1, 1, 1};
  for (int i = 0; i < ndim; ++i) {
    s[i] = srcSize[i];
    f[i] = fftSize[i];
    if (s[i] < 1 || f[i] < s[i]) return hfail(QSMI_ERR_ARG, "fftSize must be >= srcSize >= 1");
  }
  const size_t S = (size_t)s[0] * s[1] * s[2];
  const size_t N = (size_t)f[0] * f[1] * f[2];
  for (int A = 0; A < nSub; ++A) {
    const float* r = reflected[A];
    if (!r) return hfail(QSMI_ERR_ARG, "null sub-image pointer");
    float* o = out + (size_t)A * S;
    size_t inside = 0;
    for (unsigned c2 = 0; c2 < s[2]; ++c2)
      for (unsigned c1 = 0; c1 < s[1]; ++c1)
        for (unsigned c0 = 0; c0 < s[0]; ++c0) {
          const size_t flat = ((size_t)c2 * f[1] + c1) * f[0] + c0;
          o[((size_t)c2 * s[1] + c1) * s[0] + c0] = r[N - 1 - flat];
          inside++;
        }
    // the padding (reflected cells outside the TI box) must be zero
    if (inside != N) {
      for (size_t q = 0; q < N; ++q) {
        const size_t flat = N - 1 - q;
        const unsigned c0 = (unsigned)(flat % f[0]), c1 = (unsigned)((flat / f[0]) % f[1]),
                       c2 = (unsigned)(flat / ((size_t)f[0] * f[1]));
        if (c0 < s[0] && c1 < s[1] && c2 < s[2]) continue;
        if (r[q] != 0.f)
          return hfail(QSMI_ERR_ARG, "sub-image not in the reflected FFT-grid layout (non-zero padding)");
      }
    }
  }
  return QSMI_OK;
}

int qsmi_ti_create_from_reflected(int gpu, int ndim, const unsigned* srcSize, const unsigned* fftSize, int nSub,
                                  const float* const* reflected, qsmi_ti** out) {
  if (!out) return hfail(QSMI_ERR_ARG, "null argument");
  *out = nullptr;
  if (!srcSize || !fftSize || ndim < 1 || ndim > 3 || nSub < 1) return hfail(QSMI_ERR_ARG, "invalid argument");
  size_t S = 1;
  for (int i = 0; i < ndim; ++i) S *= srcSize[i];
  std::vector<float> sub((size_t)nSub * S);
  int rc = qsmi_subimages_from_reflected(ndim, srcSize, fftSize, nSub, reflected, sub.data());
  if (rc) return rc;
  return qsmi_ti_create(gpu, ndim, srcSize, fftSize, nSub, sub.data(), out);
}

int qsmi_coef_matrix_x(int nVar, const int* types, float* alpha, float* alphaX, int* convTypes, int* subVarOf,
                       int* nSubOut) {
  if (!types || !alpha || !alphaX || !convTypes || !subVarOf || !nSubOut) return hfail(QSMI_ERR_ARG, "null argument");
  for (int v = 0; v < nVar; ++v)
    if (types[v] != 0) return hfail(QSMI_ERR_UNSUPPORTED, "cross-measurement layer: continuous variables only");
  const int nSub = 3 * nVar;
  if (nSub > 16) return hfail(QSMI_ERR_UNSUPPORTED, "more than 16 sub-variables");
  std::fill(alpha, alpha + nSub * nSub, 0.f);
  std::fill(alphaX, alphaX + nSub * nSub, 0.f);
  for (int v = 0; v < nVar; ++v) {
    // [y^2 x P0] -1, [y x P1] +2, [m x P2] -1; count layer [m x P0] +1
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
  *nSubOut = nSub;
  return QSMI_OK;
}

int qsmi_ti_subimages_x(int ndim, const unsigned* s, int nVar, const int* types, const float* ti, float* out) {
  if (!s || !types || !ti || !out || ndim < 1 || ndim > 3) return hfail(QSMI_ERR_ARG, "invalid argument");
  for (int v = 0; v < nVar; ++v)
    if (types[v] != 0) return hfail(QSMI_ERR_UNSUPPORTED, "cross-measurement layer: continuous variables only");
  size_t S = 1;
  for (int i = 0; i < ndim; ++i) S *= s[i];
  for (int v = 0; v < nVar; ++v) {
    float* y2 = out + (size_t)(3 * v) * S;
    float* y1 = out + (size_t)(3 * v + 1) * S;
    float* m = out + (size_t)(3 * v + 2) * S;
    for (size_t c = 0; c < S; ++c) {
      const float y = ti[c * nVar + v];
      const bool nan = std::isnan(y);
      y2[c] = nan ? 0.f : y * y;
      y1[c] = nan ? 0.f : y;
      m[c] = nan ? 0.f : 1.f;
    }
  }
  return QSMI_OK;
}

// DataImage::genearteKernel (include/DataImage.hpp:433-480), qs -kernel /
// -ks / -alpha (src/qs.cpp:1010-1042): per variable j a uniform weight
// (type 0), exp(-alpha_j * distance) (type 2) or the unit normal density at
// alpha_j * distance (type 7) over a grid centred at (dims - 1) / 2; other
// types leave zeros, as the reference's empty cases do. The distance is
// formed in float and the exponentials in double, as the reference's calls
// of the global exp / sqrt on float arguments do.
int qsmi_kernel_generate(int ndim, const unsigned* dims, int nVar, const int* types, const float* weight,
                         const float* alphas, float* out) {
  if (ndim < 1 || ndim > 3 || nVar < 1 || !dims || !types || !weight || !alphas || !out)
    return hfail(QSMI_ERR_ARG, "invalid argument");
  size_t cells = 1;
  for (int i = 0; i < ndim; ++i) {
    if (dims[i] < 1) return hfail(QSMI_ERR_ARG, "zero kernel extent");
    cells *= dims[i];
  }
  for (size_t c = 0; c < cells; ++c) {
    float d2 = 0.f;  // squared distance to the centre, in float
    size_t rem = c;
    for (int i = 0; i < ndim; ++i) {
      const int off = std::abs((int)(rem % dims[i]) - ((int)dims[i] - 1) / 2);
      rem /= dims[i];
      d2 += (float)(off * off);
    }
    for (int j = 0; j < nVar; ++j) {
      float v = 0.f;
      if (types[j] == 0) {
        v = weight[j];
      } else if (types[j] == 2) {
        const float dist = (float)::sqrt((double)d2);
        v = (float)::exp((double)(-alphas[j] * dist));
      } else if (types[j] == 7) {
        const float a2 = d2 * alphas[j] * alphas[j];
        v = (float)(1.0 / ::sqrt(2.0 * M_PI) * ::exp((double)(-a2 / 2.f)));
      }
      out[c * nVar + j] = v;
    }
  }
  return QSMI_OK;
}

int qsmi_offsets(int ndim, const unsigned* ks, int kNVar, const float* kernel, int* out, int capacity, int* count) {
  if (!ks || !out || !count || ndim < 1 || ndim > 3) return hfail(QSMI_ERR_ARG, "invalid argument");
  // generation order of src/qs.cpp:1055-1084: dimension by dimension, each
  // existing offset followed by its +j then -j copies
  std::vector<std::vector<int>> list(1);
  for (int i = 0; i < ndim; ++i) {
    const size_t base = list.size();
    const int half = ((int)ks[i] + 1) / 2;
    for (auto& v : list) v.push_back(0);
    list.resize(base * (size_t)(2 * half - 1));
    for (int j = 1; j < half; ++j) {
      for (size_t q = 0; q < base; ++q) {
        std::vector<int> a = list[q], b = list[q];
        a[i] = j;
        b[i] = -j;
        list[base * (size_t)(2 * j - 1) + q] = a;
        list[base * (size_t)(2 * j) + q] = b;
      }
    }
  }
  if (kernel) {  // NaN kernel weights remove the offset (centre (size-1)/2)
    std::vector<std::vector<int>> kept;
    kept.reserve(list.size());
    for (auto& v : list) {
      size_t idx = 0;
      for (int i = ndim - 1; i >= 0; --i) {
        int c = ((int)ks[i] - 1) / 2 + v[i];
        c = ((c % (int)ks[i]) + (int)ks[i]) % (int)ks[i];
        idx = idx * ks[i] + (size_t)c;
      }
      if (!std::isnan(kernel[idx * kNVar])) kept.push_back(v);
    }
    list.swap(kept);
  }
  // the reference compares -|d|^2 with '>'; the same comparisons in the same
  // libstdc++ introsort give the same order of equal-distance offsets
  std::sort(list.begin(), list.end(), [&](const std::vector<int>& a, const std::vector<int>& b) {
    return -dist2_center(ndim, ks, a.data()) > -dist2_center(ndim, ks, b.data());
  });
  if ((int)list.size() > capacity) return hfail(QSMI_ERR_ARG, "offset capacity too small");
  for (size_t q = 0; q < list.size(); ++q)
    for (int i = 0; i < ndim; ++i) out[q * ndim + i] = list[q][i];
  *count = (int)list.size();
  return QSMI_OK;
}

int qsmi_path(size_t cells, const unsigned char* informed, unsigned seed, unsigned* path, float* seeds,
              unsigned* beginPath) {
  if (!informed || !path || !seeds || !beginPath) return hfail(QSMI_ERR_ARG, "null argument");
  std::mt19937 gen(seed);
  for (size_t i = 0; i < cells; ++i) path[i] = (unsigned)i;
  size_t begin = 0;
  for (size_t i = 0; i < cells; ++i)
    if (informed[i]) std::swap(path[begin++], path[i]);
  std::shuffle(path + begin, path + cells, gen);
  std::uniform_real_distribution<float> u(0.f, 1.f);
  for (size_t i = 0; i < cells; ++i) {
    seeds[i] = u(gen);
    if (seeds[i] == 1.f) seeds[i] = u(gen);
  }
  *beginPath = (unsigned)begin;
  return QSMI_OK;
}

// ---------------------------------------------------------------------------
// In-memory QS run
// ---------------------------------------------------------------------------
int qsmi_qs_run_in_memory(const qsmi_qs_request* rq, qsmi_qs_result* rs) {
  using clk = std::chrono::steady_clock;
  if (!rq || !rs || !rs->simulation || !rs->index) return hfail(QSMI_ERR_ARG, "null argument");
  const int D = rq->ndim;
  if (D < 1 || D > 3) return hfail(QSMI_ERR_UNSUPPORTED, "ndim must be 1, 2 or 3");
  const int nTI = rq->nTI > 1 ? rq->nTI : 1;
  if (rq->nVar < 1 || !rq->varTypes || !rq->di || (nTI == 1 && !rq->ti) || (nTI > 1 && !rq->tis))
    return hfail(QSMI_ERR_ARG, "invalid variables");
  if (nTI > 8) return hfail(QSMI_ERR_UNSUPPORTED, "at most 8 training images");
  if ((!rq->nImage && rq->n < 1) || !(rq->k > 0.f) || std::ceil(rq->k) > 16.f)
    return hfail(QSMI_ERR_UNSUPPORTED, "need n >= 1, 0 < k <= 16");
  if (rq->circularTI && (rq->lp || rq->recordCap > 0))
    return hfail(QSMI_ERR_UNSUPPORTED, "circular TI: not with the learned proposal or the record");
  if (rq->circularSim && (rq->lp || rq->recordCap > 0))
    return hfail(QSMI_ERR_UNSUPPORTED, "circular simulation: not with the learned proposal or the record");
  if ((rq->nImage || rq->kImage) && (rq->lp || rq->recordCap > 0))
    return hfail(QSMI_ERR_UNSUPPORTED, "n / k images: not with the learned proposal or the record");
  const auto t0 = clk::now();
  const int nVar = rq->nVar;
  unsigned s[3] = {1, 1, 1}, sd[3] = {1, 1, 1};
  size_t S = 1, SD = 1;
  for (int i = 0; i < D; ++i) {
    s[i] = rq->tiSize[i];
    sd[i] = rq->simSize[i];
    S *= s[i];
    SD *= sd[i];
  }
  // n and k images (include/simulation.hpp:337-357): the largest n sizes
  // the neighbour lists, the largest k the candidate lists
  int nMax = rq->n;
  float kMax = rq->k;
  if (rq->nImage) {
    nMax = 1;
    for (size_t c = 0; c < SD; ++c)
      if (std::isfinite(rq->nImage[c]) && rq->nImage[c] < 1e6f) nMax = std::max(nMax, (int)rq->nImage[c]);
    if (nMax > 4096) return hfail(QSMI_ERR_UNSUPPORTED, "n image: values above 4096");
  }
  if (rq->kImage) {
    for (size_t c = 0; c < SD; ++c)
      if (rq->kImage[c] > kMax) kMax = rq->kImage[c];
    if (!(std::ceil(kMax) <= 16.f)) return hfail(QSMI_ERR_UNSUPPORTED, "k image: values above 16");
  }
  // several TIs: one array, TI after TI (the order qs walks them in for
  // categories, src/qs.cpp:1498-1525, and the marginal draw)
  std::vector<float> tiCat;
  const float* tiIn = rq->ti;
  // extents and cell offsets of the TIs in the concatenation
  std::vector<std::array<unsigned, 3>> sT(nTI);
  std::vector<size_t> cumT(nTI + 1, 0);
  for (int t = 0; t < nTI; ++t) {
    size_t c = 1;
    for (int i = 0; i < 3; ++i) {
      sT[t][i] = i < D ? ((nTI > 1 && rq->tiSizes) ? rq->tiSizes[t * 3 + i] : s[i]) : 1u;
      if (sT[t][i] < 1) return hfail(QSMI_ERR_ARG, "zero TI extent");
      c *= sT[t][i];
    }
    cumT[t + 1] = cumT[t] + c;
  }
  for (int i = 0; i < D; ++i)
    if (sT[0][i] != s[i]) return hfail(QSMI_ERR_ARG, "tiSizes[0..2] must equal tiSize");
  if (nTI > 1) {
    size_t maxCells = 0;
    for (int t = 0; t < nTI; ++t) maxCells = std::max(maxCells, cumT[t + 1] - cumT[t]);
    if ((unsigned long long)maxCells * nTI > 0xFFFFFFFFull || cumT[nTI] > 0xFFFFFFFFull)
      return hfail(QSMI_ERR_UNSUPPORTED, "several TIs: cells x nTI must fit the 32-bit index (position * nTI + TI)");
    if (rq->lp || rq->recordCap > 0 || rq->noVerbatim)
      return hfail(QSMI_ERR_UNSUPPORTED, "several TIs: not with the learned proposal, the record or -nV");
    tiCat.resize(cumT[nTI] * nVar);
    for (int t = 0; t < nTI; ++t) {
      if (!rq->tis[t]) return hfail(QSMI_ERR_ARG, "null training image");
      std::memcpy(tiCat.data() + cumT[t] * nVar, rq->tis[t], (cumT[t + 1] - cumT[t]) * nVar * sizeof(float));
    }
    for (float x : tiCat)
      if (std::isnan(x)) return hfail(QSMI_ERR_UNSUPPORTED, "several TIs with NaN (cross-measurement form)");
    tiIn = tiCat.data();
  }
  // a TI with NaN runs in the cross-measurement form (src/qs.cpp:1436-1466):
  // cells with any NaN variable lose all of them
  bool cross = false;
  for (size_t c = 0; c < S * nVar && !cross; ++c) cross = std::isnan(tiIn[c]);
  std::vector<float> tiX;
  std::vector<uint8_t> centreOk;
  if (cross) {
    for (int v = 0; v < nVar; ++v)
      if (rq->varTypes[v] != 0)
        return hfail(QSMI_ERR_UNSUPPORTED,
                     "TI with NaN and a categorical variable: the reference's cross-measurement conversion reads one "
                     "value past each neighbour's expanded values (include/simulation.hpp:512-537 against "
                     "include/DataImage.hpp:681-692); continuous variables only");
    if (rq->lp || rq->recordCap > 0) return hfail(QSMI_ERR_UNSUPPORTED, "learned proposal: TI with NaN unsupported");
    tiX.assign(tiIn, tiIn + S * nVar);
    centreOk.assign(S, 1);
    for (size_t c = 0; c < S; ++c) {
      bool nan = false;
      for (int v = 0; v < nVar; ++v) nan |= std::isnan(tiX[c * nVar + v]);
      if (nan) {
        for (int v = 0; v < nVar; ++v) tiX[c * nVar + v] = NAN;
        centreOk[c] = 0;
      }
    }
  }
  const float* tiData = cross ? tiX.data() : tiIn;
  if ((rq->lp || rq->recordCap > 0) && !(D == 2 && nVar == 1 && rq->varTypes[0] == 0))
    return hfail(QSMI_ERR_UNSUPPORTED, "learned proposal and pair record: 2-D, one continuous variable");
  if (rq->recordCap > 0 && (!rq->recOffsets || !rq->recValues || !rq->recBest))
    return hfail(QSMI_ERR_ARG, "record arrays missing");
  // categories and sub-variables (src/qs.cpp:1500-1537)
  std::vector<int> ncat;
  std::vector<float> catvals;
  std::vector<int> factor(nVar, 1);
  for (int v = 0; v < nVar; ++v)
    if (rq->varTypes[v] != 0) {
      std::vector<float> cv(4096);
      int cnt = 0;
      int rc = qsmi_categories(cumT[nTI], nVar, tiIn, v, cv.data(), 4096, &cnt);
      if (rc) return rc;
      ncat.push_back(cnt);
      catvals.insert(catvals.end(), cv.begin(), cv.begin() + cnt);
      factor[v] = cnt;
    }
  int nExp = 0;
  for (int v = 0; v < nVar; ++v) nExp += factor[v];
  float alpha[256], alphaX[256];
  int convTypes[16], subVarOf[16], nSub = 0;
  {
    int rc = cross ? qsmi_coef_matrix_x(nVar, rq->varTypes, alpha, alphaX, convTypes, subVarOf, &nSub)
                   : qsmi_coef_matrix(nVar, rq->varTypes, ncat.data(), alpha, convTypes, subVarOf, &nSub);
    if (rc) return rc;
  }
  // kernel (src/qs.cpp:1010-1042) expanded per one-hot variable
  // (DataImage::offsetKernel4categories, include/DataImage.hpp:482-508)
  const int nKer = rq->nKernels >= 2 ? rq->nKernels : 1;
  if (nKer > 8) return hfail(QSMI_ERR_UNSUPPORTED, "at most 8 kernels");
  if (nKer > 1 && (!rq->kernelSizes || !rq->kernels || !rq->kernelIndexImage))
    return hfail(QSMI_ERR_ARG, "several kernels need their sizes, data and a kernel-index image");
  if (nKer > 1 && (rq->lp || rq->recordCap > 0))
    return hfail(QSMI_ERR_UNSUPPORTED, "kernel-index image: not with the learned proposal or the record");
  const float* kernel0 = nKer > 1 ? rq->kernels[0] : rq->kernel;
  const unsigned* ksize0 = nKer > 1 ? rq->kernelSizes : rq->kernelSize;
  KernelImg ker;
  ker.D = D;
  ker.nch = nExp;
  bool defaultKernel = (kernel0 == nullptr);
  size_t KC = 1;
  for (int i = 0; i < D; ++i) {
    ker.dims[i] = defaultKernel ? std::min(s[i] / 2 + 1, s[i] / 2 + 1) : ksize0[i];
    if (ker.dims[i] < 1) return hfail(QSMI_ERR_ARG, "zero kernel extent");
    KC *= ker.dims[i];
  }
  ker.data.assign(KC * nExp, 1.f);
  if (!defaultKernel) {
    int e = 0;
    for (int v = 0; v < nVar; ++v)
      for (int q = 0; q < factor[v]; ++q, ++e)
        for (size_t c = 0; c < KC; ++c) ker.data[c * nExp + e] = kernel0[c * nVar + v];
  }
  float maxw = 0.f;
  for (float w : ker.data)
    if (std::isfinite(w)) maxw = std::max(maxw, std::fabs(w));
  // offsets (src/qs.cpp:1055-1137)
  size_t cap = 1;
  for (int i = 0; i < D; ++i) cap *= (size_t)(2 * (((int)ker.dims[i] + 1) / 2) - 1);
  std::vector<int> offs(cap * D);
  int nOff = 0;
  {
    int rc = qsmi_offsets(D, ker.dims, nExp, defaultKernel ? nullptr : ker.data.data(), offs.data(), (int)cap, &nOff);
    if (rc) return rc;
  }
  // the other kernels: expanded data after kernel 0's, offsets after its
  // offsets (src/qs.cpp:1055-1137 builds one sorted list per kernel)
  std::vector<int> offStart(nKer, 0), offCount(nKer, nOff);
  unsigned kdK[8][3] = {};
  long long kerOff[8] = {};
  for (int i = 0; i < 3; ++i) kdK[0][i] = i < D ? ker.dims[i] : 1;
  offs.resize((size_t)nOff * D);
  for (int q = 1; q < nKer; ++q) {
    if (!rq->kernels[q]) return hfail(QSMI_ERR_ARG, "null kernel");
    size_t kc = 1;
    for (int i = 0; i < 3; ++i) {
      kdK[q][i] = i < D ? rq->kernelSizes[q * 3 + i] : 1;
      if (kdK[q][i] < 1) return hfail(QSMI_ERR_ARG, "zero kernel extent");
      kc *= kdK[q][i];
    }
    std::vector<float> kq(kc * nExp);
    int e = 0;
    for (int v = 0; v < nVar; ++v)
      for (int f = 0; f < factor[v]; ++f, ++e)
        for (size_t c = 0; c < kc; ++c) kq[c * nExp + e] = rq->kernels[q][c * nVar + v];
    for (float w : kq)
      if (std::isfinite(w)) maxw = std::max(maxw, std::fabs(w));
    kerOff[q] = (long long)ker.data.size();
    ker.data.insert(ker.data.end(), kq.begin(), kq.end());
    size_t capq = 1;
    for (int i = 0; i < D; ++i) capq *= (size_t)(2 * (((int)kdK[q][i] + 1) / 2) - 1);
    std::vector<int> oq(capq * D);
    int nq = 0;
    int rc = qsmi_offsets(D, kdK[q], nExp, kq.data(), oq.data(), (int)capq, &nq);
    if (rc) return rc;
    offStart[q] = (int)(offs.size() / D);
    offCount[q] = nq;
    offs.insert(offs.end(), oq.begin(), oq.begin() + (size_t)nq * D);
  }
  if (nKer > 1) nOff = (int)(offs.size() / D);
  // path and seeds (src/qs.cpp:1159-1258)
  // a cell is informed (hard data, path head) when it holds every variable;
  // a partially informed cell is simulated at its place in the path with
  // its own values as the centre neighbour, and keeps them
  // (include/simulation.hpp:324-335, 581-584, 658-666)
  std::vector<unsigned char> informed(SD, 0);
  for (size_t c = 0; c < SD; ++c) {
    int nn = 0;
    for (int v = 0; v < nVar; ++v) nn += std::isnan(rq->di[c * nVar + v]) ? 1 : 0;
    if (nn != 0 && nn != nVar && cross)
      return hfail(QSMI_ERR_UNSUPPORTED, "partially informed simulation cells with a TI holding NaN");
    informed[c] = (nn == 0);
  }
  std::vector<unsigned> path(SD);
  std::vector<float> seeds(SD);
  unsigned begin = 0;
  qsmi_path(SD, informed.data(), rq->seed, path.data(), seeds.data(), &begin);
  const size_t nPath = SD - begin;
  // device: TI spectra + scan engine
  std::vector<float> sub((size_t)nSub * S);
  if (cross)
    qsmi_ti_subimages_x(D, s, nVar, rq->varTypes, tiData, sub.data());
  else
    qsmi_ti_subimages(D, s, nVar, rq->varTypes, tiIn, ncat.data(), catvals.data(), sub.data());
  qsmi_ti* ti = nullptr;
  qsmi_device* dev = nullptr;
  int rc = qsmi_ti_create(rq->gpu, D, s, nullptr, nSub, sub.data(), &ti);
  if (rc) return rc;
  rc = qsmi_device_create(ti, alpha, rq->circularTI ? 1 : 0, &dev);
  if (!rc && cross) rc = qsmi::set_cross(dev, alphaX, centreOk.data());
  // the other TIs: spectra and a device each
  std::vector<qsmi_ti*> xti;
  std::vector<qsmi_device*> xdev;
  for (int t = 1; t < nTI && !rc; ++t) {
    std::vector<float> subT((size_t)nSub * (cumT[t + 1] - cumT[t]));
    qsmi_ti_subimages(D, sT[t].data(), nVar, rq->varTypes, tiIn + cumT[t] * nVar, ncat.data(), catvals.data(),
                      subT.data());
    qsmi_ti* tt = nullptr;
    qsmi_device* dd = nullptr;
    rc = qsmi_ti_create(rq->gpu, D, sT[t].data(), nullptr, nSub, subT.data(), &tt);
    if (!rc) xti.push_back(tt);
    if (!rc) rc = qsmi_device_create(tt, alpha, rq->circularTI ? 1 : 0, &dd);
    if (!rc) xdev.push_back(dd);
  }
  auto destroy_extra = [&]() {
    for (qsmi_device* d : xdev) qsmi_device_destroy(d);
    for (qsmi_ti* x : xti) qsmi_ti_destroy(x);
    xdev.clear();
    xti.clear();
  };
  if (rc) {
    destroy_extra();
    if (dev) qsmi_device_destroy(dev);
    qsmi_ti_destroy(ti);
    return rc;
  }
  bool anyCont = false;
  for (int v = 0; v < nVar; ++v) anyCont |= rq->varTypes[v] == 0;
  float quantum = rq->tieQuantum;
  if (quantum < 0.f) {
    quantum = 0.f;
    if (!anyCont && maxw > 0.f) {
      int e2 = (int)std::ceil(std::log2((double)nMax * (double)maxw));
      quantum = std::ldexp(1.f, e2 - 16);
    }
  }
  qsmi_set_tie_quantum(dev, quantum);
  for (qsmi_device* d : xdev) qsmi_set_tie_quantum(d, quantum);
  // output state
  float* di = rs->simulation;
  std::memcpy(di, rq->di, SD * nVar * sizeof(float));
  std::fill(rs->index, rs->index + SD, 0u);
  const auto t1 = clk::now();

  // ---- path loop -----------------------------------------------------------
  // device-resident state
  std::vector<void*> allocs;
  auto dev_upload = [&](const void* src, size_t bytes, void** dst) -> int {
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, std::max<size_t>(bytes, 4));
    if (e != hipSuccess) return hfail(QSMI_ERR_NOMEM, std::string("path loop state: ") + hipGetErrorString(e));
    allocs.push_back(p);
    if (src && bytes) {
      e = hipMemcpy(p, src, bytes, hipMemcpyHostToDevice);
      if (e != hipSuccess) return hfail(QSMI_ERR_HIP, std::string("path loop upload: ") + hipGetErrorString(e));
    }
    *dst = p;
    return QSMI_OK;
  };
  // path index of every cell, -1 for hard data (the posterior order of
  // src/qs.cpp:1615-1641: hard data first, then the path)
  const int n = nMax;
  std::vector<int> postS(SD, -1);
  for (size_t i = 0; i < nPath; ++i) postS[path[begin + i]] = (int)i;
  // neighbour sets of every node (simulation.hpp:447-511), centre first, on
  // the device; their path indices come back for the dependency levels
  void *dCells = nullptr, *dPostS = nullptr, *dOffs = nullptr, *dNb = nullptr, *dPj = nullptr, *dCnt = nullptr;
  if (!rc) rc = dev_upload(path.data() + begin, nPath * sizeof(unsigned), &dCells);
  if (!rc) rc = dev_upload(postS.data(), SD * sizeof(int), &dPostS);
  if (!rc) rc = dev_upload(offs.data(), (size_t)nOff * D * sizeof(int), &dOffs);
  if (!rc) rc = dev_upload(nullptr, nPath * (size_t)n * sizeof(int), &dNb);
  if (!rc) rc = dev_upload(nullptr, nPath * (size_t)n * sizeof(int), &dPj);
  if (!rc) rc = dev_upload(nullptr, nPath * sizeof(int), &dCnt);
  void *dNNode = nullptr, *dPathK = nullptr, *dNodeKer = nullptr, *dOffStart = nullptr, *dOffCount = nullptr;
  void* dNodeTi = nullptr;
  if (rq->tiIndexImage) {
    if (nTI < 2) rc = hfail(QSMI_ERR_ARG, "a TI-index image needs several TIs");
    std::vector<int> nt(nPath);
    for (size_t i = 0; i < nPath && !rc; ++i) {
      const float v = rq->tiIndexImage[path[begin + i]];
      if (!(v >= 0.f && v < (float)nTI)) rc = hfail(QSMI_ERR_ARG, "TI-index image: value outside [0, nTI)");
      nt[i] = rc ? 0 : (int)v;
    }
    if (!rc) rc = dev_upload(nt.data(), nPath * sizeof(int), &dNodeTi);
  }
  if (nKer > 1) {
    std::vector<int> nk(nPath);
    for (size_t i = 0; i < nPath; ++i) {
      const float v = rq->kernelIndexImage[path[begin + i]];
      if (!(v >= 0.f && v < (float)nKer)) rc = hfail(QSMI_ERR_ARG, "kernel-index image: value outside [0, nKernels)");
      nk[i] = rc ? 0 : (int)v;
    }
    if (!rc) rc = dev_upload(nk.data(), nPath * sizeof(int), &dNodeKer);
    if (!rc) rc = dev_upload(offStart.data(), nKer * sizeof(int), &dOffStart);
    if (!rc) rc = dev_upload(offCount.data(), nKer * sizeof(int), &dOffCount);
  }
  if (rq->nImage) {
    std::vector<int> nNode(nPath);
    for (size_t i = 0; i < nPath; ++i) {
      const float v = rq->nImage[path[begin + i]];
      nNode[i] = (std::isfinite(v) && v > 0.f && v < 1e6f) ? (int)v : 0;
    }
    if (!rc) rc = dev_upload(nNode.data(), nPath * sizeof(int), &dNNode);
  }
  if (rq->kImage) {
    std::vector<float> kNode(nPath);
    for (size_t i = 0; i < nPath; ++i) kNode[i] = rq->kImage[path[begin + i]];
    if (!rc) rc = dev_upload(kNode.data(), nPath * sizeof(float), &dPathK);
  }
  if (!rc)
    rc = qsmi::nb_search_device(D, sd, nPath, (const unsigned*)dCells, (const int*)dPostS, (const int*)dOffs, nOff, n,
                                (int*)dNb, (int*)dPj, (int*)dCnt, dev->stream, (const int*)dNNode,
                                (const int*)dNodeKer, (const int*)dOffStart, (const int*)dOffCount,
                                rq->circularSim ? 1 : 0);
  // learned proposal: which nodes reach the stencil coverage it needs
  const float lpMinCover = rq->lpMinCover > 0.f ? rq->lpMinCover : 0.5f;
  std::vector<unsigned char> lpOk;
  void* dOk = nullptr;
  if (!rc && rq->lp) {
    rc = dev_upload(nullptr, nPath, &dOk);
    if (!rc)
      rc = qsmi::lp_cover_all(rq->lp, (int)nPath, n, (const int*)dNb, (const int*)dCnt, (const int*)dOffs, lpMinCover,
                              (unsigned char*)dOk, dev->stream);
  }
  // dependency levels (1 + the deepest level among a node's path neighbours)
  // on the device; only the levels come back
  std::vector<int> level(nPath, 0);
  void *dLevel = nullptr, *dLvErr = nullptr;
  if (!rc) rc = dev_upload(nullptr, nPath * sizeof(int), &dLevel);
  if (!rc) rc = dev_upload(nullptr, sizeof(int), &dLvErr);
  if (!rc) rc = qsmi::levels_device(nPath, n, (const int*)dPj, (const int*)dCnt, (int*)dLevel, (int*)dLvErr,
                                    dev->stream);
  int lvErr = 0;
  if (!rc) {
    hipError_t e = hipStreamSynchronize(dev->stream);
    if (e == hipSuccess && nPath) e = hipMemcpy(level.data(), dLevel, nPath * sizeof(int), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(&lvErr, dLvErr, sizeof(int), hipMemcpyDeviceToHost);
    if (e == hipSuccess && dOk) {
      lpOk.resize(nPath);
      e = hipMemcpy(lpOk.data(), dOk, nPath, hipMemcpyDeviceToHost);
    }
    if (e != hipSuccess) rc = hfail(QSMI_ERR_HIP, std::string("neighbour search: ") + hipGetErrorString(e));
  }
  if (!rc && lvErr) {
    // a device wait hit its bound: the levels again on the host, in path order
    std::vector<int> nbPj(nPath * (size_t)n), nbCnt(nPath, 0);
    hipError_t e = hipMemcpy(nbCnt.data(), dCnt, nPath * sizeof(int), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(nbPj.data(), dPj, nbPj.size() * sizeof(int), hipMemcpyDeviceToHost);
    if (e != hipSuccess) rc = hfail(QSMI_ERR_HIP, std::string("neighbour search: ") + hipGetErrorString(e));
    for (size_t i = 0; i < nPath && !rc; ++i) {
      int lv = 0;
      const int* pj = nbPj.data() + i * (size_t)n;
      for (int j = 0; j < nbCnt[i]; ++j)
        if (pj[j] >= 0 && (size_t)pj[j] < i) lv = std::max(lv, level[pj[j]] + 1);
      level[i] = lv;
    }
  }
  const auto tNb = clk::now();
  int maxLevel = 0;
  for (size_t i = 0; i < nPath; ++i) maxLevel = std::max(maxLevel, level[i]);
  // nodes ordered by (level, path order)
  std::vector<size_t> lvStart(maxLevel + 2, 0);
  for (size_t i = 0; i < nPath; ++i) lvStart[level[i] + 1]++;
  for (int l = 0; l <= maxLevel; ++l) lvStart[l + 1] += lvStart[l];
  std::vector<unsigned> order(nPath);
  {
    std::vector<size_t> fillp(lvStart.begin(), lvStart.end() - 1);
    for (size_t i = 0; i < nPath; ++i) order[fillp[level[i]]++] = (unsigned)i;
  }
  // 512 nodes per wave measured best on the 1000x1000 path (fewer launches;
  // 768 no better); the learned proposal keeps its queries in LDS: 256
  const int waveSize = rq->waveSize > 0 ? rq->waveSize : (rq->lp ? 256 : 512);
  const auto tLv = clk::now();
  // waves: each level cut into the fewest waves of at most waveSize nodes,
  // of equal sizes (the waves of a level run side by side on two streams)
  qsmi::ResidentLoop R;
  R.waveStart.push_back(0);
  for (int l = 0; l <= maxLevel; ++l) {
    const size_t n = lvStart[l + 1] - lvStart[l];
    const size_t nw = (n + (size_t)waveSize - 1) / (size_t)waveSize;
    for (size_t q = 1; q <= nw; ++q) {
      R.waveStart.push_back((int)(lvStart[l] + n * q / nw));
      R.waveLevel.push_back(l);
    }
  }
  for (int d = 0; d < 3; ++d) {
    R.sd[d] = (int)sd[d];
    R.kd[d] = (int)ker.dims[d];
  }
  R.nVar = nVar;
  R.nExp = nExp;
  R.n = n;
  R.k = rq->k;
  R.kMax = kMax;
  R.nKer = nKer;
  for (int q = 0; q < nKer; ++q) {
    for (int i = 0; i < 3; ++i) R.kdK[q][i] = (int)kdK[q][i];
    R.kerOff[q] = kerOff[q];
  }
  for (int v = 0; v < nVar && v < 8; ++v) R.types[v] = rq->varTypes[v];
  {
    int e = 0, ci = 0, off = 0;
    for (int v = 0; v < nVar; ++v) {
      if (rq->varTypes[v] == 0) {
        R.expVar[e] = v;
        R.expCat[e] = -1;
        e++;
      } else {
        for (int q = 0; q < ncat[ci]; ++q, ++e) {
          R.expVar[e] = v;
          R.expCat[e] = off + q;
        }
        off += ncat[ci];
        ci++;
      }
    }
  }
  for (int B = 0; B < nSub; ++B) {
    R.conv[B] = convTypes[B];
    R.subOf[B] = subVarOf[B];
  }
  R.stageTiming = std::max(0, rq->stageTiming);
  R.progressSec = std::max(0, rq->progressSeconds);
  R.noVerbatim = rq->noVerbatim != 0;
  R.circSim = rq->circularSim != 0;
  std::vector<int> orderI(order.begin(), order.end());
  void *dOrder = nullptr, *dSeeds = nullptr, *dKer = nullptr, *dCat = nullptr, *dDi = nullptr, *dIdx = nullptr,
       *dTi = nullptr;
  std::vector<float> catv = catvals.empty() ? std::vector<float>(1, 0.f) : catvals;
  if (!rc) rc = dev_upload(orderI.data(), nPath * sizeof(int), &dOrder);
  if (!rc) rc = dev_upload(seeds.data() + begin, nPath * sizeof(float), &dSeeds);
  if (!rc) rc = dev_upload(ker.data.data(), ker.data.size() * sizeof(float), &dKer);
  if (!rc) rc = dev_upload(catv.data(), catv.size() * sizeof(float), &dCat);
  if (!rc) rc = dev_upload(di, SD * nVar * sizeof(float), &dDi);
  if (!rc) rc = dev_upload(rs->index, SD * sizeof(unsigned), &dIdx);
  if (!rc) rc = dev_upload(tiData, cumT[nTI] * nVar * sizeof(float), &dTi);
  const size_t recCap = rq->recordCap > 0 ? std::min<size_t>((size_t)rq->recordCap, nPath) : 0;
  void *dRecOff = nullptr, *dRecVal = nullptr, *dRecBest = nullptr;
  if (!rc && recCap) {
    rc = dev_upload(nullptr, recCap * 80 * 2 * sizeof(int), &dRecOff);
    if (!rc) rc = dev_upload(nullptr, recCap * 80 * sizeof(float), &dRecVal);
    if (!rc) rc = dev_upload(nullptr, recCap * sizeof(unsigned), &dRecBest);
  }
  R.recCap = (int)recCap;
  R.dRecOff = (int*)dRecOff;
  R.dRecVal = (float*)dRecVal;
  R.dRecBest = (uint32_t*)dRecBest;
  R.lp = rq->lp;
  R.lpShortlist = rq->lpShortlist;
  R.lpEvery = rq->lpFallbackEvery;
  R.lpMse = rq->lpFallbackMse;
  R.lpAudit = rq->lpAudit;
  R.lpMinCover = lpMinCover;
  if (rq->lp && !lpOk.empty()) {
    // a wave runs the proposal only if one of its nodes can be proposed
    // (enough stencil coverage and not an every-m-th exact node)
    const size_t nW = R.waveStart.size() - 1;
    R.waveLp.assign(nW, 0);
    for (size_t w = 0; w < nW; ++w)
      for (int q = R.waveStart[w]; q < R.waveStart[w + 1] && !R.waveLp[w]; ++q) {
        const unsigned i = order[q];
        const bool periodic = R.lpEvery > 0 && (i % (unsigned)R.lpEvery) == 0;
        if (lpOk[i] && !periodic) R.waveLp[w] = 1;
      }
  }
  const auto tUp = clk::now();
  double stageMs[4] = {0, 0, 0, 0};
  long long bandStats[4] = {0, 0, 0, 0};
  long long scanned = 0;
  int waves = 0;
  if (!rc) {
    R.dOrder = (const int*)dOrder;
    R.dPathCells = (const unsigned*)dCells;
    R.dPathSeeds = (const float*)dSeeds;
    R.dNbIdx = (const int*)dNb;
    R.dNbCnt = (const int*)dCnt;
    R.dOffs = (const int*)dOffs;
    R.dKernel = (const float*)dKer;
    R.dCatvals = (const float*)dCat;
    R.dDi = (float*)dDi;
    R.dIndex = (uint32_t*)dIdx;
    R.dTi = (const float*)dTi;
    R.extraDevs = xdev;
    for (int t = 0; t <= nTI; ++t) R.tiCum[t] = (unsigned)cumT[t];
    R.dPathK = (const float*)dPathK;
    R.dNodeKer = (const int*)dNodeKer;
    R.dNodeTi = (const int*)dNodeTi;
    rc = qsmi::run_resident_waves(dev, R, stageMs, &scanned, &waves, bandStats);
  }
  if (!rc) {
    hipError_t e = hipMemcpy(di, dDi, SD * nVar * sizeof(float), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(rs->index, dIdx, SD * sizeof(unsigned), hipMemcpyDeviceToHost);
    if (e == hipSuccess && recCap) {
      e = hipMemcpy(rq->recOffsets, dRecOff, recCap * 80 * 2 * sizeof(int), hipMemcpyDeviceToHost);
      if (e == hipSuccess) e = hipMemcpy(rq->recValues, dRecVal, recCap * 80 * sizeof(float), hipMemcpyDeviceToHost);
      if (e == hipSuccess) e = hipMemcpy(rq->recBest, dRecBest, recCap * sizeof(unsigned), hipMemcpyDeviceToHost);
    }
    if (e != hipSuccess) rc = hfail(QSMI_ERR_HIP, std::string("path loop download: ") + hipGetErrorString(e));
  }
  long long lpc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (!rc && rq->lp) rc = qsmi::lp_read_counters(rq->lp, lpc);
  for (void* p : allocs) (void)hipFree(p);
  const auto tLoop = clk::now();
  double scanMs = stageMs[0] + stageMs[1] + stageMs[2] + stageMs[3];
  rs->phaseMs[0] = std::chrono::duration<double, std::milli>(tNb - t1).count();
  rs->phaseMs[1] = std::chrono::duration<double, std::milli>(tLv - tNb).count();
  rs->phaseMs[2] = std::chrono::duration<double, std::milli>(tUp - tLv).count();
  rs->phaseMs[3] = std::chrono::duration<double, std::milli>(tLoop - tUp).count();
  const auto t2 = clk::now();
  destroy_extra();
  qsmi_device_destroy(dev);
  qsmi_ti_destroy(ti);
  if (rc) {
    return rc;
  }
  rs->setupMs = std::chrono::duration<double, std::milli>(t1 - t0).count();
  rs->pathLoopMs = std::chrono::duration<double, std::milli>(t2 - t1).count();
  rs->scanMs = scanMs;
  for (int q = 0; q < 4; ++q) rs->stageMs[q] = stageMs[q];
  rs->scannedNodes = scanned;
  rs->nodes = (long long)nPath;
  rs->waves = waves;
  rs->lpProposed = lpc[0];
  rs->lpFallbacks = lpc[1] + R.lpSkipped;
  rs->lpOverflows = lpc[2];
  rs->lpShortlistSum = lpc[3];
  rs->lpAuditNodes = lpc[4];
  rs->lpAuditHits = lpc[5];
  rs->lpAuditBestHits = lpc[6];
  rs->exactNodes = bandStats[0];
  rs->bandPositions = bandStats[1];
  rs->secondPassBlocks = bandStats[2];
  rs->fullRescoreNodes = bandStats[3];
  return QSMI_OK;
}

}  // extern "C"
