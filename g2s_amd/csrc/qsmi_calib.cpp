// auto_qs calibration above the qsmi C-ABI: auto_qs's set-up
// (src/auto_qs.cpp:449-492 offsets, :561-582 categories, :640 the sampling
// module) and the Monte Carlo loop of calibration()
// (include/calibration.hpp:30-253), its QuantileSamplingModule::distribution
// calls (include/quantileSamplingModule.hpp:128-169) batched into GPU waves.
//
// Per iteration every setting is visited in order: the skip test on the
// state at the iteration's start, then the sample's draws (TI, cell,
// neighbourhood, noise swaps, distribution seed) from one std::mt19937. The
// draws never depend on scan results, so the whole iteration's samples are
// drawn first, scanned in waves (one wave per TI device), and accumulated in
// setting order with the reference's running-best update.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/qsmi.h"
#include "qsmi_internal.hpp"

extern "C" int qsmi_internal_fail(int code, const char* msg);

namespace {

int cfail(int code, const std::string& m) { return qsmi_internal_fail(code, m.c_str()); }

using Off = std::array<int, 3>;

// src/auto_qs.cpp:449-474: every offset of kernel 0's box, dimension by
// dimension (0, +1, -1, +2, -2, ... per dimension, earlier dimensions
// fastest), then :476-492 sorted by descending weight -distance^2 to the
// kernel centre with std::sort (the same comparisons on the same sequence
// give the reference's order of equal distances).
std::vector<Off> path_positions(int D, const unsigned* kdims) {
  std::vector<Off> pp(1, Off{0, 0, 0});
  for (int i = 0; i < D; ++i) {
    const size_t orig = pp.size();
    const int size = ((int)kdims[i] + 1) / 2;
    pp.resize(orig * (size_t)(2 * size - 1));
    for (int j = 1; j < size; ++j) {
      std::copy(pp.begin(), pp.begin() + orig, pp.begin() + orig * (2 * j - 1));
      std::copy(pp.begin(), pp.begin() + orig, pp.begin() + orig * (2 * j));
      for (size_t k = orig * (2 * j - 1); k < orig * (2 * j); ++k) pp[k][i] = j;
      for (size_t k = orig * (2 * j); k < orig * (2 * j + 1); ++k) pp[k][i] = -j;
    }
  }
  std::sort(pp.begin(), pp.end(), [D](const Off& a, const Off& b) {
    float wa = 0.f, wb = 0.f;
    for (int i = 0; i < D; ++i) {
      wa += (float)(a[i] * a[i]);
      wb += (float)(b[i] * b[i]);
    }
    return -wa > -wb;
  });
  return pp;
}

// DataImage::indexWithDelta (include/DataImage.hpp:343-376) on a grid
bool index_with_delta(int D, const unsigned* dims, unsigned position, const int* d, unsigned* loc) {
  bool ok = true;
  unsigned v[3] = {0, 0, 0};
  for (int i = 0; i < D; ++i) {
    int c = (int)(position % dims[i]) + d[i];
    ok &= (c >= 0 && c < (int)dims[i]);
    c = ((c % (int)dims[i]) + (int)dims[i]) % (int)dims[i];
    v[i] = (unsigned)c;
    position /= dims[i];
  }
  unsigned fin = 0;
  for (int i = D - 1; i >= 0; --i) fin = fin * dims[i] + v[i];
  *loc = fin;
  return ok;
}

// std::uniform_real_distribution<float>(0, 1) over std::mt19937 as
// libstdc++ computes it (generate_canonical<float, 24> with one 32-bit
// word: float(word) / 2^32, values rounding to 1 replaced by the largest
// float below 1), without its per-call logarithms
inline float u01(std::mt19937& g) {
  const float r = (float)g() * (1.0f / 4294967296.0f);
  return r >= 1.f ? 0.99999994f : r;
}

struct Sample {
  int setting;
  int tiIndex;
  unsigned cell;
  float seed;
  std::vector<Off> off;
  std::vector<float> val;  // expanded (one-hot) values, nExp per neighbour
};

struct Tiv {
  unsigned s[3] = {1, 1, 1};
  size_t cells = 0;
  const float* data = nullptr;
  qsmi_ti* ti = nullptr;
  qsmi_device* dev = nullptr;
};

}  // namespace

extern "C" int qsmi_calibration(const qsmi_calib_request* rq, qsmi_calib_result* rs) {
  using clk = std::chrono::steady_clock;
  if (!rq || !rs || !rs->meanError || !rs->devError || !rs->samples) return cfail(QSMI_ERR_ARG, "null argument");
  const int D = rq->ndim;
  if (D < 1 || D > 3) return cfail(QSMI_ERR_UNSUPPORTED, "ndim must be 1, 2 or 3");
  if (rq->nTI < 1 || !rq->tiSize || !rq->ti) return cfail(QSMI_ERR_ARG, "no training image");
  if (rq->nVar < 1 || !rq->varTypes) return cfail(QSMI_ERR_ARG, "invalid variables");
  if (rq->nKernels < 1 || !rq->kernelSize || !rq->kernels) return cfail(QSMI_ERR_ARG, "no kernel");
  if (rq->nConfigs < 1 || rq->nbWidth < 1 || !rq->nbList) return cfail(QSMI_ERR_ARG, "no neighbour configuration");
  if (rq->nDensity < 1 || !rq->density) return cfail(QSMI_ERR_ARG, "no density");
  if (rq->maxK < 1 || rq->maxK > 16) return cfail(QSMI_ERR_UNSUPPORTED, "maxK must be in [1, 16]");
  if (!(rq->power > 0.f)) return cfail(QSMI_ERR_ARG, "power must be > 0");
  if (rq->logCap > 0 && (!rq->logSetting || !rq->logCell || !rq->logPos)) return cfail(QSMI_ERR_ARG, "log arrays missing");
  const int nVar = rq->nVar, nTI = rq->nTI, K = rq->maxK;
  const auto t0 = clk::now();

  std::vector<Tiv> tis(nTI);
  for (int t = 0; t < nTI; ++t) {
    if (!rq->ti[t]) return cfail(QSMI_ERR_ARG, "null training image");
    tis[t].cells = 1;
    for (int i = 0; i < D; ++i) {
      tis[t].s[i] = rq->tiSize[t * 3 + i];
      if (tis[t].s[i] < 1) return cfail(QSMI_ERR_ARG, "zero TI extent");
      tis[t].cells *= tis[t].s[i];
    }
    tis[t].data = rq->ti[t];
    for (size_t c = 0; c < tis[t].cells * nVar; ++c)
      if (std::isnan(tis[t].data[c]))
        return cfail(QSMI_ERR_UNSUPPORTED, "calibration on a TI with NaN (cross-measurement form) is not supported");
  }
  // categories over all TIs (src/auto_qs.cpp:561-582), one-hot expansion
  std::vector<std::vector<float>> catv(nVar);
  std::vector<int> factor(nVar, 1), ncat;
  std::vector<float> catflat;
  for (int v = 0; v < nVar; ++v) {
    if (rq->varTypes[v] == 0) continue;
    for (int t = 0; t < nTI; ++t)
      for (size_t c = 0; c < tis[t].cells; ++c) {
        const float x = tis[t].data[c * nVar + v];
        if (std::find(catv[v].begin(), catv[v].end(), x) == catv[v].end()) catv[v].push_back(x);
      }
    factor[v] = (int)catv[v].size();
    ncat.push_back(factor[v]);
    catflat.insert(catflat.end(), catv[v].begin(), catv[v].end());
  }
  int nExp = 0;
  for (int v = 0; v < nVar; ++v) nExp += factor[v];
  float alpha[256];
  int convTypes[16], subVarOf[16], nSub = 0;
  {
    int rc = qsmi_coef_matrix(nVar, rq->varTypes, ncat.data(), alpha, convTypes, subVarOf, &nSub);
    if (rc) return rc;
  }
  // kernels expanded per one-hot variable (DataImage::offsetKernel4categories,
  // include/DataImage.hpp:492-518)
  struct Ker {
    unsigned dims[3] = {1, 1, 1};
    std::vector<float> w;  // cells x nExp
  };
  std::vector<Ker> kers(rq->nKernels);
  float maxw = 0.f;
  for (int q = 0; q < rq->nKernels; ++q) {
    if (!rq->kernels[q]) return cfail(QSMI_ERR_ARG, "null kernel");
    size_t kc = 1;
    for (int i = 0; i < D; ++i) {
      kers[q].dims[i] = rq->kernelSize[q * 3 + i];
      if (kers[q].dims[i] < 1) return cfail(QSMI_ERR_ARG, "zero kernel extent");
      kc *= kers[q].dims[i];
    }
    kers[q].w.resize(kc * nExp);
    int e = 0;
    for (int v = 0; v < nVar; ++v)
      for (int f = 0; f < factor[v]; ++f, ++e)
        for (size_t c = 0; c < kc; ++c) {
          const float w = rq->kernels[q][c * nVar + v];
          kers[q].w[c * nExp + e] = w;
          if (std::isfinite(w)) maxw = std::max(maxw, std::fabs(w));
        }
  }
  const std::vector<Off> pp = path_positions(D, kers[0].dims);
  // devices, one per TI
  auto cleanup = [&]() {
    for (auto& t : tis) {
      if (t.dev) qsmi_device_destroy(t.dev);
      if (t.ti) qsmi_ti_destroy(t.ti);
      t.dev = nullptr;
      t.ti = nullptr;
    }
  };
  bool anyCont = false;
  unsigned maxNb = 1;
  for (int v = 0; v < nVar; ++v) anyCont |= rq->varTypes[v] == 0;
  for (int c = 0; c < rq->nConfigs * rq->nbWidth; ++c) maxNb = std::max(maxNb, rq->nbList[c]);
  const float quantum = anyCont || maxw <= 0.f
                            ? 0.f
                            : std::ldexp(1.f, (int)std::ceil(std::log2((double)maxNb * nVar * (double)maxw)) - 16);
  for (int t = 0; t < nTI; ++t) {
    std::vector<float> sub((size_t)nSub * tis[t].cells);
    int rc = qsmi_ti_subimages(D, tis[t].s, nVar, rq->varTypes, tis[t].data, ncat.data(), catflat.data(), sub.data());
    if (!rc) rc = qsmi_ti_create(rq->gpu, D, tis[t].s, nullptr, nSub, sub.data(), &tis[t].ti);
    if (!rc) rc = qsmi_device_create(tis[t].ti, alpha, 0, &tis[t].dev);
    if (!rc) rc = qsmi_set_tie_quantum(tis[t].dev, quantum);
    if (rc) {
      cleanup();
      return rc;
    }
  }

  const int nCfg = rq->nConfigs, nKer = rq->nKernels, nDen = rq->nDensity;
  const size_t settings = (size_t)nCfg * nKer * nDen;
  std::vector<double> cumErr(settings * K, 0.0), cumSq(settings * K, 0.0);
  std::vector<unsigned> nSamp(settings * K, 0u);
  std::vector<float> best(nDen, INFINITY), devBest(nDen, 0.f);
  // include/calibration.hpp:33: the exclusion radius, TI 0's cells x 0.005
  // to the power 1/D
  const float radius = (float)std::pow((double)(tis[0].cells) * 0.005, 1.f / (float)D);
  std::mt19937 gen(rq->seed);
  const int waveMax = rq->waveSize > 0 ? std::min(rq->waveSize, 1024) : 256;
  const double maxSec = rq->maxSeconds > 0 ? rq->maxSeconds : INFINITY;
  const float nbSigmas = 1.f / rq->power;
  long long logged = 0, dists = 0;
  int waves = 0, iter = 0;
  double scanMs = 0.0, drawMs = 0.0;
  int rc = QSMI_OK;
  std::vector<Sample> smp;
  // per-wave host arrays
  std::vector<int> nbStart, offs;
  std::vector<float> uvals, d0, seeds, exR, oErr, allErr;
  std::vector<unsigned> exP, oPos, allPos;
  std::vector<float> ubuf;
  std::vector<unsigned> kidx;
  std::vector<char> kin;
  std::vector<unsigned> perVar(nVar);
  std::vector<float> raw;
  double stage[4] = {0, 0, 0, 0};
  for (auto& T : tis) {
    if (!T.dev->wb) T.dev->wb = new WaveBufs();
    if (rq->stageTiming) qsmi_set_stage_timing(T.dev, 1);
  }
  const auto t1 = clk::now();

  // Draws of one setting's sample (include/calibration.hpp:101-203);
  // appends to smp unless the setting is skipped or the draw fails.
  size_t si = 0;
  auto draw_one = [&]() {
    const size_t s0 = si++;
    const int cfg = (int)(s0 % nCfg);
    const int den = (int)(s0 / ((size_t)nCfg * nKer));
    // :101-107: skip a setting clearly worse than its density's best
    const unsigned n0 = nSamp[s0 * K];
    const float eb = (float)(cumErr[s0 * K] / n0);
    const float db = (float)std::sqrt(cumSq[s0 * K] / n0 - (double)(eb * eb));
    if (n0 > rq->minIter && (eb - nbSigmas * db) > (best[den] + nbSigmas * devBest[den])) return;
    const float density = rq->density[den];
    const unsigned* nbc = rq->nbList + (size_t)cfg * rq->nbWidth;
    const int tiIndex = (int)std::floor(u01(gen) * (float)nTI);
    const Tiv& T = tis[tiIndex];
    // float arithmetic as the reference's U * dataSize() / nbVariable
    const unsigned cell = (unsigned)std::floor(u01(gen) * (float)(T.cells * nVar) / (float)nVar);
    // :115-154: offsets in order, each kept with probability `density`
    // while inside the TI, until every variable has its count
    Sample sm;
    sm.setting = (int)s0;
    sm.tiIndex = tiIndex;
    sm.cell = cell;
    std::fill(perVar.begin(), perVar.end(), 0u);
    raw.clear();
    size_t ps = 0;
    while ((rq->nbWidth > 1 || sm.off.size() < nbc[0]) && ps < pp.size()) {
      unsigned loc;
      if (u01(gen) < density && index_with_delta(D, T.s, cell, pp[ps].data(), &loc)) {
        unsigned cpt = 0;
        for (int v = 0; v < nVar; ++v) {
          if (perVar[v] < nbc[v % rq->nbWidth]) {
            raw.push_back(T.data[(size_t)loc * nVar + v]);
            cpt++;
            perVar[v]++;
          } else {
            raw.push_back(std::nanf("0"));
          }
        }
        sm.off.push_back(pp[ps]);
        if (cpt == 0) break;
      }
      ps++;
    }
    if (ps >= pp.size()) return;
    // :158-180: one value per expanded variable (a category's indicator)
    const size_t nb = sm.off.size();
    sm.val.resize(nb * nExp);
    for (size_t j = 0; j < nb; ++j) {
      int e = 0;
      for (int v = 0; v < nVar; ++v) {
        const float x = raw[j * nVar + v];
        if (rq->varTypes[v] == 0) {
          sm.val[j * nExp + e++] = x;
        } else {
          for (float c : catv[v]) sm.val[j * nExp + e++] = (x == c) ? 1.f : 0.f;
        }
      }
    }
    // :184-203: noise, offsets swapped between neighbours
    if (nb > 1 && rq->noise > 0.f) {
      float fraction = rq->noise / (float)nb / 2.f;
      if (rq->noise < 1.0f) fraction = rq->noise / 2.f;
      for (size_t j = 0; j < nb; ++j) {
        if (u01(gen) < fraction) {
          const size_t a = (size_t)std::floor(u01(gen) * (float)nb);
          const size_t b = (size_t)std::floor(u01(gen) * (float)nb);
          std::iter_swap(sm.off.begin() + a, sm.off.begin() + b);
        }
      }
    }
    if (nb > 1) {
      sm.seed = u01(gen);
      smp.push_back(std::move(sm));
    }
  };
  // samples until `target` are drawn or the iteration's settings run out
  auto draw_until = [&](size_t target) {
    const auto td = clk::now();
    while (smp.size() < target && si < settings) draw_one();
    drawMs += std::chrono::duration<double, std::milli>(clk::now() - td).count();
  };
  // host conversion of samples [w0, w0 + nn) for TI t, staged and launched
  auto launch = [&](size_t w0, int nn, int t) -> int {
    nbStart.assign(nn + 1, 0);
    offs.clear();
    uvals.clear();
    d0.assign(nn, 0.f);
    seeds.assign(nn, 0.f);
    exP.assign(nn, 0xFFFFFFFFu);
    exR.assign(nn, 0.f);
    for (int q = 0; q < nn; ++q) {
      const Sample& sm = smp[w0 + q];
      const Ker& kq = kers[(sm.setting / nCfg) % nKer];
      const int nb = (int)sm.off.size();
      nbStart[q + 1] = nbStart[q] + nb;
      // include/quantileSamplingModule.hpp:368-457 with this sample's
      // kernel: u = w, w x, w x^2 per sub-variable (0 outside the kernel or
      // for NaN), delta0 = -sum w x^2 over the continuous variables at
      // their original index
      unsigned centre = 0;
      for (int i = D - 1; i >= 0; --i) centre = centre * kq.dims[i] + kq.dims[i] / 2;
      ubuf.assign((size_t)nb * nSub, 0.f);
      kidx.resize(nb);
      kin.resize(nb);
      for (int j = 0; j < nb; ++j) {
        kin[j] = index_with_delta(D, kq.dims, centre, sm.off[j].data(), &kidx[j]);
        for (int i = 0; i < D; ++i) offs.push_back(sm.off[j][i]);
        for (int B = 0; B < nSub; ++B) {
          const int e = subVarOf[B];
          const float x = sm.val[(size_t)j * nExp + e];
          const float w = kq.w[(size_t)kidx[j] * nExp + (e % nExp)];
          float r = 0.f;
          if (kin[j] && !std::isnan(x)) r = convTypes[B] == 0 ? w * 1.f : convTypes[B] == 2 ? w * x * x : w * x;
          ubuf[(size_t)j * nSub + B] = r;
        }
      }
      uvals.insert(uvals.end(), ubuf.begin(), ubuf.end());
      float delta0 = 0.f;
      for (int v = 0; v < nVar; ++v) {
        if (rq->varTypes[v] != 0) continue;
        for (int j = 0; j < nb; ++j) {
          const float x = sm.val[(size_t)j * nExp + v];
          if (kin[j] && !std::isnan(x)) delta0 += -1.f * (kq.w[(size_t)kidx[j] * nExp + (v % nExp)] * x * x);
        }
      }
      d0[q] = delta0;
      seeds[q] = sm.seed;
      if (sm.tiIndex == t) {
        exP[q] = sm.cell;
        exR[q] = radius;
      }
    }
    qsmi_device* dv = tis[t].dev;
    int r = qsmi::stage_wave(dv, dv->wb, nn, nbStart.data(), offs.data(), uvals.data(), d0.data(), seeds.data(),
                             (float)K, exP.data(), exR.data());
    if (!r) r = qsmi::run_wave(dv, dv->wb);
    return r;
  };
  auto collect = [&](size_t w0, int nn, int t) -> int {
    oErr.resize((size_t)nn * K);
    oPos.resize((size_t)nn * K);
    qsmi_device* dv = tis[t].dev;
    int r = qsmi::read_wave(dv, dv->wb, oErr.data(), oPos.data(), nullptr);
    if (r) return r;
    if (rq->stageTiming)
      for (int i = 0; i < 4; ++i) stage[i] += dv->stageMs[i];
    allErr.resize(smp.size() * (size_t)nTI * K);
    allPos.resize(smp.size() * (size_t)nTI * K);
    for (int q = 0; q < nn; ++q)
      for (int i = 0; i < K; ++i) {
        allErr[((w0 + q) * nTI + t) * K + i] = oErr[(size_t)q * K + i];
        allPos[((w0 + q) * nTI + t) * K + i] = oPos[(size_t)q * K + i];
      }
    waves++;
    return QSMI_OK;
  };

  for (unsigned it = 0; it < rq->maxIter && !rc; ++it) {
    smp.clear();
    si = 0;
    // draws of the next wave's samples run on the host while the GPU scans
    // the current wave (one TI); with several TIs each wave visits them in turn
    draw_until((size_t)waveMax);
    const auto ts = clk::now();
    double hostInScan = 0.0;
    for (size_t w0 = 0; w0 < smp.size() && !rc; ) {
      const int nn = (int)std::min<size_t>(waveMax, smp.size() - w0);
      for (int t = 0; t < nTI && !rc; ++t) {
        rc = launch(w0, nn, t);
        if (!rc && t == nTI - 1) {
          const double before = drawMs;
          draw_until(w0 + nn + (size_t)waveMax);
          hostInScan += drawMs - before;
        }
        if (!rc) rc = collect(w0, nn, t);
      }
      w0 += nn;
    }
    scanMs += std::chrono::duration<double, std::milli>(clk::now() - ts).count() - hostInScan;
    if (rc) break;
    const size_t ns = smp.size();
    // accumulate in setting order (include/calibration.hpp:206-246)
    for (size_t q = 0; q < ns; ++q) {
      const Sample& sm = smp[q];
      const size_t sx = (size_t)sm.setting;
      const int den = (int)(sx / ((size_t)nCfg * nKer));
      const unsigned n0 = nSamp[sx * K];
      const float errBefore = (float)(cumErr[sx * K] / n0);
      // merge the TIs' lists (each best first): descending error (its tie
      // key when a tie quantum is set), earlier TI first on ties
      std::vector<std::pair<float, std::pair<int, unsigned>>> m;
      for (int t = 0; t < nTI; ++t)
        for (int i = 0; i < K; ++i) {
          const unsigned p = allPos[(q * nTI + t) * K + i];
          const float v = allErr[(q * nTI + t) * K + i];
          if (p != 0xFFFFFFFFu) m.push_back({quantum > 0.f ? std::rint(v / quantum) : v, {t, p}});
        }
      std::stable_sort(m.begin(), m.end(), [](const auto& a, const auto& b) { return a.first > b.first; });
      const float ref = tis[sm.tiIndex].data[sm.cell];
      for (int i = 0; i < K && i < (int)m.size(); ++i) {
        // the reference reads both values at the cell index of the
        // variable-interleaved array (variable 0 for one variable)
        const float x = tis[m[i].second.first].data[m[i].second.second];
        const double err = std::pow(std::fabs(x - ref), rq->power);
        cumErr[sx * K + i] += err;
        cumSq[sx * K + i] += err * err;
        nSamp[sx * K + i]++;
      }
      if (logged < rq->logCap) {
        rq->logSetting[logged] = (int)sx;
        rq->logCell[logged * 2] = (unsigned)sm.tiIndex;
        rq->logCell[logged * 2 + 1] = sm.cell;
        for (int i = 0; i < K; ++i) {
          const bool ok = i < (int)m.size();
          rq->logPos[(logged * K + i) * 2] = ok ? (unsigned)m[i].second.first : 0xFFFFFFFFu;
          rq->logPos[(logged * K + i) * 2 + 1] = ok ? m[i].second.second : 0xFFFFFFFFu;
        }
        logged++;
      }
      dists++;
      const float errAfter = (float)(cumErr[sx * K] / nSamp[sx * K]);
      if (errBefore > errAfter && best[den] > errAfter) {
        best[den] = errAfter;
        devBest[den] = (float)std::sqrt(cumSq[sx * K] / nSamp[sx * K] - (double)(errAfter * errAfter));
      }
    }
    iter = (int)it + 1;
    if (std::chrono::duration<double>(clk::now() - t0).count() > maxSec) break;
  }
  cleanup();
  if (rc) return rc;
  // :255-262
  for (size_t i = 0; i < settings * (size_t)K; ++i) {
    const float mean = (float)(cumErr[i] / nSamp[i]);
    rs->meanError[i] = std::pow(mean, 1.f / rq->power);
    rs->devError[i] = (float)std::pow(cumSq[i] / nSamp[i] - (double)(mean * mean), 0.5 / rq->power);
    rs->samples[i] = nSamp[i];
  }
  rs->distributions = dists;
  rs->iterations = iter;
  rs->waves = waves;
  rs->ms = std::chrono::duration<double, std::milli>(clk::now() - t1).count();
  rs->scanMs = scanMs;
  rs->drawMs = drawMs;
  for (int i = 0; i < 4; ++i) rs->stageMs[i] = stage[i];
  rs->setupMs = std::chrono::duration<double, std::milli>(t1 - t0).count();
  return QSMI_OK;
}
