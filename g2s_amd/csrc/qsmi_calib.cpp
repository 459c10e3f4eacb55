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
  std::uniform_real_distribution<float> U(0.f, 1.f);
  const int waveMax = rq->waveSize > 0 ? std::min(rq->waveSize, 1024) : 256;
  const double maxSec = rq->maxSeconds > 0 ? rq->maxSeconds : INFINITY;
  const float nbSigmas = 1.f / rq->power;
  long long logged = 0, dists = 0;
  int waves = 0, iter = 0;
  double scanMs = 0.0;
  int rc = QSMI_OK;
  std::vector<Sample> smp;
  // per-wave host arrays
  std::vector<int> nbStart, offs;
  std::vector<float> uvals, d0, seeds, exR, oErr;
  std::vector<unsigned> exP, oPos;
  std::vector<float> ubuf;

  for (unsigned it = 0; it < rq->maxIter && !rc; ++it) {
    smp.clear();
    for (size_t si = 0; si < settings; ++si) {
      const int cfg = (int)(si % nCfg);
      const int ker = (int)((si / nCfg) % nKer);
      const int den = (int)(si / ((size_t)nCfg * nKer));
      // include/calibration.hpp:101-107: skip a setting clearly worse than
      // its density's best
      const unsigned n0 = nSamp[si * K];
      const float eb = (float)(cumErr[si * K] / n0);
      const float db = (float)std::sqrt(cumSq[si * K] / n0 - (double)(eb * eb));
      if (n0 > rq->minIter && (eb - nbSigmas * db) > (best[den] + nbSigmas * devBest[den])) continue;
      const float density = rq->density[den];
      const unsigned* nbc = rq->nbList + (size_t)cfg * rq->nbWidth;
      const int tiIndex = (int)std::floor(U(gen) * (float)nTI);
      const Tiv& T = tis[tiIndex];
      // float arithmetic as the reference's U * dataSize() / nbVariable
      const unsigned cell = (unsigned)std::floor(U(gen) * (float)(T.cells * nVar) / (float)nVar);
      // :115-154: offsets in order, each kept with probability `density`
      // while inside the TI, until every variable has its count
      Sample sm;
      sm.setting = (int)si;
      sm.tiIndex = tiIndex;
      sm.cell = cell;
      std::vector<unsigned> perVar(nVar, 0);
      std::vector<std::vector<float>> raw;
      size_t ps = 0;
      while ((rq->nbWidth > 1 || sm.off.size() < nbc[0]) && ps < pp.size()) {
        unsigned loc;
        if (U(gen) < density && index_with_delta(D, T.s, cell, pp[ps].data(), &loc)) {
          std::vector<float> x(nVar);
          unsigned cpt = 0;
          for (int v = 0; v < nVar; ++v) {
            if (perVar[v] < nbc[v % rq->nbWidth]) {
              x[v] = T.data[(size_t)loc * nVar + v];
              cpt++;
              perVar[v]++;
            } else {
              x[v] = std::nanf("0");
            }
          }
          raw.push_back(x);
          sm.off.push_back(pp[ps]);
          if (cpt == 0) break;
        }
        ps++;
      }
      if (ps >= pp.size()) continue;
      // :158-180: one value per expanded variable (a category's indicator)
      sm.val.resize(raw.size() * nExp);
      for (size_t j = 0; j < raw.size(); ++j) {
        int e = 0;
        for (int v = 0; v < nVar; ++v) {
          if (rq->varTypes[v] == 0) {
            sm.val[j * nExp + e++] = raw[j][v];
          } else {
            for (float c : catv[v]) sm.val[j * nExp + e++] = (raw[j][v] == c) ? 1.f : 0.f;
          }
        }
      }
      // :184-203: noise, offsets swapped between neighbours
      if (sm.off.size() > 1 && rq->noise > 0.f) {
        float fraction = rq->noise / (float)sm.off.size() / 2.f;
        if (rq->noise < 1.0f) fraction = rq->noise / 2.f;
        for (size_t j = 0; j < sm.off.size(); ++j) {
          if (U(gen) < fraction) {
            const size_t a = (size_t)std::floor(U(gen) * (float)sm.off.size());
            const size_t b = (size_t)std::floor(U(gen) * (float)sm.off.size());
            std::iter_swap(sm.off.begin() + a, sm.off.begin() + b);
          }
        }
      }
      if (sm.off.size() > 1) {
        sm.seed = U(gen);
        smp.push_back(std::move(sm));
      }
    }
    // scan: every sample on every TI, in waves
    const size_t ns = smp.size();
    std::vector<float> allErr(ns * (size_t)nTI * K);
    std::vector<unsigned> allPos(ns * (size_t)nTI * K);
    const auto ts = clk::now();
    for (int t = 0; t < nTI && !rc; ++t) {
      for (size_t w0 = 0; w0 < ns && !rc; w0 += waveMax) {
        const int nn = (int)std::min<size_t>(waveMax, ns - w0);
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
          // kernel: u = w, w x, w x^2 per sub-variable (0 outside the
          // kernel or for NaN), delta0 = -sum w x^2 over the continuous
          // variables at their original index
          unsigned centre = 0;
          for (int i = D - 1; i >= 0; --i) centre = centre * kq.dims[i] + kq.dims[i] / 2;
          ubuf.assign((size_t)nb * nSub, 0.f);
          std::vector<unsigned> kidx(nb);
          std::vector<char> kin(nb);
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
        oErr.resize((size_t)nn * K);
        oPos.resize((size_t)nn * K);
        rc = qsmi_distribution_wave(tis[t].dev, nn, nbStart.data(), offs.data(), uvals.data(), d0.data(), seeds.data(),
                                    (float)K, exP.data(), exR.data(), oErr.data(), oPos.data());
        if (rc) break;
        for (int q = 0; q < nn; ++q)
          for (int i = 0; i < K; ++i) {
            allErr[((w0 + q) * nTI + t) * K + i] = oErr[(size_t)q * K + i];
            allPos[((w0 + q) * nTI + t) * K + i] = oPos[(size_t)q * K + i];
          }
        waves++;
      }
    }
    scanMs += std::chrono::duration<double, std::milli>(clk::now() - ts).count();
    if (rc) break;
    // accumulate in setting order (include/calibration.hpp:206-246)
    for (size_t q = 0; q < ns; ++q) {
      const Sample& sm = smp[q];
      const size_t si = (size_t)sm.setting;
      const int den = (int)(si / ((size_t)nCfg * nKer));
      const unsigned n0 = nSamp[si * K];
      const float errBefore = (float)(cumErr[si * K] / n0);
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
        cumErr[si * K + i] += err;
        cumSq[si * K + i] += err * err;
        nSamp[si * K + i]++;
      }
      if (logged < rq->logCap) {
        rq->logSetting[logged] = (int)si;
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
      const float errAfter = (float)(cumErr[si * K] / nSamp[si * K]);
      if (errBefore > errAfter && best[den] > errAfter) {
        best[den] = errAfter;
        devBest[den] = (float)std::sqrt(cumSq[si * K] / nSamp[si * K] - (double)(errAfter * errAfter));
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
  rs->ms = std::chrono::duration<double, std::milli>(clk::now() - t0).count();
  rs->scanMs = scanMs;
  return QSMI_OK;
}
