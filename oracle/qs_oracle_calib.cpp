// auto_qs calibration, CPU restatement — TEST INFRASTRUCTURE ONLY.
//
// The Monte Carlo loop of calibration() (include/calibration.hpp:30-253)
// with auto_qs's set-up (src/auto_qs.cpp:449-492, 561-582, 640-666), each
// QuantileSamplingModule::distribution call (include/quantileSamplingModule.
// hpp:128-169, 342-577) evaluated by brute force: the FP64 direct map of the
// sample (orc_map_direct_f64), the verbatim disk set to -inf, then the
// device selection rule on the FP64 maps (orc_select_device_f64). Written
// as the reference's single-threaded loop, with one stated change: the skip
// test of an iteration reads the running best per density as it stood at
// the start of that iteration (the order the GPU batch needs; the
// reference's parallel loop leaves the order of these updates to its OpenMP
// schedule).
//
// Pinned by tests/test_oracle_calib.py: offsets order (every box offset
// once, non-decreasing distance, std::sort's order of equal distances
// against an independent Python restatement of the construction), draw
// sequence against a Python restatement over numpy's MT19937 on the same
// seed, the exclusion disk and k-best against brute force, accumulation
// and the skip rule against the definitions on a hand-made run.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <vector>

extern "C" {
int orc_coef_matrix(int nVar, const int* types, const int* ncat, float* alpha, int* convTypes, int* subVarOf);
int orc_ti_subimages(int D, const unsigned* s, const unsigned* f, int nVar, const int* types, const float* data,
                     const int* ncat, const float* catvals, int reflected, float* out);
float orc_convert_neighbors(int D, int nb, const int* off, int nExp, const float* vals, int nSub,
                            const int* convTypes, const int* subVarOf, int nVarOrig, const int* typesOrig,
                            const unsigned* kdims, const float* kernel, float* u);
void orc_map_direct_f64(int D, const unsigned* s, int nSub, const float* Y, const float* alpha, int nb,
                        const int* off, const float* u, double delta0, int circular, double* out);
int orc_select_device_f64(int D, const unsigned* s, const double* map, const int* lo, const int* hi, int K,
                          float seed, float quantum, double* outVal, unsigned* outPos);
int orc_select_device(int D, const unsigned* s, const float* map, const int* lo, const int* hi, int K, float seed,
                      float quantum, float* outVal, unsigned* outPos);
int orc_index_with_delta(int D, const unsigned* dims, unsigned position, const int* delta, unsigned* location);
}

namespace {

// src/auto_qs.cpp:449-474 (construction) and :476-492 (sort by the weight
// kernel -distance^2, DataImage::distance2ToCenter, include/DataImage.hpp:
// 521-536, centre (dims-1)/2)
std::vector<std::vector<int>> calib_offsets(int D, const unsigned* kdims) {
  std::vector<std::vector<int>> pathPosition;
  pathPosition.push_back(std::vector<int>(0));
  for (int i = 0; i < D; ++i) {
    unsigned originalSize = pathPosition.size();
    int sizeInThisDim = (kdims[i] + 1) / 2;
    pathPosition.resize(originalSize * (2 * sizeInThisDim - 1));
    for (unsigned k = 0; k < originalSize; ++k) pathPosition[k].push_back(0);
    for (int j = 1; j < sizeInThisDim; ++j) {
      std::copy(pathPosition.begin(), pathPosition.begin() + originalSize,
                pathPosition.begin() + originalSize * (2 * j - 1));
      std::copy(pathPosition.begin(), pathPosition.begin() + originalSize, pathPosition.begin() + originalSize * (2 * j));
      for (unsigned k = originalSize * (2 * j - 1); k < originalSize * (2 * j); ++k) pathPosition[k][i] = j;
      for (unsigned k = originalSize * (2 * j); k < originalSize * (2 * j + 1); ++k) pathPosition[k][i] = -j;
    }
  }
  // the weight image: -distance2ToCenter per cell of kernel 0's grid
  size_t kc = 1;
  for (int i = 0; i < D; ++i) kc *= kdims[i];
  std::vector<float> weight(kc);
  for (size_t c = 0; c < kc; ++c) {
    size_t v = c;
    float d2 = 0.f;
    for (int i = 0; i < D; ++i) {
      int d = std::abs(int(v % kdims[i]) - (int(kdims[i]) - 1) / 2);
      d2 += d * d;
      v /= kdims[i];
    }
    weight[c] = -d2;
  }
  unsigned center = 0;
  for (int i = D - 1; i >= 0; --i) center = center * kdims[i] + (kdims[i] - 1) / 2;
  std::sort(pathPosition.begin(), pathPosition.end(), [&](std::vector<int>& a, std::vector<int>& b) {
    unsigned l1 = 0, l2 = 0;
    orc_index_with_delta(D, kdims, center, a.data(), &l1);
    orc_index_with_delta(D, kdims, center, b.data(), &l2);
    return weight[l1] > weight[l2];
  });
  return pathPosition;
}

}  // namespace

extern "C" {

// Offsets in calibration order, D ints each; returns their count.
int orc_calib_offsets(int D, const unsigned* kdims, int* out, int cap) {
  auto pp = calib_offsets(D, kdims);
  int n = 0;
  for (auto& p : pp) {
    if (n >= cap) break;
    for (int i = 0; i < D; ++i) out[n * D + i] = p[i];
    n++;
  }
  return (int)pp.size();
}

// The exclusion disk the product states (Euclidean, radius^2 inclusive, no
// wrap) applied to a position-space map.
void orc_exclude_disk(int D, const unsigned* s, float* map, unsigned origin, float radius) {
  const float r2 = radius >= 1.f ? radius * radius : 0.f;
  int o[3] = {0, 0, 0};
  unsigned rem = origin;
  for (int i = 0; i < D; ++i) {
    o[i] = int(rem % s[i]);
    rem /= s[i];
  }
  size_t S = 1;
  for (int i = 0; i < D; ++i) S *= s[i];
  for (size_t c = 0; c < S; ++c) {
    size_t v = c;
    float d2 = 0.f;
    for (int i = 0; i < D; ++i) {
      float d = float(int(v % s[i]) - o[i]);
      d2 += d * d;
      v /= s[i];
    }
    if (d2 <= r2) map[c] = -INFINITY;
  }
}

int orc_calibration(int D, int nTI, const unsigned* tiSize, const float* const* TIs, int nVar, const int* types,
                    int nKer, const unsigned* kerSize, const float* const* kernelsIn, int nCfg, int nbWidth,
                    const unsigned* nbList, int nDen, const float* densityArray, int maxK, float power,
                    unsigned maxNumberOfIteration, unsigned minNumberOfIteration, float noiseLevel, unsigned seed,
                    float quantum, float* meanOut, float* devOut, unsigned* samplesOut, int logCap, int* logSetting,
                    unsigned* logCell, unsigned* logPos, long long* nDist) {
  // categories (src/auto_qs.cpp:561-582)
  std::vector<std::vector<float>> categoriesValues;
  std::vector<unsigned> numberOfComputedVariablesPerVariable;
  for (int i = 0; i < nVar; ++i) {
    if (types[i] == 0) numberOfComputedVariablesPerVariable.push_back(1);
    if (types[i] == 1) {
      std::vector<float> cur;
      for (int im = 0; im < nTI; ++im) {
        size_t cells = 1;
        for (int d = 0; d < D; ++d) cells *= tiSize[im * 3 + d];
        for (size_t j = i; j < cells * nVar; j += nVar) {
          float x = TIs[im][j];
          if (std::isnan(x)) continue;
          bool present = false;
          for (float c : cur) present |= (x == c);
          if (!present) cur.push_back(x);
        }
      }
      categoriesValues.push_back(cur);
      numberOfComputedVariablesPerVariable.push_back(cur.size());
    }
  }
  int nExp = 0;
  for (unsigned f : numberOfComputedVariablesPerVariable) nExp += f;
  std::vector<int> ncat;
  std::vector<float> catflat;
  for (auto& c : categoriesValues) {
    ncat.push_back((int)c.size());
    catflat.insert(catflat.end(), c.begin(), c.end());
  }
  std::vector<float> alpha(64 * 64);
  std::vector<int> convTypes(64), subVarOf(64);
  const int nSub = orc_coef_matrix(nVar, types, ncat.data(), alpha.data(), convTypes.data(), subVarOf.data());
  // kernels (offsetKernel4categories, include/DataImage.hpp:492-518)
  std::vector<std::vector<float>> kernels(nKer);
  for (int q = 0; q < nKer; ++q) {
    size_t kc = 1;
    for (int d = 0; d < D; ++d) kc *= kerSize[q * 3 + d];
    kernels[q].assign(kc * nExp, 0.f);
    unsigned cp = 0;
    for (int i = 0; i < nVar; ++i)
      for (unsigned j = 0; j < numberOfComputedVariablesPerVariable[i]; ++j, ++cp)
        for (size_t k = 0; k < kc; ++k) kernels[q][k * nExp + cp] = kernelsIn[q][k * nVar + i];
  }
  auto pathPosition = calib_offsets(D, kerSize);
  // sub-images per TI, unpadded
  std::vector<std::vector<float>> Y(nTI);
  std::vector<size_t> cellsOf(nTI);
  for (int t = 0; t < nTI; ++t) {
    size_t S = 1;
    for (int d = 0; d < D; ++d) S *= tiSize[t * 3 + d];
    cellsOf[t] = S;
    Y[t].resize((size_t)nSub * S);
    orc_ti_subimages(D, tiSize + t * 3, tiSize + t * 3, nVar, types, TIs[t], ncat.data(), catflat.data(), 0,
                     Y[t].data());
  }

  const int maxK_ = maxK;
  const size_t computeArraySize = (size_t)nDen * nCfg * nKer;
  const size_t arraySize = maxK_ * computeArraySize;
  std::vector<double> cumulattedError(arraySize, 0.0), cumulattedSquaredError(arraySize, 0.0);
  std::vector<unsigned> numberOfSampling(arraySize, 0);
  std::vector<float> bestProDensity(nDen, INFINITY), devBestProDensity(nDen, 0.f);
  const float radius = std::pow((cellsOf[0]) * 0.005, 1.f / D);
  std::mt19937 randomGenerator(seed);
  std::uniform_real_distribution<float> uniformDistributionOverSource(0.f, 1.f);
  long long logged = 0, dists = 0;

  for (unsigned iteration = 0; iteration < maxNumberOfIteration; ++iteration) {
    const std::vector<float> bestSnap = bestProDensity, devSnap = devBestProDensity;
    for (size_t setupIndex = 0; setupIndex < computeArraySize; ++setupIndex) {
      unsigned val = setupIndex;
      unsigned numberOfneihbours = val % nCfg;
      val /= nCfg;
      unsigned kernelIndex = val % nKer;
      val /= nKer;
      unsigned densityIndex = val % nDen;

      float localErrorBefore = cumulattedError[setupIndex * maxK_] / numberOfSampling[setupIndex * maxK_];
      float localDevBefore = std::sqrt(cumulattedSquaredError[setupIndex * maxK_] / numberOfSampling[setupIndex * maxK_] -
                                       localErrorBefore * localErrorBefore);
      float nbSigmas = 1 / power;
      if ((numberOfSampling[setupIndex * maxK_] > minNumberOfIteration) &&
          ((localErrorBefore - nbSigmas * localDevBefore) >
           (bestSnap[densityIndex] + nbSigmas * devSnap[densityIndex])))
        continue;

      float density = densityArray[densityIndex];
      std::vector<unsigned> numberNeighbor(nbList + numberOfneihbours * nbWidth,
                                           nbList + (numberOfneihbours + 1) * nbWidth);
      int tiIndex = int(std::floor(uniformDistributionOverSource(randomGenerator) * nTI));
      const unsigned* tdims = tiSize + tiIndex * 3;
      unsigned currentCell =
          unsigned(std::floor(uniformDistributionOverSource(randomGenerator) * (cellsOf[tiIndex] * nVar) / nVar));

      std::vector<unsigned> numberOfNeighborsProVariable(nVar);
      std::vector<std::vector<int>> neighborArrayVector;
      std::vector<std::vector<float>> neighborValueArrayVector;
      unsigned positionSearch = 0;
      while ((numberNeighbor.size() > 1 || (neighborArrayVector.size() < numberNeighbor[0])) &&
             (positionSearch < pathPosition.size())) {
        unsigned dataIndex;
        std::vector<int> vectorInDi = pathPosition[positionSearch];
        vectorInDi.resize(D, 0);
        if (uniformDistributionOverSource(randomGenerator) < density &&
            orc_index_with_delta(D, tdims, currentCell, vectorInDi.data(), &dataIndex)) {
          std::vector<float> data(nVar);
          unsigned cpt = 0;
          for (int i = 0; i < nVar; ++i) {
            if (numberOfNeighborsProVariable[i] < numberNeighbor[i % numberNeighbor.size()]) {
              data[i] = TIs[tiIndex][dataIndex * nVar + i];
              cpt++;
              numberOfNeighborsProVariable[i]++;
            } else {
              data[i] = std::nanf("0");
            }
          }
          neighborValueArrayVector.push_back(data);
          neighborArrayVector.push_back(pathPosition[positionSearch]);
          if (cpt == 0) break;
        }
        positionSearch++;
      }
      if (positionSearch >= pathPosition.size()) continue;

      for (size_t j = 0; j < neighborValueArrayVector.size(); ++j) {
        std::vector<float> data(nExp);
        unsigned id = 0, idCategorie = 0;
        for (int i = 0; i < nVar; ++i) {
          if (types[i] == 0) data[id++] = neighborValueArrayVector[j][i];
          if (types[i] == 1) {
            for (size_t k = 0; k < categoriesValues[idCategorie].size(); ++k)
              data[id++] = (neighborValueArrayVector[j][i] == categoriesValues[idCategorie][k]);
            idCategorie++;
          }
        }
        neighborValueArrayVector[j] = data;
      }

      if (neighborArrayVector.size() > 1 && noiseLevel > 0) {
        float fraction = noiseLevel / neighborArrayVector.size() / 2;
        if (noiseLevel < 1.0) fraction = noiseLevel / 2;
        for (size_t j = 0; j < neighborValueArrayVector.size(); ++j) {
          if (uniformDistributionOverSource(randomGenerator) < fraction) {
            size_t a = std::floor(uniformDistributionOverSource(randomGenerator) * neighborValueArrayVector.size());
            size_t b = std::floor(uniformDistributionOverSource(randomGenerator) * neighborValueArrayVector.size());
            std::iter_swap(neighborArrayVector.begin() + a, neighborArrayVector.begin() + b);
          }
        }
      }

      if (neighborArrayVector.size() > 1) {
        const float sseed = uniformDistributionOverSource(randomGenerator);
        const int nb = (int)neighborArrayVector.size();
        std::vector<int> off((size_t)nb * D);
        std::vector<float> vals((size_t)nb * nExp);
        for (int j = 0; j < nb; ++j) {
          for (int d = 0; d < D; ++d) off[j * D + d] = neighborArrayVector[j][d];
          for (int e = 0; e < nExp; ++e) vals[j * nExp + e] = neighborValueArrayVector[j][e];
        }
        std::vector<float> u((size_t)nb * nSub);
        const unsigned* kd = kerSize + kernelIndex * 3;
        const float delta0 = orc_convert_neighbors(D, nb, off.data(), nExp, vals.data(), nSub, convTypes.data(),
                                                   subVarOf.data(), nVar, types, kd, kernels[kernelIndex].data(),
                                                   u.data());
        // distribution: every TI's k best, merged by error (stable)
        std::vector<float> errs;
        std::vector<std::pair<int, unsigned>> where;
        for (int t = 0; t < nTI; ++t) {
          const size_t S = cellsOf[t];
          std::vector<double> m64(S);
          orc_map_direct_f64(D, tiSize + t * 3, nSub, Y[t].data(), alpha.data(), nb, off.data(), u.data(), delta0, 0,
                             m64.data());
          if (t == tiIndex) {  // the disk, marked on a float copy, applied to the FP64 map
            std::vector<float> mark(S, 0.f);
            orc_exclude_disk(D, tiSize + t * 3, mark.data(), currentCell, radius);
            for (size_t c = 0; c < S; ++c)
              if (mark[c] == -INFINITY) m64[c] = -INFINITY;
          }
          int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
          for (int d = 0; d < D; ++d) hi[d] = int(tiSize[t * 3 + d]) - 1;
          std::vector<double> ov(maxK_);
          std::vector<unsigned> op(maxK_);
          orc_select_device_f64(D, tiSize + t * 3, m64.data(), lo, hi, maxK_, sseed, quantum, ov.data(), op.data());
          for (int i = 0; i < maxK_; ++i)
            if (op[i] != UINT_MAX) {
              const float v = (float)ov[i];
              errs.push_back(quantum > 0.f ? std::rint(v / quantum) : v);
              where.push_back({t, op[i]});
            }
        }
        std::vector<size_t> order(errs.size());
        for (size_t i = 0; i < order.size(); ++i) order[i] = i;
        std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return errs[a] > errs[b]; });

        for (int i = 0; i < maxK_ && i < (int)order.size(); ++i) {
          const auto& wi = where[order[i]];
          double error = std::pow(std::fabs(TIs[wi.first][wi.second] - TIs[tiIndex][currentCell]), power);
          cumulattedError[setupIndex * maxK_ + i] += error;
          cumulattedSquaredError[setupIndex * maxK_ + i] += error * error;
          numberOfSampling[setupIndex * maxK_ + i]++;
        }
        if (logged < logCap) {
          logSetting[logged] = (int)setupIndex;
          logCell[logged * 2] = tiIndex;
          logCell[logged * 2 + 1] = currentCell;
          for (int i = 0; i < maxK_; ++i) {
            const bool ok = i < (int)order.size();
            logPos[(logged * maxK_ + i) * 2] = ok ? (unsigned)where[order[i]].first : UINT_MAX;
            logPos[(logged * maxK_ + i) * 2 + 1] = ok ? where[order[i]].second : UINT_MAX;
          }
          logged++;
        }
        dists++;
        float localErrorAfter = cumulattedError[setupIndex * maxK_] / numberOfSampling[setupIndex * maxK_];
        if (localErrorBefore > localErrorAfter) {
          if (bestProDensity[densityIndex] > localErrorAfter) {
            bestProDensity[densityIndex] = localErrorAfter;
            devBestProDensity[densityIndex] =
                std::sqrt(cumulattedSquaredError[setupIndex * maxK_] / numberOfSampling[setupIndex * maxK_] -
                          localErrorAfter * localErrorAfter);
          }
        }
      }
    }
  }
  for (size_t i = 0; i < arraySize; ++i) {
    float localMeanValue = cumulattedError[i] / numberOfSampling[i];
    meanOut[i] = std::pow(localMeanValue, 1 / power);
    devOut[i] = std::pow(cumulattedSquaredError[i] / numberOfSampling[i] - localMeanValue * localMeanValue, 0.5 / power);
    samplesOut[i] = numberOfSampling[i];
  }
  *nDist = dists;
  return 0;
}

}  // extern "C"
