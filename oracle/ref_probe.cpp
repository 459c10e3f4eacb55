// Driver that calls the reference's own header-only QS pieces, compiled in
// place from /root/reference (oracle/Makefile), to produce checks for the
// oracle. Built into oracle/_ref/libqs_ref.so; TEST INFRASTRUCTURE ONLY.
//
// Functions reached (all unmodified reference code):
//   SharedMemoryManager size rule     src/sharedMemoryManager.cpp:26-58
//   fKst::findKBiggest (RNG variant)  include/fKb.hpp:58-93,121-132,907-954
//   DataImage::convertInput4Xcorr     include/DataImage.hpp:538-599
//   DataImage::generateCoefMatrix4Xcorr include/DataImage.hpp:626-710
//   DataImage::genearteKernel         include/DataImage.hpp:433-480
//   DataImage::offsetKernel4categories include/DataImage.hpp:482-508
//   DataImage::indexWithDelta         include/DataImage.hpp:343-376
//   g2s::complexAddAlphaxCxD          include/complexMulti.hpp:58-67,303-331
//   the same two with the cross-measurement layer (TIs with NaN)
//   DataImage::distance2ToCenter      include/DataImage.hpp:521-530 (the
//                                     offset sort key of src/qs.cpp:1055-1137)
//   DataImage::serialize              include/DataImage.hpp:143-170
//   DataImage::validateSerializedData include/DataImage.hpp:172-224
// The glue below only marshals plain arrays in and out.
#include <algorithm>
#include <climits>
#include <cstring>
#include <cmath>
#include <functional>
#include <random>
#include <vector>

#include "DataImage.hpp"
#include "OperationMatrix.hpp"
#include "complexMulti.hpp"
#include "fKb.hpp"
#include "sharedMemoryManager.hpp"

extern "C" {

void ref_fft_size(int D, const unsigned* src, unsigned* out) {
  std::vector<unsigned> s(src, src + D);
  SharedMemoryManager m(s);
  for (int i = 0; i < D; ++i) out[i] = m._fftSize[i];
}

// The call site of include/quantileSamplingModule.hpp:533-539 for one chunk
// (threadRatio 1): generator seeded from the unit seed, float uniform RNG.
void ref_fkst(const float* data, unsigned N, unsigned short k, float seed, float* out, unsigned* pos) {
  unsigned sd;
  if (!std::isfinite(seed) || seed <= 0.f)
    sd = 0;
  else if (seed >= 1.f)
    sd = std::numeric_limits<unsigned>::max();
  else
    sd = static_cast<unsigned>(std::floor(static_cast<double>(seed) * static_cast<double>(std::numeric_limits<unsigned>::max())));
  std::mt19937 generator;
  generator.seed(sd);
  std::uniform_real_distribution<float> distribution(0.0, 1.0);
  auto rng = std::bind(distribution, std::ref(generator));
  fKst::findKBiggest(data, N, k, out, pos, rng);
}

static g2s::DataImage make_image(int D, const unsigned* dims, int nVar, const int* types, const float* data) {
  std::vector<unsigned> d(dims, dims + D);
  g2s::DataImage im(D, d.data(), nVar);
  for (int v = 0; v < nVar; ++v)
    im._types[v] = types[v] ? g2s::DataImage::Categorical : g2s::DataImage::Continuous;
  size_t n = im.dataSize();
  for (size_t i = 0; i < n; ++i) im._data[i] = data[i];
  return im;
}

static std::vector<std::vector<float>> make_cats(int nVar, const int* types, const int* ncat, const float* catvals) {
  std::vector<std::vector<float>> cats;
  int ci = 0;
  const float* cv = catvals;
  for (int v = 0; v < nVar; ++v)
    if (types[v]) {
      cats.push_back(std::vector<float>(cv, cv + ncat[ci]));
      cv += ncat[ci++];
    }
  return cats;
}

// Writes every sub-image (f-grid, reflected) back to back; returns the count.
int ref_convert_input4xcorr(int D, const unsigned* s, const unsigned* f, int nVar, const int* types, const float* data,
                            const int* ncat, const float* catvals, float* out) {
  g2s::DataImage ti = make_image(D, s, nVar, types, data);
  std::vector<unsigned> fs(f, f + D);
  auto cats = make_cats(nVar, types, ncat, catvals);
  auto imgs = ti.convertInput4Xcorr(fs, false, cats);
  int cnt = 0;
  for (auto& grp : imgs)
    for (auto& im : grp) {
      size_t n = im.dataSize();
      for (size_t i = 0; i < n; ++i) out[cnt * n + i] = im._data[i];
      cnt++;
    }
  return cnt;
}

// alpha[A + B*nSub] of the regular layer, the main conversion types flattened
// (one entry per sub-variable) and the constant-term types/coefs per variable.
int ref_coef_matrix(int nVar, const int* types, const int* ncat, const float* catvals, float* alpha, int* convTypes,
                    int* constTypes, float* constCoefs) {
  std::vector<unsigned> dims = {1};
  std::vector<float> dummy(nVar, 0.f);
  g2s::DataImage ti = make_image(1, dims.data(), nVar, types, dummy.data());
  auto cats = make_cats(nVar, types, ncat, catvals);
  std::vector<g2s::OperationMatrix> coef;
  std::vector<std::vector<conversionType>> convMain;
  std::vector<std::vector<std::vector<conversionType>>> convConst;
  std::vector<std::vector<std::vector<float>>> coefConst;
  ti.generateCoefMatrix4Xcorr(coef, convMain, convConst, coefConst, false, cats);
  int nSub = coef[0].getNumberOfVariable();
  for (int A = 0; A < nSub; ++A)
    for (int B = 0; B < nSub; ++B) alpha[A + B * nSub] = coef[0].getVariableAt(A, B);
  int q = 0;
  for (auto& v : convMain)
    for (auto t : v) convTypes[q++] = int(t);
  for (int v = 0; v < nVar; ++v) {
    constTypes[v] = convConst[0][v].empty() ? -9 : int(convConst[0][v][0]);
    constCoefs[v] = convConst[0][v].empty() ? 0.f : coefConst[0][v][0];
  }
  return nSub;
}

void ref_generate_kernel(int D, const unsigned* dims, int nVar, const int* ktype, const float* weight,
                         const float* alphas, float* out) {
  std::vector<g2s::KernelType> kt;
  for (int v = 0; v < nVar; ++v) kt.push_back(static_cast<g2s::KernelType>(ktype[v]));
  std::vector<unsigned> ms(dims, dims + D);
  std::vector<float> w(weight, weight + nVar), a(alphas, alphas + nVar);
  g2s::DataImage k = g2s::DataImage::genearteKernel(kt, ms, w, a);
  size_t n = k.dataSize();
  for (size_t i = 0; i < n; ++i) out[i] = k._data[i];
}

// Neighbour offsets of one kernel in qs.cpp's order (src/qs.cpp:1055-1137,
// Euclidean search distance). Only the candidate generation is restated
// here (every offset with |o_i| < (k_i + 1) / 2; per dimension the previous
// list with o_i = 0, then with o_i = +1, -1, +2, -2, ...). The sort key image
// (minus the squared distance to the centre), the centre, the NaN test and
// the comparator's lookups are the compiled DataImage's (emptyCopy,
// distance2ToCenter, indexWithDelta), sorted by std::sort like qs.cpp.
// kernel: cells x kNVar (nullptr: no NaN removal). Returns the count.
int ref_offsets(int D, const unsigned* kdims, int kNVar, const float* kernel, int* out) {
  size_t cells = 1;
  for (int i = 0; i < D; ++i) cells *= kdims[i];
  std::vector<float> kdata = kernel ? std::vector<float>(kernel, kernel + cells * kNVar)
                                    : std::vector<float>(cells * kNVar, 1.f);
  std::vector<int> types(kNVar, 0);
  g2s::DataImage kim = make_image(D, kdims, kNVar, types.data(), kdata.data());
  std::vector<std::vector<int>> offs(1);
  for (int i = 0; i < D; ++i) {
    const int half = ((int)kdims[i] + 1) / 2;
    std::vector<int> order = {0};  // the blocks' o_i: 0, +1, -1, +2, -2, ...
    for (int j = 1; j < half; ++j) {
      order.push_back(j);
      order.push_back(-j);
    }
    std::vector<std::vector<int>> grown;
    for (int o : order)
      for (std::vector<int> v : offs) {
        v.push_back(o);
        grown.push_back(v);
      }
    offs.swap(grown);
  }
  g2s::DataImage w = kim.emptyCopy(true);
  for (unsigned i = 0; i < w.dataSize(); ++i) w._data[i] = -w.distance2ToCenter(i);
  unsigned centre = 0;
  for (int i = D - 1; i >= 0; --i) centre = centre * w._dims[i] + (w._dims[i] - 1) / 2;
  if (kernel) {
    std::vector<std::vector<int>> kept;
    for (auto& v : offs) {
      unsigned at;
      kim.indexWithDelta(at, centre, v);
      if (!std::isnan(kim._data[at * (unsigned)kNVar])) kept.push_back(v);
    }
    offs.swap(kept);
  }
  std::sort(offs.begin(), offs.end(), [&](std::vector<int>& a, std::vector<int>& b) {
    unsigned la, lb;
    w.indexWithDelta(la, centre, a);
    w.indexWithDelta(lb, centre, b);
    return w._data[la] > w._data[lb];
  });
  for (size_t k = 0; k < offs.size(); ++k)
    for (int i = 0; i < D; ++i) out[k * D + i] = offs[k][i];
  return (int)offs.size();
}

int ref_kernel_expand(int D, const unsigned* dims, int nVarIn, const float* kin, const int* factor, float* out) {
  std::vector<int> t(nVarIn, 0);
  g2s::DataImage k = make_image(D, dims, nVarIn, t.data(), kin);
  std::vector<unsigned> fac(factor, factor + nVarIn);
  g2s::DataImage e = g2s::DataImage::offsetKernel4categories(k, fac, false);
  size_t n = e.dataSize();
  for (size_t i = 0; i < n; ++i) out[i] = e._data[i];
  return e._nbVariable;
}

int ref_index_with_delta(int D, const unsigned* dims, unsigned position, const int* delta, unsigned* location) {
  std::vector<int> t(1, 0);
  std::vector<float> one(1, 0.f);
  std::vector<unsigned> d(dims, dims + D);
  g2s::DataImage im(D, d.data(), 1);
  std::vector<int> dv(delta, delta + D);
  return im.indexWithDelta(*location, position, dv) ? 1 : 0;
}

void ref_complex_add(float* dst, const float* C, const float* Dm, float alpha, unsigned size) {
  g2s::complexAddAlphaxCxD(dst, C, Dm, alpha, size);
}

// convertInput4Xcorr with needCrossMeasurement (mask image per variable)
int ref_convert_input4xcorr_x(int D, const unsigned* s, const unsigned* f, int nVar, const int* types,
                              const float* data, const int* ncat, const float* catvals, float* out) {
  g2s::DataImage ti = make_image(D, s, nVar, types, data);
  std::vector<unsigned> fs(f, f + D);
  auto cats = make_cats(nVar, types, ncat, catvals);
  auto imgs = ti.convertInput4Xcorr(fs, true, cats);
  int cnt = 0;
  for (auto& grp : imgs)
    for (auto& im : grp) {
      size_t n = im.dataSize();
      for (size_t i = 0; i < n; ++i) out[cnt * n + i] = im._data[i];
      cnt++;
    }
  return cnt;
}

// generateCoefMatrix4Xcorr with forXMeasurement: both layers [A + B*nSub]
// and the main conversion types flattened; returns nSub.
int ref_coef_matrix_x(int nVar, const int* types, const int* ncat, const float* catvals, float* alpha, float* alphaX,
                      int* convTypes) {
  std::vector<unsigned> dims = {1};
  std::vector<float> dummy(nVar, 0.f);
  g2s::DataImage ti = make_image(1, dims.data(), nVar, types, dummy.data());
  auto cats = make_cats(nVar, types, ncat, catvals);
  std::vector<g2s::OperationMatrix> coef;
  std::vector<std::vector<conversionType>> convMain;
  std::vector<std::vector<std::vector<conversionType>>> convConst;
  std::vector<std::vector<std::vector<float>>> coefConst;
  ti.generateCoefMatrix4Xcorr(coef, convMain, convConst, coefConst, true, cats);
  int nSub = coef[0].getNumberOfVariable();
  for (int A = 0; A < nSub; ++A)
    for (int B = 0; B < nSub; ++B) {
      alpha[A + B * nSub] = coef[0].getVariableAt(A, B);
      alphaX[A + B * nSub] = coef.size() > 1 ? coef[1].getVariableAt(A, B) : 0.f;
    }
  int q = 0;
  for (auto& v : convMain)
    for (auto t : v) convTypes[q++] = int(t);
  return nSub;
}

// serialized bytes of a grid (returns the full size; copies at most cap bytes)
size_t ref_bgrid_serialize(int nd, const unsigned* dims, int nVar, const int* types, int enc, const float* data,
                           char* out, size_t cap) {
  std::vector<unsigned> d(dims, dims + nd);
  g2s::DataImage im((unsigned)nd, d.data(), (unsigned)nVar);
  for (int v = 0; v < nVar; ++v) im._types[v] = (g2s::DataImage::VariableType)types[v];
  im.setEncoding((g2s::DataImage::EncodingType)enc);
  size_t n = nVar;
  for (int i = 0; i < nd; ++i) n *= dims[i];
  memcpy(im._data, data, n * sizeof(float));
  char* raw = im.serialize();
  size_t full = 0;
  memcpy(&full, raw, sizeof(size_t));
  memcpy(out, raw, full < cap ? full : cap);
  free(raw);
  return full;
}

int ref_bgrid_valid(const char* raw, size_t size) {
  return g2s::DataImage::validateSerializedData(raw, size) ? 1 : 0;
}

}  // extern "C"
