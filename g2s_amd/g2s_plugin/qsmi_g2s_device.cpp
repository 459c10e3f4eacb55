// G2S AcceleratorDevice over the qsmi C-ABI: the MI355X QuickSampling scan as
// the device G2S's qs and auto_qs load from their dlopen GPU slot
// (-W_CUDA: dlopen("./g2s_cuda.so"), symbols g2s_cudaGetDeviceCount and
// c_NvidiaGPUAcceleratorDevice, src/qs.cpp:861-911, src/auto_qs.cpp:364-406;
// the symbols' contract: src/NvidiaGPUAcceleratorDevice.cu:26-29, 229-260).
//
// Build (needs G2S's include/ and FFTW's headers, which G2S's
// AcceleratorDevice.hpp includes; the same C++ standard library as the
// qs executable, since std::vector crosses the dlopen boundary):
//   g++ -O2 -std=c++17 -fopenmp -shared -fPIC -I<g2s>/include -I<this repo>/include \
//       qsmi_g2s_device.cpp -L<this repo>/g2s_amd -lqsmi -Wl,-rpath,<this repo>/g2s_amd -o g2s_cuda.so
// (g2s_amd/build.py does this when it finds both.) The qs executable provides
// AcceleratorDevice's constructor, initDim and SharedMemoryManager at load
// time (it is linked with -Wl,--export-dynamic, build/Makefile:20,106).
//
// Ownership follows the reference's GPU device: every instance is one path
// thread's device; instances on one GPU share the GPU's TI spectra through
// SharedMemoryManager (the first one's allocAndInitSharedMemory builds the
// qsmi_ti, the others find it by (_deviceType, _deviceID),
// src/sharedMemoryManager.cpp:68-116), and the last one removed frees it.
// Fatal errors print the qsmi message and exit, as the reference's CUDA
// device does (src/NvidiaGPUAcceleratorDevice.cu:45-54).
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "AcceleratorDevice.hpp"
#include "sharedMemoryManager.hpp"
#include "qsmi.h"

namespace {

void check(int rc, const char* what) {
  if (rc == QSMI_OK) return;
  std::fprintf(stderr, "qsmi device: %s: %s\n", what, qsmi_last_error());
  std::exit(rc);
}

[[noreturn]] void unsupported(const char* what) {
  std::fprintf(stderr, "qsmi device: %s is not provided by this device\n", what);
  std::exit(QSMI_ERR_UNSUPPORTED);
}

}  // namespace

class QsmiAcceleratorDevice : public AcceleratorDevice {
 public:
  QsmiAcceleratorDevice(int deviceId, SharedMemoryManager* smm, std::vector<g2s::OperationMatrix> coefficientMatrix,
                        unsigned threadRatio, bool withCrossMeasurement, bool circularTI)
      : AcceleratorDevice(smm, coefficientMatrix, threadRatio, withCrossMeasurement, circularTI) {
    int count = 0;
    check(qsmi_get_device_count(&count), "device count");
    if (count < 1) check(QSMI_ERR_HIP, "no GPU");
    // set before addDevice: SharedMemoryManager::findSameHostDevice matches
    // devices of one GPU by (type, id)
    _deviceType = DT_gpuCuda;
    _deviceID = (unsigned)(deviceId % count);
#pragma omp critical(createDevices)
    { _sharedMemoryManager->addDevice(this); }
    initDim();  // _srcCplx: this GPU's shared entry, [0].space = its qsmi_ti
    qsmi_ti* ti = static_cast<qsmi_ti*>(_srcCplx.at(0).space);
    std::vector<float> alpha = layerCoefficients(0);
    check(qsmi_device_create(ti, alpha.data(), circularTI ? 1 : 0, &_dev), "device create");
    // TIs with NaN: the second layer counts the TI values under the
    // neighbours (DataImage::generateCoefMatrix4Xcorr with forXMeasurement)
    if (withCrossMeasurement) {
      if (_coefficientMatrix.size() < 2) check(QSMI_ERR_ARG, "cross measurement without a second coefficient layer");
      std::vector<float> alphaX = layerCoefficients(1);
      check(qsmi_device_set_cross_layer(_dev, alphaX.data()), "cross layer");
    }
    check(qsmi_set_true_mismatch(_dev, _trueMismatch ? 1 : 0), "true mismatch");
  }

  ~QsmiAcceleratorDevice() {
    qsmi_device_destroy(_dev);  // before the TI can be freed below
    _dev = nullptr;
#pragma omp critical(createDevices)
    { _sharedMemoryManager->removeDevice(this); }
  }

  // SharedMemoryManager::createSharedMemoryForDevice: src holds one host
  // pointer per TI sub-image, point-reflected on the FFT grid
  // (DataImage::convertInput4Xcorr via src/qs.cpp:1544-1550).
  std::vector<g2s::spaceFrequencyMemoryAddress> allocAndInitSharedMemory(std::vector<void*> src,
                                                                         std::vector<unsigned> srcSize,
                                                                         std::vector<unsigned> fftSize) override {
    std::vector<const float*> images(src.size());
    for (size_t i = 0; i < src.size(); ++i) images[i] = static_cast<const float*>(src[i]);
    qsmi_ti* ti = nullptr;
    check(qsmi_ti_create_from_reflected((int)_deviceID, (int)srcSize.size(), srcSize.data(), fftSize.data(),
                                        (int)images.size(), images.data(), &ti),
          "TI spectra");
    g2s::spaceFrequencyMemoryAddress a;
    a.space = ti;
    a.fft = nullptr;
    return std::vector<g2s::spaceFrequencyMemoryAddress>(1, a);
  }

  std::vector<g2s::spaceFrequencyMemoryAddress> freeSharedMemory(
      std::vector<g2s::spaceFrequencyMemoryAddress> shared) override {
    if (!shared.empty() && shared[0].space) check(qsmi_ti_destroy(static_cast<qsmi_ti*>(shared[0].space)), "TI free");
    return std::vector<g2s::spaceFrequencyMemoryAddress>();
  }

  // The whole scan of one node (src/AcceleratorDevice.cpp:92-136: bounding
  // box, shifted offsets, computeFreqMismatchMap,
  // computeRealMissmatchAndRemoveWrongPattern) in one call; the offsets go
  // over unshifted, the device applies the bounding-box minimum itself.
  bool candidateForPattern(std::vector<std::vector<int> >& neighborArrayVector,
                           std::vector<std::vector<float> >& neighborValueArrayVector,
                           std::vector<float>& /*variablesCoefficient*/, std::vector<float> delta0) override {
    if (neighborValueArrayVector.empty()) return false;
    const size_t nb = neighborArrayVector.size(), D = _fftSize.size();
    const size_t nSub = neighborValueArrayVector[0].size();
    std::vector<int> off(nb * D);
    std::vector<float> val(nb * nSub);
    for (size_t j = 0; j < nb; ++j) {
      for (size_t d = 0; d < D; ++d) off[j * D + d] = neighborArrayVector[j][d];
      for (size_t s = 0; s < nSub; ++s) val[j * nSub + s] = neighborValueArrayVector[j][s];
    }
    int computed = 0;
    check(qsmi_candidate_for_pattern(_dev, (int)nb, off.data(), val.data(), delta0.data(), &computed), "scan");
    _hostMap.clear();
    _hostCross.clear();
    return computed != 0;
  }

  void searchKBiggest(float* errors, unsigned* encodedPosition, unsigned extendK, float seed) override {
    check(qsmi_search_k_biggest(_dev, errors, encodedPosition, extendK, seed), "k best");
  }

  unsigned cvtIndexToPosition(unsigned index) override {
    unsigned p = 0;
    check(qsmi_cvt_index_to_position(_dev, index, &p), "index map");
    return p;
  }
  unsigned cvtPositionToIndex(unsigned position) override {
    unsigned i = 0;
    check(qsmi_cvt_position_to_index(_dev, position, &i), "index map");
    return i;
  }

  unsigned getArraySize() override {
    unsigned s = 0;
    check(qsmi_get_array_size(_dev, &s), "array size");
    return s;
  }
  unsigned getErrorsArraySize() override { return getArraySize(); }

  // map value at an index of the reference's index space (the indices
  // searchKBiggest and cvtPositionToIndex give)
  float getErrorAtPosition(unsigned index) override {
    if (_hostMap.empty()) {
      size_t cells = 1;
      for (unsigned s : _srcSize) cells *= s;
      _hostMap.resize(cells);
      check(qsmi_get_error_map(_dev, _hostMap.data()), "map read");
    }
    return _hostMap[cvtIndexToPosition(index)];
  }
  float getValueAtPosition(unsigned layer, unsigned index) override {
    if (layer == 0) return getErrorAtPosition(index);
    if (layer == 1 && _crossMeasurement) return getCrossErrorAtPosition(index);
    unsupported("a coefficient layer beyond the map and the cross layer");
  }

  void setTrueMismatch(bool value) override {
    AcceleratorDevice::setTrueMismatch(value);
    check(qsmi_set_true_mismatch(_dev, value ? 1 : 0), "true mismatch");
  }

  void setValueInErrorArray(unsigned index, float value) override {
    check(qsmi_set_value_in_error_array(_dev, index, value), "map write");
    _hostMap.clear();
  }
  void setValueInErrorArrayWithRadius(unsigned index, float value, float radius) override {
    check(qsmi_set_value_in_error_array_with_radius(_dev, index, value, radius), "map write");
    _hostMap.clear();
  }

  // Steps of the base class's candidateForPattern, which this device
  // replaces as a whole.
  void zerosFrenquencySpaceOutputArray(unsigned) override { unsupported("zerosFrenquencySpaceOutputArray"); }
  void computeFreqMismatchMap(std::vector<std::vector<int> >, std::vector<std::vector<float> >&) override {
    unsupported("computeFreqMismatchMap");
  }
  void computeRealMissmatchAndRemoveWrongPattern(float*) override {
    unsupported("computeRealMissmatchAndRemoveWrongPattern");
  }

  // The cross-measurement layer of TIs with NaN (layer 1; the base class's
  // maskCroossError / maskCroossErrorWithVariable call maskLayerWithVariable
  // on it, src/AcceleratorDevice.cpp:136-145).
  float getCrossErrorAtPosition(unsigned index) override {
    if (_hostCross.empty()) {
      size_t cells = 1;
      for (unsigned s : _srcSize) cells *= s;
      _hostCross.resize(cells);
      check(qsmi_get_cross_map(_dev, _hostCross.data()), "cross layer read");
    }
    return _hostCross[cvtIndexToPosition(index)];
  }
  void maskLayerWithVariable(unsigned layer, unsigned variable) override {
    check(qsmi_mask_layer_with_variable(_dev, layer, variable), "mask layer");
    if (layer == 0) _hostMap.clear();
    else _hostCross.clear();
  }
  void compensateMissingData() override {
    check(qsmi_compensate_missing_data(_dev), "compensate missing data");
    _hostMap.clear();
  }

 private:
  // coefficient layer `layer` as qsmi's alpha[A + B*nSub]
  std::vector<float> layerCoefficients(size_t layer) {
    g2s::OperationMatrix& m = _coefficientMatrix.at(layer);
    const unsigned n = m.getNumberOfVariable();
    std::vector<float> a((size_t)n * n);
    for (unsigned i = 0; i < n; ++i)
      for (unsigned j = 0; j < n; ++j) a[i + (size_t)j * n] = m.getVariableAt((int)i, (int)j);
    return a;
  }

  qsmi_device* _dev = nullptr;
  std::vector<float> _hostMap;    // position-space copy of the map for getErrorAtPosition
  std::vector<float> _hostCross;  // the same for the cross layer
};

extern "C" {

void g2s_cudaGetDeviceCount(int* count) {
  if (qsmi_get_device_count(count) != QSMI_OK) *count = 0;
}

AcceleratorDevice* c_NvidiaGPUAcceleratorDevice(int deviceId, SharedMemoryManager* sharedMemoryManager,
                                                std::vector<g2s::OperationMatrix> coefficientMatrix,
                                                unsigned int threadRatio, bool withCrossMeasurement,
                                                bool circularTI) {
  return new QsmiAcceleratorDevice(deviceId, sharedMemoryManager, coefficientMatrix, threadRatio,
                                   withCrossMeasurement, circularTI);
}

}  // extern "C"
