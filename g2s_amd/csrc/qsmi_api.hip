// C-ABI of the MI355X QuickSampling exact scan (include/qsmi.h).
// Host side: TI spectra set-up, per-node and per-wave scans, error codes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <mutex>
#include <string>
#include <vector>

#include "qsmi_internal.hpp"

namespace qsmi {
hipError_t set_lds_limits(int device, size_t* limitOut);
size_t fft_lds_bytes_of(int L, int nvec);
void launch_r2c_rows(const float* in, float2* out, const FftPlan& plan, const float2* tw, int lognv, long long nrows,
                     int h0, hipStream_t st);
void launch_colfft(int dir, float2* base, const FftPlan& plan, const float2* tw, int lognv, long long rowStride,
                   int ncols, long long planes, int pdiv, long long pOuter, long long pInner, const NodeInfo* info,
                   int planesPerNode, int validDim, hipStream_t st);
void launch_pencil(const ScanGeom& g, const FftPlan& planP, int lognv, const float2* Yhat, const float2* tw0,
                   const float2* tw1, const float2* twP, int nNodes, int npb, const NodeInfo* info, const int* nbOff,
                   const float* nbVal, uint32_t* ord, float2* C, hipStream_t st);
void launch_direct1d(const ScanGeom& g, const float2* Yhat, const float2* tw0, int nNodes, const NodeInfo* info,
                     const int* nbOff, const float* nbVal, float2* C, hipStream_t st);
void launch_rows_select(const ScanGeom& g, const FftPlan& plan0, int lognv, const float2* tw0, const NodeInfo* info,
                        const float2* C, int nNodes, int rowBlocks, int K, Cand* cand, float* map, hipStream_t st,
                        int crossOff = 0, const uint8_t* centreOk = nullptr, int* gcnt = nullptr);
void launch_select_map(const float* map, long long S, int K, float quantum, uint32_t seedbits, Cand* cand, int chunks,
                       hipStream_t st);
void launch_fill(float* p, long long n, float v, hipStream_t st);
void launch_mask_layer(float* layer, const float* mask, long long n, hipStream_t st);
void launch_compensate(float* map, const float* cnt, long long n, float cmin, hipStream_t st);
void launch_fill_disk(float* map, const int* s, uint32_t o, float r2, float v, hipStream_t st);
long long ct1024_stride();
int rows1024_groups();
hipError_t spectra1024(const float* dSub, int s0, int s1, long long S, int nSub, const double2* twd, float2* YT,
                       hipStream_t st);
void launch_pencil1024(const ScanGeom& g, const float2* YT, const float2* tw, int nNodes, const NodeInfo* info,
                       const int* nbOff, const float* nbVal, TapInfo1024* taps, uint16_t* tapD0, float* tapVal,
                       float2* CT, const int* count, bool tapsBuilt, hipStream_t st);
void launch_rows1024(const ScanGeom& g, const float2* tw, const NodeInfo* info, const float2* CT, int nNodes, int K,
                     bool quant, Cand* cand, float* map, const int* count, int* gcnt, hipStream_t st);
}  // namespace qsmi

using namespace qsmi;

namespace qsmi {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

namespace {
std::mutex g_tune_mu;
ScanTune g_default_tune;  // copied into every TI at creation (qsmi_set_default_tuning)
}  // namespace

ScanTune default_tune() {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  return g_default_tune;
}

#define HIPCHK(expr)                                                                                  \
  do {                                                                                                \
    hipError_t _e = (expr);                                                                           \
    if (_e != hipSuccess)                                                                             \
      return fail(QSMI_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e));                  \
  } while (0)

unsigned size_rule(unsigned src) {
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

// Radix plan of one FFT length; prefers 10, 8, 4 (fewest LDS passes), no
// radix above radixMax unless the length leaves no other factor.
bool make_plan(int L, FftPlan* p, int radixMax) {
  p->L = L;
  p->nst = 0;
  int m = L;
  static const int pref[] = {10, 8, 4, 5, 3, 2, 7, 11, 13};
  while (m > 1) {
    int pick = 0;
    for (int r : pref)
      if (r <= radixMax && m % r == 0) {
        pick = r;
        break;
      }
    if (!pick)
      for (int r : pref)
        if (m % r == 0) {
          pick = r;
          break;
        }
    if (!pick || p->nst >= kMaxStages) return false;
    p->radix[p->nst++] = pick;
    m /= pick;
  }
  if (L == 1) {  // a length-1 transform is the identity: one radix-1 stage is not needed
    p->nst = 0;
  }
  return true;
}

// largest power-of-two vector count with L*nvec <= kMaxVecElems and LDS <= 64 KiB,
// at most 2^lognvMax
int choose_lognv(int L, int lognvMax) {
  int lognv = 0;
  while (lognv < lognvMax) {
    int nv = 1 << (lognv + 1);
    if ((long long)L * nv > kMaxVecElems) break;
    if ((size_t)(L + 2 * L * nv) * sizeof(float2) > 64 * 1024) break;
    if (nv > 64) break;
    lognv++;
  }
  return lognv;
}

unsigned seed_from_unit_float(float seed) {
  if (!std::isfinite(seed) || seed <= 0.f) return 0;
  if (seed >= 1.f) return std::numeric_limits<unsigned>::max();
  return (unsigned)std::floor((double)seed * (double)std::numeric_limits<unsigned>::max());
}

template <typename T>
hipError_t dalloc(T** p, size_t count) {
  return hipMalloc((void**)p, std::max<size_t>(count, 1) * sizeof(T));
}

}  // namespace qsmi


// Scratch of qsmi_search_k_biggest_noalloc: one node's candidate lists, its
// K best, its node record and the exact selection's buffers. Every GPU the
// library runs on holds one copy until the process ends; its address on a
// GPU is taken once (hipGetSymbolAddress with that GPU current, in
// qsmi_ti_create) and used for every access from the host.
constexpr size_t kScratchBytes = 256 * 1024;
constexpr int kScratchChunks = 256;  // blocks of the FP32 top-K over the map
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace qsmi {

namespace {
std::mutex g_scratchMu;
std::vector<void*> g_scratchPtr;  // per GPU ordinal
}  // namespace

// The scratch address on GPU `gpu`, which must be the current device.
int scratch_address(int gpu, void** out) {
  std::lock_guard<std::mutex> lock(g_scratchMu);
  if (gpu < 0) return fail(QSMI_ERR_ARG, "negative GPU ordinal");
  if ((size_t)gpu >= g_scratchPtr.size()) g_scratchPtr.resize((size_t)gpu + 1, nullptr);
  if (!g_scratchPtr[gpu]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(QSMI_ERR_HIP, std::string("scratch address: ") + hipGetErrorString(e));
    g_scratchPtr[gpu] = p;
  }
  *out = g_scratchPtr[gpu];
  return QSMI_OK;
}

ScanGeom geom_of(const qsmi_device* dev) {
  const qsmi_ti* t = dev->ti;
  ScanGeom g;
  memset(&g, 0, sizeof(g));
  g.D = t->D;
  for (int i = 0; i < 3; ++i) {
    g.s[i] = t->s[i];
    g.f[i] = t->f[i];
  }
  g.h0 = t->h0;
  g.P = t->D - 1;
  g.V = (int)dev->active.size();
  for (int a = 0; a < g.V; ++a) g.A[a] = dev->active[a];
  g.circular = dev->circular;
  g.invN = (float)(1.0 / (double)t->N);
  g.H = t->H;
  return g;
}

// the device scans on the 1024 x 1024 grid (wavefront transforms)
bool use_w1k(const qsmi_device* dev) { return dev->ti->w1k && !dev->circular && !dev->cross; }

// cells of the grid the device's transforms run on
double scan_cells(const qsmi_device* dev) { return use_w1k(dev) ? 1024.0 * 1024.0 : (double)dev->ti->N; }

// map = raw * scale + delta0 for the unnormalised correlation raw on the scan
// grid: the true mismatch (trueMismatch) or N (E - delta0) with N the cells
// of the TI's FFT grid (the reference's unnormalised CPU map)
float map_scale(const qsmi_device* dev) {
  return (float)((dev->trueMismatch ? 1.0 : (double)dev->ti->N) / scan_cells(dev));
}

// rows per candidate block; 0: the merge reads every block (tiles write -inf
// candidates when they hold no valid row)
int rows_per_block(const qsmi_device* dev) { return use_w1k(dev) ? 0 : (2 << dev->ti->lognv[0]); }

int row_blocks(const qsmi_device* dev) {
  if (use_w1k(dev)) return rows1024_groups();
  const qsmi_ti* t = dev->ti;
  long long rows = (long long)t->s[1] * t->s[2];
  return (int)((rows + rows_per_block(dev) - 1) / rows_per_block(dev));
}

// spectrum-stage intermediate per node (float2 elements)
long long c_stride(const qsmi_device* dev) {
  return use_w1k(dev) ? std::max<long long>(dev->ti->H, ct1024_stride()) : dev->ti->H;
}

// Node preparation on the host: bounding box, valid range, alpha-combined
// values v_ja = sum_B alpha(A_a, B) u_jB.
void prep_node(const qsmi_device* dev, int nb, const int* off, const float* u, float delta0, float seed,
               int nbStart, NodeInfo* ni, int* offOut, float* valOut, int* mnOut, int* mxOut) {
  const qsmi_ti* t = dev->ti;
  const int D = t->D, nSub = t->nSub, V = (int)dev->active.size();
  int mn[3] = {0, 0, 0}, mx[3] = {0, 0, 0};
  for (int j = 0; j < nb; ++j)
    for (int i = 0; i < D; ++i) {
      mn[i] = std::min(mn[i], off[j * D + i]);
      mx[i] = std::max(mx[i], off[j * D + i]);
    }
  for (int i = 0; i < 3; ++i) {
    if (i >= D) {
      ni->lo[i] = ni->hi[i] = 0;
    } else if (dev->circular) {
      ni->lo[i] = 0;
      ni->hi[i] = t->s[i] - 1;
    } else {
      ni->lo[i] = -mn[i];
      ni->hi[i] = t->s[i] - 1 - mx[i];
    }
  }
  if (mnOut)
    for (int i = 0; i < 3; ++i) {
      mnOut[i] = mn[i];
      mxOut[i] = mx[i];
    }
  ni->nbStart = nbStart;
  ni->nb = nb;
  ni->scale = map_scale(dev);
  ni->delta0 = dev->trueMismatch ? delta0 : 0.f;
  ni->seed = seed;
  ni->seedbits = seed_from_unit_float(seed);
  ni->quantum = dev->quantum;
  ni->cell = 0;
  ni->cmin = 0.f;
  ni->exOn = 0;
  ni->exPos = 0;
  ni->exR2 = 0.f;
  ni->kLocal = 0.f;
  ni->tiReq = -1;
  std::vector<double> sumAbs(V, 0.0);
  for (int j = 0; j < nb; ++j) {
    for (int i = 0; i < D; ++i) offOut[j * D + i] = off[j * D + i];
    for (int a = 0; a < V; ++a) {
      const int A = dev->active[a];
      float acc = 0.f;
      for (int B = 0; B < nSub; ++B) acc += dev->alpha[A + B * nSub] * u[j * nSub + B];
      valOut[j * V + a] = acc;
      sumAbs[a] += std::fabs((double)acc);
    }
  }
  ni->eps = 0.f;
  if (dev->quantum == 0.f && !dev->cross && nb > 0) {
    double e = 0.0;
    for (int a = 0; a < V; ++a) e += (double)t->rms[dev->active[a]] * sumAbs[a];
    ni->eps = (float)((double)kMapErrC * (double)ni->scale * scan_cells(dev) * e);
  }
}

template <typename T>
int ensure(T** p, size_t* cap, size_t need) {
  if (need <= *cap) return QSMI_OK;
  if (*p) hipFree(*p);
  *p = nullptr;
  size_t n = std::max(need, *cap * 3 / 2);
  HIPCHK(dalloc(p, n));
  *cap = n;
  return QSMI_OK;
}

// Launch the spectrum stage and the dim1 inverse (3-D) for nNodes nodes.
int ensure_tap_tables(qsmi_device* dev, int nNodes, int V) {
  if (dev->capTap >= (size_t)nNodes && dev->capTapV >= V) return QSMI_OK;
  // grown rarely: hipFree waits for queued work that may still read them
  if (dev->dTap) (void)hipFree(dev->dTap);
  if (dev->dTapD0) (void)hipFree(dev->dTapD0);
  if (dev->dTapVal) (void)hipFree(dev->dTapVal);
  dev->dTap = nullptr;
  dev->dTapD0 = nullptr;
  dev->dTapVal = nullptr;
  dev->capTap = 0;
  dev->capTapV = 0;
  const size_t cap = std::max<size_t>((size_t)nNodes, 256);
  const int capV = std::max(V, 2);
  if (hipMalloc(&dev->dTap, cap * sizeof(TapInfo1024)) != hipSuccess ||
      hipMalloc(&dev->dTapD0, cap * kTapNbMax * sizeof(uint16_t)) != hipSuccess ||
      hipMalloc(&dev->dTapVal, cap * kTapNbMax * capV * sizeof(float)) != hipSuccess)
    return fail(QSMI_ERR_NOMEM, "tap tables: out of device memory");
  dev->capTap = cap;
  dev->capTapV = capV;
  return QSMI_OK;
}

int launch_spectrum(qsmi_device* dev, const ScanGeom& g, int nNodes, const NodeInfo* info, const int* off,
                    const float* val, uint32_t* ord, float2* C, const int* count, bool tapsBuilt) {
  qsmi_ti* t = dev->ti;
  hipStream_t st = dev->stream;
  if (dev->timing) hipEventRecord(dev->ev[0], st);
  if (use_w1k(dev)) {
    int rc = ensure_tap_tables(dev, nNodes, g.V);
    if (rc) return rc;
    launch_pencil1024(g, t->dYT, t->dTw1024, nNodes, info, off, val, dev->dTap, dev->dTapD0, dev->dTapVal, C, count,
                      tapsBuilt, st);
  } else if (count) {
    return fail(QSMI_ERR_UNSUPPORTED, "a device node count is supported on the 1024-grid path only");
  } else if (t->D == 1) {
    launch_direct1d(g, t->dYhat, t->dTw[0], nNodes, info, off, val, C, st);
  } else {
    const int P = t->D - 1;
    launch_pencil(g, t->plan[P], t->lognv[P], t->dYhat, t->dTw[0], t->dTw[1], t->dTw[P], nNodes, t->tune.pencilNpb,
                  info, off, val, ord, C, st);
  }
  if (dev->timing) hipEventRecord(dev->ev[1], st);
  if (t->D == 3) {
    // inverse along dim1 of the valid c2 planes: plane p = node*f2 + c2
    launch_colfft(+1, C, t->plan[1], t->dTw[1], t->lognv[1], t->h0, t->h0, (long long)nNodes * t->f[2], t->f[2],
                  t->H, (long long)t->f[1] * t->h0, info, t->f[2], 2, st);
  }
  if (dev->timing) hipEventRecord(dev->ev[2], st);
  return QSMI_OK;
}

void launch_rows(qsmi_device* dev, const ScanGeom& g, const NodeInfo* info, const float2* C, int nNodes, int rowBlocks,
                 int K, Cand* cand, int* gcnt, float* map, const int* count, int crossOff) {
  qsmi_ti* t = dev->ti;
  if (use_w1k(dev))
    launch_rows1024(g, t->dTw1024, info, C, nNodes, K, dev->quantum > 0.f, cand, map, count, gcnt, dev->stream);
  else
    launch_rows_select(g, t->plan[0], t->lognv[0], t->dTw[0], info, C, nNodes, rowBlocks, K, cand, map, dev->stream,
                       crossOff, dev->dCentreOk, gcnt);
}

int set_cross(qsmi_device* dev, const float* alphaX, const uint8_t* centreOk) {
  qsmi_ti* t = dev->ti;
  const int nSub = t->nSub;
  dev->alphaX.assign(alphaX, alphaX + nSub * nSub);
  for (int A = 0; A < nSub; ++A) {  // the count layer's images are scanned too
    bool need = false;
    for (int B = 0; B < nSub; ++B) need |= alphaX[A + B * nSub] != 0.f;
    if (need && std::find(dev->active.begin(), dev->active.end(), A) == dev->active.end()) dev->active.push_back(A);
  }
  std::sort(dev->active.begin(), dev->active.end());
  std::vector<uint32_t> complete;
  for (long long q = 0; q < t->S; ++q)
    if (centreOk[q]) complete.push_back((uint32_t)q);
  dev->nComplete = (uint32_t)complete.size();
  HIPCHK(hipSetDevice(t->gpu));
  HIPCHK(dalloc(&dev->dCentreOk, (size_t)t->S));
  HIPCHK(dalloc(&dev->dComplete, std::max<size_t>(complete.size(), 1)));
  HIPCHK(hipMemcpy(dev->dCentreOk, centreOk, (size_t)t->S, hipMemcpyHostToDevice));
  if (!complete.empty())
    HIPCHK(hipMemcpy(dev->dComplete, complete.data(), complete.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
  dev->cross = true;  // the two-layer rows stage exists on the generic path (use_w1k is false)
  return QSMI_OK;
}

int check_dev(qsmi_device* dev) {
  if (!dev || !dev->ti) return fail(QSMI_ERR_ARG, "null device");
  HIPCHK(hipSetDevice(dev->ti->gpu));
  return QSMI_OK;
}

}  // namespace qsmi

extern "C" {

// error channel shared with the host layer (not part of the public header)
__attribute__((visibility("hidden"))) int qsmi_internal_fail(int code, const char* msg) { return fail(code, msg); }

const char* qsmi_last_error(void) { return g_err.c_str(); }

int qsmi_get_device_count(int* count) {
  if (!count) return fail(QSMI_ERR_ARG, "null count");
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  if (e != hipSuccess) c = 0;
  *count = c;
  return QSMI_OK;
}

unsigned qsmi_fft_size(unsigned src) { return size_rule(src); }

int qsmi_ti_create(int gpu, int ndim, const unsigned* srcSize, const unsigned* fftSize, int nSub,
                   const float* subImages, qsmi_ti** out) {
  if (!out || !srcSize || !subImages) return fail(QSMI_ERR_ARG, "null argument");
  *out = nullptr;
  if (ndim < 1 || ndim > 3) return fail(QSMI_ERR_UNSUPPORTED, "ndim must be 1, 2 or 3");
  if (nSub < 1 || nSub > kMaxSub) return fail(QSMI_ERR_UNSUPPORTED, "nSub must be in [1, 16]");
  auto* t = new qsmi_ti();
  t->tune = default_tune();
  t->gpu = gpu;
  t->D = ndim;
  for (int i = 0; i < ndim; ++i) {
    if (srcSize[i] < 1) {
      delete t;
      return fail(QSMI_ERR_ARG, "zero TI extent");
    }
    t->s[i] = (int)srcSize[i];
    t->f[i] = fftSize ? (int)fftSize[i] : (int)size_rule(srcSize[i]);
    if (t->f[i] < t->s[i]) {
      delete t;
      return fail(QSMI_ERR_ARG, "fftSize smaller than srcSize");
    }
    if (t->f[i] > 4096) {
      delete t;
      return fail(QSMI_ERR_UNSUPPORTED, "FFT extent above 4096");
    }
  }
  if (ndim >= 2 && t->f[0] < 2) {
    delete t;
    return fail(QSMI_ERR_UNSUPPORTED, "dimension 0 of extent 1 in a multi-dimensional TI");
  }
  t->h0 = t->f[0] / 2 + 1;
  t->S = (long long)t->s[0] * t->s[1] * t->s[2];
  t->N = (long long)t->f[0] * t->f[1] * t->f[2];
  t->H = (long long)t->h0 * t->f[1] * t->f[2];
  t->nSub = nSub;
  for (int i = 0; i < 3; ++i) {
    if (!make_plan(t->f[i], &t->plan[i], t->tune.radixMax)) {
      delete t;
      return fail(QSMI_ERR_UNSUPPORTED, "FFT extent with a prime factor above 13");
    }
    t->lognv[i] = choose_lognv(t->f[i], t->tune.lognvMax);
  }
  hipError_t e = hipSetDevice(gpu);
  if (e != hipSuccess) {
    delete t;
    return fail(QSMI_ERR_HIP, std::string("hipSetDevice: ") + hipGetErrorString(e));
  }
  void* scratchPtr = nullptr;  // once per GPU, with the GPU current
  if (scratch_address(gpu, &scratchPtr) != QSMI_OK) {
    delete t;
    return QSMI_ERR_HIP;
  }
  size_t ldsLimit = 0;
  e = set_lds_limits(gpu, &ldsLimit);
  if (e != hipSuccess) {
    delete t;
    return fail(QSMI_ERR_HIP, std::string("LDS limit: ") + hipGetErrorString(e));
  }
  for (int i = 0; i < 3; ++i)
    if (fft_lds_bytes_of(t->f[i], 1 << t->lognv[i]) > ldsLimit) {
      delete t;
      return fail(QSMI_ERR_UNSUPPORTED, "FFT extent too large for one LDS block");
    }
#define TCHK(expr)                                                                  \
  do {                                                                              \
    hipError_t _e = (expr);                                                         \
    if (_e != hipSuccess) {                                                         \
      qsmi_ti_destroy(t);                                                           \
      return fail(QSMI_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    }                                                                               \
  } while (0)
  TCHK(hipStreamCreateWithFlags(&t->stream, hipStreamNonBlocking));
  // twiddle tables, FP64-generated: tw[k] = exp(-2 pi i k / f)
  for (int i = 0; i < 3; ++i) {
    std::vector<float2> tw(t->f[i]);
    for (int k = 0; k < t->f[i]; ++k) {
      long double a = -6.283185307179586476925286766559L * (long double)k / (long double)t->f[i];
      tw[k] = make_float2((float)cosl(a), (float)sinl(a));
    }
    TCHK(dalloc(&t->dTw[i], tw.size()));
    TCHK(hipMemcpy(t->dTw[i], tw.data(), tw.size() * sizeof(float2), hipMemcpyHostToDevice));
  }
  // zero-padded real images on the f-grid
  std::vector<float> pad((size_t)nSub * t->N, 0.f);
  for (int A = 0; A < nSub; ++A)
    for (int c2 = 0; c2 < t->s[2]; ++c2)
      for (int c1 = 0; c1 < t->s[1]; ++c1) {
        const float* src = subImages + (size_t)A * t->S + ((size_t)c2 * t->s[1] + c1) * t->s[0];
        float* dst = pad.data() + (size_t)A * t->N + ((size_t)c2 * t->f[1] + c1) * t->f[0];
        memcpy(dst, src, sizeof(float) * t->s[0]);
      }
  // root mean square of every sub-image over the FFT grid: the scale of the
  // FP32 map error bound
  for (int A = 0; A < nSub; ++A) {
    double acc = 0.0;
    const float* src = subImages + (size_t)A * t->S;
    for (long long q = 0; q < t->S; ++q) acc += (double)src[q] * (double)src[q];
    t->rms[A] = (float)std::sqrt(acc / (double)t->N);
  }
  // spatial copy on the TI grid (exact rescoring of band and proposed positions)
  TCHK(dalloc(&t->dSub, (size_t)nSub * t->S));
  TCHK(hipMemcpy(t->dSub, subImages, (size_t)nSub * t->S * sizeof(float), hipMemcpyHostToDevice));
  float* dReal = nullptr;
  TCHK(dalloc(&dReal, pad.size()));
  TCHK(hipMemcpy(dReal, pad.data(), pad.size() * sizeof(float), hipMemcpyHostToDevice));
  TCHK(dalloc(&t->dYhat, (size_t)nSub * t->H));
  // rows r2c along dim0: rows (A, k2, k1)
  const long long nrows = (long long)nSub * t->f[1] * t->f[2];
  launch_r2c_rows(dReal, t->dYhat, t->plan[0], t->dTw[0], t->lognv[0], nrows, t->h0, t->stream);
  TCHK(hipGetLastError());
  if (ndim >= 2) {  // along dim1: planes (A, k2), stride f1*h0
    launch_colfft(-1, t->dYhat, t->plan[1], t->dTw[1], t->lognv[1], t->h0, t->h0, (long long)nSub * t->f[2], 1,
                  (long long)t->f[1] * t->h0, 0, nullptr, 1, 0, t->stream);
    TCHK(hipGetLastError());
  }
  if (ndim == 3) {  // along dim2: planes (A, k1)
    launch_colfft(-1, t->dYhat, t->plan[2], t->dTw[2], t->lognv[2], (long long)t->f[1] * t->h0, t->h0,
                  (long long)nSub * t->f[1], t->f[1], (long long)t->f[2] * t->f[1] * t->h0, t->h0, nullptr, 1, 0,
                  t->stream);
    TCHK(hipGetLastError());
  }
  // the 1024 x 1024 grid holds any 2-D TI of extents up to 1024 (valid
  // positions never wrap); below grid1024Min the TI's own FFT grid is used
  const int g1k = t->tune.grid1024Min;
  t->w1k = (ndim == 2 && t->s[0] >= g1k && t->s[1] >= g1k && t->s[0] <= 1024 && t->s[1] <= 1024);
  if (t->w1k) {
    std::vector<float2> tw(1024);
    std::vector<double2> twd(1024);
    for (int i = 0; i < 1024; ++i) {
      const double a = -2.0 * M_PI * (double)i / 1024.0;
      twd[i] = make_double2(std::cos(a), std::sin(a));
      tw[i] = make_float2((float)twd[i].x, (float)twd[i].y);
    }
    TCHK(dalloc(&t->dTw1024, 1024));
    TCHK(hipMemcpy(t->dTw1024, tw.data(), 1024 * sizeof(float2), hipMemcpyHostToDevice));
    double2* dTwd = nullptr;
    TCHK(hipMalloc(&dTwd, 1024 * sizeof(double2)));
    TCHK(hipMemcpy(dTwd, twd.data(), 1024 * sizeof(double2), hipMemcpyHostToDevice));
    TCHK(dalloc(&t->dYT, (size_t)nSub * 513 * 1024));
    const hipError_t se = spectra1024(t->dSub, t->s[0], t->s[1], t->S, nSub, dTwd, t->dYT, t->stream);
    (void)hipFree(dTwd);
    TCHK(se);
  }
  TCHK(hipStreamSynchronize(t->stream));
  (void)hipFree(dReal);
#undef TCHK
  *out = t;
  return QSMI_OK;
}

int qsmi_ti_destroy(qsmi_ti* t) {
  if (!t) return QSMI_OK;
  if (t->refs > 0) return fail(QSMI_ERR_ARG, "TI still used by a device");
  hipSetDevice(t->gpu);
  if (t->stream) hipStreamSynchronize(t->stream);
  if (t->dYhat) (void)hipFree(t->dYhat);
  if (t->dYT) (void)hipFree(t->dYT);
  if (t->dTw1024) (void)hipFree(t->dTw1024);
  if (t->dSub) (void)hipFree(t->dSub);
  for (int i = 0; i < 3; ++i)
    if (t->dTw[i]) hipFree(t->dTw[i]);
  if (t->stream) hipStreamDestroy(t->stream);
  delete t;
  return QSMI_OK;
}

namespace {
struct TuneKey {
  const char* name;
  int ScanTune::*field;
  int lo, hi;
};
const TuneKey kTuneKeys[] = {
    {"pencil.npb", &ScanTune::pencilNpb, 1, 16},
    {"scan.grid1024_min", &ScanTune::grid1024Min, 2, 1025},
    {"fft.radix_max", &ScanTune::radixMax, 2, 13},
    {"fft.lognv_max", &ScanTune::lognvMax, 0, 6},
    {"loop.graph", &ScanTune::loopGraph, 0, 1},
};
const TuneKey* find_tune_key(const char* key) {
  if (!key) return nullptr;
  for (const TuneKey& k : kTuneKeys)
    if (std::strcmp(k.name, key) == 0) return &k;
  return nullptr;
}
}  // namespace

namespace {
int set_tune_field(ScanTune* tune, const char* key, int value) {
  const TuneKey* k = find_tune_key(key);
  if (!k) return fail(QSMI_ERR_ARG, std::string("unknown tuning key: ") + (key ? key : "(null)"));
  if (value < k->lo || value > k->hi)
    return fail(QSMI_ERR_ARG, std::string("tuning value out of range for ") + k->name);
  tune->*(k->field) = value;
  return QSMI_OK;
}
}  // namespace

int qsmi_ti_set_tuning(qsmi_ti* ti, const char* key, int value) {
  if (!ti) return fail(QSMI_ERR_ARG, "null TI");
  return set_tune_field(&ti->tune, key, value);
}

int qsmi_set_default_tuning(const char* key, int value) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  return set_tune_field(&g_default_tune, key, value);
}

int qsmi_ti_get_tuning(const qsmi_ti* ti, const char* key, int* value) {
  if (!ti || !value) return fail(QSMI_ERR_ARG, "null argument");
  const TuneKey* k = find_tune_key(key);
  if (!k) return fail(QSMI_ERR_ARG, std::string("unknown tuning key: ") + (key ? key : "(null)"));
  *value = ti->tune.*(k->field);
  return QSMI_OK;
}

int qsmi_device_create(qsmi_ti* ti, const float* alpha, int circularTI, qsmi_device** out) {
  if (!ti || !alpha || !out) return fail(QSMI_ERR_ARG, "null argument");
  *out = nullptr;
  if (circularTI)
    for (int i = 0; i < ti->D; ++i)
      if (ti->f[i] != ti->s[i]) return fail(QSMI_ERR_UNSUPPORTED, "circular TI needs fftSize == srcSize");
  auto* d = new qsmi_device();
  d->ti = ti;
  const int nSub = ti->nSub;
  d->alpha.assign(alpha, alpha + nSub * nSub);
  for (int A = 0; A < nSub; ++A) {
    bool need = false;
    for (int B = 0; B < nSub; ++B) need |= (alpha[A + B * nSub] != 0.f);
    if (need) d->active.push_back(A);
  }
  d->circular = circularTI;
  hipError_t e = hipSetDevice(ti->gpu);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = dalloc(&d->dMap, (size_t)ti->S);
  for (int i = 0; i < 5 && e == hipSuccess; ++i) e = hipEventCreate(&d->ev[i]);
  if (e != hipSuccess) {
    qsmi_device_destroy(d);
    return fail(QSMI_ERR_HIP, std::string("device create: ") + hipGetErrorString(e));
  }
  ti->refs++;
  d->registered = true;
  *out = d;
  return QSMI_OK;
}

int qsmi_device_destroy(qsmi_device* d) {
  if (!d) return QSMI_OK;
  if (d->ti) hipSetDevice(d->ti->gpu);
  if (d->stream) hipStreamSynchronize(d->stream);
  hipFree(d->dMap);
  hipFree(d->dC);
  hipFree(d->dTap);
  hipFree(d->dTapD0);
  hipFree(d->dTapVal);
  hipFree(d->dCand);
  hipFree(d->dInfo);
  hipFree(d->dOff);
  hipFree(d->dVal);
  hipFree(d->dOrd);
  hipFree(d->dOutVal);
  hipFree(d->dOutPos);
  hipFree(d->dChosen);
  hipFree(d->dCentreOk);
  hipFree(d->dComplete);
  band_free(&d->band);
  if (d->wb) {
    free_wave_bufs(d->wb);
    delete d->wb;
  }
  if (d->side.stream) hipStreamSynchronize(d->side.stream);
  if (d->side.wb) {
    free_wave_bufs(d->side.wb);
    delete d->side.wb;
  }
  hipFree(d->side.dTap);
  hipFree(d->side.dTapD0);
  hipFree(d->side.dTapVal);
  if (d->side.stream) hipStreamDestroy(d->side.stream);
  for (int i = 0; i < 5; ++i)
    if (d->ev[i]) (void)hipEventDestroy(d->ev[i]);
  for (hipEvent_t ev : d->evPool) (void)hipEventDestroy(ev);
  if (d->stream) hipStreamDestroy(d->stream);
  if (d->ti && d->registered) d->ti->refs--;
  delete d;
  return QSMI_OK;
}

int qsmi_set_true_mismatch(qsmi_device* dev, int value) {
  if (!dev) return fail(QSMI_ERR_ARG, "null device");
  dev->trueMismatch = value ? 1 : 0;
  return QSMI_OK;
}

int qsmi_set_tie_quantum(qsmi_device* dev, float quantum) {
  if (!dev || !(quantum >= 0.f)) return fail(QSMI_ERR_ARG, "invalid quantum");
  dev->quantum = quantum;
  return QSMI_OK;
}

int qsmi_set_stage_timing(qsmi_device* dev, int enable) {
  if (!dev) return fail(QSMI_ERR_ARG, "null device");
  dev->timing = enable != 0;
  return QSMI_OK;
}

int qsmi_last_stage_ms(qsmi_device* dev, float* out4) {
  if (!dev || !out4) return fail(QSMI_ERR_ARG, "null argument");
  for (int i = 0; i < 4; ++i) out4[i] = dev->stageMs[i];
  return QSMI_OK;
}

int qsmi_synchronize(qsmi_device* dev) {
  int rc = check_dev(dev);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(dev->stream));
  return QSMI_OK;
}

int qsmi_candidate_for_pattern(qsmi_device* dev, int nb, const int* offsets, const float* values,
                               const float* delta0, int* computed) {
  int rc = check_dev(dev);
  if (rc) return rc;
  if (computed) *computed = 0;
  if (nb < 0) return fail(QSMI_ERR_ARG, "negative neighbour count");
  if (nb == 0) return QSMI_OK;
  if (!offsets || !values || !delta0) return fail(QSMI_ERR_ARG, "null argument");
  qsmi_ti* t = dev->ti;
  const int V = (int)dev->active.size(), D = t->D, nSub = t->nSub;
  // with the cross layer the count layer is a second node (same neighbours,
  // alphaX-combined values) whose map follows the first one's
  const int layers = dev->xLayer ? 2 : 1;
  NodeInfo ni[2];
  std::vector<int> off((size_t)layers * nb * D);
  std::vector<float> val((size_t)layers * nb * std::max(V, 1));
  prep_node(dev, nb, offsets, values, delta0[0], 0.f, 0, &ni[0], off.data(), val.data(), dev->mn, dev->mx);
  if (dev->xLayer) {
    // the reference's devices add no delta0 with the cross layer
    // (src/CPUThreadAcceleratorDevice.cpp:348), and the count's division
    // leaves no exact-selection bound
    ni[0].delta0 = 0.f;
    ni[0].eps = 0.f;
    ni[1] = ni[0];
    ni[1].nbStart = nb;
    float wmin = INFINITY;
    for (int j = 0; j < nb; ++j) {
      for (int d = 0; d < D; ++d) off[(size_t)(nb + j) * D + d] = offsets[j * D + d];
      for (int a = 0; a < V; ++a) {
        const int A = dev->active[a];
        float acc = 0.f;
        for (int B = 0; B < nSub; ++B) acc += dev->alphaX[A + B * nSub] * values[j * nSub + B];
        val[(size_t)(nb + j) * V + a] = acc;
        if (acc > 0.f) wmin = std::min(wmin, acc);
      }
    }
    dev->xCmin = wmin < INFINITY ? (float)(0.5 * wmin * (dev->trueMismatch ? 1.0 : (double)t->N)) : 0.f;
  }
  if ((rc = ensure(&dev->dInfo, &dev->capInfo, (size_t)layers))) return rc;
  if ((rc = ensure(&dev->dOff, &dev->capOff, off.size()))) return rc;
  if ((rc = ensure(&dev->dVal, &dev->capVal, val.size()))) return rc;
  if ((rc = ensure(&dev->dOrd, &dev->capOrd, (size_t)layers * nb))) return rc;
  if ((rc = ensure(&dev->dC, &dev->capC, (size_t)layers * c_stride(dev)))) return rc;
  hipStream_t st = dev->stream;
  HIPCHK(hipMemcpyAsync(dev->dInfo, ni, layers * sizeof(NodeInfo), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(dev->dOff, off.data(), off.size() * sizeof(int), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(dev->dVal, val.data(), val.size() * sizeof(float), hipMemcpyHostToDevice, st));
  launch_fill(dev->dMap, layers * t->S, -INFINITY, st);
  ScanGeom g = geom_of(dev);
  if ((rc = launch_spectrum(dev, g, layers, dev->dInfo, dev->dOff, dev->dVal, dev->dOrd, dev->dC))) return rc;
  launch_rows(dev, g, dev->dInfo, dev->dC, layers, row_blocks(dev), 1, nullptr, nullptr, dev->dMap);
  HIPCHK(hipGetLastError());
  // the host arrays above are pageable and go out of scope: finish the copies
  HIPCHK(hipStreamSynchronize(st));
  dev->mapValid = true;
  dev->mapEdited = false;
  dev->lastNode = ni[0];
  if (computed) *computed = 1;
  return QSMI_OK;
}

int qsmi_device_set_cross_layer(qsmi_device* dev, const float* alphaX) {
  int rc = check_dev(dev);
  if (rc) return rc;
  if (!alphaX) return fail(QSMI_ERR_ARG, "null argument");
  if (dev->cross) return fail(QSMI_ERR_ARG, "device already runs the in-memory cross form");
  qsmi_ti* t = dev->ti;
  const int nSub = t->nSub;
  bool any = false;
  for (int q = 0; q < nSub * nSub; ++q) {
    if (!std::isfinite(alphaX[q])) return fail(QSMI_ERR_ARG, "cross layer: non-finite coefficient");
    any |= alphaX[q] != 0.f;
  }
  if (!any) return fail(QSMI_ERR_ARG, "cross layer: all coefficients zero");
  float* m2 = nullptr;
  HIPCHK(dalloc(&m2, (size_t)2 * t->S));
  hipFree(dev->dMap);
  dev->dMap = m2;
  dev->alphaX.assign(alphaX, alphaX + nSub * nSub);
  for (int A = 0; A < nSub; ++A) {  // the count layer's images are scanned too
    bool need = false;
    for (int B = 0; B < nSub; ++B) need |= alphaX[A + B * nSub] != 0.f;
    if (need && std::find(dev->active.begin(), dev->active.end(), A) == dev->active.end()) dev->active.push_back(A);
  }
  std::sort(dev->active.begin(), dev->active.end());
  dev->xLayer = true;
  dev->mapValid = false;
  return QSMI_OK;
}

int qsmi_mask_layer_with_variable(qsmi_device* dev, unsigned layer, unsigned variable) {
  int rc = check_dev(dev);
  if (rc) return rc;
  if (!dev->xLayer) return fail(QSMI_ERR_ARG, "no cross layer: call qsmi_device_set_cross_layer first");
  if (layer > 1) return fail(QSMI_ERR_ARG, "layer must be 0 (map) or 1 (count layer)");
  if (!dev->mapValid) return fail(QSMI_ERR_ARG, "no map: call qsmi_candidate_for_pattern first");
  // the variable-th sub-variable the count layer reads (a row of alphaX with
  // a non-zero coefficient, src/CPUThreadAcceleratorDevice.cpp:366-376)
  const qsmi_ti* t = dev->ti;
  const int nSub = t->nSub;
  int sub = -1, left = (int)variable;
  for (int A = 0; A < nSub && sub < 0; ++A) {
    bool need = false;
    for (int B = 0; B < nSub; ++B) need |= dev->alphaX[A + B * nSub] != 0.f;
    if (need && left-- == 0) sub = A;
  }
  if (sub < 0) return fail(QSMI_ERR_ARG, "mask variable out of range");
  launch_mask_layer(dev->dMap + (size_t)layer * t->S, t->dSub + (size_t)sub * t->S, t->S, dev->stream);
  HIPCHK(hipGetLastError());
  if (layer == 0) dev->mapEdited = true;
  return QSMI_OK;
}

int qsmi_compensate_missing_data(qsmi_device* dev) {
  int rc = check_dev(dev);
  if (rc) return rc;
  if (!dev->xLayer) return fail(QSMI_ERR_ARG, "no cross layer: call qsmi_device_set_cross_layer first");
  if (!dev->mapValid) return fail(QSMI_ERR_ARG, "no map: call qsmi_candidate_for_pattern first");
  const long long S = dev->ti->S;
  launch_compensate(dev->dMap, dev->dMap + S, S, dev->xCmin, dev->stream);
  HIPCHK(hipGetLastError());
  dev->mapEdited = true;
  return QSMI_OK;
}

int qsmi_get_cross_map(qsmi_device* dev, float* out) {
  int rc = check_dev(dev);
  if (rc) return rc;
  if (!out) return fail(QSMI_ERR_ARG, "null argument");
  if (!dev->xLayer) return fail(QSMI_ERR_ARG, "no cross layer: call qsmi_device_set_cross_layer first");
  if (!dev->mapValid) return fail(QSMI_ERR_ARG, "no map: call qsmi_candidate_for_pattern first");
  const long long S = dev->ti->S;
  HIPCHK(hipMemcpyAsync(out, dev->dMap + S, S * sizeof(float), hipMemcpyDeviceToHost, dev->stream));
  HIPCHK(hipStreamSynchronize(dev->stream));
  return QSMI_OK;
}

int qsmi_search_k_biggest(qsmi_device* dev, float* errors, unsigned* encodedPosition, unsigned extendK, float seed) {
  int rc = check_dev(dev);
  if (rc) return rc;
  if (!errors || !encodedPosition) return fail(QSMI_ERR_ARG, "null argument");
  if (extendK < 1 || extendK > (unsigned)kMaxK) return fail(QSMI_ERR_UNSUPPORTED, "extendK must be in [1, 16]");
  if (!dev->mapValid) return fail(QSMI_ERR_ARG, "no map: call qsmi_candidate_for_pattern first");
  qsmi_ti* t = dev->ti;
  const int K = (int)extendK;
  const int chunks = 512;
  if ((rc = ensure(&dev->dCand, &dev->capCand, (size_t)chunks * K))) return rc;
  if ((rc = ensure(&dev->dOutVal, &dev->capOut, (size_t)K))) return rc;
  if ((rc = ensure(&dev->dOutPos, &dev->capOutPos, (size_t)K))) return rc;
  hipStream_t st = dev->stream;
  // the last scanned node with this call's seed (its neighbours stay in
  // dOff / dVal for the exact selection)
  NodeInfo ni = dev->lastNode;
  ni.seed = seed;
  ni.seedbits = seed_from_unit_float(seed);
  ni.quantum = dev->quantum;
  if ((rc = ensure(&dev->dInfo, &dev->capInfo, 1))) return rc;
  HIPCHK(hipMemcpyAsync(dev->dInfo, &ni, sizeof(ni), hipMemcpyHostToDevice, st));
  launch_select_map(dev->dMap, t->S, K, dev->quantum, ni.seedbits, dev->dCand, chunks, st);
  ScanGeom g = geom_of(dev);
  launch_merge_draw(g, dev->dInfo, dev->dCand, 1, chunks, K, 0, K, (float)K, dev->dOutVal, dev->dOutPos, nullptr,
                    nullptr, nullptr, 0, nullptr, st);
  // exact selection of the FP32 map's band unless a value was set into the
  // map from outside (exclusions by -inf keep it)
  if (ni.eps > 0.f && !dev->mapEdited && exact_selection_on(dev))
    if ((rc = band_select_map(dev, dev->dInfo, ni.nb, K, ni.eps, dev->dOutVal, dev->dOutPos, &dev->band))) return rc;
  HIPCHK(hipGetLastError());
  std::vector<uint32_t> pos(K);
  HIPCHK(hipMemcpyAsync(errors, dev->dOutVal, K * sizeof(float), hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(pos.data(), dev->dOutPos, K * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  for (int i = 0; i < K; ++i) {
    if (pos[i] == 0xFFFFFFFFu) {
      encodedPosition[i] = UINT_MAX;
      continue;
    }
    unsigned idx;
    qsmi_cvt_position_to_index(dev, pos[i], &idx);
    encodedPosition[i] = idx;
  }
  return QSMI_OK;
}

int qsmi_search_k_biggest_noalloc(qsmi_device* dev, float* errors, unsigned* encodedPosition, unsigned extendK,
                                  float seed) {
  int rc = check_dev(dev);
  if (rc) return rc;
  if (!errors || !encodedPosition) return fail(QSMI_ERR_ARG, "null argument");
  if (extendK < 1 || extendK > (unsigned)kMaxK) return fail(QSMI_ERR_UNSUPPORTED, "extendK must be in [1, 16]");
  if (!dev->mapValid) return fail(QSMI_ERR_ARG, "no map: call qsmi_candidate_for_pattern first");
  qsmi_ti* t = dev->ti;
  const int K = (int)extendK;
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t candBytes = al((size_t)kScratchChunks * K * sizeof(Cand));
  const size_t need = candBytes + 2 * al((size_t)K * 4) + al(sizeof(NodeInfo)) + band_scratch_bytes(K);
  if (need > kScratchBytes)
    return fail(QSMI_ERR_UNSUPPORTED, "extendK " + std::to_string(K) + " needs " + std::to_string(need) +
                                          " bytes of scratch, above the " + std::to_string(kScratchBytes) + " it holds");
  void* base = nullptr;
  if ((rc = scratch_address(t->gpu, &base))) return rc;
  char* p = static_cast<char*>(base);
  Cand* cand = reinterpret_cast<Cand*>(p);
  p += candBytes;
  float* outVal = reinterpret_cast<float*>(p);
  p += al((size_t)K * 4);
  uint32_t* outPos = reinterpret_cast<uint32_t*>(p);
  p += al((size_t)K * 4);
  NodeInfo* info = reinterpret_cast<NodeInfo*>(p);
  p += al(sizeof(NodeInfo));
  BandBufs band;
  band_view(&band, p, K);
  hipStream_t st = dev->stream;
  // the last scanned node with this call's seed (its neighbours stay in
  // dOff / dVal for the exact selection); the copy completes before return
  NodeInfo ni = dev->lastNode;
  ni.seed = seed;
  ni.seedbits = seed_from_unit_float(seed);
  ni.quantum = dev->quantum;
  HIPCHK(hipMemcpyAsync(info, &ni, sizeof(ni), hipMemcpyHostToDevice, st));
  launch_select_map(dev->dMap, t->S, K, dev->quantum, ni.seedbits, cand, kScratchChunks, st);
  ScanGeom g = geom_of(dev);
  launch_merge_draw(g, info, cand, 1, kScratchChunks, K, 0, K, (float)K, outVal, outPos, nullptr, nullptr, nullptr,
                    0, nullptr, st);
  if (ni.eps > 0.f && !dev->mapEdited && exact_selection_on(dev))
    if ((rc = band_select_map(dev, info, ni.nb, K, ni.eps, outVal, outPos, &band))) return rc;
  HIPCHK(hipGetLastError());
  uint32_t pos[kMaxK];
  HIPCHK(hipMemcpyAsync(errors, outVal, K * sizeof(float), hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(pos, outPos, K * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  for (int i = 0; i < K; ++i) {
    if (pos[i] == 0xFFFFFFFFu) {
      encodedPosition[i] = UINT_MAX;
      continue;
    }
    unsigned idx;
    qsmi_cvt_position_to_index(dev, pos[i], &idx);
    encodedPosition[i] = idx;
  }
  return QSMI_OK;
}

// src/CPUThreadDevice.cpp:257-268
int qsmi_cvt_index_to_position(qsmi_device* dev, unsigned index, unsigned* position) {
  if (!dev || !position) return fail(QSMI_ERR_ARG, "null argument");
  const qsmi_ti* t = dev->ti;
  unsigned pos = 0;
  unsigned div = (unsigned)t->N;
  for (int i = t->D - 1; i >= 0; --i) {
    const unsigned f = t->f[i], s = t->s[i];
    div /= f;
    pos = pos * s + ((f - (index / div) % f - dev->mn[i] - 1) + s) % s;
  }
  *position = pos;
  return QSMI_OK;
}

// src/CPUThreadDevice.cpp:270-284
int qsmi_cvt_position_to_index(qsmi_device* dev, unsigned position, unsigned* index) {
  if (!dev || !index) return fail(QSMI_ERR_ARG, "null argument");
  const qsmi_ti* t = dev->ti;
  unsigned idx = 0;
  unsigned div = (unsigned)t->S;
  for (int i = t->D - 1; i >= 0; --i) {
    const unsigned f = t->f[i], s = t->s[i];
    div /= s;
    idx = idx * f + (f - (position / div + dev->mn[i]) % s - 1);
  }
  *index = idx;
  return QSMI_OK;
}

int qsmi_get_array_size(qsmi_device* dev, unsigned* size) {
  if (!dev || !size) return fail(QSMI_ERR_ARG, "null argument");
  *size = (unsigned)dev->ti->N;
  return QSMI_OK;
}

int qsmi_get_error_map(qsmi_device* dev, float* out) {
  int rc = check_dev(dev);
  if (rc) return rc;
  if (!out) return fail(QSMI_ERR_ARG, "null argument");
  if (!dev->mapValid) return fail(QSMI_ERR_ARG, "no map: call qsmi_candidate_for_pattern first");
  HIPCHK(hipMemcpyAsync(out, dev->dMap, dev->ti->S * sizeof(float), hipMemcpyDeviceToHost, dev->stream));
  HIPCHK(hipStreamSynchronize(dev->stream));
  return QSMI_OK;
}

int qsmi_set_value_in_error_array(qsmi_device* dev, unsigned index, float value) {
  int rc = check_dev(dev);
  if (rc) return rc;
  if (!dev->mapValid) return fail(QSMI_ERR_ARG, "no map: call qsmi_candidate_for_pattern first");
  if (index >= (unsigned)dev->ti->N) return fail(QSMI_ERR_ARG, "index out of range");
  unsigned pos;
  qsmi_cvt_index_to_position(dev, index, &pos);
  if (pos >= (unsigned)dev->ti->S) return fail(QSMI_ERR_ARG, "index maps outside the TI");
  HIPCHK(hipMemcpyAsync(dev->dMap + pos, &value, sizeof(float), hipMemcpyHostToDevice, dev->stream));
  HIPCHK(hipStreamSynchronize(dev->stream));
  if (!(value == -INFINITY)) dev->mapEdited = true;
  return QSMI_OK;
}

// ---------------------------------------------------------------------------
// waves
// ---------------------------------------------------------------------------
}  // extern "C"

namespace qsmi {

void free_wave_bufs(WaveBufs* b) {
  (void)hipFree(b->dInfo);
  (void)hipFree(b->dOff);
  (void)hipFree(b->dVal);
  (void)hipFree(b->dOrd);
  (void)hipFree(b->dC);
  (void)hipFree(b->dCand);
  (void)hipFree(b->dCandEx);
  (void)hipFree(b->dGcnt);
  (void)hipFree(b->dOutVal);
  (void)hipFree(b->dOutPos);
  (void)hipFree(b->dChosen);
  (void)hipHostFree(b->hInfo);
  (void)hipHostFree(b->hOff);
  (void)hipHostFree(b->hVal);
  band_free(&b->band);
  *b = WaveBufs();
}

int grow_wave_bufs(qsmi_device* dev, WaveBufs* b, int nNodes, int total, int K) {
  qsmi_ti* t = dev->ti;
  const int V = std::max((int)dev->active.size(), 1);
  const int rb = row_blocks(dev);
  if ((size_t)nNodes > b->capNodes || (size_t)nNodes * rb * K > b->capCand) {
    size_t n = std::max<size_t>(nNodes, b->capNodes);
    (void)hipFree(b->dInfo);
    (void)hipFree(b->dOutVal);
    (void)hipFree(b->dOutPos);
    (void)hipFree(b->dChosen);
    (void)hipFree(b->dCand);
    (void)hipFree(b->dCandEx);
    (void)hipFree(b->dGcnt);
    (void)hipHostFree(b->hInfo);
    b->dCandEx = nullptr;
    b->dGcnt = nullptr;
    b->dInfo = nullptr;
    b->dOutVal = nullptr;
    b->dOutPos = nullptr;
    b->dChosen = nullptr;
    b->dCand = nullptr;
    b->hInfo = nullptr;
    b->capNodes = 0;
    b->capCand = 0;
    HIPCHK(dalloc(&b->dInfo, n));
    HIPCHK(dalloc(&b->dOutVal, n * kMaxK));
    HIPCHK(dalloc(&b->dOutPos, n * kMaxK));
    HIPCHK(dalloc(&b->dChosen, n));
    HIPCHK(dalloc(&b->dCand, n * rb * kMaxK));
    HIPCHK(dalloc(&b->dCandEx, n * kMaxK));
    HIPCHK(dalloc(&b->dGcnt, n * rb));
    HIPCHK(hipHostMalloc((void**)&b->hInfo, n * sizeof(NodeInfo), hipHostMallocDefault));
    b->capNodes = n;
    b->capCand = n * rb * kMaxK;
  }
  if ((size_t)nNodes * c_stride(dev) > b->capC) {
    (void)hipFree(b->dC);
    b->dC = nullptr;
    b->capC = 0;
    size_t n = (size_t)nNodes * c_stride(dev);
    hipError_t e = dalloc(&b->dC, n);
    if (e != hipSuccess) return fail(QSMI_ERR_NOMEM, std::string("wave spectrum buffer: ") + hipGetErrorString(e));
    b->capC = n;
  }
  if ((size_t)std::max(total, 1) > b->capNb) {
    size_t n = std::max<size_t>(std::max(total, 1), b->capNb * 3 / 2);
    (void)hipFree(b->dOff);
    (void)hipFree(b->dVal);
    (void)hipFree(b->dOrd);
    (void)hipHostFree(b->hOff);
    (void)hipHostFree(b->hVal);
    b->dOff = nullptr;
    b->dVal = nullptr;
    b->dOrd = nullptr;
    b->hOff = nullptr;
    b->hVal = nullptr;
    b->capNb = 0;
    HIPCHK(dalloc(&b->dOff, n * 3));
    HIPCHK(dalloc(&b->dVal, n * V));
    HIPCHK(dalloc(&b->dOrd, n));
    HIPCHK(hipHostMalloc((void**)&b->hOff, n * 3 * sizeof(int), hipHostMallocDefault));
    HIPCHK(hipHostMalloc((void**)&b->hVal, n * V * sizeof(float), hipHostMallocDefault));
    b->capNb = n;
  }
  return QSMI_OK;
}

// Host preparation into pinned staging + asynchronous upload.
int stage_wave(qsmi_device* dev, WaveBufs* b, int nNodes, const int* nbStart, const int* offsets,
               const float* values, const float* delta0, const float* seeds, float k, const unsigned* exPos,
               const float* exRadius) {
  if (nNodes < 1) return fail(QSMI_ERR_ARG, "empty wave");
  if (!nbStart || !delta0 || !seeds) return fail(QSMI_ERR_ARG, "null argument");
  const int K = (int)std::ceil(k);
  if (!(k > 0.f) || K > kMaxK) return fail(QSMI_ERR_UNSUPPORTED, "ceil(k) must be in [1, 16]");
  qsmi_ti* t = dev->ti;
  const int D = t->D, nSub = t->nSub, V = (int)dev->active.size();
  const int total = nbStart[nNodes];
  if (nbStart[0] != 0 || total < 0) return fail(QSMI_ERR_ARG, "nbStart must start at 0");
  int nbMax = 0;
  for (int i = 0; i < nNodes; ++i) {
    if (nbStart[i + 1] < nbStart[i]) return fail(QSMI_ERR_ARG, "nbStart must be non-decreasing");
    nbMax = std::max(nbMax, nbStart[i + 1] - nbStart[i]);
  }
  if (total > 0 && (!offsets || !values)) return fail(QSMI_ERR_ARG, "null neighbour arrays");
  int rc = grow_wave_bufs(dev, b, nNodes, total, K);
  if (rc) return rc;
  for (int i = 0; i < nNodes; ++i) {
    const int s0 = nbStart[i], nb = nbStart[i + 1] - s0;
    prep_node(dev, nb, offsets + (size_t)s0 * D, values + (size_t)s0 * nSub, delta0[i], seeds[i], s0, &b->hInfo[i],
              b->hOff + (size_t)s0 * D, b->hVal + (size_t)s0 * V, nullptr, nullptr);
    if (exPos && exPos[i] != 0xFFFFFFFFu) {
      if (exPos[i] >= (unsigned)t->S) return fail(QSMI_ERR_ARG, "verbatim position outside the TI");
      // QuantileSamplingModule::determineDistribution: a radius below 1
      // removes the position alone (include/quantileSamplingModule.hpp:566-571)
      const float r = exRadius ? exRadius[i] : 0.f;
      b->hInfo[i].exOn = 1;
      b->hInfo[i].exPos = exPos[i];
      b->hInfo[i].exR2 = (r >= 1.f) ? r * r : 0.f;
    }
  }
  hipStream_t st = dev->stream;
  HIPCHK(hipMemcpyAsync(b->dInfo, b->hInfo, nNodes * sizeof(NodeInfo), hipMemcpyHostToDevice, st));
  if (total > 0) {
    HIPCHK(hipMemcpyAsync(b->dOff, b->hOff, (size_t)total * D * sizeof(int), hipMemcpyHostToDevice, st));
    if (V > 0) HIPCHK(hipMemcpyAsync(b->dVal, b->hVal, (size_t)total * V * sizeof(float), hipMemcpyHostToDevice, st));
  }
  b->nNodes = nNodes;
  b->K = K;
  b->k = k;
  b->nbMax = nbMax;
  b->rowBlocks = row_blocks(dev);
  return QSMI_OK;
}

int run_wave(qsmi_device* dev, WaveBufs* b) {
  qsmi_ti* t = dev->ti;
  hipStream_t st = dev->stream;
  ScanGeom g = geom_of(dev);
  int rc = launch_spectrum(dev, g, b->nNodes, b->dInfo, b->dOff, b->dVal, b->dOrd, b->dC);
  if (rc) return rc;
  launch_rows(dev, g, b->dInfo, b->dC, b->nNodes, b->rowBlocks, b->K, b->dCand, b->dGcnt, nullptr);
  if (dev->timing) (void)hipEventRecord(dev->ev[3], st);
  rc = band_select(dev, g, b->dInfo, b->dCand, b->dGcnt, b->nNodes, b->K, b->k, b->dC, b->dOff, b->dVal, b->nbMax,
                   b->dOutVal, b->dOutPos, nullptr, b->dChosen, nullptr, nullptr, 0, nullptr, &b->band);
  if (rc) return rc;
  if (dev->timing) (void)hipEventRecord(dev->ev[4], st);
  HIPCHK(hipGetLastError());
  return QSMI_OK;
}

int read_wave(qsmi_device* dev, WaveBufs* b, float* outErrors, unsigned* outPositions, unsigned* outChosen) {
  hipStream_t st = dev->stream;
  const size_t nk = (size_t)b->nNodes * b->K;
  if (outErrors) HIPCHK(hipMemcpyAsync(outErrors, b->dOutVal, nk * sizeof(float), hipMemcpyDeviceToHost, st));
  if (outPositions) HIPCHK(hipMemcpyAsync(outPositions, b->dOutPos, nk * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  if (outChosen) HIPCHK(hipMemcpyAsync(outChosen, b->dChosen, b->nNodes * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  if (dev->timing) {
    for (int i = 0; i < 4; ++i) {
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, dev->ev[i], dev->ev[i + 1]);
      dev->stageMs[i] = ms;
    }
  }
  return QSMI_OK;
}

}  // namespace qsmi

extern "C" {

int qsmi_wave_create(qsmi_device* dev, int nNodes, const int* nbStart, const int* offsets, const float* values,
                     const float* delta0, const float* seeds, float k, qsmi_wave** out) {
  int rc = check_dev(dev);
  if (rc) return rc;
  if (dev->xLayer) return fail(QSMI_ERR_UNSUPPORTED, "wave calls: not with the per-node cross layer");
  if (!out) return fail(QSMI_ERR_ARG, "null argument");
  *out = nullptr;
  auto* w = new qsmi_wave();
  w->dev = dev;
  rc = stage_wave(dev, &w->b, nNodes, nbStart, offsets, values, delta0, seeds, k);
  if (!rc) {
    hipError_t e = hipStreamSynchronize(dev->stream);
    if (e != hipSuccess) rc = fail(QSMI_ERR_HIP, std::string("wave upload: ") + hipGetErrorString(e));
  }
  if (rc) {
    free_wave_bufs(&w->b);
    delete w;
    return rc;
  }
  *out = w;
  return QSMI_OK;
}

int qsmi_wave_run(qsmi_wave* w) {
  if (!w) return fail(QSMI_ERR_ARG, "null wave");
  int rc = check_dev(w->dev);
  if (rc) return rc;
  return run_wave(w->dev, &w->b);
}

int qsmi_wave_read(qsmi_wave* w, float* outErrors, unsigned* outPositions, unsigned* outChosen) {
  if (!w) return fail(QSMI_ERR_ARG, "null wave");
  int rc = check_dev(w->dev);
  if (rc) return rc;
  return read_wave(w->dev, &w->b, outErrors, outPositions, outChosen);
}

int qsmi_wave_destroy(qsmi_wave* w) {
  if (!w) return QSMI_OK;
  if (w->dev && w->dev->ti) {
    (void)hipSetDevice(w->dev->ti->gpu);
    (void)hipStreamSynchronize(w->dev->stream);
  }
  free_wave_bufs(&w->b);
  delete w;
  return QSMI_OK;
}

int qsmi_distribution_wave(qsmi_device* dev, int nNodes, const int* nbStart, const int* offsets, const float* values,
                           const float* delta0, const float* seeds, float k, const unsigned* verbatimPos,
                           const float* verbatimRadius, float* outErrors, unsigned* outPositions) {
  int rc = check_dev(dev);
  if (rc) return rc;
  if (dev->xLayer) return fail(QSMI_ERR_UNSUPPORTED, "wave calls: not with the per-node cross layer");
  if (!verbatimPos || !verbatimRadius) return fail(QSMI_ERR_ARG, "null verbatim arrays");
  if (!dev->wb) dev->wb = new WaveBufs();
  rc = stage_wave(dev, dev->wb, nNodes, nbStart, offsets, values, delta0, seeds, k, verbatimPos, verbatimRadius);
  if (!rc) rc = run_wave(dev, dev->wb);
  if (!rc) rc = read_wave(dev, dev->wb, outErrors, outPositions, nullptr);
  return rc;
}

int qsmi_set_value_in_error_array_with_radius(qsmi_device* dev, unsigned index, float value, float radius) {
  int rc = check_dev(dev);
  if (rc) return rc;
  if (!dev->mapValid) return fail(QSMI_ERR_ARG, "no map: call qsmi_candidate_for_pattern first");
  if (index >= (unsigned)dev->ti->N) return fail(QSMI_ERR_ARG, "index out of range");
  if (!(radius >= 0.f) || !std::isfinite(radius)) return fail(QSMI_ERR_ARG, "radius must be finite and >= 0");
  unsigned pos;
  qsmi_cvt_index_to_position(dev, index, &pos);
  if (pos >= (unsigned)dev->ti->S) return fail(QSMI_ERR_ARG, "index maps outside the TI");
  launch_fill_disk(dev->dMap, dev->ti->s, pos, radius * radius, value, dev->stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(dev->stream));
  if (!(value == -INFINITY)) dev->mapEdited = true;
  return QSMI_OK;
}

int qsmi_scan_wave(qsmi_device* dev, int nNodes, const int* nbStart, const int* offsets, const float* values,
                   const float* delta0, const float* seeds, float k, float* outErrors, unsigned* outPositions,
                   unsigned* outChosen) {
  int rc = check_dev(dev);
  if (rc) return rc;
  if (dev->xLayer) return fail(QSMI_ERR_UNSUPPORTED, "wave calls: not with the per-node cross layer");
  if (!dev->wb) dev->wb = new WaveBufs();
  rc = stage_wave(dev, dev->wb, nNodes, nbStart, offsets, values, delta0, seeds, k);
  if (!rc) rc = run_wave(dev, dev->wb);
  if (!rc) rc = read_wave(dev, dev->wb, outErrors, outPositions, outChosen);
  return rc;
}

}  // extern "C"
