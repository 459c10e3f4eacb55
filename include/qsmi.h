/*
 * qsmi — C-ABI of the MI355X-native QuickSampling exact scan.
 *
 * Plain pointers, sizes and integer status codes; no C++ or torch types.
 * Every entry point names the reference interface it replaces
 * (paths relative to the G2S source tree).
 *
 * Conventions
 *  - Grids are flattened with dimension 0 fastest (the reference's
 *    DataImage order). ndim is 1, 2 or 3.
 *  - "position" = flattened TI cell index on the TI grid (srcSize);
 *    "index"    = the reference's encoded map index on the FFT grid
 *                 (src/CPUThreadDevice.cpp:257-284).
 *  - Neighbour offsets are relative to the simulated cell, NOT yet shifted
 *    by the bounding-box minimum (as candidateForPattern receives them).
 *  - Neighbour values are the converted sub-variable values the reference's
 *    QuantileSamplingModule passes on (P0 = w, P1 = w*x, ...,
 *    include/quantileSamplingModule.hpp:368-410), nNeighbors x nSub,
 *    sub-variable fastest.
 *  - Calls on one device are stream-ordered on that device's own HIP stream;
 *    a device must not be used from two host threads at once.
 *  - Host buffers may be pageable; the library stages them.
 */
#ifndef QSMI_H
#define QSMI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct qsmi_ti qsmi_ti;          /* TI spectra resident on one GPU */
typedef struct qsmi_device qsmi_device;  /* one scan engine over a qsmi_ti */
typedef struct qsmi_wave qsmi_wave;      /* a device-resident path wave */

enum {
  QSMI_OK = 0,
  QSMI_ERR_ARG = 1,          /* invalid argument (message in qsmi_last_error) */
  QSMI_ERR_HIP = 2,          /* HIP runtime error */
  QSMI_ERR_UNSUPPORTED = 3,  /* valid request outside what this build does */
  QSMI_ERR_NOMEM = 4
};

/* Last error message of the calling thread ("" when none). */
const char* qsmi_last_error(void);

/* Replaces g2s_cudaGetDeviceCount (src/NvidiaGPUAcceleratorDevice.cu:26-29). */
int qsmi_get_device_count(int* count);

/* FFT size rule of SharedMemoryManager (src/sharedMemoryManager.cpp:26-58):
 * smallest size >= src of the form 2^a 3^b 5^c 7^d {1, 11, 13}. */
unsigned qsmi_fft_size(unsigned src);

/* Replaces AcceleratorDevice::allocAndInitSharedMemory
 * (include/AcceleratorDevice.hpp:46; src/NvidiaGPUAcceleratorDevice.cu:350-389;
 * src/CPUThreadDevice.cpp:171-211) together with the SharedMemoryManager that
 * calls it (src/sharedMemoryManager.cpp:68-129).
 * subImages: nSub images of prod(srcSize) floats on the TI grid, NOT
 * reflected (y^2, y for a continuous variable, 1(y==c) per class; NaN -> 0,
 * include/DataImage.hpp:538-599). fftSize may be NULL (size rule applied).
 * Every FFT extent must be <= 4096. The spectra stay in HBM until destroy. */
int qsmi_ti_create(int gpu, int ndim, const unsigned* srcSize, const unsigned* fftSize, int nSub,
                   const float* subImages, qsmi_ti** out);
/* The sub-images in the layout G2S's SharedMemoryManager hands to
 * AcceleratorDevice::allocAndInitSharedMemory (src/sharedMemoryManager.cpp:
 * 103-108, fed by src/qs.cpp:1544-1550 from DataImage::convertInput4Xcorr,
 * include/DataImage.hpp:538-599): one host pointer per sub-image, each
 * prod(fftSize) floats, point-reflected on the FFT grid (TI cell c at
 * prod(fftSize) - 1 - flat_fft(c), zero elsewhere). Writes the unreflected
 * TI-grid images qsmi_ti_create takes (nSub x prod(srcSize)). Returns
 * QSMI_ERR_ARG when an image holds a non-zero value outside the reflected
 * TI box (not this layout). Host only. */
int qsmi_subimages_from_reflected(int ndim, const unsigned* srcSize, const unsigned* fftSize, int nSub,
                                  const float* const* reflected, float* out);
/* qsmi_ti_create from that layout: what an AcceleratorDevice subclass's
 * allocAndInitSharedMemory calls with the arguments it receives
 * (include/AcceleratorDevice.hpp:46). fftSize is required. */
int qsmi_ti_create_from_reflected(int gpu, int ndim, const unsigned* srcSize, const unsigned* fftSize, int nSub,
                                  const float* const* reflected, qsmi_ti** out);
/* Replaces AcceleratorDevice::freeSharedMemory (include/AcceleratorDevice.hpp:47). */
int qsmi_ti_destroy(qsmi_ti* ti);

/* Launch parameters of the scan kernels over one TI (no counterpart in the
 * reference, whose FFTW/cuFFT plans choose their own; its planner call is
 * src/CPUThreadDevice.cpp:99-141). Keys:
 *   "pencil.npb"  path nodes per block of the pencil kernel of TIs scanned
 *                 on their own FFT grid (3-D, smaller than 513 or circular): 1..16
 *   "scan.grid1024_min"  2-D TIs (not circular, no NaN) whose extents are
 *                 all in [value, 1024] scan on a 1024 x 1024 grid instead of
 *                 their own FFT grid: 2..1025 (1025: never), default 320
 *                 (a 500 x 500 TI with six spectra: 2.5x the own-grid rate);
 *                 read when the TI is created (set it with
 *                 qsmi_set_default_tuning before qsmi_ti_create)
 *   "fft.radix_max"  largest radix of the own-grid FFT plans: 2..13, default
 *                 13 (a length with no other factor keeps its larger radix);
 *                 read when the TI is created
 *   "fft.lognv_max"  at most 2^value transforms per own-grid FFT block:
 *                 0..6, default 6; read when the TI is created
 *   "loop.graph"  1: the path loop of qsmi_qs_run_in_memory replays each
 *                 wave's kernels as one captured HIP graph per stream (plain
 *                 1024-grid scans with the two-stream loop; the sampled
 *                 stage timing is not recorded then): 0..1, default 0
 * Results do not depend on them up to the FP32 map error the exact selection
 * absorbs (pencil.npb, loop.graph: every choice gives the same bytes).
 * QSMI_ERR_ARG for an unknown key or a value out of range. */
int qsmi_ti_set_tuning(qsmi_ti* ti, const char* key, int value);
int qsmi_ti_get_tuning(const qsmi_ti* ti, const char* key, int* value);
/* The value every TI created afterwards starts with (also the TIs the
 * in-memory simulation entry points build). Thread-safe. */
int qsmi_set_default_tuning(const char* key, int value);

/* Replaces the c_NvidiaGPUAcceleratorDevice factory
 * (src/NvidiaGPUAcceleratorDevice.cu:229-232) for one coefficient layer
 * (the cross-measurement layer is added by qsmi_device_set_cross_layer):
 * alpha is the OperationMatrix of
 * include/OperationMatrix.hpp:9-73, nSub x nSub, element (A,B) at
 * alpha[A + B*nSub]. circularTI as in the reference (requires
 * fftSize == srcSize). The device starts with trueMismatch = 1 (what qs sets
 * for accelerator devices, src/qs.cpp:1583). */
int qsmi_device_create(qsmi_ti* ti, const float* alpha, int circularTI, qsmi_device** out);
int qsmi_device_destroy(qsmi_device* dev);

/* AcceleratorDevice::setTrueMismatch (include/AcceleratorDevice.hpp:59):
 * 1 -> map = E (true mismatch incl. delta0); 0 -> map = N * (E - delta0),
 * N = prod(fftSize), the unnormalised CPU-device map. */
int qsmi_set_true_mismatch(qsmi_device* dev, int value);

/* Tie quantum of the device selection rule (no reference counterpart).
 * 0 (default): candidates ordered by value; > 0: by rint(value/quantum),
 * which makes the exact ties of integer-valued categorical maps exact.
 * Ties are broken by a hash of (seed, position): uniform among ties and
 * deterministic given the node seed. */
int qsmi_set_tie_quantum(qsmi_device* dev, float quantum);

/* Replaces AcceleratorDevice::candidateForPattern
 * (src/AcceleratorDevice.cpp:92-136 -> computeFreqMismatchMap +
 * computeRealMissmatchAndRemoveWrongPattern, include/AcceleratorDevice.hpp:71-72):
 * computes the node's mismatch map on the device (position space, -inf at
 * positions whose neighbourhood leaves the TI). delta0 has one entry.
 * Returns QSMI_OK; nNeighbors == 0 leaves the map untouched and returns
 * QSMI_OK with *computed = 0 (the reference returns false). */
int qsmi_candidate_for_pattern(qsmi_device* dev, int nNeighbors, const int* offsets, const float* values,
                               const float* delta0, int* computed);

/* Replaces AcceleratorDevice::searchKBiggest (include/AcceleratorDevice.hpp:68):
 * the extendK best entries of the current map, best first, as values and
 * encoded indices (reference index space). extendK <= 16.
 * Selection: the device rule (value desc, then a seeded hash of the position,
 * see qsmi_set_tie_quantum). Without a tie quantum it is exact: every map
 * position whose FP32 value lies within 2 eps of the extendK-th best
 * (eps = 4e-5 * sum_a rms(Y_a) * sum_j |v_ja|, the bound of the FP32 map
 * error: rms over the FFT grid of TI sub-image a, v_ja the neighbours'
 * alpha-combined values) is rescored in FP64 and the extendK best of those
 * values are returned (as floats). After a finite value was set with
 * qsmi_set_value_in_error_array(_with_radius) the FP32 map is used as is. */
int qsmi_search_k_biggest(qsmi_device* dev, float* errors, unsigned* encodedPosition, unsigned extendK, float seed);

/* No counterpart in the reference: qsmi_search_k_biggest for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid it). Same arguments and results; it calls no
 * allocation routine: the candidate lists, the K best and the exact
 * selection's buffers live in a 256 KiB scratch array the library holds on
 * every GPU it runs on (enough for every extendK <= 16; larger requests fail
 * with QSMI_ERR_UNSUPPORTED). Every call on a GPU shares that array, so calls
 * that overlap on one GPU (several devices' streams, several host threads)
 * race: serialise them. */
int qsmi_search_k_biggest_noalloc(qsmi_device* dev, float* errors, unsigned* encodedPosition, unsigned extendK,
                                  float seed);

/* cvtIndexToPosition / cvtPositionToIndex (include/AcceleratorDevice.hpp:57-58),
 * using the bounding box of the last candidate_for_pattern call. */
int qsmi_cvt_index_to_position(qsmi_device* dev, unsigned index, unsigned* position);
int qsmi_cvt_position_to_index(qsmi_device* dev, unsigned position, unsigned* index);

/* getArraySize / getErrorsArraySize (include/AcceleratorDevice.hpp:51-52): prod(fftSize). */
int qsmi_get_array_size(qsmi_device* dev, unsigned* size);

/* getErrorAtPosition for every position (include/AcceleratorDevice.hpp:54):
 * copies the current map, prod(srcSize) floats in position space. */
int qsmi_get_error_map(qsmi_device* dev, float* out);

/* setValueInErrorArray (include/AcceleratorDevice.hpp:65), index space. */
int qsmi_set_value_in_error_array(qsmi_device* dev, unsigned index, float value);

/* setValueInErrorArrayWithRadius (include/AcceleratorDevice.hpp:66), the
 * verbatim exclusion of QuantileSamplingModule::distribution
 * (include/quantileSamplingModule.hpp:566-571): every map position p with
 * sum_d (p_d - o_d)^2 <= radius^2, o = cvtIndexToPosition(index), is set to
 * value. The reference's devices disagree here (src/CPUThreadDevice.cpp:
 * 310-341 steps each level by an FFT extent, not by the dimension's stride,
 * src/CPUThreadAcceleratorDevice.cpp:385-407 compares an L1 distance with
 * radius^2, src/NvidiaGPUAcceleratorDevice.cu:646-669 does nothing); this is
 * the Euclidean disk their commented form states (src/CPUThreadDevice.cpp:
 * 286-308), without wrap-around. */
int qsmi_set_value_in_error_array_with_radius(qsmi_device* dev, unsigned index, float value, float radius);

/* The per-node cross-measurement layer of TIs with NaN: an AcceleratorDevice
 * constructed withCrossMeasurement (include/AcceleratorDevice.hpp:40;
 * src/AcceleratorDevice.cpp:33-37), whose second coefficient layer counts
 * the TI values under the neighbours (DataImage::generateCoefMatrix4Xcorr
 * with forXMeasurement, include/DataImage.hpp:626-710). alphaX: that
 * OperationMatrix, nSub x nSub, element (A,B) at alphaX[A + B*nSub]; the
 * TI's sub-images are convertInput4Xcorr's cross layout (qsmi_ti_subimages_x:
 * per variable y^2, y and the mask m). Afterwards qsmi_candidate_for_pattern
 * computes both layers: the map E (delta0 not added, as the reference's
 * devices skip it with the cross layer, src/CPUThreadAcceleratorDevice.cpp:348)
 * and the count layer C; qsmi_search_k_biggest selects on the FP32 map (no
 * exact selection); the wave calls are rejected. */
int qsmi_device_set_cross_layer(qsmi_device* dev, const float* alphaX);
/* maskLayerWithVariable (include/AcceleratorDevice.hpp:63,
 * src/CPUThreadAcceleratorDevice.cpp:360-383): layer 0 (the map) or 1 (the
 * count layer) times, at every valid position, the TI sub-image the count
 * layer reads for `variable` (the variable-th row of alphaX with a non-zero
 * coefficient: that variable's mask image). maskCroossError /
 * maskCroossErrorWithVariable (src/AcceleratorDevice.cpp:136-145) are
 * layer 1. */
int qsmi_mask_layer_with_variable(qsmi_device* dev, unsigned layer, unsigned variable);
/* compensateMissingData (src/CPUThreadAcceleratorDevice.cpp:414-425):
 * map = -|E / (C C C C)|, -inf where the count is zero. A zero count comes
 * out of the FP32 transforms as noise around 0, so "zero" is a count below
 * half the node's smallest positive count-layer value (the reference
 * compares with 0.0f). */
int qsmi_compensate_missing_data(qsmi_device* dev);
/* getCrossErrorAtPosition for every position (include/AcceleratorDevice.hpp:56):
 * the count layer, prod(srcSize) floats in position space. */
int qsmi_get_cross_map(qsmi_device* dev, float* out);

/* Path-wave batch: nNodes independent nodes in one launch sequence
 * (replaces nNodes x [candidateForPattern + searchKBiggest + the
 * QuantileSamplingModule::sample draw, include/quantileSamplingModule.hpp:90-126]).
 * nbStart: nNodes+1 prefix offsets into offsets (x ndim) / values (x nSub).
 * delta0, seeds: per node. k: the QS k (K = ceil(k) <= 16).
 * Outputs (host): outErrors/outPositions nNodes x K (TI positions, best
 * first; UINT32_MAX when fewer valid positions exist), outChosen (nullable)
 * the TI position drawn with rank floor(seed*k). Blocks until done. */
int qsmi_scan_wave(qsmi_device* dev, int nNodes, const int* nbStart, const int* offsets, const float* values,
                   const float* delta0, const float* seeds, float k, float* outErrors, unsigned* outPositions,
                   unsigned* outChosen);

/* nNodes x QuantileSamplingModule::distribution for one TI
 * (include/quantileSamplingModule.hpp:128-169 with determineDistribution
 * :342-577, the k best of the map without the draw): the scan_wave inputs,
 * plus per node a verbatim TI position (UINT32_MAX: none) whose disk of
 * radius verbatimRadius (below 1: the position alone) holds no candidate,
 * as qsmi_set_value_in_error_array_with_radius states it. Outputs
 * outErrors/outPositions nNodes x ceil(k), best first (UINT32_MAX past the
 * valid positions). Blocks until done. */
int qsmi_distribution_wave(qsmi_device* dev, int nNodes, const int* nbStart, const int* offsets, const float* values,
                           const float* delta0, const float* seeds, float k, const unsigned* verbatimPos,
                           const float* verbatimRadius, float* outErrors, unsigned* outPositions);

/* The same wave kept resident in HBM, for repeated timed launches:
 * upload once, run (asynchronous, stream-ordered), read results. */
int qsmi_wave_create(qsmi_device* dev, int nNodes, const int* nbStart, const int* offsets, const float* values,
                     const float* delta0, const float* seeds, float k, qsmi_wave** out);
int qsmi_wave_run(qsmi_wave* wave);
int qsmi_wave_read(qsmi_wave* wave, float* outErrors, unsigned* outPositions, unsigned* outChosen);
int qsmi_wave_destroy(qsmi_wave* wave);

/* Blocks until all work queued on the device's stream has finished. */
int qsmi_synchronize(qsmi_device* dev);

/* Device time of the last qsmi_wave_run / qsmi_scan_wave per kernel stage,
 * measured with HIP events on the device stream (milliseconds):
 * out[0] spectrum stage (pencil / direct), out[1] dim1 inverse (3-D only),
 * out[2] rows + select, out[3] merge + draw. */
int qsmi_last_stage_ms(qsmi_device* dev, float* out4);
/* Enables (1) or disables (0) the per-stage events (off by default). */
int qsmi_set_stage_timing(qsmi_device* dev, int enable);

/* ======================================================================
 * QuickSampling host layer: the qs.cpp / QuantileSamplingModule /
 * simulation() steps around the scan, restated in C++ above this C-ABI.
 * ====================================================================== */

/* Category values of variable `var` in order of first appearance, NaN
 * skipped (src/qs.cpp:1500-1522). *count receives the number found. */
int qsmi_categories(size_t cells, int nVar, const float* data, int var, float* out, int maxOut, int* count);

/* Coefficient matrix and conversions of DataImage::generateCoefMatrix4Xcorr
 * (include/DataImage.hpp:626-710), no cross-measurement layer (see
 * qsmi_coef_matrix_x).
 * types: nVar (0 continuous, 1 categorical); ncat: classes per categorical
 * variable. Outputs (capacity 16 sub-variables): alpha [A + B*nSub],
 * convTypes[B] (0 = P0 = w, 1 = P1 = w*x), subVarOf[B] (one-hot expanded
 * variable of B), *nSub. */
int qsmi_coef_matrix(int nVar, const int* types, const int* ncat, float* alpha, int* convTypes, int* subVarOf,
                     int* nSub);

/* TI sub-variable images of DataImage::convertInput4Xcorr
 * (include/DataImage.hpp:538-599), unreflected, on the TI grid:
 * out = nSub x prod(srcSize). catvals: classes of each categorical variable
 * back to back. */
int qsmi_ti_subimages(int ndim, const unsigned* srcSize, int nVar, const int* types, const float* ti,
                      const int* ncat, const float* catvals, float* out);

/* The cross-measurement form of the two set-up steps above, used when the
 * TI holds NaN (DataImage::generateCoefMatrix4Xcorr with forXMeasurement and
 * convertInput4Xcorr with needCrossMeasurement, include/DataImage.hpp:
 * 538-599, 626-710; continuous variables only): per variable the images
 * [y^2, y, m] (m = 1 where y is a number, NaN -> 0) and the conversions
 * [P0, P1, P2] (convTypes 0, 1, 2 = w, w x, w x^2); alpha is the mismatch
 * layer diag(-1, +2, -1), alphaX the count layer (m against P0). */
int qsmi_coef_matrix_x(int nVar, const int* types, float* alpha, float* alphaX, int* convTypes, int* subVarOf,
                       int* nSub);
int qsmi_ti_subimages_x(int ndim, const unsigned* srcSize, int nVar, const int* types, const float* ti, float* out);

/* Replaces DataImage::genearteKernel (include/DataImage.hpp:433-480), the
 * kernel qs builds from -kernel / -ks / -alpha (src/qs.cpp:1010-1042):
 * out = prod(dims) x nVar; types[j]: 0 uniform (weight[j]), 2 exponential
 * exp(-alphas[j] d), 7 gaussian (unit normal density at alphas[j] d), d the
 * distance to the centre (dims-1)/2; other types give zeros. */
int qsmi_kernel_generate(int ndim, const unsigned* dims, int nVar, const int* types, const float* weight,
                         const float* alphas, float* out);

/* Sorted neighbour offset list of src/qs.cpp:1055-1137: every offset of the
 * kernel box (half-width (size+1)/2-1), NaN-kernel offsets removed
 * (kernel may be NULL), std::sort by increasing |d|^2 (libstdc++, so ties
 * keep the reference's order). out: capacity x ndim ints. */
int qsmi_offsets(int ndim, const unsigned* kernelSize, int kernelNVar, const float* kernel, int* out,
                 int capacity, int* count);

/* Random path and per-node seeds of src/qs.cpp:975, 1159-1201, 1250-1258:
 * fully informed cells first, the rest std::shuffle'd with mt19937(seed),
 * then one U[0,1) float per path entry from the same generator. */
int qsmi_path(size_t cells, const unsigned char* informed, unsigned seed, unsigned* path, float* seeds,
              unsigned* beginPath);

/* In-memory QuickSampling run, the counterpart of g2s::qs::runInMemory
 * (include/qsCore.hpp:17-58, src/qsCore.cpp:86-170) for one TI:
 * qs.cpp's set-up, the simulation() path loop (include/simulation.hpp:43-710)
 * with the scan on the GPU, path nodes packed into waves of independent
 * nodes (serial-equivalent result). A TI with NaN (continuous variables)
 * runs in the cross-measurement form: the map is -|E / C^4| with E the
 * mismatch over the TI cells that hold values and C their weighted count
 * times the centre cell's completeness (CPUThreadDevice::maskLayerWithVariable
 * and compensateMissingData, src/CPUThreadDevice.cpp:500-524,
 * src/CPUThreadAcceleratorDevice.cpp:414-425). */
typedef struct {
  int gpu;
  int ndim;
  unsigned tiSize[3];
  int nVar;
  const int* varTypes;     /* nVar: 0 continuous, 1 categorical */
  const float* ti;         /* prod(tiSize) x nVar (NaN: no value, continuous variables only) */
  unsigned simSize[3];
  const float* di;         /* prod(simSize) x nVar; fully informed cells are hard
                              data, the others are simulated (a partially
                              informed cell keeps the values it holds) */
  unsigned kernelSize[3];  /* all 0: default uniform kernel (src/qs.cpp:1010-1042) */
  const float* kernel;     /* prod(kernelSize) x nVar, or NULL */
  int n;                   /* -n, centre included */
  float k;                 /* -k */
  unsigned seed;           /* -s */
  int waveSize;            /* max nodes per wave, 0 -> 512 (256 with the learned proposal) */
  float tieQuantum;        /* < 0 automatic: 0 with a continuous variable,
                              2^(ceil(log2(n*max|w|))-16) otherwise */
  int stageTiming;         /* k >= 1: time the kernels of every k-th wave with HIP
                              events (stageMs, scaled to all waves); 0: off */
  /* optional training-pair record (2-D, one continuous variable, exact scan
     only): for path node i (path order, hard data excluded) with i <
     recordCap, its neighbours 1..n-1 (the cell itself excluded, nearest
     first) and the exact best TI position */
  int recordCap;
  int* recOffsets;         /* recordCap x 80 x 2 (dx, dy) */
  float* recValues;        /* recordCap x 80, NaN in empty slots */
  unsigned* recBest;       /* recordCap; UINT32_MAX when the node drew from the marginal */
  /* optional learned proposal (qsmi_lp with keys built for this TI): */
  struct qsmi_lp* lp;      /* NULL: exact scan for every node */
  int lpShortlist;         /* target positions per node (0 -> 4096) */
  int lpFallbackEvery;     /* exact scan for every m-th path node (0: never) */
  float lpFallbackMse;     /* exact scan when the best proposed mismatch per
                              neighbour exceeds this x sigma^2 (<= 0: off) */
  int lpAudit;             /* 1: also run the exact scan for every node and
                              count how many exact k-best positions the
                              proposals held (slow; for measurements) */
  float lpMinCover;        /* exact scan when fewer than this fraction of a
                              node's neighbours lie on the key stencil (the
                              encoders cannot see the others; <= 0 -> 0.5) */
  int noVerbatim;          /* -nV (src/qs.cpp:711-715): per node, no candidate at
                              the TI position that continues the nearest
                              neighbour's copy (include/simulation.hpp:567-580);
                              not with the learned proposal */
  /* several training images (qs -ti a b ...): nTI >= 2 takes tis[0..nTI-1],
     each prod(its extent) x nVar (tiSizes, or tiSize for all), and ignores ti. Per node
     every TI's k best (QuantileSamplingModule::determineDistribution), merged
     by value and drawn with rank floor(seed k ceil(nTI/k) / nTI)
     (include/quantileSamplingModule.hpp:90-126); index = position * nTI + TI;
     the marginal draw spans the TIs in order (include/simulation.hpp:586-606).
     Not with NaN TIs, the learned proposal, the pair record or -nV. */
  int nTI;
  const float* const* tis;
  /* per simulation cell (prod(simSize) floats, NULL: none), as qs -ni / -kvi
     (include/simulation.hpp:337-357): the cell's n (centre included; n
     itself is then ignored) and its k (values <= 0: the run's k; ceil <= 16).
     Not with the learned proposal or the pair record. */
  const float* nImage;
  const float* kImage;
  /* several kernels chosen per cell (qs -ki a b ... -kii image,
     include/simulation.hpp:348-351, src/qs.cpp:1055-1137): nKernels >= 2
     takes kernels[q] (prod(kernelSizes[3q..3q+2]) x nVar) in place of
     kernel/kernelSize, each with its own offset order, and
     kernelIndexImage (prod(simSize) floats) the kernel of every simulated
     cell. Not with the learned proposal or the pair record. */
  int nKernels;
  const unsigned* kernelSizes;
  const float* const* kernels;
  const float* kernelIndexImage;
  int circularSim;         /* -csim (src/qs.cpp:797-801): neighbours wrap around the
                              simulation grid (include/simulation.hpp:394) */
  const float* tiIndexImage; /* -ii with several TIs (include/simulation.hpp:560): per
                              simulation cell the only TI its scan samples (NULL: all) */
  int circularTI;          /* -cti (src/qs.cpp:791-795): the devices' circularTI, every TI
                              position valid, neighbourhoods wrapped over the FFT grid
                              (the TI grid where its extents are FFT sizes) */
  const unsigned* tiSizes; /* several TIs of their own extents: nTI x 3 (entry 0 = tiSize);
                              NULL: tiSize for all */
  int progressSeconds;     /* > 0: a progress line on stderr about every that many
                              seconds of the path loop (simulated nodes so far; the
                              reference reports its progress the same way,
                              include/simulation.hpp:685-692); 0: quiet */
} qsmi_qs_request;

typedef struct {
  float* simulation;       /* out: prod(simSize) x nVar */
  unsigned* index;         /* out: prod(simSize), importDataIndex (0 where not simulated) */
  double setupMs;          /* out: TI spectra, offsets, path */
  double pathLoopMs;       /* out: neighbour search + waves + writes (the reference's timed loop) */
  double scanMs;           /* out: device time of the scan stages inside the loop */
  double stageMs[4];       /* out: per stage: [0] spectrum, [1] gather, [2] rows +
                              select, [3] merge + draw + write (stageTiming; with
                              k > 1 the timed waves' sum x waves / timed waves) */
  long long nodes;         /* out: simulated path nodes */
  long long scannedNodes;  /* out: nodes that went through the scan (the rest
                              drew from the marginal) */
  int waves;               /* out: waves launched */
  double phaseMs[4];       /* out: path loop phases: neighbour search, dependency
                              levels and waves, uploads, device loop */
  long long lpProposed;    /* out: nodes served by the learned proposal */
  long long lpFallbacks;   /* out: nodes sent to the exact scan by the rules */
  long long lpOverflows;   /* out: of which shortlist overflows */
  long long lpShortlistSum;  /* out: proposed positions over all proposed nodes */
  long long lpAuditNodes;  /* out: audited proposal nodes */
  long long lpAuditHits;   /* out: exact k-best positions found in the proposals */
  long long lpAuditBestHits; /* out: audited nodes whose exact best position was proposed */
  long long exactNodes;      /* out: nodes selected by the exact selection (FP64 rescoring) */
  long long bandPositions;   /* out: positions rescored in FP64 over those nodes */
  long long secondPassBlocks; /* out: rows blocks transformed again to list their band */
  long long fullRescoreNodes; /* out: nodes rescored over their whole valid box */
} qsmi_qs_result;

int qsmi_qs_run_in_memory(const qsmi_qs_request* req, qsmi_qs_result* res);

/* ======================================================================
 * Learned proposal (no counterpart in the reference): pattern encoders
 * that propose a shortlist of TI positions per path node; the shortlist is
 * rescored exactly and the FFT scan stays the fallback. Two MLP towers
 * 240 -> 256 -> 256 -> 256 -> 128 (ReLU), products on bf16 MFMA with fp32
 * accumulation, fp32 weights trained with Adam on an InfoNCE loss.
 * Features: 80 slots x (dx/16, dy/16, (x - mu)/sigma); query = a node's
 * neighbours (nearest first), key = a TI position on the 80 canonical
 * offsets. Parameters: qsmi_lp_param_count floats, query tower then key
 * tower, each W1 [240][256], b1, W2 [256][256], b2, W3 [256][256], b3,
 * W4 [256][128], b4 (W[in][out], row-major).
 * ====================================================================== */
typedef struct qsmi_lp qsmi_lp;

int qsmi_lp_create(int gpu, unsigned seed, qsmi_lp** out);
int qsmi_lp_destroy(qsmi_lp* lp);
int qsmi_lp_param_count(long long* count);
int qsmi_lp_get_params(qsmi_lp* lp, float* out);
int qsmi_lp_set_params(qsmi_lp* lp, const float* in);

/* C = op(A) op(B) on the lp GEMM (host arrays, row-major; transA: A given
 * as [K][M]; transB: B given as [N][K]). For tests of the MFMA kernel. */
int qsmi_lp_gemm(qsmi_lp* lp, int transA, int transB, int M, int N, int K, const float* A, const float* B,
                 float* C);
/* embeddings (n x 128) of n feature rows (n x 240) by tower 0 (query) or 1 (key) */
int qsmi_lp_encode(qsmi_lp* lp, int tower, int n, const float* features, float* embeddings);
/* InfoNCE loss of B queries against C >= B key rows (row b's positive is key
 * b) with logits scaled by invTau; mean loss and the gradient of every
 * parameter (qsmi_lp_param_count floats). */
int qsmi_lp_loss_grad(qsmi_lp* lp, int B, int C, const float* qfeat, const float* kfeat, float invTau,
                      float* loss, float* grad);

/* One TI and the training pairs recorded on it (qsmi_qs_request record). */
typedef struct {
  unsigned tiSize[2];
  const float* ti;          /* prod(tiSize), the continuous variable */
  const int* canon;         /* 80 x 2 canonical offsets of the key features */
  int nPairs;
  const int* offsets;       /* nPairs x 80 x 2 */
  const float* values;      /* nPairs x 80 (NaN: empty) */
  const unsigned* best;     /* nPairs (UINT32_MAX pairs are skipped) */
} qsmi_lp_train_set;

typedef struct {
  int steps;
  int batch;                /* queries per step (0 -> 256) */
  int negatives;            /* random TI positions per step (0 -> 768) */
  float lr;                 /* Adam step (0 -> 1e-3) */
  float invTau;             /* logit scale (0 -> 10) */
  unsigned seed;
} qsmi_lp_train_opts;

/* steps of Adam on batches drawn from one set; lossOut (nullable): mean loss
 * of each block of 50 steps (ceil(steps / 50) floats). */
int qsmi_lp_train(qsmi_lp* lp, const qsmi_lp_train_set* set, const qsmi_lp_train_opts* opts, float* lossOut);

/* Keys of every position of a TI (kept in HBM for proposals) and a fixed
 * sample of 8192 of them for the per-node thresholds. */
int qsmi_lp_build_keys(qsmi_lp* lp, const unsigned* tiSize, const float* ti, const int* canon);

/* ======================================================================
 * auto_qs calibration: the Monte Carlo search over (number of neighbours,
 * kernel, density) of calibration() (include/calibration.hpp:30-253) with
 * auto_qs's set-up (src/auto_qs.cpp:449-492 offsets sorted by distance,
 * :561-582 categories, :640 the QuantileSamplingModule with k = maxK, the
 * verbatim exclusion on and no kernel of its own), its distribution() calls
 * batched into waves of the GPU scan (qsmi_distribution_wave).
 *
 * One iteration visits every setting s = config + nConfigs * (kernel +
 * nKernels * density); a setting is skipped once its sample count exceeds
 * minIter and its error minus 1/power deviations exceeds the best of its
 * density plus 1/power of that best's deviations. Draws per setting, from one
 * std::mt19937 seeded with `seed` and std::uniform_real_distribution<float>
 * over [0, 1), in the reference's order: TI, cell, then per offset tried a
 * density draw, the noise swaps, the distribution seed. Every setting's
 * sample is drawn before the iteration's scans run, so the skip test of an
 * iteration sees the state at its start; the reference, a parallel loop with
 * a dynamic schedule and a clock seed, leaves this order open.
 * ====================================================================== */
typedef struct {
  int gpu;
  int ndim;                   /* 1..3 */
  int nTI;                    /* >= 1, one device each */
  const unsigned* tiSize;     /* nTI x 3 (unused dimensions 1) */
  const float* const* ti;     /* nTI pointers: prod(tiSize) x nVar, no NaN */
  int nVar;
  const int* varTypes;        /* nVar: 0 continuous, 1 categorical */
  int nKernels;               /* >= 1 */
  const unsigned* kernelSize; /* nKernels x 3; kernel 0's extent sets the offsets */
  const float* const* kernels;/* nKernels pointers: prod(kernelSize) x nVar */
  int nConfigs;               /* numbers of neighbours to try */
  int nbWidth;                /* entries per configuration (per variable, cycled) */
  const unsigned* nbList;     /* nConfigs x nbWidth */
  int nDensity;
  const float* density;       /* nDensity probabilities of keeping an offset */
  int maxK;                   /* -k of auto_qs: candidates scored per sample, <= 16 */
  float power;                /* error metric |x - y|^power */
  unsigned maxIter, minIter;
  double maxSeconds;          /* <= 0: no time limit */
  float noise;                /* offset swap level (0: none) */
  unsigned seed;
  int waveSize;               /* max samples per GPU wave, 0 -> 256, at most 1024 */
  /* optional per-sample record (samples taken in order, up to logCap) */
  int logCap;
  int* logSetting;            /* logCap */
  unsigned* logCell;          /* logCap x 2: TI index, cell */
  unsigned* logPos;           /* logCap x maxK x 2: TI index, position (UINT32_MAX past the valid ones) */
  int stageTiming;            /* 1: time every wave's kernels with HIP events (stageMs) */
} qsmi_calib_request;

typedef struct {
  float* meanError;           /* out: settings x maxK, (mean |x - y|^power)^(1/power) */
  float* devError;            /* out: settings x maxK */
  unsigned* samples;          /* out: settings x maxK */
  long long distributions;    /* out: samples scored */
  int iterations;             /* out */
  int waves;                  /* out */
  double ms;                  /* out: wall time of the loop (set-up excluded) */
  double scanMs;              /* out: of which waiting for the GPU waves */
  double drawMs;              /* out: host draws (partly overlapped with the waves) */
  double setupMs;             /* out: TI spectra, devices, offsets */
  double stageMs[4];          /* out: with stageTiming, device time per stage summed
                                 over the waves: spectrum, dim1 inverse (3-D),
                                 rows + select, merge */
} qsmi_calib_result;

int qsmi_calibration(const qsmi_calib_request* req, qsmi_calib_result* res);

/* ======================================================================
 * .bgrid files: G2S's serialized DataImage, the format its qs reads
 * training images and destination grids from and writes results to.
 *   size_t fullSize | unsigned nbDim | unsigned dims[nbDim] | unsigned
 *   nbVariable | unsigned types[nbVariable] (0 continuous, 1 categorical) |
 *   unsigned encoding (0 float, 1 integer, 2 unsigned) | float data[...]
 * (data in DataImage order: variable fastest, then dimension 0).
 * ====================================================================== */

/* Serialized size of a grid. Replaces the size computed in
 * DataImage::serialize (include/DataImage.hpp:145). */
int qsmi_bgrid_bytes(int ndim, const unsigned* dims, int nVar, size_t* bytes);

/* DataImage::serialize (include/DataImage.hpp:143-170) into a caller buffer
 * of at least qsmi_bgrid_bytes bytes. */
int qsmi_bgrid_serialize(int ndim, const unsigned* dims, int nVar, const int* varTypes, int encoding,
                         const float* data, void* out, size_t outBytes);

/* 1 when raw[0, bytes) is a well-formed serialized grid, else 0:
 * DataImage::validateSerializedData (include/DataImage.hpp:172-224). */
int qsmi_bgrid_valid(const void* raw, size_t bytes);

/* Header fields of a valid buffer (dims: up to 16 entries; varTypes:
 * typesCap entries) and the byte offset of its float data; the
 * DataImage(char*, size_t) constructor (include/DataImage.hpp:226-254). */
int qsmi_bgrid_header(const void* raw, size_t bytes, int* ndim, unsigned* dims, int* nVar, int* varTypes,
                      int typesCap, int* encoding, size_t* dataOffset);

/* Write a valid buffer to path, gzip-compressed when gzip != 0 (the file
 * forms of writeRawData, src/DataImage.cpp:126-163; the content-addressed
 * naming under /tmp/G2S/data is the server's and not reproduced). */
int qsmi_bgrid_write_file(const char* path, const void* raw, size_t bytes, int gzip);

/* Read a plain or gzip .bgrid file (loadRawData's readers,
 * src/DataImage.cpp:35-124) and validate it. out == NULL: only *bytes is
 * set (the file is read twice by a size query followed by a read). */
int qsmi_bgrid_read_file(const char* path, void* out, size_t outBytes, size_t* bytes);

#ifdef __cplusplus
}
#endif

#endif /* QSMI_H */
