// Private header of the qsmi library: device-side state behind the C-ABI
// handles and the internal entry points shared by the library's translation
// units. Not installed; include/qsmi.h is the public interface.
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <string>
#include <vector>

#include "../../include/qsmi.h"
#include "qsmi_kernels.hpp"

using qsmi::Cand;
using qsmi::FftPlan;
using qsmi::NodeInfo;
using qsmi::ScanGeom;

// Launch parameters of the scan kernels for one TI geometry: defaults, set
// by qsmi_ti_set_tuning, or chosen by measurement (qsmi_ti_autotune).
struct ScanTune {
  int pencilNpb = 1;       // generic pencil: path nodes per block
  int grid1024Min = 320;   // 2-D TIs whose extents are all in [grid1024Min, 1024] scan on the 1024 x 1024
                           // grid (read when the TI is created)
  int radixMax = 13;       // largest radix of the own-grid FFT plans (read when the TI is created)
  int lognvMax = 6;        // at most 2^lognvMax transforms per own-grid FFT block (read at creation)
  int loopGraph = 0;       // 1: the path loop replays each wave's kernels as one captured HIP graph
                           // (1024-grid scans on two lanes)
};

struct qsmi_ti {
  int gpu = 0;
  int D = 0;
  int s[3] = {1, 1, 1};
  int f[3] = {1, 1, 1};
  int h0 = 1;
  long long S = 1, N = 1, H = 1;
  int nSub = 0;
  float2* dYhat = nullptr;
  float2* dYT = nullptr;  // spectra on the 1024 x 1024 scan grid, [A][k0 < 513][k1] (w1k)
  float* dSub = nullptr;  // the sub-images on the TI grid, nSub x S
  float rms[qsmi::kMaxSub] = {};  // root mean square of each sub-image over the FFT grid
  // 2-D TIs of extent 513..1024: non-circular scans run on a 1024 x 1024
  // grid, one wavefront per transform (qsmi_scan1024.hip)
  bool w1k = false;
  float2* dTw1024 = nullptr;  // exp(-2 pi i t / 1024)
  float2* dTw[3] = {nullptr, nullptr, nullptr};
  FftPlan plan[3];
  int lognv[3] = {0, 0, 0};
  hipStream_t stream = nullptr;
  std::atomic<int> refs{0};  // devices over this TI (created from several host threads at once)
  ScanTune tune;
};

struct WaveBufs;

// Buffers of the exact selection (qsmi_band.hip), grown on demand.
struct BandBufs {
  size_t capNodes = 0, capQueue = 0;
  int K = 0;
  uint32_t* pos = nullptr;       // [node][kBandCap] positions to rescore
  int* cnt = nullptr;            // [node]
  float* theta = nullptr;        // [node] lower end of the band
  int2* queue = nullptr;         // second-pass items (node, rows block)
  int* qcnt = nullptr;           // [2]: queue items (zero between calls), unused
  double* val = nullptr;         // [node][kBandCap] FP64 values of the listed positions
  qsmi::Cand64* part = nullptr;  // [node][kFullTiles][K] full-fallback tile results
  unsigned long long* stats = nullptr;  // exact nodes, listed positions, second-pass blocks, full-fallback nodes
};

struct qsmi_device {
  qsmi_ti* ti = nullptr;
  hipStream_t stream = nullptr;
  std::vector<float> alpha;  // nSub x nSub, [A + B*nSub]
  // cross-measurement form (TI with NaN): the count layer's coefficients,
  // per TI cell 1 when complete, and the complete cells in order
  bool cross = false;
  std::vector<float> alphaX;
  uint8_t* dCentreOk = nullptr;
  uint32_t* dComplete = nullptr;
  uint32_t nComplete = 0;
  std::vector<int> active;   // TI sub-variables with a non-zero alpha row
  int circular = 0;
  int trueMismatch = 1;
  float quantum = 0.f;
  int mn[3] = {0, 0, 0}, mx[3] = {0, 0, 0};
  // per-node map state
  float* dMap = nullptr;  // the map; with xLayer the count layer follows at dMap + S
  bool mapValid = false;
  // per-node cross layer (qsmi_device_set_cross_layer): candidate_for_pattern
  // computes the count layer with alphaX too; xCmin: the last node's zero-count
  // threshold in map units
  bool xLayer = false;
  float xCmin = 0.f;
  // workspace
  float2* dC = nullptr;
  size_t capC = 0;
  qsmi::TapInfo1024* dTap = nullptr;  // 1024-grid path: distinct neighbour rows per node
  uint16_t* dTapD0 = nullptr;  // [node][kTapNbMax] d0 mod 1024 of each slot of the tap order
  float* dTapVal = nullptr;    // [node][kTapNbMax][V] values of each slot
  size_t capTap = 0;
  int capTapV = 0;
  Cand* dCand = nullptr;
  size_t capCand = 0;
  NodeInfo* dInfo = nullptr;
  size_t capInfo = 0;
  int* dOff = nullptr;
  size_t capOff = 0;
  float* dVal = nullptr;
  size_t capVal = 0;
  uint32_t* dOrd = nullptr;  // neighbours in pencil-row order (generic spectrum stage)
  size_t capOrd = 0;
  float* dOutVal = nullptr;
  uint32_t* dOutPos = nullptr;
  uint32_t* dChosen = nullptr;
  size_t capOut = 0;
  size_t capOutPos = 0;
  bool registered = false;
  // stage timing
  bool timing = false;
  hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  float stageMs[4] = {0, 0, 0, 0};
  WaveBufs* wb = nullptr;  // workspace of qsmi_scan_wave
  std::vector<hipEvent_t> evPool;  // per-wave stage events of the resident loop
  // per-node API: the last node (exact selection of its map) and whether the
  // map holds a value set from outside (then the FP32 selection is kept)
  NodeInfo lastNode{};
  bool mapEdited = false;
  BandBufs band;
  // second stream of the resident loop (consecutive waves of one dependency
  // level alternate between stream and side.stream), with its own wave
  // buffers and tap tables; swapped into the fields above while one of its
  // waves is enqueued
  struct SideLane {
    hipStream_t stream = nullptr;
    WaveBufs* wb = nullptr;
    qsmi::TapInfo1024* dTap = nullptr;
    uint16_t* dTapD0 = nullptr;
    float* dTapVal = nullptr;
    size_t capTap = 0;
    int capTapV = 0;
  } side;
};

// Device buffers of one wave (grown on demand) and pinned host staging.
struct WaveBufs {
  size_t capNodes = 0, capNb = 0, capC = 0, capCand = 0;
  NodeInfo* dInfo = nullptr;
  int* dOff = nullptr;
  float* dVal = nullptr;
  uint32_t* dOrd = nullptr;   // neighbours in pencil-row order (generic spectrum stage)
  float2* dC = nullptr;
  Cand* dCand = nullptr;
  Cand* dCandEx = nullptr;    // [node][K] exact K best as candidates (learned proposal)
  int* dGcnt = nullptr;       // [node][group] values within 2 eps of the group's K-th best
  float* dOutVal = nullptr;
  uint32_t* dOutPos = nullptr;
  uint32_t* dChosen = nullptr;
  NodeInfo* hInfo = nullptr;  // pinned
  int* hOff = nullptr;        // pinned
  float* hVal = nullptr;      // pinned
  int nNodes = 0, K = 0, rowBlocks = 0, nbMax = 0;
  float k = 0.f;
  BandBufs band;
};

struct qsmi_wave {
  qsmi_device* dev = nullptr;
  WaveBufs b;
};


namespace qsmi {

// error channel (thread-local message + code)
int fail(int code, const std::string& msg);
ScanTune default_tune();
unsigned seed_from_unit_float(float seed);

ScanGeom geom_of(const qsmi_device* dev);
bool use_w1k(const qsmi_device* dev);
double scan_cells(const qsmi_device* dev);
float map_scale(const qsmi_device* dev);
int rows_per_block(const qsmi_device* dev);
int row_blocks(const qsmi_device* dev);
long long c_stride(const qsmi_device* dev);
int grow_wave_bufs(qsmi_device* dev, WaveBufs* b, int nNodes, int total, int K);
void free_wave_bufs(WaveBufs* b);
// One wave in three steps (qsmi_scan_wave / qsmi_distribution_wave are
// stage + run + read): host preparation into pinned staging and upload,
// the asynchronous launches, the blocking read-back. exPos/exRadius: the
// optional verbatim disk per node (UINT32_MAX: none).
int stage_wave(qsmi_device* dev, WaveBufs* b, int nNodes, const int* nbStart, const int* offsets,
               const float* values, const float* delta0, const float* seeds, float k,
               const unsigned* exPos = nullptr, const float* exRadius = nullptr);
int run_wave(qsmi_device* dev, WaveBufs* b);
int read_wave(qsmi_device* dev, WaveBufs* b, float* outErrors, unsigned* outPositions, unsigned* outChosen);
// spectrum stage (+ dim1 inverse for 3-D) and rows stage of nNodes nodes
// (returns a qsmi status: tap tables may be allocated)
// count (device, optional; 1000x1000 path only): nodes [0, *count) of the
// nNodes are scanned, the rest skipped at no cost
// ord: work space of one uint32 per neighbour (the generic path sorts them)
// tapsBuilt: the wave's tap lists are already in the device's tap tables
// (the path loop's gather builds them)
int launch_spectrum(qsmi_device* dev, const ScanGeom& g, int nNodes, const NodeInfo* info, const int* off,
                    const float* val, uint32_t* ord, float2* C, const int* count = nullptr, bool tapsBuilt = false);
// tap tables of the 1000^2 spectrum stage for nNodes nodes per launch
int ensure_tap_tables(qsmi_device* dev, int nNodes, int V);
// crossOff > 0 (cross-measurement form, generic path): node q's count layer
// is node q + crossOff of C; rowBlocks then covers one row per FFT
// cand: [node][group][K] (a group is a wavefront of the 1000^2 rows kernel,
// a block of the generic one); gcnt (optional): [node][group] count of the
// group's values within 2 eps of its K-th best (exact-selection nodes)
void launch_rows(qsmi_device* dev, const ScanGeom& g, const NodeInfo* info, const float2* C, int nNodes,
                 int rowBlocks, int K, Cand* cand, int* gcnt, float* map, const int* count = nullptr,
                 int crossOff = 0);
// merge + draw; when simVals != nullptr also writes the drawn TI values and
// the index image of each node's cell (NodeInfo::cell)
// M: candidates per block
void launch_merge_draw(const ScanGeom& g, const NodeInfo* info, const Cand* cand, int nNodes, int candBlocks,
                       int M, int rowsPerBlock, int K, float kf, float* outVal, uint32_t* outPos, uint32_t* chosen,
                       float* simVals, const float* tiVals, int nVar, uint32_t* indexOut, hipStream_t st);

// Exact selection (qsmi_band.hip). exact_selection_on: the device's maps
// take it (no tie quantum, no NaN TI, a rows grid it can track).
bool exact_selection_on(const qsmi_device* dev);
int band_grow(BandBufs* b, int nNodes, int blocksPerNode, int K);
// band_grow for band_select calls of up to nNodes nodes of this device (one
// coefficient layer), so that such calls allocate nothing
int band_reserve(qsmi_device* dev, BandBufs* bb, int nNodes, int K);
void band_free(BandBufs* b);
// counters accumulated since the buffers were made: exact-selection nodes,
// listed (rescored) positions, second-pass rows blocks, full-fallback nodes
int band_read_stats(const BandBufs* b, long long out[4]);
// One node's buffers inside caller memory (no allocation): bytes needed for
// ceil(k) = K, and the view over base (256-byte aligned)
size_t band_scratch_bytes(int K);
void band_view(BandBufs* b, void* base, int K);
// After the rows stage of nNodes nodes (cand from launch_rows, C the
// spectrum-stage buffer the rows stage read): the K best of every node by the
// exact selection (nodes with eps == 0: by their FP32 candidates) into
// outVal/outPos [node][K] (and outCand when given), then the rank draw into
// chosen / the simulation write when given. nbMax: largest neighbour count.
// count (device, optional): only nodes [0, *count) (compacted lists).
int band_select(qsmi_device* dev, const ScanGeom& g, const NodeInfo* info, const Cand* cand, const int* gcnt,
                int nNodes, int K,
                float kf, const float2* C, const int* nbOff, const float* nbVal, int nbMax, float* outVal,
                uint32_t* outPos, Cand* outCand, uint32_t* chosen, float* simVals, const float* tiVals, int nVar,
                uint32_t* indexOut, BandBufs* bb, const int* count = nullptr);
// Per-node API: exact K best of the device's materialised map (dev->dMap)
// for the node in info[0] (neighbours in dev->dOff / dev->dVal), given its
// FP32 K best in outVal/outPos (overwritten).
int band_select_map(qsmi_device* dev, const NodeInfo* info, int nb, int K, float eps, float* outVal,
                    uint32_t* outPos, BandBufs* bb);
void launch_merge_ti(const NodeInfo* info, int nN, int nTI, int K, float kf, const float* val, const uint32_t* pos,
                     long long stride, float quantum, float* simVals, const float* tiVals, int nVar, TiCum cum,
                     uint32_t* indexOut, hipStream_t st);
void launch_info_shift(const NodeInfo* in, NodeInfo* out, int nN, int dh0, int dh1, int dh2, float scale,
                       float epsFactor, hipStream_t st);

// Inputs of the device-resident path loop (qsmi_sim.hip). Device pointers
// are indexed by path position after the hard data (i in [0, nPath)).
struct ResidentLoop {
  int sd[3];      // simulation grid
  int kd[3];      // expanded kernel grid
  int nVar, nExp, n;
  float k;
  int types[8];   // per original variable
  int expVar[16], expCat[16];
  int conv[16], subOf[16];
  std::vector<int> waveStart;  // wave w = order[waveStart[w] .. waveStart[w+1])
  std::vector<int> waveLevel;  // dependency level of wave w (empty: one stream)
  int progressSec = 0;          // > 0: a progress line on stderr about every that many seconds
  int stageTiming = 0;         // k >= 1: events around the stages of every k-th wave
  bool noVerbatim = false;     // -nV: no candidate at the continuation of the nearest neighbour's copy
  std::vector<qsmi_device*> extraDevs;  // TIs 1.. of a several-TI run (same extent as this device's)
  const float* dPathK = nullptr;        // per path node k from a k image (nullptr: the run's k)
  float kMax = 0.f;                     // largest k of the run (candidate list length ceil(kMax))
  // kernel-index image: per path node kernel, per kernel its grid and offset
  const int* dNodeKer = nullptr;
  bool circSim = false;  // -csim: the simulation grid wraps around
  const int* dNodeTi = nullptr;  // per path node TI of a TI-index image (several TIs)
  unsigned tiCum[9] = {};         // several TIs: first cell of each in the concatenation, then the total
  int nKer = 1;
  int kdK[8][3] = {};
  long long kerOff[8] = {};
  const int* dOrder;
  const unsigned* dPathCells;
  const float* dPathSeeds;
  const int* dNbIdx;  // nPath x n offset-list indices
  const int* dNbCnt;
  const int* dOffs;
  const float* dKernel;
  const float* dCatvals;
  float* dDi;
  uint32_t* dIndex;
  const float* dTi;
  // learned proposal (nullptr: exact scan for every node)
  qsmi_lp* lp = nullptr;
  int lpShortlist = 0, lpEvery = 0, lpAudit = 0;
  float lpMse = 0.f, lpMinCover = 0.5f;
  std::vector<unsigned char> waveLp;  // per wave: 1 = run the proposal (empty: every wave)
  mutable long long lpSkipped = 0;    // out: nodes of waves that skipped the proposal
  // training-pair record (device arrays; recCap 0: off)
  int recCap = 0;
  int* dRecOff = nullptr;
  float* dRecVal = nullptr;
  uint32_t* dRecBest = nullptr;
};

// learned proposal inside the path loop (qsmi_lp.hip)
int lp_prepare(qsmi_lp* lp, const qsmi_ti* t, int maxNodes, int K, int shortlist, const int* active, int V,
               int circular);
void lp_propose(qsmi_lp* lp, const qsmi_ti* t, int V, int nN, const NodeInfo* info, const int* nbOff,
                const float* nbVal, const float* di, int nVar, int sd0, int K, float minCover, hipStream_t st);
void lp_flags(qsmi_lp* lp, int nN, int w0, const int* order, const NodeInfo* info, int K, int every, float mseSigma2,
              float minCover, int audit, hipStream_t st);
// ok[i] = 1 for path nodes whose stencil coverage reaches minCover (the only
// nodes the learned proposal serves); 2-D offsets [nOff][2].
int lp_cover_all(qsmi_lp* lp, int nPath, int n, const int* dNbIdx, const int* dNbCnt, const int* dOffs, float minCover,
                 unsigned char* dOk, hipStream_t st);
// compacted: candEx holds the scanned nodes in lp_info_compact order
void lp_combine(qsmi_lp* lp, int nN, const NodeInfo* info, const Cand* candEx, bool compacted, int blocks,
                int rowsPerBlock, int K, int audit, hipStream_t st);
NodeInfo* lp_info_ex(qsmi_lp* lp);
NodeInfo* lp_info_compact(qsmi_lp* lp);
const int* lp_compact_count(qsmi_lp* lp);
Cand* lp_cand(qsmi_lp* lp);
// proposed, fallbacks, overflows, shortlist sum, audit nodes, audit hits, audit best hits, audit positions
int lp_read_counters(qsmi_lp* lp, long long out[8]);
void lp_record(int nN, int w0, const int* order, const NodeInfo* info, const int* nbOff, const float* di, int nVar,
               int sd0, const uint32_t* outPos, int K, int cap, int* recOff, float* recVal, uint32_t* recBest,
               hipStream_t st);

// neighbour sets of every path node on the device (stream-ordered)
// dependency levels of the path nodes from their neighbours' path indices
// (level[i] = 1 + max level of i's path neighbours); *dErr = 1 when a wait
// hit its bound (the levels are then not valid)
int levels_device(size_t nPath, int n, const int* dNbPj, const int* dNbCnt, int* dLevel, int* dErr, hipStream_t st);
int nb_search_device(int D, const unsigned* sd, size_t nPath, const unsigned* dCells, const int* dPostS,
                     const int* dOffs, int nOff, int n, int* dNbIdx, int* dNbPj, int* dNbCnt, hipStream_t st,
                     const int* dNNode = nullptr, const int* dNodeKer = nullptr, const int* dOffStart = nullptr,
                     const int* dOffCount = nullptr, int circSim = 0);

// stageMs: [0] spectrum (+ dim1 inverse), [1] gather, [2] rows + select,
// [3] merge + draw + write; filled when R.stageTiming.
// bandStats (optional): the exact selection's counters over the run
// (band_read_stats, summed over the TIs)
int run_resident_waves(qsmi_device* dev, const ResidentLoop& R, double* stageMs, long long* scanned, int* waves,
                       long long* bandStats = nullptr);
// cross-measurement form: the count layer's coefficients and the TI cells'
// completeness (forces the generic kernels)
int set_cross(qsmi_device* dev, const float* alphaX, const uint8_t* centreOk);

}  // namespace qsmi
