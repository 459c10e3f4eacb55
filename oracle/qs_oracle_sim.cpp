// QuickSampling CPU oracle, simulation driver — TEST INFRASTRUCTURE ONLY.
//
// Serial restatement of the reference QS node loop for one training image:
//   include/simulation.hpp:304-666  (node body: gather, one-hot, sample, write)
//   include/quantileSamplingModule.hpp:90-126, 342-577 (sample, conversion,
//                                                      scan, select, draw)
//   src/qs.cpp:1604-1641            (posterior path)
// The reference runs the loop over OpenMP threads and spin-waits on
// neighbours that are earlier in the path but not yet written
// (simulation.hpp:463-473), which makes its result identical to the serial
// order restated here.
//
// Supported here (and by the device): fully-NaN and partially informed cells
// to simulate (a partially informed cell is its own first neighbour with the
// values it holds and keeps them, include/simulation.hpp:324-335, 581-584,
// 658-666), fully informed hard-data cells, one value of n for all
// variables, one TI, no
// -nV / -fs / kernel images / transforms. A TI with NaN (continuous
// variables) runs in the cross-measurement form (`cross`): a cell with any
// NaN variable loses all of them (src/qs.cpp:1436-1466), the map is
// -|E / C^4| with E the masked mismatch layer and C the weighted count of
// TI values under the neighbours times the centre's validity
// (CPUThreadDevice::maskLayerWithVariable, src/CPUThreadDevice.cpp:500-524;
// compensateMissingData, src/CPUThreadAcceleratorDevice.cpp:414-425), and
// the marginal draw skips incomplete cells (simulation.hpp:586-654).
#include <algorithm>
#include <array>
#include <climits>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <vector>

#include "fft_cpu.hpp"

using namespace orc;

extern "C" {
// from qs_oracle.cpp
int orc_ti_subimages(int D, const unsigned* s, const unsigned* f, int nVar, const int* types, const float* data,
                     const int* ncat, const float* catvals, int reflected, float* out);
int orc_coef_matrix(int nVar, const int* types, const int* ncat, float* alpha, int* convTypes, int* subVarOf);
int orc_coef_matrix_x(int nVar, const int* types, float* alpha, float* alphaX, int* convTypes, int* subVarOf);
int orc_ti_subimages_x(int D, const unsigned* s, const unsigned* f, int nVar, const int* types, const float* data,
                       int reflected, float* out);
void orc_fft_size(int D, const unsigned* src, unsigned* out);
float orc_convert_neighbors(int D, int nb, const int* off, int nExp, const float* vals, int nSub,
                            const int* convTypes, const int* subVarOf, int nVarOrig, const int* typesOrig,
                            const unsigned* kdims, const float* kernel, float* u);
void orc_map_direct_f64(int D, const unsigned* s, int nSub, const float* Y, const float* alpha, int nb,
                        const int* off, const float* u, double delta0, int circular, double* out);
void orc_map_refflow_f32(int D, const unsigned* s, const unsigned* f, int nSub, const float* Yhat,
                         const float* alpha, int nb, const int* off, const float* u, float delta0, int circular,
                         int trueMismatch, float* out);
void orc_map_refflow_f64(int D, const unsigned* s, const unsigned* f, int nSub, const double* Yhat,
                         const float* alpha, int nb, const int* off, const float* u, float delta0, int circular,
                         int trueMismatch, float* out);
void orc_ti_spectra_f32(int D, const unsigned* f, int nSub, const float* imgs, float* out);
void orc_ti_spectra_f64(int D, const unsigned* f, int nSub, const float* imgs, double* out);
void orc_fkst(const float* data, unsigned N, unsigned short k, float seed, float* out, unsigned* pos);
unsigned orc_index_to_position(int D, const unsigned* f, const unsigned* s, const int* mn, unsigned index);
unsigned orc_position_to_index(int D, const unsigned* f, const unsigned* s, const int* mn, unsigned position);
int orc_select_device_f64(int D, const unsigned* s, const double* map, const int* lo, const int* hi, int K,
                          float seed, float quantum, double* outVal, unsigned* outPos);
int orc_select_device(int D, const unsigned* s, const float* map, const int* lo, const int* hi, int K, float seed,
                      float quantum, float* outVal, unsigned* outPos);
int orc_index_with_delta(int D, const unsigned* dims, unsigned position, const int* delta, unsigned* location);
unsigned orc_seed_from_unit_float(float seed);
uint32_t orc_tie_hash(uint32_t seedbits, uint32_t position);
void orc_bbox(int D, int nb, const int* off, int* mn, int* mx);
unsigned orc_draw_rank(float seed, float k, unsigned nTI);

struct OrcSim {
  int D;
  unsigned tiDims[3];
  int nVar;
  const int* types;  // 0 continuous, 1 categorical
  const float* ti;   // cells x nVar
  const int* ncat;   // per categorical variable
  const float* catvals;
  unsigned simDims[3];
  unsigned kdims[3];
  const float* kernel;  // expanded kernel, cells(kdims) x nExp
  int nOffsets;
  const int* offsets;  // sorted offset list, nOffsets x D
  int n;
  float k;
  const unsigned* path;  // full path (hard data first)
  const float* seeds;    // per path entry
  unsigned beginPath;
  unsigned pathLen;
  int mode;       // 0: direct FP64 + device rule; 1: refflow FP32 + fKst; 2: refflow FP64 + fKst
  float quantum;  // tie quantum of the device rule (0: none)
  int cross;      // 1: TI with NaN, cross-measurement form (mode 0 only)
  int noVerbatim; // 1: -nV, no candidate at the continuation of neighbour 1's copy
  int nTI;        // training images in `ti`, one after the other (<= 1: one; mode 0 only when > 1)
  const float* nImage;  // per simulation cell n (nullptr: n); n must then be the largest value
  const float* kImage;  // per simulation cell k (nullptr, or values <= 0: k)
  int circularSim;      // 1: -csim, neighbour cells wrap around the simulation grid
  int circularTI;       // 1: -cti, every TI position valid, neighbourhoods wrapped
  const float* tiImage; // several TIs: per simulation cell the only TI sampled (nullptr: all)
  const unsigned* tiDimsAll;  // several TIs: nTI x 3 extents (nullptr: tiDims for all)
  const float* di0;     // the simulation grid as given (nullptr: no partially informed cells): a
                        // node's own values when it is reached (include/simulation.hpp:463-473)
};

struct OrcSimCtx {
  OrcSim p;
  int D;
  unsigned s[3], f[3], sd[3];
  size_t S, F, SD;
  int nSub = 0, nExp = 0;
  std::vector<float> alpha, alphaX;
  std::vector<int> convTypes, subVarOf;
  std::vector<float> tiX;        // cross form: the TI with incomplete cells all-NaN
  std::vector<uint8_t> centreOk; // cross form: TI cell complete
  std::vector<float> Y;       // unreflected sub-images on the TI grid (nSub x S)
  std::vector<float> Yhat32;  // reflected spectra (FP32)
  std::vector<double> Yhat64; // reflected spectra (FP64)
  std::vector<float> catflat;
  std::vector<int> typesOrig;
  std::vector<unsigned> posterior;
  const unsigned* index = nullptr;  // importDataIndex state for -nV (orc_sim_set_index)
  int nTI = 1;
  std::vector<std::vector<float>> Yx;  // sub-images of TIs 1.. (several TIs)
  std::vector<std::array<unsigned, 3>> sT;  // extents of the TIs
  std::vector<size_t> cumT;                 // first cell of each TI, then the total
};

// include/simulation.hpp:394: with circularSim the wrapped cell of
// DataImage::indexWithDelta is taken
static inline int wrap_if(int x, int n, int circ) { return circ ? ((x % n) + n) % n : x; }

void* orc_sim_create(const OrcSim* p) {
  auto* c = new OrcSimCtx();
  c->p = *p;
  if (p->cross) c->p.mode = 0;  // the cross form has the direct FP64 map only
  c->D = p->D;
  c->S = c->F = c->SD = 1;
  for (int i = 0; i < 3; ++i) {
    c->s[i] = (i < p->D) ? p->tiDims[i] : 1;
    c->sd[i] = (i < p->D) ? p->simDims[i] : 1;
  }
  orc_fft_size(p->D, c->s, c->f);
  for (int i = p->D; i < 3; ++i) c->f[i] = 1;
  for (int i = 0; i < 3; ++i) {
    c->S *= c->s[i];
    c->F *= c->f[i];
    c->SD *= c->sd[i];
  }
  int nCatVars = 0;
  c->nExp = 0;
  for (int v = 0; v < p->nVar; ++v) {
    c->typesOrig.push_back(p->types[v]);
    if (p->types[v] == 0)
      c->nExp += 1;
    else
      c->nExp += p->ncat[nCatVars++];
  }
  c->alpha.resize(64 * 64);
  c->convTypes.resize(64);
  c->subVarOf.resize(64);
  if (p->cross) {
    // src/qs.cpp:1447-1466: a cell with any NaN variable becomes all-NaN
    c->tiX.assign(p->ti, p->ti + c->S * p->nVar);
    c->centreOk.assign(c->S, 1);
    for (size_t q = 0; q < c->S; ++q) {
      bool nan = false;
      for (int v = 0; v < p->nVar; ++v) nan |= std::isnan(c->tiX[q * p->nVar + v]);
      if (nan) {
        for (int v = 0; v < p->nVar; ++v) c->tiX[q * p->nVar + v] = NAN;
        c->centreOk[q] = 0;
      }
    }
    c->p.ti = c->tiX.data();
    c->alphaX.resize(64 * 64);
    c->nSub = orc_coef_matrix_x(p->nVar, p->types, c->alpha.data(), c->alphaX.data(), c->convTypes.data(),
                                c->subVarOf.data());
    c->alpha.resize((size_t)c->nSub * c->nSub);
    c->alphaX.resize((size_t)c->nSub * c->nSub);
    c->Y.resize((size_t)c->nSub * c->S);
    orc_ti_subimages_x(p->D, c->s, c->s, p->nVar, p->types, c->tiX.data(), 0, c->Y.data());
  } else {
    c->nSub = orc_coef_matrix(p->nVar, p->types, p->ncat, c->alpha.data(), c->convTypes.data(), c->subVarOf.data());
    c->alpha.resize((size_t)c->nSub * c->nSub);
    // unreflected sub-images on the TI grid
    c->Y.resize((size_t)c->nSub * c->S);
    orc_ti_subimages(p->D, c->s, c->s, p->nVar, p->types, p->ti, p->ncat, p->catvals, 0, c->Y.data());
    c->nTI = p->nTI > 1 ? p->nTI : 1;
    c->sT.resize(c->nTI);
    c->cumT.assign(c->nTI + 1, 0);
    for (int t = 0; t < c->nTI; ++t) {
      size_t cells = 1;
      for (int i = 0; i < 3; ++i) {
        c->sT[t][i] = (t > 0 && p->tiDimsAll && i < p->D) ? p->tiDimsAll[t * 3 + i] : c->s[i];
        cells *= c->sT[t][i];
      }
      c->cumT[t + 1] = c->cumT[t] + cells;
    }
    for (int t = 1; t < c->nTI; ++t) {
      const size_t St = c->cumT[t + 1] - c->cumT[t];
      c->Yx.emplace_back((size_t)c->nSub * St);
      orc_ti_subimages(p->D, c->sT[t].data(), c->sT[t].data(), p->nVar, p->types, p->ti + c->cumT[t] * p->nVar,
                       p->ncat, p->catvals, 0, c->Yx.back().data());
    }
    if (c->nTI > 1) c->p.mode = 0;
  }
  if (p->mode == 1 || p->mode == 2) {
    std::vector<float> refl((size_t)c->nSub * c->F);
    orc_ti_subimages(p->D, c->s, c->f, p->nVar, p->types, p->ti, p->ncat, p->catvals, 1, refl.data());
    size_t H = (size_t)(c->f[0] / 2 + 1) * c->f[1] * c->f[2];
    if (p->mode == 1) {
      c->Yhat32.resize(2 * H * c->nSub);
      orc_ti_spectra_f32(p->D, c->f, c->nSub, refl.data(), c->Yhat32.data());
    } else {
      c->Yhat64.resize(2 * H * c->nSub);
      orc_ti_spectra_f64(p->D, c->f, c->nSub, refl.data(), c->Yhat64.data());
    }
  }
  // posterior path — src/qs.cpp:1615-1641: hard data 0, path entry i -> i
  c->posterior.assign(c->SD, UINT_MAX);
  for (unsigned i = 0; i < p->beginPath; ++i) c->posterior[p->path[i]] = 0;
  for (unsigned i = p->beginPath; i < p->pathLen; ++i) c->posterior[p->path[i]] = i - p->beginPath;
  return c;
}

void orc_sim_free(void* h) { delete static_cast<OrcSimCtx*>(h); }

void orc_sim_posterior(void* h, unsigned* out) {
  auto* c = static_cast<OrcSimCtx*>(h);
  std::copy(c->posterior.begin(), c->posterior.end(), out);
}

int orc_sim_nsub(void* h) { return static_cast<OrcSimCtx*>(h)->nSub; }

// Neighbour gather of one path node — include/simulation.hpp:447-511 (the
// centre counts as the first neighbour; a cell is informed when its
// posterior order is <= the node's). Writes offset-list indices; returns the
// count.
int orc_sim_gather(void* h, unsigned pathIdx, int* outIdx) {
  auto* c = static_cast<OrcSimCtx*>(h);
  const OrcSim& p = c->p;
  unsigned cell = p.path[p.beginPath + pathIdx];
  unsigned ord = pathIdx;
  int cnt = 0;
  int cc[3];
  size_t rem = cell;
  for (int i = 0; i < c->D; ++i) {
    cc[i] = int(rem % c->sd[i]);
    rem /= c->sd[i];
  }
  // include/simulation.hpp:337-343: the cell's n from the n image
  int nCell = p.n;
  if (p.nImage) {
    const float v = p.nImage[cell];
    nCell = (std::isfinite(v) && v > 0.f && v < 1e6f) ? std::min((int)v, p.n) : 0;
  }
  for (int o = 0; o < p.nOffsets && cnt < nCell; ++o) {
    bool inside = true;
    size_t t = 0, mul = 1;
    for (int i = 0; i < c->D; ++i) {
      int x = wrap_if(cc[i] + p.offsets[o * c->D + i], int(c->sd[i]), p.circularSim);
      inside &= (x >= 0 && x < int(c->sd[i]));
      t += (size_t)x * mul;
      mul *= c->sd[i];
    }
    if (!inside) continue;
    if (c->posterior[t] <= ord) outIdx[cnt++] = o;
  }
  return cnt;
}

// Neighbours of one node as the serial loop sees them: offsets, converted
// values (QSM conversion) and delta0. Returns the neighbour count.
static int node_inputs(OrcSimCtx* c, const float* di, unsigned pathIdx, std::vector<int>& off, std::vector<float>& u,
                       float* delta0, std::vector<float>* rawOut) {
  const OrcSim& p = c->p;
  const int D = c->D;
  std::vector<int> idx(p.n > 0 ? p.n : 1);
  int nb = orc_sim_gather(c, pathIdx, idx.data());
  unsigned cell = p.path[p.beginPath + pathIdx];
  off.assign((size_t)nb * D, 0);
  std::vector<float> vals((size_t)nb * c->nExp);
  if (rawOut) rawOut->assign((size_t)nb * p.nVar, NAN);
  int cc[3];
  size_t rem = cell;
  for (int i = 0; i < D; ++i) {
    cc[i] = int(rem % c->sd[i]);
    rem /= c->sd[i];
  }
  for (int j = 0; j < nb; ++j) {
    size_t t = 0, mul = 1;
    for (int i = 0; i < D; ++i) {
      int d = p.offsets[idx[j] * D + i];
      off[(size_t)j * D + i] = d;
      t += (size_t)wrap_if(cc[i] + d, int(c->sd[i]), p.circularSim) * mul;
      mul *= c->sd[i];
    }
    // one-hot expansion — simulation.hpp:512-537. The node's own cell holds
    // what it was given when the serial loop reaches it, whatever `di` holds.
    const bool isCentre = (t == cell);
    int e = 0, catIdx = 0;
    const float* cv = p.catvals;
    for (int v = 0; v < p.nVar; ++v) {
      float x = isCentre ? (p.di0 ? p.di0[(size_t)cell * p.nVar + v] : NAN) : di[t * p.nVar + v];
      if (rawOut) (*rawOut)[(size_t)j * p.nVar + v] = x;
      if (p.types[v] == 0) {
        vals[(size_t)j * c->nExp + e++] = x;
      } else {
        for (int q = 0; q < p.ncat[catIdx]; ++q) vals[(size_t)j * c->nExp + e++] = (x == cv[q]) ? 1.f : 0.f;
        cv += p.ncat[catIdx];
        catIdx++;
      }
    }
  }
  u.assign((size_t)nb * c->nSub, 0.f);
  *delta0 = orc_convert_neighbors(D, nb, off.data(), c->nExp, vals.data(), c->nSub, c->convTypes.data(),
                                  c->subVarOf.data(), p.nVar, c->typesOrig.data(), p.kdims, p.kernel, u.data());
  return nb;
}

// Cross-measurement map of one node (FP64): E from the regular layer, C from
// the cross layer times the centre cell's completeness, out = -|E / C^4|
// (-inf where C = 0 or outside the valid box). scale (nullable) receives
// the magnitude the FP32 error of out is measured against:
// sE / C^4 + 4 |E| sC / C^5 with sE = sum_j w (x^2 m + y^2) and sC = sum_j w.
static void cross_map(OrcSimCtx* c, int nb, const std::vector<int>& off, const std::vector<float>& u, double* out,
                      double* scale) {
  const int D = c->D;
  std::vector<double> E(c->S), C(c->S);
  orc_map_direct_f64(D, c->s, c->nSub, c->Y.data(), c->alpha.data(), nb, off.data(), u.data(), 0.0, 0, E.data());
  orc_map_direct_f64(D, c->s, c->nSub, c->Y.data(), c->alphaX.data(), nb, off.data(), u.data(), 0.0, 0, C.data());
  std::vector<double> sE;
  double sC = 0.0;
  if (scale) {
    // sum_j w (x^2 m(p+d_j) + y^2(p+d_j)) via the same direct form: |P2| on
    // the mask image and |P0| on the y^2 image, per variable
    std::vector<float> al((size_t)c->nSub * c->nSub, 0.f);
    std::vector<float> ua(u.size());
    for (size_t q = 0; q < u.size(); ++q) ua[q] = std::fabs(u[q]);
    for (int v = 0; v < c->p.nVar; ++v) {
      al[(3 * v) + (3 * v) * c->nSub] = 1.f;
      al[(3 * v + 2) + (3 * v + 2) * c->nSub] = 1.f;
    }
    sE.resize(c->S);
    orc_map_direct_f64(D, c->s, c->nSub, c->Y.data(), al.data(), nb, off.data(), ua.data(), 0.0, 0, sE.data());
    for (int j = 0; j < nb; ++j)
      for (int v = 0; v < c->p.nVar; ++v) sC += std::fabs(u[(size_t)j * c->nSub + 3 * v]);
  }
  for (size_t q = 0; q < c->S; ++q) {
    const double cm = std::isfinite(C[q]) ? C[q] * (double)c->centreOk[q] : C[q];
    double v;
    if (!std::isfinite(E[q]) || !std::isfinite(cm) || cm == 0.0)
      v = -INFINITY;
    else
      v = -std::fabs(E[q] / (cm * cm * cm * cm));
    out[q] = v;
    if (scale) {
      if (std::isfinite(v)) {
        const double c4 = cm * cm * cm * cm;
        scale[q] = std::fabs(sE[q]) / c4 + 4.0 * std::fabs(E[q]) * sC / (c4 * std::fabs(cm));
      } else {
        scale[q] = 0.0;
      }
    }
  }
}

// One node's scan and draw given the DI state. Writes the K best
// (value, TI position) pairs sorted best first, and returns the TI position
// The marginal draw of one TI (include/simulation.hpp:586-654): cell
// floor(seed * cells); where that cell has a NaN variable, the walk over the
// complete cells: rank r = floor(seed * (C / nVar)) * nVar with C the number
// of complete cells, and the first cell whose running count of complete
// cells reaches r (r = 0 stops at cell 0, complete or not).
extern "C" unsigned orc_marginal_cell(const uint8_t* complete, size_t S, int nVar, float seed) {
  const unsigned position = (unsigned)int(std::floor(seed * (float)(unsigned)S));
  if (complete[position]) return position;
  unsigned total = 0;
  for (size_t q = 0; q < S; ++q) total += complete[q];
  const unsigned r = (unsigned)int(std::floor(seed * (float)(total / (unsigned)nVar))) * (unsigned)nVar;
  unsigned run = 0;
  for (size_t q = 0; q < S; ++q) {
    run += complete[q];
    if (r <= run) return (unsigned)q;
  }
  return position;
}

// chosen by the draw (or the marginal draw when only the centre is a
// neighbour, simulation.hpp:586-606). *nbOut = neighbour count.
// -nV (include/simulation.hpp:567-580): the TI position the nearest
// neighbour (entry 1 of the neighbour list) was copied from, moved back by
// its offset with DataImage::indexWithDelta (wrapped, validity ignored), as
// the reference index QuantileSamplingModule hands to setValueInErrorArray
// (include/quantileSamplingModule.hpp:566-571): cvtPositionToIndex with the
// node's bounding box.
static unsigned verbatim_index(const OrcSimCtx* c, unsigned pathIdx, const std::vector<int>& off, int nb) {
  const OrcSim& p = c->p;
  const int D = c->D;
  unsigned cell = p.path[p.beginPath + pathIdx];
  size_t rem = cell, t = 0, mul = 1;
  for (int i = 0; i < D; ++i) {
    int cc = int(rem % c->sd[i]);
    rem /= c->sd[i];
    t += (size_t)wrap_if(cc + off[(size_t)1 * D + i], int(c->sd[i]), p.circularSim) * mul;
    mul *= c->sd[i];
  }
  const unsigned from = c->index[t];
  int back[3] = {0, 0, 0};
  for (int i = 0; i < D; ++i) back[i] = -off[(size_t)1 * D + i];
  unsigned vpos = 0;
  orc_index_with_delta(D, c->s, from, back, &vpos);
  int mn[3], mx[3];
  orc_bbox(D, nb, off.data(), mn, mx);
  return orc_position_to_index(D, c->f, c->s, mn, vpos);
}

void orc_sim_set_index(void* h, const unsigned* index) { static_cast<OrcSimCtx*>(h)->index = index; }

// Several TIs (include/quantileSamplingModule.hpp:90-126 with
// determineDistribution :478-577): every TI's K best by the device rule on
// its FP64 map, then all of them ordered together — value (or its tie key)
// descending, then the hash and the position of the encoded index
// position * nTI + TI — and the rank floor(seed k ceil(nTI/k) / nTI)
// (float arithmetic) taken, clamped to the last valid entry. Returns the
// encoded index; outVal/outPos get the merged K best (encoded).
static unsigned node_multi_ti(OrcSimCtx* c, int nb, const std::vector<int>& off, const std::vector<float>& u,
                              float delta0, float seed, int K, float kLoc, int tiReq, float* outVal,
                              unsigned* outPos) {
  const OrcSim& p = c->p;
  const int D = c->D;
  const unsigned nT = (unsigned)c->nTI;
  const uint32_t sb = orc_seed_from_unit_float(seed);
  struct Cd {
    float key;
    uint32_t h, pos;
    float val;
  };
  auto better = [](const Cd& a, const Cd& b) {
    if (a.key != b.key) return a.key > b.key;
    if (a.h != b.h) return a.h > b.h;
    return a.pos < b.pos;
  };
  std::vector<Cd> all;
  for (unsigned t = 0; t < nT; ++t) {
    // a TI-index image: only that TI's device is updated
    // (include/quantileSamplingModule.hpp:460-464)
    if (tiReq >= 0 && (int)t != tiReq) continue;
    const unsigned* st = c->sT[t].data();
    const size_t St = c->cumT[t + 1] - c->cumT[t];
    int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    for (int i = 0; i < D; ++i) hi[i] = int(st[i]) - 1;
    std::vector<double> m64(St);
    orc_map_direct_f64(D, st, c->nSub, t == 0 ? c->Y.data() : c->Yx[t - 1].data(), c->alpha.data(), nb, off.data(),
                       u.data(), (double)delta0, c->p.circularTI, m64.data());
    // every TI's K best by the device rule on its FP64 map, then the merge
    // of their values (as float) across TIs
    std::vector<double> v64(K);
    std::vector<unsigned> ps(K);
    orc_select_device_f64(D, st, m64.data(), lo, hi, K, seed, p.quantum, v64.data(), ps.data());
    for (int i = 0; i < K; ++i) {
      const float vi = (float)v64[i];
      if (ps[i] == UINT_MAX || !(vi > -INFINITY)) continue;
      Cd cd;
      cd.val = vi;
      cd.key = p.quantum > 0.f ? std::rint(vi * (1.f / p.quantum)) : vi;
      cd.pos = ps[i] * nT + t;
      cd.h = orc_tie_hash(sb, cd.pos);
      all.push_back(cd);
    }
  }
  std::sort(all.begin(), all.end(), better);
  for (int i = 0; i < K; ++i) {
    outVal[i] = i < (int)all.size() ? all[i].val : -INFINITY;
    outPos[i] = i < (int)all.size() ? all[i].pos : UINT_MAX;
  }
  if (all.empty()) return UINT_MAX;
  int r = (int)std::floor(seed * kLoc * (std::ceil((float)nT / kLoc) / (float)nT));
  if (r < 0) r = 0;
  if (r > (int)all.size() - 1) r = (int)all.size() - 1;
  return all[r].pos;
}

static unsigned node_impl(OrcSimCtx* c, const float* di, unsigned pathIdx, int K, float kLoc, float* outVal,
                          unsigned* outPos, int* nbOut);

unsigned orc_sim_node(void* h, const float* di, unsigned pathIdx, int K, float* outVal, unsigned* outPos,
                      int* nbOut) {
  auto* c = static_cast<OrcSimCtx*>(h);
  const OrcSim& p = c->p;
  // include/simulation.hpp:353-357: the cell's k from the k image (values
  // <= 0 leave the run's k, quantileSamplingModule.hpp:92-93); the k best
  // lists hold ceil(k) entries, whatever K the caller reads back
  float kLoc = p.k;
  if (p.kImage) {
    const float v = p.kImage[p.path[p.beginPath + pathIdx]];
    if (v > 0.f) kLoc = v;
  }
  const int KL = (int)std::ceil(kLoc);
  std::vector<float> v(KL);
  std::vector<unsigned> ps(KL);
  const unsigned r = node_impl(c, di, pathIdx, KL, kLoc, v.data(), ps.data(), nbOut);
  for (int i = 0; i < K; ++i) {
    outVal[i] = i < KL ? v[i] : -INFINITY;
    outPos[i] = i < KL ? ps[i] : UINT_MAX;
  }
  return r;
}

static unsigned node_impl(OrcSimCtx* c, const float* di, unsigned pathIdx, int K, float kLoc, float* outVal,
                          unsigned* outPos, int* nbOut) {
  const OrcSim& p = c->p;
  const int D = c->D;
  float seed = p.seeds[p.beginPath + pathIdx];
  for (int i = 0; i < K; ++i) {
    outVal[i] = -INFINITY;
    outPos[i] = UINT_MAX;
  }
  std::vector<int> off;
  std::vector<float> u;
  float delta0 = 0.f;
  int nb = node_inputs(c, di, pathIdx, off, u, &delta0, nullptr);
  *nbOut = nb;
  // only the cell itself: the marginal unless the cell holds values
  // (include/simulation.hpp:581-584: then its own values are scanned)
  bool centreData = false;
  if (nb == 1 && p.di0) {
    const unsigned cell = p.path[p.beginPath + pathIdx];
    for (int v = 0; v < p.nVar; ++v) centreData |= !std::isnan(p.di0[(size_t)cell * p.nVar + v]);
  }
  if (nb <= 1 && !centreData) {
    // marginal draw: position = floor(seed * cells) in float, over all TIs in
    // order (include/simulation.hpp:586-606); several TIs: the encoded index
    if (c->nTI > 1) {
      const unsigned q = (unsigned)int(std::floor(seed * (float)(unsigned)c->cumT[c->nTI]));
      unsigned t = 0;
      while ((int)t + 1 < c->nTI && q >= c->cumT[t + 1]) ++t;
      return (q - (unsigned)c->cumT[t]) * (unsigned)c->nTI + t;
    }
    if (p.cross) return orc_marginal_cell(c->centreOk.data(), c->S, p.nVar, seed);
    return (unsigned)int(std::floor(seed * (float)(unsigned)c->S));
  }
  if (c->nTI > 1) {
    const int tiReq = p.tiImage ? (int)p.tiImage[p.path[p.beginPath + pathIdx]] : -1;
    return node_multi_ti(c, nb, off, u, delta0, seed, K, kLoc, tiReq, outVal, outPos);
  }
  if (p.mode == 0) {
    std::vector<double> map(c->S);
    if (p.cross)
      cross_map(c, nb, off, u, map.data(), nullptr);
    else
      orc_map_direct_f64(D, c->s, c->nSub, c->Y.data(), c->alpha.data(), nb, off.data(), u.data(), (double)delta0, c->p.circularTI,
                         map.data());
    if (p.noVerbatim && c->index) {
      // the index's entry in position space
      int mn[3], mx[3];
      orc_bbox(D, nb, off.data(), mn, mx);
      map[orc_index_to_position(D, c->f, c->s, mn, verbatim_index(c, pathIdx, off, nb))] = -INFINITY;
    }
    // device rule on the exact map
    uint32_t sb = orc_seed_from_unit_float(seed);
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
    for (size_t q = 0; q < c->S; ++q) {
      double v = map[q];
      if (!std::isfinite(v)) continue;
      Cd cd{p.quantum > 0 ? std::rint(v / (double)p.quantum) : v, orc_tie_hash(sb, (uint32_t)q), (uint32_t)q, v};
      if ((int)best.size() < K) {
        best.push_back(cd);
        std::sort(best.begin(), best.end(), better);
      } else if (better(cd, best.back())) {
        best.back() = cd;
        std::sort(best.begin(), best.end(), better);
      }
    }
    for (int i = 0; i < (int)best.size(); ++i) {
      outVal[i] = (float)best[i].val;
      outPos[i] = best[i].pos;
    }
  } else {
    std::vector<float> map(c->F);
    if (p.mode == 1)
      orc_map_refflow_f32(D, c->s, c->f, c->nSub, c->Yhat32.data(), c->alpha.data(), nb, off.data(), u.data(),
                          delta0, c->p.circularTI, 1, map.data());
    else
      orc_map_refflow_f64(D, c->s, c->f, c->nSub, c->Yhat64.data(), c->alpha.data(), nb, off.data(), u.data(),
                          delta0, c->p.circularTI, 1, map.data());
    if (p.noVerbatim && c->index) map[verbatim_index(c, pathIdx, off, nb)] = -INFINITY;  // index space
    std::vector<unsigned> pos(K);
    orc_fkst(map.data(), (unsigned)c->F, (unsigned short)K, seed, outVal, pos.data());
    int mn[3], mx[3];
    orc_bbox(D, nb, off.data(), mn, mx);
    for (int i = 0; i < K; ++i)
      outPos[i] = (pos[i] == UINT_MAX) ? UINT_MAX : orc_index_to_position(D, c->f, c->s, mn, pos[i]);
  }
  // QSM::sample — sort the candidates (already best first), rank by the seed
  unsigned r = orc_draw_rank(seed, kLoc, 1);
  if ((int)r >= K) r = K - 1;
  return outPos[r];
}

// The node's exact FP64 map (position space, -inf outside the valid box) and
// the magnitude s(p) of its cancelling terms: per continuous variable
// sum_j |w_j| (x_j^2 + y(p+d_j)^2), per categorical variable 2 sum_j |w_j|.
// Returns the neighbour count.
int orc_sim_node_map(void* h, const float* di, unsigned pathIdx, double* map, double* scale) {
  auto* c = static_cast<OrcSimCtx*>(h);
  const int D = c->D;
  std::vector<int> off;
  std::vector<float> u;
  std::vector<float> raw;
  float delta0 = 0.f;
  int nb = node_inputs(c, di, pathIdx, off, u, &delta0, &raw);
  if (c->p.cross) {
    cross_map(c, nb, off, u, map, scale);
    return nb;
  }
  orc_map_direct_f64(D, c->s, c->nSub, c->Y.data(), c->alpha.data(), nb, off.data(), u.data(), (double)delta0, c->p.circularTI,
                     map);
  if (!scale) return nb;
  std::fill(scale, scale + c->S, 0.0);
  // |w| per neighbour and sub-variable from the converted values: w = u of the
  // P0 sub-variable (continuous) or |u| summed over the classes (categorical)
  int b = 0;
  for (int v = 0; v < c->p.nVar; ++v) {
    if (c->p.types[v] == 0) {
      // s += sum_j |w_j| y^2(p+d_j) + |w_j| x_j^2 : correlate |w| with the y^2 image
      std::vector<float> us((size_t)nb * c->nSub, 0.f);
      std::vector<float> al((size_t)c->nSub * c->nSub, 0.f);
      al[b + b * c->nSub] = 1.f;
      double cst = 0.0;
      for (int j = 0; j < nb; ++j) {
        float w = std::fabs(u[(size_t)j * c->nSub + b]);  // P0 sub-variable = w (0 for NaN values)
        us[(size_t)j * c->nSub + b] = w;
        float x = raw[(size_t)j * c->p.nVar + v];
        if (!std::isnan(x)) cst += (double)w * x * x;
      }
      std::vector<double> part(c->S);
      orc_map_direct_f64(D, c->s, c->nSub, c->Y.data(), al.data(), nb, off.data(), us.data(), cst, 0, part.data());
      for (size_t q = 0; q < c->S; ++q) scale[q] += part[q];
      b += 2;
    } else {
      int ci = 0;
      for (int v2 = 0; v2 < v; ++v2) ci += (c->p.types[v2] != 0);
      int ncl = c->p.ncat[ci];
      double sw = 0.0;
      for (int j = 0; j < nb; ++j) {
        double wj = 0.0;
        for (int q = 0; q < ncl; ++q) wj = std::max(wj, (double)std::fabs(u[(size_t)j * c->nSub + b + q]));
        sw += wj;
      }
      for (size_t q = 0; q < c->S; ++q) scale[q] += 2.0 * sw;
      b += ncl;
    }
  }
  return nb;
}

// Norm that scales the FP32 map error of a node (test criterion of the GPU
// maps): sum_A rms(Y_A) * sum_j |v_jA|, v_jA = sum_B alpha_AB u_jB, rms over
// the FFT grid (zero padding included). Single TI, plain form.
double orc_sim_node_norm(void* h, const float* di, unsigned pathIdx) {
  auto* c = static_cast<OrcSimCtx*>(h);
  std::vector<int> off;
  std::vector<float> u;
  float delta0 = 0.f;
  int nb = node_inputs(c, di, pathIdx, off, u, &delta0, nullptr);
  double total = 0.0;
  for (int A = 0; A < c->nSub; ++A) {
    double ss = 0.0;
    for (size_t q = 0; q < c->S; ++q) ss += (double)c->Y[(size_t)A * c->S + q] * (double)c->Y[(size_t)A * c->S + q];
    const double rms = std::sqrt(ss / (double)c->F);
    double sv = 0.0;
    for (int j = 0; j < nb; ++j) {
      double v = 0.0;
      for (int B = 0; B < c->nSub; ++B) v += (double)c->alpha[A + B * c->nSub] * (double)u[(size_t)j * c->nSub + B];
      sv += std::fabs(v);
    }
    total += rms * sv;
  }
  return total;
}

// Serial simulation: fills di (cells x nVar) and idx (importDataIndex).
void orc_sim_run(void* h, float* di, unsigned* idxOut) {
  auto* c = static_cast<OrcSimCtx*>(h);
  const OrcSim& p = c->p;
  int K = (int)std::ceil(p.k);
  std::vector<float> v(K);
  std::vector<unsigned> pos(K);
  const unsigned* keep = c->index;
  c->index = idxOut;  // the run's importDataIndex (hard data 0, as the reference's empty copy)
  for (unsigned i = 0; i + p.beginPath < p.pathLen; ++i) {
    unsigned cell = p.path[p.beginPath + i];
    int nb;
    unsigned tpos = orc_sim_node(h, di, i, K, v.data(), pos.data(), &nb);
    if (tpos == UINT_MAX) continue;
    idxOut[cell] = tpos;  // importDataIndex = index*nTI + TI
    const float* tiv = p.cross ? c->tiX.data() : p.ti;
    const size_t src = (c->nTI > 1 ? c->cumT[tpos % (unsigned)c->nTI] : 0) + tpos / (unsigned)c->nTI;
    for (int q = 0; q < p.nVar; ++q)  // the variables the cell does not hold (simulation.hpp:658-666)
      if (std::isnan(di[(size_t)cell * p.nVar + q])) di[(size_t)cell * p.nVar + q] = tiv[src * p.nVar + q];
  }
  c->index = keep;
}

}  // extern "C"
