// HIP kernels of the MI355X QuickSampling exact scan (gfx950).
//
// Per path wave (a batch of independent path nodes) the scan is
//   [D>=2] k_pencil      : for each block of `nvec` adjacent spectrum columns
//                          and each node: sparse neighbour DFT along the
//                          leading dims (direct, <= n terms), forward FFT along
//                          the last dim, complex MAC with the TI spectra of
//                          every active sub-variable, inverse FFT along the
//                          last dim; writes only the rows whose centre
//                          coordinate is valid.
//   [D==3] k_colfft      : inverse along dim1 of the valid planes.
//   [D==1] k_direct1d    : the spectrum of a 1-D node directly.
//   k_rows_select        : two rows per complex FFT (two-for-one c2r along
//                          dim0), scale + delta0, wrap-around mask as a
//                          predicate, block top-K of the device rule
//                          (optionally the full position-space map).
//   k_merge_draw         : per node top-K of the block candidates, the rank
//                          draw floor(seed*k), outputs.
// The mismatch map of a node is never written to HBM in the wave path.
//
// Replaces the reference's per-node cuFFT R2C / complexAddAlphaxCxD /
// cuFFT C2R / removeBorder / fma / cublasIsamin sequence
// (src/NvidiaGPUAcceleratorDevice.cu:466-614, 688-717) and the CPU scan
// (src/CPUThreadDevice.cpp:349-490, src/CPUThreadAcceleratorDevice.cpp:251-357).
#include <hip/hip_runtime.h>
#include <math.h>
#include <algorithm>
#include <cmath>
#include <stdint.h>

#include "qsmi_kernels.hpp"
#include "qsmi_select.hpp"

namespace qsmi {

// ---------------------------------------------------------------------------
// TI spectra, once per run: rows r2c (two real rows per complex FFT)
// in: real images [nrows][f0]; out: half rows [nrows][h0]
// ---------------------------------------------------------------------------
__global__ void k_r2c_rows(const float* __restrict__ in, float2* __restrict__ out, FftPlan plan,
                           const float2* __restrict__ twg, int lognv, long long nrows, int h0) {
  extern __shared__ float2 smem[];
  const int L = plan.L, nvec = 1 << lognv, tid = threadIdx.x, nthr = blockDim.x;
  float2* tw = smem;
  float2* a = tw + L;
  float2* b = a + L * nvec;
  for (int i = tid; i < L; i += nthr) tw[i] = twg[i];
  const long long row0 = (long long)blockIdx.x * 2 * nvec;
  for (int e = tid; e < L * nvec; e += nthr) {
    int i = e >> lognv, v = e & (nvec - 1);
    long long ra = row0 + 2 * v, rb = ra + 1;
    float xa = (ra < nrows) ? in[ra * L + i] : 0.f;
    float xb = (rb < nrows) ? in[rb * L + i] : 0.f;
    a[e] = make_float2(xa, xb);
  }
  __syncthreads();
  float2* r = stockham<-1>(a, b, tw, plan, lognv, tid, nthr);
  for (int e = tid; e < h0 * nvec; e += nthr) {
    int k = e >> lognv, v = e & (nvec - 1);
    long long ra = row0 + 2 * v, rb = ra + 1;
    float2 z = r[e];
    float2 zc = r[(((L - k) % L) << lognv) + v];
    zc.y = -zc.y;  // conj(Z[-k])
    // X_a = (Z + conj Z-)/2 ; X_b = -i (Z - conj Z-)/2
    float2 xa = make_float2(0.5f * (z.x + zc.x), 0.5f * (z.y + zc.y));
    float2 d = make_float2(0.5f * (z.x - zc.x), 0.5f * (z.y - zc.y));
    float2 xb = make_float2(d.y, -d.x);
    if (ra < nrows) out[ra * h0 + k] = xa;
    if (rb < nrows) out[rb * h0 + k] = xb;
  }
}

// ---------------------------------------------------------------------------
// Column FFT in place along a strided dimension.
// Plane p (blockIdx.x) starts at base + (p / pdiv) * pOuter + (p % pdiv) * pInner;
// element (row r, column c) of the plane at r * rowStride + c, c < ncols.
// When info != nullptr the plane p belongs to node p / planesPerNode with
// coordinate c = p % planesPerNode along dimension validDim; planes outside
// [lo, hi] of that node are skipped.
// ---------------------------------------------------------------------------
template <int DIR>
__global__ void k_colfft(float2* __restrict__ base, FftPlan plan, const float2* __restrict__ twg, int lognv,
                         long long rowStride, int ncols, int pdiv, long long pOuter, long long pInner,
                         const NodeInfo* __restrict__ info, int planesPerNode, int validDim) {
  extern __shared__ float2 smem[];
  const int L = plan.L, nvec = 1 << lognv, tid = threadIdx.x, nthr = blockDim.x;
  const int p = blockIdx.x;
  if (info) {
    const NodeInfo& ni = info[p / planesPerNode];
    int c = p % planesPerNode;
    if (c < ni.lo[validDim] || c > ni.hi[validDim]) return;
  }
  float2* tw = smem;
  float2* a = tw + L;
  float2* b = a + L * nvec;
  for (int i = tid; i < L; i += nthr) tw[i] = twg[i];
  float2* pl = base + (long long)(p / pdiv) * pOuter + (long long)(p % pdiv) * pInner;
  const int c0 = blockIdx.y * nvec;
  for (int e = tid; e < L * nvec; e += nthr) {
    int r = e >> lognv, v = e & (nvec - 1);
    a[e] = (c0 + v < ncols) ? pl[r * rowStride + c0 + v] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  float2* res = stockham<DIR>(a, b, tw, plan, lognv, tid, nthr);
  for (int e = tid; e < L * nvec; e += nthr) {
    int r = e >> lognv, v = e & (nvec - 1);
    if (c0 + v < ncols) pl[r * rowStride + c0 + v] = res[e];
  }
}

// ---------------------------------------------------------------------------
// Fused pencil: sparse neighbour DFT + forward FFT + MAC + inverse FFT along
// the last dimension P. grid.x = column blocks (nblk0 per dim1 row for D=3),
// grid.y = node groups of `npb` nodes.
// Yhat: [A][kP][col] with col = k0 + h0*k1 (k1 only for D=3).
// C (per node, stride H): [cP][col], rows cP in [lo_P, hi_P] written.
// ---------------------------------------------------------------------------
constexpr int kPencilThreads = 256;
constexpr int kMaxEPT = kMaxVecElems / kPencilThreads;  // 16
// STAGE: the dim-0/1 twiddle tables once per block and each node's
// neighbours (nb <= kStageNb, V <= kStageV) in LDS, so the column build reads
// only LDS; the sums are the same operations in the same order either way
constexpr int kStageNb = 128, kStageV = 8;

// x mod f in [0, f): a float estimate of the quotient, corrected once
// (|x| < 2^23, so x and the product are exact to within one unit of q)
__device__ __forceinline__ int mod_small(int x, int f, float invf) {
  if (x <= -(1 << 23) || x >= (1 << 23)) return ((x % f) + f) % f;
  const int q = (int)floorf((float)x * invf);
  int r = x - q * f;
  if (r < 0) r += f;
  if (r >= f) r -= f;
  return r;
}

template <bool STAGE>
__global__ void __launch_bounds__(kPencilThreads) k_pencil(ScanGeom g, FftPlan planP, int lognv, int nblk0,
                                                           const float2* __restrict__ Yhat,
                                                           const float2* __restrict__ tw0g,
                                                           const float2* __restrict__ tw1g,
                                                           const float2* __restrict__ twPg, int nNodes, int npb,
                                                           const NodeInfo* __restrict__ info,
                                                           const int* __restrict__ nbOff,
                                                           const float* __restrict__ nbVal,
                                                           const uint32_t* __restrict__ ord,
                                                           float2* __restrict__ C) {
  extern __shared__ float2 smem[];
  const int LP = planP.L, nvec = 1 << lognv, tid = threadIdx.x, nthr = blockDim.x;
  const int D = g.D, P = g.P;
  float2* twP = smem;
  float2* bufA = twP + LP;
  float2* bufB = bufA + LP * nvec;
  float2* tw0s = bufB + LP * nvec;  // STAGE: f[0] entries, then f[1] (D == 3)
  float2* tw1s = tw0s + g.f[0];
  int* sOff = reinterpret_cast<int*>(tw1s + (D == 3 ? g.f[1] : 0));  // [kStageNb][3]
  float* sVal = reinterpret_cast<float*>(sOff + kStageNb * 3);      // [kStageNb][kStageV]
  uint32_t* sOrd = reinterpret_cast<uint32_t*>(sVal + kStageNb * kStageV);
  float2* sP1 = reinterpret_cast<float2*>(sOrd + kStageNb);  // D == 3: exp(+2 pi i k1 d1 / f1) per neighbour
  for (int i = tid; i < LP; i += nthr) twP[i] = twPg[i];
  if (STAGE) {
    for (int i = tid; i < g.f[0]; i += nthr) tw0s[i] = tw0g[i];
    if (D == 3)
      for (int i = tid; i < g.f[1]; i += nthr) tw1s[i] = tw1g[i];
  }
  const float2* __restrict__ tw0 = STAGE ? tw0s : tw0g;
  const float2* __restrict__ tw1 = STAGE ? tw1s : tw1g;

  const int blk0 = blockIdx.x % nblk0;
  const int k1 = (D == 3) ? (int)(blockIdx.x / nblk0) : 0;
  const int k0base = blk0 * nvec;
  const int ncolsP = g.h0 * ((D == 3) ? g.f[1] : 1);
  const int colbase = k0base + g.h0 * k1;
  const int elems = LP * nvec;
  const int fP = g.f[P];
  const int f0 = g.f[0], f1 = g.f[1];
  const float invf0 = 1.f / (float)f0, invf1 = 1.f / (float)f1;

  const int nodeBeg = blockIdx.y * npb;
  const int nodeEnd = min(nNodes, nodeBeg + npb);
  for (int node = nodeBeg; node < nodeEnd; ++node) {
    const NodeInfo ni = info[node];
    if (ni.lo[P] > ni.hi[P]) continue;  // no valid centre
    // block-uniform; the staged lists are read after the first barrier below
    // and not overwritten before the node's last barrier
    const bool staged = STAGE && ni.nb <= kStageNb;
    if (staged)
      for (int j = tid; j < ni.nb; j += nthr) {
        const long long jj = ni.nbStart + j;
        sOrd[j] = ord[jj];
        for (int d = 0; d < D; ++d) sOff[j * 3 + d] = nbOff[jj * D + d];
        for (int a = 0; a < g.V; ++a) sVal[j * kStageV + a] = nbVal[jj * g.V + a];
        if (D == 3) {  // the block's k1 phase of neighbour j, shared by every column and spectrum
          int t1 = (int)(((long long)k1 * nbOff[jj * D + 1]) % f1);
          if (t1 < 0) t1 += f1;
          float2 p1 = tw1g[t1];
          p1.y = -p1.y;
          sP1[j] = p1;
        }
      }
    float2 M[kMaxEPT];
#pragma unroll
    for (int i = 0; i < kMaxEPT; ++i) M[i] = make_float2(0.f, 0.f);

    for (int a = 0; a < g.V; ++a) {
      for (int e = tid; e < elems; e += nthr) bufA[e] = make_float2(0.f, 0.f);
      __syncthreads();
      // the sparse neighbour kernel column(s): the neighbours sorted by
      // their row along the pencil dimension (k_nb_sort); the head of each
      // row's run sums the run in neighbour-list order (fixed order)
      const uint32_t* op = staged ? sOrd : ord + ni.nbStart;
      for (int t = tid; t < ni.nb * nvec; t += nthr) {
        const int i = t >> lognv, v = t & (nvec - 1);
        const int len = (int)(op[i] >> 16);
        if (len == 0) continue;
        const int k0 = k0base + v;
        if (k0 >= g.h0) continue;
        float2 acc = make_float2(0.f, 0.f);
        int dP = 0;
        for (int r = 0; r < len; ++r) {
          const int j = (int)(op[i + r] & 0xFFFFu);
          float val;
          int d0, d1 = 0;
          if (staged) {
            val = sVal[j * kStageV + a];
            d0 = sOff[j * 3 + 0];
            if (D == 3) d1 = sOff[j * 3 + 1];
            dP = sOff[j * 3 + P];
          } else {
            const int jj = ni.nbStart + j;
            val = nbVal[(long long)jj * g.V + a];
            d0 = nbOff[jj * D + 0];
            if (D == 3) d1 = nbOff[jj * D + 1];
            dP = nbOff[jj * D + P];
          }
          float2 ph = tw0[mod_small(k0 * d0, f0, invf0)];
          ph.y = -ph.y;  // exp(+2 pi i k0 d0 / f0)
          if (D == 3) {
            float2 p1;
            if (staged) {
              p1 = sP1[j];
            } else {
              p1 = tw1[mod_small(k1 * d1, f1, invf1)];
              p1.y = -p1.y;
            }
            ph = cmul(ph, p1);
          }
          acc.x += val * ph.x;
          acc.y += val * ph.y;
        }
        int q = (-dP) % fP;
        if (q < 0) q += fP;
        bufA[(q << lognv) + v] = acc;
      }
      __syncthreads();
      float2* X = stockham<-1>(bufA, bufB, twP, planP, lognv, tid, nthr);
      const float2* Ya = Yhat + (long long)g.A[a] * g.H;
#pragma unroll
      for (int i = 0; i < kMaxEPT; ++i) {
        const int e = tid + i * nthr;
        if (e < elems) {
          const int kP = e >> lognv, v = e & (nvec - 1);
          if (k0base + v < g.h0) {
            const float2 y = Ya[(long long)kP * ncolsP + colbase + v];
            const float2 x = X[e];
            M[i].x += y.x * x.x - y.y * x.y;
            M[i].y += y.x * x.y + y.y * x.x;
          }
        }
      }
      __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < kMaxEPT; ++i) {
      const int e = tid + i * nthr;
      if (e < elems) bufA[e] = M[i];
    }
    __syncthreads();
    float2* R = stockham<+1>(bufA, bufB, twP, planP, lognv, tid, nthr);
    float2* Cn = C + (long long)node * g.H;
    const int lo = ni.lo[P], hi = ni.hi[P];
    for (int e = tid; e < elems; e += nthr) {
      const int cP = e >> lognv, v = e & (nvec - 1);
      if (cP >= lo && cP <= hi && k0base + v < g.h0) Cn[(long long)cP * ncolsP + colbase + v] = R[e];
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Neighbours of each node sorted by their row along the pencil dimension P
// (key (-d_P) mod f_P, ties in list order), for the fixed-order column build
// of k_pencil: ord[nbStart + i] = j | (run length << 16) at the head of each
// run of equal keys (0 elsewhere). One 64-thread block per node.
// ---------------------------------------------------------------------------
constexpr int kSortLds = 2048;
__global__ void __launch_bounds__(64) k_nb_sort(const NodeInfo* __restrict__ info, const int* __restrict__ nbOff,
                                               int D, int fP, uint32_t* __restrict__ ord) {
  __shared__ int key[kSortLds];
  __shared__ int jAt[kSortLds];
  const NodeInfo ni = info[blockIdx.x];
  const int nb = ni.nb, lane = threadIdx.x;
  if (nb <= 0) return;
  auto keyOf = [&](int j) {
    int q = (-nbOff[(long long)(ni.nbStart + j) * D + (D - 1)]) % fP;
    return q < 0 ? q + fP : q;
  };
  const bool lds = nb <= kSortLds;
  if (lds)
    for (int j = lane; j < nb; j += 64) key[j] = keyOf(j);
  __syncthreads();
  uint32_t* out = ord + ni.nbStart;
  for (int j = lane; j < nb; j += 64) {
    const int kj = lds ? key[j] : keyOf(j);
    int pos = 0;
    for (int q = 0; q < nb; ++q) {
      const int kq = lds ? key[q] : keyOf(q);
      pos += (kq < kj || (kq == kj && q < j)) ? 1 : 0;
    }
    out[pos] = (uint32_t)j;
    if (lds) jAt[pos] = j;
  }
  __syncthreads();
  // run lengths at the heads
  for (int i = lane; i < nb; i += 64) {
    const int ji = lds ? jAt[i] : (int)(out[i] & 0xFFFFu);
    const int ki = lds ? key[ji] : keyOf(ji);
    bool head = true;
    if (i > 0) {
      const int jp = lds ? jAt[i - 1] : (int)(out[i - 1] & 0xFFFFu);
      head = (lds ? key[jp] : keyOf(jp)) != ki;
    }
    if (!head) continue;
    int len = 1;
    while (i + len < nb) {
      const int jn = lds ? jAt[i + len] : (int)(out[i + len] & 0xFFFFu);
      if ((lds ? key[jn] : keyOf(jn)) != ki) break;
      ++len;
    }
    out[i] = (uint32_t)ji | ((uint32_t)len << 16);
  }
}

// ---------------------------------------------------------------------------
// 1-D nodes: the spectrum directly, M(k0) = sum_A Yhat_A(k0) sum_j v_jA e^{+2pi i k0 d_j/f0}
// ---------------------------------------------------------------------------
__global__ void k_direct1d(ScanGeom g, const float2* __restrict__ Yhat, const float2* __restrict__ tw0g,
                           const NodeInfo* __restrict__ info, const int* __restrict__ nbOff,
                           const float* __restrict__ nbVal, float2* __restrict__ C) {
  const int node = blockIdx.y;
  const int k0 = blockIdx.x * blockDim.x + threadIdx.x;
  if (k0 >= g.h0) return;
  const NodeInfo ni = info[node];
  const int f0 = g.f[0];
  float2 acc = make_float2(0.f, 0.f);
  for (int a = 0; a < g.V; ++a) {
    float2 x = make_float2(0.f, 0.f);
    for (int j = 0; j < ni.nb; ++j) {
      const int jj = ni.nbStart + j;
      const float val = nbVal[(long long)jj * g.V + a];
      if (val == 0.f) continue;
      int t0 = (int)(((long long)k0 * nbOff[jj]) % f0);
      if (t0 < 0) t0 += f0;
      float2 ph = tw0g[t0];
      x.x += val * ph.x;
      x.y -= val * ph.y;
    }
    const float2 y = Yhat[(long long)g.A[a] * g.H + k0];
    acc.x += y.x * x.x - y.y * x.y;
    acc.y += y.x * x.y + y.y * x.x;
  }
  C[(long long)node * g.H + k0] = acc;
}

// ---------------------------------------------------------------------------
// Rows of one rows block (valid rows q0 .. q0 + 2 nvec - 1 of the node, or
// nvec rows in the cross-measurement form): Hermitian completion and one
// two-for-one c2r along dim0 per row pair (Z = A + i B). smem: twiddles
// (already loaded, first read after a barrier), then two L x nvec buffers.
// Returns the LDS buffer holding the transform ([c0][v]). Every thread of
// the block calls it (barriers inside).
// ---------------------------------------------------------------------------
__device__ float2* rows_block_c2r(const ScanGeom& g, const FftPlan& plan0, int lognv, const NodeInfo& ni,
                                  const float2* __restrict__ Cn, const float2* __restrict__ Cx, bool cross,
                                  long long q0, long long nrv, int n1v, float2* smem) {
  const int L = plan0.L, nvec = 1 << lognv, tid = threadIdx.x, nthr = blockDim.x;
  float2* tw = smem;
  float2* a = tw + L;
  float2* b = a + L * nvec;
  const int h0 = g.h0, f1 = g.f[1];
  for (int e = tid; e < L * nvec; e += nthr) {
    const int k = e >> lognv, v = e & (nvec - 1);
    const long long qa = q0 + (cross ? v : 2 * v), qb = cross ? qa : qa + 1;
    const bool conjk = (k >= h0);
    const int kk = conjk ? (L - k) : k;
    float2 A = make_float2(0.f, 0.f), B = make_float2(0.f, 0.f);
    if (qa < nrv) {
      const int c1 = ni.lo[1] + (int)(qa % n1v);
      const int c2 = ni.lo[2] + (int)(qa / n1v);
      A = Cn[((long long)c2 * f1 + c1) * h0 + kk];
    }
    if (qb < nrv) {
      const int c1 = ni.lo[1] + (int)(qb % n1v);
      const int c2 = ni.lo[2] + (int)(qb / n1v);
      B = Cx[((long long)c2 * f1 + c1) * h0 + kk];
    }
    // a Hermitian c2r ignores the imaginary part of the DC and Nyquist bins
    if (kk == 0 || 2 * kk == L) {
      A.y = 0.f;
      B.y = 0.f;
    }
    if (conjk) {
      A.y = -A.y;
      B.y = -B.y;
    }
    a[e] = make_float2(A.x - B.y, A.y + B.x);
  }
  __syncthreads();
  return stockham<+1>(a, b, tw, plan0, lognv, tid, nthr);
}

// ---------------------------------------------------------------------------
// Rows: two-for-one c2r along dim0, epilogue, mask, block top-K.
// grid.x = row-pair blocks (2*nvec valid rows each), grid.y = node.
// Valid rows q enumerate (c1, c2) in the node's valid box, c1 fastest;
// lo/hi of unused dimensions are 0.
// cand (optional): [node][blockIdx.x][K]; map (optional): [node][S] position
// space (only valid positions written).
// ---------------------------------------------------------------------------
template <int KT>
__global__ void __launch_bounds__(256) k_rows_select(ScanGeom g, FftPlan plan0, int lognv,
                                                     const float2* __restrict__ tw0g,
                                                     const NodeInfo* __restrict__ info,
                                                     const float2* __restrict__ C, int K, Cand* __restrict__ cand,
                                                     int candBlocks, float* __restrict__ map, int crossOff,
                                                     const uint8_t* __restrict__ centreOk, int* __restrict__ gcnt) {
  extern __shared__ float2 smem[];
  __shared__ Cand red[16];
  __shared__ int cred[16];
  const int L = plan0.L, nvec = 1 << lognv, tid = threadIdx.x, nthr = blockDim.x;
  const int node = blockIdx.y;
  const NodeInfo ni = info[node];
  const int n1v = ni.hi[1] - ni.lo[1] + 1;
  const int n2v = ni.hi[2] - ni.lo[2] + 1;
  const int n0v = ni.hi[0] - ni.lo[0] + 1;
  const long long nrv = (n1v > 0 && n2v > 0 && n0v > 0) ? (long long)n1v * n2v : 0;
  // cross-measurement form (crossOff > 0): one row per complex FFT, the
  // mismatch layer (node) in the real part, the count layer (node +
  // crossOff) in the imaginary part; else two rows of the one layer
  const bool cross = crossOff > 0;
  const long long q0 = (long long)blockIdx.x * (cross ? 1 : 2) * nvec;
  if (q0 >= nrv) return;
  float2* tw = smem;
  for (int i = tid; i < L; i += nthr) tw[i] = tw0g[i];
  const float2* Cn = C + (long long)node * g.H;
  const float2* Cx = cross ? C + (long long)(node + crossOff) * g.H : Cn;
  float2* z = rows_block_c2r(g, plan0, lognv, ni, Cn, Cx, cross, q0, nrv, n1v, smem);
  TopK<KT> tk;
  tk.init();
  const float scale = ni.scale, d0 = ni.delta0;
  const int s0 = g.s[0], s1 = g.s[1];
  for (int e = tid; e < L * nvec; e += nthr) {
    const int c0 = e >> lognv, v = e & (nvec - 1);
    if (c0 < ni.lo[0] || c0 > ni.hi[0]) continue;
    const float2 zz = z[e];
    if (cross) {
      // CPUThreadDevice::maskLayerWithVariable + compensateMissingData:
      // count times the centre's completeness, map = -|E / C^4|, no
      // candidate where the count is zero
      const long long q = q0 + v;
      if (q >= nrv) continue;
      const int c1 = ni.lo[1] + (int)(q % n1v);
      const int c2 = ni.lo[2] + (int)(q / n1v);
      const uint32_t pos = (uint32_t)(c0 + s0 * (c1 + s1 * c2));
      // the count is a sum of neighbour weights: values below cmin (half
      // the smallest weight) are FFT noise around an exact zero
      const float E = zz.x * scale;
      const float Cc = (centreOk[pos] && zz.y * g.invN >= ni.cmin) ? zz.y * scale : 0.f;
      float val = (Cc == 0.f) ? -INFINITY : -fabsf(E / (Cc * Cc * Cc * Cc));
      if (ni.exOn && excluded(ni, s0, s1, c0, c1, c2)) val = -INFINITY;
      if (map) map[(long long)node * ((long long)s0 * s1 * g.s[2]) + pos] = val;
      if (cand && val > -INFINITY) tk.push_value(val, 0.f, pos, ni.seedbits);
      continue;
    }
#pragma unroll
    for (int w = 0; w < 2; ++w) {
      const long long q = q0 + 2 * v + w;
      if (q >= nrv) continue;
      const int c1 = ni.lo[1] + (int)(q % n1v);
      const int c2 = ni.lo[2] + (int)(q / n1v);
      float val = (w == 0 ? zz.x : zz.y) * scale + d0;
      const uint32_t pos = (uint32_t)(c0 + s0 * (c1 + s1 * c2));
      const bool ex = ni.exOn && excluded(ni, s0, s1, c0, c1, c2);
      if (ex) val = -INFINITY;
      if (map) map[(long long)node * ((long long)s0 * s1 * g.s[2]) + pos] = val;
      if (cand && !ex) tk.push_value(val, ni.quantum, pos, ni.seedbits);
    }
  }
  if (cand) {
    const Cand kth = block_topk<KT>(tk, K, cand + ((long long)node * candBlocks + blockIdx.x) * K, red);
    if (gcnt && !cross && ni.eps > 0.f && ni.quantum == 0.f) {
      // exact selection: how many of the block's values lie within 2 eps of
      // its K-th best (more than K: the band may reach beyond its candidates)
      const float th = (kth.pos == 0xFFFFFFFFu) ? -INFINITY : kth.val - 2.f * ni.eps * 1.0009765625f;
      int c = 0;
      for (int e = tid; e < L * nvec; e += nthr) {
        const int c0 = e >> lognv, v = e & (nvec - 1);
        if (c0 < ni.lo[0] || c0 > ni.hi[0]) continue;
        const float2 zz = z[e];
#pragma unroll
        for (int w = 0; w < 2; ++w) {
          const long long q = q0 + 2 * v + w;
          if (q >= nrv) continue;
          const int c1 = ni.lo[1] + (int)(q % n1v);
          const int c2 = ni.lo[2] + (int)(q / n1v);
          const float val = (w == 0 ? zz.x : zz.y) * scale + d0;
          if (val >= th && !(ni.exOn && excluded(ni, s0, s1, c0, c1, c2))) ++c;
        }
      }
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off);
      if ((tid & 63) == 0) cred[tid >> 6] = c;
      __syncthreads();
      if (tid == 0) {
        int tot = 0;
        for (int i = 0; i < (nthr >> 6); ++i) tot += cred[i];
        gcnt[(long long)node * candBlocks + blockIdx.x] = tot;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Exact selection, second pass on the generic path: for every queued
// (node, rows block) the block's rows again (same transform and epilogue as
// k_rows_select, never the cross-measurement form), appending every valid
// position whose FP32 value is >= theta[node] to the node's band list.
// Persistent grid over the device queue count.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_rows_collect(ScanGeom g, FftPlan plan0, int lognv,
                                                      const float2* __restrict__ tw0g,
                                                      const NodeInfo* __restrict__ info,
                                                      const float2* __restrict__ C, const int2* __restrict__ queue,
                                                      const int* __restrict__ qcnt, const float* __restrict__ theta,
                                                      uint32_t* __restrict__ bandPos, int* __restrict__ bandCnt) {
  extern __shared__ float2 smem[];
  const int L = plan0.L, nvec = 1 << lognv, tid = threadIdx.x, nthr = blockDim.x, lane = tid & 63;
  const int items = *qcnt;
  if ((int)blockIdx.x >= items) return;
  for (int i = tid; i < L; i += nthr) smem[i] = tw0g[i];  // first read after a barrier
  const int s0 = g.s[0], s1 = g.s[1];
  const int per = (L * nvec + nthr - 1) / nthr;
  for (int it = blockIdx.x; it < items; it += gridDim.x) {
    const int2 qi = queue[it];
    const int node = qi.x;
    const NodeInfo ni = info[node];
    const int n1v = ni.hi[1] - ni.lo[1] + 1;
    const int n2v = ni.hi[2] - ni.lo[2] + 1;
    const int n0v = ni.hi[0] - ni.lo[0] + 1;
    const long long nrv = (n1v > 0 && n2v > 0 && n0v > 0) ? (long long)n1v * n2v : 0;
    const long long q0 = (long long)qi.y * 2 * nvec;
    const float2* Cn = C + (long long)node * g.H;
    float2* z = rows_block_c2r(g, plan0, lognv, ni, Cn, Cn, false, q0, nrv, n1v, smem);
    const float th = theta[node];
    const float scale = ni.scale, d0 = ni.delta0;
    uint32_t* out = bandPos + (long long)node * kBandCap;
    // every lane runs the same number of rounds (ballots below)
    for (int r = 0; r < per; ++r) {
      const int e = tid + r * nthr;
#pragma unroll
      for (int w = 0; w < 2; ++w) {
        bool ok = false;
        uint32_t pos = 0;
        if (e < L * nvec) {
          const int c0 = e >> lognv, v = e & (nvec - 1);
          const long long q = q0 + 2 * v + w;
          if (c0 >= ni.lo[0] && c0 <= ni.hi[0] && q < nrv) {
            const int c1 = ni.lo[1] + (int)(q % n1v);
            const int c2 = ni.lo[2] + (int)(q / n1v);
            const float2 zz = z[e];
            const float val = (w == 0 ? zz.x : zz.y) * scale + d0;
            pos = (uint32_t)(c0 + s0 * (c1 + s1 * c2));
            ok = val >= th && !(ni.exOn && excluded(ni, s0, s1, c0, c1, c2));
          }
        }
        const unsigned long long m = __ballot(ok);
        if (m == 0ull) continue;
        const int leader = __ffsll((long long)m) - 1;
        int base = 0;
        if (lane == leader) base = atomicAdd(&bandCnt[node], __popcll(m));
        base = __shfl(base, leader);
        if (ok) {
          const int idx = base + __popcll(m & ((1ull << lane) - 1ull));
          if (idx < kBandCap) out[idx] = pos;
        }
      }
    }
    __syncthreads();  // the next item's transform reuses the LDS buffers
  }
}

// ---------------------------------------------------------------------------
// Top-K over a position-space map (per-node API after map edits).
// grid.x = chunks, one node; cand [blockIdx.x][K].
// ---------------------------------------------------------------------------
template <int KT>
__global__ void __launch_bounds__(256) k_select_map(const float* __restrict__ map, long long S, int K,
                                                    float quantum, uint32_t seedbits, Cand* __restrict__ cand) {
  __shared__ Cand red[16];
  const long long chunk = (S + gridDim.x - 1) / gridDim.x;
  const long long beg = (long long)blockIdx.x * chunk;
  const long long end = min(S, beg + chunk);
  TopK<KT> tk;
  tk.init();
  for (long long i = beg + threadIdx.x; i < end; i += blockDim.x) {
    const float v = map[i];
    if (isfinite(v)) tk.push_value(v, quantum, (uint32_t)i, seedbits);
  }
  block_topk<KT>(tk, K, cand + (long long)blockIdx.x * K, red);
}

// ---------------------------------------------------------------------------
// Per node: K best of the block candidates, rank draw.
// nblkOf: candidate blocks per node are ceil(validRows / rowsPerBlock)
// (rowsPerBlock <= 0: use candBlocks for every node).
// ---------------------------------------------------------------------------
template <int KT>
__global__ void __launch_bounds__(256) k_merge_draw(ScanGeom g, const NodeInfo* __restrict__ info,
                                                    const Cand* __restrict__ cand, int candBlocks, int M,
                                                    int rowsPerBlock, int K, float kf, float* __restrict__ outVal,
                                                    uint32_t* __restrict__ outPos, uint32_t* __restrict__ chosen,
                                                    float* __restrict__ simVals, const float* __restrict__ tiVals,
                                                    int nVar, uint32_t* __restrict__ indexOut) {
  __shared__ Cand red[16];
  __shared__ Cand best[kMaxK];
  __shared__ uint32_t bpos[kMaxK];
  const int node = blockIdx.x;
  const NodeInfo ni = info[node];
  if (simVals && ni.nb == 0) return;  // drawn from the marginal by the gather stage
  int nblk = candBlocks;
  if (rowsPerBlock > 0) {
    const int n1v = ni.hi[1] - ni.lo[1] + 1;
    const int n2v = ni.hi[2] - ni.lo[2] + 1;
    const int n0v = ni.hi[0] - ni.lo[0] + 1;
    const long long nrv = (n1v > 0 && n2v > 0 && n0v > 0) ? (long long)n1v * n2v : 0;
    nblk = (int)((nrv + rowsPerBlock - 1) / rowsPerBlock);
  }
  TopK<KT> tk;
  tk.init();
  const Cand* cn = cand + (long long)node * candBlocks * M;
  for (int i = threadIdx.x; i < nblk * M; i += blockDim.x) tk.push(cn[i]);
  block_topk<KT>(tk, K, best, red);
  if (threadIdx.x < K) {
    outVal[(long long)node * K + threadIdx.x] = best[threadIdx.x].val;
    outPos[(long long)node * K + threadIdx.x] = best[threadIdx.x].pos;
    bpos[threadIdx.x] = best[threadIdx.x].pos;
  }
  __syncthreads();
  if (threadIdx.x == 0 && (chosen || simVals)) draw_write(ni, bpos, K, kf, chosen, node, simVals, tiVals, nVar, indexOut);
}

// Several TIs: per node, the TIs' K best lists (val/pos [t][node][K], TI
// positions) merged as QuantileSamplingModule::sample does
// (include/quantileSamplingModule.hpp:90-126: every TI's k best sorted
// together by value, rank floor(seed k ceil(nTI/k) / nTI)), the device rule
// ordering them with positions encoded as position * nTI + TI. The rank is
// clamped to the last valid candidate. Writes the value and importDataIndex.
// One thread per node.
__global__ void k_merge_ti(const NodeInfo* __restrict__ info, int nN, int nTI, int K, float kf,
                           const float* __restrict__ val, const uint32_t* __restrict__ pos, long long stride,
                           float quantum, float* __restrict__ simVals, const float* __restrict__ tiVals, int nVar,
                           TiCum cum, uint32_t* __restrict__ indexOut) {
  const int node = blockIdx.x * blockDim.x + threadIdx.x;
  if (node >= nN) return;
  const NodeInfo ni = info[node];
  if (ni.nb == 0) return;  // drawn from the marginal by the gather stage
  const float kk = ni.kLocal > 0.f ? ni.kLocal : kf;
  int r = (int)floorf(ni.seed * kk * (ceilf((float)nTI / kk) / (float)nTI));
  if (r < 0) r = 0;
  Cand prev = worst_cand(), pick = worst_cand();
  bool any = false;
  for (int q = 0; q <= r; ++q) {
    Cand best = worst_cand();
    bool found = false;
    for (int t = 0; t < nTI; ++t) {
      // a TI-index image restricts the node to one TI (include/
      // quantileSamplingModule.hpp:460-464: only that device is updated)
      if (ni.tiReq >= 0 && t != ni.tiReq) continue;
      for (int i = 0; i < K; ++i) {
        const long long o = t * stride + (long long)node * K + i;
        const uint32_t p = pos[o];
        const float v = val[o];
        if (p == 0xFFFFFFFFu || !(v > -INFINITY)) continue;
        Cand c;
        c.val = v;
        c.key = (quantum > 0.f) ? rintf(v * (1.f / quantum)) : v;
        c.pos = p * (uint32_t)nTI + (uint32_t)t;
        c.h = tie_hash(ni.seedbits, c.pos);
        if (q > 0 && !cand_better(prev, c)) continue;
        if (!found || cand_better(c, best)) {
          best = c;
          found = true;
        }
      }
    }
    if (!found) break;
    prev = best;
    pick = best;
    any = true;
  }
  if (!any) return;
  const uint32_t t = pick.pos % (uint32_t)nTI, p = pick.pos / (uint32_t)nTI;
  for (int v = 0; v < nVar; ++v) {  // the variables the cell does not hold yet
    float* dst = simVals + (long long)ni.cell * nVar + v;
    if (isnan(*dst)) *dst = tiVals[((long long)cum.c[t] + p) * nVar + v];
  }
  indexOut[ni.cell] = pick.pos;
}

// Several TIs of their own extents: a node on another TI. The lower bound
// -min of the valid box depends on the offsets only; the upper bound
// s - 1 - max moves with the extent (dh = s_t - s_0), and the map scale is
// that TI's 1 / prod(FFT extent). Marginal nodes (nb == 0) keep their empty
// box; circular TIs span the whole extent.
__global__ void k_info_shift(const NodeInfo* __restrict__ in, NodeInfo* __restrict__ out, int nN, int dh0, int dh1,
                             int dh2, float scale, float epsFactor) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nN) return;
  NodeInfo ni = in[q];
  if (ni.nb > 0) {
    ni.hi[0] += dh0;
    ni.hi[1] += dh1;
    ni.hi[2] += dh2;
  }
  ni.scale = scale;
  ni.eps *= epsFactor;
  out[q] = ni;
}

__global__ void k_fill(float* __restrict__ p, long long n, float v) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// maskLayerWithVariable on a position-space layer: every finite value times
// the mask image at its position (-inf outside the valid box stays -inf)
__global__ void k_mask_layer(float* __restrict__ layer, const float* __restrict__ mask, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float v = layer[i];
  if (v > -INFINITY) layer[i] = v * mask[i];
}

// compensateMissingData: map = -|E / (((C C) C) C)| (the reference's
// evaluation order), -inf where the count is zero: below cmin (half the
// node's smallest positive count-layer value, in map units) the FFT value is
// noise around an exact zero; cmin == 0 compares with zero as the reference does
__global__ void k_compensate(float* __restrict__ map, const float* __restrict__ cnt, long long n, float cmin) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float e = map[i], c = cnt[i];
  float r = -INFINITY;
  const bool zero = cmin > 0.f ? !(c >= cmin) : c == 0.f;
  if (e > -INFINITY && c > -INFINITY && !zero) r = -fabsf(e / (c * c * c * c));
  map[i] = r;
}

void launch_mask_layer(float* layer, const float* mask, long long n, hipStream_t st) {
  hipLaunchKernelGGL(k_mask_layer, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, layer, mask, n);
}

void launch_compensate(float* map, const float* cnt, long long n, float cmin, hipStream_t st) {
  hipLaunchKernelGGL(k_compensate, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, map, cnt, n, cmin);
}

// Every position of the map within the disk sum_d (p_d - o_d)^2 <= r2 set to
// v; the grid covers the disk's bounding box [bl, bl + bn) (clipped to the
// TI), one thread per position.
__global__ void k_fill_disk(float* __restrict__ map, int s0, int s1, uint32_t o, float r2, int bl0, int bl1, int bl2,
                            int bn0, int bn1, long long total, float v) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const int c0 = bl0 + (int)(i % bn0);
  const int c1 = bl1 + (int)((i / bn0) % bn1);
  const int c2 = bl2 + (int)(i / ((long long)bn0 * bn1));
  NodeInfo ni;
  ni.exPos = o;
  ni.exR2 = r2;
  if (excluded(ni, s0, s1, c0, c1, c2)) map[(long long)c0 + (long long)s0 * (c1 + (long long)s1 * c2)] = v;
}

// explicit instantiations used by the host
template __global__ void k_colfft<-1>(float2*, FftPlan, const float2*, int, long long, int, int, long long, long long,
                                      const NodeInfo*, int, int);
template __global__ void k_colfft<+1>(float2*, FftPlan, const float2*, int, long long, int, int, long long, long long,
                                      const NodeInfo*, int, int);
#define QSMI_INST_K(KT)                                                                                            \
  template __global__ void k_rows_select<KT>(ScanGeom, FftPlan, int, const float2*, const NodeInfo*,              \
                                             const float2*, int, Cand*, int, float*, int, const uint8_t*, int*);  \
  template __global__ void k_select_map<KT>(const float*, long long, int, float, uint32_t, Cand*);                \
  template __global__ void k_merge_draw<KT>(ScanGeom, const NodeInfo*, const Cand*, int, int, int, int, float,     \
                                            float*, uint32_t*, uint32_t*, float*, const float*, int, uint32_t*);
QSMI_INST_K(2)
QSMI_INST_K(4)
QSMI_INST_K(8)
QSMI_INST_K(16)

}  // namespace qsmi

// ---------------------------------------------------------------------------
// Host-side launchers (same translation unit as the kernels).
// ---------------------------------------------------------------------------
namespace qsmi {

static size_t fft_lds_bytes(int L, int nvec) { return (size_t)(L + 2 * L * nvec) * sizeof(float2); }
static size_t g_ldsLimit = 0;  // dynamic LDS per block the FFT kernels may use (set_lds_limits)

// Raises the dynamic LDS limit of the FFT kernels to what the device allows
// (160 KiB per workgroup on gfx950) minus a margin for static LDS.
hipError_t set_lds_limits(int device, size_t* limitOut) {
  int maxShared = 0;
  hipError_t e = hipDeviceGetAttribute(&maxShared, hipDeviceAttributeMaxSharedMemoryPerBlock, device);
  if (e != hipSuccess) return e;
  const int lim = maxShared - 1024;
  *limitOut = (size_t)lim;
  g_ldsLimit = (size_t)lim;
  const void* fns[] = {(const void*)k_pencil<false>, (const void*)k_pencil<true>, (const void*)k_r2c_rows,
                       (const void*)k_colfft<-1>,
                       (const void*)k_colfft<+1>,     (const void*)k_rows_select<2>,  (const void*)k_rows_select<4>,
                       (const void*)k_rows_select<8>, (const void*)k_rows_select<16>, (const void*)k_rows_collect};
  for (const void* f : fns) {
    e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

size_t fft_lds_bytes_of(int L, int nvec) { return fft_lds_bytes(L, nvec); }

void launch_r2c_rows(const float* in, float2* out, const FftPlan& plan, const float2* tw, int lognv, long long nrows,
                     int h0, hipStream_t st) {
  const int nvec = 1 << lognv;
  long long blocks = (nrows + 2 * nvec - 1) / (2 * nvec);
  hipLaunchKernelGGL(k_r2c_rows, dim3((unsigned)blocks), dim3(256), fft_lds_bytes(plan.L, nvec), st, in, out, plan, tw,
                     lognv, nrows, h0);
}

void launch_colfft(int dir, float2* base, const FftPlan& plan, const float2* tw, int lognv, long long rowStride,
                   int ncols, long long planes, int pdiv, long long pOuter, long long pInner, const NodeInfo* info,
                   int planesPerNode, int validDim, hipStream_t st) {
  const int nvec = 1 << lognv;
  dim3 grid((unsigned)planes, (unsigned)((ncols + nvec - 1) / nvec));
  if (dir < 0)
    hipLaunchKernelGGL(k_colfft<-1>, grid, dim3(256), fft_lds_bytes(plan.L, nvec), st, base, plan, tw, lognv,
                       rowStride, ncols, pdiv, pOuter, pInner, info, planesPerNode, validDim);
  else
    hipLaunchKernelGGL(k_colfft<+1>, grid, dim3(256), fft_lds_bytes(plan.L, nvec), st, base, plan, tw, lognv,
                       rowStride, ncols, pdiv, pOuter, pInner, info, planesPerNode, validDim);
}

void launch_pencil(const ScanGeom& g, const FftPlan& planP, int lognv, const float2* Yhat, const float2* tw0,
                   const float2* tw1, const float2* twP, int nNodes, int npb, const NodeInfo* info, const int* nbOff,
                   const float* nbVal, uint32_t* ord, float2* C, hipStream_t st) {
  const int nvec = 1 << lognv;
  const int nblk0 = (g.h0 + nvec - 1) / nvec;
  hipLaunchKernelGGL(k_nb_sort, dim3((unsigned)nNodes), dim3(64), 0, st, info, nbOff, g.D, g.f[g.D - 1], ord);
  dim3 grid((unsigned)(nblk0 * (g.D == 3 ? g.f[1] : 1)), (unsigned)((nNodes + npb - 1) / npb));
  const size_t base = fft_lds_bytes(planP.L, nvec);
  const size_t staged = base + (size_t)(g.f[0] + (g.D == 3 ? g.f[1] : 0)) * sizeof(float2) +
                        (size_t)kStageNb * (3 * sizeof(int) + kStageV * sizeof(float) + sizeof(uint32_t) +
                                            sizeof(float2));
  if (g.V <= kStageV && staged <= g_ldsLimit)
    hipLaunchKernelGGL(k_pencil<true>, grid, dim3(kPencilThreads), staged, st, g, planP, lognv, nblk0, Yhat, tw0, tw1,
                       twP, nNodes, npb, info, nbOff, nbVal, ord, C);
  else
    hipLaunchKernelGGL(k_pencil<false>, grid, dim3(kPencilThreads), base, st, g, planP, lognv, nblk0, Yhat, tw0, tw1,
                       twP, nNodes, npb, info, nbOff, nbVal, ord, C);
}

void launch_direct1d(const ScanGeom& g, const float2* Yhat, const float2* tw0, int nNodes, const NodeInfo* info,
                     const int* nbOff, const float* nbVal, float2* C, hipStream_t st) {
  dim3 grid((unsigned)((g.h0 + 255) / 256), (unsigned)nNodes);
  hipLaunchKernelGGL(k_direct1d, grid, dim3(256), 0, st, g, Yhat, tw0, info, nbOff, nbVal, C);
}

static int kt_for(int K) { return K <= 2 ? 2 : K <= 4 ? 4 : K <= 8 ? 8 : 16; }


void launch_rows_select(const ScanGeom& g, const FftPlan& plan0, int lognv, const float2* tw0, const NodeInfo* info,
                        const float2* C, int nNodes, int rowBlocks, int K, Cand* cand, float* map, hipStream_t st,
                        int crossOff, const uint8_t* centreOk, int* gcnt) {
  const int nvec = 1 << lognv;
  dim3 grid((unsigned)rowBlocks, (unsigned)nNodes);
  size_t lds = fft_lds_bytes(plan0.L, nvec);
#define QSMI_ROWS(KT)                                                                                        \
  hipLaunchKernelGGL(k_rows_select<KT>, grid, dim3(256), lds, st, g, plan0, lognv, tw0, info, C, K, cand, rowBlocks, \
                     map, crossOff, centreOk, gcnt)
  switch (kt_for(K)) {
    case 2: QSMI_ROWS(2); break;
    case 4: QSMI_ROWS(4); break;
    case 8: QSMI_ROWS(8); break;
    default: QSMI_ROWS(16); break;
  }
#undef QSMI_ROWS
}

void launch_rows_collect(const ScanGeom& g, const FftPlan& plan0, int lognv, const float2* tw0, const NodeInfo* info,
                         const float2* C, const int2* queue, const int* qcnt, const float* theta, uint32_t* bandPos,
                         int* bandCnt, int maxItems, hipStream_t st) {
  const int nvec = 1 << lognv;
  const int blocks = std::max(1, std::min(maxItems, 2048));
  hipLaunchKernelGGL(k_rows_collect, dim3((unsigned)blocks), dim3(256), fft_lds_bytes(plan0.L, nvec), st, g, plan0,
                     lognv, tw0, info, C, queue, qcnt, theta, bandPos, bandCnt);
}

void launch_select_map(const float* map, long long S, int K, float quantum, uint32_t seedbits, Cand* cand, int chunks,
                       hipStream_t st) {
  switch (kt_for(K)) {
    case 2: hipLaunchKernelGGL(k_select_map<2>, dim3(chunks), dim3(256), 0, st, map, S, K, quantum, seedbits, cand); break;
    case 4: hipLaunchKernelGGL(k_select_map<4>, dim3(chunks), dim3(256), 0, st, map, S, K, quantum, seedbits, cand); break;
    case 8: hipLaunchKernelGGL(k_select_map<8>, dim3(chunks), dim3(256), 0, st, map, S, K, quantum, seedbits, cand); break;
    default: hipLaunchKernelGGL(k_select_map<16>, dim3(chunks), dim3(256), 0, st, map, S, K, quantum, seedbits, cand); break;
  }
}

void launch_merge_draw(const ScanGeom& g, const NodeInfo* info, const Cand* cand, int nNodes, int candBlocks,
                       int M, int rowsPerBlock, int K, float kf, float* outVal, uint32_t* outPos, uint32_t* chosen,
                       float* simVals, const float* tiVals, int nVar, uint32_t* indexOut, hipStream_t st) {
#define QSMI_MERGE(KT)                                                                                               \
  hipLaunchKernelGGL(k_merge_draw<KT>, dim3(nNodes), dim3(256), 0, st, g, info, cand, candBlocks, M, rowsPerBlock, \
                     K, kf, outVal, outPos, chosen, simVals, tiVals, nVar, indexOut)
  switch (kt_for(K)) {
    case 2: QSMI_MERGE(2); break;
    case 4: QSMI_MERGE(4); break;
    case 8: QSMI_MERGE(8); break;
    default: QSMI_MERGE(16); break;
  }
#undef QSMI_MERGE
}

void launch_fill_disk(float* map, const int* s, uint32_t o, float r2, float v, hipStream_t st) {
  const int oc[3] = {(int)(o % (uint32_t)s[0]), (int)((o / (uint32_t)s[0]) % (uint32_t)s[1]),
                     (int)(o / ((uint32_t)s[0] * (uint32_t)s[1]))};
  const int r = (int)std::ceil(std::sqrt((double)r2));
  int bl[3], bn[3];
  long long total = 1;
  for (int d = 0; d < 3; ++d) {
    bl[d] = std::max(0, oc[d] - r);
    const int bh = std::min(s[d] - 1, oc[d] + r);
    bn[d] = bh - bl[d] + 1;
    if (bn[d] <= 0) return;
    total *= bn[d];
  }
  hipLaunchKernelGGL(k_fill_disk, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, map, s[0], s[1], o, r2, bl[0],
                     bl[1], bl[2], bn[0], bn[1], total, v);
}

void launch_merge_ti(const NodeInfo* info, int nN, int nTI, int K, float kf, const float* val, const uint32_t* pos,
                     long long stride, float quantum, float* simVals, const float* tiVals, int nVar, TiCum cum,
                     uint32_t* indexOut, hipStream_t st) {
  hipLaunchKernelGGL(k_merge_ti, dim3((unsigned)((nN + 63) / 64)), dim3(64), 0, st, info, nN, nTI, K, kf, val, pos,
                     stride, quantum, simVals, tiVals, nVar, cum, indexOut);
}

void launch_info_shift(const NodeInfo* in, NodeInfo* out, int nN, int dh0, int dh1, int dh2, float scale,
                       float epsFactor, hipStream_t st) {
  hipLaunchKernelGGL(k_info_shift, dim3((unsigned)((nN + 63) / 64)), dim3(64), 0, st, in, out, nN, dh0, dh1, dh2,
                     scale, epsFactor);
}

void launch_fill(float* p, long long n, float v, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n, v);
}

}  // namespace qsmi
