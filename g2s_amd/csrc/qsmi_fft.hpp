// LDS-staged mixed-radix Stockham FFT building blocks for gfx950.
//
// A block transforms `nvec` independent vectors of length L that sit
// interleaved in LDS (element i of vector v at buf[i*nvec + v], nvec a power
// of two), ping-ponging between two LDS buffers. Each stage has radix R in
// {2,3,4,5,8,10} (hand-written butterflies) or {7,11,13} (small DFT through
// the twiddle table). Twiddles come from a table tw[t] = exp(-2*pi*i*t/L)
// generated on the host in FP64 and staged in LDS. Stockham autosort keeps
// natural order in and out, so no bit-reversal pass exists.
//
// DIR = -1: forward (exp(-i...)); DIR = +1: backward (exp(+i...)),
// unnormalised.
#pragma once
#include <hip/hip_runtime.h>

namespace qsmi {

constexpr int kMaxStages = 16;

struct FftPlan {
  int L;
  int nst;
  int radix[kMaxStages];
};

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// multiply by s*i (s = +-1)
template <int S>
__device__ __forceinline__ float2 mul_si(float2 a) {
  return (S > 0) ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}
// twiddle w (forward table entry) applied in direction DIR
template <int DIR>
__device__ __forceinline__ float2 twid(float2 w) {
  return (DIR < 0) ? w : make_float2(w.x, -w.y);
}

template <int DIR>
__device__ __forceinline__ void dft2(float2* u) {
  float2 a = u[0], b = u[1];
  u[0] = cadd(a, b);
  u[1] = csub(a, b);
}

template <int DIR>
__device__ __forceinline__ void dft3(float2* u) {
  const float h = 0.86602540378443864676f;  // sqrt(3)/2
  float2 t1 = cadd(u[1], u[2]);
  float2 t2 = csub(u[1], u[2]);
  float2 a = make_float2(u[0].x - 0.5f * t1.x, u[0].y - 0.5f * t1.y);
  float2 b = mul_si<DIR>(cscale(t2, h));
  u[0] = cadd(u[0], t1);
  u[1] = cadd(a, b);
  u[2] = csub(a, b);
}

template <int DIR>
__device__ __forceinline__ void dft4(float2* u) {
  float2 s02 = cadd(u[0], u[2]), d02 = csub(u[0], u[2]);
  float2 s13 = cadd(u[1], u[3]), d13 = mul_si<DIR>(csub(u[1], u[3]));
  u[0] = cadd(s02, s13);
  u[2] = csub(s02, s13);
  u[1] = cadd(d02, d13);
  u[3] = csub(d02, d13);
}

template <int DIR>
__device__ __forceinline__ void dft5(float2* u) {
  const float c1 = 0.30901699437494742410f;   // cos(2pi/5)
  const float c2 = -0.80901699437494742410f;  // cos(4pi/5)
  const float s1 = 0.95105651629515357212f;   // sin(2pi/5)
  const float s2 = 0.58778525229247312917f;   // sin(4pi/5)
  float2 t1 = cadd(u[1], u[4]), t3 = csub(u[1], u[4]);
  float2 t2 = cadd(u[2], u[3]), t4 = csub(u[2], u[3]);
  float2 a1 = make_float2(u[0].x + c1 * t1.x + c2 * t2.x, u[0].y + c1 * t1.y + c2 * t2.y);
  float2 a2 = make_float2(u[0].x + c2 * t1.x + c1 * t2.x, u[0].y + c2 * t1.y + c1 * t2.y);
  float2 b1 = mul_si<DIR>(make_float2(s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y));
  float2 b2 = mul_si<DIR>(make_float2(s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y));
  u[0] = cadd(u[0], cadd(t1, t2));
  u[1] = cadd(a1, b1);
  u[4] = csub(a1, b1);
  u[2] = cadd(a2, b2);
  u[3] = csub(a2, b2);
}

template <int DIR>
__device__ __forceinline__ void dft8(float2* u) {
  const float r = 0.70710678118654752440f;
  float2 e[4] = {u[0], u[2], u[4], u[6]};
  float2 o[4] = {u[1], u[3], u[5], u[7]};
  dft4<DIR>(e);
  dft4<DIR>(o);
  // o[k] *= W8^k with W8 = exp(DIR*2*pi*i/8)
  float2 o1 = o[1];
  o[1] = make_float2(r * (o1.x - (float)DIR * o1.y), r * (o1.y + (float)DIR * o1.x));
  o[2] = mul_si<DIR>(o[2]);
  float2 o3 = o[3];
  o[3] = make_float2(r * (-o3.x - (float)DIR * o3.y), r * (-o3.y + (float)DIR * o3.x));
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    u[k] = cadd(e[k], o[k]);
    u[k + 4] = csub(e[k], o[k]);
  }
}

// Prime-factor (Good-Thomas) 10 = 2 x 5: 2 and 5 are coprime, so with the
// input map n = (5 n1 + 2 n2) mod 10 and the output map k = (5 k1 + 6 k2)
// mod 10 the transform is two length-5 DFTs and five 2-point butterflies with
// no twiddle factors in between: W10^(nk) = W2^(n1 k1) W5^(n2 k2).
template <int DIR>
__device__ __forceinline__ void dft10(float2* u) {
  float2 a[5] = {u[0], u[2], u[4], u[6], u[8]};  // n1 = 0: n = 2 n2
  float2 b[5] = {u[5], u[7], u[9], u[1], u[3]};  // n1 = 1: n = 5 + 2 n2 mod 10
  dft5<DIR>(a);
  dft5<DIR>(b);
  // k2 -> outputs (6 k2) mod 10 (k1 = 0) and (5 + 6 k2) mod 10 (k1 = 1)
  u[0] = cadd(a[0], b[0]);
  u[5] = csub(a[0], b[0]);
  u[6] = cadd(a[1], b[1]);
  u[1] = csub(a[1], b[1]);
  u[2] = cadd(a[2], b[2]);
  u[7] = csub(a[2], b[2]);
  u[8] = cadd(a[3], b[3]);
  u[3] = csub(a[3], b[3]);
  u[4] = cadd(a[4], b[4]);
  u[9] = csub(a[4], b[4]);
}

// generic odd prime radix through the length-L twiddle table (R | L)
template <int R, int DIR>
__device__ __forceinline__ void dftg(float2* u, const float2* tw, int L) {
  float2 out[R];
  const int st = L / R;
#pragma unroll
  for (int s = 0; s < R; ++s) {
    float2 acc = u[0];
#pragma unroll
    for (int q = 1; q < R; ++q) {
      float2 w = twid<DIR>(tw[((q * s) % R) * st]);
      acc = cadd(acc, cmul(u[q], w));
    }
    out[s] = acc;
  }
#pragma unroll
  for (int s = 0; s < R; ++s) u[s] = out[s];
}

template <int R, int DIR>
__device__ __forceinline__ void dft(float2* u, const float2* tw, int L) {
  if constexpr (R == 2) dft2<DIR>(u);
  else if constexpr (R == 3) dft3<DIR>(u);
  else if constexpr (R == 4) dft4<DIR>(u);
  else if constexpr (R == 5) dft5<DIR>(u);
  else if constexpr (R == 8) dft8<DIR>(u);
  else if constexpr (R == 10) dft10<DIR>(u);
  else dftg<R, DIR>(u, tw, L);
}

// One Stockham stage over nvec interleaved vectors (x -> y). lognv = log2(nvec).
template <int R, int DIR>
__device__ __forceinline__ void stockham_stage(const float2* __restrict__ x, float2* __restrict__ y,
                                               const float2* __restrict__ tw, int L, int lognv, int Ns, int tid,
                                               int nthr) {
  const int m = L / R;
  const int nvec = 1 << lognv;
  const int total = m << lognv;
  const int twstride = L / (Ns * R);
  for (int b = tid; b < total; b += nthr) {
    const int j = b >> lognv;
    const int v = b & (nvec - 1);
    const int k = j % Ns;
    float2 u[R];
#pragma unroll
    for (int q = 0; q < R; ++q) u[q] = x[((j + q * m) << lognv) + v];
    if (Ns > 1) {
#pragma unroll
      for (int q = 1; q < R; ++q) u[q] = cmul(u[q], twid<DIR>(tw[k * q * twstride]));
    }
    dft<R, DIR>(u, tw, L);
    const int base = (j - k) * R + k;
#pragma unroll
    for (int q = 0; q < R; ++q) y[((base + q * Ns) << lognv) + v] = u[q];
  }
}

// Runs every stage of `plan`; returns the buffer holding the result (a or b).
// Callers must __syncthreads() before (inputs written) — the function ends
// with a barrier so the result is visible to the whole block.
template <int DIR>
__device__ float2* stockham(float2* a, float2* b, const float2* tw, const FftPlan& plan, int lognv, int tid,
                            int nthr) {
  int Ns = 1;
  const int L = plan.L;
  for (int s = 0; s < plan.nst; ++s) {
    const int R = plan.radix[s];
    switch (R) {
      case 2: stockham_stage<2, DIR>(a, b, tw, L, lognv, Ns, tid, nthr); break;
      case 3: stockham_stage<3, DIR>(a, b, tw, L, lognv, Ns, tid, nthr); break;
      case 4: stockham_stage<4, DIR>(a, b, tw, L, lognv, Ns, tid, nthr); break;
      case 5: stockham_stage<5, DIR>(a, b, tw, L, lognv, Ns, tid, nthr); break;
      case 7: stockham_stage<7, DIR>(a, b, tw, L, lognv, Ns, tid, nthr); break;
      case 8: stockham_stage<8, DIR>(a, b, tw, L, lognv, Ns, tid, nthr); break;
      case 10: stockham_stage<10, DIR>(a, b, tw, L, lognv, Ns, tid, nthr); break;
      case 11: stockham_stage<11, DIR>(a, b, tw, L, lognv, Ns, tid, nthr); break;
      case 13: stockham_stage<13, DIR>(a, b, tw, L, lognv, Ns, tid, nthr); break;
      default: break;
    }
    __syncthreads();
    float2* t = a;
    a = b;
    b = t;
    Ns *= R;
  }
  return a;
}

}  // namespace qsmi
