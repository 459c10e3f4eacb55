// CPU mixed-radix FFT used by the QS oracle (test infrastructure only).
//
// This is the oracle's own FFT. The reference computes its scan with FFTW3
// single precision (src/CPUThreadDevice.cpp:99-141, :417-422, :451), which is
// not present in this image; the oracle replaces it with a plain iterative
// Stockham autosort transform (Cooley-Tukey factorisation, natural order in
// and out) over radices 2,3,4,5,7,8,11,13 and a generic O(r^2) DFT for any
// other factor. Twiddles are generated in long double from the definition
// W_n^k = exp(-2*pi*i*k/n) and rounded once to T.
//
// Pinned in tests/test_oracle_fft.py against numpy.fft (pocketfft) on every
// size class the size rule can produce.
#pragma once
#include <cmath>
#include <complex>
#include <cstddef>
#include <vector>

namespace orc {

template <typename T>
struct FftPlan1D {
  using C = std::complex<T>;
  size_t n = 0;
  std::vector<int> radices;
  std::vector<C> tw;  // W_n^k, k in [0,n)

  explicit FftPlan1D(size_t n_) : n(n_) {
    size_t m = n;
    static const int pref[] = {8, 4, 2, 3, 5, 7, 11, 13};
    // take 8/4 first while possible, then the remaining prime factors
    while (m > 1) {
      bool found = false;
      for (int r : pref) {
        if (m % (size_t)r == 0) {
          radices.push_back(r);
          m /= r;
          found = true;
          break;
        }
      }
      if (!found) {  // generic prime factor
        size_t p = 17;
        while (m % p) p += 2;
        radices.push_back((int)p);
        m /= p;
      }
    }
    tw.resize(n);
    const long double two_pi = 6.283185307179586476925286766559L;
    for (size_t k = 0; k < n; ++k) {
      long double a = -two_pi * (long double)k / (long double)n;
      tw[k] = C((T)cosl(a), (T)sinl(a));
    }
  }

  // sign = -1 forward (exp(-i...)), +1 backward (exp(+i...)); unnormalised.
  void exec(C* data, int sign, std::vector<C>& work) const {
    if (n <= 1) return;
    work.resize(n);
    C* x = data;
    C* y = work.data();
    size_t Ns = 1;
    C u[64];
    C v[64];
    for (int R : radices) {
      size_t m = n / (size_t)R;
      size_t step = n / (Ns * (size_t)R);  // twiddle stride in the W_n table
      for (size_t j = 0; j < m; ++j) {
        size_t k = j % Ns;
        for (int q = 0; q < R; ++q) {
          C val = x[j + (size_t)q * m];
          size_t e = (k * (size_t)q * step) % n;
          C w = tw[e];
          if (sign > 0) w = std::conj(w);
          u[q] = val * w;
        }
        // small DFT of length R
        for (int s = 0; s < R; ++s) {
          C acc(0, 0);
          for (int q = 0; q < R; ++q) {
            size_t e = ((size_t)q * (size_t)s % (size_t)R) * (n / (size_t)R);
            C w = tw[e];
            if (sign > 0) w = std::conj(w);
            acc += u[q] * w;
          }
          v[s] = acc;
        }
        size_t base = (j / Ns) * Ns * (size_t)R + k;
        for (int s = 0; s < R; ++s) y[base + (size_t)s * Ns] = v[s];
      }
      std::swap(x, y);
      Ns *= (size_t)R;
    }
    if (x != data)
      for (size_t i = 0; i < n; ++i) data[i] = x[i];
  }
};

// Multi-dimensional transforms on a grid with dim0 fastest.
// dims[0..D): f0 (fastest), f1, f2. Half spectrum along dim0: h0 = f0/2+1.
template <typename T>
struct FftND {
  using C = std::complex<T>;
  int D;
  size_t f[3] = {1, 1, 1};
  size_t h0;
  std::vector<FftPlan1D<T>> plans;

  FftND(int D_, const unsigned* dims) : D(D_) {
    for (int i = 0; i < D; ++i) f[i] = dims[i];
    h0 = f[0] / 2 + 1;
    for (int i = 0; i < 3; ++i) plans.emplace_back(f[i]);
  }
  size_t real_size() const { return f[0] * f[1] * f[2]; }
  size_t half_size() const { return h0 * f[1] * f[2]; }

  // real (f-grid) -> half spectrum (h0 x f1 x f2), forward sign.
  // rowMask (optional, length f1*f2): rows along dim0 that are known to be
  // zero may be skipped, as the reference's partial r2c does
  // (src/CPUThreadDevice.cpp:411-420).
  void r2c(const T* in, C* out, const std::vector<char>* rowMask = nullptr) const {
    std::vector<C> line(f[0]), work;
    size_t rows = f[1] * f[2];
    for (size_t r = 0; r < rows; ++r) {
      C* o = out + r * h0;
      if (rowMask && !(*rowMask)[r]) {
        for (size_t k = 0; k < h0; ++k) o[k] = C(0, 0);
        continue;
      }
      for (size_t i = 0; i < f[0]; ++i) line[i] = C(in[r * f[0] + i], 0);
      plans[0].exec(line.data(), -1, work);
      for (size_t k = 0; k < h0; ++k) o[k] = line[k];
    }
    for (int d = 1; d < D; ++d) pass_along(out, d, -1);
  }

  // half spectrum -> real, backward sign, unnormalised (like FFTW c2r).
  void c2r(C* spec, T* out) const {
    for (int d = D - 1; d >= 1; --d) pass_along(spec, d, +1);
    std::vector<C> line(f[0]), work;
    size_t rows = f[1] * f[2];
    for (size_t r = 0; r < rows; ++r) {
      const C* s = spec + r * h0;
      for (size_t k = 0; k < h0; ++k) line[k] = s[k];
      for (size_t k = h0; k < f[0]; ++k) line[k] = std::conj(s[f[0] - k]);
      // the imaginary parts of the DC (and Nyquist) bins are ignored by a
      // Hermitian c2r; drop them explicitly
      line[0] = C(line[0].real(), 0);
      if (f[0] % 2 == 0) line[f[0] / 2] = C(line[f[0] / 2].real(), 0);
      plans[0].exec(line.data(), +1, work);
      for (size_t i = 0; i < f[0]; ++i) out[r * f[0] + i] = line[i].real();
    }
  }

  // complex transform of the half-spectrum array along dimension d (1 or 2)
  void pass_along(C* a, int d, int sign) const {
    size_t n = f[d];
    size_t stride = h0 * (d == 2 ? f[1] : 1);
    size_t outer = (d == 1) ? f[2] : 1;
    size_t inner = (d == 1) ? h0 : h0 * f[1];
    std::vector<C> line(n), work;
    for (size_t o = 0; o < outer; ++o) {
      for (size_t i = 0; i < inner; ++i) {
        C* base = a + o * (d == 1 ? h0 * f[1] : 0) + i;
        for (size_t t = 0; t < n; ++t) line[t] = base[t * stride];
        plans[d].exec(line.data(), sign, work);
        for (size_t t = 0; t < n; ++t) base[t * stride] = line[t];
      }
    }
  }
};

}  // namespace orc
