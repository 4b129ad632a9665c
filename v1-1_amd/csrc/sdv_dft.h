// sdv_dft.h -- small in-register DFT codelets for the pencil FFTs.
//
// Forward:  X[k] = sum_j x[j] exp(-2 pi i jk/R)   (sign of DFTPACK zfftf and
//           FFTW_FORWARD, CP/ftcore.f90:84-92 / :140-157)
// Inverse:  X[k] = sum_j x[j] exp(+2 pi i jk/R)   (unnormalised)
//
// R = 2, 3, 4, 5, 7 are written out; composites (6, 8, 9, 10, 12, 16) are
// Cooley-Tukey in registers with compile-time twiddles.  All indices are
// compile-time constants after unrolling, so the arrays live in VGPRs.
#pragma once
#include "sdv_common.h"

namespace sdv {

// ---- constexpr cos/sin(2*pi*a/b) for compile-time twiddles -------------
constexpr double kPi = 3.141592653589793238462643383279502884;

constexpr double taylor_cos(double x) {
  double term = 1.0, sum = 1.0;
  for (int k = 1; k < 24; ++k) {
    term *= -x * x / ((2.0 * k - 1.0) * (2.0 * k));
    sum += term;
  }
  return sum;
}
constexpr double taylor_sin(double x) {
  double term = x, sum = x;
  for (int k = 1; k < 24; ++k) {
    term *= -x * x / ((2.0 * k) * (2.0 * k + 1.0));
    sum += term;
  }
  return sum;
}
// x = 2 pi a / b with the angle in [0, pi/2]
constexpr double cos2pi_q(long a, long b) {
  // cos x = sin(pi/2 - x) when x > pi/4 ; pi/2 - x = 2 pi (b - 4a) / (4b)
  return (8 * a > b) ? taylor_sin(2.0 * kPi * (double)(b - 4 * a) / (4.0 * (double)b))
                     : taylor_cos(2.0 * kPi * (double)a / (double)b);
}
// cos(2 pi a / b) for integers, reduced by exact symmetries to [0, pi/2].
constexpr double cos2pi(long a, long b) {
  a %= b;
  if (a < 0) a += b;
  if (2 * a > b) a = b - a;                                // cos is even about pi
  return (4 * a > b) ? -cos2pi_q(b - 2 * a, 2 * b)          // cos x = -cos(pi - x)
                     : cos2pi_q(a, b);
}
// sin x = cos(x - pi/2) ; x - pi/2 = 2 pi (4a - b) / (4b)
constexpr double sin2pi(long a, long b) { return cos2pi(4 * a - b, 4 * b); }

// Compile-time table of cos/sin(2 pi k / R); indexed with unrolled constant
// indices, so every load folds to an immediate.
template <int R>
struct TwTab {
  double c[R];
  double s[R];
  constexpr TwTab() : c(), s() {
    for (int k = 0; k < R; ++k) {
      c[k] = cos2pi(k, R);
      s[k] = sin2pi(k, R);
    }
  }
};
template <int R>
constexpr TwTab<R> kTw = TwTab<R>();

// ---- codelets -----------------------------------------------------------
template <typename T, int R, bool INV>
struct DFT;

template <typename T, bool INV>
struct DFT<T, 1, INV> {
  static __device__ __forceinline__ void run(cplx<T>*) {}
};

template <typename T, bool INV>
struct DFT<T, 2, INV> {
  static __device__ __forceinline__ void run(cplx<T>* v) {
    cplx<T> a = v[0], b = v[1];
    v[0] = a + b;
    v[1] = a - b;
  }
};

template <typename T, bool INV>
struct DFT<T, 3, INV> {
  static __device__ __forceinline__ void run(cplx<T>* v) {
    constexpr T s = (T)0.866025403784438646763723170752936183L;  // sin(2pi/3)
    cplx<T> t1 = v[1] + v[2];
    cplx<T> t2 = mk(v[0].re - (T)0.5 * t1.re, v[0].im - (T)0.5 * t1.im);
    cplx<T> d = v[1] - v[2];
    v[0] = v[0] + t1;
    // t2 -+ i s d, each component one expression (one fused multiply-add)
    const cplx<T> p = mk(t2.re + s * d.im, t2.im - s * d.re);
    const cplx<T> q = mk(t2.re - s * d.im, t2.im + s * d.re);
    v[1] = INV ? q : p;
    v[2] = INV ? p : q;
  }
};

template <typename T, bool INV>
struct DFT<T, 4, INV> {
  static __device__ __forceinline__ void run(cplx<T>* v) {
    cplx<T> a = v[0] + v[2], b = v[0] - v[2], c = v[1] + v[3], d = v[1] - v[3];
    cplx<T> id = INV ? mul_pi(d) : mul_mi(d);
    v[0] = a + c;
    v[2] = a - c;
    v[1] = b + id;
    v[3] = b - id;
  }
};

// odd prime R: symmetric pair formula
template <typename T, int R, bool INV>
struct DFTOddPrime {
  static __device__ __forceinline__ void run(cplx<T>* v) {
    constexpr int H = (R - 1) / 2;
    cplx<T> sp[H], sm[H];
#pragma unroll
    for (int j = 1; j <= H; ++j) {
      sp[j - 1] = v[j] + v[R - j];
      sm[j - 1] = v[j] - v[R - j];
    }
    cplx<T> x0 = v[0];
    cplx<T> tot = x0;
#pragma unroll
    for (int j = 0; j < H; ++j) tot = tot + sp[j];
    cplx<T> out[R];
    out[0] = tot;
#pragma unroll
    for (int k = 1; k <= H; ++k) {
      cplx<T> a = x0, b = mk((T)0, (T)0);
#pragma unroll
      for (int j = 1; j <= H; ++j) {
        const T c = (T)kTw<R>.c[(j * k) % R];
        const T s = (T)kTw<R>.s[(j * k) % R];
        a.re += sp[j - 1].re * c;
        a.im += sp[j - 1].im * c;
        b.re += sm[j - 1].re * s;
        b.im += sm[j - 1].im * s;
      }
      // forward: X_k = a - i b ; X_{R-k} = a + i b
      cplx<T> ib = mul_pi(b);
      if (!INV) {
        out[k] = a - ib;
        out[R - k] = a + ib;
      } else {
        out[k] = a + ib;
        out[R - k] = a - ib;
      }
    }
#pragma unroll
    for (int k = 0; k < R; ++k) v[k] = out[k];
  }
};
template <typename T, bool INV>
struct DFT<T, 5, INV> : DFTOddPrime<T, 5, INV> {};
template <typename T, bool INV>
struct DFT<T, 7, INV> : DFTOddPrime<T, 7, INV> {};

// composite R = A*B : X[k1 + A k2] = sum_j2 w_B^{j2 k2} w_R^{j2 k1} sum_j1 w_A^{j1 k1} x[j1 B + j2]
template <typename T, int A, int B, bool INV>
struct DFTComposite {
  static constexpr int R = A * B;
  static __device__ __forceinline__ void run(cplx<T>* v) {
    cplx<T> y[A][B];
#pragma unroll
    for (int j2 = 0; j2 < B; ++j2) {
      cplx<T> t[A];
#pragma unroll
      for (int j1 = 0; j1 < A; ++j1) t[j1] = v[j1 * B + j2];
      DFT<T, A, INV>::run(t);
#pragma unroll
      for (int k1 = 0; k1 < A; ++k1) {
        const T c = (T)kTw<R>.c[(j2 * k1) % R];
        const T s = (T)(INV ? kTw<R>.s[(j2 * k1) % R] : -kTw<R>.s[(j2 * k1) % R]);
        if (j2 * k1 == 0)
          y[k1][j2] = t[k1];
        else
          y[k1][j2] = t[k1] * mk(c, s);
      }
    }
#pragma unroll
    for (int k1 = 0; k1 < A; ++k1) {
      DFT<T, B, INV>::run(y[k1]);
#pragma unroll
      for (int k2 = 0; k2 < B; ++k2) v[k1 + A * k2] = y[k1][k2];
    }
  }
};
template <typename T, bool INV>
struct DFT<T, 6, INV> : DFTComposite<T, 3, 2, INV> {};
template <typename T, bool INV>
struct DFT<T, 8, INV> : DFTComposite<T, 2, 4, INV> {};
template <typename T, bool INV>
struct DFT<T, 9, INV> : DFTComposite<T, 3, 3, INV> {};
template <typename T, bool INV>
struct DFT<T, 10, INV> : DFTComposite<T, 2, 5, INV> {};
template <typename T, bool INV>
struct DFT<T, 12, INV> : DFTComposite<T, 3, 4, INV> {};
template <typename T, bool INV>
struct DFT<T, 16, INV> : DFTComposite<T, 4, 4, INV> {};

}  // namespace sdv
