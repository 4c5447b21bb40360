// Register-resident FFT codelets for gfx950 (HIP device code).
//
// fft<N, S>(v) computes the unnormalised DFT of the N values in v in place,
// v[n] <- sum_k v[k] exp(S * 2 pi i k n / N), S = -1 (forward) or +1
// (backward), in natural order. The transform is a mixed-radix Stockham
// sequence fully unrolled at compile time: every index is a constant, so the
// array stays in VGPRs, and every twiddle is a compile-time constant folded
// into the instruction stream (no table loads). Radix order: 4s, then 2,
// then odd primes; radix 2/3/4/5 butterflies are specialised, other primes
// use a direct butterfly with constant roots.
#pragma once

#include <hip/hip_runtime.h>

namespace agdfft {

constexpr double kPi = 3.14159265358979323846264338327950288;

constexpr double sin_poly(double x) {
  double x2 = x * x, t = x, s = x;
  for (int i = 1; i < 16; ++i) {
    t *= -x2 / ((2.0 * i) * (2.0 * i + 1));
    s += t;
  }
  return s;
}
constexpr double cos_poly(double x) {
  double x2 = x * x, t = 1.0, s = 1.0;
  for (int i = 1; i < 16; ++i) {
    t *= -x2 / ((2.0 * i - 1) * (2.0 * i));
    s += t;
  }
  return s;
}
// cos and sin of 2 pi m / n, reduced to |angle| <= pi/4 around a multiple
// of pi/2 so that the series is accurate to double rounding
constexpr double cos2pi(long m, long n) {
  m %= n;
  if (m < 0) m += n;
  long const q = (8 * m + n) / (2 * n);  // nearest multiple of pi/2
  double const r = 2.0 * kPi * ((double)m / (double)n - (double)q / 4.0);
  double const c = cos_poly(r), s = sin_poly(r);
  int const k = (int)(q & 3);
  return k == 0 ? c : k == 1 ? -s : k == 2 ? -c : s;
}
constexpr double sin2pi(long m, long n) {
  m %= n;
  if (m < 0) m += n;
  long const q = (8 * m + n) / (2 * n);
  double const r = 2.0 * kPi * ((double)m / (double)n - (double)q / 4.0);
  double const c = cos_poly(r), s = sin_poly(r);
  int const k = (int)(q & 3);
  return k == 0 ? s : k == 1 ? c : k == 2 ? -s : -c;
}

// exp(S * 2 pi i e / N) as a constant
template <int N, int S>
struct Roots {
  float re[N];
  float im[N];
  constexpr Roots() : re(), im() {
    for (int m = 0; m < N; ++m) {
      re[m] = (float)cos2pi(m, N);
      im[m] = (float)(S * sin2pi(m, N));
    }
  }
};

template <int N, int S>
__device__ __forceinline__ float2 root(int e) {
  constexpr Roots<N, S> t{};
  int const m = ((e % N) + N) % N;
  return make_float2(t.re[m], t.im[m]);
}

__device__ __forceinline__ float2 c_mul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 c_add(float2 a, float2 b) {
  return make_float2(a.x + b.x, a.y + b.y);
}
__device__ __forceinline__ float2 c_sub(float2 a, float2 b) {
  return make_float2(a.x - b.x, a.y - b.y);
}
// multiply by S*i
template <int S>
__device__ __forceinline__ float2 mul_si(float2 a) {
  return S < 0 ? make_float2(a.y, -a.x) : make_float2(-a.y, a.x);
}

constexpr int first_radix(int n) {
  if (n % 4 == 0) return 4;
  if (n % 2 == 0) return 2;
  for (int f = 3; f <= n; f += 2)
    if (n % f == 0) return f;
  return n;
}

template <int R, int S>
__device__ __forceinline__ void butterfly(float2 (&v)[R]) {
  if constexpr (R == 1) {
  } else if constexpr (R == 2) {
    float2 const a = v[0], b = v[1];
    v[0] = c_add(a, b);
    v[1] = c_sub(a, b);
  } else if constexpr (R == 4) {
    float2 const s02 = c_add(v[0], v[2]), d02 = c_sub(v[0], v[2]);
    float2 const s13 = c_add(v[1], v[3]), d13 = mul_si<S>(c_sub(v[1], v[3]));
    v[0] = c_add(s02, s13);
    v[1] = c_add(d02, d13);
    v[2] = c_sub(s02, s13);
    v[3] = c_sub(d02, d13);
  } else if constexpr (R == 3) {
    constexpr float c = -0.5f;
    constexpr float s = (float)(S * 0.866025403784438646763723170752936183);
    float2 const t1 = c_add(v[1], v[2]);
    float2 const d = c_sub(v[1], v[2]);
    float2 const t2 = make_float2(v[0].x + c * t1.x, v[0].y + c * t1.y);
    float2 const t3 = make_float2(-s * d.y, s * d.x);
    v[0] = c_add(v[0], t1);
    v[1] = c_add(t2, t3);
    v[2] = c_sub(t2, t3);
  } else if constexpr (R == 5) {
    constexpr float c1 = (float)0.309016994374947424102293417182819059;   // cos 72
    constexpr float c2 = (float)-0.809016994374947424102293417182819059;  // cos 144
    constexpr float s1 = (float)(S * 0.951056516295153572116439333379382143);
    constexpr float s2 = (float)(S * 0.587785252292473129168705954639072769);
    float2 const a1 = c_add(v[1], v[4]), b1 = c_sub(v[1], v[4]);
    float2 const a2 = c_add(v[2], v[3]), b2 = c_sub(v[2], v[3]);
    float2 const r1 = make_float2(v[0].x + c1 * a1.x + c2 * a2.x, v[0].y + c1 * a1.y + c2 * a2.y);
    float2 const r2 = make_float2(v[0].x + c2 * a1.x + c1 * a2.x, v[0].y + c2 * a1.y + c1 * a2.y);
    // i * (s1 b1 + s2 b2) and i * (s2 b1 - s1 b2)
    float2 const q1 = make_float2(s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y);
    float2 const q2 = make_float2(s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y);
    float2 const i1 = make_float2(-q1.y, q1.x);
    float2 const i2 = make_float2(-q2.y, q2.x);
    v[0] = make_float2(v[0].x + a1.x + a2.x, v[0].y + a1.y + a2.y);
    v[1] = c_add(r1, i1);
    v[4] = c_sub(r1, i1);
    v[2] = c_add(r2, i2);
    v[3] = c_sub(r2, i2);
  } else {
    float2 o[R];
#pragma unroll
    for (int q = 0; q < R; ++q) {
      float2 a = v[0];
#pragma unroll
      for (int r = 1; r < R; ++r) {
        if ((q * r) % R == 0) {
          a = c_add(a, v[r]);
        } else {
          a = c_add(a, c_mul(v[r], root<R, S>(q * r)));
        }
      }
      o[q] = a;
    }
#pragma unroll
    for (int q = 0; q < R; ++q) v[q] = o[q];
  }
}

template <int N, int R, int Ns, int S>
__device__ __forceinline__ void stockham_pass(float2 (&x)[N]) {
  float2 y[N];
#pragma unroll
  for (int j = 0; j < N / R; ++j) {
    int const k = j % Ns;
    float2 v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      v[r] = x[j + r * (N / R)];
      if (r > 0 && k > 0) v[r] = c_mul(v[r], root<Ns * R, S>(r * k));
    }
    butterfly<R, S>(v);
    int const base = (j / Ns) * Ns * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) y[base + r * Ns] = v[r];
  }
#pragma unroll
  for (int i = 0; i < N; ++i) x[i] = y[i];
}

template <int N, int Ns, int S>
__device__ __forceinline__ void stockham(float2 (&x)[N]) {
  if constexpr (Ns < N) {
    constexpr int R = first_radix(N / Ns);
    stockham_pass<N, R, Ns, S>(x);
    stockham<N, Ns * R, S>(x);
  }
}

template <int N, int S>
__device__ __forceinline__ void fft(float2 (&v)[N]) {
  stockham<N, 1, S>(v);
}

}  // namespace agdfft
