// fft_reg.hpp -- register-resident small DFTs and the two-step (L = A x B)
// pencil transform used by the y and z passes.
//
// rdft<N, DIR>(v) transforms N complex values held in registers in place
// (natural-order output).  N is split recursively as N = R x M (R a base
// radix 2, 3, 4, 5, 8): by the prime-factor algorithm where R and M are
// coprime (no twiddles between the steps), by Cooley-Tukey otherwise, whose
// twiddles are compile-time constants (computed by constexpr series in double,
// rounded to float); the index permutations are register renamings that the
// compiler resolves.
#pragma once
#include "fft_core.hpp"

namespace mvr {

// ------------------------------------------------------------------ constexpr cos/sin of 2*pi*p/q
constexpr double kTwoPi = 6.283185307179586476925286766559;
constexpr double kPiD = 3.141592653589793238462643383279;

constexpr double cx_sin_series(double x) {
  double x2 = x * x, term = x, sum = x;
  for (int i = 1; i < 16; ++i) {
    term *= -x2 / ((2.0 * i) * (2.0 * i + 1.0));
    sum += term;
  }
  return sum;
}
constexpr double cx_cos_series(double x) {
  double x2 = x * x, term = 1.0, sum = 1.0;
  for (int i = 1; i < 16; ++i) {
    term *= -x2 / ((2.0 * i - 1.0) * (2.0 * i));
    sum += term;
  }
  return sum;
}
// angle 2*pi*p/q reduced to [-pi, pi]
constexpr double cx_angle(long long p, long long q) {
  long long r = p % q;
  if (r < 0) r += q;
  double a = kTwoPi * (double)r / (double)q;
  return a > kPiD ? a - kTwoPi : a;
}
constexpr double cx_cos2pi(long long p, long long q) { return cx_cos_series(cx_angle(p, q)); }
constexpr double cx_sin2pi(long long p, long long q) { return cx_sin_series(cx_angle(p, q)); }

// W_N^e for direction DIR: exp(DIR * 2 pi i e / N)
template <int N, int DIR>
struct TwTab {
  float re[N];
  float im[N];
  constexpr TwTab() : re(), im() {
    for (int e = 0; e < N; ++e) {
      re[e] = (float)cx_cos2pi(e, N);
      im[e] = (float)(DIR * cx_sin2pi(e, N));
    }
  }
};

__host__ __device__ constexpr bool is_base_radix(int n) { return n == 2 || n == 3 || n == 4 || n == 5 || n == 8; }
__host__ __device__ constexpr int cx_gcd(int a, int b) { return b == 0 ? a : cx_gcd(b, a % b); }
// a base radix R with gcd(R, n / R) = 1 (n / R > 1), 0 if there is none
__host__ __device__ constexpr int coprime_radix(int n) {
  for (int R : {8, 4, 2, 5, 3})
    if (n % R == 0 && n / R > 1 && cx_gcd(R, n / R) == 1) return R;
  return 0;
}
// x with a x = 1 mod m (gcd(a, m) = 1, m > 1)
__host__ __device__ constexpr int cx_modinv(int a, int m) {
  for (int x = 1; x < m; ++x)
    if ((a * x) % m == 1) return x;
  return 1;
}
__host__ __device__ constexpr int split_radix(int n) {
  return n % 4 == 0 ? 4 : n % 8 == 0 ? 8 : n % 2 == 0 ? 2 : n % 5 == 0 ? 5 : n % 3 == 0 ? 3 : 0;
}
__host__ __device__ constexpr bool is_smooth(int n) {
  while (n % 2 == 0) n /= 2;
  while (n % 3 == 0) n /= 3;
  while (n % 5 == 0) n /= 5;
  return n == 1;
}

// PF: coprime splits by the prime-factor algorithm (false: Cooley-Tukey throughout)
template <int N, int DIR, bool PF = true>
__device__ __forceinline__ void rdft(float2* v);

template <int N, int DIR, bool PF, bool BASE = is_base_radix(N)>
struct RDFT;

template <int N, int DIR, bool PF>
struct RDFT<N, DIR, PF, true> {
  __device__ __forceinline__ static void run(float2* v) { dft<N, DIR>(v); }
};

template <int N, int DIR, bool PF>
struct RDFT<N, DIR, PF, false> {
  // N = N1 x N2 with coprime factors (N1 a base radix): the prime-factor (Good-Thomas) algorithm, no twiddles
  // between the steps.  Input x[(N2 n1 + N1 n2) mod N] (n1 < N1, n2 < N2); N1-point DFTs over n1, N2-point
  // DFTs over n2; X[k1, k2] is X[(N2 a k1 + N1 b k2) mod N] with a = N2^-1 mod N1, b = N1^-1 mod N2, since then
  // W_N^{n k} = W_N1^{n1 k1} W_N2^{n2 k2}.  Both index maps are register renamings.
  static constexpr int P = PF ? coprime_radix(N) : 0;
  // otherwise Cooley-Tukey N = R x M with inter-step twiddles
  static constexpr int R = P > 0 ? P : split_radix(N);
  static constexpr int M = N / R;
  static_assert(R > 0 && M > 1, "register DFT length must be 2^a 3^b 5^c");
  __device__ __forceinline__ static void run(float2* v) {
    if constexpr (P > 0) {
      constexpr int N1 = P, N2 = N / P;
      constexpr int a = cx_modinv(N2 % N1, N1), b = cx_modinv(N1 % N2, N2);
      float2 t[N];
#pragma unroll
      for (int n2 = 0; n2 < N2; ++n2) {
        float2 u[N1];
#pragma unroll
        for (int n1 = 0; n1 < N1; ++n1) u[n1] = v[(N2 * n1 + N1 * n2) % N];
        dft<N1, DIR>(u);
#pragma unroll
        for (int k1 = 0; k1 < N1; ++k1) t[N2 * k1 + n2] = u[k1];
      }
#pragma unroll
      for (int k1 = 0; k1 < N1; ++k1) RDFT<N2, DIR, PF>::run(t + N2 * k1);
#pragma unroll
      for (int k1 = 0; k1 < N1; ++k1)
#pragma unroll
        for (int k2 = 0; k2 < N2; ++k2) v[(N2 * a * k1 + N1 * b * k2) % N] = t[N2 * k1 + k2];
      return;
    }
    constexpr TwTab<N, DIR> tw{};
    // step A: R-point DFTs over stride-M subsequences, twiddle W_N^{j k1}
#pragma unroll
    for (int j = 0; j < M; ++j) {
      float2 t[R];
#pragma unroll
      for (int a = 0; a < R; ++a) t[a] = v[j + M * a];
      dft<R, DIR>(t);
#pragma unroll
      for (int k1 = 0; k1 < R; ++k1) {
        const int e = (j * k1) % N;
        v[j + M * k1] = e == 0 ? t[k1] : cmul(t[k1], make_float2(tw.re[e], tw.im[e]));
      }
    }
    // step B: M-point DFTs over the contiguous blocks [M k1, M k1 + M)
#pragma unroll
    for (int k1 = 0; k1 < R; ++k1) RDFT<M, DIR, PF>::run(v + M * k1);
    // block position M k1 + k2 holds X[k1 + R k2]: rename into natural order
    float2 o[N];
#pragma unroll
    for (int k1 = 0; k1 < R; ++k1)
#pragma unroll
      for (int k2 = 0; k2 < M; ++k2) o[k1 + R * k2] = v[M * k1 + k2];
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] = o[i];
  }
};

template <int N, int DIR, bool PF>
__device__ __forceinline__ void rdft(float2* v) {
  RDFT<N, DIR, PF>::run(v);
}

// ------------------------------------------------------------------ two-step split of a pencil length
struct Split2 {
  int A, B;
};
// L = A x B with A >= B, both smooth and <= amax, max(A, B) as small as possible; {0, 0} if none
__host__ __device__ constexpr Split2 split2_max(int L, int amax) {
  Split2 best{0, 0};
  for (int b = 2; b <= 32; ++b) {
    if (L % b) continue;
    const int a = L / b;
    if (a < b || a > amax || !is_smooth(a) || !is_smooth(b)) continue;
    if (best.A == 0 || a < best.A) best = Split2{a, b};
  }
  return best;
}
// factors up to 32; lengths above 1024 (no such split) take A up to 40 (1080 = 36 x 30, 1152 = 36 x 32,
// 1280 = 40 x 32)
__host__ __device__ constexpr Split2 split2(int L) {
  return split2_max(L, 32).A ? split2_max(L, 32) : split2_max(L, 40);
}

}  // namespace mvr
