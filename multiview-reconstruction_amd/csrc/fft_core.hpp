// fft_core.hpp -- complex arithmetic and the base-radix DFTs (2, 3, 4, 5, 8) held in registers, the building
// blocks of the register FFTs in fft_reg.hpp.
//
// Sign convention: DIR = -1 forward (exp(-2 pi i nk/L)), DIR = +1 inverse (unnormalised).
#pragma once
#include <hip/hip_runtime.h>

namespace mvr {

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
// (DIR * i) * a
template <int DIR>
__device__ __forceinline__ float2 mul_i(float2 a) {
  return DIR < 0 ? make_float2(a.y, -a.x) : make_float2(-a.y, a.x);
}

template <int R, int DIR>
__device__ __forceinline__ void dft(float2* v);

template <>
__device__ __forceinline__ void dft<2, -1>(float2* v) { float2 a = v[0], b = v[1]; v[0] = cadd(a, b); v[1] = csub(a, b); }
template <>
__device__ __forceinline__ void dft<2, 1>(float2* v) { float2 a = v[0], b = v[1]; v[0] = cadd(a, b); v[1] = csub(a, b); }

template <int DIR>
__device__ __forceinline__ void dft4(float2* v) {
  const float2 a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
  const float2 a2 = cadd(v[1], v[3]), a3 = mul_i<DIR>(csub(v[1], v[3]));
  v[0] = cadd(a0, a2);
  v[2] = csub(a0, a2);
  v[1] = cadd(a1, a3);
  v[3] = csub(a1, a3);
}
template <>
__device__ __forceinline__ void dft<4, -1>(float2* v) { dft4<-1>(v); }
template <>
__device__ __forceinline__ void dft<4, 1>(float2* v) { dft4<1>(v); }

template <int DIR>
__device__ __forceinline__ void dft8(float2* v) {
  // even/odd split into two DFT4, then combine with W8^k
  float2 e[4] = {v[0], v[2], v[4], v[6]};
  float2 o[4] = {v[1], v[3], v[5], v[7]};
  dft4<DIR>(e);
  dft4<DIR>(o);
  const float r = 0.70710678118654752440f;
  const float2 o1 = cmul(o[1], make_float2(r, DIR * r));
  const float2 o2 = mul_i<DIR>(o[2]);
  const float2 o3 = cmul(o[3], make_float2(-r, DIR * r));
  v[0] = cadd(e[0], o[0]);
  v[4] = csub(e[0], o[0]);
  v[1] = cadd(e[1], o1);
  v[5] = csub(e[1], o1);
  v[2] = cadd(e[2], o2);
  v[6] = csub(e[2], o2);
  v[3] = cadd(e[3], o3);
  v[7] = csub(e[3], o3);
}
template <>
__device__ __forceinline__ void dft<8, -1>(float2* v) { dft8<-1>(v); }
template <>
__device__ __forceinline__ void dft<8, 1>(float2* v) { dft8<1>(v); }

template <int DIR>
__device__ __forceinline__ void dft3(float2* v) {
  const float s = DIR * 0.86602540378443864676f;
  const float2 t1 = cadd(v[1], v[2]), t2 = csub(v[1], v[2]);
  const float2 m1 = make_float2(v[0].x - 0.5f * t1.x, v[0].y - 0.5f * t1.y);
  const float2 m2 = make_float2(-s * t2.y, s * t2.x);
  v[0] = cadd(v[0], t1);
  v[1] = cadd(m1, m2);
  v[2] = csub(m1, m2);
}
template <>
__device__ __forceinline__ void dft<3, -1>(float2* v) { dft3<-1>(v); }
template <>
__device__ __forceinline__ void dft<3, 1>(float2* v) { dft3<1>(v); }

template <int DIR>
__device__ __forceinline__ void dft5(float2* v) {
  const float c1 = 0.30901699437494742410f, c2 = -0.80901699437494742410f;
  const float s1 = 0.95105651629515357212f, s2 = 0.58778525229247312917f;
  const float2 t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
  const float2 t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
  const float2 a1 = make_float2(v[0].x + c1 * t1.x + c2 * t2.x, v[0].y + c1 * t1.y + c2 * t2.y);
  const float2 a2 = make_float2(v[0].x + c2 * t1.x + c1 * t2.x, v[0].y + c2 * t1.y + c1 * t2.y);
  const float2 b1 = mul_i<DIR>(make_float2(s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y));
  const float2 b2 = mul_i<DIR>(make_float2(s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y));
  v[0] = cadd(v[0], cadd(t1, t2));
  v[1] = cadd(a1, b1);
  v[4] = csub(a1, b1);
  v[2] = cadd(a2, b2);
  v[3] = csub(a2, b2);
}
template <>
__device__ __forceinline__ void dft<5, -1>(float2* v) { dft5<-1>(v); }
template <>
__device__ __forceinline__ void dft<5, 1>(float2* v) { dft5<1>(v); }

}  // namespace mvr
