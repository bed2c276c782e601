// xpasses.hpp -- x (real axis) passes with the two-step register FFT.
//
// A work-group owns RW rows of one x-pass; each row is served by T lanes, T the power of two >= A, so that a
// row never straddles a wave.  The length-M complex transform of a row (M = P0/2, real FFT by the even/odd
// split) runs as M = A x B with one LDS exchange.  The R2C post-processing and the C2R pre-processing combine
// the values k and M - k of a row; after the second forward step (before the first inverse step) lane q holds
// index q + A k2 in register k2, and M - (q + A k2) = (A - q) + A (B - 1 - k2), so the partner is lane A - q,
// register B - 1 - k2 of the same row: the pairs are swapped with wave shuffles, and spectra go between HBM
// and registers directly, coalesced across the row's lanes.  Real-valued data meets the transforms in
// registers too: psi is read into the first forward step, and the quotient and the psi update work on the
// values the inverse leaves in lane q < B (x[2n], x[2n + 1] for n = q + B a).  Index arithmetic that depends on
// runtime sizes is done once per row (row tables in LDS), never per element with '/' or '%'.
#pragma once
#include "passes.hpp"

namespace mvr {

// exact n / d for n < 2^32 / d (d >= 1)
struct FastDiv {
  unsigned d, m;
  __host__ __device__ FastDiv() : d(1), m(0) {}
  __host__ __device__ explicit FastDiv(unsigned dd) : d(dd), m((unsigned)((0x100000000ull + dd - 1) / dd)) {}
  __device__ __forceinline__ unsigned div(unsigned n) const {
    return d == 1 ? n : (unsigned)(((unsigned long long)n * m) >> 32);
  }
};

template <int NT>
constexpr int reduce_width() {
  return NT <= 256 ? 256 : NT <= 512 ? 512 : 1024;  // power of two >= NT
}

// {sum, max} of a work-group of NT threads (NT <= 1024) in the caller's LDS: rs and rm hold
// reduce_width<NT>() doubles each.  Every thread must be done with that memory before the call.
template <int NT>
__device__ __forceinline__ void block_reduce_stats_nt(double s, double mx, double2* out, double* rs, double* rm) {
  constexpr int P = reduce_width<NT>();
  rs[threadIdx.x] = s;
  rm[threadIdx.x] = mx;
  for (int i = NT + (int)threadIdx.x; i < P; i += NT) {
    rs[i] = 0.0;
    rm[i] = -1.0;
  }
  __syncthreads();
#pragma unroll
  for (int off = P / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) {
      rs[threadIdx.x] += rs[threadIdx.x + off];
      rm[threadIdx.x] = java_max(rm[threadIdx.x], rm[threadIdx.x + off]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = make_double2(rs[0], rm[0]);
}

constexpr long long kOuterRow = -(1ll << 62);
constexpr size_t kNoRow = ~(size_t)0;

template <int M>
struct XCfg {
  static constexpr Split2 S = split2(M);
  static constexpr int A = S.A, B = S.B;
  static constexpr int T = A <= 16 ? 16 : A <= 32 ? 32 : 64;  // lanes per row (a power of two, divides 64)
  static constexpr int RW = 256 / T;                          // rows per work-group
  static constexpr int NT = RW * T;                           // threads per work-group
  static constexpr int RS = M + 5;                 // LDS row stride (complex): M used, +5 against bank conflicts
  static constexpr int PX = 2 * M;
  // the exchange rows, at least as large as the stats reduction that reuses them (2 x 256 doubles)
  static constexpr int BUF = RW * RS > 512 ? RW * RS : 512;
  // Twiddle tables (2M + 1 complex) stay in LDS unless leaving them in global memory (read through L1) fits one
  // more work-group per CU into the 160 KB of LDS; the other LDS is the exchange rows and 16-24 B of row tables
  // per row (the stats reduction reuses the rows), plus the x gather map in the forward pass.
  static constexpr int LDS_ROWS = RW * RS * 8 + RW * 24;
  static constexpr int LDS_TW = (2 * M + 1) * 8;
  static constexpr bool TWL = (160 * 1024) / (LDS_ROWS + LDS_TW) >= (160 * 1024) / LDS_ROWS;
  // The forward and quotient passes follow the same rule on long rows (B > 16; at M = 576 global twiddles gave
  // both a third work-group per CU and 2-5 %) and keep LDS twiddles on shorter rows (at M = 150 global twiddles
  // were slower despite the extra work-group).
  static constexpr bool long_rows = B > 16;
  // The forward pass exchanges in two rounds (half of the second step's lanes per round), so its rows hold
  // half a transform: at M = 150 8 instead of 6 work-groups per CU, at M = 576 4 instead of 3.
  static constexpr int AH = (A + 1) / 2;
  static constexpr int RS_FWD = AH * B + 5;
  static constexpr int BUF_FWD = RW * RS_FWD;
  static constexpr int LDS_ROWS_FWD = RW * RS_FWD * 8 + RW * 24;
  static constexpr bool TWL_FWD = !long_rows || (160 * 1024) / (LDS_ROWS_FWD + PX * 4 + LDS_TW) >=
                                                   (160 * 1024) / (LDS_ROWS_FWD + PX * 4);
  static constexpr bool TWL_RATIO = !long_rows || TWL;
  // The quotient pass at M = 100, 120, 150, 200 compiles to 81 VGPRs (this ROCm), one over what 6 waves per
  // SIMD allow; bounded to 6 waves it takes 80 without spilling (x_ratio 4 % faster at M = 150).  Other lengths
  // either fit already or would spill under the bound.
  static constexpr int MINW_RATIO = (M == 100 || M == 120 || M == 150 || M == 200) ? 6 : 1;
};

template <int M>
__host__ __device__ constexpr bool x2_ok() {
  return split2(M).A > 0 && split2(M).B > 1;
}

__device__ __forceinline__ float2 shfl2(float2 v, int src) {
  return make_float2(__shfl(v.x, src, 64), __shfl(v.y, src, 64));
}

// For values V[q + A k2] held as v[k2] by lane q < A of a row (and V[M] as vM in lane 0): out[k2] = V[M - (q + A k2)].
// Lane q > 0 gets them from lane A - q (register B - 1 - k2), lane 0 from itself (M - A k2 = A (B - k2)).
// Every lane of the wave must call it (wave shuffle); idle lanes (q >= A) read themselves.
template <int M>
__device__ __forceinline__ void x2_mirror_pairs(const float2* v, float2 vM, int q, float2* out) {
  using C = XCfg<M>;
  const int lane = (int)(threadIdx.x & 63);
  const int src = (lane & ~(C::T - 1)) + (q > 0 && q < C::A ? C::A - q : q);
#pragma unroll
  for (int k2 = 0; k2 < C::B; ++k2) {
    const float2 t = shfl2(v[C::B - 1 - k2], src);
    out[k2] = q != 0 ? t : (k2 == 0 ? vM : v[(C::B - k2) % C::B]);
  }
}

// forward transform whose first-step inputs v[a] = z[q + B a] (lane q < B) are in registers; LDS row `rb` is
// only written -> out[k2] = Z[q + A k2] (lane q < A)
template <int M>
__device__ __forceinline__ void x2_forward_regs(float2* rb, const float2* tw, int q, float2* v, float2* out) {
  using C = XCfg<M>;
  constexpr int A = C::A, B = C::B;
  if (q < B) {
    rdft<A, -1>(v);
#pragma unroll
    for (int k1 = 0; k1 < A; ++k1) rb[q + B * k1] = k1 == 0 ? v[k1] : cmul(v[k1], tw[q * k1]);
  }
  __syncthreads();
  if (q < A) {
#pragma unroll
    for (int b = 0; b < B; ++b) out[b] = rb[b + B * q];
    rdft<B, -1>(out);
  } else {
#pragma unroll
    for (int b = 0; b < B; ++b) out[b] = make_float2(0.f, 0.f);
  }
}

// x2_forward_regs with the exchange in two rounds through a row `rb` of AH * B values: round r carries the
// first-step results k1 in [r AH, r AH + AH), read by the lanes q = k1 of that range
template <int M>
__device__ __forceinline__ void x2_forward_regs_half(float2* rb, const float2* tw, int q, float2* v, float2* out) {
  using C = XCfg<M>;
  constexpr int A = C::A, B = C::B, AH = C::AH;
  if (q < B) {
    rdft<A, -1>(v);
#pragma unroll
    for (int k1 = 1; k1 < A; ++k1) v[k1] = cmul(v[k1], tw[q * k1]);
  }
#pragma unroll
  for (int b = 0; b < B; ++b) out[b] = make_float2(0.f, 0.f);
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    const int k0 = r * AH, kn = r == 0 ? AH : A - AH;
    if (r > 0) __syncthreads();  // the first round's readers are done
    if (q < B) {
#pragma unroll
      for (int j = 0; j < AH; ++j)
        if (j < kn) rb[q + B * j] = v[k0 + j];
    }
    __syncthreads();
    if (q >= k0 && q < k0 + kn) {
#pragma unroll
      for (int b = 0; b < B; ++b) out[b] = rb[b + B * (q - k0)];
    }
  }
  if (q < A) rdft<B, -1>(out);
}

// forward transform of LDS row `rb` (natural order) -> out[k2] = Z[q + A k2] (lane q < A)
template <int M>
__device__ __forceinline__ void x2_forward(float2* rb, const float2* tw, int q, float2* out) {
  using C = XCfg<M>;
  float2 v[C::A];
  if (q < C::B) {
#pragma unroll
    for (int a = 0; a < C::A; ++a) v[a] = rb[q + C::B * a];
  }
  __syncthreads();  // every lane has read its values before the row is overwritten
  x2_forward_regs<M>(rb, tw, q, v, out);
}

// R2C post-processing of a row whose complex transform Z lane q holds as z[k2] = Z[q + A k2]:
// X[k] = (Z[k] + Z*[M - k]) / 2 + W_P0^k (Z[k] - Z*[M - k]) / 2i for k = q + A k2, and X[M] (lane 0), written
// to dst (the row's start in HBM) when rowOk
template <int M>
__device__ __forceinline__ void x2_post_store_regs(const float2* z, int q, const float2* twr, float2* __restrict__ dst,
                                                   bool rowOk) {
  using C = XCfg<M>;
  float2 zc[C::B];
  x2_mirror_pairs<M>(z, z[0], q, zc);  // Z[M] = Z[0]
  if (!rowOk || q >= C::A) return;
#pragma unroll
  for (int k2 = 0; k2 < C::B; ++k2) {
    const int k = q + C::A * k2;
    const float2 zk = z[k2];
    const float2 c = cconj(zc[k2]);
    const float2 e = make_float2(0.5f * (zk.x + c.x), 0.5f * (zk.y + c.y));
    const float2 dd = csub(zk, c);
    const float2 od = make_float2(0.5f * dd.y, -0.5f * dd.x);
    dst[k] = cadd(e, cmul(twr[k], od));
  }
  if (q == 0) {
    const float2 zk = z[0], c = cconj(z[0]);
    const float2 e = make_float2(0.5f * (zk.x + c.x), 0.5f * (zk.y + c.y));
    const float2 dd = csub(zk, c);
    const float2 od = make_float2(0.5f * dd.y, -0.5f * dd.x);
    dst[M] = cadd(e, cmul(twr[M], od));
  }
}

// inverse of the half spectrum X[0..M] of one row (src: its start in HBM; null: a row of zeros), read straight
// into registers: the C2R pre-processing Z[k] = E[k] + i O[k] combines X[k] and X*[M - k]; returns
// z[q + B n1] in out[n1] (lane q < B)
// PF: the transforms' coprime splits by the prime-factor algorithm (false: Cooley-Tukey throughout, which the
// update pass on short rows runs 3 % faster at M = 150)
template <int M, bool PF = true>
__device__ __forceinline__ void x2_inverse_half_regs(const float2* __restrict__ src, float2* rb, const float2* tw,
                                                     const float2* twr, int q, float2* out) {
  using C = XCfg<M>;
  constexpr int A = C::A, B = C::B;
  float2 x[B];
  float2 xM = make_float2(0.f, 0.f);
  const bool ld = src != nullptr && q < A;
#pragma unroll
  for (int k2 = 0; k2 < B; ++k2) x[k2] = ld ? src[q + A * k2] : make_float2(0.f, 0.f);
  if (ld && q == 0) xM = src[M];
  float2 xc[B];
  x2_mirror_pairs<M>(x, xM, q, xc);
  float2 u[B];
  const int qq = q < A ? q : 0;  // idle lanes compute on a valid twiddle index and store nothing
#pragma unroll
  for (int k2 = 0; k2 < B; ++k2) {
    const int k = qq + A * k2;
    const float2 a = x[k2];
    const float2 b = cconj(xc[k2]);
    const float2 e = make_float2(0.5f * (a.x + b.x), 0.5f * (a.y + b.y));
    const float2 o = cmul(make_float2(0.5f * (a.x - b.x), 0.5f * (a.y - b.y)), cconj(twr[k]));
    u[k2] = make_float2(e.x - o.y, e.y + o.x);
  }
  rdft<B, 1, PF>(u);
  __syncthreads();  // the row buffer is free (previous readers are done)
  if (q < A) {
#pragma unroll
    for (int n2 = 0; n2 < B; ++n2) rb[q + A * n2] = n2 == 0 ? u[n2] : cmul(u[n2], cconj(tw[n2 * q]));
  }
  __syncthreads();
  if (q < B) {
#pragma unroll
    for (int k1 = 0; k1 < A; ++k1) out[k1] = rb[k1 + A * q];
    rdft<A, 1, PF>(out);
  }
}

template <int M>
__device__ __forceinline__ void x2_load_twiddles(float2* tw, float2* twr, const float2* twM, const float2* wr) {
  using C = XCfg<M>;
  for (int i = threadIdx.x; i < M; i += C::NT) tw[i] = twM[i];
  for (int i = threadIdx.x; i <= M; i += C::NT) twr[i] = wr[i];
}

// the inverse result z[q + B n1] (lane q < B, register n1) into LDS row `rb` in natural order
template <int M>
__device__ __forceinline__ void x2_park_inverse(float2* rb, int q, const float2* z) {
  using C = XCfg<M>;
  __syncthreads();  // every lane has read the row's exchange values
  if (q < C::B) {
#pragma unroll
    for (int n1 = 0; n1 < C::A; ++n1) rb[q + C::B * n1] = z[n1];
  }
  __syncthreads();
}

// ------------------------------------------------------------------ forward with gather
template <int M>
__global__ __launch_bounds__(XCfg<M>::NT) void k_x2_fwd(Dom d, GatherParams g, float2* __restrict__ S,
                                                         const float2* __restrict__ twM,
                                                         const float2* __restrict__ wr) {
  using C = XCfg<M>;
  __shared__ float2 buf[C::BUF_FWD];
  __shared__ float2 twS[C::TWL_FWD ? M : 1];
  __shared__ float2 twrS[C::TWL_FWD ? M + 1 : 1];
  const float2* tw = C::TWL_FWD ? twS : twM;
  const float2* twr = C::TWL_FWD ? twrS : wr;
  __shared__ int xmap[C::PX];
  __shared__ long long srcRow[C::RW];
  __shared__ size_t dstRow[C::RW];
  if (C::TWL_FWD) x2_load_twiddles<M>(twS, twrS, twM, wr);
  const int nrows = d.P[1] * d.P[2];
  // batched blocks: this work-group's block, its gather origin (output region origin - m) and spectrum
  int wg = blockIdx.x;
  int org[3] = {g.org[0], g.org[1], g.org[2]};
  if (d.btab) {
    const int b = blockIdx.x / d.wgPerBlock;
    wg -= b * d.wgPerBlock;
    const int* t = d.btab + kBlockTabInts * b;
    for (int k = 0; k < 3; ++k) org[k] = t[k] - d.m[k];
    S += (size_t)b * d.blockStride;
  }
  const int row0 = wg * C::RW;
  const int r = threadIdx.x / C::T, q = threadIdx.x % C::T;
  float2* rb = buf + r * C::RS_FWD;
  // unique rows only (mirror gather): row index -> grid row (-org1 + i % N1, -org2 + i / N1)
  const bool uniq = g.uniqueRows && g.mode == GATHER_MIRROR;
  const int nrowsRun = uniq ? g.N[1] * g.N[2] : nrows;
  // source index of FFT position x of a row: row start (srcRow, -1: no source row) + xmap[x] (-1: outside),
  // else g.value.  Mirror: the volume mirrored at its faces; value: the volume, `value` outside; kernel
  // placement: K[(p + K/2) mod P], 0 beyond K (GATHER_KERNEL, value 0).
  const bool kern = g.mode == GATHER_KERNEL;
  const float fill = kern ? 0.f : g.value;
  for (int x = threadIdx.x; x < C::PX; x += C::NT) {
    int gx;
    if (kern) {
      gx = (x + g.N[0] / 2) % d.P[0];
      if (gx >= g.N[0]) gx = -1;
    } else {
      const int p = org[0] + x;
      gx = g.mode == GATHER_MIRROR ? mirror_index(p, g.N[0]) : ((unsigned)p < (unsigned)g.N[0] ? p : -1);
    }
    xmap[x] = gx;
  }
  if (threadIdx.x < C::RW) {
    int row = row0 + threadIdx.x;
    long long off = -1;
    const bool live = row < nrowsRun;
    if (uniq && live) row = (-org[2] + row / g.N[1]) * d.P[1] + (-org[1] + row % g.N[1]);
    if (live) {
      const int y = row % d.P[1], zz = row / d.P[1];
      if (kern) {
        const int jy = (y + g.N[1] / 2) % d.P[1], jz = (zz + g.N[2] / 2) % d.P[2];
        if (jy < g.N[1] && jz < g.N[2]) off = ((long long)jz * g.N[1] + jy) * g.N[0];
      } else {
        int gy = org[1] + y, gz = org[2] + zz;
        if (g.mode == GATHER_MIRROR) {
          gy = mirror_index(gy, g.N[1]);
          gz = mirror_index(gz, g.N[2]);
          off = ((long long)gz * g.N[1] + gy) * g.N[0];
        } else if ((unsigned)gy < (unsigned)g.N[1] && (unsigned)gz < (unsigned)g.N[2]) {
          off = ((long long)gz * g.N[1] + gy) * g.N[0];
        }
      }
    }
    srcRow[threadIdx.x] = off;   // row start in src (>= 0), -1: no source row
    dstRow[threadIdx.x] = live ? (size_t)row * d.HxP : kNoRow;
  }
  __syncthreads();
  // the first transform step reads its values z[q + B a] = (x[2n], x[2n + 1]) straight from the source
  float2 v[C::A];
  if (q < C::B) {
    const long long off = srcRow[r];
#pragma unroll
    for (int a = 0; a < C::A; ++a) {
      const int x = 2 * (q + C::B * a);
      const int g0 = xmap[x], g1 = xmap[x + 1];
      const bool ok0 = off >= 0 && g0 >= 0, ok1 = off >= 0 && g1 >= 0;
      const float* p0 = g.src + (ok0 ? off + g0 : 0);
      if (ok0 && g1 == g0 + 1 && !(reinterpret_cast<uintptr_t>(p0) & 7)) {
        v[a] = *reinterpret_cast<const float2*>(p0);  // an aligned contiguous pair: one 8-byte load
      } else {
        v[a].x = ok0 ? *p0 : fill;
        v[a].y = ok1 ? g.src[off + g1] : fill;
      }
    }
  }
  float2 z[C::B];
  x2_forward_regs_half<M>(rb, tw, q, v, z);
  const size_t o = dstRow[r];
  x2_post_store_regs<M>(z, q, twr, S + (o == kNoRow ? 0 : o), o != kNoRow);
}

// ------------------------------------------------------------------ inverse, quotient, forward
// ratio = img > 0 ? img / blurred : 1 on the quotient region, 1 elsewhere
// (DeconvolutionMethods.computeQuotient, D/iteration/sequential/DeconvolutionMethods.java:46-98)
template <int M>
__global__ __launch_bounds__(XCfg<M>::NT, XCfg<M>::MINW_RATIO) void k_x2_ratio(Dom d, RatioParams q_, const float2* __restrict__ Sin,
                                                           float2* __restrict__ Sout,
                                                           const float2* __restrict__ twM,
                                                           const float2* __restrict__ wr) {
  using C = XCfg<M>;
  __shared__ float2 buf[C::BUF];
  __shared__ float2 twS[C::TWL_RATIO ? M : 1];
  __shared__ float2 twrS[C::TWL_RATIO ? M + 1 : 1];
  const float2* tw = C::TWL_RATIO ? twS : twM;
  const float2* twr = C::TWL_RATIO ? twrS : wr;
  // img index of FFT position x is imgRow + x (may be negative: x starts at m0); kOuterRow: ratio row of ones
  __shared__ long long imgRow[C::RW];
  __shared__ size_t outOff[C::RW];   // spectrum row (read and written; kNoRow: beyond the grid)
  if (C::TWL_RATIO) x2_load_twiddles<M>(twS, twrS, twM, wr);
  const int nrows = d.P[1] * d.P[2];
  // batched blocks: this work-group's block, its output region and spectra
  int wg = blockIdx.x;
  int o[3] = {q_.o[0], q_.o[1], q_.o[2]}, D[3] = {d.D[0], d.D[1], d.D[2]};
  if (d.btab) {
    const int b = blockIdx.x / d.wgPerBlock;
    wg -= b * d.wgPerBlock;
    const int* t = d.btab + kBlockTabInts * b;
    for (int k = 0; k < 3; ++k) { o[k] = t[k]; D[k] = t[3 + k]; }
    Sin += (size_t)b * d.blockStride;
    Sout += (size_t)b * d.blockStride;
  }
  const int row0 = wg * C::RW;
  if (threadIdx.x < C::RW) {
    // skipOnes: the work-groups cover the quotient rows only, index i -> grid row (rowLo1 + i % rowN1,
    // rowLo2 + i / rowN1)
    int row = row0 + threadIdx.x;
    const bool live = row < (q_.skipOnes ? q_.rowN[1] * q_.rowN[2] : nrows);
    if (q_.skipOnes && live) row = (q_.rowLo[2] + row / q_.rowN[1]) * d.P[1] + q_.rowLo[1] + row % q_.rowN[1];
    long long ir = kOuterRow;
    if (live) {
      // quotient rows: the output region widened by qlo / qhi, inside the volume
      const int ly = row % d.P[1] - d.m[1], lz = row / d.P[1] - d.m[2];
      if (ly >= -q_.qlo[1] && ly < D[1] + q_.qhi[1] && lz >= -q_.qlo[2] && lz < D[2] + q_.qhi[2] &&
          (unsigned)(o[1] + ly) < (unsigned)q_.N[1] && (unsigned)(o[2] + lz) < (unsigned)q_.N[2])
        ir = ((long long)(o[2] + lz) * q_.N[1] + (o[1] + ly)) * q_.N[0] + o[0] - d.m[0];
    }
    imgRow[threadIdx.x] = ir;
    outOff[threadIdx.x] = live ? (size_t)row * d.HxP : kNoRow;
  }
  __syncthreads();
  const int r = threadIdx.x / C::T, q = threadIdx.x % C::T;
  float2* rb = buf + r * C::RS;
  const long long ir = imgRow[r];
  const size_t so = outOff[r];
  // FFT positions x of the row whose image value is read: the quotient region, inside the volume
  const int x0f = d.m[0] + max(-q_.qlo[0], -o[0]), x1f = d.m[0] + min(D[0] + q_.qhi[0], q_.N[0] - o[0]);
  // rows outside the quotient region are a row of ones after the quotient: their blurred values are not read
  float2 z[C::A];
  x2_inverse_half_regs<M>(ir != kOuterRow ? Sin + so : nullptr, rb, tw, twr, q, z);
  float2 zf[C::B];
  if constexpr (!C::long_rows) {
    // short rows (only B of T lanes hold values): park the blurred rows in LDS, form the quotient there with the
    // row's T lanes over positions x = q + T i, all of a lane's image values in flight at once (offsets from the
    // lane's row pointers known at compile time; at M = 150 4 % faster than a flat mapping over the work-group,
    // 9 % faster than a wave walking its rows one by one), and run the forward transform from LDS
    x2_park_inverse<M>(rb, q, z);
    float* fr = reinterpret_cast<float*>(rb);
    constexpr int NI = (C::PX + C::T - 1) / C::T;
    const bool rowIn = ir != kOuterRow;
    const float* imr = q_.img + (rowIn ? ir : 0);  // image value of FFT position x: imr[x], x in [x0f, x1f)
    float im[NI];
#pragma unroll
    for (int i = 0; i < NI; ++i) {
      const int x = q + C::T * i;
      im[i] = rowIn && x >= x0f && x < x1f ? imr[x] : 0.f;
    }
#pragma unroll
    for (int i = 0; i < NI; ++i) {
      const int x = q + C::T * i;
      if (x < C::PX) fr[x] = im[i] > 0.f ? im[i] / fr[x] : 1.f;
    }
    __syncthreads();
    x2_forward<M>(rb, tw, q, zf);
  } else {
    // long rows: lane q < B holds blurred values z[a] = (x[2n], x[2n + 1]), n = q + B a, exactly the inputs of its
    // first forward step: the quotient is formed there, with the image read straight from HBM
    if (q < C::B) {
      constexpr int H = (C::A + 1) / 2;  // two halves of loads in flight
#pragma unroll
      for (int a0 = 0; a0 < C::A; a0 += H) {
        // pair h: image values of FFT positions x, x + 1 (x = 2 (q + B a)); one 8-byte load where both are read
        // and the address is aligned
        float2 im[H];
#pragma unroll
        for (int h = 0; h < H; ++h) {
          const int a = a0 + h;
          const int x = 2 * (q + C::B * a);
          const bool ok0 = a < C::A && ir != kOuterRow && x >= x0f && x < x1f;
          const bool ok1 = a < C::A && ir != kOuterRow && x + 1 >= x0f && x + 1 < x1f;
          const float* p = q_.img + (ok0 ? ir + x : 0);
          if (ok0 && ok1 && !(reinterpret_cast<uintptr_t>(p) & 7)) {
            im[h] = *reinterpret_cast<const float2*>(p);
          } else {
            im[h].x = ok0 ? *p : 0.f;
            im[h].y = ok1 ? q_.img[ir + x + 1] : 0.f;
          }
        }
#pragma unroll
        for (int h = 0; h < H; ++h) {
          const int a = a0 + h;
          if (a < C::A) {
            z[a].x = im[h].x > 0.f ? im[h].x / z[a].x : 1.f;
            z[a].y = im[h].y > 0.f ? im[h].y / z[a].y : 1.f;
          }
        }
      }
    }
    __syncthreads();  // the inverse's reads of the row are done before the forward writes it
    x2_forward_regs<M>(rb, tw, q, z, zf);
  }
  x2_post_store_regs<M>(zf, q, twr, Sout + (so == kNoRow ? 0 : so),
                        so != kNoRow && !(q_.skipOnes && ir == kOuterRow));
}

// ------------------------------------------------------------------ inverse + psi update (output rows only)
template <int M>
__global__ __launch_bounds__(XCfg<M>::NT) void k_x2_update(Dom d, UpdateParams u, FastDiv divD0,
                                                            const float2* __restrict__ S,
                                                            const float2* __restrict__ twM,
                                                            const float2* __restrict__ wr) {
  using C = XCfg<M>;
  __shared__ float2 buf[C::BUF];
  __shared__ float2 twS[C::TWL ? M : 1];
  __shared__ float2 twrS[C::TWL ? M + 1 : 1];
  const float2* tw = C::TWL ? twS : twM;
  const float2* twr = C::TWL ? twrS : wr;
  __shared__ long long psiRow[C::RW];
  __shared__ size_t rowOff[C::RW];
  if (C::TWL) x2_load_twiddles<M>(twS, twrS, twM, wr);
  // batched blocks: this work-group's block, its output region and spectrum
  int wg = blockIdx.x;
  int o[3] = {u.o[0], u.o[1], u.o[2]}, D[3] = {d.D[0], d.D[1], d.D[2]};
  if (d.btab) {
    const int b = blockIdx.x / d.wgPerBlock;
    wg -= b * d.wgPerBlock;
    const int* t = d.btab + kBlockTabInts * b;
    for (int k = 0; k < 3; ++k) { o[k] = t[k]; D[k] = t[3 + k]; }
    S += (size_t)b * d.blockStride;
    if (D[0] != d.D[0]) divD0 = FastDiv((unsigned)D[0]);  // a tail block's row length
  }
  const int nrows = D[1] * D[2];
  const int row0 = wg * C::RW;
  if (threadIdx.x < C::RW) {
    const int row = row0 + threadIdx.x;
    long long pr = -1;
    size_t so = kNoRow;
    if (row < nrows) {
      const int ly = row % D[1], lz = row / D[1];
      pr = ((long long)(o[2] + lz) * u.N[1] + (o[1] + ly)) * u.N[0] + o[0];
      so = ((size_t)(d.m[2] + lz) * d.P[1] + (d.m[1] + ly)) * d.HxP;
    }
    psiRow[threadIdx.x] = pr;   // >= 0 for output rows (the output region lies inside the volume)
    rowOff[threadIdx.x] = so;
  }
  __syncthreads();
  const int r = threadIdx.x / C::T, q = threadIdx.x % C::T;
  float2* rb = buf + r * C::RS;
  const size_t so = rowOff[r];
  float2 z[C::A];
  x2_inverse_half_regs<M, C::long_rows>(so != kNoRow ? S + so : nullptr, rb, tw, twr, q, z);
  double sum = 0.0;
  float mx = -1.f;
  const RLConsts rc = rl_consts(u.lambda, u.minValue, u.maxI);
  const int DX = D[0];
  if constexpr (C::long_rows) {
    // long rows: lane q < B holds the values z[a] = (x[2n], x[2n + 1]), n = q + B a, of its row and updates psi
    // there (at M = 576 faster than a flat pass)
    const long long pr = psiRow[r];
    if (q < C::B && pr >= 0) {
      const int m0 = d.m[0];
      float* psiRowPtr = u.psi + pr;
      const float* wRowPtr = u.w + pr;
      constexpr int H = (C::A + 3) / 4;  // loads of a quarter of the row's values in flight
#pragma unroll
      for (int a0 = 0; a0 < C::A; a0 += H) {
        float last[2 * H], wt[2 * H];
        int lx[2 * H];
#pragma unroll
        for (int j = 0; j < 2 * H; ++j) {
          const int a = a0 + (j >> 1);
          lx[j] = 2 * (q + C::B * a) + (j & 1) - m0;
          if (a >= C::A || lx[j] < 0 || lx[j] >= DX) lx[j] = -1;
          last[j] = psiRowPtr[max(lx[j], 0)];
          wt[j] = wRowPtr[max(lx[j], 0)];
        }
#pragma unroll
        for (int j = 0; j < 2 * H; ++j) {
          if (lx[j] < 0) continue;
          const int a = min(a0 + (j >> 1), C::A - 1);
          const float next = rl_next_value(last[j], (j & 1) ? z[a].y : z[a].x, wt[j], rc);
          psiRowPtr[lx[j]] = next;
          const float ch = next - last[j];
          sum += (double)ch;
          mx = java_maxf(mx, ch);
        }
      }
    }
  } else {
    // short rows (only B of T lanes hold values): park the rows in LDS and update them with a flat mapping over
    // every thread, six psi / weight pairs in flight (at M = 150 35 % faster than the register update; the
    // row's lanes over positions q + T j with 8 pairs in flight were 8 % slower, a wave walking its rows one by
    // one 12-37 % slower, fewer instructions notwithstanding: the loop is bound by its rounds of loads)
    x2_park_inverse<M>(rb, q, z);
    const float* fb = reinterpret_cast<const float*>(buf);
    const unsigned total = (unsigned)(C::RW * DX);
    constexpr int U = 6;
    if (!(DX & 1) && !(u.N[0] & 1) && !(o[0] & 1)) {
      // even rows in an even-width volume at an even x origin: every pair of positions (2k, 2k + 1) lies in one
      // row and is 8-byte aligned in psi and the weights, so a thread takes U2 such pairs per round (one row
      // lookup and one address per pair, 8-byte loads and stores)
      constexpr int U2 = 4;
      const unsigned total2 = total / 2;
      for (unsigned base = threadIdx.x; base < total2; base += U2 * C::NT) {
        float2 last[U2], wt[U2];
        size_t gi[U2];
        unsigned rr[U2];
        bool ok[U2];
#pragma unroll
        for (int j = 0; j < U2; ++j) {
          const unsigned e = 2 * (base + j * C::NT);  // first position of the pair
          rr[j] = min(divD0.div(e), (unsigned)C::RW - 1);
          const int lx = (int)(e - rr[j] * (unsigned)DX);
          const long long prr = psiRow[rr[j]];
          ok[j] = e < total && prr >= 0;
          gi[j] = ok[j] ? (size_t)(prr + lx) : 0;
          last[j] = *reinterpret_cast<const float2*>(u.psi + gi[j]);
          wt[j] = *reinterpret_cast<const float2*>(u.w + gi[j]);
        }
#pragma unroll
        for (int j = 0; j < U2; ++j) {
          if (!ok[j]) continue;
          const unsigned e = 2 * (base + j * C::NT);
          const int lx = (int)(e - rr[j] * (unsigned)DX);
          const float* f = fb + rr[j] * 2 * C::RS + d.m[0] + lx;
          float2 next;
          next.x = rl_next_value(last[j].x, f[0], wt[j].x, rc);
          next.y = rl_next_value(last[j].y, f[1], wt[j].y, rc);
          *reinterpret_cast<float2*>(u.psi + gi[j]) = next;
          const float c0 = next.x - last[j].x, c1 = next.y - last[j].y;
          sum += (double)c0;
          mx = java_maxf(mx, c0);
          sum += (double)c1;
          mx = java_maxf(mx, c1);
        }
      }
    } else {
      for (unsigned base = threadIdx.x; base < total; base += U * C::NT) {
        float last[U], wt[U];
        size_t gi[U];
        unsigned rr[U];
        bool ok[U];
#pragma unroll
        for (int j = 0; j < U; ++j) {
          const unsigned idx = base + j * C::NT;
          rr[j] = min(divD0.div(idx), (unsigned)C::RW - 1);
          const int lx = (int)(idx - rr[j] * (unsigned)DX);
          const long long prr = psiRow[rr[j]];
          ok[j] = idx < total && prr >= 0;
          gi[j] = ok[j] ? (size_t)(prr + lx) : 0;
          last[j] = u.psi[gi[j]];
          wt[j] = u.w[gi[j]];
        }
#pragma unroll
        for (int j = 0; j < U; ++j) {
          if (!ok[j]) continue;
          const unsigned idx = base + j * C::NT;
          const int lx = (int)(idx - rr[j] * (unsigned)DX);
          const float next = rl_next_value(last[j], fb[rr[j] * 2 * C::RS + d.m[0] + lx], wt[j], rc);
          u.psi[gi[j]] = next;
          const float ch = next - last[j];
          sum += (double)ch;
          mx = java_maxf(mx, ch);
        }
      }
    }
  }
  // the stats reduction reuses the row buffer once every thread has read its rows
  static_assert(C::BUF * sizeof(float2) >= 2 * reduce_width<C::NT>() * sizeof(double),
                "the row buffer holds the reduction");
  __syncthreads();
  double* red = reinterpret_cast<double*>(buf);
  block_reduce_stats_nt<C::NT>(sum, mx, u.partial + blockIdx.x, red, red + reduce_width<C::NT>());
}

// ------------------------------------------------------------------ inverse + store of the output region
template <int M>
__global__ __launch_bounds__(XCfg<M>::NT) void k_x2_store(Dom d, StoreParams st, const float2* __restrict__ S,
                                                           const float2* __restrict__ twM,
                                                           const float2* __restrict__ wr) {
  using C = XCfg<M>;
  __shared__ float2 buf[C::BUF];
  __shared__ float2 twS[C::TWL ? M : 1];
  __shared__ float2 twrS[C::TWL ? M + 1 : 1];
  const float2* tw = C::TWL ? twS : twM;
  const float2* twr = C::TWL ? twrS : wr;
  __shared__ size_t rowOff[C::RW];
  if (C::TWL) x2_load_twiddles<M>(twS, twrS, twM, wr);
  const int nrows = d.D[1] * d.D[2];
  const int row0 = blockIdx.x * C::RW;
  if (threadIdx.x < C::RW) {
    const int row = row0 + threadIdx.x;
    size_t so = kNoRow;
    if (row < nrows) {
      const int ly = row % d.D[1], lz = row / d.D[1];
      so = ((size_t)(d.m[2] + lz) * d.P[1] + (d.m[1] + ly)) * d.HxP;
    }
    rowOff[threadIdx.x] = so;
  }
  __syncthreads();
  const int r = threadIdx.x / C::T, q = threadIdx.x % C::T;
  float2* rb = buf + r * C::RS;
  const size_t so = rowOff[r];
  float2 z[C::A];
  x2_inverse_half_regs<M>(so != kNoRow ? S + so : nullptr, rb, tw, twr, q, z);
  const int row = row0 + r;
  if (q < C::B && row < nrows) {
    const int DX = d.D[0], m0 = d.m[0];
    float* out = st.out + (size_t)row * DX;
#pragma unroll
    for (int a = 0; a < C::A; ++a) {
      const int lx = 2 * (q + C::B * a) - m0;
      if (lx >= 0 && lx < DX) out[lx] = z[a].x;
      if (lx + 1 >= 0 && lx + 1 < DX) out[lx + 1] = z[a].y;
    }
  }
}

// ------------------------------------------------------------------ host launchers
// Batched launches (d.btab set) run d.nb blocks of wgPerBlock work-groups each: work-group w serves block
// w / wgPerBlock; rows of a block beyond its own extent are skipped inside the kernels.
template <int M>
void launch_x_fwd(const Dom& d, const GatherParams& g, float2* S, const float2* twM, const float2* wr,
                  hipStream_t s) {
  static_assert(x2_ok<M>(), "every compiled x length has a two-step split");
  using C = XCfg<M>;
  Dom dd = d;
  const int rows = g.uniqueRows && g.mode == GATHER_MIRROR ? g.N[1] * g.N[2] : d.P[1] * d.P[2];
  dd.wgPerBlock = (rows + C::RW - 1) / C::RW;
  const int nb = d.btab ? d.nb : 1;
  hipLaunchKernelGGL(k_x2_fwd<M>, dim3(dd.wgPerBlock * nb), dim3(C::NT), 0, s, dd, g, S, twM, wr);
}
template <int M>
void launch_x_ratio(const Dom& d, const RatioParams& q, const float2* Sin, float2* Sout, const float2* twM,
                    const float2* wr, hipStream_t s) {
  static_assert(x2_ok<M>(), "every compiled x length has a two-step split");
  using C = XCfg<M>;
  Dom dd = d;
  const int rows = q.skipOnes ? q.rowN[1] * q.rowN[2] : d.P[1] * d.P[2];
  if (rows == 0) return;  // no quotient rows: every row is a row of ones, put in by the z pass
  dd.wgPerBlock = (rows + C::RW - 1) / C::RW;
  const int nb = d.btab ? d.nb : 1;
  hipLaunchKernelGGL(k_x2_ratio<M>, dim3(dd.wgPerBlock * nb), dim3(C::NT), 0, s, dd, q, Sin, Sout, twM, wr);
}
template <int M>
int launch_x_update(const Dom& d, const UpdateParams& u, const float2* S, const float2* twM, const float2* wr,
                    hipStream_t s) {
  static_assert(x2_ok<M>(), "every compiled x length has a two-step split");
  using C = XCfg<M>;
  Dom dd = d;
  const int rows = d.D[1] * d.D[2];
  dd.wgPerBlock = (rows + C::RW - 1) / C::RW;
  const int nwg = dd.wgPerBlock * (d.btab ? d.nb : 1);
  hipLaunchKernelGGL(k_x2_update<M>, dim3(nwg), dim3(C::NT), 0, s, dd, u, FastDiv((unsigned)d.D[0]), S, twM, wr);
  return nwg;
}
template <int M>
void launch_x_store(const Dom& d, const StoreParams& st, const float2* S, const float2* twM, const float2* wr,
                    hipStream_t s) {
  static_assert(x2_ok<M>(), "every compiled x length has a two-step split");
  using C = XCfg<M>;
  const int rows = d.D[1] * d.D[2];
  hipLaunchKernelGGL(k_x2_store<M>, dim3((rows + C::RW - 1) / C::RW), dim3(C::NT), 0, s, d, st, S, twM, wr);
}

}  // namespace mvr
