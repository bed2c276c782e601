// passes.hpp -- the HBM passes of one Richardson-Lucy view update.
//
// One view of one iteration (Seq scheme, ComputeBlockSeqThreadCPU.runIteration,
// D/iteration/sequential/ComputeBlockSeqThreadCPU.java:79-169) is nine launches:
//   x_fwd(gather psi, mirror)  z_fwd  y(*K1^)  z_inv  x_ratio(C2R, img/blurred, R2C)
//   z_fwd  y(*K2^)  z_inv  x_update(C2R, Tikhonov update of psi, stats)
// Every pass reads and writes its data once; the pointwise work of the
// reference (quotient, update, spectrum multiply) is fused into the passes.
// Each pass moves only the rows its neighbours use: the first convolution
// transforms psi's own rows (its mirror copies are read from their sources),
// the second does not store or read the quotient's rows and y planes of ones
// (their spectra are known and put in by the pass that reads them), and the
// inverse passes produce only the rows the next x pass reads (the StridedGeom
// and RatioParams fields the plan sets per launch: Plan::restrict_rows,
// skip_ones_rows, mirror_rows).
//
// x passes work on whole rows (real axis, half-length complex FFT with the
// even/odd split); y and z passes work on tiles of W consecutive spectrum
// columns so that every global access is a run of W*8 contiguous bytes.
#pragma once
#include <hip/hip_ext.h>

#include "engine.hpp"
#include "fft_reg.hpp"

namespace mvr {

constexpr int kThreads = 256;

__device__ __forceinline__ int mirror_index(int p, int n) {
  if (n == 1) return 0;
  const int period = 2 * n - 2;
  p %= period;
  if (p < 0) p += period;
  return p >= n ? period - p : p;
}

__device__ __forceinline__ float gather_value(const GatherParams& g, const Dom& d, int x, int y, int z) {
  if (g.mode == GATHER_MIRROR) {
    const int gx = mirror_index(g.org[0] + x, g.N[0]);
    const int gy = mirror_index(g.org[1] + y, g.N[1]);
    const int gz = mirror_index(g.org[2] + z, g.N[2]);
    return g.src[((size_t)gz * g.N[1] + gy) * g.N[0] + gx];
  } else if (g.mode == GATHER_VALUE) {
    const int gx = g.org[0] + x, gy = g.org[1] + y, gz = g.org[2] + z;
    if (gx < 0 || gy < 0 || gz < 0 || gx >= g.N[0] || gy >= g.N[1] || gz >= g.N[2]) return g.value;
    return g.src[((size_t)gz * g.N[1] + gy) * g.N[0] + gx];
  } else {
    const int jx = (x + g.N[0] / 2) % d.P[0];
    const int jy = (y + g.N[1] / 2) % d.P[1];
    const int jz = (z + g.N[2] / 2) % d.P[2];
    if (jx >= g.N[0] || jy >= g.N[1] || jz >= g.N[2]) return 0.f;
    return g.src[((size_t)jz * g.N[1] + jy) * g.N[0] + jx];
  }
}

// ------------------------------------------------------------------ update of psi
// DeconvolutionMethods.computeNextValue (D/iteration/sequential/DeconvolutionMethods.java:320-358),
// float operations in Java order; Tikhonov in double.  Java rounds every float and double operation, so the
// function bodies forbid FMA contraction (also where they are inlined), and every result below is the one the
// corresponding Java operation rounds to:
//   * value / maxI (float): the double product value * RN(1/maxI) is within 2^-52 (relative) of the quotient,
//     and a quotient of two floats is never that close to a float rounding midpoint without being on it (which
//     it cannot be), so rounding the product to float gives the correctly rounded float quotient;
//   * sqrt(y), y >= 1: Markstein's iteration from the hardware reciprocal square root with two residual
//     corrections, the correctly rounded double square root (the sequence of the device library's sqrt, whose
//     operand scaling is an identity for y >= 1; y = inf, from an overflowing value / maxI, is the one special
//     value it can meet, and is passed through as the library does);
//   * a / lambda: the quotient estimate a * R and one residual correction, R the reciprocal of lambda refined by
//     two Newton steps (the device division's sequence, whose operand scaling is an identity for the normal a
//     and lambda here).
// The per-launch constants come from rl_consts.
struct RLConsts {
  double lambda, twoLambda, rLambda, rMax;
  float minI, maxI;
};

__device__ __forceinline__ RLConsts rl_consts(double lambda, float minI, float maxI) {
  RLConsts c;
  c.lambda = lambda;
  c.twoLambda = 2.0 * lambda;
  c.minI = minI;
  c.maxI = maxI;
  c.rMax = 1.0 / (double)maxI;
  double y = __builtin_amdgcn_rcp(lambda);
  y = __builtin_fma(y, __builtin_fma(-lambda, y, 1.0), y);
  y = __builtin_fma(y, __builtin_fma(-lambda, y, 1.0), y);
  c.rLambda = y;
  return c;
}

__device__ __forceinline__ float rl_tikhonov_adjust(float value, const RLConsts& c) {
#pragma clang fp contract(off)
  float adjusted;
  if (value > 0.f) {
    if (c.lambda > 0.0) {
      const double vn = (double)(float)((double)value * c.rMax);  // (double)(value / maxI)
      const double y = 1.0 + c.twoLambda * vn;                     // >= 1
      const double r = __builtin_amdgcn_rsq(y);
      double g = y * r, h = 0.5 * r;
      const double e = __builtin_fma(-h, g, 0.5);
      g = __builtin_fma(g, e, g);
      h = __builtin_fma(h, e, h);
      g = __builtin_fma(__builtin_fma(-g, g, y), h, g);
      g = __builtin_fma(__builtin_fma(-g, g, y), h, g);  // sqrt(y)
      const double a = g - 1.0;
      const double q0 = a * c.rLambda;
      double t = __builtin_fma(__builtin_fma(-c.lambda, q0, a), c.rLambda, q0);  // a / lambda
      if (y == __builtin_inf()) t = y;  // value / maxI overflowed: sqrt and quotient are infinite
      adjusted = (float)t * c.maxI;
    } else {
      adjusted = value;
    }
  } else {
    adjusted = c.minI;
  }
  return adjusted > c.minI ? adjusted : c.minI;  // Math.max(minI, adjusted), NaN -> minI
}

__device__ __forceinline__ float rl_next_value(float last, float integral, float weight, const RLConsts& c) {
#pragma clang fp contract(off)
  const float next = rl_tikhonov_adjust(last * integral, c);
  return last + (next - last) * weight;
}

// Math.max of Java on doubles for the maxChange statistic: a NaN change stays NaN (fmax would drop it)
__device__ __forceinline__ double java_max(double a, double b) { return a != a ? a : b != b ? b : fmax(a, b); }
// the same over float changes (their maximum is the same number in float; a NaN stays; of equal zeros the
// first is kept)
__device__ __forceinline__ float java_maxf(float m, float ch) { return (ch > m || ch != ch) ? ch : m; }

__device__ __forceinline__ void block_reduce_stats(double s, double mx, double2* out) {
  __shared__ double rs[kThreads / 64], rm[kThreads / 64];
  for (int off = 32; off > 0; off >>= 1) {
    s += __shfl_down(s, off);
    mx = java_max(mx, __shfl_down(mx, off));
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) { rs[wv] = s; rm[wv] = mx; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0.0, b = -1.0;
    for (int i = 0; i < kThreads / 64; ++i) { a += rs[i]; b = java_max(b, rm[i]); }
    *out = make_double2(a, b);
  }
}

// ------------------------------------------------------------------ y / z passes, two-step register FFT
// Pencil length L = A x B.  Forward: thread (w, b) loads x[b + B a] (a < A) straight
// from HBM, does the A-point DFT in registers, multiplies by W_L^{b k1} and parks
// the values in LDS; thread (w, k1) then reads its B values and does the
// B-point DFT, which yields X[k1 + A k2].  The inverse runs the same two steps
// in the opposite order, so in the convolution pass the spectrum product and
// the first inverse step stay in the registers of the forward's second step:
// one LDS round trip per transform.
template <int L, int W>
__host__ __device__ constexpr bool s2_fits();
template <int L>
__host__ __device__ constexpr int s2_width() {
  // 16 columns = 128-byte runs (32 was slower at L = 300, 8 slower at L = 400); 8 for 720..1024 (two work-groups
  // per CU); 16 again for 1080 and 1152, whose 16-column tile fills the LDS of one CU and still runs 12 % faster
  // than two 8-column work-groups (y pass of 1152-long pencils: 3.24 -> 2.86 ms per 4-iteration run)
  return L <= 640 ? 16 : (L > 1024 && s2_fits<L, 16>()) ? 16 : 8;
}
constexpr bool kStridedXcdOrder = true;
// tile widths compiled per length: 8 and 16 columns where the tile fits 96 KB of LDS; above L = 1024 also
// 16-column tiles of up to 152 KB (one work-group per CU, twiddles in LDS beside the tile)
template <int L, int W>
__host__ __device__ constexpr bool s2_fits() {
  return split2(L).A > 0 && L * W * 8 <= (L > 1024 ? 152 : 96) * 1024;
}

// The exchange between the two register steps goes through LDS in R rounds: round r carries the values of
// one part of the second step's threads (a half for R = 2), so the tile holds about L / R values per column
// and more work-groups share a CU's 160 KB of LDS, for two more barriers per exchange.  The plain forward and
// inverse passes (the z passes) use two rounds up to L = 1024: at L = 300 z_fwd goes from 3 to 6 work-groups
// per CU (2 % faster), at L = 576 from 2 to 3 (z_fwd and z_inv 5-16 % faster).  The convolution pass keeps one
// (its kernel spectrum values, requested before the exchange, would stay live across the rounds and take back
// the occupancy in registers), and so do the longest pencils, whose registers hold one work-group per CU
// either way; their code is the one-round code.
template <int L, int W, int MODE>
__host__ __device__ constexpr int s2_rounds() {
  return MODE != 2 && L <= 1024 ? 2 : 1;
}
// complex values per column the exchange tile holds
template <int L, int W, int MODE>
__host__ __device__ constexpr int s2_tile_rows() {
  constexpr Split2 S = split2(L);
  constexpr int R = s2_rounds<L, W, MODE>();
  constexpr int fa = (S.A + R - 1) / R * S.B, fb = S.A * ((S.B + R - 1) / R);
  return R == 1 ? L : (fa > fb ? fa : fb);
}

// The twiddles stay in global memory (read through L1) when that lets a second work-group share the CU's
// 160 KB of LDS; such kernels are then compiled for the waves per SIMD that two work-groups need.
template <int L, int W, int MODE>
__host__ __device__ constexpr bool s2_global_tw() {
  constexpr int T = s2_tile_rows<L, W, MODE>() * W * 8;
  return 2 * (T + L * 8) > 160 * 1024 && 2 * T <= 160 * 1024;
}
template <int L, int W, int MODE>
__host__ __device__ constexpr int s2_min_waves() {
  return s2_global_tw<L, W, MODE>() ? (2 * ((W * split2(L).A + 63) / 64) + 3) / 4 : 1;
}

// Inverse exchange in R >= 2 rounds: thread q = k1 < A holds u[n2] for n2 < B, to be multiplied by
// conj(W_L^{n2 k1}); afterwards thread q = n2 < B holds v[k1] = the twiddled u[n2] of thread k1, for k1 < A.
// Round r carries n2 in [r BH, r BH + BH).  Starts with a barrier: every thread is done with the tile before it.
template <int A, int B, int W, int R>
__device__ __forceinline__ void s2_inverse_exchange(float2* tile, float2* u, float2* v, int q, int w,
                                                    const float2* tw) {
  {
    constexpr int BH = (B + R - 1) / R;
#pragma unroll
    for (int n2 = 1; n2 < B; ++n2) u[n2] = cmul(u[n2], cconj(tw[n2 * q]));
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int n0 = r * BH, nn = (r + 1) * BH < B ? BH : B - r * BH;
      __syncthreads();
#pragma unroll
      for (int j = 0; j < BH; ++j)
        if (j < nn) tile[(q + A * j) * W + w] = u[n0 + j];
      __syncthreads();
      if (q >= n0 && q < n0 + nn) {
#pragma unroll
        for (int k1 = 0; k1 < A; ++k1) v[k1] = tile[(k1 + A * (q - n0)) * W + w];
      }
    }
  }
}

// Global accesses of a tile: a base that is uniform over the work-group (the tile's first column on its plane,
// plus a multiple of the pencil stride, formed in scalar registers) and the lane's 32-bit byte offset (its column
// and pencil position q).  They go through a raw buffer resource on the uniform base (4 GiB range, gfx9 format
// dword 0x00020000), so that no per-access 64-bit address is formed in vector registers: the memory
// instruction adds the lane's offset itself.
constexpr int kBufferFormatGfx9 = 0x00020000;
__device__ __forceinline__ float2 s2_ld(const float2* base, unsigned off) {
  const __amdgpu_buffer_rsrc_t r =
      __builtin_amdgcn_make_buffer_rsrc(const_cast<float2*>(base), 0, 0xffffffff, kBufferFormatGfx9);
  typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
  const u32x2 v = __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0);
  float2 x;
  __builtin_memcpy(&x, &v, sizeof(x));
  return x;
}
__device__ __forceinline__ void s2_st(float2* base, unsigned off, float2 x) {
  const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc(base, 0, 0xffffffff, kBufferFormatGfx9);
  typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
  u32x2 v;
  __builtin_memcpy(&v, &x, sizeof(v));
  __builtin_amdgcn_raw_buffer_store_b64(v, r, off, 0, 0);
}
// the largest pencil position q a lane's offset covers, plus one (the host checks that the offsets fit 32 bits)
template <int L>
int strided_span() {
  constexpr Split2 S = split2(L);
  return (S.A > S.B ? S.A : S.B) + 1;
}

template <int L, int W, int MODE>
__global__ __launch_bounds__(W * split2(L).A, (s2_min_waves<L, W, MODE>())) void k_strided2(
    StridedGeom gm, const float2* __restrict__ in, float2* __restrict__ out, const float2* __restrict__ spec,
    float scale, const float2* __restrict__ twg) {
  constexpr Split2 S = split2(L);
  constexpr int A = S.A, B = S.B;
  constexpr int NT = W * A;
  constexpr bool TWG = s2_global_tw<L, W, MODE>();
  constexpr int R = s2_rounds<L, W, MODE>();
  constexpr int AH = (A + R - 1) / R, BH = (B + R - 1) / R;  // second-step threads (k1, resp. n2) per round
  __shared__ float2 tile[s2_tile_rows<L, W, MODE>() * W];
  __shared__ float2 twS[TWG ? 1 : L];
  const float2* tw = TWG ? twg : twS;
  if (!TWG)
    for (int i = threadIdx.x; i < L; i += NT) twS[i] = twg[i];
  const int w = threadIdx.x % W, q = threadIdx.x / W;
  // Work-groups are dealt to the 8 XCDs in turn; the tile order is remapped so that each XCD sweeps a
  // contiguous range of (column tile, plane) tiles: neighbouring column tiles, which share cache lines
  // when a tile row is shorter than a line, meet in one L2.
  unsigned tid = blockIdx.y * gridDim.x + blockIdx.x;
  const unsigned ntile = gridDim.x * gridDim.y;
  if (kStridedXcdOrder && ntile % 8 == 0) tid = (tid % 8) * (ntile / 8) + tid / 8;
  int plane = (int)(tid / gridDim.x);
  const int col0 = (int)(tid % gridDim.x) * W;
  const bool colOk = w < gm.ncols - col0;
  size_t blockOff = 0;  // batched blocks: the block's data offset; `plane` becomes the plane within the block
  if (gm.blockPlanes) {
    const int b = plane / gm.blockPlanes;
    plane -= b * gm.blockPlanes;
    blockOff = (size_t)b * gm.blockStride;
  }
  plane += gm.plane0;
  // Lanes of columns beyond ncols (the last tile) read the row pitch's padding (pitch = 16-column multiple) and
  // compute values that are never stored.
  const size_t ls = gm.lineStride;
  const float2* sb = in + blockOff + (size_t)plane * gm.planeStride + col0;
  float2* db = out + blockOff + (size_t)plane * gm.planeStride + col0;
  const unsigned lo = (unsigned)((w + (size_t)q * ls) * sizeof(float2));
  float2 v[A];
  if constexpr (MODE != 1) {
    // ---- forward step 1: thread q = b < B
    if (q < B) {
      if (gm.onesDC != 0.f) {
        // rows of ones outside the written region: their spectrum is known and not read
        const bool planeIn = plane >= gm.onesPlaneLo && plane < gm.onesPlaneHi;
        const bool dc = col0 + w == 0 && (gm.onesDCPlane < 0 || plane == gm.onesDCPlane);
        const float2 one = make_float2(dc ? gm.onesDC : 0.f, 0.f);
#pragma unroll
        for (int a = 0; a < A; ++a) {
          const int n = q + B * a;
          const bool rd = planeIn && n >= gm.onesLo && n < gm.onesHi;
          v[a] = rd ? s2_ld(sb + (size_t)(B * a) * ls, lo) : one;
        }
      } else if (gm.mirN > 0) {
        // a mirror extension along the pencil: the copies are read from their sources.  One reflection at
        // each face suffices when the pencil reaches less than mirN - 1 beyond either face (the usual case).
        // The positions differ per lane: 32-bit byte offsets from the plane (the host enables this mode only
        // where the pencil's span fits them).
        const unsigned w8 = (unsigned)(w * sizeof(float2)), ls8 = (unsigned)(ls * sizeof(float2));
        if (gm.mirM < gm.mirN && L - gm.mirM <= 2 * gm.mirN - 1) {
#pragma unroll
          for (int a = 0; a < A; ++a) {
            int n = q + B * a - gm.mirM;
            n = n < 0 ? -n : n;
            n = n >= gm.mirN ? 2 * gm.mirN - 2 - n : n;
            v[a] = s2_ld(sb, w8 + (unsigned)(gm.mirM + n) * ls8);
          }
        } else {
#pragma unroll
          for (int a = 0; a < A; ++a) {
            const int n = mirror_index(q + B * a - gm.mirM, gm.mirN);
            v[a] = s2_ld(sb, w8 + (unsigned)(gm.mirM + n) * ls8);
          }
        }
      } else {
#pragma unroll
        for (int a = 0; a < A; ++a) v[a] = s2_ld(sb + (size_t)(B * a) * ls, lo);
      }
    }
    __syncthreads();  // twiddle table
    if (q < B) {
      rdft<A, -1>(v);
      if constexpr (R == 1) {
#pragma unroll
        for (int k1 = 0; k1 < A; ++k1) {
          const float2 t = k1 == 0 ? v[k1] : cmul(v[k1], tw[q * k1]);
          tile[(q + B * k1) * W + w] = t;
        }
      } else {
#pragma unroll
        for (int k1 = 1; k1 < A; ++k1) v[k1] = cmul(v[k1], tw[q * k1]);
      }
    }
    // convolution: the kernel spectrum values of this thread's outputs are requested before the exchange,
    // so that their latency overlaps it
    float2 kv[MODE == 2 ? B : 1];
    if constexpr (MODE == 2) {
      const float2* spb = spec + (size_t)plane * gm.planeStride + col0;
#pragma unroll
      for (int k2 = 0; k2 < B; ++k2) kv[k2] = s2_ld(spb + (size_t)(A * k2) * ls, lo);
    }
    // ---- exchange (round r: the values of k1 in [r AH, r AH + AH)), then forward step 2: thread q = k1 < A
    float2 u[B];
    if constexpr (R == 1) {
      __syncthreads();
#pragma unroll
      for (int b = 0; b < B; ++b) u[b] = tile[(b + B * q) * W + w];
    } else {
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int k0 = r * AH, kn = (r + 1) * AH < A ? AH : A - r * AH;
        if (r > 0) __syncthreads();  // the previous round's readers are done
        if (q < B) {
#pragma unroll
          for (int j = 0; j < AH; ++j)
            if (j < kn) tile[(q + B * j) * W + w] = v[k0 + j];
        }
        __syncthreads();
        if (q >= k0 && q < k0 + kn) {
#pragma unroll
          for (int b = 0; b < B; ++b) u[b] = tile[(b + B * (q - k0)) * W + w];
        }
      }
    }
    rdft<B, -1>(u);  // u[k2] = X[q + A k2]
    if constexpr (MODE == 0 || MODE == 3) {
      if (colOk) {
#pragma unroll
        for (int k2 = 0; k2 < B; ++k2) {
          float2 x = u[k2];
          if constexpr (MODE == 3) x = cscale(x, scale);
          s2_st(db + (size_t)(A * k2) * ls, lo, x);
        }
      }
      return;
    } else {
      // spectrum product, then inverse step 1 over k2 in registers
#pragma unroll
      for (int k2 = 0; k2 < B; ++k2) u[k2] = cmul(u[k2], kv[k2]);
      rdft<B, 1>(u);  // u[n2] = sum_k2 X[q + A k2] W_B^{-n2 k2}
      if constexpr (R == 1) {
        __syncthreads();  // every thread has read its forward values from the tile
#pragma unroll
        for (int n2 = 0; n2 < B; ++n2) {
          const float2 t = n2 == 0 ? u[n2] : cmul(u[n2], cconj(tw[n2 * q]));
          tile[(q + A * n2) * W + w] = t;
        }
      } else {
        s2_inverse_exchange<A, B, W, R>(tile, u, v, q, w, tw);
      }
    }
  } else {
    // ---- inverse only: thread q = k1 < A holds X[q + A k2]
    float2 u[B];
#pragma unroll
    for (int k2 = 0; k2 < B; ++k2) u[k2] = s2_ld(sb + (size_t)(A * k2) * ls, lo);
    __syncthreads();  // twiddle table
    rdft<B, 1>(u);
    if constexpr (R == 1) {
#pragma unroll
      for (int n2 = 0; n2 < B; ++n2) {
        const float2 t = n2 == 0 ? u[n2] : cmul(u[n2], cconj(tw[n2 * q]));
        tile[(q + A * n2) * W + w] = t;
      }
    } else {
      s2_inverse_exchange<A, B, W, R>(tile, u, v, q, w, tw);
    }
  }
  if constexpr (MODE == 1 || MODE == 2) {
    if constexpr (R == 1) __syncthreads();
    // ---- inverse step 2: thread q = n2 < B does the A-point DFT over k1 -> x[n2 + B n1]
    if (q < B) {
      if constexpr (R == 1) {
#pragma unroll
        for (int k1 = 0; k1 < A; ++k1) v[k1] = tile[(k1 + A * q) * W + w];
      }
      rdft<A, 1>(v);
      if (colOk) {
        // positions the next pass reads (storeLo <= q + B n1 < storeHi)
#pragma unroll
        for (int n1 = 0; n1 < A; ++n1) {
          const int n = q + B * n1;
          if (n >= gm.storeLo && n < gm.storeHi) s2_st(db + (size_t)(B * n1) * ls, lo, v[n1]);
        }
      }
    }
  }
}

// ------------------------------------------------------------------ host launchers
template <int L, int W, int MODE>
void launch_strided2_mode(const StridedGeom& gm, const float2* in, float2* out, const float2* spec, float scale,
                          const float2* tw, hipStream_t s) {
  constexpr int NT = W * split2(L).A;
  const dim3 g2((gm.ncols + W - 1) / W, gm.nplanes);
  if (gm.ncols <= 0 || gm.nplanes <= 0) {  // nothing to transform (an empty row range)
    if (gm.evStart) {
      (void)hipEventRecord(gm.evStart, s);
      (void)hipEventRecord(gm.evStop, s);
    }
    return;
  }
  if (gm.evStart)  // timed: the events hold the kernel's own start and stop
    hipExtLaunchKernelGGL((k_strided2<L, W, MODE>), g2, dim3(NT), 0, s, gm.evStart, gm.evStop, 0u, gm, in, out, spec,
                          scale, tw);
  else
    hipLaunchKernelGGL((k_strided2<L, W, MODE>), g2, dim3(NT), 0, s, gm, in, out, spec, scale, tw);
}

template <int L, int W>
void launch_strided2(const StridedGeom& gm, int mode, const float2* in, float2* out, const float2* spec, float scale,
                     const float2* tw, hipStream_t s) {
  switch (mode) {
    case 0: return launch_strided2_mode<L, W, 0>(gm, in, out, spec, scale, tw, s);
    case 1: return launch_strided2_mode<L, W, 1>(gm, in, out, spec, scale, tw, s);
    case 2: return launch_strided2_mode<L, W, 2>(gm, in, out, spec, scale, tw, s);
    default: return launch_strided2_mode<L, W, 3>(gm, in, out, spec, scale, tw, s);
  }
}

// gm.width picks the tile width (0 or a width not compiled for L: s2_width<L>())
template <int L>
void launch_strided(const StridedGeom& gm, int mode, const float2* in, float2* out, const float2* spec, float scale,
                    const float2* tw, hipStream_t s) {
  static_assert(s2_fits<L, s2_width<L>()>(), "every compiled length has a two-step split whose tile fits LDS");
  if constexpr (s2_fits<L, 16>())
    if (gm.width == 16) return launch_strided2<L, 16>(gm, mode, in, out, spec, scale, tw, s);
  if constexpr (s2_fits<L, 8>())
    if (gm.width == 8) return launch_strided2<L, 8>(gm, mode, in, out, spec, scale, tw, s);
  return launch_strided2<L, s2_width<L>()>(gm, mode, in, out, spec, scale, tw, s);
}

}  // namespace mvr
