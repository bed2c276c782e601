// passes.hpp -- the HBM passes of one Richardson-Lucy view update.
//
// One view of one iteration (Seq scheme, ComputeBlockSeqThreadCPU.runIteration,
// D/iteration/sequential/ComputeBlockSeqThreadCPU.java:79-169) is nine launches:
//   x_fwd(gather psi, mirror)  y_fwd  z(*K1^)  y_inv  x_ratio(C2R, img/blurred, R2C)
//   y_fwd  z(*K2^)  y_inv  x_update(C2R, Tikhonov update of psi, stats)
// Every pass reads and writes its data once; the pointwise work of the
// reference (quotient, update, spectrum multiply) is fused into the passes.
//
// x passes work on whole rows (real axis, half-length complex FFT with the
// even/odd split); y and z passes work on tiles of W consecutive spectrum
// columns so that every global access is a run of W*8 contiguous bytes.
#pragma once
#include "engine.hpp"
#include "fft_core.hpp"
#include "fft_reg.hpp"

namespace mvr {

constexpr int kThreads = 256;

// rows per work-group of an x pass: about four butterflies per thread and stage
__host__ __device__ constexpr int x_threads_per_row(int M) {
  return M >= 256 ? 64 : M >= 128 ? 32 : M >= 48 ? 16 : M >= 16 ? 8 : 4;
}
__host__ __device__ constexpr int strided_width(int L) { return L > 640 ? 8 : 16; }

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

// R2C post-processing of row r: X[k] = E[k] + W^k O[k], k = 0..M, from the
// length-M transform Z of z[m] = x[2m] + i x[2m+1]; written straight to HBM.
template <int M, int TP, int RS>
__device__ __forceinline__ void r2c_post_store(const float2* rowb, const float2* twr, int tp, bool valid,
                                               float2* __restrict__ dst) {
#pragma unroll
  for (int k0 = 0; k0 <= M; k0 += TP) {
    const int k = k0 + tp;
    if (k <= M) {
      const float2 zk = rowb[k == M ? 0 : k];
      const float2 zc = cconj(rowb[k == 0 ? 0 : M - k]);
      const float2 e = make_float2(0.5f * (zk.x + zc.x), 0.5f * (zk.y + zc.y));
      const float2 dd = csub(zk, zc);
      const float2 od = make_float2(0.5f * dd.y, -0.5f * dd.x);
      const float2 X = cadd(e, cmul(twr[k], od));
      if (valid) dst[k] = X;
    }
  }
}

// C2R pre-processing of row r in place: Z[k] = E[k] + i O[k], k = 0..M-1.
template <int M, int TP>
__device__ __forceinline__ void c2r_pre(float2* rowb, const float2* twr, int tp) {
  constexpr int IT = (M + TP - 1) / TP;
  float2 z[IT];
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const int k = tp + it * TP;
    if (k < M) {
      const float2 a = rowb[k];
      const float2 b = cconj(rowb[M - k]);
      const float2 e = make_float2(0.5f * (a.x + b.x), 0.5f * (a.y + b.y));
      const float2 o = cmul(make_float2(0.5f * (a.x - b.x), 0.5f * (a.y - b.y)), cconj(twr[k]));
      z[it] = make_float2(e.x - o.y, e.y + o.x);
    }
  }
  __syncthreads();
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const int k = tp + it * TP;
    if (k < M) rowb[k] = z[it];
  }
  __syncthreads();
}

template <int M>
__device__ __forceinline__ void load_twiddles(float2* tw, float2* twr, const float2* twM, const float2* wr) {
  for (int i = threadIdx.x; i < M; i += kThreads) tw[i] = twM[i];
  for (int i = threadIdx.x; i <= M; i += kThreads) twr[i] = wr[i];
}

// ------------------------------------------------------------------ x forward with gather
template <int M>
__global__ __launch_bounds__(kThreads) void k_x_fwd(Dom d, GatherParams g, float2* __restrict__ S,
                                                    const float2* __restrict__ twM, const float2* __restrict__ wr) {
  constexpr int TP = x_threads_per_row(M);
  constexpr int RW = kThreads / TP;
  constexpr int PX = 2 * M;
  constexpr int RS = M + 1;
  __shared__ float2 buf[RW * RS];
  __shared__ float2 tw[M];
  __shared__ float2 twr[M + 1];
  load_twiddles<M>(tw, twr, twM, wr);
  const int nrows = d.P[1] * d.P[2];
  const int row0 = blockIdx.x * RW;
  float* fb = reinterpret_cast<float*>(buf);
  for (int idx = threadIdx.x; idx < RW * PX; idx += kThreads) {
    const int r = idx / PX, x = idx - r * PX;
    const int row = row0 + r;
    float v = 0.f;
    if (row < nrows) v = gather_value(g, d, x, row % d.P[1], row / d.P[1]);
    fb[r * 2 * RS + x] = v;
  }
  __syncthreads();
  const int r = threadIdx.x / TP, tp = threadIdx.x % TP;
  fft_lds<M, TP, 1, RS, -1>(buf, tw, r, tp);
  const int row = row0 + r;
  r2c_post_store<M, TP, RS>(buf + r * RS, twr, tp, row < nrows, S + (size_t)row * d.HxP);
}

// ------------------------------------------------------------------ x inverse + quotient + x forward
// ratio = img > 0 ? img / blurred : 1 on the output region, 1 elsewhere
// (DeconvolutionMethods.computeQuotient, D/iteration/sequential/DeconvolutionMethods.java:46-98;
// ratio outside the block is extendValue(1), ComputeBlockSeqThreadCPU.java:191-209).
template <int M>
__global__ __launch_bounds__(kThreads) void k_x_ratio(Dom d, RatioParams q, const float2* __restrict__ Sin,
                                                      float2* __restrict__ Sout, const float2* __restrict__ twM,
                                                      const float2* __restrict__ wr) {
  constexpr int TP = x_threads_per_row(M);
  constexpr int RW = kThreads / TP;
  constexpr int PX = 2 * M;
  constexpr int RS = M + 1;
  __shared__ float2 buf[RW * RS];
  __shared__ float2 tw[M];
  __shared__ float2 twr[M + 1];
  load_twiddles<M>(tw, twr, twM, wr);
  const int nrows = d.P[1] * d.P[2];
  const int row0 = blockIdx.x * RW;
  for (int idx = threadIdx.x; idx < RW * RS; idx += kThreads) {
    const int r = idx / RS, k = idx - r * RS;
    const int row = row0 + r;
    float2 v = make_float2(0.f, 0.f);
    if (row < nrows) {
      const int y = row % d.P[1], z = row / d.P[1];
      const bool inner = (unsigned)(y - d.m[1]) < (unsigned)d.D[1] && (unsigned)(z - d.m[2]) < (unsigned)d.D[2];
      if (inner) v = Sin[(size_t)row * d.HxP + k];
    }
    buf[idx] = v;
  }
  __syncthreads();
  const int r = threadIdx.x / TP, tp = threadIdx.x % TP;
  c2r_pre<M, TP>(buf + r * RS, twr, tp);
  fft_lds<M, TP, 1, RS, 1>(buf, tw, r, tp);
  float* fb = reinterpret_cast<float*>(buf);
  for (int idx = threadIdx.x; idx < RW * PX; idx += kThreads) {
    const int rr = idx / PX, x = idx - rr * PX;
    const int row = row0 + rr;
    float rv = 1.f;
    if (row < nrows) {
      const int y = row % d.P[1], z = row / d.P[1];
      const int ly = y - d.m[1], lz = z - d.m[2], lx = x - d.m[0];
      if ((unsigned)ly < (unsigned)d.D[1] && (unsigned)lz < (unsigned)d.D[2] && (unsigned)lx < (unsigned)d.D[0]) {
        const float im = q.img[((size_t)(q.o[2] + lz) * q.N[1] + (q.o[1] + ly)) * q.N[0] + (q.o[0] + lx)];
        if (im > 0.f) rv = im / fb[rr * 2 * RS + x];
      }
    }
    fb[rr * 2 * RS + x] = rv;
  }
  __syncthreads();
  fft_lds<M, TP, 1, RS, -1>(buf, tw, r, tp);
  const int row = row0 + r;
  r2c_post_store<M, TP, RS>(buf + r * RS, twr, tp, row < nrows, Sout + (size_t)row * d.HxP);
}

// ------------------------------------------------------------------ update of psi
// DeconvolutionMethods.computeNextValue (D/iteration/sequential/DeconvolutionMethods.java:320-358),
// float operations in Java order without contraction; Tikhonov in double.
__device__ __forceinline__ float rl_next_value(float last, float integral, float weight, double lambda, float minI,
                                               float maxI) {
  const float value = __fmul_rn(last, integral);
  float adjusted;
  if (value > 0.f) {
    if (lambda > 0.0) {
      const double vn = (double)__fdiv_rn(value, maxI);
      const double t = __ddiv_rn(__dadd_rn(__dsqrt_rn(__dadd_rn(1.0, __dmul_rn(__dmul_rn(2.0, lambda), vn))), -1.0),
                                 lambda);
      adjusted = __fmul_rn((float)t, maxI);
    } else {
      adjusted = value;
    }
  } else {
    adjusted = minI;
  }
  const float next = isnan(adjusted) ? minI : fmaxf(minI, adjusted);
  return __fadd_rn(last, __fmul_rn(__fsub_rn(next, last), weight));
}

__device__ __forceinline__ void block_reduce_stats(double s, double mx, double2* out) {
  __shared__ double rs[kThreads / 64], rm[kThreads / 64];
  for (int off = 32; off > 0; off >>= 1) {
    s += __shfl_down(s, off);
    mx = fmax(mx, __shfl_down(mx, off));
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) { rs[wv] = s; rm[wv] = mx; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0.0, b = -1.0;
    for (int i = 0; i < kThreads / 64; ++i) { a += rs[i]; b = fmax(b, rm[i]); }
    *out = make_double2(a, b);
  }
}

template <int M>
__global__ __launch_bounds__(kThreads) void k_x_update(Dom d, UpdateParams u, const float2* __restrict__ S,
                                                       const float2* __restrict__ twM, const float2* __restrict__ wr) {
  constexpr int TP = x_threads_per_row(M);
  constexpr int RW = kThreads / TP;
  constexpr int RS = M + 1;
  __shared__ float2 buf[RW * RS];
  __shared__ float2 tw[M];
  __shared__ float2 twr[M + 1];
  load_twiddles<M>(tw, twr, twM, wr);
  const int nrows = d.D[1] * d.D[2];
  const int row0 = blockIdx.x * RW;
  for (int idx = threadIdx.x; idx < RW * RS; idx += kThreads) {
    const int r = idx / RS, k = idx - r * RS;
    const int row = row0 + r;
    float2 v = make_float2(0.f, 0.f);
    if (row < nrows) {
      const int y = d.m[1] + row % d.D[1], z = d.m[2] + row / d.D[1];
      v = S[((size_t)z * d.P[1] + y) * d.HxP + k];
    }
    buf[idx] = v;
  }
  __syncthreads();
  const int r = threadIdx.x / TP, tp = threadIdx.x % TP;
  c2r_pre<M, TP>(buf + r * RS, twr, tp);
  fft_lds<M, TP, 1, RS, 1>(buf, tw, r, tp);
  const float* fb = reinterpret_cast<const float*>(buf);
  double sum = 0.0, mx = -1.0;
  const int DX = d.D[0];
  for (int idx = threadIdx.x; idx < RW * DX; idx += kThreads) {
    const int rr = idx / DX, lx = idx - rr * DX;
    const int row = row0 + rr;
    if (row < nrows) {
      const int ly = row % d.D[1], lz = row / d.D[1];
      const size_t gi = ((size_t)(u.o[2] + lz) * u.N[1] + (u.o[1] + ly)) * u.N[0] + (u.o[0] + lx);
      const float last = u.psi[gi];
      const float next = rl_next_value(last, fb[rr * 2 * RS + d.m[0] + lx], u.w[gi], u.lambda, u.minValue, u.maxI);
      u.psi[gi] = next;
      const float ch = __fsub_rn(next, last);
      sum += (double)ch;
      mx = fmax(mx, (double)ch);
    }
  }
  block_reduce_stats(sum, mx, u.partial + blockIdx.x);
}

// ------------------------------------------------------------------ inverse + store of the output region
template <int M>
__global__ __launch_bounds__(kThreads) void k_x_store(Dom d, StoreParams st, const float2* __restrict__ S,
                                                      const float2* __restrict__ twM, const float2* __restrict__ wr) {
  constexpr int TP = x_threads_per_row(M);
  constexpr int RW = kThreads / TP;
  constexpr int RS = M + 1;
  __shared__ float2 buf[RW * RS];
  __shared__ float2 tw[M];
  __shared__ float2 twr[M + 1];
  load_twiddles<M>(tw, twr, twM, wr);
  const int nrows = d.D[1] * d.D[2];
  const int row0 = blockIdx.x * RW;
  for (int idx = threadIdx.x; idx < RW * RS; idx += kThreads) {
    const int r = idx / RS, k = idx - r * RS;
    const int row = row0 + r;
    float2 v = make_float2(0.f, 0.f);
    if (row < nrows) {
      const int y = d.m[1] + row % d.D[1], z = d.m[2] + row / d.D[1];
      v = S[((size_t)z * d.P[1] + y) * d.HxP + k];
    }
    buf[idx] = v;
  }
  __syncthreads();
  const int r = threadIdx.x / TP, tp = threadIdx.x % TP;
  c2r_pre<M, TP>(buf + r * RS, twr, tp);
  fft_lds<M, TP, 1, RS, 1>(buf, tw, r, tp);
  const float* fb = reinterpret_cast<const float*>(buf);
  const int DX = d.D[0];
  for (int idx = threadIdx.x; idx < RW * DX; idx += kThreads) {
    const int rr = idx / DX, lx = idx - rr * DX;
    const int row = row0 + rr;
    if (row < nrows) st.out[(size_t)row * DX + lx] = fb[rr * 2 * RS + d.m[0] + lx];
  }
}

// ------------------------------------------------------------------ y / z passes
template <int L, int MODE>
__global__ __launch_bounds__(kThreads) void k_strided(StridedGeom gm, const float2* __restrict__ in,
                                                      float2* __restrict__ out, const float2* __restrict__ spec,
                                                      float scale, const float2* __restrict__ twg) {
  constexpr int W = strided_width(L);
  constexpr int TP = kThreads / W;
  constexpr int NE = L * W;
  constexpr int PER = (NE + kThreads - 1) / kThreads;
  __shared__ float2 tile[NE];
  __shared__ float2 tw[L];
  for (int i = threadIdx.x; i < L; i += kThreads) tw[i] = twg[i];
  const int col0 = blockIdx.x * W;
  const int wv = min(W, gm.ncols - col0);
  const size_t base = (size_t)blockIdx.y * gm.planeStride + col0;
  float2 sp[MODE == 2 ? PER : 1];
#pragma unroll
  for (int e = 0; e < PER; ++e) {
    const int idx = threadIdx.x + e * kThreads;
    if ((NE % kThreads) == 0 || idx < NE) {
      const int n = idx / W, w = idx % W;
      float2 v = make_float2(0.f, 0.f);
      if (w < wv) v = in[base + (size_t)n * gm.lineStride + w];
      tile[idx] = v;
      if constexpr (MODE == 2) {
        sp[e] = make_float2(0.f, 0.f);
        if (w < wv) sp[e] = spec[base + (size_t)n * gm.lineStride + w];
      }
    }
  }
  __syncthreads();
  const int p = threadIdx.x % W, tp = threadIdx.x / W;
  if constexpr (MODE == 1) {
    fft_lds<L, TP, W, 1, 1>(tile, tw, p, tp);
  } else {
    fft_lds<L, TP, W, 1, -1>(tile, tw, p, tp);
  }
  if constexpr (MODE == 2) {
#pragma unroll
    for (int e = 0; e < PER; ++e) {
      const int idx = threadIdx.x + e * kThreads;
      if ((NE % kThreads) == 0 || idx < NE) tile[idx] = cmul(tile[idx], sp[e]);
    }
    __syncthreads();
    fft_lds<L, TP, W, 1, 1>(tile, tw, p, tp);
  }
#pragma unroll
  for (int e = 0; e < PER; ++e) {
    const int idx = threadIdx.x + e * kThreads;
    if ((NE % kThreads) == 0 || idx < NE) {
      const int n = idx / W, w = idx % W;
      float2 v = tile[idx];
      if constexpr (MODE == 3) v = cscale(v, scale);
      if (w < wv) out[base + (size_t)n * gm.lineStride + w] = v;
    }
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
template <int L>
__host__ __device__ constexpr int s2_width() {
  // 16 columns = 128-byte runs (32 was slower at L = 300, 8 slower at L = 400); 8 above L = 640 (LDS)
  return L <= 640 ? 16 : 8;
}
constexpr bool kStridedXcdOrder = true;
// tile widths compiled per length: 8 and 16 columns where the tile fits 96 KB of LDS
template <int L, int W>
__host__ __device__ constexpr bool s2_fits() {
  return split2(L).A > 0 && L * W * 8 <= 96 * 1024;
}

template <int L, int W, int MODE>
__global__ __launch_bounds__(W * split2(L).A) void k_strided2(StridedGeom gm, const float2* __restrict__ in,
                                                              float2* __restrict__ out,
                                                              const float2* __restrict__ spec, float scale,
                                                              const float2* __restrict__ twg) {
  constexpr Split2 S = split2(L);
  constexpr int A = S.A, B = S.B;
  constexpr int NT = W * A;
  __shared__ float2 tile[L * W];
  __shared__ float2 tw[L];
  for (int i = threadIdx.x; i < L; i += NT) tw[i] = twg[i];
  const int w = threadIdx.x % W, q = threadIdx.x / W;
  // Work-groups are dealt to the 8 XCDs in turn; the tile order is remapped so that each XCD sweeps a
  // contiguous range of (column tile, plane) tiles: neighbouring column tiles, which share cache lines
  // when a tile row is shorter than a line, meet in one L2.
  unsigned tid = blockIdx.y * gridDim.x + blockIdx.x;
  const unsigned ntile = gridDim.x * gridDim.y;
  if (kStridedXcdOrder && ntile % 8 == 0) tid = (tid % 8) * (ntile / 8) + tid / 8;
  const int plane = (int)(tid / gridDim.x);
  const int col0 = (int)(tid % gridDim.x) * W;
  const bool colOk = w < gm.ncols - col0;
  const float2* src = in + (size_t)plane * gm.planeStride + col0 + w;
  float2* dst = out + (size_t)plane * gm.planeStride + col0 + w;
  const size_t ls = gm.lineStride;
  float2 v[A];
  if constexpr (MODE != 1) {
    // ---- forward step 1: thread q = b < B
    if (q < B) {
#pragma unroll
      for (int a = 0; a < A; ++a) v[a] = colOk ? src[(size_t)(q + B * a) * ls] : make_float2(0.f, 0.f);
    }
    __syncthreads();  // twiddle table
    if (q < B) {
      rdft<A, -1>(v);
#pragma unroll
      for (int k1 = 0; k1 < A; ++k1) {
        const float2 t = k1 == 0 ? v[k1] : cmul(v[k1], tw[q * k1]);
        tile[(q + B * k1) * W + w] = t;
      }
    }
    __syncthreads();
    // ---- forward step 2: thread q = k1 < A
    float2 u[B];
#pragma unroll
    for (int b = 0; b < B; ++b) u[b] = tile[(b + B * q) * W + w];
    rdft<B, -1>(u);  // u[k2] = X[q + A k2]
    if constexpr (MODE == 0 || MODE == 3) {
      if (colOk) {
#pragma unroll
        for (int k2 = 0; k2 < B; ++k2) {
          float2 x = u[k2];
          if constexpr (MODE == 3) x = cscale(x, scale);
          dst[(size_t)(q + A * k2) * ls] = x;
        }
      }
      return;
    } else {
      // spectrum product, then inverse step 1 over k2 in registers
      const float2* sp = spec + (size_t)plane * gm.planeStride + col0 + w;
#pragma unroll
      for (int k2 = 0; k2 < B; ++k2) {
        const float2 kv = colOk ? sp[(size_t)(q + A * k2) * ls] : make_float2(0.f, 0.f);
        u[k2] = cmul(u[k2], kv);
      }
      rdft<B, 1>(u);  // u[n2] = sum_k2 X[q + A k2] W_B^{-n2 k2}
      __syncthreads();  // every thread has read its forward values from the tile
#pragma unroll
      for (int n2 = 0; n2 < B; ++n2) {
        const float2 t = n2 == 0 ? u[n2] : cmul(u[n2], cconj(tw[n2 * q]));
        tile[(q + A * n2) * W + w] = t;
      }
    }
  } else {
    // ---- inverse only: thread q = k1 < A holds X[q + A k2]
    float2 u[B];
#pragma unroll
    for (int k2 = 0; k2 < B; ++k2) u[k2] = colOk ? src[(size_t)(q + A * k2) * ls] : make_float2(0.f, 0.f);
    __syncthreads();  // twiddle table
    rdft<B, 1>(u);
#pragma unroll
    for (int n2 = 0; n2 < B; ++n2) {
      const float2 t = n2 == 0 ? u[n2] : cmul(u[n2], cconj(tw[n2 * q]));
      tile[(q + A * n2) * W + w] = t;
    }
  }
  if constexpr (MODE == 1 || MODE == 2) {
    __syncthreads();
    // ---- inverse step 2: thread q = n2 < B does the A-point DFT over k1 -> x[n2 + B n1]
    if (q < B) {
#pragma unroll
      for (int k1 = 0; k1 < A; ++k1) v[k1] = tile[(k1 + A * q) * W + w];
      rdft<A, 1>(v);
      if (colOk) {
#pragma unroll
        for (int n1 = 0; n1 < A; ++n1) dst[(size_t)(q + B * n1) * ls] = v[n1];
      }
    }
  }
}

// ------------------------------------------------------------------ host launchers
template <int M>
void launch_x1_fwd(const Dom& d, const GatherParams& g, float2* S, const float2* twM, const float2* wr,
                  hipStream_t s) {
  constexpr int RW = kThreads / x_threads_per_row(M);
  const int rows = d.P[1] * d.P[2];
  hipLaunchKernelGGL(k_x_fwd<M>, dim3((rows + RW - 1) / RW), dim3(kThreads), 0, s, d, g, S, twM, wr);
}
template <int M>
void launch_x1_ratio(const Dom& d, const RatioParams& q, const float2* Sin, float2* Sout, const float2* twM,
                    const float2* wr, hipStream_t s) {
  constexpr int RW = kThreads / x_threads_per_row(M);
  const int rows = d.P[1] * d.P[2];
  hipLaunchKernelGGL(k_x_ratio<M>, dim3((rows + RW - 1) / RW), dim3(kThreads), 0, s, d, q, Sin, Sout, twM, wr);
}
template <int M>
int launch_x1_update(const Dom& d, const UpdateParams& u, const float2* S, const float2* twM, const float2* wr,
                    hipStream_t s) {
  constexpr int RW = kThreads / x_threads_per_row(M);
  const int rows = d.D[1] * d.D[2];
  const int nb = (rows + RW - 1) / RW;
  hipLaunchKernelGGL(k_x_update<M>, dim3(nb), dim3(kThreads), 0, s, d, u, S, twM, wr);
  return nb;
}
template <int M>
void launch_x1_store(const Dom& d, const StoreParams& st, const float2* S, const float2* twM, const float2* wr,
                    hipStream_t s) {
  constexpr int RW = kThreads / x_threads_per_row(M);
  const int rows = d.D[1] * d.D[2];
  hipLaunchKernelGGL(k_x_store<M>, dim3((rows + RW - 1) / RW), dim3(kThreads), 0, s, d, st, S, twM, wr);
}
template <int L, int W>
void launch_strided2(const StridedGeom& gm, int mode, const float2* in, float2* out, const float2* spec, float scale,
                     const float2* tw, hipStream_t s) {
  constexpr int NT = W * split2(L).A;
  const dim3 g2((gm.ncols + W - 1) / W, gm.nplanes);
  switch (mode) {
    case 0: hipLaunchKernelGGL((k_strided2<L, W, 0>), g2, dim3(NT), 0, s, gm, in, out, spec, scale, tw); break;
    case 1: hipLaunchKernelGGL((k_strided2<L, W, 1>), g2, dim3(NT), 0, s, gm, in, out, spec, scale, tw); break;
    case 2: hipLaunchKernelGGL((k_strided2<L, W, 2>), g2, dim3(NT), 0, s, gm, in, out, spec, scale, tw); break;
    default: hipLaunchKernelGGL((k_strided2<L, W, 3>), g2, dim3(NT), 0, s, gm, in, out, spec, scale, tw); break;
  }
}

// gm.width picks the tile width (0 or a width not compiled for L: s2_width<L>())
template <int L>
void launch_strided(const StridedGeom& gm, int mode, const float2* in, float2* out, const float2* spec, float scale,
                    const float2* tw, hipStream_t s) {
  if constexpr (s2_fits<L, s2_width<L>()>()) {
    if constexpr (s2_fits<L, 16>())
      if (gm.width == 16) return launch_strided2<L, 16>(gm, mode, in, out, spec, scale, tw, s);
    if constexpr (s2_fits<L, 8>())
      if (gm.width == 8) return launch_strided2<L, 8>(gm, mode, in, out, spec, scale, tw, s);
    return launch_strided2<L, s2_width<L>()>(gm, mode, in, out, spec, scale, tw, s);
  }
  constexpr int W = strided_width(L);
  const dim3 grid((gm.ncols + W - 1) / W, gm.nplanes);
  switch (mode) {
    case 0: hipLaunchKernelGGL((k_strided<L, 0>), grid, dim3(kThreads), 0, s, gm, in, out, spec, scale, tw); break;
    case 1: hipLaunchKernelGGL((k_strided<L, 1>), grid, dim3(kThreads), 0, s, gm, in, out, spec, scale, tw); break;
    case 2: hipLaunchKernelGGL((k_strided<L, 2>), grid, dim3(kThreads), 0, s, gm, in, out, spec, scale, tw); break;
    default: hipLaunchKernelGGL((k_strided<L, 3>), grid, dim3(kThreads), 0, s, gm, in, out, spec, scale, tw); break;
  }
}

}  // namespace mvr
