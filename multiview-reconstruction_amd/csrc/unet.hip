// unet.hip -- kernels of the learned initial estimate: a 3-level 3D U-Net
// (16/32/64 channels, 3^3 convolutions, ReLU) in BF16 on MFMA with FP32
// accumulation, forward and backward.
//
// Tensors are channels-last bf16 [n][z][y][x][c] with the channel count padded
// to a multiple of 16; a tensor view is (pointer, channel stride, channel
// offset), so a skip concatenation is two producers writing disjoint channel
// ranges of one buffer.  Convolution weights are [tap][cout][cin] (cin
// contiguous), master copies FP32, compute copies bf16.
//
// MFMA layout (v_mfma_f32_16x16x16_bf16): A is 16(m) x 16(k), lane l holds
// A[l%16][4(l/16)..4(l/16)+3]; B is 16(k) x 16(n), lane l holds
// B[4(l/16)..+3][l%16]; D lane l holds D[4(l/16)+r][l%16], r = 0..3.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "unet.hpp"

namespace mvr {

typedef __attribute__((ext_vector_type(4))) short s16x4;
typedef __attribute__((ext_vector_type(4))) float f32x4;

__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }
__device__ __forceinline__ uint16_t f2bf(float f) {
  const uint32_t u = __float_as_uint(f);
  return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);  // round to nearest even
}

// ------------------------------------------------------------------ convolution (forward and data gradient)
// The 3^3 convolutions run in the LDS-tiled kernel of unet_conv.hip (launch_conv_tiled).

// ------------------------------------------------------------------ weight gradient
// dw[tap][co][ci] = sum_v dout[v][co] * in[v + tap][ci], accumulated per work-group
// over a contiguous range of 16-voxel x-segments; partial[blockIdx.x][tap][co][ci].
// Grid: x = number of partitions, y = 27 (or 1) taps.  The 16 x 16 tiles
// (co block, ci block) are spread over the four waves.
__global__ __launch_bounds__(256) void k_conv_wgrad(WgradArgs a) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int i = lane & 15, g = lane >> 4;
  const int tap = blockIdx.y;
  int dz = 0, dy = 0, dx = 0;
  if (a.taps == 27) { dz = tap / 9 - 1; dy = (tap / 3) % 3 - 1; dx = tap % 3 - 1; }
  const int ncb = a.cout / 16, nib = a.cin / 16, ntiles = ncb * nib;
  // tiles handled by this wave: t = wave, wave + 4, ... (at most kMaxWgradTiles)
  f32x4 acc[kMaxWgradTiles];
#pragma unroll
  for (int t = 0; t < kMaxWgradTiles; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int segsPerRow = (a.W + 15) / 16;
  const long long nseg = (long long)a.N * a.D * a.H * segsPerRow;
  const long long s0 = nseg * blockIdx.x / gridDim.x, s1 = nseg * (blockIdx.x + 1) / gridDim.x;
  for (long long s = s0; s < s1; ++s) {
    const int sx = (int)(s % segsPerRow);
    const long long row = s / segsPerRow;
    const int y = (int)(row % a.H);
    const long long nz = row / a.H;
    const int z = (int)(nz % a.D), n = (int)(nz / a.D);
    const int x0 = sx * 16;
    // A operand: dout^T, A[m = co][k = v] -> lane holds dout[v = 4g + r][co = i]
    // B operand: in shifted by the tap, B[k = v][n = ci] -> lane holds in[v = 4g + r][ci = i]
    const int zz = z + dz, yy = y + dy;
    const bool rowok = (unsigned)zz < (unsigned)a.D && (unsigned)yy < (unsigned)a.H;
    for (int t = wave, slot = 0; t < ntiles; t += 4, ++slot) {
      const int cb = t / nib, ib = t % nib;
      s16x4 av, bv;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int x = x0 + 4 * g + r;
        uint16_t dv = 0, iv = 0;
        if (x < a.W) {
          const size_t vo = (((size_t)n * a.D + z) * a.H + y) * a.W + x;
          dv = a.dout[vo * a.dout_cs + a.dout_co + cb * 16 + i];
          const int xx = x + dx;
          if (rowok && (unsigned)xx < (unsigned)a.W) {
            const size_t vi = (((size_t)n * a.D + zz) * a.H + yy) * a.W + xx;
            iv = a.in[vi * a.in_cs + ib * 16 + i];
          }
        }
        av[r] = (short)dv;
        bv[r] = (short)iv;
      }
#pragma unroll
      for (int q = 0; q < kMaxWgradTiles; ++q)
        if (q == slot) acc[q] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc[q], 0, 0, 0);
    }
  }
  // D lane l holds dw[co = cb*16 + 4g + r][ci = ib*16 + i]
  float* part = a.partial + ((size_t)blockIdx.x * a.taps + tap) * a.cout * a.cin;
  for (int t = wave, slot = 0; t < ntiles; t += 4, ++slot) {
    const int cb = t / nib, ib = t % nib;
#pragma unroll
    for (int q = 0; q < kMaxWgradTiles; ++q)
      if (q == slot)
#pragma unroll
        for (int r = 0; r < 4; ++r) part[(size_t)(cb * 16 + 4 * g + r) * a.cin + ib * 16 + i] = acc[q][r];
  }
}

// sum of partials: grad[j] = sum_p partial[p][j], j < count
__global__ void k_sum_partials(const float* __restrict__ partial, int nparts, size_t count, float* __restrict__ grad) {
  for (size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x; j < count; j += (size_t)gridDim.x * blockDim.x) {
    float s = 0.f;
    for (int p = 0; p < nparts; ++p) s += partial[(size_t)p * count + j];
    grad[j] = s;
  }
}

// bias gradient partials: db[co] = sum_v dout[v][co]; partial[blockIdx.x][co]
__global__ __launch_bounds__(256) void k_bias_grad(const uint16_t* __restrict__ dout, int cs, int co0, int cout,
                                                   size_t nvox, float* __restrict__ partial) {
  __shared__ float red[256];
  const int c = threadIdx.x % cout;
  const int lanesPerC = 256 / cout;
  const int sub = threadIdx.x / cout;
  float s = 0.f;
  if (sub < lanesPerC) {
    const size_t v0 = nvox * blockIdx.x / gridDim.x, v1 = nvox * (blockIdx.x + 1) / gridDim.x;
    for (size_t v = v0 + sub; v < v1; v += lanesPerC) s += bf2f(dout[v * cs + co0 + c]);
  }
  red[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x < cout) {
    float t = 0.f;
    for (int k = 0; k < lanesPerC; ++k) t += red[k * cout + threadIdx.x];
    partial[(size_t)blockIdx.x * cout + threadIdx.x] = t;
  }
}

// ------------------------------------------------------------------ elementwise pieces
// All work on units of 8 channels (16 bytes); channel counts, strides and offsets are multiples of 16.
typedef __attribute__((ext_vector_type(8))) unsigned short u16x8;

__device__ __forceinline__ void unpack8(const u16x8& u, float f[8]) {
#pragma unroll
  for (int k = 0; k < 8; ++k) f[k] = bf2f(u[k]);
}
__device__ __forceinline__ u16x8 pack8(const float f[8]) {
  u16x8 u;
#pragma unroll
  for (int k = 0; k < 8; ++k) u[k] = f2bf(f[k]);
  return u;
}
template <typename T>
__device__ __forceinline__ const u16x8& at8(const T* p) { return *reinterpret_cast<const u16x8*>(p); }
__device__ __forceinline__ u16x8& at8w(uint16_t* p) { return *reinterpret_cast<u16x8*>(p); }

// Index spaces: row kernels run on grid (ceil(W * U / 256), H, N * D) of the tensor they iterate
// over, U = C / 8 units per voxel, so all index arithmetic is 32-bit.
struct RowIdx {
  int x, c, y, z, n;
  bool ok;
};
__device__ __forceinline__ RowIdx row_idx(int U, int W, int D) {
  RowIdx r;
  const int xu = blockIdx.x * blockDim.x + threadIdx.x;
  r.ok = xu < W * U;
  r.x = xu / U;
  r.c = (xu - r.x * U) * 8;
  r.y = blockIdx.y;
  r.z = blockIdx.z % D;
  r.n = blockIdx.z / D;
  return r;
}
__device__ __forceinline__ size_t vox(int n, int z, int y, int x, int D, int H, int W) {
  return (((size_t)n * D + z) * H + y) * W + x;
}

// ReLU backward in place: g[v][c] = post[v][c] > 0 ? g : 0 (C / 8 a power of two: ushift = log2(C / 8))
__global__ void k_relu_bwd(uint16_t* __restrict__ grad, int gcs, int gco, const uint16_t* __restrict__ post, int pcs,
                           int pco, int ushift, size_t nvox) {
  const size_t total = nvox << ushift;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    const size_t v = e >> ushift;
    const int c = (int)(e & ((1u << ushift) - 1)) * 8;
    const u16x8 pv = at8(post + v * pcs + pco + c);
    u16x8 gv = at8(grad + v * gcs + gco + c);
#pragma unroll
    for (int k = 0; k < 8; ++k)
      if (bf2f(pv[k]) <= 0.f) gv[k] = 0;
    at8w(grad + v * gcs + gco + c) = gv;
  }
}

// 2x2x2 average pooling (forward) over channels [0, C) of in -> out; grid over the output (D, H, W)
__global__ void k_pool_fwd(const uint16_t* __restrict__ in, int ics, int ico, uint16_t* __restrict__ out, int ocs,
                           int C, int D, int H, int W) {
  const RowIdx r = row_idx(C / 8, W, D);
  if (!r.ok) return;
  float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const size_t vi = vox(r.n, 2 * r.z + (k >> 2), 2 * r.y + ((k >> 1) & 1), 2 * r.x + (k & 1), 2 * D, 2 * H, 2 * W);
    float f[8];
    unpack8(at8(in + vi * ics + ico + r.c), f);
#pragma unroll
    for (int q = 0; q < 8; ++q) s[q] += f[q];
  }
#pragma unroll
  for (int q = 0; q < 8; ++q) s[q] *= 0.125f;
  at8w(out + vox(r.n, r.z, r.y, r.x, D, H, W) * ocs + r.c) = pack8(s);
}

// pooling backward: din[2v + k] (+)= dout[v] / 8; grid over the fine tensor (2D, 2H, 2W)
__global__ void k_pool_bwd(const uint16_t* __restrict__ dout, int ocs, uint16_t* __restrict__ din, int ics, int ico,
                           int C, int D, int H, int W, int accumulate) {  // D, H, W of the pooled tensor
  const RowIdx r = row_idx(C / 8, 2 * W, 2 * D);
  if (!r.ok) return;
  float gval[8];
  unpack8(at8(dout + vox(r.n, r.z / 2, r.y / 2, r.x / 2, D, H, W) * ocs + r.c), gval);
  uint16_t* p = din + vox(r.n, r.z, r.y, r.x, 2 * D, 2 * H, 2 * W) * ics + ico + r.c;
  float cur[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  if (accumulate) unpack8(at8(p), cur);
#pragma unroll
  for (int q = 0; q < 8; ++q) cur[q] += 0.125f * gval[q];
  at8w(p) = pack8(cur);
}

// nearest-neighbour x2 upsampling into channels [oco, oco + C) of out; grid over the fine tensor
__global__ void k_up_fwd(const uint16_t* __restrict__ in, int ics, uint16_t* __restrict__ out, int ocs, int oco, int C,
                         int D, int H, int W) {  // D, H, W of the coarse tensor
  const RowIdx r = row_idx(C / 8, 2 * W, 2 * D);
  if (!r.ok) return;
  at8w(out + vox(r.n, r.z, r.y, r.x, 2 * D, 2 * H, 2 * W) * ocs + oco + r.c) =
      at8(in + vox(r.n, r.z / 2, r.y / 2, r.x / 2, D, H, W) * ics + r.c);
}

// upsampling backward: dcoarse[v] = sum of the 8 fine gradients; grid over the coarse tensor
__global__ void k_up_bwd(const uint16_t* __restrict__ dfine, int fcs, int fco, uint16_t* __restrict__ dcoarse, int ccs,
                         int C, int D, int H, int W) {  // coarse dims
  const RowIdx r = row_idx(C / 8, W, D);
  if (!r.ok) return;
  float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const size_t vf = vox(r.n, 2 * r.z + (k >> 2), 2 * r.y + ((k >> 1) & 1), 2 * r.x + (k & 1), 2 * D, 2 * H, 2 * W);
    float f[8];
    unpack8(at8(dfine + vf * fcs + fco + r.c), f);
#pragma unroll
    for (int q = 0; q < 8; ++q) s[q] += f[q];
  }
  at8w(dcoarse + vox(r.n, r.z, r.y, r.x, D, H, W) * ccs + r.c) = pack8(s);
}

// network input: channels 0..V-1 = w_v * img_v / scale, channel V = their weighted mean
// sum_v w_v img_v / (scale * sum_v w_v) (0 where no view has weight), rest 0; one sample per
// volume region [x0, x0 + W) x [y0, ...) x [z0, ...) of volumes of dims N3.  Positions past the
// volume's last voxel (a region padded up to a multiple of 4) repeat that voxel.
__global__ void k_pack_input(InputArgs a) {
  const size_t total = (size_t)a.D * a.H * a.W;
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y, z = blockIdx.z;
  if (x >= a.W) return;
  {
    const size_t v = ((size_t)z * a.H + y) * a.W + x;
    const int gx = min(a.x0 + x, a.N3[0] - 1), gy = min(a.y0 + y, a.N3[1] - 1), gz = min(a.z0 + z, a.N3[2] - 1);
    const size_t gi = ((size_t)gz * a.N3[1] + gy) * a.N3[0] + gx;
    float s = 0.f, sw = 0.f;
    float val[16];
#pragma unroll
    for (int c = 0; c < 16; ++c) val[c] = 0.f;
#pragma unroll
    for (int c = 0; c < kNetMaxViews; ++c) {  // unrolled over the packer's view slots
      if (c < a.V) {
        const float wv = a.w[c][gi];
        val[c] = a.img[c][gi] * wv * a.inv_scale;
        s += val[c];
        sw += wv;
      }
    }
    const float mean = sw > 0.f ? s / sw : 0.f;
#pragma unroll
    for (int c = 0; c < 16; ++c)
      if (c == a.V) val[c] = mean;
    uint16_t* o = a.out + ((size_t)a.sample * total + v) * 16;
    at8w(o) = pack8(val);
    at8w(o + 8) = pack8(val + 8);
    if (a.avg_out) a.avg_out[(size_t)a.sample * total + v] = mean;
  }
}

// output head: r[v] = sum_c wo[c] d1[v][c] + bo; pred = avg[v] + r[v]
// training: loss partials of (pred - target)^2 and dr[v] = 2 (pred - target) / count
// inference: psi[v] = max(minValue, scale * pred)
__global__ __launch_bounds__(256) void k_head(HeadArgs a) {
  __shared__ double red[256];
  double ls = 0.0;
  for (size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x; v < a.nvox; v += (size_t)gridDim.x * blockDim.x) {
    float r = a.bias[0];
    const uint16_t* p = a.d1 + v * a.d1_cs;
#pragma unroll
    for (int c = 0; c < 16; ++c) r += a.wo[c] * bf2f(p[c]);
    const float pred = a.avg[v] + r;
    if (a.target) {
      const float e = pred - a.target[v];
      ls += (double)e * e;
      a.dr[v] = 2.f * e / (float)a.nvox;
    }
    if (a.psi) {
      // net dims (W, H, D) may be the output dims (ox, oy, oz) padded up; only inner voxels are written
      const int x = (int)(v % a.W), y = (int)((v / a.W) % a.H), z = (int)(v / ((size_t)a.W * a.H));
      if (x < a.ox && y < a.oy && z >= a.zlo && z < a.oz)
        a.psi[((size_t)z * a.oy + y) * a.ox + x] = fmaxf(a.minValue, a.scale * pred);
    }
  }
  red[threadIdx.x] = ls;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0 && a.loss_partial) a.loss_partial[blockIdx.x] = red[0];
}

// head backward: d_d1[v][c] = dr[v] * wo[c]; partials of dwo[c] = sum dr d1[v][c], dbo = sum dr
__global__ __launch_bounds__(256) void k_head_bwd(HeadArgs a, uint16_t* __restrict__ dd1, float* __restrict__ partial) {
  __shared__ float red[256][17];
  float acc[17];
#pragma unroll
  for (int c = 0; c < 17; ++c) acc[c] = 0.f;
  for (size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x; v < a.nvox; v += (size_t)gridDim.x * blockDim.x) {
    const float dr = a.dr[v];
    const uint16_t* p = a.d1 + v * a.d1_cs;
#pragma unroll
    for (int c = 0; c < 16; ++c) {
      acc[c] += dr * bf2f(p[c]);
      dd1[v * 16 + c] = f2bf(dr * a.wo[c]);
    }
    acc[16] += dr;
  }
#pragma unroll
  for (int c = 0; c < 17; ++c) red[threadIdx.x][c] = acc[c];
  __syncthreads();
  if (threadIdx.x < 17) {
    float s = 0.f;
    for (int k = 0; k < 256; ++k) s += red[k][threadIdx.x];
    partial[(size_t)blockIdx.x * 17 + threadIdx.x] = s;
  }
}

// Adam on FP32 master weights; refreshes the bf16 compute copy and, for 3^3 layers,
// the flipped/transposed copy used by the data gradient (wt[26 - tap][ci][co] = w[tap][co][ci]).
__global__ void k_adam(AdamArgs a) {
  for (size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x; j < a.count; j += (size_t)gridDim.x * blockDim.x) {
    const float g = a.grad[j];
    const float m = a.beta1 * a.m[j] + (1.f - a.beta1) * g;
    const float v = a.beta2 * a.v[j] + (1.f - a.beta2) * g * g;
    a.m[j] = m;
    a.v[j] = v;
    const float mh = m * a.c1, vh = v * a.c2;
    const float w = a.w[j] - a.lr * mh / (sqrtf(vh) + a.eps);
    a.w[j] = w;
    if (a.wbf) a.wbf[j] = f2bf(w);
    if (a.wt) {
      const int per = a.cout * a.cin;
      const int tap = (int)(j / per), rem = (int)(j % per);
      const int co = rem / a.cin, ci = rem % a.cin;
      a.wt[((size_t)(a.taps - 1 - tap) * a.cin + ci) * a.cout + co] = f2bf(w);
    }
  }
}

__global__ void k_f32_to_bf16(const float* __restrict__ src, uint16_t* __restrict__ dst, size_t n) {
  for (size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (size_t)gridDim.x * blockDim.x)
    dst[j] = f2bf(src[j]);
}

__global__ void k_bf16_to_f32(const uint16_t* __restrict__ src, float* __restrict__ dst, size_t n) {
  for (size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (size_t)gridDim.x * blockDim.x)
    dst[j] = bf2f(src[j]);
}

__global__ void k_zero_u16(uint16_t* __restrict__ p, size_t n) {
  for (size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (size_t)gridDim.x * blockDim.x) p[j] = 0;
}

// normalisation of the network input: partial[b] = {sum w_v img_v, sum w_v} (FP64) over a
// grid-stride share of the voxels and all views; the scale is the ratio of the totals
__global__ __launch_bounds__(256) void k_wsum_partials(InputArgs a, size_t n, double2* __restrict__ partial) {
  __shared__ double red[256], redw[256];
  double s = 0.0, sw = 0.0;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  const size_t t0 = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  // Four voxels per step with 16-byte loads when every volume is 16-byte aligned: all views' loads are issued
  // before the sums (views beyond V read view 0 and are not summed), and each voxel's float sums run over the
  // views in order, as in the scalar loop below that takes the rest.
  bool vec = true;
  for (int c = 0; c < a.V; ++c) vec = vec && !(((uintptr_t)a.img[c] | (uintptr_t)a.w[c]) & 15);
  const size_t n4 = vec ? n / 4 : 0;
  for (size_t i = t0; i < n4; i += stride) {
    float4 im[kNetMaxViews], wv[kNetMaxViews];
#pragma unroll
    for (int c = 0; c < kNetMaxViews; ++c) {
      const int cc = c < a.V ? c : 0;
      im[c] = reinterpret_cast<const float4*>(a.img[cc])[i];
      wv[c] = reinterpret_cast<const float4*>(a.w[cc])[i];
    }
    float4 t = make_float4(0.f, 0.f, 0.f, 0.f), tw = t;
#pragma unroll
    for (int c = 0; c < kNetMaxViews; ++c) {
      if (c >= a.V) break;
      t.x += im[c].x * wv[c].x;
      t.y += im[c].y * wv[c].y;
      t.z += im[c].z * wv[c].z;
      t.w += im[c].w * wv[c].w;
      tw.x += wv[c].x;
      tw.y += wv[c].y;
      tw.z += wv[c].z;
      tw.w += wv[c].w;
    }
    s += (double)t.x;
    s += (double)t.y;
    s += (double)t.z;
    s += (double)t.w;
    sw += (double)tw.x;
    sw += (double)tw.y;
    sw += (double)tw.z;
    sw += (double)tw.w;
  }
  for (size_t i = 4 * n4 + t0; i < n; i += stride) {
    float t = 0.f, tw = 0.f;
    for (int c = 0; c < a.V; ++c) {
      t += a.img[c][i] * a.w[c][i];
      tw += a.w[c][i];
    }
    s += (double)t;
    sw += (double)tw;
  }
  red[threadIdx.x] = s;
  redw[threadIdx.x] = sw;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) {
      red[threadIdx.x] += red[threadIdx.x + off];
      redw[threadIdx.x] += redw[threadIdx.x + off];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = make_double2(red[0], redw[0]);
}

// target patch of a training volume, normalised: dst[v] = src[o + v] * invScale
__global__ void k_gather_target(const float* __restrict__ src, int3 N, int3 o, int P, float invScale,
                                float* __restrict__ dst) {
  const size_t total = (size_t)P * P * P;
  for (size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x; v < total; v += (size_t)gridDim.x * blockDim.x) {
    const int x = (int)(v % P), y = (int)((v / P) % P), z = (int)(v / ((size_t)P * P));
    dst[v] = src[((size_t)(o.z + z) * N.y + (o.y + y)) * N.x + (o.x + x)] * invScale;
  }
}

// ------------------------------------------------------------------ launchers
static inline int grid1(size_t n, int cap = 8192) {
  const size_t b = (n + 255) / 256;
  return (int)(b < 1 ? 1 : (b > (size_t)cap ? (size_t)cap : b));
}

bool launch_conv_fwd(const ConvArgs& a, hipStream_t s) {
  // every layer of the network has cin, cout and channel strides that are multiples of 16
  return launch_conv_tiled(a, s);
}
void launch_conv_wgrad(const WgradArgs& a, int nparts, hipStream_t s) {
  hipLaunchKernelGGL(k_conv_wgrad, dim3(nparts, a.taps), dim3(256), 0, s, a);
}
void launch_sum_partials(const float* partial, int nparts, size_t count, float* grad, hipStream_t s) {
  hipLaunchKernelGGL(k_sum_partials, dim3(grid1(count, 4096)), dim3(256), 0, s, partial, nparts, count, grad);
}
void launch_bias_grad(const uint16_t* dout, int cs, int co0, int cout, size_t nvox, float* partial, int nparts,
                      hipStream_t s) {
  hipLaunchKernelGGL(k_bias_grad, dim3(nparts), dim3(256), 0, s, dout, cs, co0, cout, nvox, partial);
}
void launch_relu_bwd(uint16_t* g, int gcs, int gco, const uint16_t* post, int pcs, int pco, int C, size_t nvox,
                     hipStream_t s) {
  const int ushift = C == 16 ? 1 : C == 32 ? 2 : 3;  // C in {16, 32, 64}
  hipLaunchKernelGGL(k_relu_bwd, dim3(grid1(nvox * C / 8)), dim3(256), 0, s, g, gcs, gco, post, pcs, pco, ushift, nvox);
}
void launch_pool_fwd(const uint16_t* in, int ics, int ico, uint16_t* out, int ocs, int C, int N, int D, int H, int W,
                     hipStream_t s) {
  const dim3 grid((W * (C / 8) + 255) / 256, H, N * D);
  hipLaunchKernelGGL(k_pool_fwd, grid, dim3(256), 0, s, in, ics, ico, out, ocs, C, D, H, W);
}
void launch_pool_bwd(const uint16_t* dout, int ocs, uint16_t* din, int ics, int ico, int C, int N, int D, int H, int W,
                     int accumulate, hipStream_t s) {
  const dim3 grid((2 * W * (C / 8) + 255) / 256, 2 * H, N * 2 * D);
  hipLaunchKernelGGL(k_pool_bwd, grid, dim3(256), 0, s, dout, ocs, din, ics, ico, C, D, H, W, accumulate);
}
void launch_up_fwd(const uint16_t* in, int ics, uint16_t* out, int ocs, int oco, int C, int N, int D, int H, int W,
                   hipStream_t s) {
  const dim3 grid((2 * W * (C / 8) + 255) / 256, 2 * H, N * 2 * D);
  hipLaunchKernelGGL(k_up_fwd, grid, dim3(256), 0, s, in, ics, out, ocs, oco, C, D, H, W);
}
void launch_up_bwd(const uint16_t* dfine, int fcs, int fco, uint16_t* dcoarse, int ccs, int C, int N, int D, int H,
                   int W, hipStream_t s) {
  const dim3 grid((W * (C / 8) + 255) / 256, H, N * D);
  hipLaunchKernelGGL(k_up_bwd, grid, dim3(256), 0, s, dfine, fcs, fco, dcoarse, ccs, C, D, H, W);
}
void launch_pack_input(const InputArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_pack_input, dim3((a.W + 255) / 256, a.H, a.D), dim3(256), 0, s, a);
}
int launch_head(const HeadArgs& a, hipStream_t s) {
  const int nb = grid1(a.nvox, 1024);
  hipLaunchKernelGGL(k_head, dim3(nb), dim3(256), 0, s, a);
  return nb;
}
int launch_head_bwd(const HeadArgs& a, uint16_t* dd1, float* partial, hipStream_t s) {
  const int nb = grid1(a.nvox, 512);
  hipLaunchKernelGGL(k_head_bwd, dim3(nb), dim3(256), 0, s, a, dd1, partial);
  return nb;
}
void launch_adam(const AdamArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_adam, dim3(grid1(a.count, 2048)), dim3(256), 0, s, a);
}
void launch_f32_to_bf16(const float* src, uint16_t* dst, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_f32_to_bf16, dim3(grid1(n)), dim3(256), 0, s, src, dst, n);
}
void launch_bf16_to_f32(const uint16_t* src, float* dst, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_bf16_to_f32, dim3(grid1(n)), dim3(256), 0, s, src, dst, n);
}
void launch_gather_target(const float* src, const int N[3], const int o[3], int P, float invScale, float* dst,
                          hipStream_t s) {
  hipLaunchKernelGGL(k_gather_target, dim3(grid1((size_t)P * P * P, 2048)), dim3(256), 0, s, src,
                     make_int3(N[0], N[1], N[2]), make_int3(o[0], o[1], o[2]), P, invScale, dst);
}
int launch_wsum_partials(const InputArgs& a, size_t n, double2* partial, hipStream_t s) {
  const int nb = grid1(n, 1024);
  hipLaunchKernelGGL(k_wsum_partials, dim3(nb), dim3(256), 0, s, a, n, partial);
  return nb;
}
void launch_zero_u16(uint16_t* p, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_zero_u16, dim3(grid1(n)), dim3(256), 0, s, p, n);
}

}  // namespace mvr
