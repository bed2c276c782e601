// unet_conv.hip -- LDS-tiled 3^3 convolution of the U-Net on BF16 MFMA (forward and data gradient).
//
// out[v][co] = act( sum_tap sum_ci in[v + off(tap)][ci] * w[tap][co][ci] + bias[co] ), zero padding,
// channels-last bf16 tensors (see unet.hip).  One work-group computes an output tile of TY rows x TX
// voxels of one z plane for NG x 16 output channels (grid.y selects the channel group):
//   - an input may be a concatenation [up(coarse), fine] whose first part is read from the coarse
//     tensor at half the coordinates (nearest x2 upsampling fused into the staging);
//   - per 16-channel input chunk, the halo tile 3 x (TY+2) x (TX+2) x 16 bf16 and the chunk's weights
//     [28 taps][NG*16 cout][16 cin] (tap 27 = zeros) are staged into LDS with 16-byte accesses, each
//     stored as two 8-channel halves [half][...][8] so that the 16 lanes reading one half of 16
//     consecutive voxels (or output channels) touch 256 contiguous bytes (no bank conflicts);
//   - v_mfma_f32_16x16x32_bf16 takes K = 32 as two taps x 16 channels: lanes 0-31 carry tap 2p,
//     lanes 32-63 tap 2p+1, eight channels per lane;
//   - the tile's 16-voxel row segments are the M-tiles (16 per work-group, 4 per wave); a wave reads
//     each A fragment from LDS once per tap pair and reuses it for its NG output tiles, and each B
//     fragment once for its 4 M-tiles.
// The sum over K pairs A and B slots the same way, so the kernel relies only on the M/N placement:
// lane l holds rows/columns l%16 of A/B, and D[4(l/16)+r][l%16].
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "unet.hpp"

namespace mvr {

namespace {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(4))) float f32x4;

__device__ __forceinline__ uint16_t f2bf_rne(float f) {
  const uint32_t u = __float_as_uint(f);
  return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

// planes per work-group: a work-group walks zp consecutive z planes of its (y, x) tile (single-chunk layers
// stage their weights once); launches with fewer than kConvZPTiles plane tiles use one plane per work-group
// so that the grid still fills the chip
constexpr int kConvZP = 4;
constexpr long long kConvZPTiles = 16384;

template <int TX, int NG>
__global__ __launch_bounds__(256) void k_conv_tile(ConvArgs a, int zp) {
  constexpr int TY = 256 / TX;          // tile rows
  constexpr int HX = TX + 2, HY = TY + 2;
  constexpr int HV = 3 * HY * HX;       // halo voxels
  constexpr int SEGX = TX / 16;         // 16-voxel segments per row
  constexpr int WC = NG * 16;           // output channels of this work-group
  __shared__ __attribute__((aligned(16))) uint16_t tile[HV * 16];
  __shared__ __attribute__((aligned(16))) uint16_t wl[28 * WC * 16];

  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int i = lane & 15, g = lane >> 4;
  const int hi = g >> 1;                // tap of the pair carried by this lane
  const int ch8 = (g & 1) * 8;          // first of the lane's eight channels
  // work-group -> (n, z group, y block, x block); x block fastest.  Consecutive work-groups go to the 8 XCDs
  // in turn, so the tile order is remapped to give each XCD a contiguous range of tiles (whole z planes
  // and their neighbours stay in that XCD's L2).
  const int nbx = (a.W + TX - 1) / TX, nby = (a.H + TY - 1) / TY, nzg = (a.D + zp - 1) / zp;
  const int nbt = (int)gridDim.x;
  int b = nbt % 8 == 0 ? (int)(blockIdx.x % 8) * (nbt / 8) + (int)(blockIdx.x / 8) : (int)blockIdx.x;
  const int bx = b % nbx; b /= nbx;
  const int by = b % nby; b /= nby;
  const int zg = b % nzg, n = b / nzg;
  const int x0 = bx * TX, y0 = by * TY;
  const int zs = zg * zp, ze = min(a.D, zs + zp);
  const int co0 = blockIdx.y * WC;

  f32x4 acc[4][NG];
#pragma unroll
  for (int m = 0; m < 4; ++m)
#pragma unroll
    for (int t = 0; t < NG; ++t) acc[m][t] = f32x4{0.f, 0.f, 0.f, 0.f};

  // this wave's M-tiles: segments 4 * wave .. 4 * wave + 3 of the tile (row-major)
  int segBase[4];
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    const int s = 4 * wave + m;
    segBase[m] = (s / SEGX) * HX + (s % SEGX) * 16 + i;  // halo row offset (before the tap shift)
  }

  // Staging of one (plane, 16-channel chunk) item: every thread issues all of its 16-byte loads (branch-free,
  // invalid lanes read a valid dummy address) before the first LDS write, so the loads are in flight together.
  constexpr int NL = (HV * 2 + 255) / 256;         // halo loads per thread
  constexpr int NW = (28 * WC * 2 + 255) / 256;    // weight loads per thread
  uint4 hbuf[NL], wbuf[NW];
  // per-thread halo coordinates of unit k: v = (threadIdx.x >> 1) + 128 k, decomposed incrementally
  const int half = threadIdx.x & 1;
  auto issue = [&](int z, int cb, bool weights) {
    // the chunk comes from the upsampled source (half coordinates) or from in
    const bool up = cb * 16 < a.up_c;
    const int sh = up ? 1 : 0;
    const int sD = a.D >> sh, sH = a.H >> sh, sW = a.W >> sh;
    const int cs = up ? a.up_cs : a.in_cs;
    // base of the first source plane the halo touches, in 64 bits; offsets from it span at most three
    // planes and fit 32 bits (checked at launch).  zb = max(z - 1, 0) >> sh is a valid plane, so the
    // dummy address of an invalid lane (offset 0) is inside the tensor.
    const int zb = max(z - 1, 0) >> sh;
    const uint16_t* src = (up ? a.up_in + cb * 16 : a.in + (cb * 16 - a.up_c)) +
                          ((size_t)n * sD + zb) * sH * sW * cs + half * 8;
    int v0 = threadIdx.x >> 1;
    int hx = v0 % HX, hy = (v0 / HX) % HY, hz = v0 / (HX * HY);
    constexpr int DX = 128 % HX, DY = (128 / HX) % HY, DZ = 128 / (HX * HY);
#pragma unroll
    for (int k = 0; k < NL; ++k) {
      if (k) {  // advance by 128 halo voxels
        hx += DX;
        int cy = hx >= HX;
        hx -= cy ? HX : 0;
        hy += DY + cy;
        int cz = hy >= HY;
        hy -= cz ? HY : 0;
        hz += DZ + cz;
      }
      const int xx = x0 + hx - 1, yy = y0 + hy - 1, zz = z + hz - 1;
      const bool ok = (threadIdx.x + 256 * k) < HV * 2 && (unsigned)xx < (unsigned)a.W &&
                      (unsigned)yy < (unsigned)a.H && (unsigned)zz < (unsigned)a.D;
      const int gv = (((zz >> sh) - zb) * sH + (yy >> sh)) * sW + (xx >> sh);
      hbuf[k] = *reinterpret_cast<const uint4*>(src + (ok ? gv * cs : 0));
      if (!ok) hbuf[k] = make_uint4(0u, 0u, 0u, 0u);
    }
    if (!weights) return;
#pragma unroll
    for (int k = 0; k < NW; ++k) {
      const int e = threadIdx.x + 256 * k;
      const int c = (e >> 1) % WC, tap = (e >> 1) / WC;
      const bool ok = tap < 27;
      wbuf[k] = *reinterpret_cast<const uint4*>(a.w + ((ok ? tap : 0) * a.cout + co0 + c) * a.cin + cb * 16 +
                                               half * 8);
      if (!ok) wbuf[k] = make_uint4(0u, 0u, 0u, 0u);
    }
  };
  auto park = [&](bool weights) {
#pragma unroll
    for (int k = 0; k < NL; ++k) {
      const int e = threadIdx.x + 256 * k;
      if (e < HV * 2) *reinterpret_cast<uint4*>(tile + ((e & 1) * HV + (e >> 1)) * 8) = hbuf[k];
    }
    if (!weights) return;
#pragma unroll
    for (int k = 0; k < NW; ++k) {
      const int e = threadIdx.x + 256 * k;
      if (e < 28 * WC * 2) *reinterpret_cast<uint4*>(wl + ((e & 1) * 28 * WC + (e >> 1)) * 8) = wbuf[k];
    }
  };

  const int nchunk = a.cin / 16;
  const int nitems = (ze - zs) * nchunk;  // (plane, chunk) items, chunks fastest
  const int hsel = ch8 ? 1 : 0;  // the lane's channel half
  // a single-chunk layer stages its weights once per work-group
  for (int it = 0; it < nitems; ++it) {
    const int z = zs + it / nchunk, cb = it % nchunk;
    const bool wts = nchunk > 1 || it == 0;
    if (it) __syncthreads();  // the previous item's LDS reads are done
    issue(z, cb, wts);
    park(wts);
    __syncthreads();
    // full unroll folds the tap arithmetic into constants; single-group layers keep registers for occupancy
#pragma unroll(NG == 1 ? 2 : 14)
    for (int p = 0; p < 14; ++p) {
      const int tap = 2 * p + hi;                 // 27 -> zero weights; its A reads the centre tap
      const int tv = tap < 27 ? tap : 13;
      const int dz = tv / 9, dy = (tv / 3) % 3, dx = tv % 3;
      const int shift = (dz * HY + dy) * HX + dx;
      bf16x8 bfr[NG];
#pragma unroll
      for (int t = 0; t < NG; ++t)
        bfr[t] = *reinterpret_cast<const bf16x8*>(wl + (hsel * 28 * WC + tap * WC + t * 16 + i) * 8);
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        const bf16x8 afr = *reinterpret_cast<const bf16x8*>(tile + (hsel * HV + segBase[m] + shift) * 8);
#pragma unroll
        for (int t = 0; t < NG; ++t) acc[m][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(afr, bfr[t], acc[m][t], 0, 0, 0);
      }
    }
    if (cb != nchunk - 1) continue;

    // epilogue of plane z: bias, ReLU, bf16 stores (voxel 4g + r of the segment, channel co0 + t * 16 + i)
    float bias[NG];
#pragma unroll
    for (int t = 0; t < NG; ++t) bias[t] = a.bias ? a.bias[co0 + t * 16 + i] : 0.f;
    uint16_t* outPlane = a.out + (((size_t)n * a.D + z) * a.H) * a.W * a.out_cs + a.out_co + co0 + i;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const int s = 4 * wave + m;
      const int y = y0 + s / SEGX;
      const int xs = x0 + (s % SEGX) * 16 + 4 * g;
      if (y < a.H) {
        uint16_t* orow = outPlane + ((size_t)y * a.W + xs) * a.out_cs;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          if (xs + r >= a.W) continue;
#pragma unroll
          for (int t = 0; t < NG; ++t) {
            float v = acc[m][t][r] + bias[t];
            if (a.relu) v = fmaxf(v, 0.f);
            orow[r * a.out_cs + t * 16] = f2bf_rne(v);
          }
        }
      }
#pragma unroll
      for (int t = 0; t < NG; ++t) acc[m][t] = f32x4{0.f, 0.f, 0.f, 0.f};
    }
  }
}

template <int TX, int NG>
void launch_tile(const ConvArgs& a, hipStream_t s) {
  constexpr int TY = 256 / TX;
  const long long tiles = (long long)((a.W + TX - 1) / TX) * ((a.H + TY - 1) / TY) * a.D * a.N;
  const int zp = tiles >= kConvZPTiles ? kConvZP : 1;
  const unsigned nb = (unsigned)((a.W + TX - 1) / TX) * (unsigned)((a.H + TY - 1) / TY) *
                      (unsigned)((a.D + zp - 1) / zp) * a.N;
  hipLaunchKernelGGL((k_conv_tile<TX, NG>), dim3(nb, a.cout / (16 * NG)), dim3(256), 0, s, a, zp);
}

template <int TX>
void launch_tx(const ConvArgs& a, hipStream_t s) {
  if ((a.cout / 16) % 2 == 0)
    launch_tile<TX, 2>(a, s);
  else
    launch_tile<TX, 1>(a, s);
}

}  // namespace

// 3^3 convolutions with cin and cout multiples of 16
bool launch_conv_tiled(const ConvArgs& a, hipStream_t s) {
  if (a.taps != 27 || a.cin % 16 != 0 || a.cout % 16 != 0 || a.in_cs % 16 != 0) return false;
  // the staging's offsets from its first halo plane cover three planes in 32 bits
  if (3ll * a.H * a.W * std::max(a.in_cs, a.up_cs) >= (1ll << 31)) return false;
  if (a.up_c && (a.up_c % 16 != 0 || a.up_cs % 16 != 0 || !a.up_in || (a.W | a.H | a.D) & 1)) return false;
  // 16 x 16 voxel tiles: the smallest halo per output voxel (3 x 18 x 18 staged per 256 outputs; 64 x 4 and
  // 32 x 8 tiles stage 1188 and 1020 halo voxels for the same outputs and measured 2.4 % / 0.5 % slower at 256^3)
  launch_tx<16>(a, s);
  return true;
}

}  // namespace mvr
