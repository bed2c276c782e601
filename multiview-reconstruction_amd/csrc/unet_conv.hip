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

// Tile geometry of the 3^3 convolutions (16 x 16 output voxels per work-group at TX = 16) and the halo offsets of
// the taps; HVP pads each channel half of the halo to whole 256-voxel DMA rounds.
template <int TX>
struct ConvTile {
  static constexpr int TY = 256 / TX;
  static constexpr int HX = TX + 2, HY = TY + 2;
  static constexpr int HV = 3 * HY * HX;          // halo voxels
  static constexpr int NR = (HV + 255) / 256;     // DMA rounds per half
  static constexpr int HVP = NR * 256;            // halo slots per half
  static constexpr int WC = 16;
  // halo offset of tap t (t = 27: the zero-weight tap, whose A reads the centre tap 13)
  __host__ __device__ static constexpr int shift(int t) {
    const int tv = t < 27 ? t : 13;
    return (tv / 9 * HY + tv / 3 % 3) * HX + tv % 3;
  }
  // the odd tap of pair p lies one of four offsets beyond the even one
  static constexpr int DOFF[4] = {1, HX - 2, HX * HY - 2 * HX - 2, shift(27) - shift(26)};
  __host__ __device__ static constexpr int dkind(int p) {
    const int d = shift(2 * p + 1) - shift(2 * p);
    return d == DOFF[0] ? 0 : d == DOFF[1] ? 1 : d == DOFF[2] ? 2 : d == DOFF[3] ? 3 : -1;
  }
  __host__ __device__ static constexpr bool pairs_ok() {
    for (int p = 0; p < 14; ++p)
      if (dkind(p) < 0) return false;
    return true;
  }
};

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
    // Every fragment address is one of four per-lane bases plus a constant: the odd tap of a pair lies one of four
    // offsets beyond the even one, and the B fragment's tap parity is the lane's (no tap arithmetic in the loop).
    {
      using G = ConvTile<TX>;
      static_assert(TX == 16 && G::pairs_ok(), "one 16-voxel segment per tile row, four odd-tap offsets");
      const int lb = hsel * HV + 4 * wave * HX + i;
      const uint16_t* ab[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) ab[k] = tile + (lb + hi * G::DOFF[k]) * 8;
      const uint16_t* wb = wl + (hsel * 28 * WC + hi * WC + i) * 8;
#pragma unroll
      for (int p = 0; p < 14; ++p) {
        bf16x8 bfr[NG];
#pragma unroll
        for (int t = 0; t < NG; ++t) bfr[t] = *reinterpret_cast<const bf16x8*>(wb + (2 * p * WC + t * 16) * 8);
        const uint16_t* a0 = ab[G::dkind(p)] + G::shift(2 * p) * 8;
#pragma unroll
        for (int m = 0; m < 4; ++m) {
          const bf16x8 afr = *reinterpret_cast<const bf16x8*>(a0 + m * HX * 8);
#pragma unroll
          for (int t = 0; t < NG; ++t)
            acc[m][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(afr, bfr[t], acc[m][t], 0, 0, 0);
        }
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

// Single-chunk layers with 16 output channels (the full-resolution 16 -> 16 layers) on large launches: the
// work-group walks kConvZPDma planes and stages the next plane's halo into a second LDS buffer with LDS-DMA loads
// (global_load_lds_dwordx4: 16 bytes per lane into LDS at the wave's base + 16 lane, no registers) while the
// MFMAs of the current plane run.  One DMA instruction per 8-channel half writes 256 consecutive voxels of
// that half, so the LDS layout is the same half-major [half][voxel][8] as above (halves padded to HVP voxels,
// whole DMA rounds; the padding slots take the dummy loads and are never read).  Halo voxels outside the
// volume load a valid dummy address and are zeroed after the DMA has landed.  The arithmetic is that of
// k_conv_tile<TX, 1>: the results are bitwise equal.
constexpr int kConvZPDma = 8;

// per-lane state of the DMA staging: the lane's halo voxels v0 + 256 r (r < NR), their element offsets from the
// plane before the output plane (the same for every plane), which are inside the xy extent, and their z offsets
template <int TX>
struct DmaLane {
  const uint16_t* nbase;  // this sample's first plane
  long long planeElems;
  int x0, y0;
  int off[ConvTile<TX>::NR];
  int hz[ConvTile<TX>::NR];
  unsigned okxy;          // bit r: voxel r inside the halo and the xy extent
};

// DMA of plane z's halo into `dst` (both halves); returns the mask of rounds whose voxel lies outside the volume
template <int TX>
__device__ __forceinline__ unsigned conv_dma_issue(const ConvArgs& a, const DmaLane<TX>& L, int z,
                                                   uint16_t* __restrict__ dst) {
  using T = ConvTile<TX>;
  const int wave = threadIdx.x >> 6;
  const long long zoff = (long long)(z - 1) * L.planeElems;  // the lane offsets count from plane z - 1
  unsigned bad = 0;
#pragma unroll
  for (int r = 0; r < T::NR; ++r) {
    const bool ok = ((L.okxy >> r) & 1u) && (unsigned)(z + L.hz[r] - 1) < (unsigned)a.D;
    // a voxel outside the volume loads the sample's first voxel (a valid address) and is zeroed afterwards
    const uint16_t* p = L.nbase + (ok ? zoff + L.off[r] : 0ll);
    bad |= ok ? 0u : 1u << r;
#pragma unroll
    for (int h = 0; h < 2; ++h)
      __builtin_amdgcn_global_load_lds((const void*)(p + h * 8),
                                       (__attribute__((address_space(3))) void*)(dst + (h * T::HVP + wave * 64 +
                                                                                       256 * r) * 8),
                                       16, 0, 0);
  }
  return bad;
}

// after the lane's DMA has landed: zeros into the slots of its voxels outside the volume
template <int TX>
__device__ __forceinline__ void conv_dma_fix(unsigned bad, uint16_t* __restrict__ dst) {
  using T = ConvTile<TX>;
  if (!bad) return;
#pragma unroll
  for (int r = 0; r < T::NR; ++r) {
    if (!((bad >> r) & 1u)) continue;
#pragma unroll
    for (int h = 0; h < 2; ++h)
      *reinterpret_cast<uint4*>(dst + (h * T::HVP + (int)threadIdx.x + 256 * r) * 8) = make_uint4(0u, 0u, 0u, 0u);
  }
}

// One plane's MFMAs: the next plane's DMA into `nxt` (when `more`), then the products of plane z from `cur` with
// the weights `w` into acc.  The LDS buffers are distinct restrict parameters, so that the compiler can tell the
// DMA's writes from this plane's reads and does not wait for the DMA before them.  The loop is unrolled and every
// fragment address is one of four per-lane bases plus a constant (the A bases differ by the odd tap's offset
// kind, the B base by the lane's tap parity), so the loop issues no address arithmetic.  Returns the mask of the
// next plane's voxels outside the volume.
template <int TX>
__device__ __forceinline__ unsigned conv_dma_mfma(const ConvArgs& a, const DmaLane<TX>& L, int z, bool more,
                                                  f32x4* acc, const uint16_t* __restrict__ cur,
                                                  uint16_t* __restrict__ nxt, const uint16_t* __restrict__ w) {
  using T = ConvTile<TX>;
  static_assert(TX == 16 && T::pairs_ok(), "one 16-voxel segment per tile row, four odd-tap offsets");
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int i = lane & 15, g = lane >> 4;
  const int hi = g >> 1;       // tap of the pair carried by this lane
  const int hsel = g & 1;      // the lane's channel half
  const unsigned bad = more ? conv_dma_issue<TX>(a, L, z + 1, nxt) : 0u;
  // A: segment m of this wave at halo row 4 wave + m, voxel i; B: output channel i of the lane's tap
  const int lb = hsel * T::HVP + 4 * wave * T::HX + i;
  const uint16_t* ab[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) ab[k] = cur + (lb + hi * T::DOFF[k]) * 8;
  const uint16_t* wb = w + (hsel * 28 * T::WC + hi * T::WC + i) * 8;
#pragma unroll
  for (int m = 0; m < 4; ++m) acc[m] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int p = 0; p < 14; ++p) {
    const bf16x8 bfr = *reinterpret_cast<const bf16x8*>(wb + 2 * p * T::WC * 8);
    const uint16_t* a0 = ab[T::dkind(p)] + T::shift(2 * p) * 8;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const bf16x8 afr = *reinterpret_cast<const bf16x8*>(a0 + m * T::HX * 8);
      acc[m] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(afr, bfr, acc[m], 0, 0, 0);
    }
  }
  return bad;
}

// bias (read before any DMA: a load issued after one would wait for it), ReLU and bf16 stores of plane z: voxel
// 4g + r of the segment, channel i.  With a pooled output, pool[ky][kx] (the lane's 2 x 2 output voxels of the
// plane pair) sums the stored values in the order of k_pool_fwd: z, then y, then x (even plane first: rows
// m = 2 ky, 2 ky + 1, voxels r = 2 kx, 2 kx + 1), so the fused pooling rounds exactly as the pooling kernel.
template <int TX>
__device__ __forceinline__ void conv_dma_store(const ConvArgs& a, const DmaLane<TX>& L, int n, int z,
                                               const f32x4* acc, float bias, float (*pool)[2]) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int i = lane & 15, g = lane >> 4;
  uint16_t* outPlane = a.out + (((size_t)n * a.D + z) * a.H) * a.W * a.out_cs + a.out_co + i;
  const bool pooling = a.pool_out != nullptr;
  if (pooling && !(z & 1)) {
#pragma unroll
    for (int k = 0; k < 4; ++k) pool[k >> 1][k & 1] = 0.f;
  }
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    const int y = L.y0 + 4 * wave + m;
    const int xs = L.x0 + 4 * g;
    if (y < a.H) {
      uint16_t* orow = outPlane + ((size_t)y * a.W + xs) * a.out_cs;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        if (xs + r >= a.W) continue;
        float v = acc[m][r] + bias;
        if (a.relu) v = fmaxf(v, 0.f);
        const uint16_t b = f2bf_rne(v);
        orow[r * a.out_cs] = b;
        if (pooling) pool[m >> 1][r >> 1] += __uint_as_float((uint32_t)b << 16);
      }
    }
  }
  if (pooling && (z & 1)) {
    // the plane pair is complete: output voxel (z / 2, y / 2, x / 2) of the pooled tensor (H, W even)
    const int H2 = a.H / 2, W2 = a.W / 2;
    uint16_t* pp = a.pool_out + (((size_t)n * (a.D / 2) + z / 2) * H2) * W2 * a.pool_cs + a.pool_co + i;
#pragma unroll
    for (int ky = 0; ky < 2; ++ky) {
      const int y2 = (L.y0 + 4 * wave) / 2 + ky;
#pragma unroll
      for (int kx = 0; kx < 2; ++kx) {
        const int x2 = (L.x0 + 4 * g) / 2 + kx;
        if (y2 < H2 && x2 < W2) pp[((size_t)y2 * W2 + x2) * a.pool_cs] = f2bf_rne(pool[ky][kx] * 0.125f);
      }
    }
  }
}

template <int TX>
__global__ __launch_bounds__(256) void k_conv_tile_dma(ConvArgs a, int zp) {
  using T = ConvTile<TX>;
  __shared__ __attribute__((aligned(16))) uint16_t tileA[2 * T::HVP * 8];
  __shared__ __attribute__((aligned(16))) uint16_t tileB[2 * T::HVP * 8];
  __shared__ __attribute__((aligned(16))) uint16_t wl[28 * T::WC * 16];

  const int i = threadIdx.x & 15;
  // work-group -> (n, z group, y block, x block), XCD-remapped as in k_conv_tile
  const int nbx = (a.W + TX - 1) / TX, nby = (a.H + T::TY - 1) / T::TY, nzg = (a.D + zp - 1) / zp;
  const int nbt = (int)gridDim.x;
  int b = nbt % 8 == 0 ? (int)(blockIdx.x % 8) * (nbt / 8) + (int)(blockIdx.x / 8) : (int)blockIdx.x;
  const int bx = b % nbx; b /= nbx;
  const int by = b % nby; b /= nby;
  const int zg = b % nzg, n = b / nzg;
  const int zs = zg * zp, ze = min(a.D, zs + zp);

  DmaLane<TX> L;
  L.planeElems = (long long)a.H * a.W * a.in_cs;
  L.nbase = a.in + (size_t)n * a.D * L.planeElems;
  L.x0 = bx * TX;
  L.y0 = by * T::TY;
  L.okxy = 0;
  {
    // the lane's halo voxels v = threadIdx.x + 256 r, decomposed incrementally; offsets fit 32 bits (at most
    // three planes, checked at launch)
    constexpr int DX = 256 % T::HX, DY = (256 / T::HX) % T::HY, DZ = 256 / (T::HX * T::HY);
    const int v0 = threadIdx.x;
    int hx = v0 % T::HX, hy = (v0 / T::HX) % T::HY, hz = v0 / (T::HX * T::HY);
#pragma unroll
    for (int r = 0; r < T::NR; ++r) {
      if (r) {
        hx += DX;
        const int cy = hx >= T::HX;
        hx -= cy ? T::HX : 0;
        hy += DY + cy;
        const int cz = hy >= T::HY;
        hy -= cz ? T::HY : 0;
        hz += DZ + cz;
      }
      const int xx = L.x0 + hx - 1, yy = L.y0 + hy - 1;
      const bool ok = v0 + 256 * r < T::HV && (unsigned)xx < (unsigned)a.W && (unsigned)yy < (unsigned)a.H;
      L.okxy |= ok ? 1u << r : 0u;
      L.off[r] = ok ? ((hz * a.H + yy) * a.W + xx) * a.in_cs : 0;
      L.hz[r] = hz;
    }
  }

  // weights [28 taps][16 cout][16 cin] as two halves, staged once through registers
  {
    constexpr int NW = (28 * T::WC * 2 + 255) / 256;
    uint4 wbuf[NW];
    const int half = threadIdx.x & 1;
#pragma unroll
    for (int k = 0; k < NW; ++k) {
      const int e = threadIdx.x + 256 * k;
      const int c = (e >> 1) % T::WC, tap = (e >> 1) / T::WC;
      const bool ok = tap < 27;
      wbuf[k] = *reinterpret_cast<const uint4*>(a.w + ((ok ? tap : 0) * a.cout + c) * a.cin + half * 8);
      if (!ok) wbuf[k] = make_uint4(0u, 0u, 0u, 0u);
    }
#pragma unroll
    for (int k = 0; k < NW; ++k) {
      const int e = threadIdx.x + 256 * k;
      if (e < 28 * T::WC * 2) *reinterpret_cast<uint4*>(wl + ((e & 1) * 28 * T::WC + (e >> 1)) * 8) = wbuf[k];
    }
  }
  const float bias = a.bias ? a.bias[i] : 0.f;
  unsigned bad = conv_dma_issue<TX>(a, L, zs, tileA);
  __builtin_amdgcn_s_waitcnt(0);
  conv_dma_fix<TX>(bad, tileA);
  __syncthreads();
  // Planes alternate between the buffers.  A step: the next plane's DMA, this plane's MFMAs, the wait for this
  // lane's DMA (and for the stores of the plane before, issued a step ago), the zeros, a barrier (the next plane is
  // in LDS for every wave, every wave is done with the buffer the step after refills), and only then this plane's
  // stores, which drain during the next step.
  f32x4 acc[4];
  float pool[2][2];
  for (int z = zs; z < ze; ++z) {
    const bool more = z + 1 < ze;
    const bool odd = (z - zs) & 1;
    uint16_t* cur = odd ? tileB : tileA;
    uint16_t* nxt = odd ? tileA : tileB;
    bad = conv_dma_mfma<TX>(a, L, z, more, acc, cur, nxt, wl);
    if (more) {
      __builtin_amdgcn_s_waitcnt(0);
      conv_dma_fix<TX>(bad, nxt);
      __syncthreads();
    }
    conv_dma_store<TX>(a, L, n, z, acc, bias, pool);
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

// the 16 -> 16 single-chunk layers of large launches take the DMA kernel (a pooled output needs even dimensions
// and the work-groups' even-aligned plane pairs, which kConvZPDma gives)
template <int TX>
bool takes_dma(const ConvArgs& a) {
  constexpr int TY = 256 / TX;
  const long long tiles = (long long)((a.W + TX - 1) / TX) * ((a.H + TY - 1) / TY) * a.D * a.N;
  return a.cin == 16 && a.cout == 16 && a.up_c == 0 && tiles >= kConvZPTiles &&
         (!a.pool_out || !((a.D | a.H | a.W) & 1));
}
static_assert(kConvZPDma % 2 == 0, "plane pairs of the fused pooling stay in one work-group");

template <int TX>
void launch_tx(const ConvArgs& a, hipStream_t s) {
  constexpr int TY = 256 / TX;
  if (takes_dma<TX>(a)) {
    const unsigned nb = (unsigned)((a.W + TX - 1) / TX) * (unsigned)((a.H + TY - 1) / TY) *
                        (unsigned)((a.D + kConvZPDma - 1) / kConvZPDma) * a.N;
    hipLaunchKernelGGL((k_conv_tile_dma<TX>), dim3(nb), dim3(256), 0, s, a, kConvZPDma);
    return;
  }
  if ((a.cout / 16) % 2 == 0)
    launch_tile<TX, 2>(a, s);
  else
    launch_tile<TX, 1>(a, s);
}

}  // namespace

bool conv_fuses_pool(const ConvArgs& a) {
  return a.pool_out && a.taps == 27 && a.in_cs % 16 == 0 && takes_dma<16>(a) &&
         3ll * a.H * a.W * a.in_cs < (1ll << 31);
}

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
