// prep.hip -- input preparation for the deconvolution on the GPU: affine resampling of a view into the
// deconvolution frame, its blending weights, and the normalisation of the weights across views.
//
// References (src/main/java/net/preibisch/mvrecon/process/):
//   fusion/transformed/TransformView.java (transformView, linear or nearest interpolation),
//   fusion/transformed/TransformedInputRandomAccess.java:61-84 + AbstractTransformedImgRandomAccess.java
//     (world -> input pixel coordinates by the inverse affine; strictly-inside test; minValue clamp;
//     outside value), AbstractTransformedIntervalRandomAccess.intersectsLinearInterpolation,
//   fusion/transformed/TransformWeight.transformBlending + weights/BlendingRealRandomAccess.java:46-130,
//   deconvolution/normalization/NormalizingRandomAccess.java:76-108, 183-214 (hard / smooth weights, OSEM),
//   psf/PSFExtraction.java:367-451 (transformPSF / transform: odd-sized, centre-preserving PSF resampling).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "prep.hpp"

namespace mvr {

namespace {

// world -> input pixel coordinates: t = A s + b (the inverse of the view's registration)
__device__ __forceinline__ void apply_inv(const Affine3& a, double sx, double sy, double sz, double& tx, double& ty,
                                          double& tz) {
  tx = a.m[0] * sx + a.m[1] * sy + a.m[2] * sz + a.m[3];
  ty = a.m[4] * sx + a.m[5] * sy + a.m[6] * sz + a.m[7];
  tz = a.m[8] * sx + a.m[9] * sy + a.m[10] * sz + a.m[11];
}

// input pixel value, optionally intensity-adjusted (IntensityAdjuster.convert: (float)(v * m00 + m01))
__device__ __forceinline__ double pix(const float* p, bool adjust, double a, double b) {
  return adjust ? (double)(float)((double)*p * a + b) : (double)*p;
}

// TransformedInputRandomAccess.get at input location t: strictly inside (0, n - 1) in every dimension ->
// linear or nearest interpolation clamped to >= minValue; otherwise the outside value
__device__ __forceinline__ float sample_view(const float* __restrict__ img, int nx, int ny, int nz, double tx,
                                             double ty, double tz, int hasMin, float minValue, float outside,
                                             int linear, bool adjust = false, double a = 1.0, double b = 0.0) {
  if (!(tx > 0.0 && ty > 0.0 && tz > 0.0 && tx < nx - 1 && ty < ny - 1 && tz < nz - 1)) return outside;
  double val;
  if (linear) {
    const int x0 = (int)floor(tx), y0 = (int)floor(ty), z0 = (int)floor(tz);
    const double fx = tx - x0, fy = ty - y0, fz = tz - z0;
    const float* p = img + ((size_t)z0 * ny + y0) * nx + x0;
    const size_t sy = nx, sz = (size_t)nx * ny;
    const double c00 = (1 - fx) * pix(p, adjust, a, b) + fx * pix(p + 1, adjust, a, b);
    const double c10 = (1 - fx) * pix(p + sy, adjust, a, b) + fx * pix(p + sy + 1, adjust, a, b);
    const double c01 = (1 - fx) * pix(p + sz, adjust, a, b) + fx * pix(p + sz + 1, adjust, a, b);
    const double c11 = (1 - fx) * pix(p + sz + sy, adjust, a, b) + fx * pix(p + sz + sy + 1, adjust, a, b);
    val = (1 - fz) * ((1 - fy) * c00 + fy * c10) + fz * ((1 - fy) * c01 + fy * c11);
  } else {
    // nearest neighbour: the closest pixel, ties rounded up (floor(t + 0.5))
    const int xi = (int)floor(tx + 0.5), yi = (int)floor(ty + 0.5), zi = (int)floor(tz + 0.5);
    val = pix(img + ((size_t)zi * ny + yi) * nx + xi, adjust, a, b);
  }
  float v = (float)val;
  if (hasMin) v = fmaxf(minValue, v);
  return v;
}

__global__ void k_transform_view(const float* __restrict__ img, int3 in, Affine3 inv, long3 off, int3 outDim,
                                 int hasMin, float minValue, float outside, int linear, float* __restrict__ out) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y, z = blockIdx.z;
  if (x >= outDim.x) return;
  double tx, ty, tz;
  apply_inv(inv, (double)(x + off.x), (double)(y + off.y), (double)(z + off.z), tx, ty, tz);
  out[((size_t)z * outDim.y + y) * outDim.x + x] =
      sample_view(img, in.x, in.y, in.z, tx, ty, tz, hasMin, minValue, outside, linear);
}

__constant__ double c_blendLut[1001];

// BlendingRealRandomAccess.computeWeight at the (float) input location l, interval [0, dm]
__device__ __forceinline__ float blend_weight(const float l[3], const int dm[3], const float bo[3],
                                              const float bl[3]) {
  float dist[3];
  float w = 1.f;
  bool zero = false;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    dist[d] = fminf(__fsub_rn(l[d], bo[d]), __fsub_rn(__fsub_rn((float)dm[d], l[d]), bo[d]));
    zero |= dist[d] <= 0.f;
  }
  if (!zero) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const float rel = __fdiv_rn(dist[d], bl[d]);
      if (rel < 1.f) w = (float)((double)w * c_blendLut[(int)((double)rel * 1000.0 + 0.5)]);
    }
  } else {
    w = 0.f;
  }
  return w;
}

__global__ void k_transform_blending(int3 dimMinus1, float3 border, float3 blending, Affine3 inv, long3 off,
                                     int3 outDim, float* __restrict__ out) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y, z = blockIdx.z;
  if (x >= outDim.x) return;
  double tx, ty, tz;
  apply_inv(inv, (double)(x + off.x), (double)(y + off.y), (double)(z + off.z), tx, ty, tz);
  const float l[3] = {(float)tx, (float)ty, (float)tz};
  const int dm[3] = {dimMinus1.x, dimMinus1.y, dimMinus1.z};
  const float bo[3] = {border.x, border.y, border.z};
  const float bl[3] = {blending.x, blending.y, blending.z};
  out[((size_t)z * outDim.y + y) * outDim.x + x] = blend_weight(l, dm, bo, bl);
}

// ProcessInputImages.fuseGroups (deconvolution/util/ProcessInputImages.java:309-396) for one group, fused in
// one pass: each view is resampled (TransformView, minValueImg 1, outside 0, linear) after its intensity
// adjustment; img = FusedRandomAccess AVG (fusion/transformed/FusedRandomAccess.java:67-88: views whose fusion
// weight is 0 are skipped, sum I w / sum w in double, 0 when no weight) and weight = CombineWeightsSum
// (weightcombination/CombineWeightsSumRandomAccess.java:40-50: (float) of the double sum of decon weights)
__global__ void k_fuse_group(const FuseView* __restrict__ views, int V, long3 off, int3 outDim,
                             float* __restrict__ img, float* __restrict__ weight) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y, z = blockIdx.z;
  if (x >= outDim.x) return;
  const double sx = (double)(x + off.x), sy = (double)(y + off.y), sz = (double)(z + off.z);
  double sumI = 0.0, sumW = 0.0, sumD = 0.0;
  for (int j = 0; j < V; ++j) {
    const FuseView& v = views[j];
    double tx, ty, tz;
    apply_inv(v.inv, sx, sy, sz, tx, ty, tz);
    const float l[3] = {(float)tx, (float)ty, (float)tz};
    const int dm[3] = {v.in[0] - 1, v.in[1] - 1, v.in[2] - 1};
    sumD += v.deconBlend ? (double)blend_weight(l, dm, v.dBorder, v.dRange) : 1.0;
    const float wf = v.fusionBlend ? blend_weight(l, dm, v.fBorder, v.fRange) : 1.f;
    if (wf == 0.f) continue;
    const float iv = sample_view(v.img, v.in[0], v.in[1], v.in[2], tx, ty, tz, 1, 1.0f, 0.0f, 1, v.adjust != 0,
                                 v.adjA, v.adjB);
    sumI += (double)iv * (double)wf;
    sumW += (double)wf;
  }
  const size_t o = ((size_t)z * outDim.y + y) * outDim.x + x;
  img[o] = sumW > 0.0 ? (float)(sumI / sumW) : 0.f;
  weight[o] = (float)sumD;
}

// PSFExtraction.transform (psf/PSFExtraction.java:411-451): out[p] = linear interpolation of the zero-extended
// PSF at inverse(p + offset)
__global__ void k_transform_psf(const float* __restrict__ psf, int3 in, Affine3 inv, double3 off, int3 outDim,
                                float* __restrict__ out) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y, z = blockIdx.z;
  if (x >= outDim.x) return;
  double tx, ty, tz;
  apply_inv(inv, x + off.x, y + off.y, z + off.z, tx, ty, tz);
  const int x0 = (int)floor(tx), y0 = (int)floor(ty), z0 = (int)floor(tz);
  const double fx = tx - x0, fy = ty - y0, fz = tz - z0;
  double val = 0.0;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const int xi = x0 + (c & 1), yi = y0 + ((c >> 1) & 1), zi = z0 + (c >> 2);
    if ((unsigned)xi < (unsigned)in.x && (unsigned)yi < (unsigned)in.y && (unsigned)zi < (unsigned)in.z) {
      const double wgt = ((c & 1) ? fx : 1 - fx) * (((c >> 1) & 1) ? fy : 1 - fy) * ((c >> 2) ? fz : 1 - fz);
      val += wgt * (double)psf[((size_t)zi * in.y + yi) * in.x + xi];
    }
  }
  out[((size_t)z * outDim.y + y) * outDim.x + x] = (float)val;
}

struct WeightPtrs { float* p[kPrepMaxViews]; };

// NormalizingRandomAccess.get for every view at once (in place)
__global__ void k_normalize_weights(WeightPtrs w, int V, size_t n, double osem, int smooth, float maxDiffRange,
                                    float scalingRange) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    double sumW = 0.0;
    for (int v = 0; v < V; ++v) sumW += fmin(1.0, (double)w.p[v][i]);
    for (int v = 0; v < V; ++v) {
      const float my = (float)fmin(1.0, (double)w.p[v][i]);
      double val;
      if (smooth) {
        if (sumW <= 0) {
          val = 0.0;
        } else {
          const float ideal = (float)((double)my / sumW);
          const float diff = __fsub_rn(my, ideal);
          const float yv = fmaxf(0.f, __fmul_rn(__fsub_rn(maxDiffRange, fabsf(diff)), __fdiv_rn(1.0f, maxDiffRange)));
          const float scale = __fmul_rn(__fmul_rn(yv, my), scalingRange);
          val = (double)__fsub_rn(fminf(my, ideal), scale);
        }
      } else if (sumW > 1) {
        val = (double)(float)((double)my / sumW);
      } else {
        val = (double)my;
      }
      w.p[v][i] = (float)fmin(1.0, val * osem);
    }
  }
}

}  // namespace

int launch_transform_view(const float* img, const int inXYZ[3], const Affine3& inv, const long long off[3],
                          const int outXYZ[3], int hasMin, float minValue, float outside, int linear, float* out,
                          hipStream_t s) {
  const dim3 grid((outXYZ[0] + 255) / 256, outXYZ[1], outXYZ[2]);
  hipLaunchKernelGGL(k_transform_view, grid, dim3(256), 0, s, img, make_int3(inXYZ[0], inXYZ[1], inXYZ[2]), inv,
                     make_long3(off[0], off[1], off[2]), make_int3(outXYZ[0], outXYZ[1], outXYZ[2]), hasMin, minValue,
                     outside, linear, out);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_transform_psf(const float* psf, const int inXYZ[3], const Affine3& inv, const double off[3],
                         const int outXYZ[3], float* out, hipStream_t s) {
  const dim3 grid((outXYZ[0] + 255) / 256, outXYZ[1], outXYZ[2]);
  hipLaunchKernelGGL(k_transform_psf, grid, dim3(256), 0, s, psf, make_int3(inXYZ[0], inXYZ[1], inXYZ[2]), inv,
                     make_double3(off[0], off[1], off[2]), make_int3(outXYZ[0], outXYZ[1], outXYZ[2]), out);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int upload_blend_lut(const double* lut) {
  return hipMemcpyToSymbol(HIP_SYMBOL(c_blendLut), lut, sizeof(double) * 1001) == hipSuccess ? 0 : -1;
}

int launch_transform_blending(const int inXYZ[3], const float border[3], const float blending[3], const Affine3& inv,
                              const long long off[3], const int outXYZ[3], float* out, hipStream_t s) {
  const dim3 grid((outXYZ[0] + 255) / 256, outXYZ[1], outXYZ[2]);
  hipLaunchKernelGGL(k_transform_blending, grid, dim3(256), 0, s, make_int3(inXYZ[0] - 1, inXYZ[1] - 1, inXYZ[2] - 1),
                     make_float3(border[0], border[1], border[2]), make_float3(blending[0], blending[1], blending[2]),
                     inv, make_long3(off[0], off[1], off[2]), make_int3(outXYZ[0], outXYZ[1], outXYZ[2]), out);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_fuse_group(const FuseView* views, int V, const long long off[3], const int outXYZ[3], float* img,
                      float* weight, hipStream_t s) {
  const dim3 grid((outXYZ[0] + 255) / 256, outXYZ[1], outXYZ[2]);
  hipLaunchKernelGGL(k_fuse_group, grid, dim3(256), 0, s, views, V, make_long3(off[0], off[1], off[2]),
                     make_int3(outXYZ[0], outXYZ[1], outXYZ[2]), img, weight);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_normalize_weights(float* const* w, int V, size_t n, double osem, int smooth, float maxDiffRange,
                             float scalingRange, hipStream_t s) {
  WeightPtrs p{};
  for (int v = 0; v < V && v < kPrepMaxViews; ++v) p.p[v] = w[v];
  const int nb = (int)std::max<size_t>(1, std::min<size_t>((n + 255) / 256, 8192));
  hipLaunchKernelGGL(k_normalize_weights, dim3(nb), dim3(256), 0, s, p, V, n, osem, smooth, maxDiffRange,
                     scalingRange);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace mvr
