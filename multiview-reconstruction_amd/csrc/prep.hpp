// prep.hpp -- input preparation kernels (affine resampling, blending weights, weight normalisation).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>

namespace mvr {

constexpr int kPrepMaxViews = 32;

// 3 x 4 row-major affine map: t = A s + b with m = {A row 0, b0, A row 1, b1, A row 2, b2}
struct Affine3 {
  double m[12];
};

int launch_transform_view(const float* img, const int inXYZ[3], const Affine3& inv, const long long off[3],
                          const int outXYZ[3], int hasMin, float minValue, float outside, int linear, float* out,
                          hipStream_t s);
int upload_blend_lut(const double* lut);
int launch_transform_psf(const float* psf, const int inXYZ[3], const Affine3& inv, const double off[3],
                         const int outXYZ[3], float* out, hipStream_t s);
int launch_transform_blending(const int inXYZ[3], const float border[3], const float blending[3], const Affine3& inv,
                              const long long off[3], const int outXYZ[3], float* out, hipStream_t s);
// one view of a fused group: its image, inverse registration, optional intensity adjustment and blending
struct FuseView {
  const float* img;
  int in[3];                 // image dims x, y, z
  Affine3 inv;               // world -> input pixel
  double adjA, adjB;         // intensity v -> (float)(v * adjA + adjB) when adjust != 0
  int adjust;
  int fusionBlend, deconBlend;  // 0: constant weight 1
  float fBorder[3], fRange[3], dBorder[3], dRange[3];
};

// ProcessInputImages.fuseGroups for one group: per output voxel, AVG fusion of the resampled views weighted
// by their fusion blending, and the sum of their deconvolution blending weights (views: device array)
int launch_fuse_group(const FuseView* views, int V, const long long off[3], const int outXYZ[3], float* img,
                      float* weight, hipStream_t s);
int launch_normalize_weights(float* const* w, int V, size_t n, double osem, int smooth, float maxDiffRange,
                             float scalingRange, hipStream_t s);

}  // namespace mvr
