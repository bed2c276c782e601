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
int launch_normalize_weights(float* const* w, int V, size_t n, double osem, int smooth, float maxDiffRange,
                             float scalingRange, hipStream_t s);

}  // namespace mvr
