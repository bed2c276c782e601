// engine.hpp -- shared declarations of the HIP deconvolution engine.
//
// Geometry of one convolution plan ("domain"):
//   * the output region is D[0] x D[1] x D[2] voxels of a source volume,
//     starting at volume coordinate o;
//   * the FFT grid is P[0] x P[1] x P[2] (P[d] >= D[d] + K[d] - 1); FFT
//     position p corresponds to volume coordinate o + p - m, m[d] = K[d]-1-K[d]/2
//     (the left reach of the kernel), so the output region is p in [m, m + D);
//   * the half spectrum is stored [z][y][hx], hx < P[0]/2 + 1, with row pitch HxP.
// A linear convolution evaluated this way is exact on the output region for
// any extension of the input outside it (util/FFTConvolution.java:508-543 pads
// to at least N + K - 1 for the same reason).
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

namespace mvr {

struct Dom {
  int P[3];
  int HxP;
  int D[3];
  int m[3];
  // Batched blocks: when btab is set, one launch runs nb blocks that share P, HxP and m.  Entry b of btab
  // (device memory, 6 ints) is {o[3], D[3]}: the volume coordinate of the block's output region and its size
  // (at most D above); the block's spectrum starts at b * blockStride complex values.  The launchers set
  // wgPerBlock (work-groups per block of an x pass).
  const int* btab = nullptr;
  int nb = 1;
  int wgPerBlock = 0;
  size_t blockStride = 0;
};
constexpr int kBlockTabInts = 6;

enum GatherMode : int { GATHER_MIRROR = 0, GATHER_VALUE = 1, GATHER_KERNEL = 2 };

// Source of the forward transform: value at FFT position p is
//   MIRROR: src[mirror(org + p)]                (extendMirrorSingle)
//   VALUE : src[org + p] inside the volume, `value` outside (extendValue / extendZero)
//   KERNEL: K[(p + K/2) mod P] if < K, else 0    (kernel centre moved to the origin)
struct GatherParams {
  const float* src;
  int N[3];
  int org[3];
  int mode;
  float value;
  // MIRROR only: transform only the rows (y, z) that are not mirror copies, FFT positions [-org, -org + N)
  // per axis (the passes after it read the copies from their sources)
  int uniqueRows;
};

// Inverse-transform epilogues.  `o` is the volume coordinate of the output
// region's first voxel; N are the dims of the volume that img/psi/w index.
// The quotient is formed on the output region widened by qlo / qhi voxels per side (the reach of the second
// kernel, for domains whose neighbourhood lies inside the volume: blocks), within the volume; 1 elsewhere.
// skipOnes: rows outside the quotient region (a row of ones, whose x spectrum is P0 at frequency 0 and 0
// elsewhere) are not written; the z pass that reads the result puts them in itself (StridedGeom ones*), and
// the work-groups are dealt over the quotient rows only: grid rows (y, z) in [rowLo, rowLo + rowN) per axis
// (k = 1, 2), which are exactly the rows the quotient region selects.
struct RatioParams {
  const float* img;
  int N[3];
  int o[3];
  int qlo[3];
  int qhi[3];
  int skipOnes;
  int rowLo[3], rowN[3];
};

struct UpdateParams {
  float* psi;
  const float* w;
  int N[3];
  int o[3];
  double lambda;
  float minValue;
  float maxI;
  double2* partial;  // per work-group {sum change, max change}
};

struct StoreParams {
  float* out;        // compact D[0]*D[1]*D[2] array
};

typedef void (*XFwdFn)(const Dom&, const GatherParams&, float2* S, const float2* twM, const float2* wr,
                       hipStream_t s);
typedef void (*XRatioFn)(const Dom&, const RatioParams&, const float2* Sin, float2* Sout, const float2* twM,
                         const float2* wr, hipStream_t s);
typedef int (*XUpdateFn)(const Dom&, const UpdateParams&, const float2* S, const float2* twM, const float2* wr,
                         hipStream_t s);  // returns #work-groups (= #partials)
typedef void (*XStoreFn)(const Dom&, const StoreParams&, const float2* S, const float2* twM, const float2* wr,
                         hipStream_t s);
// strided passes: y (plane = z) or z (plane = y)
struct StridedGeom {
  size_t planeStride;  // complex elements between consecutive planes
  size_t lineStride;   // complex elements between consecutive pencil elements
  int ncols;           // valid columns (hx) per line
  int nplanes;
  int width = 0;       // tile width in columns (8 or 16; 0 = the default for the length)
  // batched blocks: planes [b * blockPlanes, (b + 1) * blockPlanes) belong to block b, whose data starts at
  // b * blockStride; the kernel spectrum is shared (indexed by the plane within the block).  0 = one block.
  int blockPlanes = 0;
  size_t blockStride = 0;
  // Only what the next pass reads: planes [plane0, plane0 + nplanes) (per block) are transformed, and the inverse
  // modes (1, 2) store only the pencil positions [storeLo, storeHi).  The defaults cover the whole grid.
  int plane0 = 0;
  int storeLo = 0, storeHi = 1 << 30;
  // Forward mode over data whose rows outside planes [onesPlaneLo, onesPlaneHi) x positions [onesLo, onesHi) are
  // spectra of rows (or planes) of ones, never written: the pass reads none of them and uses onesDC at column 0,
  // 0 at the other columns; with onesDCPlane >= 0 onesDC is on that plane only, 0 on the others.  Off when onesDC
  // is 0.
  float onesDC = 0.f;
  int onesPlaneLo = 0, onesPlaneHi = 0, onesLo = 0, onesHi = 0, onesDCPlane = -1;
  // Forward modes over data that is a mirror extension along the pencil (mirN > 0): position n holds what
  // position mirM + mirror(n - mirM, mirN) holds, and only those positions were written; the loads read them
  // (byte offsets from the plane up to L lineStride 8 must fit 32 bits).
  int mirM = 0, mirN = 0;
  // timing: when set, the launch records the kernel's own start and stop in these events
  hipEvent_t evStart = nullptr, evStop = nullptr;
};
// mode: 0 forward, 1 inverse, 2 forward * spec * inverse (spec layout = data layout),
// 3 forward then scale
typedef void (*StridedFn)(const StridedGeom&, int mode, const float2* in, float2* out, const float2* spec,
                          float scale, const float2* tw, hipStream_t s);

XFwdFn get_x_fwd(int M);
XRatioFn get_x_ratio(int M);
XUpdateFn get_x_update(int M);
XStoreFn get_x_store(int M);
StridedFn get_strided(int L);
// the largest pencil position a lane of the length-L strided pass addresses with its 32-bit offset, plus one
int get_strided_span(int L);
// lengths with compiled kernels, ascending; terminated by 0
const int* strided_lengths();
const int* half_lengths();

// misc kernels (misc.hip)
void launch_fill(float* dst, size_t n, float v, hipStream_t s);
void launch_reduce_stats(const double2* partial, int n, double2* out, hipStream_t s);
// nUpdates <= 32 reductions in one launch: update v's partials at partial + v * stride, counts[v] of them
void launch_reduce_stats_batched(const double2* partial, size_t stride, const int* counts, int nUpdates, double2* out,
                                 hipStream_t s);
int launch_psi_init_partials(const float* const* imgs, int V, size_t n, double* psum, unsigned long long* pcount,
                             float* pmax, hipStream_t s);  // returns #partials
void launch_update_mul(float* psi, const float* const* T2, const float* const* w, int V, const int N[3],
                       const int o[3], const int D[3], double lambda, float minValue, float maxI, double2* partial,
                       int* nPartialOut, hipStream_t s);
void launch_copy_block_zero(const float* src, const int N[3], const long long off[3], float* dst, const int B[3],
                            hipStream_t s);
void launch_blocks_with_content(const float* w, const int N[3], const int* tab, int nblocks, int* flags,
                                hipStream_t s);
void launch_quotient(const float* img, float* blurred, size_t n, hipStream_t s);
int launch_final_values(float* psi, const float* integral, const float* weight, size_t n, double lambda,
                        float minValue, float maxI, double2* partial, int maxPartials, hipStream_t s);
void launch_scale_copy_kernel_spectrum(const float2* in, float2* out, size_t n, float scale, hipStream_t s);
void launch_rel_l2(const float* a, const float* b, size_t n, double2* partial, double2* out, hipStream_t s);
int launch_psi_approx_slice(const float* img, const int N[3], double2* partial, hipStream_t s);
void launch_fused_nonzero(const float* const* imgs, const float* const* ws, int V, size_t n, float* out,
                          hipStream_t s);
void launch_sym_conv_axis(const float* in, float* out, const int N[3], int axis, const double* half, int H,
                          hipStream_t s);

}  // namespace mvr
