// methods.cpp -- the elementwise DeconvolutionMethods over whole arrays (C ABI).
//
// The deconvolution passes fuse these operations into their FFT passes; these entry points run them alone, on
// arrays the caller holds (host or device memory of `dev`), for callers that keep the reference's own
// convolutions and for exact checks of the fused arithmetic.
#include <algorithm>
#include <vector>

#include "../../include/mvr_hip.h"
#include "host.hpp"

using namespace mvr;

namespace {

struct Buf {
  void* p = nullptr;
  Buf() = default;
  Buf(const Buf&) = delete;
  Buf& operator=(const Buf&) = delete;
  ~Buf() { if (p) (void)hipFree(p); }
  float* f() const { return static_cast<float*>(p); }
};

constexpr int kMaxPartials = 2048;

int alloc_copy(Buf& b, const float* src, size_t n, hipStream_t s) {
  MVR_HIP(hipMalloc(&b.p, std::max<size_t>(n, 1) * 4));
  if (src && n) MVR_HIP(hipMemcpyAsync(b.p, src, n * 4, hipMemcpyDefault, s));
  return 0;
}

struct Stream {
  hipStream_t s = nullptr;
  ~Stream() { if (s) (void)hipStreamDestroy(s); }
};

}  // namespace

extern "C" {

int mvr_compute_quotient(int dev, const float* img, float* psiBlurred, long long n) {
  if (!img || !psiBlurred || n < 0) { set_error("null array or negative size"); return -1; }
  MVR_HIP(hipSetDevice(dev));
  Stream st;
  MVR_HIP(hipStreamCreateWithFlags(&st.s, hipStreamNonBlocking));
  Buf di, db;
  if (alloc_copy(di, img, (size_t)n, st.s) || alloc_copy(db, psiBlurred, (size_t)n, st.s)) return -1;
  launch_quotient(di.f(), db.f(), (size_t)n, st.s);
  MVR_HIP(hipGetLastError());
  MVR_HIP(hipMemcpyAsync(psiBlurred, db.p, (size_t)n * 4, hipMemcpyDefault, st.s));
  MVR_HIP(hipStreamSynchronize(st.s));
  return 0;
}

int mvr_compute_final_values(int dev, float* psi, const float* integral, const float* weight, long long n,
                             double lambda, float minIntensity, float maxIntensity, double stats[2]) {
  if (!psi || !integral || !weight || n < 0) { set_error("null array or negative size"); return -1; }
  if (n == 0) {
    if (stats) { stats[0] = 0.0; stats[1] = -1.0; }
    return 0;
  }
  MVR_HIP(hipSetDevice(dev));
  Stream st;
  MVR_HIP(hipStreamCreateWithFlags(&st.s, hipStreamNonBlocking));
  Buf dp, di, dw, part, out;
  if (alloc_copy(dp, psi, (size_t)n, st.s) || alloc_copy(di, integral, (size_t)n, st.s) ||
      alloc_copy(dw, weight, (size_t)n, st.s))
    return -1;
  MVR_HIP(hipMalloc(&part.p, sizeof(double2) * kMaxPartials));
  MVR_HIP(hipMalloc(&out.p, sizeof(double2)));
  const int np = launch_final_values(dp.f(), di.f(), dw.f(), (size_t)n, lambda, minIntensity, maxIntensity,
                                     static_cast<double2*>(part.p), kMaxPartials, st.s);
  launch_reduce_stats(static_cast<double2*>(part.p), np, static_cast<double2*>(out.p), st.s);
  MVR_HIP(hipGetLastError());
  double2 h;
  MVR_HIP(hipMemcpyAsync(psi, dp.p, (size_t)n * 4, hipMemcpyDefault, st.s));
  MVR_HIP(hipMemcpyAsync(&h, out.p, sizeof(double2), hipMemcpyDeviceToHost, st.s));
  MVR_HIP(hipStreamSynchronize(st.s));
  if (stats) { stats[0] = h.x; stats[1] = h.y; }
  return 0;
}

int mvr_compute_final_values_mul(int dev, float* psi, int numViews, const float* const* integral,
                                 const float* const* weight, long long n, double lambda, float minIntensity,
                                 float maxIntensity, double stats[2]) {
  if (!psi || !integral || !weight || n < 0) { set_error("null array or negative size"); return -1; }
  if (numViews < 1 || numViews > 32) { set_error("numViews must be in [1, 32]"); return -1; }
  if (n >= (1ll << 31)) { set_error("at most 2^31 - 1 values per call"); return -1; }
  for (int v = 0; v < numViews; ++v)
    if (!integral[v] || !weight[v]) { set_error("null view array"); return -1; }
  if (n == 0) {
    if (stats) { stats[0] = 0.0; stats[1] = -1.0; }
    return 0;
  }
  MVR_HIP(hipSetDevice(dev));
  Stream st;
  MVR_HIP(hipStreamCreateWithFlags(&st.s, hipStreamNonBlocking));
  Buf dp, part, out;
  std::vector<Buf> di(numViews), dw(numViews);
  if (alloc_copy(dp, psi, (size_t)n, st.s)) return -1;
  std::vector<const float*> ip(numViews), wp(numViews);
  for (int v = 0; v < numViews; ++v) {
    if (alloc_copy(di[v], integral[v], (size_t)n, st.s) || alloc_copy(dw[v], weight[v], (size_t)n, st.s)) return -1;
    ip[v] = di[v].f();
    wp[v] = dw[v].f();
  }
  MVR_HIP(hipMalloc(&part.p, sizeof(double2) * kMaxPartials));
  MVR_HIP(hipMalloc(&out.p, sizeof(double2)));
  const int N[3] = {(int)n, 1, 1}, o[3] = {0, 0, 0}, D[3] = {(int)n, 1, 1};
  int np = 0;
  launch_update_mul(dp.f(), ip.data(), wp.data(), numViews, N, o, D, lambda, minIntensity, maxIntensity,
                    static_cast<double2*>(part.p), &np, st.s);
  launch_reduce_stats(static_cast<double2*>(part.p), np, static_cast<double2*>(out.p), st.s);
  MVR_HIP(hipGetLastError());
  double2 h;
  MVR_HIP(hipMemcpyAsync(psi, dp.p, (size_t)n * 4, hipMemcpyDefault, st.s));
  MVR_HIP(hipMemcpyAsync(&h, out.p, sizeof(double2), hipMemcpyDeviceToHost, st.s));
  MVR_HIP(hipStreamSynchronize(st.s));
  if (stats) { stats[0] = h.x; stats[1] = h.y; }
  return 0;
}

}  // extern "C"
