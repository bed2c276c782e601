// noalloc.hip -- mvr_fft_convolve_noalloc: one FFT convolution (util.FFTConvolution semantics, as
// mvr_fft_convolve) that allocates no device memory.
//
// The twiddle tables, the kernel spectrum and the work spectrum live in `scratch`, a static device array of
// MVR_NOALLOC_SCRATCH_BYTES that every GPU the library runs on holds until the process ends.  In host code the
// array's name is only a stand-in for the symbol: its device address is looked up once per device
// (hipGetSymbolAddress with that device current) and that address is what the kernels receive.  Nothing in the
// array has to start at zero: every byte a call reads was written earlier in the same call (twiddles by
// k_twiddles, both spectra by the forward x pass over every row and half-spectrum column).  All calls on one GPU
// share its array, so calls that overlap on one GPU race; the caller serialises them (one stream).
#include <mutex>
#include <string>

#include "../../include/mvr_hip.h"
#include "host.hpp"

static __device__ __align__(256) unsigned char scratch[MVR_NOALLOC_SCRATCH_BYTES];

namespace mvr {
namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
unsigned char* g_scratch[kMaxDevices] = {};  // device address of `scratch` per device, set on first use

// the device address of `scratch` on `dev` (which must be current)
int scratch_address(int dev, unsigned char** out) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr) {
      set_error(std::string("scratch address lookup failed: ") + hipGetErrorString(e));
      return -1;
    }
    g_scratch[dev] = static_cast<unsigned char*>(p);
  }
  *out = g_scratch[dev];
  return 0;
}

// exp(-2 pi i k / period), k < n (the same table Plan::create uploads from the host)
__global__ void k_twiddles(float2* __restrict__ out, int n, int period) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) {
    double s, c;
    sincospi(-2.0 * (double)k / (double)period, &s, &c);
    out[k] = make_float2((float)c, (float)s);
  }
}

void launch_twiddles(float2* out, int n, int period, hipStream_t s) {
  hipLaunchKernelGGL(k_twiddles, dim3((n + 255) / 256), dim3(256), 0, s, out, n, period);
}

// device memory of `dev` (device or managed); host memory would fault the kernels
bool on_device(const void* p, int dev) {
  hipPointerAttribute_t a{};
  if (hipPointerGetAttributes(&a, p) != hipSuccess) {
    (void)hipGetLastError();  // an unknown (pageable host) pointer: clear the error it leaves
    return false;
  }
  return (a.type == hipMemoryTypeDevice && a.device == dev) || a.type == hipMemoryTypeManaged || a.isManaged;
}

constexpr size_t align256(size_t b) { return (b + 255) / 256 * 256; }

}  // namespace
}  // namespace mvr

using namespace mvr;

extern "C" int mvr_fft_convolve_noalloc(int dev, const float* img, const int n[3], int ext, float extValue,
                                        const float* kernel, const int k[3], float* out, void* stream) {
  if (!img || !kernel || !out || !n || !k) { set_error("null argument"); return -1; }
  if (ext != MVR_EXT_MIRROR_SINGLE && ext != MVR_EXT_VALUE) { set_error("unknown extension"); return -1; }
  if (dev < 0 || dev >= kMaxDevices) { set_error("bad device"); return -1; }
  // geometry and the scratch layout first, on the host: sizes beyond the array are refused before device work
  Plan plan;
  if (plan.create_geometry(n, k)) return -1;
  const size_t oWr = align256((size_t)plan.M * 8), oTwY = oWr + align256((size_t)(plan.M + 1) * 8);
  const size_t oTwZ = oTwY + align256((size_t)plan.dom.P[1] * 8), oSpec = oTwZ + align256((size_t)plan.dom.P[2] * 8);
  const size_t oWork = oSpec + align256(plan.ncpx() * 8), total = oWork + align256(plan.ncpx() * 8);
  if (total > (size_t)MVR_NOALLOC_SCRATCH_BYTES) {
    set_error("mvr_fft_convolve_noalloc: the FFT grid " + std::to_string(plan.dom.P[0]) + "x" +
              std::to_string(plan.dom.P[1]) + "x" + std::to_string(plan.dom.P[2]) + " needs " +
              std::to_string(total) + " bytes of scratch, more than the " +
              std::to_string((long long)MVR_NOALLOC_SCRATCH_BYTES) + " available (use mvr_fft_convolve)");
    return -1;
  }
  MVR_HIP(hipSetDevice(dev));
  if (!on_device(img, dev) || !on_device(kernel, dev) || !on_device(out, dev)) {
    set_error("mvr_fft_convolve_noalloc: img, kernel and out must be device memory of the given device");
    return -1;
  }
  unsigned char* base = nullptr;
  if (scratch_address(dev, &base)) return -1;
  hipStream_t s = static_cast<hipStream_t>(stream);
  plan.twM = reinterpret_cast<float2*>(base);
  plan.wr = reinterpret_cast<float2*>(base + oWr);
  plan.twY = reinterpret_cast<float2*>(base + oTwY);
  plan.twZ = reinterpret_cast<float2*>(base + oTwZ);
  float2* spec = reinterpret_cast<float2*>(base + oSpec);
  float2* work = reinterpret_cast<float2*>(base + oWork);
  launch_twiddles(plan.twM, plan.M, plan.M, s);
  launch_twiddles(plan.wr, plan.M + 1, plan.dom.P[0], s);
  launch_twiddles(plan.twY, plan.dom.P[1], plan.dom.P[1], s);
  launch_twiddles(plan.twZ, plan.dom.P[2], plan.dom.P[2], s);
  int rc = plan.kernel_spectrum(kernel, spec, s);
  if (!rc) {
    GatherParams g{};
    g.src = img;
    for (int d = 0; d < 3; ++d) { g.N[d] = n[d]; g.org[d] = -plan.dom.m[d]; }
    g.mode = ext == MVR_EXT_MIRROR_SINGLE ? GATHER_MIRROR : GATHER_VALUE;
    g.value = extValue;
    rc = plan.convolve(g, spec, work, out, s);
  }
  // the tables belong to the scratch array: the plan must not free them
  plan.twM = plan.wr = plan.twY = plan.twZ = nullptr;
  return rc;
}
