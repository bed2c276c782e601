// capi.cpp -- extern "C" entry points declared in include/mvr_hip.h.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/mvr_hip.h"
#include "host.hpp"
#include "prep.hpp"
#include "unet_host.hpp"

using namespace mvr;

struct mvr_block_thread { BlockThread t; };
struct mvr_decon { Decon d; };
struct mvr_unet { UNet n; };

#define NULLCHECK(p)                              \
  do {                                            \
    if (!(p)) {                                   \
      set_error("null handle or argument: " #p);  \
      return -1;                                  \
    }                                             \
  } while (0)

extern "C" {

const char* mvr_last_error(void) { return last_error(); }

int mvr_get_num_devices(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return -1;
  return n;
}

int mvr_get_device_name(int dev, char* name, int len) {
  NULLCHECK(name);
  hipDeviceProp_t p;
  MVR_HIP(hipGetDeviceProperties(&p, dev));
  const char* src = p.name[0] ? p.name : p.gcnArchName;
  std::strncpy(name, src, len > 0 ? len - 1 : 0);
  if (len > 0) name[len - 1] = 0;
  return 0;
}

long long mvr_get_mem_device(int dev) {
  hipDeviceProp_t p;
  if (hipGetDeviceProperties(&p, dev) != hipSuccess) return -1;
  return (long long)p.totalGlobalMem;
}

long long mvr_get_free_mem_device(int dev) {
  if (hipSetDevice(dev) != hipSuccess) return -1;
  size_t f = 0, t = 0;
  if (hipMemGetInfo(&f, &t) != hipSuccess) return -1;
  return (long long)f;
}

int mvr_get_arch_major(int dev) {
  hipDeviceProp_t p;
  if (hipGetDeviceProperties(&p, dev) != hipSuccess) return -1;
  return p.major;
}

int mvr_get_arch_minor(int dev) {
  hipDeviceProp_t p;
  if (hipGetDeviceProperties(&p, dev) != hipSuccess) return -1;
  return p.minor;
}

int mvr_fft_lengths(int* out, int cap) {
  int n = 0;
  for (const int* p = strided_lengths(); *p; ++p, ++n)
    if (out && n < cap) out[n] = *p;
  return n;
}

int mvr_fft_convolve(int dev, const float* img, const int n[3], int ext, float extValue, const float* kernel,
                     const int k[3], float* out) {
  NULLCHECK(img);
  NULLCHECK(kernel);
  NULLCHECK(out);
  if (ext != MVR_EXT_MIRROR_SINGLE && ext != MVR_EXT_VALUE) { set_error("unknown extension"); return -1; }
  MVR_HIP(hipSetDevice(dev));
  Plan plan;
  if (plan.create(n, k)) return -1;
  const size_t nv = (size_t)n[0] * n[1] * n[2], nk = (size_t)k[0] * k[1] * k[2];
  hipStream_t s = nullptr;
  float *dImg = nullptr, *dK = nullptr, *dOut = nullptr;
  float2 *spec = nullptr, *work = nullptr;
  int rc = -1;
  do {
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) || hipMalloc(&dImg, nv * 4) || hipMalloc(&dK, nk * 4) ||
        hipMalloc(&dOut, nv * 4) || hipMalloc(&spec, plan.ncpx() * 8) || hipMalloc(&work, plan.ncpx() * 8)) {
      set_error("mvr_fft_convolve: allocation failed");
      break;
    }
    if (hipMemcpyAsync(dImg, img, nv * 4, hipMemcpyDefault, s) || hipMemcpyAsync(dK, kernel, nk * 4, hipMemcpyDefault, s))
      break;
    if (plan.kernel_spectrum(dK, spec, s)) break;
    GatherParams g{};
    g.src = dImg;
    for (int d = 0; d < 3; ++d) { g.N[d] = n[d]; g.org[d] = -plan.dom.m[d]; }
    g.mode = ext == MVR_EXT_MIRROR_SINGLE ? GATHER_MIRROR : GATHER_VALUE;
    g.value = extValue;
    if (plan.convolve(g, spec, work, dOut, s)) break;
    if (hipMemcpyAsync(out, dOut, nv * 4, hipMemcpyDefault, s) || hipStreamSynchronize(s)) {
      set_error("mvr_fft_convolve: readback failed");
      break;
    }
    rc = 0;
  } while (0);
  for (void* p : {(void*)dImg, (void*)dK, (void*)dOut, (void*)spec, (void*)work})
    if (p) (void)hipFree(p);
  if (s) (void)hipStreamDestroy(s);
  plan.destroy();
  return rc;
}

int mvr_convolution3d_fft_inplace(float* im, const int* imDimZYX, const float* kernel, const int* kernelDimZYX,
                                  int dev) {
  NULLCHECK(imDimZYX);
  NULLCHECK(kernelDimZYX);
  const int n[3] = {imDimZYX[2], imDimZYX[1], imDimZYX[0]};
  const int k[3] = {kernelDimZYX[2], kernelDimZYX[1], kernelDimZYX[0]};
  return mvr_fft_convolve(dev, im, n, MVR_EXT_MIRROR_SINGLE, 0.f, kernel, k, im);
}

int mvr_init_psfs(int dev, const float* kernels, int V, const int k[3], int psfType, int nT, float* k1Out,
                  float* k2Out) {
  NULLCHECK(kernels);
  NULLCHECK(k);
  NULLCHECK(k1Out);
  NULLCHECK(k2Out);
  if (V < 1 || psfType < 0 || psfType > 3) { set_error("bad view count or PSF type"); return -1; }
  MVR_HIP(hipSetDevice(dev));
  const size_t nk = (size_t)k[0] * k[1] * k[2];
  std::vector<std::vector<float>> k1(V), k2;
  for (int v = 0; v < V; ++v) k1[v].assign(kernels + nk * v, kernels + nk * (v + 1));
  hipStream_t s = nullptr;
  MVR_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  const int rc = init_psfs(dev, k1, k, psfType, nT, k2, s);
  (void)hipStreamDestroy(s);
  if (rc) return rc;
  for (int v = 0; v < V; ++v) {
    std::memcpy(k1Out + nk * v, k1[v].data(), nk * 4);
    std::memcpy(k2Out + nk * v, k2[v].data(), nk * 4);
  }
  return 0;
}

// ---------------------------------------------------------------- block thread
int mvr_block_thread_create(mvr_block_thread** out, int dev, const int b[3], float minValue, float lambda, int id) {
  NULLCHECK(out);
  NULLCHECK(b);
  auto* h = new (std::nothrow) mvr_block_thread();
  if (!h) { set_error("out of host memory"); return -1; }
  if (h->t.create(dev, b, minValue, lambda, id)) {
    const std::string why = last_error();  // the clean-up's own device calls must not replace the reason
    h->t.destroy();
    delete h;
    set_error(why);
    return -1;
  }
  *out = h;
  return 0;
}

int mvr_block_thread_set_view_kernels(mvr_block_thread* t, int view, const float* k1, const float* k2,
                                      const int k[3]) {
  NULLCHECK(t);
  NULLCHECK(k1);
  NULLCHECK(k2);
  NULLCHECK(k);
  return t->t.set_view_kernels(view, k1, k2, k);
}

int mvr_block_run_iteration_seq(mvr_block_thread* t, int view, float* psi, const float* img, const float* w,
                                float maxI, double stats[2]) {
  NULLCHECK(t);
  NULLCHECK(psi);
  NULLCHECK(img);
  NULLCHECK(w);
  return t->t.run_seq(view, psi, img, w, maxI, stats);
}

int mvr_block_run_iteration_mul(mvr_block_thread* t, int nv, float* psi, const float* const* imgs,
                                const float* const* ws, const float* maxIs, double stats[2]) {
  NULLCHECK(t);
  NULLCHECK(psi);
  NULLCHECK(imgs);
  NULLCHECK(ws);
  NULLCHECK(maxIs);
  return t->t.run_mul(nv, psi, imgs, ws, maxIs, stats);
}

int mvr_block_thread_set_view_data(mvr_block_thread* t, int view, const float* img, const float* weight,
                                   const int nXYZ[3]) {
  NULLCHECK(t);
  return t->t.set_view_data(view, img, weight, nXYZ);
}

int mvr_block_run_iteration_seq_at(mvr_block_thread* t, int view, float* psiBlock, const long long offsetXYZ[3],
                                   float maxIntensityView, double stats[2]) {
  NULLCHECK(t);
  return t->t.run_seq_at(view, psiBlock, offsetXYZ, maxIntensityView, stats);
}

int mvr_block_run_iteration_mul_at(mvr_block_thread* t, int numViews, float* psiBlock, const long long offsetXYZ[3],
                                   const float* maxIntensityViews, double stats[2]) {
  NULLCHECK(t);
  return t->t.run_mul_at(numViews, psiBlock, offsetXYZ, maxIntensityViews, stats);
}

void mvr_block_thread_destroy(mvr_block_thread* t) {
  if (!t) return;
  t->t.destroy();
  delete t;
}

// ---------------------------------------------------------------- resident driver
int mvr_decon_create(mvr_decon** out, int dev, const int n[3], int V, int psfType, int scheme, float lambda,
                     float minValue, int nT) {
  NULLCHECK(out);
  NULLCHECK(n);
  auto* h = new (std::nothrow) mvr_decon();
  if (!h) { set_error("out of host memory"); return -1; }
  if (h->d.create(dev, n, V, psfType, scheme, lambda, minValue, nT)) {
    const std::string why = last_error();  // the clean-up's own device calls must not replace the reason
    h->d.destroy();
    delete h;
    set_error(why);
    return -1;
  }
  *out = h;
  return 0;
}

int mvr_decon_set_view(mvr_decon* d, int v, const float* img, const float* w, const float* k, const int kd[3]) {
  NULLCHECK(d);
  NULLCHECK(kd);
  return d->d.set_view(v, img, w, k, kd);
}

int mvr_decon_set_view_kernels(mvr_decon* d, int view, const float* kernel1, const float* kernel2,
                               const int kXYZ[3]) {
  NULLCHECK(d);
  return d->d.set_view_kernels(view, kernel1, kernel2, kXYZ);
}

int mvr_decon_set_max(mvr_decon* d, const float* maxPerView) {
  NULLCHECK(d);
  return d->d.set_max(maxPerView);
}

int mvr_decon_set_block_size(mvr_decon* d, const int blockSizeXYZ[3]) {
  NULLCHECK(d);
  NULLCHECK(blockSizeXYZ);
  return d->d.set_block_size(blockSizeXYZ);
}

int mvr_decon_get_blocks(mvr_decon* d, int* blocksPerView, int cap) {
  NULLCHECK(d);
  if (!d->d.ready) { set_error("init first"); return -1; }
  if (!d->d.blocked()) {
    if (blocksPerView) for (int v = 0; v < d->d.V && v < cap; ++v) blocksPerView[v] = 1;
    return 1;
  }
  if (blocksPerView) for (int v = 0; v < d->d.V && v < cap; ++v) blocksPerView[v] = d->d.nbView[v];
  return d->d.nbAll;
}

int mvr_decon_init(mvr_decon* d, int psiInit, const float* psi0) {
  NULLCHECK(d);
  if (psiInit != MVR_PSI_INIT_AVG_PRECISE && psiInit != MVR_PSI_INIT_FROM_ARRAY && psiInit != MVR_PSI_INIT_AVG_APPROX &&
      psiInit != MVR_PSI_INIT_BLURRED_FUSED) {
    set_error("unknown psi init");
    return -1;
  }
  return d->d.init(psiInit, psi0);
}

int mvr_decon_run(mvr_decon* d, int iterations) {
  NULLCHECK(d);
  return d->d.run(iterations);
}

int mvr_decon_run_view(mvr_decon* d, int view) {
  NULLCHECK(d);
  return d->d.run_view(view);
}

int mvr_decon_get_psi(mvr_decon* d, float* out) {
  NULLCHECK(d);
  NULLCHECK(out);
  if (!d->d.ready) { set_error("init first"); return -1; }
  MVR_HIP(hipSetDevice(d->d.dev));
  MVR_HIP(hipMemcpyAsync(out, d->d.psi, d->d.nvox() * 4, hipMemcpyDefault, d->d.stream));
  MVR_HIP(hipStreamSynchronize(d->d.stream));
  return 0;
}

int mvr_decon_get_max(mvr_decon* d, float* maxOut, double* avgOut) {
  NULLCHECK(d);
  if (!d->d.ready) { set_error("init first"); return -1; }
  if (maxOut) for (int v = 0; v < d->d.V; ++v) maxOut[v] = d->d.maxI[v];
  if (avgOut) *avgOut = d->d.avg;
  return 0;
}

int mvr_decon_get_kernels(mvr_decon* d, int v, float* k1, float* k2) {
  NULLCHECK(d);
  if (!d->d.ready || v < 0 || v >= d->d.V) { set_error("init first / bad view"); return -1; }
  if (k1) std::memcpy(k1, d->d.k1[v].data(), d->d.k1[v].size() * 4);
  if (k2) std::memcpy(k2, d->d.k2[v].data(), d->d.k2[v].size() * 4);
  return 0;
}

int mvr_decon_get_stats(mvr_decon* d, double* out, int cap) {
  NULLCHECK(d);
  const int n = (int)d->d.stats.size();
  if (out) for (int i = 0; i < n && i < cap; ++i) out[i] = d->d.stats[i];
  return n;
}

int mvr_decon_reset_psi(mvr_decon* d) {
  NULLCHECK(d);
  return d->d.reset_psi();
}

int mvr_decon_get_geometry(mvr_decon* d, int out[4]) {
  NULLCHECK(d);
  NULLCHECK(out);
  if (!d->d.ready) { set_error("init first"); return -1; }
  for (int k = 0; k < 3; ++k) out[k] = d->d.plan.dom.P[k];
  out[3] = d->d.plan.dom.HxP;
  return 0;
}

int mvr_decon_profile(mvr_decon* d, unsigned mask) {
  NULLCHECK(d);
  d->d.timer.reset(mask);
  return 0;
}

int mvr_decon_profile_sampled(mvr_decon* d, unsigned mask, int every) {
  NULLCHECK(d);
  if (every < 1) { set_error("mvr_decon_profile_sampled: every must be >= 1"); return -1; }
  d->d.timer.reset(mask, every);
  return 0;
}

int mvr_decon_pass_times(mvr_decon* d, double ms[9], long long n[9]) {
  NULLCHECK(d);
  return d->d.timer.collect(ms, n);
}

void mvr_decon_destroy(mvr_decon* d) {
  if (!d) return;
  d->d.destroy();
  delete d;
}

int mvr_decon_autotune(mvr_decon* d, int candidatesPerAxis, int chosenXYZ[3], double* msPerUpdate) {
  NULLCHECK(d);
  return d->d.autotune(candidatesPerAxis, chosenXYZ, msPerUpdate);
}

int mvr_decon_set_tile_widths(mvr_decon* d, int wy, int wz) {
  NULLCHECK(d);
  for (const int w : {wy, wz})
    if (w != 0 && w != 8 && w != 16) { set_error("tile width must be 0, 8 or 16"); return -1; }
  Decon& x = d->d;
  x.tunedW[0] = wy;
  x.tunedW[1] = wz;
  x.plan.wy = wy;  // the spectra's layout does not depend on the tiling
  x.plan.wz = wz;
  x.drop_graph();
  return 0;
}

int mvr_decon_get_tile_widths(mvr_decon* d, int out[2]) {
  NULLCHECK(d);
  NULLCHECK(out);
  out[0] = d->d.plan.wy;
  out[1] = d->d.plan.wz;
  return 0;
}

int mvr_decon_use_graph(mvr_decon* d, int enable) {
  NULLCHECK(d);
  return d->d.set_graph(enable != 0);
}

int mvr_decon_init_learned(mvr_decon* d, mvr_unet* net) {
  NULLCHECK(d);
  NULLCHECK(net);
  if (net->n.V != d->d.V) { set_error("the network was built for a different number of views"); return -1; }
  if (net->n.dev != d->d.dev) { set_error("the network lives on another device"); return -1; }
  // PSFs, spectra and the AvgPrecise pass (its max[] and avg are kept, as PsiInitFromRAI does);
  // a handle that is initialised already keeps them and only its psi is replaced
  if (!d->d.ready && d->d.init(MVR_PSI_INIT_AVG_PRECISE, nullptr)) return -1;
  MVR_HIP(hipStreamSynchronize(d->d.stream));
  if (net->n.infer(d->d.img, d->d.w, d->d.N, d->d.minValue, d->d.psi)) return -1;
  return d->d.commit_psi0();
}

int mvr_decon_set_reference(mvr_decon* d, const float* ref) {
  NULLCHECK(d);
  NULLCHECK(ref);
  return d->d.set_reference(ref);
}

int mvr_decon_rel_l2(mvr_decon* d, double* out) {
  NULLCHECK(d);
  NULLCHECK(out);
  return d->d.rel_l2(out);
}

int mvr_decon_run_until(mvr_decon* d, int maxIterations, double tol, int* iterationsOut, double* relOut) {
  NULLCHECK(d);
  return d->d.run_until(maxIterations, tol, iterationsOut, relOut);
}

// ---------------------------------------------------------------- input preparation
namespace {

// inverse of a 3 x 4 affine (input pixel -> world) in double, as AffineTransform3D keeps it
int invert_affine(const double* f, Affine3& inv) {
  const double a = f[0], b = f[1], c = f[2], d = f[4], e = f[5], g = f[6], h = f[8], i = f[9], k = f[10];
  const double det = a * (e * k - g * i) - b * (d * k - g * h) + c * (d * i - e * h);
  if (!(std::fabs(det) > 0)) { set_error("affine transform is singular"); return -1; }
  const double r[9] = {(e * k - g * i) / det, (c * i - b * k) / det, (b * g - c * e) / det,
                       (g * h - d * k) / det, (a * k - c * h) / det, (c * d - a * g) / det,
                       (d * i - e * h) / det, (b * h - a * i) / det, (a * e - b * d) / det};
  for (int row = 0; row < 3; ++row) {
    inv.m[4 * row + 0] = r[3 * row + 0];
    inv.m[4 * row + 1] = r[3 * row + 1];
    inv.m[4 * row + 2] = r[3 * row + 2];
    inv.m[4 * row + 3] = -(r[3 * row + 0] * f[3] + r[3 * row + 1] * f[7] + r[3 * row + 2] * f[11]);
  }
  return 0;
}

// BlendingRealRandomAccess's static table (fusion/transformed/weights/BlendingRealRandomAccess.java:51-57),
// uploaded once per device
int ensure_blend_lut(int dev) {
  static bool done[64] = {false};
  if (dev < 0 || dev >= 64) { set_error("bad device"); return -1; }
  if (done[dev]) return 0;
  std::vector<double> lut(1001, 0.0);
  for (double x = 0; x <= 1.0001; x = x + 0.001) lut[(int)(x * 1000.0 + 0.5)] = (std::cos((1 - x) * M_PI) + 1) / 2;
  if (upload_blend_lut(lut.data())) { set_error("blend table upload failed"); return -1; }
  done[dev] = true;
  return 0;
}

struct DevBuf {
  void* p = nullptr;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { if (p) (void)hipFree(p); }
};

}  // namespace

int mvr_transform_view(int dev, const float* img, const int imgXYZ[3], const double affine[12], const long long bbMin[3],
                       const int bbXYZ[3], int hasMinValue, float minValue, float outsideValue, int interpolation,
                       float* out) {
  NULLCHECK(img);
  NULLCHECK(imgXYZ);
  NULLCHECK(affine);
  NULLCHECK(bbMin);
  NULLCHECK(bbXYZ);
  NULLCHECK(out);
  for (int d = 0; d < 3; ++d)
    if (imgXYZ[d] < 1 || bbXYZ[d] < 1) { set_error("empty image or bounding box"); return -1; }
  if (interpolation != 0 && interpolation != 1) { set_error("interpolation must be 0 or 1"); return -1; }
  Affine3 inv;
  if (invert_affine(affine, inv)) return -1;
  MVR_HIP(hipSetDevice(dev));
  const size_t ni = (size_t)imgXYZ[0] * imgXYZ[1] * imgXYZ[2], no = (size_t)bbXYZ[0] * bbXYZ[1] * bbXYZ[2];
  DevBuf di, dout;
  MVR_HIP(hipMalloc(&di.p, ni * 4));
  MVR_HIP(hipMalloc(&dout.p, no * 4));
  MVR_HIP(hipMemcpy(di.p, img, ni * 4, hipMemcpyDefault));
  if (launch_transform_view((const float*)di.p, imgXYZ, inv, bbMin, bbXYZ, hasMinValue, minValue, outsideValue,
                            interpolation, (float*)dout.p, nullptr)) {
    set_error("transform kernel launch failed");
    return -1;
  }
  MVR_HIP(hipMemcpy(out, dout.p, no * 4, hipMemcpyDefault));
  return 0;
}

int mvr_transform_blending(int dev, const int imgXYZ[3], const float border[3], const float blending[3],
                           const double affine[12], const long long bbMin[3], const int bbXYZ[3], float* out) {
  NULLCHECK(imgXYZ);
  NULLCHECK(border);
  NULLCHECK(blending);
  NULLCHECK(affine);
  NULLCHECK(bbMin);
  NULLCHECK(bbXYZ);
  NULLCHECK(out);
  for (int d = 0; d < 3; ++d)
    if (imgXYZ[d] < 1 || bbXYZ[d] < 1) { set_error("empty image or bounding box"); return -1; }
  Affine3 inv;
  if (invert_affine(affine, inv)) return -1;
  MVR_HIP(hipSetDevice(dev));
  if (ensure_blend_lut(dev)) return -1;
  const size_t no = (size_t)bbXYZ[0] * bbXYZ[1] * bbXYZ[2];
  DevBuf dout;
  MVR_HIP(hipMalloc(&dout.p, no * 4));
  if (launch_transform_blending(imgXYZ, border, blending, inv, bbMin, bbXYZ, (float*)dout.p, nullptr)) {
    set_error("blending kernel launch failed");
    return -1;
  }
  MVR_HIP(hipMemcpy(out, dout.p, no * 4, hipMemcpyDefault));
  return 0;
}

int mvr_fuse_group(int dev, int numViews, const float* const* imgs, const int* imgXYZ, const double* affines,
                   const double* intensityAdjust, const float* fusionBlending, const float* deconBlending,
                   const long long bbMin[3], const int bbXYZ[3], float* imgOut, float* weightOut) {
  NULLCHECK(imgs);
  NULLCHECK(imgXYZ);
  NULLCHECK(affines);
  NULLCHECK(bbMin);
  NULLCHECK(bbXYZ);
  NULLCHECK(imgOut);
  NULLCHECK(weightOut);
  if (numViews < 1) { set_error("a group needs at least one view"); return -1; }
  for (int d = 0; d < 3; ++d)
    if (bbXYZ[d] < 1) { set_error("empty bounding box"); return -1; }
  std::vector<FuseView> fv(numViews);
  for (int j = 0; j < numViews; ++j) {
    NULLCHECK(imgs[j]);
    FuseView& v = fv[j];
    for (int d = 0; d < 3; ++d) {
      v.in[d] = imgXYZ[3 * j + d];
      if (v.in[d] < 1) { set_error("empty view image"); return -1; }
    }
    if (invert_affine(affines + 12 * j, v.inv)) return -1;
    v.adjust = intensityAdjust ? 1 : 0;
    v.adjA = intensityAdjust ? intensityAdjust[2 * j] : 1.0;
    v.adjB = intensityAdjust ? intensityAdjust[2 * j + 1] : 0.0;
    v.fusionBlend = fusionBlending ? 1 : 0;
    v.deconBlend = deconBlending ? 1 : 0;
    for (int d = 0; d < 3; ++d) {
      v.fBorder[d] = fusionBlending ? fusionBlending[6 * j + d] : 0.f;
      v.fRange[d] = fusionBlending ? fusionBlending[6 * j + 3 + d] : 0.f;
      v.dBorder[d] = deconBlending ? deconBlending[6 * j + d] : 0.f;
      v.dRange[d] = deconBlending ? deconBlending[6 * j + 3 + d] : 0.f;
    }
  }
  MVR_HIP(hipSetDevice(dev));
  if (ensure_blend_lut(dev)) return -1;
  std::vector<DevBuf> dimg(numViews);
  for (int j = 0; j < numViews; ++j) {
    const size_t ni = (size_t)fv[j].in[0] * fv[j].in[1] * fv[j].in[2];
    MVR_HIP(hipMalloc(&dimg[j].p, ni * 4));
    MVR_HIP(hipMemcpy(dimg[j].p, imgs[j], ni * 4, hipMemcpyDefault));
    fv[j].img = (const float*)dimg[j].p;
  }
  const size_t no = (size_t)bbXYZ[0] * bbXYZ[1] * bbXYZ[2];
  DevBuf dviews, dout, dw;
  MVR_HIP(hipMalloc(&dviews.p, sizeof(FuseView) * numViews));
  MVR_HIP(hipMemcpy(dviews.p, fv.data(), sizeof(FuseView) * numViews, hipMemcpyHostToDevice));
  MVR_HIP(hipMalloc(&dout.p, no * 4));
  MVR_HIP(hipMalloc(&dw.p, no * 4));
  if (launch_fuse_group((const FuseView*)dviews.p, numViews, bbMin, bbXYZ, (float*)dout.p, (float*)dw.p, nullptr)) {
    set_error("group fusion kernel launch failed");
    return -1;
  }
  MVR_HIP(hipMemcpy(imgOut, dout.p, no * 4, hipMemcpyDefault));
  MVR_HIP(hipMemcpy(weightOut, dw.p, no * 4, hipMemcpyDefault));
  return 0;
}

// PSFExtraction.transformPSF geometry (psf/PSFExtraction.java:367-409): the transformed bounds of [0, dim - 1],
// size rounded to an odd count, offset so that the transformed centre dim / 2 is the new centre
static void psf_geometry(const int psfXYZ[3], const double* f, int outXYZ[3], double off[3]) {
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int c = 0; c < 8; ++c) {
    const double p[3] = {(c & 1) ? psfXYZ[0] - 1.0 : 0.0, (c & 2) ? psfXYZ[1] - 1.0 : 0.0, (c & 4) ? psfXYZ[2] - 1.0 : 0.0};
    for (int d = 0; d < 3; ++d) {
      const double t = f[4 * d] * p[0] + f[4 * d + 1] * p[1] + f[4 * d + 2] * p[2] + f[4 * d + 3];
      lo[d] = std::min(lo[d], t);
      hi[d] = std::max(hi[d], t);
    }
  }
  const double ctr[3] = {(double)(psfXYZ[0] / 2), (double)(psfXYZ[1] / 2), (double)(psfXYZ[2] / 2)};
  for (int d = 0; d < 3; ++d) {
    const double tc = f[4 * d] * ctr[0] + f[4 * d + 1] * ctr[1] + f[4 * d + 2] * ctr[2] + f[4 * d + 3];
    long long n = (long long)(int)(hi[d] - lo[d]) + 1;
    if (n % 2 == 0) ++n;
    outXYZ[d] = (int)n;
    off[d] = tc - (double)(n / 2);
  }
}

int mvr_transform_psf_size(const int psfXYZ[3], const double affine[12], int outXYZ[3]) {
  NULLCHECK(psfXYZ);
  NULLCHECK(affine);
  NULLCHECK(outXYZ);
  double off[3];
  psf_geometry(psfXYZ, affine, outXYZ, off);
  return 0;
}

int mvr_transform_psf(int dev, const float* psf, const int psfXYZ[3], const double affine[12], float* out) {
  NULLCHECK(psf);
  NULLCHECK(psfXYZ);
  NULLCHECK(affine);
  NULLCHECK(out);
  for (int d = 0; d < 3; ++d)
    if (psfXYZ[d] < 1) { set_error("empty PSF"); return -1; }
  Affine3 inv;
  if (invert_affine(affine, inv)) return -1;
  int outXYZ[3];
  double off[3];
  psf_geometry(psfXYZ, affine, outXYZ, off);
  const size_t ni = (size_t)psfXYZ[0] * psfXYZ[1] * psfXYZ[2], no = (size_t)outXYZ[0] * outXYZ[1] * outXYZ[2];
  // PSFExtraction.normalize (psf/PSFExtraction.java:453-470): (v - min) / (max - min), in double
  std::vector<float> h(ni);
  MVR_HIP(hipSetDevice(dev));
  MVR_HIP(hipMemcpy(h.data(), psf, ni * 4, hipMemcpyDefault));
  double mn = 1e308, mx = -1e308;
  for (float v : h) { mn = std::min(mn, (double)v); mx = std::max(mx, (double)v); }
  for (float& v : h) v = (float)(((double)v - mn) / (mx - mn));
  DevBuf di, dout;
  MVR_HIP(hipMalloc(&di.p, ni * 4));
  MVR_HIP(hipMalloc(&dout.p, no * 4));
  MVR_HIP(hipMemcpy(di.p, h.data(), ni * 4, hipMemcpyHostToDevice));
  if (launch_transform_psf((const float*)di.p, psfXYZ, inv, off, outXYZ, (float*)dout.p, nullptr)) {
    set_error("PSF transform launch failed");
    return -1;
  }
  MVR_HIP(hipMemcpy(out, dout.p, no * 4, hipMemcpyDefault));
  return 0;
}

int mvr_normalize_weights(int dev, float* const* weights, int numViews, const int nXYZ[3], double osemSpeedup,
                          int additionalSmoothBlending, float maxDiffRange, float scalingRange) {
  NULLCHECK(weights);
  NULLCHECK(nXYZ);
  if (numViews < 1 || numViews > kPrepMaxViews) { set_error("numViews must be in [1, 32]"); return -1; }
  for (int v = 0; v < numViews; ++v)
    if (!weights[v]) { set_error("null weight array"); return -1; }
  MVR_HIP(hipSetDevice(dev));
  const size_t n = (size_t)nXYZ[0] * nXYZ[1] * nXYZ[2];
  std::vector<DevBuf> bufs(numViews);
  std::vector<float*> dp(numViews);
  for (int v = 0; v < numViews; ++v) {
    MVR_HIP(hipMalloc(&bufs[v].p, n * 4));
    MVR_HIP(hipMemcpy(bufs[v].p, weights[v], n * 4, hipMemcpyDefault));
    dp[v] = (float*)bufs[v].p;
  }
  if (launch_normalize_weights(dp.data(), numViews, n, osemSpeedup, additionalSmoothBlending, maxDiffRange,
                               scalingRange, nullptr)) {
    set_error("normalisation kernel launch failed");
    return -1;
  }
  for (int v = 0; v < numViews; ++v) MVR_HIP(hipMemcpy(weights[v], dp[v], n * 4, hipMemcpyDefault));
  return 0;
}

// ---------------------------------------------------------------- learned initial estimate
}  // extern "C"

namespace {

// One array of an uncompressed .npz (the zip of .npy files that np.savez writes): walks the zip's local file
// headers (stored entries only; sizes from the header or its zip64 extra field) to member `name`.npy and
// parses its .npy header.  Returns false with a message on anything else.
bool npz_array(const std::vector<unsigned char>& z, const std::string& name, std::string& descr,
               std::vector<long long>& shape, const unsigned char*& data, size_t& bytes, std::string& err) {
  auto u16 = [&](size_t o) { return (unsigned)z[o] | ((unsigned)z[o + 1] << 8); };
  auto u32 = [&](size_t o) { return (uint32_t)u16(o) | ((uint32_t)u16(o + 2) << 16); };
  auto u64 = [&](size_t o) { return (uint64_t)u32(o) | ((uint64_t)u32(o + 4) << 32); };
  size_t o = 0;
  while (o + 30 <= z.size() && u32(o) == 0x04034b50u) {
    const unsigned flags = u16(o + 6), method = u16(o + 8), nl = u16(o + 26), el = u16(o + 28);
    uint64_t size = u32(o + 18);
    if (o + 30 + nl + el > z.size()) break;
    const std::string entry(reinterpret_cast<const char*>(&z[o + 30]), nl);
    for (size_t e = o + 30 + nl; e + 4 <= o + 30 + nl + el;) {  // zip64 extra: original size, compressed size
      const unsigned id = u16(e), len = u16(e + 2);
      if (id == 1 && size == 0xFFFFFFFFu && len >= 16) size = u64(e + 12);
      e += 4 + len;
    }
    const size_t d = o + 30 + nl + el;
    if (flags & 8) { err = "npz entries with data descriptors are not supported"; return false; }
    if (d + size > z.size()) break;
    if (entry == name + ".npy") {
      if (method != 0) { err = "compressed npz entries are not supported (write with np.savez)"; return false; }
      const unsigned char* p = &z[d];
      if (size < 10 || std::memcmp(p, "\x93NUMPY", 6) != 0) { err = "not an .npy array: " + entry; return false; }
      const size_t hl = p[6] == 1 ? u16(d + 8) : (size_t)u32(d + 8), h0 = p[6] == 1 ? 10 : 12;
      if (h0 + hl > size) { err = "truncated .npy header"; return false; }
      const std::string h(reinterpret_cast<const char*>(p + h0), hl);
      auto field = [&](const char* key) {
        const size_t k = h.find(key);
        return k == std::string::npos ? std::string() : h.substr(k + std::strlen(key));
      };
      const std::string ds = field("'descr': '");
      descr = ds.substr(0, ds.find('\''));
      if (field("'fortran_order': ").rfind("False", 0) != 0) { err = "fortran-ordered arrays are not supported"; return false; }
      std::string sh = field("'shape': (");
      sh = sh.substr(0, sh.find(')'));
      shape.clear();
      for (size_t q = 0; q < sh.size();) {
        while (q < sh.size() && (sh[q] == ' ' || sh[q] == ',')) ++q;
        if (q >= sh.size()) break;
        shape.push_back(std::stoll(sh.substr(q)));
        while (q < sh.size() && sh[q] != ',') ++q;
      }
      data = p + h0 + hl;
      bytes = size - h0 - hl;
      return true;
    }
    o = d + size;
  }
  err = "no array '" + name + "' in the npz file";
  return false;
}

int read_weights(const char* path, std::vector<unsigned char>& buf, const float*& params, long long& count,
                 int& views) {
  FILE* f = std::fopen(path, "rb");
  if (!f) { set_error(std::string("cannot open ") + path); return -1; }
  std::fseek(f, 0, SEEK_END);
  const long n = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  buf.resize(n > 0 ? (size_t)n : 0);
  const size_t got = n > 0 ? std::fread(buf.data(), 1, buf.size(), f) : 0;
  std::fclose(f);
  if (got != buf.size()) { set_error("short read"); return -1; }
  std::string descr, err;
  std::vector<long long> shape;
  const unsigned char* data = nullptr;
  size_t bytes = 0;
  if (!npz_array(buf, "params", descr, shape, data, bytes, err)) { set_error(err); return -1; }
  if (descr != "<f4" || shape.size() != 1 || (size_t)shape[0] * 4 != bytes) {
    set_error("params must be a float32 vector");
    return -1;
  }
  params = reinterpret_cast<const float*>(data);  // callers copy out with memcpy (no alignment assumed)
  count = shape[0];
  if (!npz_array(buf, "views", descr, shape, data, bytes, err)) { set_error(err); return -1; }
  if (!shape.empty() || !(descr == "<i4" || descr == "<i8")) { set_error("views must be an integer scalar"); return -1; }
  int64_t v = 0;
  std::memcpy(&v, data, descr == "<i4" ? 4 : 8);
  if (descr == "<i4") v = (int32_t)v;
  views = (int)v;
  return 0;
}

}  // namespace

extern "C" {

long long mvr_unet_read_weights(const char* path, float* params, long long cap, int* views) {
  NULLCHECK(path);
  std::vector<unsigned char> buf;
  const float* p = nullptr;
  long long count = 0;
  int v = 0;
  if (read_weights(path, buf, p, count, v)) return -1;
  if (params) {
    if (cap < count) { set_error("params buffer too small"); return -1; }
    std::memcpy(params, p, (size_t)count * 4);
  }
  if (views) *views = v;
  return count;
}

int mvr_unet_load(mvr_unet** out, int dev, const char* path) {
  NULLCHECK(out);
  NULLCHECK(path);
  std::vector<unsigned char> buf;
  const float* p = nullptr;
  long long count = 0;
  int v = 0;
  if (read_weights(path, buf, p, count, v)) return -1;
  std::vector<float> params((size_t)count);
  std::memcpy(params.data(), p, (size_t)count * 4);
  mvr_unet* h = nullptr;
  if (mvr_unet_create(&h, dev, v, 0)) return -1;
  if (count != mvr_unet_param_count(h)) {
    mvr_unet_destroy(h);
    set_error("the weights do not match the network's parameter count");
    return -1;
  }
  if (mvr_unet_set_params(h, params.data())) { mvr_unet_destroy(h); return -1; }
  *out = h;
  return 0;
}

int mvr_unet_create(mvr_unet** out, int dev, int views, unsigned long long seed) {
  NULLCHECK(out);
  auto* h = new (std::nothrow) mvr_unet();
  if (!h) { set_error("out of host memory"); return -1; }
  if (h->n.create(dev, views, seed)) {
    const std::string why = last_error();  // the clean-up's own device calls must not replace the reason
    h->n.destroy();
    delete h;
    set_error(why);
    return -1;
  }
  *out = h;
  return 0;
}

long long mvr_unet_param_count(mvr_unet* u) {
  NULLCHECK(u);
  return (long long)u->n.param_count();
}

int mvr_unet_get_params(mvr_unet* u, float* out) {
  NULLCHECK(u);
  NULLCHECK(out);
  return u->n.copy_params(out, nullptr, false);
}

int mvr_unet_set_params(mvr_unet* u, const float* in) {
  NULLCHECK(u);
  NULLCHECK(in);
  return u->n.copy_params(nullptr, in, false);
}

int mvr_unet_get_grads(mvr_unet* u, float* out) {
  NULLCHECK(u);
  NULLCHECK(out);
  return u->n.copy_params(out, nullptr, true);
}

int mvr_unet_add_volume(mvr_unet* u, const float* const* imgs, const float* const* weights, const float* target,
                        const int n[3]) {
  NULLCHECK(u);
  NULLCHECK(imgs);
  NULLCHECK(weights);
  NULLCHECK(target);
  NULLCHECK(n);
  for (int v = 0; v < u->n.V; ++v)
    if (!imgs[v] || !weights[v]) { set_error("null view image or weight"); return -1; }
  if (n[0] < 4 || n[1] < 4 || n[2] < 4) { set_error("training volumes need extents >= 4"); return -1; }
  return u->n.add_volume(imgs, weights, target, n);
}

int mvr_unet_clear_volumes(mvr_unet* u) {
  NULLCHECK(u);
  return u->n.clear_volumes();
}

int mvr_unet_train_step(mvr_unet* u, const int* volIdx, const int* origins, int batch, int patch, float lr,
                        int update, double* loss) {
  NULLCHECK(u);
  NULLCHECK(volIdx);
  NULLCHECK(origins);
  return u->n.train_step(volIdx, origins, batch, patch, lr, loss, update != 0);
}

int mvr_unet_infer(mvr_unet* u, const float* const* imgs, const float* const* weights, const int n[3],
                   float minValue, float* psiOut) {
  NULLCHECK(u);
  NULLCHECK(imgs);
  NULLCHECK(weights);
  NULLCHECK(n);
  NULLCHECK(psiOut);
  UNet& net = u->n;
  MVR_HIP(hipSetDevice(net.dev));
  const size_t nv = (size_t)n[0] * n[1] * n[2];
  // borrowed arrays may be host memory: stage views, weights and psi on the device
  std::vector<float*> di(net.V, nullptr), dw(net.V, nullptr);
  float* dpsi = nullptr;
  int rc = -1;
  do {
    bool ok = true;
    for (int v = 0; v < net.V && ok; ++v) {
      ok = hipMalloc(&di[v], nv * 4) == hipSuccess && hipMalloc(&dw[v], nv * 4) == hipSuccess &&
           hipMemcpy(di[v], imgs[v], nv * 4, hipMemcpyDefault) == hipSuccess &&
           hipMemcpy(dw[v], weights[v], nv * 4, hipMemcpyDefault) == hipSuccess;
    }
    if (!ok || hipMalloc(&dpsi, nv * 4) != hipSuccess) { set_error("device allocation or copy failed"); break; }
    if (net.infer(di, dw, n, minValue, dpsi)) break;
    if (hipMemcpy(psiOut, dpsi, nv * 4, hipMemcpyDefault) != hipSuccess) { set_error("copy-out failed"); break; }
    rc = 0;
  } while (0);
  for (auto p : di) if (p) (void)hipFree(p);
  for (auto p : dw) if (p) (void)hipFree(p);
  if (dpsi) (void)hipFree(dpsi);
  return rc;
}

int mvr_unet_inference_halo(void) { return kInferHalo; }

int mvr_unet_set_inference_window(mvr_unet* u, long long voxels) {
  NULLCHECK(u);
  if (voxels <= 0) voxels = 1ll << 28;
  u->n.inferWindow = voxels;
  return 0;
}

void mvr_unet_destroy(mvr_unet* u) {
  if (!u) return;
  u->n.destroy();
  delete u;
}

}  // extern "C"
