// psf.cpp -- kernel normalisation and the four iteration schemes' second kernels.
#include <algorithm>
#include <cmath>

#include "host.hpp"

namespace mvr {

namespace {

// FusionTools.divideIntoPortions (F/FusionTools.java:1287-1329); nT stands for
// Threads.numThreads() (Threads.java:40), a numeric parameter of the result.
struct Portion { size_t start, len; };
std::vector<Portion> portions(size_t n, int nT) {
  std::vector<Portion> out;
  if (n == 0) return out;
  long long np = (long long)n <= nT ? (long long)n : std::max<long long>(nT, (long long)(n / (64ull * 64ull * 64ull)));
  size_t chunk = n / np;
  while (chunk == 0) { --np; chunk = n / np; }
  const size_t mod = n % np;
  for (long long i = 0; i < np; ++i) out.push_back({(size_t)i * chunk, i == np - 1 ? chunk + mod : chunk});
  return out;
}

// compensated double accumulation (imglib2 RealSum)
struct Kahan {
  double s = 0, c = 0;
  void add(double v) { double y = v - c; double t = s + y; c = (t - s) - y; s = t; }
};

}  // namespace

// AdjustInput.sumImg (D/normalization/AdjustInput.java:71-122): the total
// starts with the first portion's sum and then adds every portion's sum.
double sum_img_quirk(const float* a, size_t n, int nT) {
  auto ps = portions(n, nT);
  std::vector<double> sums;
  for (auto& p : ps) {
    Kahan k;
    for (size_t j = 0; j < p.len; ++j) k.add((double)a[p.start + j]);
    sums.push_back(k.s);
  }
  Kahan total;
  if (!sums.empty()) total.add(sums[0]);
  for (double s : sums) total.add(s);
  return total.s;
}

// AdjustInput.normToSum1 (D/normalization/AdjustInput.java:52-58)
void norm_to_sum1(float* a, size_t n, int nT) {
  const double sum = sum_img_quirk(a, n, nT);
  for (size_t i = 0; i < n; ++i) a[i] = (float)((double)a[i] / sum);
}

// Mirror.mirror on every axis (D/util/Mirror.java:55-132), sequential order:
// pairs (i, n-1-i) with i <= n/2 are swapped in flat order, so an even axis
// leaves its middle pair in place.
void mirror_all(float* a, const int dim[3]) {
  const size_t sx = dim[0], sy = dim[1];
  for (int d = 0; d < 3; ++d) {
    const int n = dim[d];
    std::vector<int> partner(n);
    for (int i = 0; i < n; ++i) partner[i] = i;
    for (int i = 0; i <= n / 2 && i < n; ++i) std::swap(partner[i], partner[n - 1 - i]);
    // partner[] now holds the final source of each position along the axis
    std::vector<float> tmp((size_t)dim[0] * dim[1] * dim[2]);
    for (int z = 0; z < dim[2]; ++z)
      for (int y = 0; y < dim[1]; ++y)
        for (int x = 0; x < dim[0]; ++x) {
          int p[3] = {x, y, z};
          p[d] = partner[p[d]];
          tmp[((size_t)z * sy + y) * sx + x] = a[((size_t)p[2] * sy + p[1]) * sx + p[0]];
        }
    std::copy(tmp.begin(), tmp.end(), a);
  }
}

// zero-extended, same-size convolution of two K^3 kernels on the GPU
// (the FFTConvolution calls of DeconViewPSF.init, :152-178, :215-225)
static int conv_same(const std::vector<float>& input, const std::vector<float>& kernel, const int K[3],
                     std::vector<float>& out, hipStream_t s) {
  Plan plan;
  if (plan.create(K, K)) return -1;
  const size_t nk = (size_t)K[0] * K[1] * K[2];
  float *dIn = nullptr, *dK = nullptr, *dOut = nullptr;
  float2 *spec = nullptr, *work = nullptr;
  int rc = -1;
  do {
    if (hipMalloc(&dIn, nk * 4) || hipMalloc(&dK, nk * 4) || hipMalloc(&dOut, nk * 4) ||
        hipMalloc(&spec, plan.ncpx() * 8) || hipMalloc(&work, plan.ncpx() * 8)) {
      set_error("conv_same: hipMalloc failed");
      break;
    }
    if (hipMemcpyAsync(dIn, input.data(), nk * 4, hipMemcpyHostToDevice, s) ||
        hipMemcpyAsync(dK, kernel.data(), nk * 4, hipMemcpyHostToDevice, s)) {
      set_error("conv_same: copy failed");
      break;
    }
    if (plan.kernel_spectrum(dK, spec, s)) break;
    GatherParams g{};
    g.src = dIn;
    for (int d = 0; d < 3; ++d) { g.N[d] = K[d]; g.org[d] = -plan.dom.m[d]; }
    g.mode = GATHER_VALUE;
    g.value = 0.f;
    if (plan.convolve(g, spec, work, dOut, s)) break;
    out.resize(nk);
    if (hipMemcpyAsync(out.data(), dOut, nk * 4, hipMemcpyDeviceToHost, s) || hipStreamSynchronize(s)) {
      set_error("conv_same: readback failed");
      break;
    }
    rc = 0;
  } while (0);
  for (void* p : {(void*)dIn, (void*)dK, (void*)dOut, (void*)spec, (void*)work})
    if (p) (void)hipFree(p);
  plan.destroy();
  return rc;
}

int init_psfs(int dev, std::vector<std::vector<float>>& k1, const int K[3], int psfType, int nT,
              std::vector<std::vector<float>>& k2, hipStream_t s) {
  (void)dev;
  const int V = (int)k1.size();
  const size_t nk = (size_t)K[0] * K[1] * K[2];
  k2.assign(V, {});
  auto inverted = [&](const std::vector<float>& a) {
    std::vector<float> r(a);
    mirror_all(r.data(), K);
    return r;
  };
  for (int v = 0; v < V; ++v) {
    norm_to_sum1(k1[v].data(), nk, nT);
    if (V == 1 || psfType == PSF_INDEPENDENT) {
      k2[v] = inverted(k1[v]);
    } else if (psfType == PSF_EFFICIENT_BAYESIAN) {
      // P_v^compound = P_v^* prod_{w != v} (P_v^* (*) P_w (*) P_w^*)   (:132-195)
      std::vector<float> tmp = inverted(k1[v]);
      for (int w = 0; w < V; ++w) {
        if (w == v) continue;
        std::vector<float> o1, o2;
        if (conv_same(inverted(k1[v]), k1[w], K, o1, s)) return -1;
        if (conv_same(o1, inverted(k1[w]), K, o2, s)) return -1;
        for (size_t i = 0; i < nk; ++i) tmp[i] = o2[i] * tmp[i];
      }
      norm_to_sum1(tmp.data(), nk, nT);
      k2[v] = tmp;
    } else if (psfType == PSF_OPT_I) {
      // (P_v prod_{w != v} (P_v (*) P_w^*))^*   (:196-242)
      std::vector<float> tmp(k1[v]);
      for (int w = 0; w < V; ++w) {
        if (w == v) continue;
        std::vector<float> o;
        if (conv_same(k1[v], inverted(k1[w]), K, o, s)) return -1;
        for (size_t i = 0; i < nk; ++i) tmp[i] = o[i] * tmp[i];
      }
      norm_to_sum1(tmp.data(), nk, nT);
      k2[v] = inverted(tmp);
    } else {
      // (P_v^V)^*, iterated float multiply   (:243-253, :276-284)
      std::vector<float> e(k1[v]);
      for (size_t i = 0; i < nk; ++i) {
        float r = k1[v][i];
        for (int p = 1; p < V; ++p) r *= k1[v][i];
        e[i] = r;
      }
      norm_to_sum1(e.data(), nk, nT);
      k2[v] = inverted(e);
    }
  }
  return 0;
}

}  // namespace mvr
