// decon.cpp -- resident multiview driver and per-block compute threads.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>

#include "host.hpp"

namespace mvr {

#define MVR_CHECK(cond, msg)  \
  do {                        \
    if (!(cond)) {            \
      set_error(msg);         \
      return -1;              \
    }                         \
  } while (0)

// ============================================================ resident driver

int Decon::create(int dev_, const int N_[3], int V_, int psfType_, int scheme_, float lambda_, float minValue_,
                  int nT_) {
  dev = dev_;
  MVR_CHECK(V_ >= 1 && V_ <= 32, "numViews must be in [1, 32]");
  MVR_CHECK(psfType_ >= 0 && psfType_ <= 3, "unknown PSF type");
  MVR_CHECK(scheme_ == SCHEME_SEQ || scheme_ == SCHEME_MUL, "unknown iteration scheme");
  for (int d = 0; d < 3; ++d) {
    MVR_CHECK(N_[d] >= 1, "psi dimensions must be positive");
    N[d] = N_[d];
  }
  V = V_;
  psfType = psfType_;
  scheme = scheme_;
  lambda = lambda_;
  minValue = minValue_;
  nT = nT_;
  MVR_HIP(hipSetDevice(dev));
  MVR_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  img.assign(V, nullptr);
  w.assign(V, nullptr);
  kraw.assign(V, {});
  kfin1.assign(V, {});
  kfin2.assign(V, {});
  for (int v = 0; v < V; ++v) {
    MVR_HIP(hipMalloc(&img[v], nvox() * 4));
    MVR_HIP(hipMalloc(&w[v], nvox() * 4));
  }
  MVR_HIP(hipMalloc(&psi, nvox() * 4));
  MVR_HIP(hipMalloc(&psi0, nvox() * 4));
  return 0;
}

int Decon::set_view(int v, const float* im, const float* wt, const float* kernel, const int Kd[3]) {
  MVR_CHECK(v >= 0 && v < V, "view index out of range");
  MVR_CHECK(im && wt && kernel, "null input");
  for (int d = 0; d < 3; ++d) {
    MVR_CHECK(Kd[d] >= 1 && Kd[d] <= N[d] + 64 * 1024, "bad kernel size");
    if (K[0] != 0) MVR_CHECK(K[d] == Kd[d], "all views must share one PSF size");
  }
  for (int d = 0; d < 3; ++d) K[d] = Kd[d];
  MVR_HIP(hipSetDevice(dev));
  MVR_HIP(hipMemcpyAsync(img[v], im, nvox() * 4, hipMemcpyDefault, stream));
  MVR_HIP(hipMemcpyAsync(w[v], wt, nvox() * 4, hipMemcpyDefault, stream));
  const size_t nk = (size_t)K[0] * K[1] * K[2];
  kraw[v].resize(nk);
  MVR_HIP(hipMemcpyAsync(kraw[v].data(), kernel, nk * 4, hipMemcpyDefault, stream));
  MVR_HIP(hipStreamSynchronize(stream));
  ready = false;
  return 0;
}

// DeconViewPSF.getKernel1/getKernel2 of a view whose DeconViewPSF.init has run on the caller's side (kernel1
// normalised, kernel2 per PSF type): init then uses them as they are
int Decon::set_view_kernels(int v, const float* k1v, const float* k2v, const int Kd[3]) {
  MVR_CHECK(v >= 0 && v < V, "view index out of range");
  MVR_CHECK(k1v && k2v && Kd, "null kernel");
  MVR_CHECK(K[0] != 0, "set_view first");
  for (int d = 0; d < 3; ++d) MVR_CHECK(K[d] == Kd[d], "the kernels must have the size given to set_view");
  const size_t nk = (size_t)K[0] * K[1] * K[2];
  kfin1[v].resize(nk);
  kfin2[v].resize(nk);
  MVR_HIP(hipSetDevice(dev));
  MVR_HIP(hipMemcpy(kfin1[v].data(), k1v, nk * 4, hipMemcpyDefault));
  MVR_HIP(hipMemcpy(kfin2[v].data(), k2v, nk * 4, hipMemcpyDefault));
  ready = false;
  return 0;
}

// PsiInit.getMax of an initialiser that ran on the caller's side (MultiViewDeconvolution.java:124-134)
int Decon::set_max(const float* mx) {
  MVR_CHECK(ready, "init first");
  MVR_CHECK(mx != nullptr, "null max");
  for (int v = 0; v < V; ++v) maxI[v] = mx[v];
  drop_graph();  // a captured iteration holds the old maxima as kernel arguments
  return 0;
}

int Decon::set_block_size(const int B[3]) {
  bool on = false;
  for (int d = 0; d < 3; ++d) {
    MVR_CHECK(B[d] >= 0, "negative block size");
    on = on || B[d] > 0;
  }
  if (on) {
    for (int d = 0; d < 3; ++d) MVR_CHECK(B[d] > 0, "give a block size on every axis (or 0 on all for one frame)");
    MVR_CHECK(scheme == SCHEME_SEQ, "batched blocks run the Seq scheme");
  }
  for (int d = 0; d < 3; ++d) blockB[d] = on ? B[d] : 0;
  tunedP[0] = tunedP[1] = tunedP[2] = 0;
  ready = false;
  return 0;
}

void Decon::free_blocks() {
  for (auto p : btab) if (p) (void)hipFree(p);
  btab.clear();
  nbView.clear();
  nbAll = 0;
}

// Blocks of BlockGeneratorFixedSizePrecise (C/BlockGeneratorFixedSizePrecise.java:59-131): effective regions on a
// grid of E, x fastest, uneven tail blocks; each view keeps the blocks whose interval (effective region plus a
// margin of K - 1, D/DeconView.java:155-159) holds non-zero weights (DeconView.filterBlocksForContent,
// D/DeconView.java:204-274).
int Decon::build_blocks() {
  free_blocks();
  int nbx[3];
  for (int d = 0; d < 3; ++d) nbx[d] = (N[d] + blockE[d] - 1) / blockE[d];
  nbAll = nbx[0] * nbx[1] * nbx[2];
  std::vector<int> out, interval;   // {o, D} and {offset, size} per block
  for (int k = 0; k < nbx[2]; ++k)
    for (int j = 0; j < nbx[1]; ++j)
      for (int i = 0; i < nbx[0]; ++i) {
        const int c[3] = {i, j, k};
        for (int d = 0; d < 3; ++d) out.push_back(c[d] * blockE[d]);
        for (int d = 0; d < 3; ++d) out.push_back(std::min(blockE[d], N[d] - c[d] * blockE[d]));
        for (int d = 0; d < 3; ++d) interval.push_back(c[d] * blockE[d] - (K[d] - 1));
        for (int d = 0; d < 3; ++d) interval.push_back(blockB[d]);
      }
  int* dInt = nullptr;
  int* dFlags = nullptr;
  MVR_HIP(hipMalloc(&dInt, interval.size() * sizeof(int)));
  MVR_HIP(hipMalloc(&dFlags, sizeof(int) * nbAll));
  MVR_HIP(hipMemcpyAsync(dInt, interval.data(), interval.size() * sizeof(int), hipMemcpyHostToDevice, stream));
  btab.assign(V, nullptr);
  nbView.assign(V, 0);
  std::vector<int> flags(nbAll);
  int rc = 0;
  for (int v = 0; v < V && !rc; ++v) {
    launch_fill(reinterpret_cast<float*>(dFlags), (size_t)nbAll, 0.f, stream);
    launch_blocks_with_content(w[v], N, dInt, nbAll, dFlags, stream);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(flags.data(), dFlags, sizeof(int) * nbAll, hipMemcpyDeviceToHost, stream) != hipSuccess ||
        hipStreamSynchronize(stream) != hipSuccess) {
      set_error("block content scan failed");
      rc = -1;
      break;
    }
    std::vector<int> tab;
    for (int b = 0; b < nbAll; ++b)
      if (flags[b]) tab.insert(tab.end(), out.begin() + kBlockTabInts * b, out.begin() + kBlockTabInts * (b + 1));
    nbView[v] = (int)(tab.size() / kBlockTabInts);
    if (nbView[v] == 0) continue;
    if (hipMalloc(&btab[v], tab.size() * sizeof(int)) != hipSuccess ||
        hipMemcpy(btab[v], tab.data(), tab.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess) {
      set_error("block table upload failed");
      rc = -1;
    }
  }
  (void)hipFree(dInt);
  (void)hipFree(dFlags);
  return rc;
}

// partial sums of one x_update launch: per work-group, at most D1 D2 / 4 + 1 per block
size_t Decon::partials_needed() const {
  const size_t per = plan.rows_inner() / 4 + 1;
  return std::max<size_t>(2048, blocked() ? per * std::max(1, nbAll) : per);
}

int Decon::ensure_partials(int n) {
  if (n <= partialCap) return 0;
  if (partial) (void)hipFree(partial);
  MVR_HIP(hipMalloc(&partial, sizeof(double2) * n));
  partialCap = n;
  return 0;
}

int Decon::init(int psiInit, const float* psiIn) {
  MVR_HIP(hipSetDevice(dev));
  for (int v = 0; v < V; ++v) MVR_CHECK(!kraw[v].empty(), "every view needs set_view before init");
  // PSFs (DeconViews ctor -> DeconViewPSF.init), unless the caller handed over every view's final kernels
  bool allFinal = true;
  for (int v = 0; v < V; ++v) allFinal = allFinal && !kfin1[v].empty();
  if (allFinal) {
    k1 = kfin1;
    k2 = kfin2;
  } else {
    k1 = kraw;
    if (init_psfs(dev, k1, K, psfType, nT, k2, stream)) return -1;
  }
  // plan over the whole psi volume, mirror-single extension at its faces
  drop_graph();
  if (build_plan_and_spectra(tunedP[0] ? tunedP : nullptr, V)) return -1;
  // PsiInitAvgPrecise (D/init/PsiInitAvgPrecise.java:52-112)
  {
    const int maxParts = 1024;
    double* psum = nullptr;
    unsigned long long* pcnt = nullptr;
    float* pmax = nullptr;
    MVR_HIP(hipMalloc(&psum, maxParts * 8));
    MVR_HIP(hipMalloc(&pcnt, maxParts * 8));
    MVR_HIP(hipMalloc(&pmax, (size_t)maxParts * V * 4));
    std::vector<const float*> ptrs(img.begin(), img.end());
    const int np = launch_psi_init_partials(ptrs.data(), V, nvox(), psum, pcnt, pmax, stream);
    MVR_HIP(hipGetLastError());
    std::vector<double> hs(np);
    std::vector<unsigned long long> hc(np);
    std::vector<float> hm((size_t)np * V);
    MVR_HIP(hipMemcpyAsync(hs.data(), psum, np * 8, hipMemcpyDeviceToHost, stream));
    MVR_HIP(hipMemcpyAsync(hc.data(), pcnt, np * 8, hipMemcpyDeviceToHost, stream));
    MVR_HIP(hipMemcpyAsync(hm.data(), pmax, (size_t)np * V * 4, hipMemcpyDeviceToHost, stream));
    MVR_HIP(hipStreamSynchronize(stream));
    (void)hipFree(psum);
    (void)hipFree(pcnt);
    (void)hipFree(pmax);
    double s = 0.0;
    unsigned long long c = 0;
    maxI.assign(V, 0.f);
    for (int i = 0; i < np; ++i) {
      s += hs[i];
      c += hc[i];
      for (int v = 0; v < V; ++v) maxI[v] = std::max(maxI[v], hm[(size_t)i * V + v]);
    }
    avg = s / (double)c;
    if (std::isnan(avg)) avg = 1.0;
    psiCount = c;
  }
  if (psiInit == PSI_INIT_AVG_PRECISE) {
    launch_fill(psi, nvox(), (float)avg, stream);
  } else if (psiInit == PSI_INIT_FROM_ARRAY) {
    MVR_CHECK(psiIn != nullptr, "psi init from array needs an array");
    MVR_HIP(hipMemcpyAsync(psi, psiIn, nvox() * 4, hipMemcpyDefault, stream));
  } else if (psiInit == PSI_INIT_AVG_APPROX) {
    // PsiInitAvgApprox.runInitialization (D/init/PsiInitAvgApprox.java:47-99): per view the mean and max
    // of its middle x hyperslice; avg = mean of the view means, max[v] = the slice max
    double2* part = nullptr;
    MVR_HIP(hipMalloc(&part, sizeof(double2) * 256));
    double a = 0.0;
    for (int v = 0; v < V; ++v) {
      const int np = launch_psi_approx_slice(img[v], N, part, stream);
      MVR_HIP(hipGetLastError());
      std::vector<double2> h(np);
      MVR_HIP(hipMemcpyAsync(h.data(), part, sizeof(double2) * np, hipMemcpyDeviceToHost, stream));
      MVR_HIP(hipStreamSynchronize(stream));
      double sum = 0.0, mx = -1.0e308;
      for (const double2& p : h) { sum += p.x; mx = std::max(mx, p.y); }
      a += sum / ((double)N[1] * N[2]);
      maxI[v] = (float)mx;
    }
    (void)hipFree(part);
    avg = a / (double)V;
    if (std::isnan(avg)) avg = 1.0;
    launch_fill(psi, nvox(), (float)avg, stream);
  } else if (psiInit == PSI_INIT_BLURRED_FUSED) {
    // PsiInitBlurredFused.runInitialization (D/init/PsiInitBlurredFused.java:62-127): psi = weighted mean
    // of the positive views, blurred by a Gaussian of sigma 5 with mirror-single borders (separable, the
    // normalised half kernel of ImgLib2's Gauss3: H = max(2, (int)(3 sigma + 0.5) + 1) taps,
    // k[x] = exp(-x^2 / (2 sigma^2)) / (2 (0.5 + sum_{x>0} exp(...)))).  max[] and avg are the AvgPrecise ones.
    MVR_CHECK(psiCount > 0, "none of the views covers the deconvolved area");
    const double sigma = 5.0;
    const int H = std::max(2, (int)(3 * sigma + 0.5) + 1);
    std::vector<double> half(H);
    half[0] = 1.0;
    double sum = 0.5;
    for (int x = 1; x < H; ++x) {
      half[x] = std::exp(-(double)(x * x) / (2 * sigma * sigma));
      sum += half[x];
    }
    sum *= 2;
    for (double& k : half) k /= sum;
    double* dh = nullptr;
    MVR_HIP(hipMalloc(&dh, sizeof(double) * H));
    MVR_HIP(hipMemcpyAsync(dh, half.data(), sizeof(double) * H, hipMemcpyHostToDevice, stream));
    float* tmp = reinterpret_cast<float*>(S);  // the work spectra hold >= 4 bytes per voxel, unless blocks are dropped
    float* own = nullptr;
    if (blocked()) {
      size_t have = plan.ncpx() * 8;
      int nS = 1;
      for (int v = 0; v < V; ++v) nS = std::max(nS, nbView[v]);
      have *= (size_t)nS;
      if (have < nvox() * 4) {
        MVR_HIP(hipMalloc(&own, nvox() * 4));
        tmp = own;
      }
    }
    std::vector<const float*> ip(img.begin(), img.end()), wp(w.begin(), w.end());
    launch_fused_nonzero(ip.data(), wp.data(), V, nvox(), tmp, stream);
    launch_sym_conv_axis(tmp, psi, N, 0, dh, H, stream);
    launch_sym_conv_axis(psi, tmp, N, 1, dh, H, stream);
    launch_sym_conv_axis(tmp, psi, N, 2, dh, H, stream);
    MVR_HIP(hipGetLastError());
    MVR_HIP(hipStreamSynchronize(stream));
    (void)hipFree(dh);
    if (own) (void)hipFree(own);
  } else {
    MVR_CHECK(false, "unknown psi init");
  }
  MVR_HIP(hipMemcpyAsync(psi0, psi, nvox() * 4, hipMemcpyDeviceToDevice, stream));
  MVR_HIP(hipStreamSynchronize(stream));
  stats.clear();
  ready = true;
  return 0;
}

// Plan for the whole volume (FFT lengths >= Pmin where given), kernel spectra of the first nviews views,
// work spectra.
int Decon::build_plan_and_spectra(const int* Pmin, int nviews) {
  plan.destroy();
  int nS = 1;  // work spectra: one per block of the largest view list
  if (blocked()) {
    // a block's domain holds its output region plus both kernels' reach on either side, so that both chained
    // convolutions are exact on the effective region: P >= E + 2 (K - 1), output region at m = m1 + m2
    int D[3], Pm[3];
    for (int d = 0; d < 3; ++d) {
      blockE[d] = blockB[d] - 2 * K[d] + 2;
      MVR_CHECK(blockE[d] >= 1, "the block size must exceed 2 x kernel size - 2 (BlockGeneratorFixedSizePrecise)");
      MVR_CHECK(blockE[d] >= K[d] - 1 || blockE[d] >= N[d],
                "batched blocks need an effective block size of at least kernel size - 1");
      D[d] = std::min(blockE[d], N[d]);
      Pm[d] = std::max(D[d] + 2 * (K[d] - 1), Pmin ? Pmin[d] : 0);
    }
    if (plan.create(D, K, Pm)) return -1;
    for (int d = 0; d < 3; ++d) plan.dom.m[d] = 2 * (K[d] - 1 - K[d] / 2);
    if (build_blocks()) return -1;
    for (int v = 0; v < V; ++v) nS = std::max(nS, nbView[v]);
  } else {
    if (plan.create(N, K, Pmin)) return -1;
  }
  plan.wy = tunedW[0];
  plan.wz = tunedW[1];
  for (auto p : spec1) if (p) (void)hipFree(p);
  for (auto p : spec2) if (p) (void)hipFree(p);
  spec1.assign(V, nullptr);
  spec2.assign(V, nullptr);
  const size_t nk = (size_t)K[0] * K[1] * K[2];
  float* dk = nullptr;
  MVR_HIP(hipMalloc(&dk, nk * 4));
  for (int v = 0; v < nviews; ++v) {
    MVR_HIP(hipMalloc(&spec1[v], plan.ncpx() * 8));
    MVR_HIP(hipMalloc(&spec2[v], plan.ncpx() * 8));
    MVR_HIP(hipMemcpyAsync(dk, k1[v].data(), nk * 4, hipMemcpyHostToDevice, stream));
    if (plan.kernel_spectrum(dk, spec1[v], stream)) return -1;
    MVR_HIP(hipMemcpyAsync(dk, k2[v].data(), nk * 4, hipMemcpyHostToDevice, stream));
    if (plan.kernel_spectrum(dk, spec2[v], stream)) return -1;
    MVR_HIP(hipStreamSynchronize(stream));
  }
  (void)hipFree(dk);
  if (S) (void)hipFree(S);
  MVR_HIP(hipMalloc(&S, plan.ncpx() * 8 * (size_t)nS));
  if (scheme == SCHEME_MUL) {
    if (Spsi) (void)hipFree(Spsi);
    if (T2) (void)hipFree(T2);
    MVR_HIP(hipMalloc(&Spsi, plan.ncpx() * 8));
    MVR_HIP(hipMalloc(&T2, nvox() * 4 * V));
  }
  return 0;
}

void Decon::drop_graph() {
  if (graphExec) (void)hipGraphExecDestroy(graphExec);
  graphExec = nullptr;
}

int Decon::set_graph(bool on) {
  useGraph = on;
  if (!on) drop_graph();
  return 0;
}

// Coordinate descent over the compiled FFT lengths of each axis.  A trial rebuilds the plan and view 0's
// spectra and times view updates (Seq) or iterations (Mul) of view 0 with HIP events; psi is restored
// from psi0 afterwards.  On any failure the handle is put back on its previous schedule (plan, every view's
// spectra, work buffers), so that later runs use a complete plan; if even that fails, run() refuses until the
// next init.
int Decon::autotune(int candidates, int chosen[3], double* msOut) {
  MVR_CHECK(ready, "init first");
  MVR_CHECK(!blocked(), "autotuning runs on the one-frame schedule (set the block size to 0)");
  MVR_CHECK(candidates >= 1 && candidates <= 8, "candidates per axis must be in [1, 8]");
  MVR_HIP(hipSetDevice(dev));
  drop_graph();
  const int Vsave = V;
  const int wSave[2] = {tunedW[0], tunedW[1]};
  const int pSave[3] = {tunedP[0], tunedP[1], tunedP[2]};
  double2* st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int best[3] = {0, 0, 0};
  double bestMs = 1e30;
  // a host-side failure injection for tests of this recovery: MVR_TEST_FAIL_AUTOTUNE_TRIAL=k fails trial k
  const char* inj = std::getenv("MVR_TEST_FAIL_AUTOTUNE_TRIAL");
  const int failTrial = inj ? std::atoi(inj) : -1;
  int trialNo = 0;
  // a trial is the fastest of 3 separately timed view updates after a warm-up one (a transient slowdown of
  // the device does not pick the schedule); a candidate replaces the incumbent only when it is faster by
  // more than kMargin, so that timing noise does not trade a smaller grid for a larger one
  constexpr double kMargin = 0.01;
  auto trial = [&](const int P[3], double* ms) -> int {
    if (trialNo++ == failTrial) {
      set_error("autotune: injected trial failure");
      return -1;
    }
    if (build_plan_and_spectra(P, 1)) return -1;
    const bool timing = timer.on;
    timer.on = false;
    int rc = scheme == SCHEME_SEQ ? view_seq(0, st) : iteration_mul(st);
    double best = 1e30;
    for (int rep = 0; rep < 3 && !rc; ++rep) {
      float t = 0;
      if (hipEventRecord(e0, stream) != hipSuccess) {
        set_error("autotune: event record failed");
        rc = -1;
        break;
      }
      rc = scheme == SCHEME_SEQ ? view_seq(0, st) : iteration_mul(st);  // sets its own error
      if (rc) break;
      if (hipEventRecord(e1, stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
          hipEventElapsedTime(&t, e0, e1) != hipSuccess) {
        set_error("autotune: event timing failed");
        rc = -1;
        break;
      }
      best = std::min(best, (double)t);
    }
    timer.on = timing;
    if (rc) return -1;
    *ms = best;
    return 0;
  };
  int rc = 0;
  if (ensure_partials((int)std::max<size_t>(2048, (size_t)N[1] * N[2] / 4 + 1)) ||
      hipMalloc(&st, sizeof(double2) * 4) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
      hipEventCreate(&e1) != hipSuccess) {
    if (st == nullptr || e0 == nullptr || e1 == nullptr) set_error("autotune: device allocation failed");
    rc = -1;
  }
  if (!rc) {
    plan.destroy();
    if (plan.create(N, K, nullptr)) rc = -1;
  }
  if (!rc) {
    for (int d = 0; d < 3; ++d) best[d] = plan.dom.P[d];
    if (scheme == SCHEME_MUL) V = 1;  // time the shared forward transform + one view
    tunedW[0] = tunedW[1] = 0;
    if (trial(best, &bestMs)) rc = -1;
  }
  for (int d = 0; d < 3 && !rc; ++d) {
    const int need = N[d] + K[d] - 1;
    int tried = 1;
    for (const int* p = strided_lengths(); *p && tried < candidates && !rc; ++p) {
      if (*p <= best[d] || *p < need) continue;
      if (d == 0) {
        bool ok = false;
        for (const int* h = half_lengths(); *h; ++h) ok |= (*h * 2 == *p);
        if (!ok) continue;
      }
      int P[3] = {best[0], best[1], best[2]};
      P[d] = *p;
      double ms = 0;
      if (trial(P, &ms)) { rc = -1; break; }
      ++tried;
      if (ms < bestMs * (1.0 - kMargin)) {
        bestMs = ms;
        best[d] = *p;
      }
    }
  }
  // tile schedule of the strided passes: 8- or 16-column tiles per axis (lengths whose 16-column tile does
  // not fit the LDS budget run 8 either way)
  for (int a = 0; a < 2 && !rc; ++a) {
    int bestW = 0;
    for (const int wdt : {8, 16}) {
      tunedW[a] = wdt;
      double ms = 0;
      if (trial(best, &ms)) { rc = -1; break; }
      if (ms < bestMs * (1.0 - kMargin)) {
        bestMs = ms;
        bestW = wdt;
      }
    }
    tunedW[a] = bestW;
  }
  V = Vsave;
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (st) (void)hipFree(st);
  if (rc) {
    // back to the schedule in use before the call (the reason of the failure is kept)
    const std::string why = last_error();
    tunedW[0] = wSave[0];
    tunedW[1] = wSave[1];
    for (int d = 0; d < 3; ++d) tunedP[d] = pSave[d];
    if (build_plan_and_spectra(tunedP[0] ? tunedP : nullptr, V) ||
        hipMemcpyAsync(psi, psi0, nvox() * 4, hipMemcpyDeviceToDevice, stream) != hipSuccess ||
        hipStreamSynchronize(stream) != hipSuccess) {
      ready = false;
      set_error(why + "; restoring the previous schedule failed too: " + last_error() + " (init again)");
      return -1;
    }
    set_error(why);
    return -1;
  }
  for (int d = 0; d < 3; ++d) tunedP[d] = best[d];
  if (build_plan_and_spectra(tunedP, V)) {
    ready = false;
    return -1;
  }
  MVR_HIP(hipMemcpyAsync(psi, psi0, nvox() * 4, hipMemcpyDeviceToDevice, stream));
  MVR_HIP(hipStreamSynchronize(stream));
  if (chosen) for (int d = 0; d < 3; ++d) chosen[d] = best[d];
  if (msOut) *msOut = bestMs;
  return 0;
}

int Decon::reset_psi() {
  MVR_CHECK(ready, "init first");
  MVR_HIP(hipSetDevice(dev));
  MVR_HIP(hipMemcpyAsync(psi, psi0, nvox() * 4, hipMemcpyDeviceToDevice, stream));
  MVR_HIP(hipStreamSynchronize(stream));
  stats.clear();
  return 0;
}

// geometry of a strided launch whose kernel execution is timed under category c (untimed when c is off)
static StridedGeom timed(Timer& t, StridedGeom g, int c) {
  t.kernel(c, &g.evStart, &g.evStop);
  return g;
}

// One view of one Seq iteration over the whole psi volume.
int Decon::view_seq(int v, double2* statOut, int slot) {
  Dom d = plan.dom;
  StridedGeom gy = plan.geom_y(), gz = plan.geom_z();
  double2* part = slot >= 0 ? partial + (size_t)slot * partials_needed() : partial;
  if (blocked()) {
    // every block of this view in one launch per pass
    if (nbView[v] == 0) {  // no block with content: the view's statistics stay {0, -1}
      if (slot >= 0) {
        slotCount[slot] = 0;
      } else {
        launch_reduce_stats(partial, 0, statOut, stream);
        MVR_HIP(hipGetLastError());
      }
      return 0;
    }
    d.btab = btab[v];
    d.nb = nbView[v];
    d.blockStride = plan.ncpx();
    for (StridedGeom* gm : {&gy, &gz}) {
      gm->blockPlanes = gm->nplanes;
      gm->nplanes *= d.nb;
      gm->blockStride = plan.ncpx();
    }
  }
  GatherParams g{};
  g.src = psi;
  for (int k = 0; k < 3; ++k) { g.N[k] = N[k]; g.org[k] = -d.m[k]; }
  g.mode = GATHER_MIRROR;
  RatioParams q{};
  q.img = img[v];
  for (int k = 0; k < 3; ++k) {
    q.N[k] = N[k];
    q.o[k] = 0;
    if (blocked()) {  // the quotient around a block: the second kernel's reach beyond its output region
      q.qlo[k] = K[k] - 1 - K[k] / 2;
      q.qhi[k] = K[k] / 2;
    }
  }
  // the first convolution is read on the quotient rows, the second on the output rows; on one frame the
  // quotient pass leaves the rows of ones to the z pass after it
  const int zero[3] = {0, 0, 0};
  StridedGeom gy1 = gy, gz1 = gz, gy2 = gy, gz2 = gz, gzq = gz, gzf = gz;
  plan.restrict_rows(q.qlo, q.qhi, gy1, gz1);
  plan.restrict_rows(zero, zero, gy2, gz2);
  if (!blocked()) {
    plan.skip_ones_rows(q, gzq, &gy2);
    plan.mirror_rows(g, gzf, gy1);  // psi's mirror copies are read from their sources
  }
  timer.begin(stream, PASS_X_FWD);
  plan.xf(d, g, S, plan.twM, plan.wr, stream);
  timer.end(stream);
  plan.sz(timed(timer, gzf, PASS_Z_FWD), 0, S, S, nullptr, 1.f, plan.twZ, stream);
  plan.sy(timed(timer, gy1, PASS_Y_CONV), 2, S, S, spec1[v], 1.f, plan.twY, stream);
  plan.sz(timed(timer, gz1, PASS_Z_INV), 1, S, S, nullptr, 1.f, plan.twZ, stream);
  timer.begin(stream, PASS_X_RATIO);
  plan.xr(d, q, S, S, plan.twM, plan.wr, stream);
  timer.end(stream);
  plan.sz(timed(timer, gzq, PASS_Z_FWD), 0, S, S, nullptr, 1.f, plan.twZ, stream);
  plan.sy(timed(timer, gy2, PASS_Y_CONV), 2, S, S, spec2[v], 1.f, plan.twY, stream);
  plan.sz(timed(timer, gz2, PASS_Z_INV), 1, S, S, nullptr, 1.f, plan.twZ, stream);
  UpdateParams u{};
  u.psi = psi;
  u.w = w[v];
  for (int k = 0; k < 3; ++k) { u.N[k] = N[k]; u.o[k] = 0; }
  u.lambda = (double)lambda;
  u.minValue = minValue;
  u.maxI = maxI[v];
  u.partial = part;
  timer.begin(stream, PASS_X_UPDATE);
  const int np = plan.xu(d, u, S, plan.twM, plan.wr, stream);
  timer.end(stream);
  if (slot >= 0) {
    slotCount[slot] = np;
  } else {
    timer.begin(stream, PASS_REDUCE);
    launch_reduce_stats(partial, np, statOut, stream);
    timer.end(stream);
  }
  MVR_HIP(hipGetLastError());
  return 0;
}

// the statistics of the view updates in slots 0..n-1, in one launch
void Decon::reduce_slots(int n, double2* out) {
  timer.begin(stream, PASS_REDUCE);
  launch_reduce_stats_batched(partial, partials_needed(), slotCount, n, out, stream);
  timer.end(stream);
}

// One Mul iteration (ComputeBlockMulThreadCPU.runIteration semantics on the whole volume).
int Decon::iteration_mul(double2* statOut) {
  const Dom& d = plan.dom;
  const StridedGeom gy = plan.geom_y(), gz = plan.geom_z();
  // both convolutions are read on the output rows (the quotient and the stored result)
  const int zero[3] = {0, 0, 0};
  StridedGeom gyo = gy, gzo = gz;
  plan.restrict_rows(zero, zero, gyo, gzo);
  GatherParams g{};
  g.src = psi;
  for (int k = 0; k < 3; ++k) { g.N[k] = N[k]; g.org[k] = -d.m[k]; }
  g.mode = GATHER_MIRROR;
  StridedGeom gzf = gz, gyf = gyo;  // psi's mirror copies are read from their sources
  plan.mirror_rows(g, gzf, gyf);
  // the forward transform of psi is shared by all views
  timer.begin(stream, PASS_X_FWD);
  plan.xf(d, g, Spsi, plan.twM, plan.wr, stream);
  timer.end(stream);
  plan.sz(timed(timer, gzf, PASS_Z_FWD), 0, Spsi, Spsi, nullptr, 1.f, plan.twZ, stream);
  std::vector<const float*> t2p(V), wp(V);
  for (int v = 0; v < V; ++v) {
    plan.sy(timed(timer, gyf, PASS_Y_CONV), 2, Spsi, S, spec1[v], 1.f, plan.twY, stream);
    plan.sz(timed(timer, gzo, PASS_Z_INV), 1, S, S, nullptr, 1.f, plan.twZ, stream);
    RatioParams q{};
    q.img = img[v];
    for (int k = 0; k < 3; ++k) { q.N[k] = N[k]; q.o[k] = 0; }
    StridedGeom gzq = gz, gyq = gyo;
    plan.skip_ones_rows(q, gzq, &gyq);
    timer.begin(stream, PASS_X_RATIO);
    plan.xr(d, q, S, S, plan.twM, plan.wr, stream);
    timer.end(stream);
    plan.sz(timed(timer, gzq, PASS_Z_FWD), 0, S, S, nullptr, 1.f, plan.twZ, stream);
    plan.sy(timed(timer, gyq, PASS_Y_CONV), 2, S, S, spec2[v], 1.f, plan.twY, stream);
    plan.sz(timed(timer, gzo, PASS_Z_INV), 1, S, S, nullptr, 1.f, plan.twZ, stream);
    StoreParams st{T2 + nvox() * v};
    timer.begin(stream, PASS_X_STORE);
    plan.xs(d, st, S, plan.twM, plan.wr, stream);
    timer.end(stream);
    t2p[v] = T2 + nvox() * v;
    wp[v] = w[v];
  }
  double m = 0;
  for (int v = 0; v < V; ++v) m += maxI[v];
  m /= (double)V;
  const int o[3] = {0, 0, 0};
  int np = 0;
  if (ensure_partials(2048)) return -1;
  timer.begin(stream, PASS_UPDATE_MUL);
  launch_update_mul(psi, t2p.data(), wp.data(), V, N, o, N, (double)lambda, minValue, (float)m, partial, &np, stream);
  timer.end(stream);
  timer.begin(stream, PASS_REDUCE);
  launch_reduce_stats(partial, np, statOut, stream);
  timer.end(stream);
  MVR_HIP(hipGetLastError());
  return 0;
}

int Decon::run(int iterations) {
  MVR_CHECK(ready, "init first");
  MVR_CHECK(iterations >= 0, "negative iteration count");
  MVR_HIP(hipSetDevice(dev));
  const int perIt = scheme == SCHEME_SEQ ? V : 1;
  const int need = iterations * perIt;
  if (need > statsCap) {
    if (statsDev) (void)hipFree(statsDev);
    MVR_HIP(hipMalloc(&statsDev, sizeof(double2) * std::max(1, need)));
    statsCap = std::max(1, need);
  }
  // one slot of partials per view of an iteration, reduced in one launch per iteration
  if (ensure_partials((int)(partials_needed() * (size_t)(scheme == SCHEME_SEQ ? V : 1)))) return -1;
  if (useGraph && !timer.on && iterations > 0) {
    if (!graphStats) MVR_HIP(hipMalloc(&graphStats, sizeof(double2) * V));
    if (!graphExec) {
      // capture one iteration (all views): kernel launches only, no allocation or host synchronisation
      MVR_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
      int rc = 0;
      if (scheme == SCHEME_SEQ) {
        for (int v = 0; v < V && !rc; ++v) rc = view_seq(v, nullptr, v);
        if (!rc) reduce_slots(V, graphStats);
      } else {
        rc = iteration_mul(graphStats);
      }
      hipGraph_t g = nullptr;
      const hipError_t e = hipStreamEndCapture(stream, &g);
      if (rc || e != hipSuccess || !g) {
        if (g) (void)hipGraphDestroy(g);
        if (!rc) set_error(std::string("graph capture failed: ") + hipGetErrorString(e));
        return -1;
      }
      const hipError_t ie = hipGraphInstantiate(&graphExec, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (ie != hipSuccess) {
        graphExec = nullptr;
        set_error(std::string("graph instantiation failed: ") + hipGetErrorString(ie));
        return -1;
      }
    }
    for (int it = 0; it < iterations; ++it) {
      MVR_HIP(hipGraphLaunch(graphExec, stream));
      MVR_HIP(hipMemcpyAsync(statsDev + (size_t)it * perIt, graphStats, sizeof(double2) * perIt,
                             hipMemcpyDeviceToDevice, stream));
    }
  } else {
    for (int it = 0; it < iterations; ++it) {
      if (scheme == SCHEME_SEQ) {
        for (int v = 0; v < V; ++v)
          if (view_seq(v, nullptr, v)) return -1;
        reduce_slots(V, statsDev + (size_t)it * V);
        MVR_HIP(hipGetLastError());
      } else {
        if (iteration_mul(statsDev + it)) return -1;
      }
    }
  }
  std::vector<double2> h(std::max(need, 1));
  if (need > 0) MVR_HIP(hipMemcpyAsync(h.data(), statsDev, sizeof(double2) * need, hipMemcpyDeviceToHost, stream));
  MVR_HIP(hipStreamSynchronize(stream));
  for (int i = 0; i < need; ++i) { stats.push_back(h[i].x); stats.push_back(h[i].y); }
  timer.collect(nullptr, nullptr);
  return 0;
}

// One view pass of the Seq scheme (MultiViewDeconvolutionSeq.runNextIteration's view loop body,
// D/MultiViewDeconvolutionSeq.java:69-162) on its own; the statistics are appended like a run's.
int Decon::run_view(int v) {
  MVR_CHECK(ready, "init first");
  MVR_CHECK(scheme == SCHEME_SEQ, "single view updates belong to the Seq scheme");
  MVR_CHECK(v >= 0 && v < V, "view index out of range");
  MVR_HIP(hipSetDevice(dev));
  if (statsCap < 1) {
    if (statsDev) (void)hipFree(statsDev);
    statsDev = nullptr;
    MVR_HIP(hipMalloc(&statsDev, sizeof(double2)));
    statsCap = 1;
  }
  if (ensure_partials((int)partials_needed())) return -1;
  if (view_seq(v, statsDev)) return -1;
  double2 h;
  MVR_HIP(hipMemcpyAsync(&h, statsDev, sizeof(double2), hipMemcpyDeviceToHost, stream));
  MVR_HIP(hipStreamSynchronize(stream));
  stats.push_back(h.x);
  stats.push_back(h.y);
  timer.collect(nullptr, nullptr);
  return 0;
}

int Decon::commit_psi0() {
  MVR_CHECK(ready, "init first");
  MVR_HIP(hipSetDevice(dev));
  MVR_HIP(hipMemcpyAsync(psi0, psi, nvox() * 4, hipMemcpyDeviceToDevice, stream));
  MVR_HIP(hipStreamSynchronize(stream));
  stats.clear();
  return 0;
}

int Decon::set_reference(const float* r) {
  MVR_CHECK(r != nullptr, "null reference volume");
  MVR_HIP(hipSetDevice(dev));
  if (!ref) MVR_HIP(hipMalloc(&ref, nvox() * 4));
  if (!relDev) MVR_HIP(hipMalloc(&relDev, sizeof(double2) * 1025));
  MVR_HIP(hipMemcpyAsync(ref, r, nvox() * 4, hipMemcpyDefault, stream));
  MVR_HIP(hipStreamSynchronize(stream));
  return 0;
}

// rel-L2 = ||psi - ref||_2 / ||ref||_2, reduced in FP64 on the device
int Decon::rel_l2(double* out) {
  MVR_CHECK(ready && ref, "init and set a reference first");
  MVR_HIP(hipSetDevice(dev));
  launch_rel_l2(psi, ref, nvox(), relDev + 1, relDev, stream);
  MVR_HIP(hipGetLastError());
  double2 h;
  MVR_HIP(hipMemcpyAsync(&h, relDev, sizeof(double2), hipMemcpyDeviceToHost, stream));
  MVR_HIP(hipStreamSynchronize(stream));
  *out = h.y > 0 ? std::sqrt(h.x / h.y) : std::sqrt(h.x);
  return 0;
}

// Iterations until rel-L2 to the reference is <= tol (checked after every iteration), at most maxIterations.
int Decon::run_until(int maxIterations, double tol, int* iterationsOut, double* relOut) {
  double r = 0;
  if (rel_l2(&r)) return -1;
  int it = 0;
  while (r > tol && it < maxIterations) {
    if (run(1) || rel_l2(&r)) return -1;
    ++it;
  }
  if (iterationsOut) *iterationsOut = it;
  if (relOut) *relOut = r;
  return 0;
}

void Decon::destroy() {
  (void)hipSetDevice(dev);
  if (stream) (void)hipStreamSynchronize(stream);
  for (auto p : img) if (p) (void)hipFree(p);
  for (auto p : w) if (p) (void)hipFree(p);
  for (auto p : spec1) if (p) (void)hipFree(p);
  for (auto p : spec2) if (p) (void)hipFree(p);
  free_blocks();
  drop_graph();
  for (void* p : {(void*)psi, (void*)psi0, (void*)S, (void*)Spsi, (void*)T2, (void*)partial, (void*)statsDev,
                  (void*)ref, (void*)relDev, (void*)graphStats})
    if (p) (void)hipFree(p);
  graphStats = nullptr;
  img.clear(); w.clear(); spec1.clear(); spec2.clear();
  psi = psi0 = nullptr; S = Spsi = nullptr; T2 = nullptr; partial = nullptr; statsDev = nullptr;
  ref = nullptr; relDev = nullptr;
  plan.destroy();
  timer.release();
  if (stream) (void)hipStreamDestroy(stream);
  stream = nullptr;
}

// ============================================================ per-block compute thread

int BlockThread::create(int dev_, const int B_[3], float minValue_, float lambda_, int id_) {
  dev = dev_;
  id = id_;
  for (int d = 0; d < 3; ++d) {
    MVR_CHECK(B_[d] >= 1, "block size must be positive");
    B[d] = B_[d];
  }
  minValue = minValue_;
  lambda = lambda_;
  MVR_HIP(hipSetDevice(dev));
  MVR_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  const size_t nb = (size_t)B[0] * B[1] * B[2];
  MVR_HIP(hipMalloc(&psiB, nb * 4));
  MVR_HIP(hipMalloc(&imgB, nb * 4));
  MVR_HIP(hipMalloc(&wB, nb * 4));
  MVR_HIP(hipMalloc(&statDev, sizeof(double2)));
  return 0;
}

int BlockThread::ensure_partials(int n) {
  if (n <= partialCap) return 0;
  if (partial) (void)hipFree(partial);
  MVR_HIP(hipMalloc(&partial, sizeof(double2) * n));
  partialCap = n;
  return 0;
}

// DeconViewPSF.getKernel{1,2}FFT: spectra at block size, computed once per view
int BlockThread::set_view_kernels(int v, const float* K1, const float* K2, const int Kd[3]) {
  MVR_CHECK(v >= 0 && v < 1024, "view index out of range");
  MVR_HIP(hipSetDevice(dev));
  if (plan.xf == nullptr) {
    for (int d = 0; d < 3; ++d) K[d] = Kd[d];
    if (plan.create(B, K)) return -1;
    MVR_HIP(hipMalloc(&S, plan.ncpx() * 8));
  } else {
    for (int d = 0; d < 3; ++d) MVR_CHECK(K[d] == Kd[d], "all views must share one PSF size");
  }
  if ((int)spec1.size() <= v) {
    spec1.resize(v + 1, nullptr);
    spec2.resize(v + 1, nullptr);
  }
  const size_t nk = (size_t)K[0] * K[1] * K[2];
  float* dk = nullptr;
  MVR_HIP(hipMalloc(&dk, nk * 4));
  for (int which = 0; which < 2; ++which) {
    float2*& sp = which == 0 ? spec1[v] : spec2[v];
    if (!sp) MVR_HIP(hipMalloc(&sp, plan.ncpx() * 8));
    MVR_HIP(hipMemcpyAsync(dk, which == 0 ? K1 : K2, nk * 4, hipMemcpyDefault, stream));
    if (plan.kernel_spectrum(dk, sp, stream)) return -1;
  }
  MVR_HIP(hipStreamSynchronize(stream));
  (void)hipFree(dk);
  return 0;
}

// The whole image and weights of a view, resident for the thread's lifetime: the drivers cut every block of
// them again in every iteration (MultiViewDeconvolutionSeq.java:115-126), so they cross the bus once here.
int BlockThread::set_view_data(int v, const float* im, const float* wt, const int n[3]) {
  MVR_CHECK(v >= 0 && v < 1024, "view index out of range");
  MVR_CHECK(im && wt && n, "null view data");
  for (int d = 0; d < 3; ++d) {
    MVR_CHECK(n[d] >= 1, "view dimensions must be positive");
    if (VN[0] != 0) MVR_CHECK(VN[d] == n[d], "all views must share one size");
  }
  MVR_HIP(hipSetDevice(dev));
  for (int d = 0; d < 3; ++d) VN[d] = n[d];
  const size_t nv = (size_t)VN[0] * VN[1] * VN[2];
  if ((int)viewImg.size() <= v) {
    viewImg.resize(v + 1, nullptr);
    viewW.resize(v + 1, nullptr);
  }
  if (!viewImg[v]) MVR_HIP(hipMalloc(&viewImg[v], nv * 4));
  if (!viewW[v]) MVR_HIP(hipMalloc(&viewW[v], nv * 4));
  MVR_HIP(hipMemcpyAsync(viewImg[v], im, nv * 4, hipMemcpyDefault, stream));
  MVR_HIP(hipMemcpyAsync(viewW[v], wt, nv * 4, hipMemcpyDefault, stream));
  MVR_HIP(hipStreamSynchronize(stream));
  return 0;
}

// ComputeBlockSeqThreadCPU.runIteration on the device buffers psiB, imgB, wB (stats left in statDev)
int BlockThread::iterate_seq(int v, float maxI) {
  const Dom& d = plan.dom;
  const StridedGeom gy = plan.geom_y(), gz = plan.geom_z();
  // both convolutions are read on the block's rows (the quotient over the whole block, the update)
  const int zero[3] = {0, 0, 0};
  StridedGeom gyo = gy, gzo = gz;
  plan.restrict_rows(zero, zero, gyo, gzo);
  // convolve1: psi block with mirror-single extension of the block itself (ComputeBlockSeqThreadCPU.java:171-189)
  GatherParams g{};
  g.src = psiB;
  for (int k = 0; k < 3; ++k) { g.N[k] = B[k]; g.org[k] = -d.m[k]; }
  g.mode = GATHER_MIRROR;
  StridedGeom gzf = gz, gyf = gyo;  // the block's mirror copies are read from their sources
  plan.mirror_rows(g, gzf, gyf);
  plan.xf(d, g, S, plan.twM, plan.wr, stream);
  plan.sz(gzf, 0, S, S, nullptr, 1.f, plan.twZ, stream);
  plan.sy(gyf, 2, S, S, spec1[v], 1.f, plan.twY, stream);
  plan.sz(gzo, 1, S, S, nullptr, 1.f, plan.twZ, stream);
  RatioParams q{};
  q.img = imgB;
  for (int k = 0; k < 3; ++k) { q.N[k] = B[k]; q.o[k] = 0; }
  StridedGeom gzq = gz, gyq = gyo;
  plan.skip_ones_rows(q, gzq, &gyq);
  plan.xr(d, q, S, S, plan.twM, plan.wr, stream);
  plan.sz(gzq, 0, S, S, nullptr, 1.f, plan.twZ, stream);
  plan.sy(gyq, 2, S, S, spec2[v], 1.f, plan.twY, stream);
  plan.sz(gzo, 1, S, S, nullptr, 1.f, plan.twZ, stream);
  if (ensure_partials((int)(plan.rows_inner() / 4 + 1))) return -1;
  UpdateParams u{};
  u.psi = psiB;
  u.w = wB;
  for (int k = 0; k < 3; ++k) { u.N[k] = B[k]; u.o[k] = 0; }
  u.lambda = (double)lambda;
  u.minValue = minValue;
  u.maxI = maxI;
  u.partial = partial;
  const int np = plan.xu(d, u, S, plan.twM, plan.wr, stream);
  launch_reduce_stats(partial, np, statDev, stream);
  MVR_HIP(hipGetLastError());
  return 0;
}

int BlockThread::finish(float* psiBlock, double stats[2]) {
  const size_t nb = (size_t)B[0] * B[1] * B[2];
  double2 h;
  MVR_HIP(hipMemcpyAsync(psiBlock, psiB, nb * 4, hipMemcpyDefault, stream));
  MVR_HIP(hipMemcpyAsync(&h, statDev, sizeof(double2), hipMemcpyDeviceToHost, stream));
  MVR_HIP(hipStreamSynchronize(stream));
  if (stats) { stats[0] = h.x; stats[1] = h.y; }
  return 0;
}

int BlockThread::run_seq(int v, float* psiBlock, const float* imgBlock, const float* wBlock, float maxI,
                         double stats[2]) {
  MVR_CHECK(plan.xf != nullptr && v >= 0 && v < (int)spec1.size() && spec1[v], "set_view_kernels first");
  MVR_CHECK(psiBlock && imgBlock && wBlock, "null block");
  MVR_HIP(hipSetDevice(dev));
  const size_t nb = (size_t)B[0] * B[1] * B[2];
  MVR_HIP(hipMemcpyAsync(psiB, psiBlock, nb * 4, hipMemcpyDefault, stream));
  MVR_HIP(hipMemcpyAsync(imgB, imgBlock, nb * 4, hipMemcpyDefault, stream));
  MVR_HIP(hipMemcpyAsync(wB, wBlock, nb * 4, hipMemcpyDefault, stream));
  if (iterate_seq(v, maxI)) return -1;
  return finish(psiBlock, stats);
}

// the same iteration with the image and weight blocks cut on the device from the view's resident data
int BlockThread::run_seq_at(int v, float* psiBlock, const long long off[3], float maxI, double stats[2]) {
  MVR_CHECK(plan.xf != nullptr && v >= 0 && v < (int)spec1.size() && spec1[v], "set_view_kernels first");
  MVR_CHECK(v < (int)viewImg.size() && viewImg[v], "set_view_data for this view first");
  MVR_CHECK(psiBlock && off, "null block or offset");
  MVR_HIP(hipSetDevice(dev));
  const size_t nb = (size_t)B[0] * B[1] * B[2];
  MVR_HIP(hipMemcpyAsync(psiB, psiBlock, nb * 4, hipMemcpyDefault, stream));
  launch_copy_block_zero(viewImg[v], VN, off, imgB, B, stream);
  launch_copy_block_zero(viewW[v], VN, off, wB, B, stream);
  MVR_HIP(hipGetLastError());
  if (iterate_seq(v, maxI)) return -1;
  return finish(psiBlock, stats);
}

// ComputeBlockMulThreadCPU.runIteration on psiB and the per-view blocks imgV / wV (stats left in statDev)
int BlockThread::iterate_mul(int nv, const float* maxIs) {
  const size_t nb = (size_t)B[0] * B[1] * B[2];
  const Dom& d = plan.dom;
  const StridedGeom gy = plan.geom_y(), gz = plan.geom_z();
  const int zero[3] = {0, 0, 0};
  StridedGeom gyo = gy, gzo = gz;
  plan.restrict_rows(zero, zero, gyo, gzo);
  GatherParams g{};
  g.src = psiB;
  for (int k = 0; k < 3; ++k) { g.N[k] = B[k]; g.org[k] = -d.m[k]; }
  g.mode = GATHER_MIRROR;
  StridedGeom gzf = gz, gyf = gyo;  // the block's mirror copies are read from their sources
  plan.mirror_rows(g, gzf, gyf);
  plan.xf(d, g, Spsi, plan.twM, plan.wr, stream);
  plan.sz(gzf, 0, Spsi, Spsi, nullptr, 1.f, plan.twZ, stream);
  std::vector<const float*> t2p(nv), wp(nv);
  for (int v = 0; v < nv; ++v) {
    plan.sy(gyf, 2, Spsi, S, spec1[v], 1.f, plan.twY, stream);
    plan.sz(gzo, 1, S, S, nullptr, 1.f, plan.twZ, stream);
    RatioParams q{};
    q.img = imgV[v];
    for (int k = 0; k < 3; ++k) { q.N[k] = B[k]; q.o[k] = 0; }
    StridedGeom gzq = gz, gyq = gyo;
    plan.skip_ones_rows(q, gzq, &gyq);
    plan.xr(d, q, S, S, plan.twM, plan.wr, stream);
    plan.sz(gzq, 0, S, S, nullptr, 1.f, plan.twZ, stream);
    plan.sy(gyq, 2, S, S, spec2[v], 1.f, plan.twY, stream);
    plan.sz(gzo, 1, S, S, nullptr, 1.f, plan.twZ, stream);
    StoreParams st{T2 + nb * v};
    plan.xs(d, st, S, plan.twM, plan.wr, stream);
    t2p[v] = T2 + nb * v;
    wp[v] = wV[v];
  }
  // mean of the per-view maxima, rounded to float (ComputeBlockMulThreadCPU.java:144-149)
  double m = 0;
  for (int v = 0; v < nv; ++v) m += maxIs[v];
  m /= (double)nv;
  const int o[3] = {0, 0, 0};
  int np = 0;
  if (ensure_partials(2048)) return -1;
  launch_update_mul(psiB, t2p.data(), wp.data(), nv, B, o, B, (double)lambda, minValue, (float)m, partial, &np,
                    stream);
  launch_reduce_stats(partial, np, statDev, stream);
  MVR_HIP(hipGetLastError());
  return 0;
}

int BlockThread::prepare_mul(int nv) {
  MVR_CHECK(nv >= 1 && nv <= 32 && nv <= (int)spec1.size(), "set_view_kernels for every view first");
  for (int v = 0; v < nv; ++v) MVR_CHECK(spec1[v] != nullptr, "set_view_kernels for every view first");
  MVR_HIP(hipSetDevice(dev));
  const size_t nb = (size_t)B[0] * B[1] * B[2];
  while ((int)imgV.size() < nv) {
    float *a = nullptr, *b = nullptr;
    MVR_HIP(hipMalloc(&a, nb * 4));
    imgV.push_back(a);
    MVR_HIP(hipMalloc(&b, nb * 4));
    wV.push_back(b);
  }
  if (!T2) MVR_HIP(hipMalloc(&T2, nb * 4 * 32));
  if (!Spsi) MVR_HIP(hipMalloc(&Spsi, plan.ncpx() * 8));
  return 0;
}

int BlockThread::run_mul(int nv, float* psiBlock, const float* const* imgBlocks, const float* const* wBlocks,
                         const float* maxIs, double stats[2]) {
  MVR_CHECK(psiBlock && imgBlocks && wBlocks && maxIs, "null block");
  if (prepare_mul(nv)) return -1;
  const size_t nb = (size_t)B[0] * B[1] * B[2];
  MVR_HIP(hipMemcpyAsync(psiB, psiBlock, nb * 4, hipMemcpyDefault, stream));
  for (int v = 0; v < nv; ++v) {
    MVR_HIP(hipMemcpyAsync(imgV[v], imgBlocks[v], nb * 4, hipMemcpyDefault, stream));
    MVR_HIP(hipMemcpyAsync(wV[v], wBlocks[v], nb * 4, hipMemcpyDefault, stream));
  }
  if (iterate_mul(nv, maxIs)) return -1;
  return finish(psiBlock, stats);
}

int BlockThread::run_mul_at(int nv, float* psiBlock, const long long off[3], const float* maxIs, double stats[2]) {
  MVR_CHECK(psiBlock && off && maxIs, "null block, offset or maxima");
  for (int v = 0; v < nv; ++v)
    MVR_CHECK(v < (int)viewImg.size() && viewImg[v], "set_view_data for every view first");
  if (prepare_mul(nv)) return -1;
  const size_t nb = (size_t)B[0] * B[1] * B[2];
  MVR_HIP(hipMemcpyAsync(psiB, psiBlock, nb * 4, hipMemcpyDefault, stream));
  for (int v = 0; v < nv; ++v) {
    launch_copy_block_zero(viewImg[v], VN, off, imgV[v], B, stream);
    launch_copy_block_zero(viewW[v], VN, off, wV[v], B, stream);
  }
  MVR_HIP(hipGetLastError());
  if (iterate_mul(nv, maxIs)) return -1;
  return finish(psiBlock, stats);
}

void BlockThread::destroy() {
  (void)hipSetDevice(dev);
  if (stream) (void)hipStreamSynchronize(stream);
  for (auto p : spec1) if (p) (void)hipFree(p);
  for (auto p : spec2) if (p) (void)hipFree(p);
  for (auto p : imgV) if (p) (void)hipFree(p);
  for (auto p : wV) if (p) (void)hipFree(p);
  for (auto p : viewImg) if (p) (void)hipFree(p);
  for (auto p : viewW) if (p) (void)hipFree(p);
  for (void* p : {(void*)psiB, (void*)imgB, (void*)wB, (void*)T2, (void*)S, (void*)Spsi, (void*)partial,
                  (void*)statDev})
    if (p) (void)hipFree(p);
  spec1.clear(); spec2.clear(); imgV.clear(); wV.clear(); viewImg.clear(); viewW.clear();
  psiB = imgB = wB = T2 = nullptr; S = Spsi = nullptr; partial = nullptr; statDev = nullptr;
  plan.destroy();
  if (stream) (void)hipStreamDestroy(stream);
  stream = nullptr;
}

}  // namespace mvr
