// plan.cpp -- FFT-convolution plans, error state and pass timers.
#include <algorithm>
#include <cmath>
#include <vector>

#include "host.hpp"

namespace mvr {

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }
const char* last_error() { return g_err.c_str(); }

static int pick_length(int need, bool half) {
  for (const int* p = strided_lengths(); *p; ++p) {
    if (*p < need) continue;
    if (half) {
      bool ok = false;
      for (const int* h = half_lengths(); *h; ++h) ok |= (*h * 2 == *p);
      if (!ok) continue;
    }
    return *p;
  }
  return -1;
}

static int upload_twiddles(float2** dst, int n, int period) {
  std::vector<float2> h(n);
  for (int k = 0; k < n; ++k) {
    const double a = -2.0 * M_PI * (double)k / (double)period;
    h[k] = make_float2((float)std::cos(a), (float)std::sin(a));
  }
  MVR_HIP(hipMalloc(dst, sizeof(float2) * n));
  MVR_HIP(hipMemcpy(*dst, h.data(), sizeof(float2) * n, hipMemcpyHostToDevice));
  return 0;
}

int Plan::create(const int D[3], const int Kd[3], const int* Pmin) {
  if (create_geometry(D, Kd, Pmin)) return -1;
  if (upload_twiddles(&twM, M, M) || upload_twiddles(&wr, M + 1, dom.P[0]) || upload_twiddles(&twY, dom.P[1], dom.P[1]) ||
      upload_twiddles(&twZ, dom.P[2], dom.P[2]))
    return -1;
  return 0;
}

int Plan::create_geometry(const int D[3], const int Kd[3], const int* Pmin) {
  for (int d = 0; d < 3; ++d) {
    if (D[d] < 1 || Kd[d] < 1) { set_error("plan: empty domain or kernel"); return -1; }
    K[d] = Kd[d];
    dom.D[d] = D[d];
    dom.m[d] = K[d] - 1 - K[d] / 2;
    int need = D[d] + K[d] - 1;
    if (Pmin && Pmin[d] > need) need = Pmin[d];
    const int P = pick_length(need, d == 0);
    if (P < 0) { set_error("plan: no compiled FFT length >= " + std::to_string(need)); return -1; }
    dom.P[d] = P;
  }
  M = dom.P[0] / 2;
  Hx = M + 1;
  dom.HxP = (Hx + 15) / 16 * 16;
  xf = get_x_fwd(M);
  xr = get_x_ratio(M);
  xu = get_x_update(M);
  xs = get_x_store(M);
  sy = get_strided(dom.P[1]);
  sz = get_strided(dom.P[2]);
  if (!xf || !xr || !xu || !xs || !sy || !sz) { set_error("plan: missing kernel instantiation"); return -1; }
  // the strided passes address a lane's column and pencil position with a 32-bit byte offset from the plane
  if (((size_t)get_strided_span(dom.P[1]) * dom.HxP + 16) * sizeof(float2) > 0xffffffffull ||
      ((size_t)get_strided_span(dom.P[2]) * dom.P[1] * dom.HxP + 16) * sizeof(float2) > 0xffffffffull) {
    set_error("plan: xy plane of the FFT grid too large for 32-bit lane offsets");
    return -1;
  }
  return 0;
}

void Plan::destroy() {
  for (float2** p : {&twM, &wr, &twY, &twZ})
    if (*p) { (void)hipFree(*p); *p = nullptr; }
}

StridedGeom Plan::geom_y() const {
  StridedGeom g;
  g.planeStride = (size_t)dom.P[1] * dom.HxP;
  g.lineStride = dom.HxP;
  g.ncols = Hx;
  g.nplanes = dom.P[2];
  g.width = wy;
  return g;
}

StridedGeom Plan::geom_z() const {
  StridedGeom g;
  g.planeStride = dom.HxP;
  g.lineStride = (size_t)dom.P[1] * dom.HxP;
  g.ncols = Hx;
  g.nplanes = dom.P[1];
  g.width = wz;
  return g;
}

void Plan::restrict_rows(const int lo[3], const int hi[3], StridedGeom& gy, StridedGeom& gz) const {
  int a[3], b[3];
  for (int k = 1; k < 3; ++k) {
    a[k] = std::max(0, dom.m[k] - lo[k]);
    b[k] = std::min(dom.P[k], dom.m[k] + dom.D[k] + hi[k]);
  }
  gy.storeLo = a[1];
  gy.storeHi = b[1];
  const int nb = gz.blockPlanes ? gz.nplanes / gz.blockPlanes : 1;
  gz.plane0 = a[1];
  gz.nplanes = (b[1] - a[1]) * nb;
  if (gz.blockPlanes) gz.blockPlanes = b[1] - a[1];
  gz.storeLo = a[2];
  gz.storeHi = b[2];
}

void Plan::skip_ones_rows(RatioParams& q, StridedGeom& gz, StridedGeom* gy) const {
  // quotient rows (y, z) in grid coordinates, as the quotient pass selects them
  int lo[3], hi[3];
  for (int k = 1; k < 3; ++k) {
    lo[k] = std::max(dom.m[k] - q.qlo[k], dom.m[k] - q.o[k]);
    hi[k] = std::min(dom.m[k] + dom.D[k] + q.qhi[k], dom.m[k] - q.o[k] + q.N[k]);
  }
  q.skipOnes = 1;
  for (int k = 1; k < 3; ++k) {
    q.rowLo[k] = lo[k];
    q.rowN[k] = std::max(0, hi[k] - lo[k]);
  }
  gz.onesDC = (float)dom.P[0];  // the x transform of a row of ones
  gz.onesPlaneLo = lo[1];
  gz.onesPlaneHi = hi[1];
  gz.onesLo = lo[2];
  gz.onesHi = hi[2];
  if (gy) {
    gz.plane0 = lo[1];
    gz.nplanes = std::max(0, hi[1] - lo[1]);
    gy->onesDC = (float)dom.P[0] * (float)dom.P[2];  // the xz transform of a plane of ones
    gy->onesDCPlane = 0;
    gy->onesPlaneLo = 0;
    gy->onesPlaneHi = dom.P[2];
    gy->onesLo = lo[1];
    gy->onesHi = hi[1];
  }
}

void Plan::mirror_rows(GatherParams& g, StridedGeom& gzf, StridedGeom& gyc) const {
  // the strided passes read the mirror copies with 32-bit byte offsets from the plane: the z pencil's span, the
  // whole spectrum, must fit them (otherwise every row is transformed and read as written)
  if ((ncpx() + 16) * sizeof(float2) > 0xffffffffull) return;
  g.uniqueRows = 1;
  // unique rows: FFT positions [-org, -org + N) per axis, inside the grid
  gzf.plane0 = -g.org[1];
  gzf.nplanes = g.N[1];
  gzf.mirM = -g.org[2];
  gzf.mirN = g.N[2];
  gyc.mirM = -g.org[1];
  gyc.mirN = g.N[1];
}

int Plan::kernel_spectrum(const float* dKernel, float2* spec, hipStream_t s) const {
  GatherParams g{};
  g.src = dKernel;
  for (int d = 0; d < 3; ++d) { g.N[d] = K[d]; g.org[d] = 0; }
  g.mode = GATHER_KERNEL;
  xf(dom, g, spec, twM, wr, s);
  sy(geom_y(), 0, spec, spec, nullptr, 1.f, twY, s);
  sz(geom_z(), 3, spec, spec, nullptr, inv_scale(), twZ, s);
  MVR_HIP(hipGetLastError());
  return 0;
}

int Plan::convolve(const GatherParams& g, const float2* spec, float2* work, float* out, hipStream_t s) const {
  // the store reads the output rows only
  const int zero[3] = {0, 0, 0};
  StridedGeom gy = geom_y(), gz = geom_z(), gzo = gz;
  restrict_rows(zero, zero, gy, gzo);
  GatherParams gg = g;
  if (g.mode == GATHER_MIRROR) mirror_rows(gg, gz, gy);
  xf(dom, gg, work, twM, wr, s);
  sz(gz, 0, work, work, nullptr, 1.f, twZ, s);
  sy(gy, 2, work, work, spec, 1.f, twY, s);
  sz(gzo, 1, work, work, nullptr, 1.f, twZ, s);
  StoreParams st{out};
  xs(dom, st, work, twM, wr, s);
  MVR_HIP(hipGetLastError());
  return 0;
}

// ------------------------------------------------------------------ timers
void Timer::begin(hipStream_t s, int c) {
  pending = on && ((mask >> c) & 1u) && seen[c]++ % every == 0;
  if (!pending) return;
  if (used + 2 > ev.size()) {
    size_t old = ev.size();
    ev.resize(old + 256);
    for (size_t i = old; i < ev.size(); ++i) (void)hipEventCreate(&ev[i]);
  }
  (void)hipEventRecord(ev[used], s);
  cat.push_back(c);
  used++;
}
void Timer::kernel(int c, hipEvent_t* start, hipEvent_t* stop) {
  *start = *stop = nullptr;
  if (!on || !((mask >> c) & 1u) || seen[c]++ % every != 0) return;
  if (used + 2 > ev.size()) {
    size_t old = ev.size();
    ev.resize(old + 256);
    for (size_t i = old; i < ev.size(); ++i) (void)hipEventCreate(&ev[i]);
  }
  *start = ev[used];
  *stop = ev[used + 1];
  cat.push_back(c);
  used += 2;
}
void Timer::end(hipStream_t s) {
  if (!pending) return;
  (void)hipEventRecord(ev[used], s);
  used++;
  pending = false;
}
void Timer::reset(unsigned m, int e) {
  mask = m;
  on = m != 0;
  every = e < 1 ? 1 : e;
  for (int c = 0; c < PASS_COUNT; ++c) { ms[c] = 0; cnt[c] = 0; seen[c] = 0; }
}
int Timer::collect(double* outMs, long long* outN) {
  if (used > 0) {
    (void)hipEventSynchronize(ev[used - 1]);
    for (size_t i = 0; i + 1 < used; i += 2) {
      float t = 0.f;
      (void)hipEventElapsedTime(&t, ev[i], ev[i + 1]);
      ms[cat[i / 2]] += t;
      cnt[cat[i / 2]] += 1;
    }
    used = 0;
    cat.clear();
  }
  for (int c = 0; c < PASS_COUNT; ++c) {
    if (outMs) outMs[c] = ms[c];
    if (outN) outN[c] = cnt[c];
  }
  return 0;
}
void Timer::release() {
  for (auto& e : ev) (void)hipEventDestroy(e);
  ev.clear();
  cat.clear();
  used = 0;
}

}  // namespace mvr
