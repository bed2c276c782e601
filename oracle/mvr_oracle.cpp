// mvr_oracle.cpp -- CPU restatement of the reference's multiview Richardson-Lucy
// deconvolution path (PreibischLab/multiview-reconstruction, Java).
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.  The HIP
// product never links or calls it.
//
// Every routine is templated on the real type T so that the same algorithm runs
// in FP32 (the reference's arithmetic: float data, double sums) and in FP64
// (the "FP64 oracle": identical steps, every float replaced by double).
//
// Paths cited below are relative to src/main/java/ of the reference;
// D/ = net/preibisch/mvrecon/process/deconvolution/, C/ = .../process/cuda/,
// F/ = .../process/fusion/.
//
// Third-party arithmetic that the reference delegates to jars which are not
// vendored (imglib2-algorithm-fft + Mines JTK FFT, imglib2 RealSum) is
// restated from its published definition:
//   * FFT: any exact DFT; here a self-written Stockham mixed-radix FFT
//     (radix 2,3,4,5 and a direct DFT for larger prime factors) with
//     twiddles computed in double.  The padded length is the smallest even
//     2^a 3^b 5^c 7^d >= the required extent (JTK picks its own "fast" length;
//     the result is independent of the length in exact arithmetic).
//   * RealSum: Kahan compensated summation in double.
// Parity status of each function is listed in the header of oracle.py.

#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace mvo {

typedef int64_t i64;

// ---------------------------------------------------------------------------
// FusionTools.divideIntoPortions (F/FusionTools.java:1287-1329) with
// Threads.numThreads() replaced by the pinned numeric thread count nT.
// ---------------------------------------------------------------------------
struct Portion { i64 start, len; };

static std::vector<Portion> divide_into_portions(i64 n, int nT) {
  std::vector<Portion> out;
  int np;
  if (n <= nT) np = (int)n;
  else np = std::max(nT, (int)(n / (64LL * 64LL * 64LL)));
  if (n == 0) return out;
  i64 chunk = n / np;
  while (chunk == 0) { --np; chunk = n / np; }
  i64 mod = n % np;
  for (int i = 0; i < np; ++i)
    out.push_back({(i64)i * chunk, i == np - 1 ? chunk + mod : chunk});
  return out;
}

// imglib2 RealSum: Kahan-compensated double sum.
struct RealSum {
  double sum = 0, c = 0;
  void add(double v) { double y = v - c; double t = sum + y; c = (t - sum) - y; sum = t; }
};

// AdjustInput.sumImg (D/normalization/AdjustInput.java:71-122): per-portion
// RealSums, then total = sums[0] + sum_i sums[i] (first portion counted twice).
template <typename T>
static double sum_img_quirk(const T* a, i64 n, int nT) {
  auto portions = divide_into_portions(n, nT);
  std::vector<double> sums(portions.size());
  for (size_t p = 0; p < portions.size(); ++p) {
    RealSum s;
    for (i64 j = 0; j < portions[p].len; ++j) s.add((double)a[portions[p].start + j]);
    sums[p] = s.sum;
  }
  RealSum total;
  if (!sums.empty()) total.add(sums[0]);
  for (double s : sums) total.add(s);
  return total.sum;
}

// AdjustInput.normToSum1 (D/normalization/AdjustInput.java:52-58).
template <typename T>
static void norm_to_sum1(T* a, i64 n, int nT) {
  double sum = sum_img_quirk(a, n, nT);
  for (i64 i = 0; i < n; ++i) a[i] = (T)((double)a[i] / sum);
}

// Mirror.mirror over all three axes (D/util/Mirror.java:55-132), as executed
// sequentially in flat (x-fastest) order.  For odd sizes this is a flip; for
// even sizes the middle pair is swapped twice and therefore stays in place.
template <typename T>
static void mirror_all(T* a, const int dim[3]) {
  const i64 sx = dim[0], sy = dim[1], sz = dim[2];
  for (int d = 0; d < 3; ++d) {
    const i64 n = dim[d], half = n / 2, maxm = n - 1;
    for (i64 z = 0; z < sz; ++z)
      for (i64 y = 0; y < sy; ++y)
        for (i64 x = 0; x < sx; ++x) {
          i64 p[3] = {x, y, z};
          if (p[d] <= half) {
            i64 q[3] = {x, y, z};
            q[d] = maxm - p[d];
            std::swap(a[(p[2] * sy + p[1]) * sx + p[0]], a[(q[2] * sy + q[1]) * sx + q[0]]);
          }
        }
    (void)half;
  }
}

// ---------------------------------------------------------------------------
// FFT (restated third-party arithmetic; see file header)
// ---------------------------------------------------------------------------
template <typename T> using cpx = std::complex<T>;

static int fast_length(int n) {
  // smallest even 2^a 3^b 5^c 7^d >= n
  for (int m = std::max(2, n + (n & 1));; m += 2) {
    int r = m;
    for (int p : {2, 3, 5, 7}) while (r % p == 0) r /= p;
    if (r == 1) return m;
  }
}

template <typename T>
struct Plan1D {
  int n = 1;
  std::vector<int> radix;
  std::vector<cpx<T>> tw;  // exp(-2 pi i k / n)
  Plan1D() {}
  explicit Plan1D(int n_) : n(n_) {
    int m = n;
    while (m % 4 == 0) { radix.push_back(4); m /= 4; }
    while (m % 2 == 0) { radix.push_back(2); m /= 2; }
    for (int p : {3, 5, 7}) while (m % p == 0) { radix.push_back(p); m /= p; }
    for (int p = 11; m > 1; p += 2) while (m % p == 0) { radix.push_back(p); m /= p; }
    tw.resize(n);
    for (int k = 0; k < n; ++k) {
      double a = -2.0 * M_PI * (double)k / (double)n;
      tw[k] = cpx<T>((T)std::cos(a), (T)std::sin(a));
    }
  }
  // Stockham autosort; sign = -1 forward, +1 inverse (unnormalised).
  // Result is left in `a`; `b` is scratch of length n.
  void exec(cpx<T>* a, cpx<T>* b, int sign) const {
    cpx<T>* in = a;
    cpx<T>* out = b;
    int Ns = 1;
    cpx<T> v[64];
    for (int R : radix) {
      const int m = n / R;
      const int twstep = n / (Ns * R);
      for (int j = 0; j < m; ++j) {
        const int k = j % Ns;
        for (int r = 0; r < R; ++r) v[r] = in[j + r * m];
        if (Ns > 1)
          for (int r = 1; r < R; ++r) {
            cpx<T> w = tw[(i64)k * r * twstep % n];
            if (sign > 0) w = std::conj(w);
            v[r] = cmul(v[r], w);
          }
        dft(v, R, sign);
        const int base = (j / Ns) * Ns * R + k;
        for (int r = 0; r < R; ++r) out[base + r * Ns] = v[r];
      }
      std::swap(in, out);
      Ns *= R;
    }
    if (in != a) std::copy(in, in + n, a);
  }
  static inline cpx<T> cmul(cpx<T> x, cpx<T> y) {
    return cpx<T>(x.real() * y.real() - x.imag() * y.imag(), x.real() * y.imag() + x.imag() * y.real());
  }
  void dft(cpx<T>* v, int R, int sign) const {
    const T s = (T)sign;
    if (R == 2) {
      cpx<T> a = v[0], b = v[1];
      v[0] = a + b; v[1] = a - b;
    } else if (R == 4) {
      cpx<T> a0 = v[0] + v[2], a1 = v[0] - v[2], a2 = v[1] + v[3], a3 = v[1] - v[3];
      // multiply a3 by (sign * i): forward (-i)
      cpx<T> a3r(-s * a3.imag(), s * a3.real());
      v[0] = a0 + a2; v[2] = a0 - a2; v[1] = a1 + a3r; v[3] = a1 - a3r;
    } else if (R == 3) {
      const T c1 = (T)-0.5, s1 = (T)(0.86602540378443864676 * sign);
      cpx<T> t1 = v[1] + v[2], t2 = v[1] - v[2];
      cpx<T> m1 = v[0] + c1 * t1;
      cpx<T> m2(-s1 * t2.imag(), s1 * t2.real());
      v[0] = v[0] + t1; v[1] = m1 + m2; v[2] = m1 - m2;
    } else {
      cpx<T> r[64];
      const int step = n / R;
      for (int q = 0; q < R; ++q) {
        cpx<T> acc = 0;
        for (int p = 0; p < R; ++p) {
          cpx<T> w = tw[(i64)((p * q) % R) * step];
          if (sign > 0) w = std::conj(w);
          acc += cmul(v[p], w);
        }
        r[q] = acc;
      }
      for (int q = 0; q < R; ++q) v[q] = r[q];
    }
  }
};

// 3D real FFT on a px*py*pz grid (x fastest).  Half spectrum layout
// [z][y][hx], hx = px/2+1.  The real axis uses the half-length complex
// transform with the standard even/odd split.
template <typename T>
struct Plan3D {
  int px, py, pz, mx, hx;
  Plan1D<T> fx, fy, fz;
  std::vector<cpx<T>> wr;  // exp(-2 pi i k / px), k = 0..mx
  Plan3D(int px_, int py_, int pz_) : px(px_), py(py_), pz(pz_), mx(px_ / 2), hx(px_ / 2 + 1),
    fx(px_ / 2), fy(py_), fz(pz_) {
    wr.resize(mx + 1);
    for (int k = 0; k <= mx; ++k) {
      double a = -2.0 * M_PI * (double)k / (double)px;
      wr[k] = cpx<T>((T)std::cos(a), (T)std::sin(a));
    }
  }
  i64 nreal() const { return (i64)px * py * pz; }
  i64 ncpx() const { return (i64)hx * py * pz; }

  void r2c(const T* in, cpx<T>* out) const {
    const i64 rows = (i64)py * pz;
#pragma omp parallel
    {
      std::vector<cpx<T>> z(mx), s(mx);
#pragma omp for schedule(static)
      for (i64 row = 0; row < rows; ++row) {
        const T* r = in + row * px;
        for (int m = 0; m < mx; ++m) z[m] = cpx<T>(r[2 * m], r[2 * m + 1]);
        fx.exec(z.data(), s.data(), -1);
        cpx<T>* o = out + row * hx;
        const T h = (T)0.5;
        for (int k = 0; k <= mx; ++k) {
          cpx<T> zk = z[k % mx], zc = std::conj(z[(mx - k) % mx]);
          cpx<T> e = h * (zk + zc);
          cpx<T> d = zk - zc;                         // 2 i O[k]
          cpx<T> od(h * d.imag(), -h * d.real());     // O[k] = d / (2i)
          o[k] = e + Plan1D<T>::cmul(wr[k], od);
        }
      }
    }
    axis_pass(out, -1);
  }

  // in is destroyed; out gets the real result scaled by 1/(px*py*pz)
  void c2r(cpx<T>* in, T* out) const {
    axis_pass(in, +1);
    const i64 rows = (i64)py * pz;
    const T scale = (T)(1.0 / ((double)mx * py * pz)) ;
#pragma omp parallel
    {
      std::vector<cpx<T>> z(mx), s(mx);
#pragma omp for schedule(static)
      for (i64 row = 0; row < rows; ++row) {
        const cpx<T>* X = in + row * hx;
        const T h = (T)0.5;
        for (int k = 0; k < mx; ++k) {
          cpx<T> a = X[k], b = std::conj(X[mx - k]);
          cpx<T> e = h * (a + b);
          cpx<T> o = Plan1D<T>::cmul(h * (a - b), std::conj(wr[k]));
          z[k] = e + cpx<T>(-o.imag(), o.real());
        }
        fx.exec(z.data(), s.data(), +1);
        T* r = out + row * px;
        for (int m = 0; m < mx; ++m) { r[2 * m] = z[m].real() * scale; r[2 * m + 1] = z[m].imag() * scale; }
      }
    }
  }

  // complex transforms along y and z of a [pz][py][hx] array
  void axis_pass(cpx<T>* a, int sign) const {
    const int CW = 16;
    const int nchunks = (hx + CW - 1) / CW;
    // y axis (order: forward y then z; inverse z then y -- any order is exact)
    auto do_y = [&]() {
#pragma omp parallel
      {
        std::vector<cpx<T>> buf((size_t)CW * py), s(py);
#pragma omp for collapse(2) schedule(static)
        for (int z = 0; z < pz; ++z)
          for (int ch = 0; ch < nchunks; ++ch) {
            const int x0 = ch * CW, w = std::min(CW, hx - x0);
            cpx<T>* base = a + (i64)z * py * hx + x0;
            for (int y = 0; y < py; ++y)
              for (int c = 0; c < w; ++c) buf[(size_t)c * py + y] = base[(i64)y * hx + c];
            for (int c = 0; c < w; ++c) fy.exec(&buf[(size_t)c * py], s.data(), sign);
            for (int y = 0; y < py; ++y)
              for (int c = 0; c < w; ++c) base[(i64)y * hx + c] = buf[(size_t)c * py + y];
          }
      }
    };
    auto do_z = [&]() {
#pragma omp parallel
      {
        std::vector<cpx<T>> buf((size_t)CW * pz), s(pz);
#pragma omp for collapse(2) schedule(static)
        for (int y = 0; y < py; ++y)
          for (int ch = 0; ch < nchunks; ++ch) {
            const int x0 = ch * CW, w = std::min(CW, hx - x0);
            cpx<T>* base = a + (i64)y * hx + x0;
            const i64 zs = (i64)py * hx;
            for (int z = 0; z < pz; ++z)
              for (int c = 0; c < w; ++c) buf[(size_t)c * pz + z] = base[z * zs + c];
            for (int c = 0; c < w; ++c) fz.exec(&buf[(size_t)c * pz], s.data(), sign);
            for (int z = 0; z < pz; ++z)
              for (int c = 0; c < w; ++c) base[z * zs + c] = buf[(size_t)c * pz + z];
          }
      }
    };
    if (sign < 0) { do_y(); do_z(); } else { do_z(); do_y(); }
  }
};

// ---------------------------------------------------------------------------
// util.FFTConvolution semantics (util/FFTConvolution.java:490-666).
// out[x] = sum_j K[j] * ext(img)[x + c - j], c = floor(K/2) per axis;
// padded length P >= N + K - 1, image centred in the padded window, kernel
// centre moved to the origin by periodic extension.
// ---------------------------------------------------------------------------
enum Ext { EXT_MIRROR = 0, EXT_VALUE = 1 };  // EXT_VALUE with value 0 = extendZero

static inline i64 mirror_index(i64 p, i64 n) {
  if (n == 1) return 0;
  const i64 period = 2 * n - 2;
  p %= period;
  if (p < 0) p += period;
  if (p >= n) p = period - p;
  return p;
}

template <typename T>
struct KernelFFT {
  int P[3];
  std::vector<cpx<T>> spec;
};

// FFTConvolution.setupFFTs + computeKernelFFT (util/FFTConvolution.java:508-582)
template <typename T>
static void kernel_fft(const T* k, const int kd[3], const int imgd[3], KernelFFT<T>& out) {
  for (int d = 0; d < 3; ++d) out.P[d] = fast_length(imgd[d] + kd[d] - 1);
  Plan3D<T> plan(out.P[0], out.P[1], out.P[2]);
  std::vector<T> buf(plan.nreal(), (T)0);
  int c[3] = {kd[0] / 2, kd[1] / 2, kd[2] / 2};
  for (int z = 0; z < kd[2]; ++z)
    for (int y = 0; y < kd[1]; ++y)
      for (int x = 0; x < kd[0]; ++x) {
        i64 px = ((x - c[0]) % out.P[0] + out.P[0]) % out.P[0];
        i64 py = ((y - c[1]) % out.P[1] + out.P[1]) % out.P[1];
        i64 pz = ((z - c[2]) % out.P[2] + out.P[2]) % out.P[2];
        buf[(pz * out.P[1] + py) * out.P[0] + px] = k[((i64)z * kd[1] + y) * kd[0] + x];
      }
  out.spec.resize(plan.ncpx());
  plan.r2c(buf.data(), out.spec.data());
}

// FFTConvolution.convolve with a precomputed kernel spectrum
// (util/FFTConvolution.java:490-506, 584-603, 659-666).
template <typename T>
static void fft_convolve(const T* img, const int nd[3], int ext, T extValue,
                         const KernelFFT<T>& kf, T* out) {
  const int* P = kf.P;
  Plan3D<T> plan(P[0], P[1], P[2]);
  std::vector<T> buf(plan.nreal());
  int off[3];
  for (int d = 0; d < 3; ++d) off[d] = (P[d] - nd[d]) / 2;
#pragma omp parallel for schedule(static)
  for (int z = 0; z < P[2]; ++z)
    for (int y = 0; y < P[1]; ++y) {
      const i64 gz = z - off[2], gy = y - off[1];
      T* row = buf.data() + ((i64)z * P[1] + y) * P[0];
      for (int x = 0; x < P[0]; ++x) {
        const i64 gx = x - off[0];
        T v;
        if (ext == EXT_MIRROR) {
          v = img[(mirror_index(gz, nd[2]) * nd[1] + mirror_index(gy, nd[1])) * nd[0] + mirror_index(gx, nd[0])];
        } else {
          if (gx < 0 || gy < 0 || gz < 0 || gx >= nd[0] || gy >= nd[1] || gz >= nd[2]) v = extValue;
          else v = img[(gz * nd[1] + gy) * nd[0] + gx];
        }
        row[x] = v;
      }
    }
  std::vector<cpx<T>> spec(plan.ncpx());
  plan.r2c(buf.data(), spec.data());
  const i64 nc = plan.ncpx();
#pragma omp parallel for schedule(static)
  for (i64 i = 0; i < nc; ++i) spec[i] = Plan1D<T>::cmul(spec[i], kf.spec[i]);
  plan.c2r(spec.data(), buf.data());
#pragma omp parallel for schedule(static)
  for (int z = 0; z < nd[2]; ++z)
    for (int y = 0; y < nd[1]; ++y)
      for (int x = 0; x < nd[0]; ++x)
        out[((i64)z * nd[1] + y) * nd[0] + x] =
            buf[((i64)(z + off[2]) * P[1] + (y + off[1])) * P[0] + (x + off[0])];
}

// zero-extended same-size convolution of two kernels (used by DeconViewPSF.init)
template <typename T>
static std::vector<T> conv_same_zero(const std::vector<T>& input, const int id[3],
                                     const std::vector<T>& kernel, const int kd[3]) {
  KernelFFT<T> kf;
  kernel_fft(kernel.data(), kd, id, kf);
  std::vector<T> out(input.size());
  fft_convolve(input.data(), id, EXT_VALUE, (T)0, kf, out.data());
  return out;
}

// ---------------------------------------------------------------------------
// DeconvolutionMethods (D/iteration/sequential/DeconvolutionMethods.java)
// ---------------------------------------------------------------------------
// Math.max(double, double): a NaN argument is the result (std::max would drop a NaN second argument)
static inline double jmax(double a, double b) { return std::isnan(a) ? a : std::isnan(b) ? b : std::max(a, b); }

static inline double tikhonov(double value, double lambda) {
  return (std::sqrt(1.0 + 2.0 * lambda * value) - 1.0) / lambda;
}

// computeNextValue (:320-358)
template <typename T>
static inline T next_value(T last, T integral, T weight, double lambda, T minI, T maxI) {
  const T value = last * integral;
  T adjusted;
  if (value > 0) {
    if (lambda > 0) adjusted = (T)tikhonov((double)(value / maxI), lambda) * maxI;
    else adjusted = value;
  } else {
    adjusted = minI;
  }
  T next;
  if (std::isnan((double)adjusted)) next = minI;
  else next = std::max(minI, adjusted);
  return last + ((next - last) * weight);
}

// computeNextValueMul (:370-419)
template <typename T>
static inline T next_value_mul(T last, const T* integral, const T* weight, int nv, double lambda, T minI, T maxI) {
  double sumW = 0, prod = 1;
  for (int i = 0; i < nv; ++i) { prod *= (double)integral[i]; sumW += (double)weight[i]; }
  prod = std::pow(prod, 1.0 / nv);
  sumW = std::min(1.0, sumW);
  const T value = last * (T)prod;
  T adjusted;
  if (value > 0) {
    if (lambda > 0) adjusted = (T)tikhonov((double)(value / maxI), lambda) * maxI;
    else adjusted = value;
  } else {
    adjusted = minI;
  }
  T next;
  if (std::isnan((double)adjusted)) next = minI;
  else next = std::max(minI, adjusted);
  return last + ((next - last) * (T)sumW);
}

// ---------------------------------------------------------------------------
// Blocks: BlockGeneratorFixedSizePrecise (C/BlockGeneratorFixedSizePrecise.java:59-131),
// BlockSorter (C/BlockSorter.java:55-143), DeconView.filterBlocksForContent
// (D/DeconView.java:204-274), Block copy/paste (C/Block.java:158-403).
// ---------------------------------------------------------------------------
struct Block {
  i64 blockSize[3], offset[3], effSize[3], effOffset[3], effLocal[3];
};

static bool divide_into_blocks(const i64 img[3], const i64 kernelSize[3], const i64 blockSize[3],
                               std::vector<Block>& out) {
  i64 eff[3], loc[3], nb[3];
  for (int d = 0; d < 3; ++d) {
    eff[d] = blockSize[d] - kernelSize[d] + 1;
    if (eff[d] <= 0) return false;
    loc[d] = kernelSize[d] / 2;
  }
  for (int d = 0; d < 3; ++d) { nb[d] = img[d] / eff[d]; if (img[d] % eff[d] != 0) ++nb[d]; }
  out.clear();
  for (i64 bz = 0; bz < nb[2]; ++bz)
    for (i64 by = 0; by < nb[1]; ++by)
      for (i64 bx = 0; bx < nb[0]; ++bx) {
        i64 cb[3] = {bx, by, bz};
        Block b;
        for (int d = 0; d < 3; ++d) {
          b.blockSize[d] = blockSize[d];
          b.effSize[d] = eff[d];
          b.effOffset[d] = cb[d] * eff[d];
          b.offset[d] = b.effOffset[d] - kernelSize[d] / 2;
          if (b.effOffset[d] + b.effSize[d] > img[d]) b.effSize[d] = img[d] - b.effOffset[d];
          b.effLocal[d] = loc[d];
        }
        out.push_back(b);
      }
  return true;
}

static std::vector<std::vector<int>> sort_blocks(const std::vector<Block>& blocks, const i64 img[3], int minRequired,
                                                 int* chosenDim) {
  const int n = 3;
  i64 nb[3];
  for (int d = 0; d < n; ++d) {
    nb[d] = img[d] / blocks[0].effSize[d];
    if (img[d] % blocks[0].effSize[d] != 0) ++nb[d];
  }
  // HashMap<Long,Integer>: later dimensions overwrite earlier ones with the same product
  std::vector<std::pair<i64, int>> map;
  std::vector<i64> sizes;
  for (int d = 0; d < n; ++d) {
    i64 size = 1;
    for (int e = 0; e < n; ++e) if (e != d) size *= nb[e];
    sizes.push_back(size);
    bool found = false;
    for (auto& kv : map) if (kv.first == size) { kv.second = d; found = true; }
    if (!found) map.push_back({size, d});
  }
  std::sort(sizes.begin(), sizes.end());
  int minDim = -1;
  for (int i = 0; i < n && minDim == -1; ++i)
    if (sizes[i] >= minRequired || (i == n - 1 && minDim == -1))
      for (auto& kv : map) if (kv.first == sizes[i]) minDim = kv.second;
  if (chosenDim) *chosenDim = minDim;
  std::vector<std::vector<int>> batches;
  const i64 minOffset = blocks[0].offset[minDim];
  size_t sum = 0;
  for (i64 i = 0; i < nb[minDim]; ++i) {
    std::vector<int> layer;
    const i64 off = minOffset + i * blocks[0].effSize[minDim];
    for (size_t b = 0; b < blocks.size(); ++b)
      if (blocks[b].offset[minDim] == off) { layer.push_back((int)b); ++sum; }
    batches.push_back(layer);
  }
  if (sum != blocks.size()) {
    batches.clear();
    std::vector<int> all;
    for (size_t b = 0; b < blocks.size(); ++b) all.push_back((int)b);
    batches.push_back(all);
  }
  return batches;
}

static bool block_contains_content(const Block& b, const float* w, const i64 N[3]) {
  for (i64 z = 0; z < b.blockSize[2]; ++z)
    for (i64 y = 0; y < b.blockSize[1]; ++y)
      for (i64 x = 0; x < b.blockSize[0]; ++x) {
        i64 gx = b.offset[0] + x, gy = b.offset[1] + y, gz = b.offset[2] + z;
        if (gx < 0 || gy < 0 || gz < 0 || gx >= N[0] || gy >= N[1] || gz >= N[2]) continue;
        if (w[(gz * N[1] + gy) * N[0] + gx] != 0.0f) return true;
      }
  return false;
}

// Block.copyBlock with source = extendMirrorSingle(psi) (D/MultiViewDeconvolutionSeq.java:115)
template <typename T>
static void copy_block_mirror(const T* psi, const i64 N[3], const Block& b, T* blk) {
  const i64 bx = b.blockSize[0], by = b.blockSize[1], bz = b.blockSize[2];
#pragma omp parallel for schedule(static)
  for (i64 z = 0; z < bz; ++z)
    for (i64 y = 0; y < by; ++y) {
      const i64 gz = mirror_index(b.offset[2] + z, N[2]), gy = mirror_index(b.offset[1] + y, N[1]);
      for (i64 x = 0; x < bx; ++x)
        blk[(z * by + y) * bx + x] = psi[(gz * N[1] + gy) * N[0] + mirror_index(b.offset[0] + x, N[0])];
    }
}

// Views.zeroMin(Views.interval(Views.extendZero(img), block)) as a dense array
template <typename T>
static void copy_block_zero(const float* src, const i64 N[3], const Block& b, T* blk) {
  const i64 bx = b.blockSize[0], by = b.blockSize[1], bz = b.blockSize[2];
#pragma omp parallel for schedule(static)
  for (i64 z = 0; z < bz; ++z)
    for (i64 y = 0; y < by; ++y)
      for (i64 x = 0; x < bx; ++x) {
        const i64 gx = b.offset[0] + x, gy = b.offset[1] + y, gz = b.offset[2] + z;
        T v = 0;
        if (gx >= 0 && gy >= 0 && gz >= 0 && gx < N[0] && gy < N[1] && gz < N[2])
          v = (T)src[(gz * N[1] + gy) * N[0] + gx];
        blk[(z * by + y) * bx + x] = v;
      }
}

// Block.pasteBlock (C/Block.java:357-403)
template <typename T>
static void paste_block(T* psi, const i64 N[3], const Block& b, const T* blk) {
  const i64 bx = b.blockSize[0], by = b.blockSize[1];
#pragma omp parallel for schedule(static)
  for (i64 z = 0; z < b.effSize[2]; ++z)
    for (i64 y = 0; y < b.effSize[1]; ++y)
      for (i64 x = 0; x < b.effSize[0]; ++x) {
        const i64 lz = z + b.effLocal[2], ly = y + b.effLocal[1], lx = x + b.effLocal[0];
        const i64 gz = z + b.effOffset[2], gy = y + b.effOffset[1], gx = x + b.effOffset[0];
        psi[(gz * N[1] + gy) * N[0] + gx] = blk[(lz * by + ly) * bx + lx];
      }
  (void)N;
}

// ---------------------------------------------------------------------------
// DeconViewPSF.init (D/DeconViewPSF.java:119-284) for all views in list order
// (D/DeconViews.java:69-70).  psfType: 0 OPTIMIZATION_II, 1 OPTIMIZATION_I,
// 2 EFFICIENT_BAYESIAN, 3 INDEPENDENT (enum order of DeconViewPSF.PSFTYPE).
// ---------------------------------------------------------------------------
enum { OPT_II = 0, OPT_I = 1, EFFICIENT_BAYESIAN = 2, INDEPENDENT = 3 };

template <typename T>
static void init_psfs(std::vector<std::vector<T>>& k1, const std::vector<std::array<int, 3>>& kd, int psfType, int nT,
                      std::vector<std::vector<T>>& k2) {
  const int V = (int)k1.size();
  k2.assign(V, {});
  for (int v = 0; v < V; ++v) {
    const int* d = kd[v].data();
    const i64 n = (i64)d[0] * d[1] * d[2];
    norm_to_sum1(k1[v].data(), n, nT);
    auto inverted = [&](const std::vector<T>& a, const int* dd) {
      std::vector<T> r(a);
      mirror_all(r.data(), dd);
      return r;
    };
    if (V == 1 || psfType == INDEPENDENT) {
      k2[v] = inverted(k1[v], d);
    } else if (psfType == EFFICIENT_BAYESIAN) {
      std::vector<T> tmp = inverted(k1[v], d);
      for (int w = 0; w < V; ++w) {
        if (w == v) continue;
        std::vector<T> input = inverted(k1[v], d);
        std::vector<T> out1 = conv_same_zero(input, d, k1[w], kd[w].data());
        std::vector<T> kinv = inverted(k1[w], kd[w].data());
        std::vector<T> out2 = conv_same_zero(out1, d, kinv, kd[w].data());
        for (i64 i = 0; i < n; ++i) tmp[i] = out2[i] * tmp[i];
      }
      norm_to_sum1(tmp.data(), n, nT);
      k2[v] = tmp;
    } else if (psfType == OPT_I) {
      std::vector<T> tmp(k1[v]);
      for (int w = 0; w < V; ++w) {
        if (w == v) continue;
        std::vector<T> kinv = inverted(k1[w], kd[w].data());
        std::vector<T> out = conv_same_zero(k1[v], d, kinv, kd[w].data());
        for (i64 i = 0; i < n; ++i) tmp[i] = out[i] * tmp[i];
      }
      norm_to_sum1(tmp.data(), n, nT);
      k2[v] = inverted(tmp, d);
    } else {  // OPT_II
      std::vector<T> e(k1[v]);
      for (i64 i = 0; i < n; ++i) {
        T r = k1[v][i];
        for (int p = 1; p < V; ++p) r *= k1[v][i];
        e[i] = r;
      }
      norm_to_sum1(e.data(), n, nT);
      k2[v] = inverted(e, d);
    }
  }
}

// ---------------------------------------------------------------------------
// PsiInitAvgPrecise (D/init/PsiInitAvgPrecise.java:52-112,
// D/init/PsiInitAvgPreciseThread.java:127-155)
// ---------------------------------------------------------------------------
static double psi_init_avg_precise(const float* const* imgs, int V, i64 n, int nT, float* maxOut) {
  auto portions = divide_into_portions(n, nT);
  RealSum total;
  i64 count = 0;
  for (int v = 0; v < V; ++v) maxOut[v] = 0.0f;
  for (auto& p : portions) {
    RealSum rs;
    i64 cnt = 0;
    std::vector<float> mx(V, 0.0f);
    for (i64 j = 0; j < p.len; ++j) {
      const i64 i = p.start + j;
      double sum = 0;
      int c = 0;
      for (int v = 0; v < V; ++v) {
        const float val = imgs[v][i];
        if (val > 0) { mx[v] = std::max(mx[v], val); sum += val; ++c; }
      }
      if (c > 0) { rs.add(sum / c); ++cnt; }
    }
    total.add(rs.sum);
    count += cnt;
    for (int v = 0; v < V; ++v) maxOut[v] = std::max(maxOut[v], mx[v]);
  }
  double avg = total.sum / (double)count;
  if (std::isnan(avg)) avg = 1.0;
  return avg;
}

// ---------------------------------------------------------------------------
// Block iteration of one view (ComputeBlockSeqThreadCPU.runIteration,
// D/iteration/sequential/ComputeBlockSeqThreadCPU.java:79-209)
// ---------------------------------------------------------------------------
template <typename T>
static void block_iteration_seq(T* psiBlk, const int bd[3], const T* imgBlk, const T* wBlk,
                                const KernelFFT<T>& k1f, const KernelFFT<T>& k2f, double lambda, T minValue,
                                T maxI, int nT, double stats[2]) {
  const i64 n = (i64)bd[0] * bd[1] * bd[2];
  std::vector<T> tmp1(n), tmp2(n);
  fft_convolve(psiBlk, bd, EXT_MIRROR, (T)0, k1f, tmp1.data());
  // computeQuotient (DeconvolutionMethods.java:46-98)
#pragma omp parallel for schedule(static)
  for (i64 i = 0; i < n; ++i) {
    const T imgV = imgBlk[i];
    tmp1[i] = imgV > 0 ? imgV / tmp1[i] : (T)1;
  }
  fft_convolve(tmp1.data(), bd, EXT_VALUE, (T)1, k2f, tmp2.data());
  // computeFinalValues per portion (DeconvolutionMethods.java:103-182)
  auto portions = divide_into_portions(n, nT);
  std::vector<double> ps(portions.size()), pm(portions.size());
#pragma omp parallel for schedule(static)
  for (size_t p = 0; p < portions.size(); ++p) {
    double sumChange = 0, maxChange = -1;
    for (i64 j = 0; j < portions[p].len; ++j) {
      const i64 i = portions[p].start + j;
      const T last = psiBlk[i];
      const T next = next_value(last, tmp2[i], wBlk[i], lambda, minValue, maxI);
      psiBlk[i] = next;
      const T change = next - last;
      sumChange += (double)change;
      maxChange = jmax(maxChange, (double)change);
    }
    ps[p] = sumChange;
    pm[p] = maxChange;
  }
  stats[0] = 0;
  stats[1] = -1;
  for (size_t p = 0; p < portions.size(); ++p) { stats[0] += ps[p]; stats[1] = jmax(stats[1], pm[p]); }
}

// ComputeBlockMulThreadCPU.runIteration (D/iteration/mul/ComputeBlockMulThreadCPU.java:87-188)
template <typename T>
static void block_iteration_mul(T* psiBlk, const int bd[3], const std::vector<std::vector<T>>& imgBlk,
                                const std::vector<std::vector<T>>& wBlk, const std::vector<KernelFFT<T>>& k1f,
                                const std::vector<KernelFFT<T>>& k2f, double lambda, T minValue, const float* maxI,
                                int nT, double stats[2]) {
  const int V = (int)imgBlk.size();
  const i64 n = (i64)bd[0] * bd[1] * bd[2];
  std::vector<std::vector<T>> tmp1(V, std::vector<T>(n)), tmp2(V, std::vector<T>(n));
  for (int v = 0; v < V; ++v) fft_convolve(psiBlk, bd, EXT_MIRROR, (T)0, k1f[v], tmp1[v].data());
  for (int v = 0; v < V; ++v) {
#pragma omp parallel for schedule(static)
    for (i64 i = 0; i < n; ++i) {
      const T imgV = imgBlk[v][i];
      tmp1[v][i] = imgV > 0 ? imgV / tmp1[v][i] : (T)1;
    }
  }
  for (int v = 0; v < V; ++v) fft_convolve(tmp1[v].data(), bd, EXT_VALUE, (T)1, k2f[v], tmp2[v].data());
  double mi = 0;
  for (int v = 0; v < V; ++v) mi += maxI[v];
  mi /= (double)V;
  const T miv = (T)(float)mi;
  auto portions = divide_into_portions(n, nT);
  std::vector<double> ps(portions.size()), pm(portions.size());
#pragma omp parallel for schedule(static)
  for (size_t p = 0; p < portions.size(); ++p) {
    double sumChange = 0, maxChange = -1;
    std::vector<T> iv(V), wv(V);
    for (i64 j = 0; j < portions[p].len; ++j) {
      const i64 i = portions[p].start + j;
      for (int v = 0; v < V; ++v) { iv[v] = tmp2[v][i]; wv[v] = wBlk[v][i]; }
      const T last = psiBlk[i];
      const T next = next_value_mul(last, iv.data(), wv.data(), V, lambda, minValue, miv);
      psiBlk[i] = next;
      const T change = next - last;
      sumChange += (double)change;
      maxChange = jmax(maxChange, (double)change);
    }
    ps[p] = sumChange;
    pm[p] = maxChange;
  }
  stats[0] = 0;
  stats[1] = -1;
  for (size_t p = 0; p < portions.size(); ++p) { stats[0] += ps[p]; stats[1] = jmax(stats[1], pm[p]); }
}

// ---------------------------------------------------------------------------
// Full driver: MultiViewDeconvolution{,Seq,Mul} (D/MultiViewDeconvolution.java:90-212,
// D/MultiViewDeconvolutionSeq.java:58-180, D/MultiViewDeconvolutionMul.java:65-245)
// ---------------------------------------------------------------------------
struct Params {
  int N[3];
  int V;
  int kdim[3];           // all PSFs share one size here (the configs do)
  int psfType;
  int scheme;            // 0 = Seq (OSEM-style), 1 = Mul
  int blockSize[3];
  float lambda;
  float minValue;
  int iterations;
  int nT;
  int psiInit;           // 0 = AvgPrecise, 1 = from array (PsiInitFromRAI, precise)
  int filterBlocks;      // DeconView filterBlocksForContent
  int sampleViews;       // > 0: stop after this many view passes (timing samples of the Seq scheme)
};

template <typename T>
static int deconvolve(const Params& p, const float* const* imgs, const float* const* weights,
                      const float* const* kernels, const float* psi0, T* psiOut, float* maxOut,
                      double* statsOut, T* k2Out, double* avgOut, double* loopSeconds) {
  const int V = p.V;
  const i64 N[3] = {p.N[0], p.N[1], p.N[2]};
  const i64 nvox = N[0] * N[1] * N[2];
  const i64 nk = (i64)p.kdim[0] * p.kdim[1] * p.kdim[2];
  // DeconView: blocks with kernelSize = 2K-1 (D/DeconView.java:150-182)
  std::vector<std::vector<Block>> blocks(V);
  std::vector<std::vector<std::vector<int>>> batches(V);
  for (int v = 0; v < V; ++v) {
    i64 ks[3], bs[3];
    for (int d = 0; d < 3; ++d) { ks[d] = 2 * (i64)p.kdim[d] - 1; bs[d] = p.blockSize[d]; }
    if (!divide_into_blocks(N, ks, bs, blocks[v])) return -1;
    batches[v] = sort_blocks(blocks[v], N, 1, nullptr);
    if (p.filterBlocks) {
      for (int j = (int)batches[v].size() - 1; j >= 0; --j) {
        auto& bl = batches[v][j];
        for (int i = (int)bl.size() - 1; i >= 0; --i)
          if (!block_contains_content(blocks[v][bl[i]], weights[v], N)) bl.erase(bl.begin() + i);
        if (bl.empty()) batches[v].erase(batches[v].begin() + j);
      }
    }
  }
  // DeconViews -> DeconViewPSF.init
  std::vector<std::vector<T>> k1(V), k2;
  std::vector<std::array<int, 3>> kd(V);
  for (int v = 0; v < V; ++v) {
    k1[v].assign(kernels[v], kernels[v] + nk);
    kd[v] = {p.kdim[0], p.kdim[1], p.kdim[2]};
  }
  init_psfs(k1, kd, p.psfType, p.nT, k2);
  if (k2Out)
    for (int v = 0; v < V; ++v) std::copy(k2[v].begin(), k2[v].end(), k2Out + v * nk);
  // kernel spectra at block size (DeconViewPSF.getKernel{1,2}FFT, :79-111), for the views that run
  const int bd[3] = {p.blockSize[0], p.blockSize[1], p.blockSize[2]};
  std::vector<KernelFFT<T>> k1f(V), k2f(V);
  const int usedViews = (p.sampleViews > 0 && p.scheme == 0) ? std::min(V, p.sampleViews) : V;
  for (int v = 0; v < usedViews; ++v) {
    kernel_fft(k1[v].data(), kd[v].data(), bd, k1f[v]);
    kernel_fft(k2[v].data(), kd[v].data(), bd, k2f[v]);
  }
  // psi + PsiInit
  std::vector<T> psi(nvox);
  double avg = psi_init_avg_precise(imgs, V, nvox, p.nT, maxOut);
  if (avgOut) *avgOut = avg;
  if (p.psiInit == 0) {
    for (i64 i = 0; i < nvox; ++i) psi[i] = (T)avg;  // (float)avg in the FP32 restatement
  } else {
    for (i64 i = 0; i < nvox; ++i) psi[i] = (T)psi0[i];
  }
  const i64 nb = (i64)bd[0] * bd[1] * bd[2];
  std::vector<T> psiBlk(nb);
  int viewPasses = 0;
#ifdef _OPENMP
  const double t0 = omp_get_wtime();
#endif
  for (int it = 0; it < p.iterations; ++it) {
    if (p.scheme == 0) {
      for (int v = 0; v < V; ++v) {
        if (p.sampleViews > 0 && viewPasses >= p.sampleViews) break;
        ++viewPasses;
        const size_t totalBlocks = blocks[v].size();
        double sumChange = 0, maxChange = -1;
        std::vector<std::pair<int, std::vector<T>>> prevQ, curQ;
        std::vector<T> imgBlk(nb), wBlk(nb);
        for (auto& batch : batches[v]) {
          for (int bi : batch) {
            const Block& b = blocks[v][bi];
            copy_block_mirror(psi.data(), N, b, psiBlk.data());
            copy_block_zero(imgs[v], N, b, imgBlk.data());
            copy_block_zero(weights[v], N, b, wBlk.data());
            double st[2];
            block_iteration_seq(psiBlk.data(), bd, imgBlk.data(), wBlk.data(), k1f[v], k2f[v], (double)p.lambda,
                                (T)p.minValue, (T)maxOut[v], p.nT, st);
            sumChange += st[0];
            maxChange = jmax(maxChange, st[1]);
            if (totalBlocks == 1) paste_block(psi.data(), N, b, psiBlk.data());
            else curQ.push_back({bi, psiBlk});
          }
          for (auto& q : prevQ) paste_block(psi.data(), N, blocks[v][q.first], q.second.data());
          prevQ.swap(curQ);
          curQ.clear();
        }
        for (auto& q : prevQ) paste_block(psi.data(), N, blocks[v][q.first], q.second.data());
        if (statsOut) { statsOut[2 * ((i64)it * V + v)] = sumChange; statsOut[2 * ((i64)it * V + v) + 1] = maxChange; }
      }
    } else {
      // Mul: blocks of view 0 drive the loop (testBlockIntegrity requires them identical)
      const size_t totalBlocks = blocks[0].size();
      double sumChange = 0, maxChange = -1;
      std::vector<std::pair<int, std::vector<T>>> prevQ, curQ;
      std::vector<std::vector<T>> imgBlk(V, std::vector<T>(nb)), wBlk(V, std::vector<T>(nb));
      for (auto& batch : batches[0]) {
        for (int bi : batch) {
          const Block& b = blocks[0][bi];
          copy_block_mirror(psi.data(), N, b, psiBlk.data());
          for (int v = 0; v < V; ++v) {
            copy_block_zero(imgs[v], N, b, imgBlk[v].data());
            copy_block_zero(weights[v], N, b, wBlk[v].data());
          }
          double st[2];
          block_iteration_mul(psiBlk.data(), bd, imgBlk, wBlk, k1f, k2f, (double)p.lambda, (T)p.minValue, maxOut,
                              p.nT, st);
          sumChange += st[0];
          maxChange = jmax(maxChange, st[1]);
          if (totalBlocks == 1) paste_block(psi.data(), N, b, psiBlk.data());
          else curQ.push_back({bi, psiBlk});
        }
        for (auto& q : prevQ) paste_block(psi.data(), N, blocks[0][q.first], q.second.data());
        prevQ.swap(curQ);
        curQ.clear();
      }
      for (auto& q : prevQ) paste_block(psi.data(), N, blocks[0][q.first], q.second.data());
      if (statsOut) { statsOut[2 * it] = sumChange; statsOut[2 * it + 1] = maxChange; }
    }
  }
#ifdef _OPENMP
  if (loopSeconds) *loopSeconds = omp_get_wtime() - t0;
#else
  if (loopSeconds) *loopSeconds = 0;
#endif
  std::copy(psi.begin(), psi.end(), psiOut);
  return 0;
}

}  // namespace mvo

// ---------------------------------------------------------------------------
// C entry points for ctypes (oracle.py)
// ---------------------------------------------------------------------------
using namespace mvo;

extern "C" {

int oracle_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
#else
  (void)n;
  return 1;
#endif
}

int oracle_fast_length(int n) { return fast_length(n); }

int oracle_divide_into_portions(long long n, int nT, long long* starts, long long* lens, int cap) {
  auto p = divide_into_portions(n, nT);
  for (size_t i = 0; i < p.size() && (int)i < cap; ++i) { starts[i] = p[i].start; lens[i] = p[i].len; }
  return (int)p.size();
}

double oracle_sum_img_quirk_f32(const float* a, long long n, int nT) { return sum_img_quirk(a, n, nT); }
void oracle_norm_to_sum1_f32(float* a, long long n, int nT) { norm_to_sum1(a, n, nT); }
void oracle_norm_to_sum1_f64(double* a, long long n, int nT) { norm_to_sum1(a, n, nT); }
void oracle_mirror_f32(float* a, const int* dim) { mirror_all(a, dim); }

// 1D complex FFT (for pinning the FFT against numpy)
void oracle_fft1d_f64(double* data, int n, int sign) {
  Plan1D<double> pl(n);
  std::vector<cpx<double>> a(n), s(n);
  for (int i = 0; i < n; ++i) a[i] = cpx<double>(data[2 * i], data[2 * i + 1]);
  pl.exec(a.data(), s.data(), sign);
  for (int i = 0; i < n; ++i) { data[2 * i] = a[i].real(); data[2 * i + 1] = a[i].imag(); }
}

// 3D real forward FFT -> half spectrum [pz][py][px/2+1] interleaved complex
void oracle_rfft3d_f64(const double* in, const int* p, double* out) {
  Plan3D<double> pl(p[0], p[1], p[2]);
  pl.r2c(in, reinterpret_cast<cpx<double>*>(out));
}

#define CONV_ENTRY(SUF, T)                                                                                  \
  int oracle_fft_convolve_##SUF(const T* img, const int* nd, int ext, T extValue, const T* k, const int* kd,  \
                                T* out) {                                                                    \
    KernelFFT<T> kf;                                                                                         \
    kernel_fft(k, kd, nd, kf);                                                                               \
    fft_convolve(img, nd, ext, extValue, kf, out);                                                           \
    return 0;                                                                                                \
  }
CONV_ENTRY(f32, float)
CONV_ENTRY(f64, double)

float oracle_next_value_f32(float last, float integral, float weight, double lambda, float minI, float maxI) {
  return next_value<float>(last, integral, weight, lambda, minI, maxI);
}
double oracle_next_value_f64(double last, double integral, double weight, double lambda, double minI, double maxI) {
  return next_value<double>(last, integral, weight, lambda, minI, maxI);
}
float oracle_next_value_mul_f32(float last, const float* integral, const float* weight, int nv, double lambda,
                                float minI, float maxI) {
  return next_value_mul<float>(last, integral, weight, nv, lambda, minI, maxI);
}

// DeconvolutionMethods.computeQuotient / computeFinalValues / computeFinalValuesMul over flat arrays
// (D/iteration/sequential/DeconvolutionMethods.java:46-98, 103-182, 187-305): the same per-value functions as
// the block iterations, statistics summed per portion of divideIntoPortions(n, nT) as the block threads do.
void oracle_compute_quotient_f32(const float* img, float* blurred, long long n) {
  for (i64 i = 0; i < n; ++i) blurred[i] = img[i] > 0 ? img[i] / blurred[i] : 1.0f;
}

void oracle_compute_final_values_f32(float* psi, const float* integral, const float* weight, long long n,
                                     double lambda, float minI, float maxI, int nT, double* stats) {
  stats[0] = 0;
  stats[1] = -1;
  for (const Portion& p : divide_into_portions(n, nT)) {
    double sumChange = 0, maxChange = -1;
    for (i64 j = 0; j < p.len; ++j) {
      const i64 i = p.start + j;
      const float last = psi[i];
      const float next = next_value<float>(last, integral[i], weight[i], lambda, minI, maxI);
      psi[i] = next;
      const float change = next - last;
      sumChange += (double)change;
      maxChange = jmax(maxChange, (double)change);
    }
    stats[0] += sumChange;
    stats[1] = jmax(stats[1], maxChange);
  }
}

// integral / weight: V arrays of n values back to back
void oracle_compute_final_values_mul_f32(float* psi, int V, const float* integral, const float* weight, long long n,
                                         double lambda, float minI, float maxI, int nT, double* stats) {
  stats[0] = 0;
  stats[1] = -1;
  std::vector<float> iv(V), wv(V);
  for (const Portion& p : divide_into_portions(n, nT)) {
    double sumChange = 0, maxChange = -1;
    for (i64 j = 0; j < p.len; ++j) {
      const i64 i = p.start + j;
      for (int v = 0; v < V; ++v) { iv[v] = integral[(i64)v * n + i]; wv[v] = weight[(i64)v * n + i]; }
      const float last = psi[i];
      const float next = next_value_mul<float>(last, iv.data(), wv.data(), V, lambda, minI, maxI);
      psi[i] = next;
      const float change = next - last;
      sumChange += (double)change;
      maxChange = jmax(maxChange, (double)change);
    }
    stats[0] += sumChange;
    stats[1] = jmax(stats[1], maxChange);
  }
}

// blocks: out array of 15 longs per block
int oracle_divide_into_blocks(const long long* img, const long long* kernelSize, const long long* blockSize,
                              long long* out, int cap) {
  i64 im[3] = {img[0], img[1], img[2]}, ks[3] = {kernelSize[0], kernelSize[1], kernelSize[2]},
      bs[3] = {blockSize[0], blockSize[1], blockSize[2]};
  std::vector<Block> blocks;
  if (!divide_into_blocks(im, ks, bs, blocks)) return -1;
  for (size_t i = 0; i < blocks.size() && (int)i < cap; ++i)
    for (int d = 0; d < 3; ++d) {
      out[15 * i + d] = blocks[i].blockSize[d];
      out[15 * i + 3 + d] = blocks[i].offset[d];
      out[15 * i + 6 + d] = blocks[i].effSize[d];
      out[15 * i + 9 + d] = blocks[i].effOffset[d];
      out[15 * i + 12 + d] = blocks[i].effLocal[d];
    }
  return (int)blocks.size();
}

// batches as a flat list: batchOf[i] = batch index of block i (in block order), returns #batches
int oracle_sort_blocks(const long long* img, const long long* kernelSize, const long long* blockSize,
                       int minRequired, int* batchOf, int* order, int* chosenDim) {
  i64 im[3] = {img[0], img[1], img[2]}, ks[3] = {kernelSize[0], kernelSize[1], kernelSize[2]},
      bs[3] = {blockSize[0], blockSize[1], blockSize[2]};
  std::vector<Block> blocks;
  if (!divide_into_blocks(im, ks, bs, blocks)) return -1;
  auto batches = sort_blocks(blocks, im, minRequired, chosenDim);
  int k = 0;
  for (size_t b = 0; b < batches.size(); ++b)
    for (int i : batches[b]) { batchOf[i] = (int)b; order[k++] = i; }
  return (int)batches.size();
}

int oracle_init_psfs_f32(const float* kernels, int V, const int* kd, int psfType, int nT, float* k1Out,
                         float* k2Out) {
  const i64 nk = (i64)kd[0] * kd[1] * kd[2];
  std::vector<std::vector<float>> k1(V), k2;
  std::vector<std::array<int, 3>> kdv(V);
  for (int v = 0; v < V; ++v) { k1[v].assign(kernels + v * nk, kernels + (v + 1) * nk); kdv[v] = {kd[0], kd[1], kd[2]}; }
  init_psfs(k1, kdv, psfType, nT, k2);
  for (int v = 0; v < V; ++v) {
    std::copy(k1[v].begin(), k1[v].end(), k1Out + v * nk);
    std::copy(k2[v].begin(), k2[v].end(), k2Out + v * nk);
  }
  return 0;
}

int oracle_init_psfs_f64(const double* kernels, int V, const int* kd, int psfType, int nT, double* k1Out,
                         double* k2Out) {
  const i64 nk = (i64)kd[0] * kd[1] * kd[2];
  std::vector<std::vector<double>> k1(V), k2;
  std::vector<std::array<int, 3>> kdv(V);
  for (int v = 0; v < V; ++v) { k1[v].assign(kernels + v * nk, kernels + (v + 1) * nk); kdv[v] = {kd[0], kd[1], kd[2]}; }
  init_psfs(k1, kdv, psfType, nT, k2);
  for (int v = 0; v < V; ++v) {
    std::copy(k1[v].begin(), k1[v].end(), k1Out + v * nk);
    std::copy(k2[v].begin(), k2[v].end(), k2Out + v * nk);
  }
  return 0;
}

double oracle_psi_init_avg_precise(const float* imgs, int V, long long n, int nT, float* maxOut) {
  std::vector<const float*> p(V);
  for (int v = 0; v < V; ++v) p[v] = imgs + (i64)v * n;
  return psi_init_avg_precise(p.data(), V, n, nT, maxOut);
}

// One block iteration of one view (the ComputeBlockSeqThread SPI), kernels already final.
int oracle_block_seq_f32(float* psiBlk, const int* bd, const float* img, const float* w, const float* k1,
                         const float* k2, const int* kd, float lambda, float minValue, float maxI, int nT,
                         double* stats) {
  KernelFFT<float> f1, f2;
  kernel_fft(k1, kd, bd, f1);
  kernel_fft(k2, kd, bd, f2);
  block_iteration_seq(psiBlk, bd, img, w, f1, f2, (double)lambda, minValue, maxI, nT, stats);
  return 0;
}

// One ComputeBlockMulThreadCPU.runIteration on a block over V views (kernels already final).
int oracle_block_mul_f32(float* psiBlk, const int* bd, int V, const float* imgs, const float* ws, const float* k1,
                         const float* k2, const int* kd, float lambda, float minValue, const float* maxI, int nT,
                         double* stats) {
  const i64 nb = (i64)bd[0] * bd[1] * bd[2], nk = (i64)kd[0] * kd[1] * kd[2];
  std::vector<std::vector<float>> ib(V), wb(V);
  std::vector<KernelFFT<float>> f1(V), f2(V);
  for (int v = 0; v < V; ++v) {
    ib[v].assign(imgs + v * nb, imgs + (v + 1) * nb);
    wb[v].assign(ws + v * nb, ws + (v + 1) * nb);
    kernel_fft(k1 + v * nk, kd, bd, f1[v]);
    kernel_fft(k2 + v * nk, kd, bd, f2[v]);
  }
  block_iteration_mul(psiBlk, bd, ib, wb, f1, f2, (double)lambda, minValue, maxI, nT, stats);
  return 0;
}

#define DECON_ENTRY(SUF, T)                                                                                        \
  int oracle_deconvolve_##SUF(const int* ip, const float* fp, const float* imgs, const float* weights,              \
                              const float* kernels, const float* psi0, T* psiOut, float* maxOut, double* statsOut,  \
                              T* k2Out, double* avgOut, double* loopSeconds) {                                      \
    Params p;                                                                                                      \
    for (int d = 0; d < 3; ++d) { p.N[d] = ip[d]; p.kdim[d] = ip[3 + d]; p.blockSize[d] = ip[6 + d]; }             \
    p.V = ip[9]; p.psfType = ip[10]; p.scheme = ip[11]; p.iterations = ip[12]; p.nT = ip[13];                      \
    p.psiInit = ip[14]; p.filterBlocks = ip[15]; p.sampleViews = ip[16];                                           \
    p.lambda = fp[0]; p.minValue = fp[1];                                                                          \
    const i64 nvox = (i64)p.N[0] * p.N[1] * p.N[2], nk = (i64)p.kdim[0] * p.kdim[1] * p.kdim[2];                  \
    std::vector<const float*> pi(p.V), pw(p.V), pk(p.V);                                                           \
    for (int v = 0; v < p.V; ++v) { pi[v] = imgs + v * nvox; pw[v] = weights + v * nvox; pk[v] = kernels + v * nk; } \
    return deconvolve<T>(p, pi.data(), pw.data(), pk.data(), psi0, psiOut, maxOut, statsOut, k2Out, avgOut,       \
                         loopSeconds);                                                                             \
  }
DECON_ENTRY(f32, float)
DECON_ENTRY(f64, double)

}  // extern "C"
