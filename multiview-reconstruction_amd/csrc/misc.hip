// misc.hip -- pointwise and reduction kernels of the deconvolution engine.
#include "passes.hpp"

namespace mvr {

__global__ void k_fill(float* __restrict__ dst, size_t n, float v) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) dst[i] = v;
}

void launch_fill(float* dst, size_t n, float v, hipStream_t s) {
  const int nb = (int)std::min<size_t>((n + kThreads - 1) / kThreads, 4096);
  hipLaunchKernelGGL(k_fill, dim3(nb > 0 ? nb : 1), dim3(kThreads), 0, s, dst, n, v);
}

// Deterministic final reduction of per-work-group {sum, max} partials
// (IterationStatistics accumulation, ComputeBlockSeqThreadCPU.java:159-166).
__global__ void k_reduce_stats(const double2* __restrict__ partial, int n, double2* __restrict__ out) {
  __shared__ double ss[kThreads], sm[kThreads];
  double s = 0.0, m = -1.0;
  for (int i = threadIdx.x; i < n; i += kThreads) { s += partial[i].x; m = java_max(m, partial[i].y); }
  ss[threadIdx.x] = s;
  sm[threadIdx.x] = m;
  __syncthreads();
  for (int off = kThreads / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) {
      ss[threadIdx.x] += ss[threadIdx.x + off];
      sm[threadIdx.x] = java_max(sm[threadIdx.x], sm[threadIdx.x + off]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = make_double2(ss[0], sm[0]);
}

void launch_reduce_stats(const double2* partial, int n, double2* out, hipStream_t s) {
  hipLaunchKernelGGL(k_reduce_stats, dim3(1), dim3(kThreads), 0, s, partial, n, out);
}

// The same reduction for up to 32 view updates at once (work-group v: partials [v * stride, v * stride + n[v]),
// result out[v]): one launch per iteration instead of one per view update.
struct PartialCounts { int n[32]; };
__global__ void k_reduce_stats_batched(const double2* __restrict__ partial, size_t stride, PartialCounts c,
                                       double2* __restrict__ out) {
  __shared__ double ss[kThreads], sm[kThreads];
  const double2* p = partial + (size_t)blockIdx.x * stride;
  const int n = c.n[blockIdx.x];
  double s = 0.0, m = -1.0;
  for (int i = threadIdx.x; i < n; i += kThreads) { s += p[i].x; m = java_max(m, p[i].y); }
  ss[threadIdx.x] = s;
  sm[threadIdx.x] = m;
  __syncthreads();
  for (int off = kThreads / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) {
      ss[threadIdx.x] += ss[threadIdx.x + off];
      sm[threadIdx.x] = java_max(sm[threadIdx.x], sm[threadIdx.x + off]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = make_double2(ss[0], sm[0]);
}

void launch_reduce_stats_batched(const double2* partial, size_t stride, const int* counts, int nUpdates, double2* out,
                                 hipStream_t s) {
  PartialCounts c{};
  for (int v = 0; v < nUpdates && v < 32; ++v) c.n[v] = counts[v];
  hipLaunchKernelGGL(k_reduce_stats_batched, dim3(nUpdates), dim3(kThreads), 0, s, partial, stride, c, out);
}

// PsiInitAvgPrecise per voxel (D/init/PsiInitAvgPreciseThread.java:127-155): the
// mean of the positive view values is summed (double) over voxels with any
// positive value; max of the positive values per view.  Per-work-group
// partials are reduced on the host in a fixed order.
constexpr int kMaxViews = 32;
struct ViewPtrs { const float* p[kMaxViews]; };

__global__ __launch_bounds__(kThreads) void k_psi_init(ViewPtrs imgs, int V, size_t n, double* __restrict__ psum,
                                                       unsigned long long* __restrict__ pcount,
                                                       float* __restrict__ pmax) {
  __shared__ double ssum[kThreads / 64];
  __shared__ unsigned long long scnt[kThreads / 64];
  __shared__ float smax[kThreads / 64][kMaxViews];
  double sum = 0.0;
  unsigned long long cnt = 0;
  float mx[kMaxViews];
  for (int v = 0; v < V; ++v) mx[v] = 0.f;
  for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (size_t)gridDim.x * kThreads) {
    double s = 0.0;
    int c = 0;
    for (int v = 0; v < V; ++v) {
      const float val = imgs.p[v][i];
      if (val > 0.f) { mx[v] = fmaxf(mx[v], val); s += (double)val; ++c; }
    }
    if (c > 0) { sum += s / (double)c; ++cnt; }
  }
  for (int off = 32; off > 0; off >>= 1) {
    sum += __shfl_down(sum, off);
    cnt += __shfl_down(cnt, off);
    for (int v = 0; v < V; ++v) mx[v] = fmaxf(mx[v], __shfl_down(mx[v], off));
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) {
    ssum[wv] = sum;
    scnt[wv] = cnt;
    for (int v = 0; v < V; ++v) smax[wv][v] = mx[v];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0.0;
    unsigned long long b = 0;
    for (int i = 0; i < kThreads / 64; ++i) { a += ssum[i]; b += scnt[i]; }
    psum[blockIdx.x] = a;
    pcount[blockIdx.x] = b;
    for (int v = 0; v < V; ++v) {
      float m = 0.f;
      for (int i = 0; i < kThreads / 64; ++i) m = fmaxf(m, smax[i][v]);
      pmax[(size_t)blockIdx.x * V + v] = m;
    }
  }
}

int launch_psi_init_partials(const float* const* imgs, int V, size_t n, double* psum, unsigned long long* pcount,
                             float* pmax, hipStream_t s) {
  ViewPtrs vp{};
  for (int v = 0; v < V && v < kMaxViews; ++v) vp.p[v] = imgs[v];
  const int nb = (int)std::min<size_t>((n + kThreads - 1) / kThreads, 1024);
  hipLaunchKernelGGL(k_psi_init, dim3(nb), dim3(kThreads), 0, s, vp, V, n, psum, pcount, pmax);
  return nb;
}

// PsiInitAvgApproxThread.call (D/init/PsiInitAvgApproxThread.java:57-85): min/max/sum over the hyperslice
// x = nx / 2 of one view (all values, zeros included).  The reference visits that slice once per
// dimension and divides by the tripled count, which is the slice mean.  Partials per work-group:
// {sum, max}; the host reduces them in a fixed order.
__global__ __launch_bounds__(kThreads) void k_psi_approx_slice(const float* __restrict__ img, int3 N,
                                                               double2* __restrict__ partial) {
  __shared__ double ss[kThreads], sm[kThreads];
  const int x = N.x / 2;
  const size_t n = (size_t)N.y * N.z;
  double sum = 0.0, mx = -1.0e308;
  for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (size_t)gridDim.x * kThreads) {
    const double v = (double)img[i * N.x + x];  // i = z * ny + y
    sum += v;
    mx = fmax(mx, v);
  }
  ss[threadIdx.x] = sum;
  sm[threadIdx.x] = mx;
  __syncthreads();
  for (int off = kThreads / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) {
      ss[threadIdx.x] += ss[threadIdx.x + off];
      sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + off]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = make_double2(ss[0], sm[0]);
}

int launch_psi_approx_slice(const float* img, const int N[3], double2* partial, hipStream_t s) {
  const size_t n = (size_t)N[1] * N[2];
  const int nb = (int)std::max<size_t>(1, std::min<size_t>((n + kThreads - 1) / kThreads, 256));
  hipLaunchKernelGGL(k_psi_approx_slice, dim3(nb), dim3(kThreads), 0, s, img, make_int3(N[0], N[1], N[2]), partial);
  return nb;
}

// FusedNonZeroRandomAccess.get (D/util/FusedNonZeroRandomAccess.java:57-96): the weighted mean of the
// positive view values, sum(I w) / sum(w) in double, 0 where no positive view has weight.
struct ViewPtrs2 { const float* img[kMaxViews]; const float* w[kMaxViews]; };
__global__ void k_fused_nonzero(ViewPtrs2 vp, int V, size_t n, float* __restrict__ out) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    double sI = 0.0, sW = 0.0;
    for (int v = 0; v < V; ++v) {
      const double val = (double)vp.img[v][i];
      if (val > 0.0) {
        const double wt = (double)vp.w[v][i];
        sI += val * wt;
        sW += wt;
      }
    }
    out[i] = sW > 0.0 ? (float)(sI / sW) : 0.f;
  }
}

void launch_fused_nonzero(const float* const* imgs, const float* const* ws, int V, size_t n, float* out,
                          hipStream_t s) {
  ViewPtrs2 vp{};
  for (int v = 0; v < V && v < kMaxViews; ++v) { vp.img[v] = imgs[v]; vp.w[v] = ws[v]; }
  const int nb = (int)std::max<size_t>(1, std::min<size_t>((n + kThreads - 1) / kThreads, 8192));
  hipLaunchKernelGGL(k_fused_nonzero, dim3(nb), dim3(kThreads), 0, s, vp, V, n, out);
}

// One axis of a separable symmetric convolution with mirror-single extension (the Gaussian blur of
// PsiInitBlurredFused, D/init/PsiInitBlurredFused.java:113-114): out[p] = sum_{|j| < H} k[|j|] in[mirror(p + j)]
// along `axis`, accumulated in double; mirror-single has period 2n - 2 along the axis.
__global__ void k_sym_conv_axis(const float* __restrict__ in, float* __restrict__ out, int3 N, int axis,
                                const double* __restrict__ half, int H) {
  const size_t n = (size_t)N.x * N.y * N.z;
  const int len = axis == 0 ? N.x : axis == 1 ? N.y : N.z;
  const size_t stride = axis == 0 ? 1 : axis == 1 ? (size_t)N.x : (size_t)N.x * N.y;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const int pos = (int)((i / stride) % (size_t)len);
    const float* base = in + (i - (size_t)pos * stride);
    double acc = half[0] * (double)base[(size_t)pos * stride];
    for (int j = 1; j < H; ++j) {
      const int a = mirror_index(pos - j, len), b = mirror_index(pos + j, len);
      acc += half[j] * ((double)base[(size_t)a * stride] + (double)base[(size_t)b * stride]);
    }
    out[i] = (float)acc;
  }
}

void launch_sym_conv_axis(const float* in, float* out, const int N[3], int axis, const double* half, int H,
                          hipStream_t s) {
  const size_t n = (size_t)N[0] * N[1] * N[2];
  const int nb = (int)std::max<size_t>(1, std::min<size_t>((n + kThreads - 1) / kThreads, 8192));
  hipLaunchKernelGGL(k_sym_conv_axis, dim3(nb), dim3(kThreads), 0, s, in, out, make_int3(N[0], N[1], N[2]), axis, half,
                     H);
}

// Mul-scheme update (DeconvolutionMethods.computeNextValueMul,
// D/iteration/sequential/DeconvolutionMethods.java:370-419): geometric mean of the
// per-view integrals (double product, pow 1/V), min(1, sum of weights).
struct ViewPtrsW { const float* t2[kMaxViews]; const float* w[kMaxViews]; };

__global__ __launch_bounds__(kThreads) void k_update_mul(float* __restrict__ psi, ViewPtrsW vp, int V, int3 N,
                                                         int3 o, int3 D, double lambda, float minI, float maxI,
                                                         double2* __restrict__ partial) {
  // Java rounds every float and double operation: no FMA contraction in the update
#pragma clang fp contract(off)
  const size_t nvox = (size_t)D.x * D.y * D.z;
  double sumc = 0.0;
  float mxc = -1.f;
  const double invV = 1.0 / (double)V;
  const RLConsts rc = rl_consts(lambda, minI, maxI);
  for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < nvox; i += (size_t)gridDim.x * kThreads) {
    const int lx = (int)(i % D.x);
    const size_t r = i / D.x;
    const int ly = (int)(r % D.y), lz = (int)(r / D.y);
    const size_t gi = ((size_t)(o.z + lz) * N.y + (o.y + ly)) * N.x + (o.x + lx);
    double sumW = 0.0, prod = 1.0;
    for (int v = 0; v < V; ++v) {
      prod = prod * (double)vp.t2[v][i];
      sumW = sumW + (double)vp.w[v][gi];
    }
    prod = pow(prod, invV);
    sumW = fmin(1.0, sumW);
    const float last = psi[gi];
    const float next = rl_tikhonov_adjust(last * (float)prod, rc);
    const float res = last + (next - last) * (float)sumW;
    psi[gi] = res;
    const float ch = res - last;
    sumc += (double)ch;
    mxc = java_maxf(mxc, ch);
  }
  block_reduce_stats(sumc, mxc, partial + blockIdx.x);
}

void launch_update_mul(float* psi, const float* const* T2, const float* const* w, int V, const int N[3],
                       const int o[3], const int D[3], double lambda, float minValue, float maxI, double2* partial,
                       int* nPartialOut, hipStream_t s) {
  ViewPtrsW vp{};
  for (int v = 0; v < V && v < kMaxViews; ++v) { vp.t2[v] = T2[v]; vp.w[v] = w[v]; }
  const size_t nvox = (size_t)D[0] * D[1] * D[2];
  const int nb = (int)std::min<size_t>((nvox + kThreads - 1) / kThreads, 2048);
  hipLaunchKernelGGL(k_update_mul, dim3(nb), dim3(kThreads), 0, s, psi, vp, V, make_int3(N[0], N[1], N[2]),
                     make_int3(o[0], o[1], o[2]), make_int3(D[0], D[1], D[2]), lambda, minValue, maxI, partial);
  *nPartialOut = nb;
}

// Block.copyBlock from a zero-extended volume (C/Block.java:158-197 with Views.extendZero, as the image and weight
// blocks of MultiViewDeconvolutionSeq.java:115-126): dst (B voxels) = src (N voxels) at volume coordinate
// off + local position, 0 outside the volume.  One thread per destination voxel, rows of B.x consecutive
// threads so that the loads of a row are contiguous.
__global__ void k_copy_block_zero(const float* __restrict__ src, int3 N, int3 off, float* __restrict__ dst, int3 B) {
  const size_t n = (size_t)B.x * B.y * B.z;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const int x = (int)(i % (size_t)B.x);
    const size_t r = i / (size_t)B.x;
    const int y = (int)(r % (size_t)B.y), z = (int)(r / (size_t)B.y);
    const long long gx = (long long)off.x + x, gy = (long long)off.y + y, gz = (long long)off.z + z;
    float v = 0.f;
    if (gx >= 0 && gy >= 0 && gz >= 0 && gx < N.x && gy < N.y && gz < N.z)
      v = src[((size_t)gz * N.y + (size_t)gy) * N.x + (size_t)gx];
    dst[i] = v;
  }
}

void launch_copy_block_zero(const float* src, const int N[3], const long long off[3], float* dst, const int B[3],
                            hipStream_t s) {
  const size_t n = (size_t)B[0] * B[1] * B[2];
  const int nb = (int)std::max<size_t>(1, std::min<size_t>((n + kThreads - 1) / kThreads, 16384));
  // offsets beyond +-2^31 would place the block wholly outside any volume the library holds: clamp to a value
  // that is still outside (the copy is then all zeros)
  auto c = [](long long v) { return (int)std::max<long long>(-(1ll << 30), std::min<long long>(v, 1ll << 30)); };
  hipLaunchKernelGGL(k_copy_block_zero, dim3(nb), dim3(kThreads), 0, s, src, make_int3(N[0], N[1], N[2]),
                     make_int3(c(off[0]), c(off[1]), c(off[2])), dst, make_int3(B[0], B[1], B[2]));
}

// DeconView.blockContainsContent (D/DeconView.java:232-274) for many blocks at once: flags[b] = 1 when the weights
// hold a non-zero value inside block b's interval (entry b of tab: {offset[3], size[3]}, zero outside the
// volume); flags must be zero before the launch.  Work-group (x, b) scans a strided share of block b.
__global__ void k_blocks_with_content(const float* __restrict__ w, int3 N, const int* __restrict__ tab,
                                      int* __restrict__ flags) {
  const int* t = tab + 6 * blockIdx.y;
  const int x0 = max(t[0], 0), y0 = max(t[1], 0), z0 = max(t[2], 0);
  const int x1 = min(t[0] + t[3], N.x), y1 = min(t[1] + t[4], N.y), z1 = min(t[2] + t[5], N.z);
  if (x1 <= x0 || y1 <= y0 || z1 <= z0) return;
  const int nx = x1 - x0, ny = y1 - y0;
  const size_t n = (size_t)nx * ny * (z1 - z0);
  bool any = false;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n && !any; i += (size_t)gridDim.x * blockDim.x) {
    const int x = (int)(i % nx);
    const size_t r = i / nx;
    const int y = (int)(r % ny), z = (int)(r / ny);
    any = w[((size_t)(z0 + z) * N.y + (y0 + y)) * N.x + (x0 + x)] != 0.f;
  }
  if (any) flags[blockIdx.y] = 1;  // every writer stores the same value
}

void launch_blocks_with_content(const float* w, const int N[3], const int* tab, int nblocks, int* flags,
                                hipStream_t s) {
  if (nblocks <= 0) return;
  hipLaunchKernelGGL(k_blocks_with_content, dim3(64, nblocks), dim3(kThreads), 0, s, w, make_int3(N[0], N[1], N[2]),
                     tab, flags);
}

// DeconvolutionMethods.computeQuotient (D/iteration/sequential/DeconvolutionMethods.java:46-98) over a flat array:
// blurred[i] = img[i] > 0 ? img[i] / blurred[i] : 1.
__global__ void k_quotient(const float* __restrict__ img, float* __restrict__ blurred, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const float im = img[i];
    blurred[i] = im > 0.f ? im / blurred[i] : 1.f;
  }
}

void launch_quotient(const float* img, float* blurred, size_t n, hipStream_t s) {
  const int nb = (int)std::max<size_t>(1, std::min<size_t>((n + kThreads - 1) / kThreads, 8192));
  hipLaunchKernelGGL(k_quotient, dim3(nb), dim3(kThreads), 0, s, img, blurred, n);
}

// DeconvolutionMethods.computeFinalValues (D/iteration/sequential/DeconvolutionMethods.java:103-182) over a flat
// array: psi[i] = computeNextValue(psi[i], integral[i], weight[i]); per work-group {sum, max} of the change.
__global__ __launch_bounds__(kThreads) void k_final_values(float* __restrict__ psi, const float* __restrict__ integral,
                                                           const float* __restrict__ weight, size_t n, double lambda,
                                                           float minI, float maxI, double2* __restrict__ partial) {
  double sumc = 0.0;
  float mxc = -1.f;
  const RLConsts rc = rl_consts(lambda, minI, maxI);
  for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (size_t)gridDim.x * kThreads) {
    const float last = psi[i];
    const float next = rl_next_value(last, integral[i], weight[i], rc);
    psi[i] = next;
    const float ch = next - last;
    sumc += (double)ch;
    mxc = java_maxf(mxc, ch);
  }
  block_reduce_stats(sumc, mxc, partial + blockIdx.x);
}

int launch_final_values(float* psi, const float* integral, const float* weight, size_t n, double lambda,
                        float minValue, float maxI, double2* partial, int maxPartials, hipStream_t s) {
  const int nb = (int)std::max<size_t>(1, std::min<size_t>((n + kThreads - 1) / kThreads, (size_t)maxPartials));
  hipLaunchKernelGGL(k_final_values, dim3(nb), dim3(kThreads), 0, s, psi, integral, weight, n, lambda, minValue, maxI,
                     partial);
  return nb;
}

// Distance of psi to a reference volume in FP64: per work-group {sum (a - b)^2, sum b^2},
// then one work-group sums the partials in a fixed order.
__global__ __launch_bounds__(kThreads) void k_rel_l2_partials(const float* __restrict__ a, const float* __restrict__ b,
                                                              size_t n, double2* __restrict__ partial) {
  __shared__ double sd[kThreads], sr[kThreads];
  double d = 0.0, r = 0.0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const double e = (double)a[i] - (double)b[i];
    d += e * e;
    r += (double)b[i] * (double)b[i];
  }
  sd[threadIdx.x] = d;
  sr[threadIdx.x] = r;
  __syncthreads();
  for (int off = kThreads / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) {
      sd[threadIdx.x] += sd[threadIdx.x + off];
      sr[threadIdx.x] += sr[threadIdx.x + off];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = make_double2(sd[0], sr[0]);
}

__global__ void k_sum2(const double2* __restrict__ partial, int n, double2* __restrict__ out) {
  __shared__ double sd[kThreads], sr[kThreads];
  double d = 0.0, r = 0.0;
  for (int i = threadIdx.x; i < n; i += kThreads) { d += partial[i].x; r += partial[i].y; }
  sd[threadIdx.x] = d;
  sr[threadIdx.x] = r;
  __syncthreads();
  for (int off = kThreads / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) {
      sd[threadIdx.x] += sd[threadIdx.x + off];
      sr[threadIdx.x] += sr[threadIdx.x + off];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = make_double2(sd[0], sr[0]);
}

// partial must hold 1024 entries; out receives {sum (a - b)^2, sum b^2}
void launch_rel_l2(const float* a, const float* b, size_t n, double2* partial, double2* out, hipStream_t s) {
  const int nb = (int)std::max<size_t>(1, std::min<size_t>((n + kThreads - 1) / kThreads, 1024));
  hipLaunchKernelGGL(k_rel_l2_partials, dim3(nb), dim3(kThreads), 0, s, a, b, n, partial);
  hipLaunchKernelGGL(k_sum2, dim3(1), dim3(kThreads), 0, s, partial, nb, out);
}

__global__ void k_scale_copy(const float2* __restrict__ in, float2* __restrict__ out, size_t n, float scale) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    out[i] = cscale(in[i], scale);
}

void launch_scale_copy_kernel_spectrum(const float2* in, float2* out, size_t n, float scale, hipStream_t s) {
  const int nb = (int)std::min<size_t>((n + kThreads - 1) / kThreads, 4096);
  hipLaunchKernelGGL(k_scale_copy, dim3(nb > 0 ? nb : 1), dim3(kThreads), 0, s, in, out, n, scale);
}

}  // namespace mvr
