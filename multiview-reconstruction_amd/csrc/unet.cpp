// unet.cpp -- host side of the learned initial estimate (U-Net training and inference).
#include "unet_host.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>

namespace mvr {

#define UCHECK(cond, msg) \
  do {                    \
    if (!(cond)) {        \
      set_error(msg);     \
      return -1;          \
    }                     \
  } while (0)

// layer table: input channels (padded), output channels, taps
static const int kCin[kLayers] = {16, 16, 16, 32, 32, 64, 96, 32, 48, 16};
static const int kCout[kLayers] = {16, 16, 32, 32, 64, 64, 32, 32, 16, 16};

size_t UNet::wcount(int l) const { return (size_t)27 * kCout[l] * kCin[l]; }

int UNet::create(int dev_, int V_, unsigned long long seed) {
  dev = dev_;
  UCHECK(V_ >= 1 && V_ <= kNetMaxViews, "the network takes 1..8 views");
  V = V_;
  MVR_HIP(hipSetDevice(dev));
  MVR_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  std::mt19937_64 rng(seed);
  for (int l = 0; l < kLayers; ++l) {
    Param& p = layer[l];
    p.count = wcount(l);
    std::vector<float> h(p.count, 0.f);
    const int fanin = 27 * (l == 0 ? (V + 1) : kCin[l]);
    std::normal_distribution<float> nd(0.f, std::sqrt(2.f / (float)fanin));
    for (int t = 0; t < 27; ++t)
      for (int co = 0; co < kCout[l]; ++co)
        for (int ci = 0; ci < kCin[l]; ++ci) {
          const bool real = l != 0 || ci < V + 1;
          h[((size_t)t * kCout[l] + co) * kCin[l] + ci] = real ? nd(rng) : 0.f;
        }
    if (p.alloc(true)) return -1;
    MVR_HIP(hipMemcpy(p.w, h.data(), p.count * 4, hipMemcpyHostToDevice));
    Param& b = bias[l];
    b.count = kCout[l];
    if (b.alloc(false)) return -1;
  }
  head.count = 17;  // 16 weights + bias
  if (head.alloc(false)) return -1;
  {
    std::vector<float> h(17, 0.f);
    std::normal_distribution<float> nd(0.f, 0.01f);
    for (int c = 0; c < 16; ++c) h[c] = nd(rng);
    MVR_HIP(hipMemcpy(head.w, h.data(), 17 * 4, hipMemcpyHostToDevice));
  }
  return refresh_copies();
}

int UNet::Param::alloc(bool conv) {
  MVR_HIP(hipMalloc(&w, count * 4));
  MVR_HIP(hipMalloc(&m, count * 4));
  MVR_HIP(hipMalloc(&v, count * 4));
  MVR_HIP(hipMalloc(&g, count * 4));
  MVR_HIP(hipMemset(w, 0, count * 4));
  MVR_HIP(hipMemset(m, 0, count * 4));
  MVR_HIP(hipMemset(v, 0, count * 4));
  MVR_HIP(hipMemset(g, 0, count * 4));
  if (conv) {
    MVR_HIP(hipMalloc(&wbf, count * 2));
    MVR_HIP(hipMalloc(&wt, count * 2));
  }
  return 0;
}

void UNet::Param::release() {
  for (void* p : {(void*)w, (void*)m, (void*)v, (void*)g, (void*)wbf, (void*)wt})
    if (p) (void)hipFree(p);
  w = m = v = g = nullptr;
  wbf = wt = nullptr;
}

// bf16 compute copies of the FP32 masters (an Adam step with learning rate 0 does exactly that)
int UNet::refresh_copies() {
  for (int l = 0; l < kLayers; ++l) {
    AdamArgs a{};
    a.w = layer[l].w; a.m = layer[l].m; a.v = layer[l].v; a.grad = layer[l].g;
    a.wbf = layer[l].wbf; a.wt = layer[l].wt; a.count = layer[l].count;
    a.taps = 27; a.cout = kCout[l]; a.cin = kCin[l];
    a.lr = 0.f; a.beta1 = 1.f; a.beta2 = 1.f; a.eps = 1.f; a.c1 = 0.f; a.c2 = 0.f;
    launch_adam(a, stream);
  }
  MVR_HIP(hipGetLastError());
  MVR_HIP(hipStreamSynchronize(stream));
  return 0;
}

size_t UNet::param_count() const {
  size_t n = head.count;
  for (int l = 0; l < kLayers; ++l) n += layer[l].count + bias[l].count;
  return n;
}

int UNet::copy_params(float* hostOut, const float* hostIn, bool grads) {
  MVR_HIP(hipSetDevice(dev));
  size_t off = 0;
  auto one = [&](Param& p) -> int {
    float* dev = grads ? p.g : p.w;
    if (hostOut) MVR_HIP(hipMemcpy(hostOut + off, dev, p.count * 4, hipMemcpyDeviceToHost));
    if (hostIn) MVR_HIP(hipMemcpy(dev, hostIn + off, p.count * 4, hipMemcpyHostToDevice));
    off += p.count;
    return 0;
  };
  for (int l = 0; l < kLayers; ++l) {
    if (one(layer[l]) || one(bias[l])) return -1;
  }
  if (one(head)) return -1;
  if (hostIn && !grads) return refresh_copies();
  return 0;
}

// ------------------------------------------------------------------ activations
int UNet::Acts::alloc(int N_, int D_, int H_, int W_, bool grads) {
  N = N_; D = D_; H = H_; W = W_;
  Dcap = D_;
  const size_t v0 = (size_t)N * D * H * W, v1 = v0 / 8, v2 = v0 / 64;
  struct B { uint16_t** p; uint16_t** gp; size_t n; };
  const B bufs[] = {{&x0, nullptr, v0 * 16},   {&a1, &g_a1, v0 * 16},  {&cat1, &g_cat1, v0 * 48},
                    {&p1, &g_p1, v1 * 16},     {&a2, &g_a2, v1 * 32},  {&cat2, &g_cat2, v1 * 96},
                    {&p2, &g_p2, v2 * 32},     {&a3, &g_a3, v2 * 64},  {&b3, &g_b3, v2 * 64},
                    {&a4, &g_a4, v1 * 32},     {&d2, &g_d2, v1 * 32},  {&a5, &g_a5, v0 * 16},
                    {&d1, &g_d1, v0 * 16}};
  for (const B& b : bufs) {
    MVR_HIP(hipMalloc(b.p, b.n * 2));
    if (grads && b.gp) MVR_HIP(hipMalloc(b.gp, b.n * 2));
  }
  MVR_HIP(hipMalloc(&avg, v0 * 4));
  if (grads) {
    MVR_HIP(hipMalloc(&target, v0 * 4));
    MVR_HIP(hipMalloc(&dr, v0 * 4));
  }
  return 0;
}

void UNet::Acts::release() {
  uint16_t** all[] = {&x0, &a1, &cat1, &p1, &a2, &cat2, &p2, &a3, &b3, &a4, &d2, &a5, &d1,
                      &g_a1, &g_cat1, &g_p1, &g_a2, &g_cat2, &g_p2, &g_a3, &g_b3, &g_a4, &g_d2, &g_a5, &g_d1};
  for (auto p : all)
    if (*p) { (void)hipFree(*p); *p = nullptr; }
  for (float** p : {&avg, &target, &dr})
    if (*p) { (void)hipFree(*p); *p = nullptr; }
  N = D = H = W = Dcap = 0;
}

int UNet::ensure_acts(Acts& a, int N_, int D_, int H_, int W_, bool grads) {
  UCHECK(D_ % 4 == 0 && H_ % 4 == 0 && W_ % 4 == 0 && D_ >= 4 && H_ >= 4 && W_ >= 4,
         "network input dims must be multiples of 4");
  if (a.N == N_ && a.Dcap >= D_ && a.H == H_ && a.W == W_ && (!grads || a.g_a1)) {
    a.D = D_;  // the buffers hold Dcap planes per sample; the tensors use the first D
    return 0;
  }
  a.release();
  return a.alloc(N_, D_, H_, W_, grads);
}

bool UNet::conv(int l, const uint16_t* in, int in_cs, int in_co, uint16_t* out, int out_cs, int out_co, int N, int D,
                int H, int W, bool relu, const uint16_t* up, int up_cs, int up_c, uint16_t* pool, int pool_cs) {
  ConvArgs c{};
  c.up_in = up;
  c.up_cs = up_cs;
  c.up_c = up ? up_c : 0;
  c.in = in + in_co;
  c.in_cs = in_cs;
  c.cin = kCin[l];
  c.w = layer[l].wbf;
  c.bias = bias[l].w;
  c.out = out;
  c.out_cs = out_cs;
  c.out_co = out_co;
  c.cout = kCout[l];
  c.N = N; c.D = D; c.H = H; c.W = W;
  c.taps = 27;
  c.relu = relu ? 1 : 0;
  c.pool_out = pool;
  c.pool_cs = pool_cs;
  const bool fused = pool && conv_fuses_pool(c);
  if (!fused) c.pool_out = nullptr;
  if (!launch_conv_fwd(c, stream)) convFailed = true;
  return fused;
}

// data gradient of layer l: din = conv(dout, flipped transposed weights), no bias
void UNet::dgrad(int l, const uint16_t* dout, int dout_cs, int dout_co, uint16_t* din, int din_cs, int din_co, int N,
                 int D, int H, int W) {
  ConvArgs c{};
  c.in = dout + dout_co;
  c.in_cs = dout_cs;
  c.cin = kCout[l];
  c.w = layer[l].wt;
  c.bias = nullptr;
  c.out = din;
  c.out_cs = din_cs;
  c.out_co = din_co;
  c.cout = kCin[l];
  c.N = N; c.D = D; c.H = H; c.W = W;
  c.taps = 27;
  c.relu = 0;
  if (!launch_conv_fwd(c, stream)) convFailed = true;
}

int UNet::wgrad(int l, const uint16_t* in, int in_cs, const uint16_t* dout, int dout_cs, int dout_co, int N, int D,
                int H, int W) {
  const long long nseg = (long long)N * D * H * ((W + 15) / 16);
  const int nparts = (int)std::max<long long>(1, std::min<long long>(kWgradParts, nseg));
  WgradArgs a{};
  a.in = in; a.in_cs = in_cs; a.cin = kCin[l];
  a.dout = dout; a.dout_cs = dout_cs; a.dout_co = dout_co; a.cout = kCout[l];
  a.N = N; a.D = D; a.H = H; a.W = W;
  a.taps = 27;
  a.partial = wpart;
  launch_conv_wgrad(a, nparts, stream);
  launch_sum_partials(wpart, nparts, layer[l].count, layer[l].g, stream);
  const size_t nvox = (size_t)N * D * H * W;
  const int bparts = (int)std::min<size_t>(256, std::max<size_t>(1, nvox / 256));
  launch_bias_grad(dout, dout_cs, dout_co, kCout[l], nvox, bpart, bparts, stream);
  launch_sum_partials(bpart, bparts, kCout[l], bias[l].g, stream);
  return 0;
}

// ------------------------------------------------------------------ forward / backward
int UNet::forward(Acts& a, bool training, const float* psiScale, float* psiOut, const int* outDims, float minValue,
                  double* loss) {
  const int N = a.N, D0 = a.D, H0 = a.H, W0 = a.W;
  const int D1 = D0 / 2, H1 = H0 / 2, W1 = W0 / 2, D2 = D0 / 4, H2 = H0 / 4, W2 = W0 / 4;
  convFailed = false;
  conv(0, a.x0, 16, 0, a.a1, 16, 0, N, D0, H0, W0, true);
  // e1 -> cat1[32:48), pooled into p1 by the convolution where it can
  if (!conv(1, a.a1, 16, 0, a.cat1, 48, 32, N, D0, H0, W0, true, nullptr, 0, 0, a.p1, 16))
    launch_pool_fwd(a.cat1, 48, 32, a.p1, 16, 16, N, D1, H1, W1, stream);
  conv(2, a.p1, 16, 0, a.a2, 32, 0, N, D1, H1, W1, true);
  conv(3, a.a2, 32, 0, a.cat2, 96, 64, N, D1, H1, W1, true);                 // e2 -> cat2[64:96)
  launch_pool_fwd(a.cat2, 96, 64, a.p2, 32, 32, N, D2, H2, W2, stream);
  conv(4, a.p2, 32, 0, a.a3, 64, 0, N, D2, H2, W2, true);
  conv(5, a.a3, 64, 0, a.b3, 64, 0, N, D2, H2, W2, true);
  // decoder: training materialises the concatenations (the weight gradients read them); inference
  // reads the upsampled part straight from the coarse tensor inside the convolution's staging
  if (training) {
    launch_up_fwd(a.b3, 64, a.cat2, 96, 0, 64, N, D2, H2, W2, stream);      // up(b3) -> cat2[0:64)
    conv(6, a.cat2, 96, 0, a.a4, 32, 0, N, D1, H1, W1, true);
  } else {
    conv(6, a.cat2, 96, 64, a.a4, 32, 0, N, D1, H1, W1, true, a.b3, 64, 64);
  }
  conv(7, a.a4, 32, 0, a.d2, 32, 0, N, D1, H1, W1, true);
  if (training) {
    launch_up_fwd(a.d2, 32, a.cat1, 48, 0, 32, N, D1, H1, W1, stream);      // up(d2) -> cat1[0:32)
    conv(8, a.cat1, 48, 0, a.a5, 16, 0, N, D0, H0, W0, true);
  } else {
    conv(8, a.cat1, 48, 32, a.a5, 16, 0, N, D0, H0, W0, true, a.d2, 32, 32);
  }
  conv(9, a.a5, 16, 0, a.d1, 16, 0, N, D0, H0, W0, true);
  UCHECK(!convFailed, "U-Net convolution refused: three x-y planes of the input exceed 32-bit offsets");
  HeadArgs h{};
  h.d1 = a.d1; h.d1_cs = 16;
  h.wo = head.w; h.bias = head.w + 16;
  h.avg = a.avg;
  h.nvox = (size_t)N * D0 * H0 * W0;
  h.minValue = minValue;
  h.scale = psiScale ? *psiScale : 1.f;
  if (training) {
    h.target = a.target;
    h.dr = a.dr;
    h.loss_partial = lossPart;
  }
  h.psi = psiOut;
  h.W = W0; h.H = H0; h.D = D0;
  if (outDims) { h.ox = outDims[0]; h.oy = outDims[1]; h.oz = outDims[2]; h.zlo = outDims[3]; }
  const int nb = launch_head(h, stream);
  MVR_HIP(hipGetLastError());
  if (training && loss) {
    std::vector<double> lp(nb);
    MVR_HIP(hipMemcpyAsync(lp.data(), lossPart, nb * 8, hipMemcpyDeviceToHost, stream));
    MVR_HIP(hipStreamSynchronize(stream));
    double s = 0;
    for (double x : lp) s += x;
    *loss = s / (double)h.nvox;
  }
  return 0;
}

int UNet::backward(Acts& a) {
  const int N = a.N, D0 = a.D, H0 = a.H, W0 = a.W;
  const int D1 = D0 / 2, H1 = H0 / 2, W1 = W0 / 2, D2 = D0 / 4, H2 = H0 / 4, W2 = W0 / 4;
  const size_t v0 = (size_t)N * D0 * H0 * W0, v1 = v0 / 8, v2 = v0 / 64;
  convFailed = false;
  HeadArgs h{};
  h.d1 = a.d1; h.d1_cs = 16; h.wo = head.w; h.bias = head.w + 16; h.avg = a.avg; h.dr = a.dr; h.nvox = v0;
  const int hb = launch_head_bwd(h, a.g_d1, bpart, stream);
  launch_sum_partials(bpart, hb, 17, head.g, stream);
  // decoder level 1
  launch_relu_bwd(a.g_d1, 16, 0, a.d1, 16, 0, 16, v0, stream);
  wgrad(9, a.a5, 16, a.g_d1, 16, 0, N, D0, H0, W0);
  dgrad(9, a.g_d1, 16, 0, a.g_a5, 16, 0, N, D0, H0, W0);
  launch_relu_bwd(a.g_a5, 16, 0, a.a5, 16, 0, 16, v0, stream);
  wgrad(8, a.cat1, 48, a.g_a5, 16, 0, N, D0, H0, W0);
  dgrad(8, a.g_a5, 16, 0, a.g_cat1, 48, 0, N, D0, H0, W0);
  // decoder level 2
  launch_up_bwd(a.g_cat1, 48, 0, a.g_d2, 32, 32, N, D1, H1, W1, stream);
  launch_relu_bwd(a.g_d2, 32, 0, a.d2, 32, 0, 32, v1, stream);
  wgrad(7, a.a4, 32, a.g_d2, 32, 0, N, D1, H1, W1);
  dgrad(7, a.g_d2, 32, 0, a.g_a4, 32, 0, N, D1, H1, W1);
  launch_relu_bwd(a.g_a4, 32, 0, a.a4, 32, 0, 32, v1, stream);
  wgrad(6, a.cat2, 96, a.g_a4, 32, 0, N, D1, H1, W1);
  dgrad(6, a.g_a4, 32, 0, a.g_cat2, 96, 0, N, D1, H1, W1);
  // bottleneck
  launch_up_bwd(a.g_cat2, 96, 0, a.g_b3, 64, 64, N, D2, H2, W2, stream);
  launch_relu_bwd(a.g_b3, 64, 0, a.b3, 64, 0, 64, v2, stream);
  wgrad(5, a.a3, 64, a.g_b3, 64, 0, N, D2, H2, W2);
  dgrad(5, a.g_b3, 64, 0, a.g_a3, 64, 0, N, D2, H2, W2);
  launch_relu_bwd(a.g_a3, 64, 0, a.a3, 64, 0, 64, v2, stream);
  wgrad(4, a.p2, 32, a.g_a3, 64, 0, N, D2, H2, W2);
  dgrad(4, a.g_a3, 64, 0, a.g_p2, 32, 0, N, D2, H2, W2);
  // encoder level 2: e2 receives the skip gradient (g_cat2[64:96)) plus the pooled path
  launch_pool_bwd(a.g_p2, 32, a.g_cat2, 96, 64, 32, N, D2, H2, W2, 1, stream);
  launch_relu_bwd(a.g_cat2, 96, 64, a.cat2, 96, 64, 32, v1, stream);
  wgrad(3, a.a2, 32, a.g_cat2, 96, 64, N, D1, H1, W1);
  dgrad(3, a.g_cat2, 96, 64, a.g_a2, 32, 0, N, D1, H1, W1);
  launch_relu_bwd(a.g_a2, 32, 0, a.a2, 32, 0, 32, v1, stream);
  wgrad(2, a.p1, 16, a.g_a2, 32, 0, N, D1, H1, W1);
  dgrad(2, a.g_a2, 32, 0, a.g_p1, 16, 0, N, D1, H1, W1);
  // encoder level 1
  launch_pool_bwd(a.g_p1, 16, a.g_cat1, 48, 32, 16, N, D1, H1, W1, 1, stream);
  launch_relu_bwd(a.g_cat1, 48, 32, a.cat1, 48, 32, 16, v0, stream);
  wgrad(1, a.a1, 16, a.g_cat1, 48, 32, N, D0, H0, W0);
  dgrad(1, a.g_cat1, 48, 32, a.g_a1, 16, 0, N, D0, H0, W0);
  launch_relu_bwd(a.g_a1, 16, 0, a.a1, 16, 0, 16, v0, stream);
  wgrad(0, a.x0, 16, a.g_a1, 16, 0, N, D0, H0, W0);
  UCHECK(!convFailed, "U-Net convolution refused: three x-y planes of the input exceed 32-bit offsets");
  MVR_HIP(hipGetLastError());
  return 0;
}

int UNet::adam_step(float lr) {
  ++step;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f / (1.f - std::pow(b1, (float)step)), c2 = 1.f / (1.f - std::pow(b2, (float)step));
  auto one = [&](Param& p, int l) {
    AdamArgs a{};
    a.w = p.w; a.m = p.m; a.v = p.v; a.grad = p.g; a.wbf = p.wbf; a.wt = p.wt; a.count = p.count;
    a.taps = 27; a.cout = l >= 0 ? kCout[l] : 0; a.cin = l >= 0 ? kCin[l] : 0;
    a.lr = lr; a.beta1 = b1; a.beta2 = b2; a.eps = 1e-8f; a.c1 = c1; a.c2 = c2;
    launch_adam(a, stream);
  };
  for (int l = 0; l < kLayers; ++l) {
    one(layer[l], l);
    one(bias[l], -1);
  }
  one(head, -1);
  MVR_HIP(hipGetLastError());
  return 0;
}

int UNet::ensure_scratch() {
  if (!wpart) {
    size_t maxc = 0;
    for (int l = 0; l < kLayers; ++l) maxc = std::max(maxc, wcount(l));
    MVR_HIP(hipMalloc(&wpart, (size_t)kWgradParts * maxc * 4));
    MVR_HIP(hipMalloc(&bpart, (size_t)1024 * 64 * 4));
    MVR_HIP(hipMalloc(&lossPart, 2048 * 8));  // 1024 double2
  }
  return 0;
}

// ------------------------------------------------------------------ training volumes and steps
int UNet::add_volume(const float* const* imgs, const float* const* ws, const float* target, const int N3[3]) {
  MVR_HIP(hipSetDevice(dev));
  TrainVolume tv;
  for (int d = 0; d < 3; ++d) tv.N[d] = N3[d];
  const size_t n = (size_t)N3[0] * N3[1] * N3[2];
  for (int v = 0; v < V; ++v) {
    float *a = nullptr, *b = nullptr;
    MVR_HIP(hipMalloc(&a, n * 4));
    MVR_HIP(hipMalloc(&b, n * 4));
    MVR_HIP(hipMemcpy(a, imgs[v], n * 4, hipMemcpyDefault));
    MVR_HIP(hipMemcpy(b, ws[v], n * 4, hipMemcpyDefault));
    tv.img.push_back(a);
    tv.w.push_back(b);
  }
  double mean = 0;
  if (weighted_mean(tv.img, tv.w, n, &mean)) return -1;
  tv.scale = (float)std::max(1e-6, mean);
  MVR_HIP(hipMalloc(&tv.target, n * 4));
  MVR_HIP(hipMemcpy(tv.target, target, n * 4, hipMemcpyDefault));
  vols.push_back(tv);
  return (int)vols.size() - 1;
}

int UNet::pack(Acts& a, int sample, const std::vector<float*>& img, const std::vector<float*>& w, const int N3[3],
               const int o[3], float invScale) {
  InputArgs in{};
  for (int v = 0; v < V; ++v) { in.img[v] = img[v]; in.w[v] = w[v]; }
  in.V = V;
  for (int d = 0; d < 3; ++d) in.N3[d] = N3[d];
  in.x0 = o[0]; in.y0 = o[1]; in.z0 = o[2];
  in.W = a.W; in.H = a.H; in.D = a.D;
  in.sample = sample;
  in.inv_scale = invScale;
  in.out = a.x0;
  in.avg_out = a.avg;
  launch_pack_input(in, stream);
  return 0;
}

int UNet::train_step(const int* volIdx, const int* origins, int batch, int P, float lr, double* loss, bool update) {
  MVR_HIP(hipSetDevice(dev));
  UCHECK(batch >= 1 && P >= 4 && P % 4 == 0, "bad batch or patch size");
  if (ensure_scratch() || ensure_acts(train, batch, P, P, P, true)) return -1;
  const size_t pv = (size_t)P * P * P;
  for (int b = 0; b < batch; ++b) {
    UCHECK(volIdx[b] >= 0 && volIdx[b] < (int)vols.size(), "bad training volume index");
    TrainVolume& tv = vols[volIdx[b]];
    const int* o = origins + 3 * b;
    for (int d = 0; d < 3; ++d) UCHECK(o[d] >= 0 && o[d] + P <= tv.N[d], "patch outside its volume");
    pack(train, b, tv.img, tv.w, tv.N, o, 1.f / tv.scale);
    launch_gather_target(tv.target, tv.N, o, P, 1.f / tv.scale, train.target + pv * b, stream);
  }
  if (forward(train, true, nullptr, nullptr, nullptr, 0.f, loss)) return -1;
  if (backward(train)) return -1;
  if (update && adam_step(lr)) return -1;
  MVR_HIP(hipStreamSynchronize(stream));
  return 0;
}

// weighted mean intensity sum w_v img_v / sum w_v over voxels and views (FP64 device partials,
// summed on the host); 1 when no voxel has weight
int UNet::weighted_mean(const std::vector<float*>& img, const std::vector<float*>& w, size_t n, double* mean) {
  if (ensure_scratch()) return -1;
  InputArgs in{};
  for (int v = 0; v < V; ++v) { in.img[v] = img[v]; in.w[v] = w[v]; }
  in.V = V;
  double2* part = reinterpret_cast<double2*>(lossPart);
  const int nb = launch_wsum_partials(in, n, part, stream);
  MVR_HIP(hipGetLastError());
  std::vector<double2> h(nb);
  MVR_HIP(hipMemcpyAsync(h.data(), part, nb * sizeof(double2), hipMemcpyDeviceToHost, stream));
  MVR_HIP(hipStreamSynchronize(stream));
  double s = 0, sw = 0;
  for (const double2& x : h) { s += x.x; sw += x.y; }
  *mean = sw > 0 ? s / sw : 1.0;
  return 0;
}

int UNet::infer(const std::vector<float*>& img, const std::vector<float*>& w, const int N3[3], float minValue,
                float* psiDev) {
  MVR_HIP(hipSetDevice(dev));
  UCHECK((int)img.size() == V && (int)w.size() == V, "the network was built for a different number of views");
  const size_t n = (size_t)N3[0] * N3[1] * N3[2];
  double mean = 0;
  if (weighted_mean(img, w, n, &mean)) return -1;
  const float scale = (float)std::max(1e-6, mean);
  auto up4 = [](int v) { return (v + 3) / 4 * 4; };
  const int X = up4(N3[0]), Y = up4(N3[1]), Z = up4(N3[2]);
  const long long plane = (long long)X * Y;
  // The volume runs as one window when it fits the voxel budget, else as z slabs of `slab` planes, each
  // with kInferHalo planes of context on either side (clipped at the volume's faces).  Slab starts and
  // the halo are multiples of 4, so every window keeps the volume's pooling grid, and the halo exceeds
  // the network's receptive-field radius: the kept planes are the ones a single window computes.
  int slab = Z;
  if ((long long)Z * plane > inferWindow) {
    slab = (int)((inferWindow / plane - 2 * kInferHalo) / 4 * 4);
    UCHECK(slab >= 4, "inference window budget too small for one x-y plane plus its halo");
  }
  int maxDepth = 0;
  for (int zs = 0; zs < N3[2]; zs += slab)
    maxDepth = std::max(maxDepth, up4(std::min(N3[2], zs + slab + kInferHalo) - std::max(0, zs - kInferHalo)));
  if (ensure_acts(inf, 1, maxDepth, Y, X, false)) return -1;  // one allocation for every window
  for (int zs = 0; zs < N3[2]; zs += slab) {
    const int ze = std::min(N3[2], zs + slab);
    const int ws = std::max(0, zs - kInferHalo), we = std::min(N3[2], ze + kInferHalo);
    if (ensure_acts(inf, 1, up4(we - ws), Y, X, false)) return -1;
    const int o[3] = {0, 0, ws};
    pack(inf, 0, img, w, N3, o, 1.f / scale);
    // psi planes [zs, ze) = window planes [zs - ws, ze - ws)
    const int od[4] = {N3[0], N3[1], ze - ws, zs - ws};
    if (forward(inf, false, &scale, psiDev + (size_t)ws * N3[1] * N3[0], od, minValue, nullptr)) return -1;
  }
  MVR_HIP(hipStreamSynchronize(stream));
  return 0;
}

int UNet::clear_volumes() {
  MVR_HIP(hipSetDevice(dev));
  for (auto& tv : vols) {
    for (auto p : tv.img) (void)hipFree(p);
    for (auto p : tv.w) (void)hipFree(p);
    (void)hipFree(tv.target);
  }
  vols.clear();
  train.release();
  return 0;
}

void UNet::destroy() {
  (void)hipSetDevice(dev);
  if (stream) (void)hipStreamSynchronize(stream);
  for (int l = 0; l < kLayers; ++l) {
    layer[l].release();
    bias[l].release();
  }
  head.release();
  clear_volumes();
  inf.release();
  for (void* p : {(void*)wpart, (void*)bpart, (void*)lossPart})
    if (p) (void)hipFree(p);
  wpart = nullptr; bpart = nullptr; lossPart = nullptr;
  if (stream) (void)hipStreamDestroy(stream);
  stream = nullptr;
}

}  // namespace mvr
