// unet_host.hpp -- host object of the learned initial estimate.
//
// Network (per sample, dims multiple of 4): c1a 3^3 (V+1 -> 16) ReLU, c1b (16 -> 16) ReLU = e1;
// avg-pool; c2a (16 -> 32), c2b (32 -> 32) = e2; avg-pool; c3a (32 -> 64), c3b (64 -> 64);
// nearest x2 up, concat e2 -> c4a (96 -> 32), c4b (32 -> 32); up, concat e1 -> c5a (48 -> 16),
// c5b (16 -> 16); head: 1x1 (16 -> 1) added to the view mean.  Inputs are w_v * img_v and their
// mean, divided by a per-volume scale (mean of w * img); psi = max(minValue, scale * out).
// Volume extents and patch origins are given x-fastest: N3 = {nx, ny, nz}, o = {x0, y0, z0}.
#pragma once
#include <vector>

#include "host.hpp"
#include "unet.hpp"

namespace mvr {

constexpr int kLayers = 10;
constexpr int kWgradParts = 256;

// planes of context on either side of an inference slab: a multiple of 4 above the network's receptive-field
// radius of 23 planes (4 at full resolution, 8 at half, 8 at quarter resolution, plus the pooling grid's offset)
constexpr int kInferHalo = 32;

struct UNet {
  struct Param {
    size_t count = 0;
    float *w = nullptr, *m = nullptr, *v = nullptr, *g = nullptr;
    uint16_t *wbf = nullptr, *wt = nullptr;
    int alloc(bool conv);
    void release();
  };
  struct Acts {
    int N = 0, D = 0, H = 0, W = 0;
    int Dcap = 0;  // depth the buffers were allocated for (D <= Dcap)
    uint16_t *x0 = nullptr, *a1 = nullptr, *cat1 = nullptr, *p1 = nullptr, *a2 = nullptr, *cat2 = nullptr,
             *p2 = nullptr, *a3 = nullptr, *b3 = nullptr, *a4 = nullptr, *d2 = nullptr, *a5 = nullptr,
             *d1 = nullptr;
    uint16_t *g_a1 = nullptr, *g_cat1 = nullptr, *g_p1 = nullptr, *g_a2 = nullptr, *g_cat2 = nullptr,
             *g_p2 = nullptr, *g_a3 = nullptr, *g_b3 = nullptr, *g_a4 = nullptr, *g_d2 = nullptr,
             *g_a5 = nullptr, *g_d1 = nullptr;
    float *avg = nullptr, *target = nullptr, *dr = nullptr;
    int alloc(int N, int D, int H, int W, bool grads);
    void release();
  };
  struct TrainVolume {
    int N[3];
    std::vector<float*> img, w;
    float* target = nullptr;
    float scale = 1.f;
  };

  int dev = 0, V = 0;
  long long step = 0;
  hipStream_t stream = nullptr;
  Param layer[kLayers], bias[kLayers], head;
  Acts train, inf;
  std::vector<TrainVolume> vols;
  float *wpart = nullptr, *bpart = nullptr;
  double* lossPart = nullptr;
  bool convFailed = false;         // a convolution launch was refused since the last check
  long long inferWindow = 1ll << 28;  // voxel budget of one inference window (z slab + halos)

  int create(int dev, int V, unsigned long long seed);
  void destroy();
  size_t wcount(int l) const;
  size_t param_count() const;
  int copy_params(float* hostOut, const float* hostIn, bool grads);
  int refresh_copies();
  int add_volume(const float* const* imgs, const float* const* ws, const float* target, const int N3[3]);
  int train_step(const int* volIdx, const int* origins, int batch, int P, float lr, double* loss, bool update);
  // psiDev[N3] = network estimate (device memory); extents need not be multiples of 4
  int infer(const std::vector<float*>& img, const std::vector<float*>& w, const int N3[3], float minValue,
            float* psiDev);
  int clear_volumes();
  int weighted_mean(const std::vector<float*>& img, const std::vector<float*>& w, size_t n, double* mean);

  int ensure_acts(Acts& a, int N, int D, int H, int W, bool grads);
  int ensure_scratch();
  int pack(Acts& a, int sample, const std::vector<float*>& img, const std::vector<float*>& w, const int N3[3],
           const int o[3], float invScale);
  // up != nullptr: input channels [0, up_c) are nearest x2 upsampled from up (channel stride up_cs), the
  // rest come from in + in_co
  // pool != nullptr: the 2x2x2 average pooling of the output into pool (channel stride pool_cs) where the
  // convolution fuses it; returns whether it did (else the caller pools)
  bool conv(int l, const uint16_t* in, int in_cs, int in_co, uint16_t* out, int out_cs, int out_co, int N, int D,
            int H, int W, bool relu, const uint16_t* up = nullptr, int up_cs = 0, int up_c = 0,
            uint16_t* pool = nullptr, int pool_cs = 0);
  void dgrad(int l, const uint16_t* dout, int dout_cs, int dout_co, uint16_t* din, int din_cs, int din_co, int N,
             int D, int H, int W);
  int wgrad(int l, const uint16_t* in, int in_cs, const uint16_t* dout, int dout_cs, int dout_co, int N, int D, int H,
            int W);
  int forward(Acts& a, bool training, const float* psiScale, float* psiOut, const int* outDims, float minValue,
              double* loss);
  int backward(Acts& a);
  int adam_step(float lr);
};

}  // namespace mvr
