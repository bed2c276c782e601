// unet.hpp -- argument blocks and host declarations of the U-Net kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace mvr {

constexpr int kMaxWgradTiles = 4;  // (cout/16)*(cin/16) <= 16 tiles spread over 4 waves

struct ConvArgs {
  const uint16_t* in;  // [n][z][y][x][in_cs], channels [0, cin)
  int in_cs;
  int cin;             // multiple of 16
  const uint16_t* w;   // [taps][cout][cin]
  const float* bias;   // [cout] or null
  uint16_t* out;       // [n][z][y][x][out_cs], channels [out_co, out_co + cout)
  int out_cs, out_co;
  int cout;            // multiple of 16
  int N, D, H, W;
  int taps;            // 27 (3^3, zero padding 1) or 1
  int relu;
  // optional fused nearest x2 upsampling: input channels [0, up_c) are read from up_in (channel stride
  // up_cs) at half the coordinates, channels [up_c, cin) from in (its channel 0 = input channel up_c)
  const uint16_t* up_in;
  int up_cs, up_c;
  // optional fused 2x2x2 average pooling of the stored (bf16) output into pool_out [n][D/2][H/2][W/2][pool_cs],
  // channels [pool_co, pool_co + cout): done where conv_fuses_pool(args) holds, else left to the caller
  uint16_t* pool_out;
  int pool_cs, pool_co;
};

struct WgradArgs {
  const uint16_t* in;    // forward input of the layer
  int in_cs, cin;
  const uint16_t* dout;  // gradient of the layer output (after the ReLU mask)
  int dout_cs, dout_co, cout;
  int N, D, H, W;
  int taps;
  float* partial;        // [parts][taps][cout][cin]
};

// views the network takes: V view channels + their weighted mean fit the 16 input channels; the input
// packer holds up to 8 view pointers
constexpr int kNetMaxViews = 8;

struct InputArgs {
  const float* img[kNetMaxViews];
  const float* w[kNetMaxViews];
  int V;
  int N3[3];             // dims of the source volumes (x, y, z)
  int x0, y0, z0;        // origin of the sample region
  int W, H, D;           // sample region
  int sample;            // index in the batch
  float inv_scale;
  uint16_t* out;         // [batch][D][H][W][16]
  float* avg_out;        // [batch][D][H][W] (normalised average channel) or null
};

struct HeadArgs {
  const uint16_t* d1;    // [nvox][d1_cs], 16 channels
  int d1_cs;
  const float* wo;       // 16
  const float* bias;     // 1
  const float* avg;      // normalised average channel
  const float* target;   // normalised target (training) or null
  float* dr;             // gradient of r (training)
  double* loss_partial;  // per work-group sum of squared errors (training) or null
  float* psi;            // inference output or null, dims (ox, oy, oz)
  float scale, minValue;
  size_t nvox;
  int W, H, D;           // net dims of one sample (psi output only)
  int ox, oy, oz;        // output volume dims, <= net dims; psi holds planes [zlo, oz) of them
  int zlo;
};

struct AdamArgs {
  float* w;
  float* m;
  float* v;
  const float* grad;
  uint16_t* wbf;   // bf16 copy [taps][cout][cin] or null
  uint16_t* wt;    // flipped transpose [taps][cin][cout] or null
  size_t count;
  int taps, cout, cin;
  float lr, beta1, beta2, eps, c1, c2;  // c1 = 1/(1-beta1^t), c2 = 1/(1-beta2^t)
};

bool launch_conv_fwd(const ConvArgs& a, hipStream_t s);  // false: shape not supported, nothing launched
// LDS-tiled 3^3 convolution (unet_conv.hip); false when the shape is not one it covers
bool launch_conv_tiled(const ConvArgs& a, hipStream_t s);
bool conv_fuses_pool(const ConvArgs& a);  // the launch also writes the pooled output (pool_out set, DMA kernel)
void launch_conv_wgrad(const WgradArgs& a, int nparts, hipStream_t s);
void launch_sum_partials(const float* partial, int nparts, size_t count, float* grad, hipStream_t s);
void launch_bias_grad(const uint16_t* dout, int cs, int co0, int cout, size_t nvox, float* partial, int nparts,
                      hipStream_t s);
void launch_relu_bwd(uint16_t* g, int gcs, int gco, const uint16_t* post, int pcs, int pco, int C, size_t nvox,
                     hipStream_t s);
void launch_pool_fwd(const uint16_t* in, int ics, int ico, uint16_t* out, int ocs, int C, int N, int D, int H, int W,
                     hipStream_t s);
void launch_pool_bwd(const uint16_t* dout, int ocs, uint16_t* din, int ics, int ico, int C, int N, int D, int H, int W,
                     int accumulate, hipStream_t s);
void launch_up_fwd(const uint16_t* in, int ics, uint16_t* out, int ocs, int oco, int C, int N, int D, int H, int W,
                   hipStream_t s);
void launch_up_bwd(const uint16_t* dfine, int fcs, int fco, uint16_t* dcoarse, int ccs, int C, int N, int D, int H,
                   int W, hipStream_t s);
void launch_pack_input(const InputArgs& a, hipStream_t s);
int launch_head(const HeadArgs& a, hipStream_t s);
int launch_head_bwd(const HeadArgs& a, uint16_t* dd1, float* partial, hipStream_t s);
void launch_adam(const AdamArgs& a, hipStream_t s);
void launch_f32_to_bf16(const float* src, uint16_t* dst, size_t n, hipStream_t s);
void launch_bf16_to_f32(const uint16_t* src, float* dst, size_t n, hipStream_t s);
void launch_zero_u16(uint16_t* p, size_t n, hipStream_t s);
int launch_wsum_partials(const InputArgs& a, size_t n, double2* partial, hipStream_t s);
void launch_gather_target(const float* src, const int N[3], const int o[3], int P, float invScale, float* dst,
                          hipStream_t s);

}  // namespace mvr
