// Learned 3-D convolutional preconditioner (BF16 MFMA, FP32 accumulation).
//
// The network maps the induced source of a residual to a correction:
//   z = r + N((s / t_max) T0 r) / s,  s = sqrt(n) / ||r||, t_max = max |T0 params|
// with N four 3x3x3 "same" convolutions 18 -> h -> h -> h -> 18 over the
// (N_z, M, M) voxel grid (channels = real and imaginary parts of the 9
// components, interleaved), leaky-ReLU between layers.  Activations are
// channels-last BF16, channel counts padded to multiples of 16.
#pragma once

#include <cstdint>

#include "msc_common.hpp"
#include "msc_plan.hpp"

namespace msc {

constexpr int kNetLayers = 4;
constexpr int kNetIO = 18;           // real channels in and out
constexpr int kNetIOPad = 32;        // padded to a multiple of 16
constexpr float kNetSlope = 0.1f;    // leaky-ReLU negative slope

struct Net {
  int hidden = 0;
  int cin[kNetLayers] = {}, cout[kNetLayers] = {};    // real channel counts
  int pin[kNetLayers] = {}, pout[kNetLayers] = {};    // padded counts
  size_t woff[kNetLayers] = {}, boff[kNetLayers] = {};  // offsets in the flat fp32 vector
  size_t nparams = 0;
  size_t wpad_off[kNetLayers] = {}, bpad_off[kNetLayers] = {};  // packed (padded) offsets
  size_t nwpad = 0, nbpad = 0;
  int device = 0;
  // parameters: fp32 master copy, Adam moments, gradient accumulator
  float* params = nullptr;
  float* adam_m = nullptr;
  float* adam_v = nullptr;
  float* grads = nullptr;
  long long adam_t = 0;
  // BF16 packed weights [27][pout][pin] and transposed [27][pin][pout]; fp32 padded biases
  uint16_t* wpk = nullptr;
  uint16_t* wtr = nullptr;
  float* bpk = nullptr;
  // activation workspace for one grid
  int D = 0, H = 0, W = 0;
  size_t nvox = 0;
  uint16_t* act[kNetLayers] = {};  // act[l] = output of layer l-1 (act[0] unused: the input is
                                   // the caller's, or computed by an application's first layer)
  float* outf = nullptr;           // last layer output, fp32 [vox][pout[3]]
  uint16_t* gbuf[2] = {};          // gradient ping-pong (bf16), sized for the widest layer
  double* loss = nullptr;          // device loss accumulator (+ partials)
  double* scal = nullptr;          // device (n, sum |r|^2) of the current apply: s = sqrt(n/sum)
  float scale_arg = 0.f;           // host: a given s of the current apply, 0 = read scal
  bool packed = false;
  // changes whenever device buffers are (re)allocated; captured graphs that
  // reference this network are valid only for one generation
  unsigned long long gen = 0;

  ~Net();
  void release_grid();
  int ensure_grid(int D, int H, int W);
};

int net_init(Net& net, int hidden);
// pack fp32 master parameters into the BF16 / padded copies
int net_pack(Net& net, hipStream_t s);
// forward pass of a [D][H][W][32] bf16 input; out = first 18 channels, fp32 [v][18]
int net_forward_out(Net& net, int D, int H, int W, const uint16_t* in, float* out, hipStream_t s);
// gradient of the mean-square loss over `batch` contiguous samples into net.grads
int net_gradients(Net& net, int D, int H, int W, int batch, const uint16_t* in,
                  const float* target, double* loss_host, hipStream_t s);
// one Adam step from net.grads (then re-pack)
int net_adam(Net& net, double lr, hipStream_t s);
// network input (s / t_max) T0 r as bf16 [v][32]; scale <= 0 selects s = sqrt(n) / ||r||
template <typename R>
int net_input(const Plan& p, const cx<R>* r, double scale, double* scal, uint16_t* out,
              hipStream_t s, const int* done = nullptr);
// z = r + N((s / t_max) T0 r) / s on the plan's grid (scale as for net_input); every
// kernel returns at once while *done is set (stream-ordered early exit)
template <typename R>
int net_apply(Net& net, const Plan& p, const cx<R>* r, cx<R>* z, double scale, hipStream_t s,
              const int* done = nullptr);

}  // namespace msc
