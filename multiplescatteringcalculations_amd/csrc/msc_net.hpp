// Learned 3-D convolutional preconditioner (BF16 MFMA, FP32 accumulation).
//
// The network maps the induced source of a residual to a correction:
//   z = r + N((s / t_max) T0 r) / s,  s = sqrt(n) / ||r||, t_max = max |T0 params|
// over the (N_z, M, M) voxel grid (channels = real and imaginary parts of the
// 9 components, interleaved).  N is a small layer graph of three kinds of
// layer: 3x3x3 "same" convolutions, 2x2x2 stride-2 convolutions to a grid
// half as fine in each axis ("down"), and 2x2x2 stride-2 transposed
// convolutions back ("up", whose output is added to a skip tensor of the
// finer grid).  Two architectures:
//   * local: four 3x3x3 convolutions 18 -> h -> h -> h -> 18 (h 16..64);
//   * unet: four resolution levels with widths (c0, c1, c2, c3):
//       x0 = conv(conv(in))                      level 0, c0
//       x1 = conv(down(x0)), x2 = conv(down(x1)) levels 1, 2
//       x3 = conv(down(x2))                      level 3
//       y2 = conv(up(x3) + x2), y1 = conv(up(y2) + x1)
//       out = conv(conv(up(y1) + x0)) -> 18
//     with leaky ReLU after every convolution and down layer but the last,
//     and linear up layers.
// Activations are channels-last BF16, channel counts padded to multiples of 16.
#pragma once

#include <cstdint>

#include "msc_common.hpp"
#include "msc_plan.hpp"

namespace msc {

constexpr int kNetIO = 18;           // real channels in and out
constexpr int kNetIOPad = 32;        // padded to a multiple of 16
constexpr float kNetSlope = 0.1f;    // leaky-ReLU negative slope
constexpr int kNetMaxLayers = 16;
constexpr int kNetMaxTensors = kNetMaxLayers + 1;
constexpr int kNetLevels = 4;

enum NetArch : int { kArchLocal = 0, kArchUNet = 1 };
enum LayerKind : int { kConv3 = 0, kDown2 = 1, kUp2 = 2 };

struct NetLayer {
  int kind = kConv3;
  int in = 0, out = 0;  // tensor ids (tensor 0 is the network input)
  int skip = -1;        // kUp2: tensor added to the output (-1: none)
  int cin = 0, cout = 0, pin = 0, pout = 0, taps = 27;
  bool act = true;      // leaky ReLU on the output
  size_t woff = 0, boff = 0;        // offsets in the flat fp32 vector ([cout][cin][taps], [cout])
  size_t wpad_off = 0, bpad_off = 0;  // offsets in the packed copies
};

struct NetTensor {
  int level = 0;       // resolution level (0 = the voxel grid)
  int ch = 0;          // padded channels
  bool act = false;    // produced by a leaky-ReLU layer
  int skip_user = -1;  // layer that adds this tensor as a skip (-1: none)
};

// Device scalars of an application's input scale: (n, sum |r|^2, 2 spare),
// then kNormParts per-workgroup partial sums of |r|^2.
constexpr int kNormParts = 1024;
constexpr int kScalDoubles = 4 + kNormParts;

struct Net {
  int arch = kArchLocal;
  int hidden = 0;               // local: h; unet: c0
  int widths[kNetLevels] = {};  // unet level widths
  int nlayers = 0, ntensors = 0;
  NetLayer L[kNetMaxLayers];
  NetTensor T[kNetMaxTensors];
  size_t nparams = 0, nwpad = 0, nbpad = 0;
  int device = 0;
  // parameters: fp32 master copy, Adam moments, gradient accumulator
  float* params = nullptr;
  float* adam_m = nullptr;
  float* adam_v = nullptr;
  float* grads = nullptr;
  long long adam_t = 0;
  // BF16 packed weights [taps][pout][pin] and transposed [taps][pin][pout]; fp32 padded biases
  uint16_t* wpk = nullptr;
  uint16_t* wtr = nullptr;
  float* bpk = nullptr;
  // activation workspace for one grid
  int D = 0, H = 0, W = 0;
  int LD[kNetLevels] = {}, LH[kNetLevels] = {}, LW[kNetLevels] = {};
  size_t lvox[kNetLevels] = {};
  size_t nvox = 0;
  uint16_t* act[kNetMaxTensors] = {};   // tensor values (act[0] unused: the input is the
                                        // caller's, or computed by an application's first
                                        // layer; the last tensor goes to outf)
  uint16_t* grad[kNetMaxTensors] = {};  // gradient of each tensor (of its pre-activation for
                                        // leaky-ReLU tensors); grad[last] is the loss gradient
  float* outf = nullptr;                // last layer output, fp32 [vox][kNetIOPad]
  uint16_t* inbuf = nullptr;            // gathered input of an application with a general T
  double* loss = nullptr;               // device loss accumulator (+ partials)
  double* scal = nullptr;  // device (n, sum |r|^2, ..., partials) of the current apply:
                           // s = sqrt(n/sum); kScalDoubles doubles
  float scale_arg = 0.f;   // host: a given s of the current apply, 0 = read scal
  bool packed = false;
  // changes whenever device buffers are (re)allocated; captured graphs that
  // reference this network are valid only for one generation
  unsigned long long gen = 0;

  ~Net();
  void release_grid();
  int ensure_grid(int D, int H, int W);
};

int net_init(Net& net, int hidden);
// the 4-level network of widths c[0..3] (c0 = 16; c1..c3 in {16, 32, 48})
int net_init_unet(Net& net, const int* c);
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
