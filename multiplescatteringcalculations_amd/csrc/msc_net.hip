// Learned 3-D convolutional preconditioner on gfx950: BF16 MFMA convolution
// kernels (forward, data gradient, weight gradient) with FP32 accumulation,
// the mean-square loss, Adam, and the T0-gather / correction-scatter kernels
// that connect the network to the slab operator.
//
// Convolutions are implicit GEMMs over a channels-last BF16 grid
// [D][H][W][C] (C a multiple of 16) with 3x3x3 taps and zero padding 1:
//   out[v][o] = b[o] + sum_{tap, i} W[tap][o][i] in[v + off(tap)][i]
// One v_mfma_f32_16x16x16_bf16 multiplies a 16 (out channel) x 16 (in
// channel) weight block by 16 (in channel) x 16 (voxel) activations.  Lane l
// of a wave holds A[l%16][4(l/16)+0..3], B[4(l/16)+0..3][l%16] and
// D[4(l/16)+0..3][l%16].  The forward layers of the preconditioner use
// v_mfma_f32_16x16x32_bf16 over (tap, channel) instead (conv_plane32).  A
// workgroup stages its input halo in LDS; each wave owns 64 consecutive voxels
// of one row and every output channel, and loads the weight fragments of a tap
// once for its four voxel tiles.  The last layer of a preconditioner
// application writes z = r + N / s directly.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <new>
#include <type_traits>

#include "msc_net.hpp"

namespace msc {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef short v8s __attribute__((ext_vector_type(8)));
typedef __bf16 v8b __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int kNetNT = 256;  // threads per workgroup (4 waves)
constexpr int kVT = 4;       // 16-voxel tiles per wave

__device__ __forceinline__ uint16_t f2bf(float f) {
  // round to nearest even
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// The input scale lives on the device as (n, sum |r|^2): s = sqrt(n / sum).
// A caller-given scale is stored as (s^2, 1); a measured one is accumulated by
// k_norm_scale into the second slot.
__device__ __forceinline__ float scale_of(const double* scal) {
  const double sum = scal[1];
  return sum > 0.0 ? (float)sqrt(scal[0] / sum) : 1.f;
}

enum ConvEpi : int {
  kEpiHidden = 0,    // out bf16 = lrelu(acc + bias)
  kEpiLast = 1,      // out fp32 = acc + bias
  kEpiBackward = 2,  // out bf16 = acc * lrelu'(act)  (data gradient through act)
  kEpiApply64 = 3,   // out complex64  z = r + (acc + bias) / s  (preconditioner output)
  kEpiApply128 = 4   // out complex128 z = r + (acc + bias) / s
};

struct ConvArgs {
  const uint16_t* in;   // [D][H][W][NI] bf16
  const uint16_t* w;    // [27][NO][NI] bf16
  const float* bias;    // [NO]
  const uint16_t* act;  // kEpiBackward: [D][H][W][NO] activation of the layer input
  void* out;            // [D][H][W][NO] (or [D][H][W][9] complex for kEpiApply*)
  const void* r;        // kEpiApply*: the preconditioned vector [D][H][W][9] complex
  const double* scal;   // kEpiApply*: input scale (n, sum |r|^2)
  float scale;          // kEpiApply*: the input scale s when given (> 0), else read scal
  const void* tp;       // fused input: T0 parameters [D][H][W][4] complex (precision of r)
  const unsigned int* tmax;  // fused input: max |t| over the slab (float bits)
  const int* done;      // early exit while *done (may be NULL)
  int D, H, W;
};

// Output tile of one workgroup: TH rows (x) of TW voxels (y) in one plane (z);
// wave i computes row i as TW/16 tiles of 16 voxels.  The input halo
// (3 planes, TH+2 rows, TW+2 columns, NI channels) is staged in LDS once, with
// the voxel stride padded to NI+4 channels so that the 8-byte fragment reads of
// 16 consecutive voxels fall on distinct banks; every tap then reads its B
// fragments from LDS.  MIRROR negates the tap offsets (data gradient).
constexpr int kTH = kNetNT / 64;  // rows per workgroup (one per wave)
constexpr int kTW = 16 * kVT;     // voxels per row

// Channels staged per voxel: the 32-channel network input and the last
// layer's gradient carry only 18 non-zero channels, so 20 are staged and the
// fragments of channels 20..31 are zero registers.  The LDS voxel stride is an
// odd multiple of 4 elements (8 bytes), which puts the 8-byte reads of 16
// consecutive voxels on 16 distinct bank pairs.
constexpr int kIOStaged = 20;  // staged channels of an 18-channel (padded to 32) input

template <int NS_>
struct Halo {
  static constexpr int NS = NS_;
  static constexpr int LDV = ((NS / 4) % 2 == 1) ? NS : NS + 4;
  static constexpr int ROWS = kTH + 2, COLS = kTW + 2;
  static constexpr int ELEMS = 3 * ROWS * COLS * LDV;
};

// Weight fragments of one tap: lane l holds W[tap][o*16 + l%16][ib*16 + 4(l/16) ..].
template <int NI, int NO>
__device__ __forceinline__ void load_wf(v4s (&wf)[NO / 16][NI / 16], const uint16_t* w, int tap) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int o = 0; o < NO / 16; ++o)
#pragma unroll
    for (int ib = 0; ib < NI / 16; ++ib)
      wf[o][ib] = *reinterpret_cast<const v4s*>(
          w + ((size_t)tap * NO + o * 16 + (lane & 15)) * NI + ib * 16 + (lane >> 4) * 4);
}

// The lane's first B-fragment element for tap offset (dz, dx, dy) in the
// staged halo whose plane d sits in slot (sbase + 1) mod 3.
template <int NS>
__device__ __forceinline__ const uint16_t* tap_base(const uint16_t* sh, int sbase, int dz, int dx,
                                                    int dy) {
  constexpr int LDV = Halo<NS>::LDV, COLS = Halo<NS>::COLS;
  constexpr int PLANE = Halo<NS>::ROWS * COLS * LDV;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  return sh + ((sbase + 1 + dz) % 3) * PLANE +
         ((1 + wave + dx) * COLS + (1 + dy + (lane & 15))) * LDV + (lane >> 4) * 4;
}

// The MFMAs of one tap over the wave's kVT voxel tiles.
template <int NI, int NO, int NS>
__device__ __forceinline__ void tap_mfma(v4f (&acc)[NO / 16][kVT],
                                         const v4s (&wf)[NO / 16][NI / 16],
                                         const uint16_t* base) {
  constexpr int LDV = Halo<NS>::LDV;
  const int grp = (threadIdx.x & 63) >> 4;
#pragma unroll
  for (int t = 0; t < kVT; ++t) {
#pragma unroll
    for (int ib = 0; ib < NI / 16; ++ib) {
      v4s bf = v4s{0, 0, 0, 0};
      if (ib * 16 + 16 <= NS || ib * 16 + grp * 4 < NS)
        bf = *reinterpret_cast<const v4s*>(base + t * 16 * LDV + ib * 16);
#pragma unroll
      for (int o = 0; o < NO / 16; ++o)
        acc[o][t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(wf[o][ib], bf, acc[o][t], 0, 0, 0);
    }
  }
}

// Epilogue EPI of one output plane d: rows h0 + wave, voxels w0 .. w0 + 63.
template <int NO, int EPI>
__device__ __forceinline__ void conv_store(const ConvArgs& a, const v4f (&acc)[NO / 16][kVT],
                                           int d, int h0, int w0) {
  constexpr int NOT = NO / 16;
  const int H = a.H, W = a.W;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int col = lane & 15, grp = lane >> 4;
  const int h = h0 + wave;
  if (h >= H) return;
  float inv_s = 1.f;
  if constexpr (EPI == kEpiApply64 || EPI == kEpiApply128)
    inv_s = 1.f / (a.scale > 0.f ? a.scale : scale_of(a.scal));
#pragma unroll
  for (int t = 0; t < kVT; ++t) {
    const int w = w0 + t * 16 + col;
    if (w >= W) continue;
    const size_t v = ((size_t)d * H + h) * W + w;
#pragma unroll
    for (int o = 0; o < NOT; ++o) {
      const int c0 = o * 16 + grp * 4;
      if constexpr (EPI == kEpiLast) {
        float4 r;
        r.x = acc[o][t][0] + a.bias[c0 + 0];
        r.y = acc[o][t][1] + a.bias[c0 + 1];
        r.z = acc[o][t][2] + a.bias[c0 + 2];
        r.w = acc[o][t][3] + a.bias[c0 + 3];
        *reinterpret_cast<float4*>(reinterpret_cast<float*>(a.out) + v * NO + c0) = r;
      } else if constexpr (EPI == kEpiApply64 || EPI == kEpiApply128) {
        // channels c0..c0+3 are (Re, Im) of components c0/2 and c0/2 + 1
        using R = typename std::conditional<EPI == kEpiApply64, float, double>::type;
#pragma unroll
        for (int k = 0; k < 2; ++k) {
          const int comp = c0 / 2 + k;
          if (comp < 9) {
            const cx<R> rv = reinterpret_cast<const cx<R>*>(a.r)[v * 9 + comp];
            const float re = acc[o][t][2 * k] + a.bias[c0 + 2 * k];
            const float im = acc[o][t][2 * k + 1] + a.bias[c0 + 2 * k + 1];
            reinterpret_cast<cx<R>*>(a.out)[v * 9 + comp] =
                mk<R>(rv.x + R(inv_s * re), rv.y + R(inv_s * im));
          }
        }
      } else {
        float r[4];
        if constexpr (EPI == kEpiHidden) {
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const float z = acc[o][t][i] + a.bias[c0 + i];
            r[i] = z > 0.f ? z : kNetSlope * z;
          }
        } else {
          const v4s av = *reinterpret_cast<const v4s*>(a.act + v * NO + c0);
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const float x = bf2f((uint16_t)av[i]);
            r[i] = x > 0.f ? acc[o][t][i] : kNetSlope * acc[o][t][i];
          }
        }
        v4s pk;
#pragma unroll
        for (int i = 0; i < 4; ++i) pk[i] = (short)f2bf(r[i]);
        *reinterpret_cast<v4s*>(reinterpret_cast<uint16_t*>(a.out) + v * NO + c0) = pk;
      }
    }
  }
}

// One output plane d of a workgroup's tile (rows h0..h0+3, voxels w0..w0+63)
// from the three staged input planes d-1, d, d+1, which sit in LDS plane
// slots (sbase + 0, 1, 2) mod 3 of sh.  MFMA over the 27 taps, then the
// epilogue EPI.  The weight fragments are loaded per tap.
template <int NI, int NO, int EPI, bool MIRROR, int NS>
__device__ __forceinline__ void conv_plane(const ConvArgs& a, const uint16_t* sh, int sbase,
                                           int d, int h0, int w0) {
  constexpr int NOT = NO / 16, NIB = NI / 16;
  v4f acc[NOT][kVT];
#pragma unroll
  for (int o = 0; o < NOT; ++o)
#pragma unroll
    for (int t = 0; t < kVT; ++t) acc[o][t] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll 9
  for (int tap = 0; tap < 27; ++tap) {
    int dz = tap / 9 - 1, dx = (tap / 3) % 3 - 1, dy = tap % 3 - 1;
    if (MIRROR) {
      dz = -dz;
      dx = -dx;
      dy = -dy;
    }
    v4s wf[NOT][NIB];
    load_wf<NI, NO>(wf, a.w, tap);
    tap_mfma<NI, NO, NS>(acc, wf, tap_base<NS>(sh, sbase, dz, dx, dy));
  }
  conv_store<NO, EPI>(a, acc, d, h0, w0);
}

// Forward layers contract over K = (tap, input channel), 27 NI long, in steps
// of 32 with v_mfma_f32_16x16x32_bf16 (lane l holds A[l%16][8(l/16) + 0..7],
// B[8(l/16) + 0..7][l%16]).  Lane group g of step s covers the 8 channels at
// K index 32 s + 8 g: one tap per step for NI = 32, two taps for NI = 16; K
// beyond 27 NI (the second half of the last step when 27 NI is an odd
// multiple of 16) is zero on both operands.
template <int NI>
constexpr int k32_steps() { return (27 * NI + 31) / 32; }

// Weight fragments of step s (zero beyond the 27th tap).
template <int NI, int NO>
__device__ __forceinline__ void load_wf32(v8s (&wf)[NO / 16], const uint16_t* w, int s) {
  const int lane = threadIdx.x & 63;
  const int k = 32 * s + 8 * (lane >> 4);
  const int tap = k / NI, ch = k - tap * NI;
#pragma unroll
  for (int o = 0; o < NO / 16; ++o) {
    wf[o] = v8s{0, 0, 0, 0, 0, 0, 0, 0};
    if (tap < 27)
      wf[o] = *reinterpret_cast<const v8s*>(w + ((size_t)tap * NO + o * 16 + (lane & 15)) * NI +
                                            ch);
  }
}

// The MFMAs of step s over the wave's kVT voxel tiles; B fragments from the
// staged halo (two 8-byte reads; channels at or above NS are zero).
template <int NI, int NO, int NS>
__device__ __forceinline__ void step_mfma32(v4f (&acc)[NO / 16][kVT], const v8s (&wf)[NO / 16],
                                            const uint16_t* sh, int sbase, int s) {
  constexpr int LDV = Halo<NS>::LDV;
  const int lane = threadIdx.x & 63;
  const int k = 32 * s + 8 * (lane >> 4);
  const int tap = k / NI, ch = k - tap * NI;
  const bool live = tap < 27;
  const int tq = live ? tap : 26;  // an address inside the halo for the zero half
  const uint16_t* base =
      tap_base<NS>(sh, sbase, tq / 9 - 1, (tq / 3) % 3 - 1, tq % 3 - 1) - (lane >> 4) * 4 + ch;
  const bool lo_ok = live && ch < NS, hi_ok = live && ch + 4 < NS;
#pragma unroll
  for (int t = 0; t < kVT; ++t) {
    v4s lo = v4s{0, 0, 0, 0}, hi = v4s{0, 0, 0, 0};
    if (lo_ok) lo = *reinterpret_cast<const v4s*>(base + t * 16 * LDV);
    if (hi_ok) hi = *reinterpret_cast<const v4s*>(base + t * 16 * LDV + 4);
    const v8s bf = v8s{lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
#pragma unroll
    for (int o = 0; o < NO / 16; ++o)
      acc[o][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
          __builtin_bit_cast(v8b, wf[o]), __builtin_bit_cast(v8b, bf), acc[o][t], 0, 0, 0);
  }
}

// Where a forward layer's weight fragments live during the plane loop.
enum WeightSrc : int {
  kWGlobal = 0,  // loaded per step
  kWReg = 1,     // registers, loaded once per workgroup
  kWLds = 2      // LDS in lane order, staged once per workgroup
};

// One output plane of the forward (z-marching) convolution, K = 32 steps.
template <int NI, int NO, int EPI, int NS, int WSRC>
__device__ __forceinline__ void conv_plane32(const ConvArgs& a,
                                             const v8s (&wr)[WSRC == kWReg ? k32_steps<NI>() : 1]
                                                            [NO / 16],
                                             const v8s* wl, const uint16_t* sh, int sbase, int d,
                                             int h0, int w0) {
  constexpr int NOT = NO / 16, S = k32_steps<NI>();
  const int lane = threadIdx.x & 63;
  v4f acc[NOT][kVT];
#pragma unroll
  for (int o = 0; o < NOT; ++o)
#pragma unroll
    for (int t = 0; t < kVT; ++t) acc[o][t] = v4f{0.f, 0.f, 0.f, 0.f};
  if constexpr (WSRC == kWReg) {
#pragma unroll
    for (int st = 0; st < S; ++st) step_mfma32<NI, NO, NS>(acc, wr[st], sh, sbase, st);
  } else {
#pragma unroll 9
    for (int st = 0; st < S; ++st) {
      v8s wf[NOT];
      if constexpr (WSRC == kWLds) {
#pragma unroll
        for (int o = 0; o < NOT; ++o) wf[o] = wl[(st * NOT + o) * 64 + lane];
      } else {
        load_wf32<NI, NO>(wf, a.w, st);
      }
      step_mfma32<NI, NO, NS>(acc, wf, sh, sbase, st);
    }
  }
  conv_store<NO, EPI>(a, acc, d, h0, w0);
}

// NS: channels of the input that can be non-zero (staged); NI: channels stored.
template <int NI, int NO, int EPI, bool MIRROR, int NS = NI>
__global__ __launch_bounds__(kNetNT) void k_conv(ConvArgs a) {
  static_assert(NS % 4 == 0 && NS <= NI, "staged channels must be a multiple of 4");
  constexpr int LDV = Halo<NS>::LDV;
  constexpr int ROWS = Halo<NS>::ROWS, COLS = Halo<NS>::COLS;
  __shared__ __align__(16) uint16_t sh[Halo<NS>::ELEMS];
  if (a.done != nullptr && *a.done) return;
  const int D = a.D, H = a.H, W = a.W;
  const int nbw = (W + kTW - 1) / kTW, nbh = (H + kTH - 1) / kTH;
  int b = blockIdx.x;
  const int bw = b % nbw;
  b /= nbw;
  const int bh = b % nbh;
  const int d = b / nbh;
  const int h0 = bh * kTH, w0 = bw * kTW;
  // stage the halo: 4-channel groups (8 bytes), zero outside the grid.  The
  // loads go out in batches of SB before their LDS stores, so a thread waits
  // for memory once per batch rather than once per element.
  constexpr int G4 = NS / 4;
  constexpr int NSTAGE = 3 * ROWS * COLS * G4;
  constexpr int SB = 8;
  for (int e0 = 0; e0 < NSTAGE; e0 += SB * kNetNT) {
    v4s val[SB];
    int off[SB];
#pragma unroll
    for (int k = 0; k < SB; ++k) {
      const int e = e0 + k * kNetNT + threadIdx.x;
      val[k] = v4s{0, 0, 0, 0};
      off[k] = -1;
      if (e < NSTAGE) {
        const int g = e % G4;
        int q = e / G4;
        const int c = q % COLS;
        q /= COLS;
        const int rr = q % ROWS;
        const int pl = q / ROWS;
        const int gd = d + pl - 1, gh = h0 + rr - 1, gw = w0 + c - 1;
        off[k] = ((pl * ROWS + rr) * COLS + c) * LDV + g * 4;
        if (gd >= 0 && gd < D && gh >= 0 && gh < H && gw >= 0 && gw < W)
          val[k] = *reinterpret_cast<const v4s*>(a.in + (((size_t)gd * H + gh) * W + gw) * NI +
                                                 g * 4);
      }
    }
#pragma unroll
    for (int k = 0; k < SB; ++k)
      if (off[k] >= 0) *reinterpret_cast<v4s*>(&sh[off[k]]) = val[k];
  }
  __syncthreads();
  conv_plane<NI, NO, EPI, MIRROR, NS>(a, sh, 0, d, h0, w0);
}

// The network's input at voxel v: (s / t_max) T0 r as kNetIO bf16 channels,
// channel 2c = Re, 2c+1 = Im of component c (s / t_max passed as st).
template <typename R>
__device__ __forceinline__ void t0_input(const cx<R>* __restrict__ r,
                                         const cx<R>* __restrict__ tp, size_t v, float st,
                                         uint16_t* pk) {
  const cx<R>* t = tp + v * 4;
  const cx<R> tu = t[0], td = t[1], to = t[2], ts = t[3];
  cx<R> x[9];
#pragma unroll
  for (int c = 0; c < 9; ++c) x[c] = r[v * 9 + c];
  const cx<R> sum = cadd<R>(cadd<R>(x[3], x[4]), x[5]);
  const cx<R> dmo = csub<R>(td, to);
  const cx<R> tos = cmul<R>(to, sum);
  cx<R> y[9];
#pragma unroll
  for (int c = 0; c < 3; ++c) y[c] = cmul<R>(tu, x[c]);
#pragma unroll
  for (int c = 3; c < 6; ++c) y[c] = cadd<R>(cmul<R>(dmo, x[c]), tos);
#pragma unroll
  for (int c = 6; c < 9; ++c) y[c] = cmul<R>(ts, x[c]);
#pragma unroll
  for (int c = 0; c < 9; ++c) {
    pk[2 * c] = f2bf(st * (float)y[c].x);
    pk[2 * c + 1] = f2bf(st * (float)y[c].y);
  }
}

// s / t_max of the current application (0 when T0 vanishes everywhere).
__device__ __forceinline__ float input_scale(const double* scal, float scale,
                                             const unsigned int* tmax) {
  const float tm = __uint_as_float(*tmax);
  const float s0 = scale > 0.f ? scale : scale_of(scal);
  return tm > 0.f ? s0 / tm : 0.f;
}

// Forward convolution marching along z: a workgroup computes planes
// d0 .. d0+zc-1 of its tile with a ring of three staged input planes in LDS,
// and the loads of plane d+2 are in flight while plane d is computed.  Each
// output plane stages one new input plane (three in k_conv).  SRC 0 stages
// the bf16 grid a.in; SRC 1 / 2 (the first layer of a preconditioner
// application) stages the network input computed from the complex64 /
// complex128 vector a.r and T0 parameters a.tp, as k_net_gather would
// write it, so the application needs no separate gather pass.
template <int NI, int NO, int EPI, int NS = NI, int SRC = 0>
__global__ __launch_bounds__(kNetNT) void k_conv_zm(ConvArgs a, int zc) {
  static_assert(NS % 4 == 0 && NS <= NI, "staged channels must be a multiple of 4");
  static_assert(SRC == 0 || (NS == 20 && NI == kNetIOPad), "the fused input stages 20 channels");
  using RS = typename std::conditional<SRC == 2, double, float>::type;
  constexpr int LDV = Halo<NS>::LDV;
  constexpr int ROWS = Halo<NS>::ROWS, COLS = Halo<NS>::COLS;
  constexpr int PLANE = ROWS * COLS * LDV;
  constexpr int G4 = NS / 4;
  constexpr int NPL = ROWS * COLS * G4;
  constexpr int PT = (NPL + kNetNT - 1) / kNetNT;
  __shared__ __align__(16) uint16_t sh[3 * PLANE];
  if (a.done != nullptr && *a.done) return;
  // With one output block and at most 27 steps (NI <= 32), the weight
  // fragments (<= 108 VGPRs) are loaded once, before the plane loop, at no
  // cost in resident waves.  Global loads complete in order, so a weight load
  // issued inside the loop would also wait for the next plane's halo loads
  // issued before it.  Otherwise the fragments go to LDS once, where that
  // still leaves room for two workgroups per CU (the plane loop's first
  // barrier publishes them), or are loaded per step.
  // Layers of 48 input channels, whose steps straddle taps, keep the
  // 16x16x16 form per tap (the K = 32 steps spill there).
  constexpr bool K32 = NI % 32 == 0 || NI == 16;
  constexpr int NOT = NO / 16, S = k32_steps<NI>();
  constexpr int NWF = S * NOT * 64;
  constexpr int WSRC = !K32                                        ? kWGlobal
                       : (NOT == 1 && S <= 27)                     ? kWReg
                       : (3 * PLANE * 2 + NWF * 16 <= 80 * 1024) ? kWLds
                                                                 : kWGlobal;
  v8s wr[WSRC == kWReg ? S : 1][NOT];
  __shared__ v8s wl[WSRC == kWLds ? NWF : 1];
  if constexpr (WSRC == kWReg) {
#pragma unroll
    for (int st = 0; st < S; ++st) load_wf32<NI, NO>(wr[st], a.w, st);
  } else if constexpr (WSRC == kWLds) {
    for (int f = (int)threadIdx.x; f < NWF; f += kNetNT) {
      const int lane = f & 63, r = f >> 6;
      const int o = r % NOT, st = r / NOT;
      const int k = 32 * st + 8 * (lane >> 4);
      const int tap = k / NI, ch = k - tap * NI;
      wl[f] = tap < 27 ? *reinterpret_cast<const v8s*>(
                             a.w + ((size_t)tap * NO + o * 16 + (lane & 15)) * NI + ch)
                       : v8s{0, 0, 0, 0, 0, 0, 0, 0};
    }
  }
  const int D = a.D, H = a.H, W = a.W;
  const int nbw = (W + kTW - 1) / kTW, nbh = (H + kTH - 1) / kTH;
  int b = blockIdx.x;
  const int bw = b % nbw;
  b /= nbw;
  const int bh = b % nbh;
  const int d0 = (b / nbh) * zc;
  const int d1 = min(D, d0 + zc);
  const int h0 = bh * kTH, w0 = bw * kTW;
  // per thread: PT groups of 4 channels (bf16 grid) or PV voxels of 5 groups
  constexpr int NVX = ROWS * COLS, PV = (NVX + kNetNT - 1) / kNetNT;
  v4s val[SRC == 0 ? PT : PV * 5];
  float st = 0.f;
  if constexpr (SRC != 0) st = input_scale(a.scal, a.scale, a.tmax);
  auto load = [&](int gd) {
    if constexpr (SRC == 0) {
#pragma unroll
      for (int k = 0; k < PT; ++k) {
        const int e = k * kNetNT + (int)threadIdx.x;
        val[k] = v4s{0, 0, 0, 0};
        if (e < NPL && gd >= 0 && gd < D) {
          const int g = e % G4;
          const int q = e / G4;
          const int c = q % COLS, rr = q / COLS;
          const int gh = h0 + rr - 1, gw = w0 + c - 1;
          if (gh >= 0 && gh < H && gw >= 0 && gw < W)
            val[k] = *reinterpret_cast<const v4s*>(a.in + (((size_t)gd * H + gh) * W + gw) * NI +
                                                   g * 4);
        }
      }
    } else {
#pragma unroll
      for (int k = 0; k < PV; ++k) {
        const int q = k * kNetNT + (int)threadIdx.x;
#pragma unroll
        for (int g = 0; g < 5; ++g) val[5 * k + g] = v4s{0, 0, 0, 0};
        if (q < NVX && gd >= 0 && gd < D) {
          const int c = q % COLS, rr = q / COLS;
          const int gh = h0 + rr - 1, gw = w0 + c - 1;
          if (gh >= 0 && gh < H && gw >= 0 && gw < W) {
            uint16_t pk[kNetIO];
            t0_input<RS>(reinterpret_cast<const cx<RS>*>(a.r),
                         reinterpret_cast<const cx<RS>*>(a.tp), ((size_t)gd * H + gh) * W + gw,
                         st, pk);
#pragma unroll
            for (int g = 0; g < 4; ++g)
              val[5 * k + g] = v4s{(short)pk[4 * g], (short)pk[4 * g + 1], (short)pk[4 * g + 2],
                                   (short)pk[4 * g + 3]};
            val[5 * k + 4] = v4s{(short)pk[16], (short)pk[17], 0, 0};
          }
        }
      }
    }
  };
  auto store = [&](int slot) {
    if constexpr (SRC == 0) {
#pragma unroll
      for (int k = 0; k < PT; ++k) {
        const int e = k * kNetNT + (int)threadIdx.x;
        if (e < NPL) {
          const int g = e % G4;
          const int q = e / G4;
          *reinterpret_cast<v4s*>(&sh[slot * PLANE + q * LDV + g * 4]) = val[k];
        }
      }
    } else {
#pragma unroll
      for (int k = 0; k < PV; ++k) {
        const int q = k * kNetNT + (int)threadIdx.x;
        if (q < NVX) {
#pragma unroll
          for (int g = 0; g < 5; ++g)
            *reinterpret_cast<v4s*>(&sh[slot * PLANE + q * LDV + g * 4]) = val[5 * k + g];
        }
      }
    }
  };
  // plane gd >= -1 lives in slot (gd + 3) % 3
  load(d0 - 1);
  store((d0 + 2) % 3);
  load(d0);
  store(d0 % 3);
  load(d0 + 1);
  for (int d = d0; d < d1; ++d) {
    store((d + 1) % 3);  // its slot held plane d-2, last read before the barrier below
    __syncthreads();
    if (d + 1 < d1) load(d + 2);
    if constexpr (K32)
      conv_plane32<NI, NO, EPI, NS, WSRC>(a, wr, wl, sh, (d + 2) % 3, d, h0, w0);
    else
      conv_plane<NI, NO, EPI, false, NS>(a, sh, (d + 2) % 3, d, h0, w0);
    __syncthreads();
  }
}

// Weight gradient of one tap (blockIdx.y): dW[o][i][tap] += sum_v g[v][o] in[v+off][i]
// over the block's voxel range.  Both operands are staged transposed in LDS so
// that a lane's four consecutive voxels are one 8-byte read.  The four waves
// share the (NO/16) x (NI/16) output tiles; partial tiles are added to the
// FP32 gradient with atomics (real channels only).
constexpr int kWgVox = 64;       // voxels staged per iteration
constexpr int kWgLd = kWgVox + 4;  // LDS row stride (elements)

template <int NI, int NO>
__global__ __launch_bounds__(kNetNT) void k_conv_wgrad(const uint16_t* __restrict__ in,
                                                       const uint16_t* __restrict__ g,
                                                       float* __restrict__ dW, int cin, int cout,
                                                       int D, int H, int W, int vox_per_block) {
  constexpr int NOT = NO / 16, NIT = NI / 16, NTILE = NOT * NIT;
  constexpr int TPW = (NTILE + 3) / 4;
  __shared__ __align__(16) uint16_t sg[NO * kWgLd];
  __shared__ __align__(16) uint16_t sx[NI * kWgLd];
  const int tap = blockIdx.y;
  const int dz = tap / 9 - 1, dx = (tap / 3) % 3 - 1, dy = tap % 3 - 1;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int col = lane & 15, grp = lane >> 4;
  const size_t nvox = (size_t)D * H * W;
  const size_t vbeg = (size_t)blockIdx.x * vox_per_block;
  const size_t vend = std::min(nvox, vbeg + (size_t)vox_per_block);
  v4f acc[TPW];
#pragma unroll
  for (int q = 0; q < TPW; ++q) acc[q] = v4f{0.f, 0.f, 0.f, 0.f};
  // staging moves 8 channels (16 bytes) per load and transposes them into LDS
  static_assert(NI % 8 == 0 && NO % 8 == 0, "channel counts must be multiples of 8");
  typedef short v8s __attribute__((ext_vector_type(8)));
  for (size_t c0 = vbeg; c0 < vend; c0 += kWgVox) {
    for (int e = threadIdx.x; e < kWgVox * (NO / 8); e += kNetNT) {
      const int v = e / (NO / 8), c = (e - v * (NO / 8)) * 8;
      const size_t gv = c0 + v;
      v8s val = v8s{0, 0, 0, 0, 0, 0, 0, 0};
      if (gv < vend) val = *reinterpret_cast<const v8s*>(g + gv * NO + c);
#pragma unroll
      for (int k = 0; k < 8; ++k) sg[(c + k) * kWgLd + v] = (uint16_t)val[k];
    }
    for (int e = threadIdx.x; e < kWgVox * (NI / 8); e += kNetNT) {
      const int v = e / (NI / 8), c = (e - v * (NI / 8)) * 8;
      const size_t gv = c0 + v;
      v8s val = v8s{0, 0, 0, 0, 0, 0, 0, 0};
      if (gv < vend) {
        const int gw = (int)(gv % W);
        const size_t q = gv / W;
        const int nh = (int)(q % H) + dx, nd = (int)(q / H) + dz, nw = gw + dy;
        if (nd >= 0 && nd < D && nh >= 0 && nh < H && nw >= 0 && nw < W)
          val = *reinterpret_cast<const v8s*>(in + (((size_t)nd * H + nh) * W + nw) * NI + c);
      }
#pragma unroll
      for (int k = 0; k < 8; ++k) sx[(c + k) * kWgLd + v] = (uint16_t)val[k];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < TPW; ++q) {
      const int tile = wave + 4 * q;
      if (tile < NTILE) {
        const int ot = tile / NIT, it = tile - (tile / NIT) * NIT;
#pragma unroll
        for (int ks = 0; ks < kWgVox / 16; ++ks) {
          const v4s a =
              *reinterpret_cast<const v4s*>(&sg[(ot * 16 + col) * kWgLd + ks * 16 + grp * 4]);
          const v4s b =
              *reinterpret_cast<const v4s*>(&sx[(it * 16 + col) * kWgLd + ks * 16 + grp * 4]);
          acc[q] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[q], 0, 0, 0);
        }
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int q = 0; q < TPW; ++q) {
    const int tile = wave + 4 * q;
    if (tile >= NTILE) continue;
    const int ot = tile / NIT, it = tile - (tile / NIT) * NIT;
    const int ci = it * 16 + col;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int co = ot * 16 + grp * 4 + r;
      if (co < cout && ci < cin) atomicAdd(&dW[((size_t)co * cin + ci) * 27 + tap], acc[q][r]);
    }
  }
}

// db[o] += sum_v g[v][o]
constexpr int kBiasGrid = 256;  // blocks: few global atomics per channel
template <int NO>
__global__ __launch_bounds__(kNetNT) void k_bias_grad(const uint16_t* __restrict__ g,
                                                      float* __restrict__ db, int cout,
                                                      size_t nvox) {
  // Each thread sums one 8-channel group (16-byte loads) in registers: the
  // thread stride S is a multiple of NO/8, so a thread's group never changes.
  // The block's sums meet in LDS (no atomics there); one global atomic per
  // channel and block.
  static_assert(NO % 8 == 0, "channel count must be a multiple of 8");
  constexpr int G8 = NO / 8;
  typedef short v8s __attribute__((ext_vector_type(8)));
  __shared__ float red[kNetNT][9];
  const size_t T = (size_t)gridDim.x * kNetNT;
  const size_t S = T - T % G8;
  const size_t tid = (size_t)blockIdx.x * kNetNT + threadIdx.x;
  float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  if (tid < S) {
    const size_t tot = nvox * G8;
    for (size_t e = tid; e < tot; e += S) {
      const v8s val = *reinterpret_cast<const v8s*>(g + e * 8);
#pragma unroll
      for (int k = 0; k < 8; ++k) acc[k] += bf2f((uint16_t)val[k]);
    }
  }
#pragma unroll
  for (int k = 0; k < 8; ++k) red[threadIdx.x][k] = acc[k];
  __syncthreads();
  // thread t's group is (blockIdx.x * kNetNT + t) % G8
  const int gb = (int)(((size_t)blockIdx.x * kNetNT) % G8);
  for (int c = threadIdx.x; c < cout; c += kNetNT) {
    const int grp = c / 8, k = c - grp * 8;
    float sum = 0.f;
    for (int t = (grp - gb + G8) % G8; t < kNetNT; t += G8) sum += red[t][k];
    atomicAdd(&db[c], sum);
  }
}

// Mean-square loss of the first kNetIO output channels against target
// [v][kNetIO]; writes d loss / d out (scaled by gscale) as bf16 [v][kNetIOPad]
// and adds the sum of squares to *loss.
__global__ __launch_bounds__(kNetNT) void k_loss_grad(const float* __restrict__ out,
                                                      const float* __restrict__ target,
                                                      uint16_t* __restrict__ g, size_t nvox,
                                                      float gscale, double* __restrict__ loss) {
  __shared__ double red[kNetNT / 64];
  double acc = 0.0;
  const size_t tot = nvox * kNetIOPad;
  for (size_t e = (size_t)blockIdx.x * kNetNT + threadIdx.x; e < tot;
       e += (size_t)gridDim.x * kNetNT) {
    const size_t v = e / kNetIOPad;
    const int c = (int)(e - v * kNetIOPad);
    float d = 0.f;
    if (c < kNetIO) d = out[e] - target[v * kNetIO + c];
    acc += (double)d * d;
    g[e] = f2bf(gscale * d);
  }
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int i = 0; i < kNetNT / 64; ++i) s += red[i];
    atomicAdd(loss, s);
  }
}

__global__ __launch_bounds__(kNetNT) void k_zero_f(float* __restrict__ p, size_t n) {
  for (size_t i = (size_t)blockIdx.x * kNetNT + threadIdx.x; i < n;
       i += (size_t)gridDim.x * kNetNT)
    p[i] = 0.f;
}

__global__ void k_zero_d(double* p, int n) {
  if (threadIdx.x < n) p[threadIdx.x] = 0.0;
}

// Adam on the flat fp32 parameters.
__global__ __launch_bounds__(kNetNT) void k_adam(float* __restrict__ p, float* __restrict__ m,
                                                 float* __restrict__ v,
                                                 const float* __restrict__ g, size_t n, float lr,
                                                 float b1, float b2, float eps, float c1,
                                                 float c2) {
  for (size_t i = (size_t)blockIdx.x * kNetNT + threadIdx.x; i < n;
       i += (size_t)gridDim.x * kNetNT) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  }
}

// Pack layer parameters (fp32 W [cout][cin][27], b [cout]) into BF16
// wpk [27][pout][pin], wtr [27][pin][pout] and fp32 bpk [pout] (zero padded).
__global__ __launch_bounds__(kNetNT) void k_pack(const float* __restrict__ W,
                                                 const float* __restrict__ b, int cin, int cout,
                                                 int pin, int pout, uint16_t* __restrict__ wpk,
                                                 uint16_t* __restrict__ wtr,
                                                 float* __restrict__ bpk) {
  const size_t tot = (size_t)27 * pout * pin;
  for (size_t e = (size_t)blockIdx.x * kNetNT + threadIdx.x; e < tot;
       e += (size_t)gridDim.x * kNetNT) {
    const int i = (int)(e % pin);
    const size_t q = e / pin;
    const int o = (int)(q % pout);
    const int tap = (int)(q / pout);
    const float val = (o < cout && i < cin) ? W[((size_t)o * cin + i) * 27 + tap] : 0.f;
    const uint16_t bv = f2bf(val);
    wpk[e] = bv;
    wtr[((size_t)tap * pin + i) * pout + o] = bv;
  }
  if (blockIdx.x == 0)
    for (int o = threadIdx.x; o < pout; o += kNetNT) bpk[o] = o < cout ? b[o] : 0.f;
}

// Network input: (s / t_max) * (T0 r) as bf16 [v][kNetIOPad], channel 2c = Re,
// 2c+1 = Im of component c; channels kNetIO.. are zero.  t_max = max |t| over
// the slab brings the induced sources to order one whatever the frequency and
// contrast units.
template <typename R>
__global__ __launch_bounds__(kNetNT) void k_net_gather(const cx<R>* __restrict__ r,
                                                       const cx<R>* __restrict__ tp,
                                                       const double* __restrict__ scal,
                                                       float scale,
                                                       const unsigned int* __restrict__ tmax,
                                                       uint16_t* __restrict__ out, size_t nvox,
                                                       const int* __restrict__ done) {
  if (done != nullptr && *done) return;
  const float s = input_scale(scal, scale, tmax);
  for (size_t v = (size_t)blockIdx.x * kNetNT + threadIdx.x; v < nvox;
       v += (size_t)gridDim.x * kNetNT) {
    uint16_t pk[kNetIOPad];
    t0_input<R>(r, tp, v, s, pk);
#pragma unroll
    for (int c = kNetIO; c < kNetIOPad; ++c) pk[c] = 0;
    uint4* dst = reinterpret_cast<uint4*>(out + v * kNetIOPad);
#pragma unroll
    for (int q = 0; q < kNetIOPad / 8; ++q) {
      uint4 w4;
      w4.x = pk[8 * q + 0] | ((uint32_t)pk[8 * q + 1] << 16);
      w4.y = pk[8 * q + 2] | ((uint32_t)pk[8 * q + 3] << 16);
      w4.z = pk[8 * q + 4] | ((uint32_t)pk[8 * q + 5] << 16);
      w4.w = pk[8 * q + 6] | ((uint32_t)pk[8 * q + 7] << 16);
      dst[q] = w4;
    }
  }
}

// first kNetIO channels of [v][kNetIOPad] -> [v][kNetIO]
__global__ __launch_bounds__(kNetNT) void k_compact_out(const float* __restrict__ in,
                                                        float* __restrict__ out, size_t nvox) {
  const size_t tot = nvox * kNetIO;
  for (size_t e = (size_t)blockIdx.x * kNetNT + threadIdx.x; e < tot;
       e += (size_t)gridDim.x * kNetNT) {
    const size_t v = e / kNetIO;
    out[e] = in[v * kNetIOPad + (e - v * kNetIO)];
  }
}

__global__ void k_set_scale(double* scal, double a, double b) {
  if (threadIdx.x == 0) {
    scal[0] = a;
    scal[1] = b;
  }
}

template <typename R>
__global__ __launch_bounds__(kNetNT) void k_norm_scale(const cx<R>* __restrict__ r, size_t n,
                                                       double* __restrict__ scal) {
  __shared__ double red[kNetNT / 64];
  double acc = 0.0;
  for (size_t e = (size_t)blockIdx.x * kNetNT + threadIdx.x; e < n;
       e += (size_t)gridDim.x * kNetNT) {
    const double2 a = to_c128(r[e]);
    acc += a.x * a.x + a.y * a.y;
  }
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int i = 0; i < kNetNT / 64; ++i) s += red[i];
    atomicAdd(&scal[1], s);
  }
}

// ── host side ───────────────────────────────────────────────────────────────

static int grid_of(size_t work, int cap = 16384) {
  return (int)std::max<size_t>(1, std::min<size_t>((work + kNetNT - 1) / kNetNT, cap));
}

Net::~Net() {
  release_grid();
  for (void* p : {(void*)params, (void*)adam_m, (void*)adam_v, (void*)grads, (void*)wpk,
                  (void*)wtr, (void*)bpk, (void*)loss, (void*)scal})
    if (p) (void)hipFree(p);
}

void Net::release_grid() {
  for (int l = 0; l < kNetLayers; ++l) {
    if (act[l]) (void)hipFree(act[l]);
    act[l] = nullptr;
  }
  for (int i = 0; i < 2; ++i) {
    if (gbuf[i]) (void)hipFree(gbuf[i]);
    gbuf[i] = nullptr;
  }
  if (outf) (void)hipFree(outf);
  outf = nullptr;
  D = H = W = 0;
  nvox = 0;
}

static std::atomic<unsigned long long> g_net_gen{0};

int Net::ensure_grid(int D_, int H_, int W_) {
  if (D_ == D && H_ == H && W_ == W && outf != nullptr) return kOk;
  release_grid();
  gen = ++g_net_gen;
  const size_t nv = (size_t)D_ * H_ * W_;
  for (int l = 1; l < kNetLayers; ++l)
    MSC_HIP_CHECK(hipMalloc(&act[l], std::max<size_t>(nv, 1) * pin[l] * sizeof(uint16_t)));
  const int widest = std::max(hidden, kNetIOPad);
  for (int i = 0; i < 2; ++i)
    MSC_HIP_CHECK(hipMalloc(&gbuf[i], std::max<size_t>(nv, 1) * widest * sizeof(uint16_t)));
  MSC_HIP_CHECK(hipMalloc(&outf, std::max<size_t>(nv, 1) * kNetIOPad * sizeof(float)));
  D = D_;
  H = H_;
  W = W_;
  nvox = nv;
  return kOk;
}

int net_init(Net& net, int hidden) {
  MSC_REQUIRE(hidden == 16 || hidden == 32 || hidden == 48 || hidden == 64,
              "hidden width must be 16, 32, 48 or 64");
  net.hidden = hidden;
  net.gen = ++g_net_gen;
  const int real[kNetLayers + 1] = {kNetIO, hidden, hidden, hidden, kNetIO};
  const int pad[kNetLayers + 1] = {kNetIOPad, hidden, hidden, hidden, kNetIOPad};
  size_t off = 0, wp = 0, bp = 0;
  for (int l = 0; l < kNetLayers; ++l) {
    net.cin[l] = real[l];
    net.cout[l] = real[l + 1];
    net.pin[l] = pad[l];
    net.pout[l] = pad[l + 1];
    net.woff[l] = off;
    off += (size_t)27 * net.cin[l] * net.cout[l];
    net.boff[l] = off;
    off += net.cout[l];
    net.wpad_off[l] = wp;
    wp += (size_t)27 * net.pin[l] * net.pout[l];
    net.bpad_off[l] = bp;
    bp += net.pout[l];
  }
  net.nparams = off;
  net.nwpad = wp;
  net.nbpad = bp;
  MSC_HIP_CHECK(hipGetDevice(&net.device));
  MSC_HIP_CHECK(hipMalloc(&net.params, off * sizeof(float)));
  MSC_HIP_CHECK(hipMalloc(&net.adam_m, off * sizeof(float)));
  MSC_HIP_CHECK(hipMalloc(&net.adam_v, off * sizeof(float)));
  MSC_HIP_CHECK(hipMalloc(&net.grads, off * sizeof(float)));
  MSC_HIP_CHECK(hipMemset(net.params, 0, off * sizeof(float)));
  MSC_HIP_CHECK(hipMemset(net.adam_m, 0, off * sizeof(float)));
  MSC_HIP_CHECK(hipMemset(net.adam_v, 0, off * sizeof(float)));
  MSC_HIP_CHECK(hipMemset(net.grads, 0, off * sizeof(float)));
  MSC_HIP_CHECK(hipMalloc(&net.wpk, wp * sizeof(uint16_t)));
  MSC_HIP_CHECK(hipMalloc(&net.wtr, wp * sizeof(uint16_t)));
  MSC_HIP_CHECK(hipMalloc(&net.bpk, bp * sizeof(float)));
  MSC_HIP_CHECK(hipMalloc(&net.loss, 4 * sizeof(double)));
  MSC_HIP_CHECK(hipMalloc(&net.scal, 4 * sizeof(double)));
  return kOk;
}

int net_pack(Net& net, hipStream_t s) {
  for (int l = 0; l < kNetLayers; ++l) {
    const size_t tot = (size_t)27 * net.pin[l] * net.pout[l];
    hipLaunchKernelGGL(k_pack, dim3(grid_of(tot)), dim3(kNetNT), 0, s, net.params + net.woff[l],
                       net.params + net.boff[l], net.cin[l], net.cout[l], net.pin[l],
                       net.pout[l], net.wpk + net.wpad_off[l], net.wtr + net.wpad_off[l],
                       net.bpk + net.bpad_off[l]);
  }
  MSC_LAUNCH_CHECK();
  net.packed = true;
  return kOk;
}

namespace {

ConvArgs conv_args(const Net& net, const uint16_t* in, const uint16_t* w, const float* bias,
                   const uint16_t* act, void* out, const int* done) {
  ConvArgs a{};
  a.in = in;
  a.w = w;
  a.bias = bias;
  a.act = act;
  a.out = out;
  a.done = done;
  a.D = net.D;
  a.H = net.H;
  a.W = net.W;
  return a;
}

dim3 conv_grid(const Net& net) {
  const int nbw = (net.W + kTW - 1) / kTW, nbh = (net.H + kTH - 1) / kTH;
  return dim3((unsigned)(net.D * nbh * nbw));
}

// Planes per workgroup of the z-marching forward convolution: 4 stages 1.5
// input planes per output plane and leaves 512 workgroups on a 128x128x32
// grid (2 and 8 planes measured slower there).
constexpr int kConvZc = 4;
int conv_zc() { return kConvZc; }

dim3 conv_grid_zm(const Net& net, int zc) {
  const int nbw = (net.W + kTW - 1) / kTW, nbh = (net.H + kTH - 1) / kTH;
  return dim3((unsigned)(((net.D + zc - 1) / zc) * nbh * nbw));
}

// The four-layer pass per hidden width.  forward() ends in outf (fp32);
// apply() ends in z = r + N/s (complex, R) with the scale of net.scal.
template <int HID>
struct NetOps {
  static void hidden_layers(Net& net, const uint16_t* in, hipStream_t s, const int* done) {
    const int zc = conv_zc();
    hipLaunchKernelGGL((k_conv_zm<kNetIOPad, HID, kEpiHidden, kIOStaged>), conv_grid_zm(net, zc),
                       dim3(kNetNT), 0, s,
                       conv_args(net, in, net.wpk + net.wpad_off[0], net.bpk + net.bpad_off[0],
                                 nullptr, net.act[1], done),
                       zc);
    hidden_2_3(net, s, done);
  }

  // The second and third layers (act[1] -> act[2] -> act[3]).
  static void hidden_2_3(Net& net, hipStream_t s, const int* done) {
    const int zc = conv_zc();
    const dim3 G = conv_grid_zm(net, zc);
    hipLaunchKernelGGL((k_conv_zm<HID, HID, kEpiHidden>), G, dim3(kNetNT), 0, s,
                       conv_args(net, net.act[1], net.wpk + net.wpad_off[1],
                                 net.bpk + net.bpad_off[1], nullptr, net.act[2], done),
                       zc);
    hipLaunchKernelGGL((k_conv_zm<HID, HID, kEpiHidden>), G, dim3(kNetNT), 0, s,
                       conv_args(net, net.act[2], net.wpk + net.wpad_off[2],
                                 net.bpk + net.bpad_off[2], nullptr, net.act[3], done),
                       zc);
  }

  static void forward(Net& net, const uint16_t* in, hipStream_t s, const int* done) {
    hidden_layers(net, in, s, done);
    const int zc = conv_zc();
    hipLaunchKernelGGL((k_conv_zm<HID, kNetIOPad, kEpiLast>), conv_grid_zm(net, zc),
                       dim3(kNetNT), 0, s,
                       conv_args(net, net.act[3], net.wpk + net.wpad_off[3],
                                 net.bpk + net.bpad_off[3], nullptr, net.outf, done),
                       zc);
  }

  // z = r + N((s / t_max) T0 r) / s; the first layer computes its input
  // from r and the T0 parameters tp itself.
  template <typename R>
  static void apply(Net& net, const cx<R>* r, const cx<R>* tp, const unsigned int* tmax,
                    cx<R>* z, hipStream_t s, const int* done) {
    const int zc = conv_zc();
    ConvArgs a0 = conv_args(net, nullptr, net.wpk + net.wpad_off[0], net.bpk + net.bpad_off[0],
                            nullptr, net.act[1], done);
    a0.r = r;
    a0.tp = tp;
    a0.tmax = tmax;
    a0.scal = net.scal;
    a0.scale = net.scale_arg;
    constexpr int SRC = sizeof(R) == 4 ? 1 : 2;
    hipLaunchKernelGGL((k_conv_zm<kNetIOPad, HID, kEpiHidden, kIOStaged, SRC>),
                       conv_grid_zm(net, zc), dim3(kNetNT), 0, s, a0, zc);
    hidden_2_3(net, s, done);
    ConvArgs a = conv_args(net, net.act[3], net.wpk + net.wpad_off[3], net.bpk + net.bpad_off[3],
                           nullptr, z, done);
    a.r = r;
    a.scal = net.scal;
    a.scale = net.scale_arg;
    constexpr int EPI = sizeof(R) == 4 ? kEpiApply64 : kEpiApply128;
    hipLaunchKernelGGL((k_conv_zm<HID, kNetIOPad, EPI>), conv_grid_zm(net, zc), dim3(kNetNT), 0,
                       s, a, zc);
  }

  // Backward of the loss gradient held in gbuf[0] ([v][kNetIOPad]); `in` is
  // the network input.  Accumulates into net.grads.
  static void backward(Net& net, const uint16_t* in, hipStream_t s) {
    const dim3 G = conv_grid(net);
    const int D = net.D, H = net.H, W = net.W;
    const size_t nv = net.nvox;
    const int vpb = 4096;
    const dim3 wg((unsigned)((nv + vpb - 1) / vpb), 27);
    const int gb = grid_of(nv * 64, 2048);
    uint16_t* g4 = net.gbuf[0];
    uint16_t* g3 = net.gbuf[1];
    uint16_t* g2 = net.gbuf[0];
    uint16_t* g1 = net.gbuf[1];
    float* Gr = net.grads;
    // layer 4 (HID -> IO)
    hipLaunchKernelGGL((k_conv_wgrad<HID, kNetIOPad>), wg, dim3(kNetNT), 0, s, net.act[3], g4,
                       Gr + net.woff[3], net.cin[3], net.cout[3], D, H, W, vpb);
    hipLaunchKernelGGL((k_bias_grad<kNetIOPad>), dim3(kBiasGrid), dim3(kNetNT), 0, s, g4,
                       Gr + net.boff[3], net.cout[3], nv);
    hipLaunchKernelGGL((k_conv<kNetIOPad, HID, kEpiBackward, true, kIOStaged>), G, dim3(kNetNT), 0,
                       s,
                       conv_args(net, g4, net.wtr + net.wpad_off[3], nullptr, net.act[3], g3,
                                 nullptr));
    // layer 3
    hipLaunchKernelGGL((k_conv_wgrad<HID, HID>), wg, dim3(kNetNT), 0, s, net.act[2], g3,
                       Gr + net.woff[2], net.cin[2], net.cout[2], D, H, W, vpb);
    hipLaunchKernelGGL((k_bias_grad<HID>), dim3(kBiasGrid), dim3(kNetNT), 0, s, g3, Gr + net.boff[2],
                       net.cout[2], nv);
    hipLaunchKernelGGL((k_conv<HID, HID, kEpiBackward, true>), G, dim3(kNetNT), 0, s,
                       conv_args(net, g3, net.wtr + net.wpad_off[2], nullptr, net.act[2], g2,
                                 nullptr));
    // layer 2
    hipLaunchKernelGGL((k_conv_wgrad<HID, HID>), wg, dim3(kNetNT), 0, s, net.act[1], g2,
                       Gr + net.woff[1], net.cin[1], net.cout[1], D, H, W, vpb);
    hipLaunchKernelGGL((k_bias_grad<HID>), dim3(kBiasGrid), dim3(kNetNT), 0, s, g2, Gr + net.boff[1],
                       net.cout[1], nv);
    hipLaunchKernelGGL((k_conv<HID, HID, kEpiBackward, true>), G, dim3(kNetNT), 0, s,
                       conv_args(net, g2, net.wtr + net.wpad_off[1], nullptr, net.act[1], g1,
                                 nullptr));
    // layer 1 (IO -> HID): weights and bias only
    hipLaunchKernelGGL((k_conv_wgrad<kNetIOPad, HID>), wg, dim3(kNetNT), 0, s, in, g1,
                       Gr + net.woff[0], net.cin[0], net.cout[0], D, H, W, vpb);
    hipLaunchKernelGGL((k_bias_grad<HID>), dim3(kBiasGrid), dim3(kNetNT), 0, s, g1, Gr + net.boff[0],
                       net.cout[0], nv);
  }
};

template <template <int> class F, typename... A>
void by_width(int hidden, A&&... args) {
  switch (hidden) {
    case 16: F<16>::run(args...); break;
    case 32: F<32>::run(args...); break;
    case 48: F<48>::run(args...); break;
    default: F<64>::run(args...); break;
  }
}

template <int HID>
struct Fwd {
  static void run(Net& n, const uint16_t* in, hipStream_t s, const int* done) {
    NetOps<HID>::forward(n, in, s, done);
  }
};
template <int HID>
struct Bwd {
  static void run(Net& n, const uint16_t* in, hipStream_t s) { NetOps<HID>::backward(n, in, s); }
};
template <int HID>
struct Apply64 {
  static void run(Net& n, const float2* r, const float2* tp, const unsigned int* tmax, float2* z,
                  hipStream_t s, const int* done) {
    NetOps<HID>::template apply<float>(n, r, tp, tmax, z, s, done);
  }
};
template <int HID>
struct Apply128 {
  static void run(Net& n, const double2* r, const double2* tp, const unsigned int* tmax,
                  double2* z, hipStream_t s, const int* done) {
    NetOps<HID>::template apply<double>(n, r, tp, tmax, z, s, done);
  }
};

}  // namespace

int net_forward(Net& net, const uint16_t* in, hipStream_t s) {
  MSC_REQUIRE(net.packed, "network parameters are not set");
  by_width<Fwd>(net.hidden, net, in, s, nullptr);
  MSC_LAUNCH_CHECK();
  return kOk;
}

int net_forward_out(Net& net, int D, int H, int W, const uint16_t* in, float* out,
                    hipStream_t s) {
  int rc = net.ensure_grid(D, H, W);
  if (rc) return rc;
  rc = net_forward(net, in, s);
  if (rc) return rc;
  hipLaunchKernelGGL(k_compact_out, dim3(grid_of(net.nvox * kNetIO)), dim3(kNetNT), 0, s,
                     net.outf, out, net.nvox);
  MSC_LAUNCH_CHECK();
  return kOk;
}

// Gradients of loss = mean over samples, voxels and channels of (N(in) - target)^2,
// summed over `batch` samples laid out contiguously.  Leaves the gradient in
// net.grads and the loss (sum of squares / count) in *loss_host.
int net_gradients(Net& net, int D, int H, int W, int batch, const uint16_t* in,
                  const float* target, double* loss_host, hipStream_t s) {
  MSC_REQUIRE(net.packed, "network parameters are not set");
  MSC_REQUIRE(batch >= 1, "batch must be >= 1");
  int rc = net.ensure_grid(D, H, W);
  if (rc) return rc;
  const size_t nv = net.nvox;
  hipLaunchKernelGGL(k_zero_f, dim3(grid_of(net.nparams)), dim3(kNetNT), 0, s, net.grads,
                     net.nparams);
  hipLaunchKernelGGL(k_zero_d, dim3(1), dim3(64), 0, s, net.loss, 4);
  const double count = (double)batch * (double)nv * kNetIO;
  const float gscale = (float)(2.0 / count);
  for (int b = 0; b < batch; ++b) {
    const uint16_t* xin = in + (size_t)b * nv * kNetIOPad;
    const float* tgt = target + (size_t)b * nv * kNetIO;
    by_width<Fwd>(net.hidden, net, xin, s, nullptr);
    hipLaunchKernelGGL(k_loss_grad, dim3(grid_of(nv * kNetIOPad, 4096)), dim3(kNetNT), 0, s,
                       net.outf, tgt, net.gbuf[0], nv, gscale, net.loss);
    by_width<Bwd>(net.hidden, net, xin, s);
  }
  MSC_LAUNCH_CHECK();
  if (loss_host != nullptr) {
    double sum = 0.0;
    MSC_HIP_CHECK(hipMemcpyAsync(&sum, net.loss, sizeof(double), hipMemcpyDeviceToHost, s));
    MSC_HIP_CHECK(hipStreamSynchronize(s));
    *loss_host = sum / count;
  }
  return kOk;
}

int net_adam(Net& net, double lr, hipStream_t s) {
  net.adam_t += 1;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - std::pow(b1, (float)net.adam_t);
  const float c2 = 1.f - std::pow(b2, (float)net.adam_t);
  hipLaunchKernelGGL(k_adam, dim3(grid_of(net.nparams)), dim3(kNetNT), 0, s, net.params,
                     net.adam_m, net.adam_v, net.grads, net.nparams, (float)lr, b1, b2, 1e-8f,
                     c1, c2);
  MSC_LAUNCH_CHECK();
  return net_pack(net, s);
}

// The float kernel argument for a given scale, or 0 when the scale is measured
// (scale <= 0) or not a positive normal float.
static float given_scale(double scale) {
  const float f = (float)scale;
  return scale > 0.0 && std::isnormal(f) ? f : 0.f;
}

// A given scale that is a positive normal float travels as a kernel argument
// (returned); otherwise it (or the measured one, scale <= 0) is stored in scal
// and 0 is returned.
template <typename R>
static float set_input_scale(const Plan& p, const cx<R>* r, double scale, double* scal,
                             hipStream_t s) {
  const float sarg = given_scale(scale);
  if (sarg == 0.f && scale > 0.0) {
    hipLaunchKernelGGL(k_set_scale, dim3(1), dim3(64), 0, s, scal, scale * scale, 1.0);
  } else if (sarg == 0.f) {
    hipLaunchKernelGGL(k_set_scale, dim3(1), dim3(64), 0, s, scal, (double)p.n, 0.0);
    hipLaunchKernelGGL((k_norm_scale<R>), dim3(grid_of(p.n, 2048)), dim3(kNetNT), 0, s, r, p.n,
                       scal);
  }
  return sarg;
}

template <typename R>
int net_input(const Plan& p, const cx<R>* r, double scale, double* scal, uint16_t* out,
              hipStream_t s, const int* done) {
  MSC_REQUIRE(p.have_tparams, "T0 is not set on this plan");
  const float sarg = set_input_scale<R>(p, r, scale, scal, s);
  const cx<R>* tp;
  if constexpr (sizeof(R) == 8) {
    tp = p.tp128;
  } else {
    tp = p.tp64;
  }
  hipLaunchKernelGGL((k_net_gather<R>), dim3(grid_of(p.nvox)), dim3(kNetNT), 0, s, r, tp, scal,
                     sarg, p.tmax, out, p.nvox, done);
  MSC_LAUNCH_CHECK();
  return kOk;
}

template <typename R>
int net_apply(Net& net, const Plan& p, const cx<R>* r, cx<R>* z, double scale, hipStream_t s,
              const int* done) {
  MSC_REQUIRE(net.packed, "network parameters are not set");
  int rc = net.ensure_grid(p.Nz, p.M, p.M);
  if (rc) return rc;
  MSC_REQUIRE(p.have_tparams, "T0 is not set on this plan");
  net.scale_arg = set_input_scale<R>(p, r, scale, net.scal, s);
  if constexpr (sizeof(R) == 4) {
    by_width<Apply64>(net.hidden, net, r, p.tp64, p.tmax, z, s, done);
  } else {
    by_width<Apply128>(net.hidden, net, r, p.tp128, p.tmax, z, s, done);
  }
  MSC_LAUNCH_CHECK();
  return kOk;
}

template int net_input<double>(const Plan&, const double2*, double, double*, uint16_t*,
                               hipStream_t, const int*);
template int net_input<float>(const Plan&, const float2*, double, double*, uint16_t*,
                              hipStream_t, const int*);
template int net_apply<double>(Net&, const Plan&, const double2*, double2*, double, hipStream_t,
                               const int*);
template int net_apply<float>(Net&, const Plan&, const float2*, float2*, double, hipStream_t,
                              const int*);

}  // namespace msc
