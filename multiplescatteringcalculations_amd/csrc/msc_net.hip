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

#ifndef MSC_NET_ZC
#define MSC_NET_ZC 4
#endif
constexpr int kNetNT = 256;         // threads per workgroup (4 waves)
constexpr int kVT = 4;              // default 16-voxel tiles per wave of the row-tiled kernels

__device__ __forceinline__ uint16_t f2bf(float f) {
  // round to nearest even
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// The input scale lives on the device as (n, sum |r|^2): s = sqrt(n / sum).
// A caller-given scale is stored as (s^2, 1); a measured one is summed into
// the second slot by k_norm_part and k_norm_finish.
__device__ __forceinline__ float scale_of(const double* scal) {
  const double sum = scal[1];
  return sum > 0.0 ? (float)sqrt(scal[0] / sum) : 1.f;
}

enum ConvEpi : int {
  kEpiHidden = 0,    // out bf16 = lrelu(acc + bias)
  kEpiLast = 1,      // out fp32 = acc + bias
  kEpiBackward = 2,  // out bf16 = acc * lrelu'(act)  (data gradient through act)
  kEpiApply64 = 3,   // out complex64  z = r + (acc + bias) / s  (preconditioner output)
  kEpiApply128 = 4,  // out complex128 z = r + (acc + bias) / s
  kEpiBackwardLin = 5  // out bf16 = acc  (data gradient into a linear node)
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

// Channels staged per voxel: the 32-channel network input and the last
// layer's gradient carry only 18 non-zero channels, so 20 are staged and the
// fragments of channels 20..31 are zero registers.  The LDS voxel stride is an
// odd multiple of 4 elements (8 bytes), which puts the 8-byte reads of 16
// consecutive voxels on 16 distinct bank pairs.
constexpr int kIOStaged = 20;  // staged channels of an 18-channel (padded to 32) input

template <int NS_, int VT = kVT>
struct Halo {
  static constexpr int NS = NS_;
  static constexpr int LDV = ((NS / 4) % 2 == 1) ? NS : NS + 4;
  static constexpr int ROWS = kTH + 2, COLS = 16 * VT + 2;
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
template <int NS, int VT = kVT>
__device__ __forceinline__ const uint16_t* tap_base(const uint16_t* sh, int sbase, int dz, int dx,
                                                    int dy) {
  constexpr int LDV = Halo<NS, VT>::LDV, COLS = Halo<NS, VT>::COLS;
  constexpr int PLANE = Halo<NS, VT>::ROWS * COLS * LDV;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  return sh + ((sbase + 1 + dz) % 3) * PLANE +
         ((1 + wave + dx) * COLS + (1 + dy + (lane & 15))) * LDV + (lane >> 4) * 4;
}

// The MFMAs of one tap over the wave's kVT voxel tiles.
template <int NI, int NO, int NS, int VT = kVT>
__device__ __forceinline__ void tap_mfma(v4f (&acc)[NO / 16][VT],
                                         const v4s (&wf)[NO / 16][NI / 16],
                                         const uint16_t* base) {
  constexpr int LDV = Halo<NS>::LDV;
  const int grp = (threadIdx.x & 63) >> 4;
#pragma unroll
  for (int t = 0; t < VT; ++t) {
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
template <int NO, int EPI, int VT = kVT>
__device__ __forceinline__ void conv_store(const ConvArgs& a, const v4f (&acc)[NO / 16][VT],
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
  for (int t = 0; t < VT; ++t) {
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
        } else if constexpr (EPI == kEpiBackwardLin) {
#pragma unroll
          for (int i = 0; i < 4; ++i) r[i] = acc[o][t][i];
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
template <int NI, int NO, int EPI, bool MIRROR, int NS, int VT = kVT>
__device__ __forceinline__ void conv_plane(const ConvArgs& a, const uint16_t* sh, int sbase,
                                           int d, int h0, int w0) {
  constexpr int NOT = NO / 16, NIB = NI / 16;
  v4f acc[NOT][VT];
#pragma unroll
  for (int o = 0; o < NOT; ++o)
#pragma unroll
    for (int t = 0; t < VT; ++t) acc[o][t] = v4f{0.f, 0.f, 0.f, 0.f};
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
    tap_mfma<NI, NO, NS, VT>(acc, wf, tap_base<NS, VT>(sh, sbase, dz, dx, dy));
  }
  conv_store<NO, EPI, VT>(a, acc, d, h0, w0);
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
template <int NI, int NO, int NS, int VT = kVT>
__device__ __forceinline__ void step_mfma32(v4f (&acc)[NO / 16][VT], const v8s (&wf)[NO / 16],
                                            const uint16_t* sh, int sbase, int s) {
  constexpr int LDV = Halo<NS>::LDV;
  const int lane = threadIdx.x & 63;
  const int k = 32 * s + 8 * (lane >> 4);
  const int tap = k / NI, ch = k - tap * NI;
  const bool live = tap < 27;
  const int tq = live ? tap : 26;  // an address inside the halo for the zero half
  const uint16_t* base =
      tap_base<NS, VT>(sh, sbase, tq / 9 - 1, (tq / 3) % 3 - 1, tq % 3 - 1) - (lane >> 4) * 4 + ch;
  const bool lo_ok = live && ch < NS, hi_ok = live && ch + 4 < NS;
#pragma unroll
  for (int t = 0; t < VT; ++t) {
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
template <int NI, int NO, int EPI, int NS, int WSRC, int VT = kVT>
__device__ __forceinline__ void conv_plane32(const ConvArgs& a,
                                             const v8s (&wr)[WSRC == kWReg ? k32_steps<NI>() : 1]
                                                            [NO / 16],
                                             const v8s* wl, const uint16_t* sh, int sbase, int d,
                                             int h0, int w0) {
  constexpr int NOT = NO / 16, S = k32_steps<NI>();
  const int lane = threadIdx.x & 63;
  v4f acc[NOT][VT];
#pragma unroll
  for (int o = 0; o < NOT; ++o)
#pragma unroll
    for (int t = 0; t < VT; ++t) acc[o][t] = v4f{0.f, 0.f, 0.f, 0.f};
  if constexpr (WSRC == kWReg) {
#pragma unroll
    for (int st = 0; st < S; ++st) step_mfma32<NI, NO, NS, VT>(acc, wr[st], sh, sbase, st);
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
      step_mfma32<NI, NO, NS, VT>(acc, wf, sh, sbase, st);
    }
  }
  conv_store<NO, EPI, VT>(a, acc, d, h0, w0);
}

// NS: channels of the input that can be non-zero (staged); NI: channels stored.
template <int NI, int NO, int EPI, bool MIRROR, int NS = NI, int VT = kVT>
__global__ __launch_bounds__(kNetNT) void k_conv(ConvArgs a) {
  static_assert(NS % 4 == 0 && NS <= NI, "staged channels must be a multiple of 4");
  constexpr int LDV = Halo<NS, VT>::LDV;
  constexpr int ROWS = Halo<NS, VT>::ROWS, COLS = Halo<NS, VT>::COLS;
  __shared__ __align__(16) uint16_t sh[Halo<NS, VT>::ELEMS];
  if (a.done != nullptr && *a.done) return;
  const int D = a.D, H = a.H, W = a.W;
  const int nbw = (W + 16 * VT - 1) / (16 * VT), nbh = (H + kTH - 1) / kTH;
  int b = blockIdx.x;
  const int bw = b % nbw;
  b /= nbw;
  const int bh = b % nbh;
  const int d = b / nbh;
  const int h0 = bh * kTH, w0 = bw * 16 * VT;
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
  conv_plane<NI, NO, EPI, MIRROR, NS, VT>(a, sh, 0, d, h0, w0);
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

// The same from a general 9x9 T per voxel (tgen[tidx[v]], row-major).
template <typename R>
__device__ __forceinline__ void t0_input_gen(const cx<R>* __restrict__ r,
                                             const int* __restrict__ tidx,
                                             const cx<R>* __restrict__ tgen, size_t v, float st,
                                             uint16_t* pk) {
  const cx<R>* T = tgen + (size_t)tidx[v] * 81;
  cx<R> x[9];
#pragma unroll
  for (int c = 0; c < 9; ++c) x[c] = r[v * 9 + c];
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    cx<R> acc = czero<R>();
#pragma unroll
    for (int j = 0; j < 9; ++j) acc = cadd<R>(acc, cmul<R>(T[i * 9 + j], x[j]));
    pk[2 * i] = f2bf(st * (float)acc.x);
    pk[2 * i + 1] = f2bf(st * (float)acc.y);
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
template <int NI, int NO, int EPI, int NS = NI, int SRC = 0, int VT = kVT>
__global__ __launch_bounds__(kNetNT) void k_conv_zm(ConvArgs a, int zc) {
  static_assert(NS % 4 == 0 && NS <= NI, "staged channels must be a multiple of 4");
  static_assert(SRC == 0 || (NS == 20 && NI == kNetIOPad), "the fused input stages 20 channels");
  using RS = typename std::conditional<SRC == 2, double, float>::type;
  constexpr int LDV = Halo<NS, VT>::LDV;
  constexpr int ROWS = Halo<NS, VT>::ROWS, COLS = Halo<NS, VT>::COLS;
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
  const int nbw = (W + 16 * VT - 1) / (16 * VT), nbh = (H + kTH - 1) / kTH;
  int b = blockIdx.x;
  const int bw = b % nbw;
  b /= nbw;
  const int bh = b % nbh;
  const int d0 = (b / nbh) * zc;
  const int d1 = min(D, d0 + zc);
  const int h0 = bh * kTH, w0 = bw * 16 * VT;
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
      conv_plane32<NI, NO, EPI, NS, WSRC, VT>(a, wr, wl, sh, (d + 2) % 3, d, h0, w0);
    else
      conv_plane<NI, NO, EPI, false, NS, VT>(a, sh, (d + 2) % 3, d, h0, w0);
    __syncthreads();
  }
}

// Weight gradient of one tap (blockIdx.y) over the block's voxel range.  Both
// operands are staged transposed in LDS so that a lane's four consecutive
// voxels are one 8-byte read.  The four waves share the (NO/16) x (NI/16)
// output tiles; partial tiles are added to the FP32 gradient with atomics
// (real channels only), dW laid out [o][i][tap].
//   MODE 0 (3x3x3 conv, 27 taps): dW[o][i][tap] += sum_v g[v][o] in[v + off(tap)][i]
//     over the voxels v of the (D, H, W) grid;
//   MODE 1 (2x2x2 stride-2 down, 8 taps): dW[o][i][s] += sum_c g[c][o] in[f(c, s)][i]
//     over the voxels c of the coarse (D, H, W) grid, `in` on the fine
//     (Df, Hf, Wf) grid, f(c, s) = 2c + (s >> 2, (s >> 1) & 1, s & 1);
//   MODE 2 (2x2x2 transposed up, 8 taps): dW[o][i][s] += sum_c g[f(c, s)][o] in[c][i]
//     with g on the fine grid.
// Fine voxels outside the fine grid contribute zero.
constexpr int kWgVox = 64;       // voxels staged per iteration
constexpr int kWgLd = kWgVox + 4;  // LDS row stride (elements)

struct WgradGrid {
  int D, H, W;     // the grid the voxel loop runs over (coarse for MODE 1, 2)
  int Df, Hf, Wf;  // fine grid of MODE 1, 2
};

// Fine voxel f(c, s) of coarse voxel c, or -1 outside the fine grid.
__device__ __forceinline__ long long fine_of(size_t c, int s, const WgradGrid& q) {
  const int cw = (int)(c % q.W);
  const size_t t = c / q.W;
  const int ch = (int)(t % q.H), cd = (int)(t / q.H);
  const int fd = 2 * cd + (s >> 2), fh = 2 * ch + ((s >> 1) & 1), fw = 2 * cw + (s & 1);
  if (fd >= q.Df || fh >= q.Hf || fw >= q.Wf) return -1;
  return ((long long)fd * q.Hf + fh) * q.Wf + fw;
}

template <int NI, int NO, int MODE = 0>
__global__ __launch_bounds__(kNetNT) void k_conv_wgrad(const uint16_t* __restrict__ in,
                                                       const uint16_t* __restrict__ g,
                                                       float* __restrict__ dW, int cin, int cout,
                                                       WgradGrid q, int vox_per_block) {
  constexpr int NOT = NO / 16, NIT = NI / 16, NTILE = NOT * NIT;
  constexpr int TPW = (NTILE + 3) / 4;
  constexpr int TAPS = MODE == 0 ? 27 : 8;
  __shared__ __align__(16) uint16_t sg[NO * kWgLd];
  __shared__ __align__(16) uint16_t sx[NI * kWgLd];
  const int tap = blockIdx.y;
  const int dz = tap / 9 - 1, dx = (tap / 3) % 3 - 1, dy = tap % 3 - 1;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int col = lane & 15, grp = lane >> 4;
  const int D = q.D, H = q.H, W = q.W;
  const size_t nvox = (size_t)D * H * W;
  const size_t vbeg = (size_t)blockIdx.x * vox_per_block;
  const size_t vend = std::min(nvox, vbeg + (size_t)vox_per_block);
  v4f acc[TPW];
#pragma unroll
  for (int k = 0; k < TPW; ++k) acc[k] = v4f{0.f, 0.f, 0.f, 0.f};
  // staging moves 8 channels (16 bytes) per load and transposes them into LDS
  static_assert(NI % 8 == 0 && NO % 8 == 0, "channel counts must be multiples of 8");
  typedef short v8s __attribute__((ext_vector_type(8)));
  for (size_t c0 = vbeg; c0 < vend; c0 += kWgVox) {
    for (int e = threadIdx.x; e < kWgVox * (NO / 8); e += kNetNT) {
      const int v = e / (NO / 8), c = (e - v * (NO / 8)) * 8;
      const size_t gv = c0 + v;
      v8s val = v8s{0, 0, 0, 0, 0, 0, 0, 0};
      if (gv < vend) {
        if constexpr (MODE == 2) {
          const long long f = fine_of(gv, tap, q);
          if (f >= 0) val = *reinterpret_cast<const v8s*>(g + (size_t)f * NO + c);
        } else {
          val = *reinterpret_cast<const v8s*>(g + gv * NO + c);
        }
      }
#pragma unroll
      for (int k = 0; k < 8; ++k) sg[(c + k) * kWgLd + v] = (uint16_t)val[k];
    }
    for (int e = threadIdx.x; e < kWgVox * (NI / 8); e += kNetNT) {
      const int v = e / (NI / 8), c = (e - v * (NI / 8)) * 8;
      const size_t gv = c0 + v;
      v8s val = v8s{0, 0, 0, 0, 0, 0, 0, 0};
      if (gv < vend) {
        if constexpr (MODE == 0) {
          const int gw = (int)(gv % W);
          const size_t t = gv / W;
          const int nh = (int)(t % H) + dx, nd = (int)(t / H) + dz, nw = gw + dy;
          if (nd >= 0 && nd < D && nh >= 0 && nh < H && nw >= 0 && nw < W)
            val = *reinterpret_cast<const v8s*>(in + (((size_t)nd * H + nh) * W + nw) * NI + c);
        } else if constexpr (MODE == 1) {
          const long long f = fine_of(gv, tap, q);
          if (f >= 0) val = *reinterpret_cast<const v8s*>(in + (size_t)f * NI + c);
        } else {
          val = *reinterpret_cast<const v8s*>(in + gv * NI + c);
        }
      }
#pragma unroll
      for (int k = 0; k < 8; ++k) sx[(c + k) * kWgLd + v] = (uint16_t)val[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < TPW; ++k) {
      const int tile = wave + 4 * k;
      if (tile < NTILE) {
        const int ot = tile / NIT, it = tile - (tile / NIT) * NIT;
#pragma unroll
        for (int ks = 0; ks < kWgVox / 16; ++ks) {
          const v4s a =
              *reinterpret_cast<const v4s*>(&sg[(ot * 16 + col) * kWgLd + ks * 16 + grp * 4]);
          const v4s b =
              *reinterpret_cast<const v4s*>(&sx[(it * 16 + col) * kWgLd + ks * 16 + grp * 4]);
          acc[k] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[k], 0, 0, 0);
        }
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int k = 0; k < TPW; ++k) {
    const int tile = wave + 4 * k;
    if (tile >= NTILE) continue;
    const int ot = tile / NIT, it = tile - (tile / NIT) * NIT;
    const int ci = it * 16 + col;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int co = ot * 16 + grp * 4 + r;
      if (co < cout && ci < cin)
        atomicAdd(&dW[((size_t)co * cin + ci) * TAPS + tap], acc[k][r]);
    }
  }
}

// ── small-grid convolution and the 2x2x2 stride-2 down / transposed up layers ──
// On the coarse levels of the multi-resolution network a grid has a few
// hundred to a few ten thousand voxels: the halo-staged row tiles above would
// leave most of the chip idle.  These kernels give each wave one tile of 16
// consecutive voxels (flattened over the whole grid) and one 16-channel output
// block, and load the B fragments straight from global memory (the grids are
// L2-resident): lane l holds voxel v0 + l%16 and channels 4(l/16)..+3 of each
// 16-channel input block, as v_mfma_f32_16x16x16_bf16 expects.

struct DirArgs {
  const uint16_t* in;    // [input voxels][NI]
  const uint16_t* w;     // [taps][NO][NI] bf16
  const float* bias;     // [NO] (forward epilogues)
  const uint16_t* skip;  // up layers: [fine voxels][NO] added to the output (or NULL)
  const uint16_t* act;   // gradient epilogues: [output voxels][NO]
  uint16_t* out;         // [output voxels][NO]
  const int* done;       // early exit while *done (may be NULL)
  int D, H, W;           // grid of the convolution / coarse grid of a resampling layer
  int Df, Hf, Wf;        // fine grid of a resampling layer
};

enum DirEpi : int {
  kDirHidden = 0,    // out = lrelu(acc + bias)
  kDirBackward = 1,  // out = acc * lrelu'(act)
  kDirBackLin = 2,   // out = acc
  kDirUp = 3,        // out = acc + bias + skip   (transposed up layer)
  kDirDownBack = 4   // out = (acc + skip) * lrelu'(act)  (data gradient of a down layer)
};

// Epilogue of 4 channels c0..c0+3 of output voxel v.
template <int NO, int EPI>
__device__ __forceinline__ void dir_store(const DirArgs& a, const v4f& acc, size_t v, int c0) {
  float r[4];
  if constexpr (EPI == kDirHidden || EPI == kDirUp) {
    float sk[4] = {0.f, 0.f, 0.f, 0.f};
    if (EPI == kDirUp && a.skip != nullptr) {
      const v4s sv = *reinterpret_cast<const v4s*>(a.skip + v * NO + c0);
#pragma unroll
      for (int i = 0; i < 4; ++i) sk[i] = bf2f((uint16_t)sv[i]);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const float z = acc[i] + a.bias[c0 + i];
      r[i] = EPI == kDirUp ? z + sk[i] : (z > 0.f ? z : kNetSlope * z);
    }
  } else if constexpr (EPI == kDirBackLin) {
#pragma unroll
    for (int i = 0; i < 4; ++i) r[i] = acc[i];
  } else {
    float sk[4] = {0.f, 0.f, 0.f, 0.f};
    if (EPI == kDirDownBack && a.skip != nullptr) {
      const v4s sv = *reinterpret_cast<const v4s*>(a.skip + v * NO + c0);
#pragma unroll
      for (int i = 0; i < 4; ++i) sk[i] = bf2f((uint16_t)sv[i]);
    }
    const v4s av = *reinterpret_cast<const v4s*>(a.act + v * NO + c0);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const float gsum = acc[i] + sk[i];
      r[i] = bf2f((uint16_t)av[i]) > 0.f ? gsum : kNetSlope * gsum;
    }
  }
  v4s pk;
#pragma unroll
  for (int i = 0; i < 4; ++i) pk[i] = (short)f2bf(r[i]);
  *reinterpret_cast<v4s*>(a.out + v * NO + c0) = pk;
}

// 3x3x3 "same" convolution (MIRROR: negated taps, the data gradient with
// transposed weights).  One workgroup per (16-voxel tile, 16-channel output
// block); its four waves split the 27 taps (7, 7, 7, 6), issue all their
// weight and activation loads before the MFMAs (one memory round trip per
// wave instead of a chain of them), and wave 0 adds the four partial tiles in
// a fixed order (LDS) before the epilogue.
template <int NI, int NO, int EPI, bool MIRROR>
__global__ __launch_bounds__(kNetNT) void k_conv_dir(DirArgs a) {
  if (a.done != nullptr && *a.done) return;
  constexpr int NOT = NO / 16, NIB = NI / 16, TPW = 7;
  __shared__ v4f part[kNetNT / 64][64];
  const int lane = threadIdx.x & 63, col = lane & 15, grp = lane >> 4;
  const int wave = threadIdx.x >> 6;
  const size_t nv = (size_t)a.D * a.H * a.W;
  const int ob = (int)(blockIdx.x % NOT);
  const size_t v = (size_t)(blockIdx.x / NOT) * 16 + col;
  const bool live = v < nv;
  int vw = 0, vh = 0, vd = 0;
  if (live) {
    vw = (int)(v % a.W);
    const size_t t = v / a.W;
    vh = (int)(t % a.H);
    vd = (int)(t / a.H);
  }
  v4s bf[TPW][NIB], af[TPW][NIB];
#pragma unroll
  for (int k = 0; k < TPW; ++k) {
    const int tap = wave * TPW + k;
    int dz = tap / 9 - 1, dx = (tap / 3) % 3 - 1, dy = tap % 3 - 1;
    if (MIRROR) {
      dz = -dz;
      dx = -dx;
      dy = -dy;
    }
    const int nd = vd + dz, nh = vh + dx, nw = vw + dy;
    const bool ok = tap < 27 && live && nd >= 0 && nd < a.D && nh >= 0 && nh < a.H && nw >= 0 &&
                    nw < a.W;
    const uint16_t* src = a.in + (((size_t)(ok ? nd : 0) * a.H + (ok ? nh : 0)) * a.W +
                                  (ok ? nw : 0)) * NI + grp * 4;
    const uint16_t* wsrc = a.w + ((size_t)(tap < 27 ? tap : 0) * NO + ob * 16 + col) * NI + grp * 4;
#pragma unroll
    for (int ib = 0; ib < NIB; ++ib) {
      bf[k][ib] = v4s{0, 0, 0, 0};
      af[k][ib] = v4s{0, 0, 0, 0};
      if (ok) bf[k][ib] = *reinterpret_cast<const v4s*>(src + ib * 16);
      if (tap < 27) af[k][ib] = *reinterpret_cast<const v4s*>(wsrc + ib * 16);
    }
  }
  v4f acc = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int k = 0; k < TPW; ++k)
#pragma unroll
    for (int ib = 0; ib < NIB; ++ib)
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(af[k][ib], bf[k][ib], acc, 0, 0, 0);
  part[wave][lane] = acc;
  __syncthreads();
  if (wave != 0) return;
  v4f sum = part[0][lane];
#pragma unroll
  for (int w = 1; w < kNetNT / 64; ++w) {
    const v4f q = part[w][lane];
    sum = v4f{sum[0] + q[0], sum[1] + q[1], sum[2] + q[2], sum[3] + q[3]};
  }
  // D[4 grp + i][col]: output channel ob*16 + 4 grp + i of voxel v
  if (live) dir_store<NO, EPI>(a, sum, v, ob * 16 + grp * 4);
}

// Fine voxel f(c, s) of coarse voxel c (coords cd, ch, cw), or -1 outside.
__device__ __forceinline__ long long fine_at(int cd, int ch, int cw, int s, const DirArgs& a) {
  const int fd = 2 * cd + (s >> 2), fh = 2 * ch + ((s >> 1) & 1), fw = 2 * cw + (s & 1);
  if (fd >= a.Df || fh >= a.Hf || fw >= a.Wf) return -1;
  return ((long long)fd * a.Hf + fh) * a.Wf + fw;
}

// fine -> coarse: out[c][o] = epi(sum_{s, i} w[s][o][i] in[f(c, s)][i]); EPI
// kDirHidden (a down layer) or kDirBackward (the data gradient of an up layer).
// One workgroup per (16 coarse voxels, 16-channel output block); wave w takes
// the sub-positions s = 2w, 2w + 1 and wave 0 adds the partial tiles in order.
template <int NI, int NO, int EPI>
__global__ __launch_bounds__(kNetNT) void k_down_mfma(DirArgs a) {
  if (a.done != nullptr && *a.done) return;
  constexpr int NOT = NO / 16, NIB = NI / 16;
  __shared__ v4f part[kNetNT / 64][64];
  const int lane = threadIdx.x & 63, col = lane & 15, grp = lane >> 4;
  const int wave = threadIdx.x >> 6;
  const size_t nc = (size_t)a.D * a.H * a.W;
  const int ob = (int)(blockIdx.x % NOT);
  const size_t c = (size_t)(blockIdx.x / NOT) * 16 + col;
  const bool live = c < nc;
  int cw = 0, ch = 0, cd = 0;
  if (live) {
    cw = (int)(c % a.W);
    const size_t t = c / a.W;
    ch = (int)(t % a.H);
    cd = (int)(t / a.H);
  }
  v4s bf[2][NIB], af[2][NIB];
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const int s = 2 * wave + k;
    const long long f = live ? fine_at(cd, ch, cw, s, a) : -1;
    const uint16_t* src = a.in + (size_t)(f >= 0 ? f : 0) * NI + grp * 4;
    const uint16_t* wsrc = a.w + ((size_t)s * NO + ob * 16 + col) * NI + grp * 4;
#pragma unroll
    for (int ib = 0; ib < NIB; ++ib) {
      bf[k][ib] = v4s{0, 0, 0, 0};
      if (f >= 0) bf[k][ib] = *reinterpret_cast<const v4s*>(src + ib * 16);
      af[k][ib] = *reinterpret_cast<const v4s*>(wsrc + ib * 16);
    }
  }
  v4f acc = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int k = 0; k < 2; ++k)
#pragma unroll
    for (int ib = 0; ib < NIB; ++ib)
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(af[k][ib], bf[k][ib], acc, 0, 0, 0);
  part[wave][lane] = acc;
  __syncthreads();
  if (wave != 0) return;
  v4f sum = part[0][lane];
#pragma unroll
  for (int w = 1; w < kNetNT / 64; ++w) {
    const v4f q = part[w][lane];
    sum = v4f{sum[0] + q[0], sum[1] + q[1], sum[2] + q[2], sum[3] + q[3]};
  }
  if (live) dir_store<NO, EPI>(a, sum, c, ob * 16 + grp * 4);
}

// coarse -> fine: out[f(c, s)][o] = epi(sum_i w[s][o][i] in[c][i]); EPI
// kDirUp (an up layer) or kDirDownBack (the data gradient of a down layer).
// One workgroup per (16 coarse voxels, 16-channel output block); wave w
// writes the sub-positions s = 2w, 2w + 1 of its 16 coarse voxels.
template <int NI, int NO, int EPI>
__global__ __launch_bounds__(kNetNT) void k_up_mfma(DirArgs a) {
  if (a.done != nullptr && *a.done) return;
  constexpr int NOT = NO / 16, NIB = NI / 16;
  const int lane = threadIdx.x & 63, col = lane & 15, grp = lane >> 4;
  const int wave = threadIdx.x >> 6;
  const size_t nc = (size_t)a.D * a.H * a.W;
  const int ob = (int)(blockIdx.x % NOT);
  const size_t c = (size_t)(blockIdx.x / NOT) * 16 + col;
  const bool live = c < nc;
  int cw = 0, ch = 0, cd = 0;
  if (live) {
    cw = (int)(c % a.W);
    const size_t t = c / a.W;
    ch = (int)(t % a.H);
    cd = (int)(t / a.H);
  }
  v4s bf[NIB], af[2][NIB];
#pragma unroll
  for (int ib = 0; ib < NIB; ++ib) {
    bf[ib] = v4s{0, 0, 0, 0};
    if (live) bf[ib] = *reinterpret_cast<const v4s*>(a.in + c * NI + ib * 16 + grp * 4);
#pragma unroll
    for (int k = 0; k < 2; ++k)
      af[k][ib] = *reinterpret_cast<const v4s*>(
          a.w + ((size_t)(2 * wave + k) * NO + ob * 16 + col) * NI + ib * 16 + grp * 4);
  }
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const int s = 2 * wave + k;
    v4f acc = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ib = 0; ib < NIB; ++ib)
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(af[k][ib], bf[ib], acc, 0, 0, 0);
    const long long f = live ? fine_at(cd, ch, cw, s, a) : -1;
    if (f >= 0) dir_store<NO, EPI>(a, acc, (size_t)f, ob * 16 + grp * 4);
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

// Pack layer parameters (fp32 W [cout][cin][taps], b [cout]) into BF16
// wpk [taps][pout][pin], wtr [taps][pin][pout] and fp32 bpk [pout] (zero padded).
__global__ __launch_bounds__(kNetNT) void k_pack(const float* __restrict__ W,
                                                 const float* __restrict__ b, int cin, int cout,
                                                 int pin, int pout, int taps,
                                                 uint16_t* __restrict__ wpk,
                                                 uint16_t* __restrict__ wtr,
                                                 float* __restrict__ bpk) {
  const size_t tot = (size_t)taps * pout * pin;
  for (size_t e = (size_t)blockIdx.x * kNetNT + threadIdx.x; e < tot;
       e += (size_t)gridDim.x * kNetNT) {
    const int i = (int)(e % pin);
    const size_t q = e / pin;
    const int o = (int)(q % pout);
    const int tap = (int)(q / pout);
    const float val = (o < cout && i < cin) ? W[((size_t)o * cin + i) * taps + tap] : 0.f;
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
template <typename R, bool GEN = false>
__global__ __launch_bounds__(kNetNT) void k_net_gather(const cx<R>* __restrict__ r,
                                                       const cx<R>* __restrict__ tp,
                                                       const int* __restrict__ tidx,
                                                       const cx<R>* __restrict__ tgen,
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
    if constexpr (GEN)
      t0_input_gen<R>(r, tidx, tgen, v, s, pk);
    else
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

// sum |r|^2 in two fixed-order stages: per-workgroup partials into
// scal[4 + block] (four loads in flight per thread), then one workgroup adds
// them into scal[1] (no atomics: the same scale on every run).
template <typename R>
__global__ __launch_bounds__(kNetNT) void k_norm_part(const cx<R>* __restrict__ r, size_t n,
                                                      double* __restrict__ scal) {
  __shared__ double red[kNetNT / 64];
  const size_t stride = (size_t)gridDim.x * kNetNT;
  size_t e = (size_t)blockIdx.x * kNetNT + threadIdx.x;
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  for (; e + 3 * stride < n; e += 4 * stride) {
    const double2 x0 = to_c128(r[e]), x1 = to_c128(r[e + stride]);
    const double2 x2 = to_c128(r[e + 2 * stride]), x3 = to_c128(r[e + 3 * stride]);
    a0 += x0.x * x0.x + x0.y * x0.y;
    a1 += x1.x * x1.x + x1.y * x1.y;
    a2 += x2.x * x2.x + x2.y * x2.y;
    a3 += x3.x * x3.x + x3.y * x3.y;
  }
  for (; e < n; e += stride) {
    const double2 x0 = to_c128(r[e]);
    a0 += x0.x * x0.x + x0.y * x0.y;
  }
  double acc = (a0 + a1) + (a2 + a3);
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int i = 0; i < kNetNT / 64; ++i) t += red[i];
    scal[4 + blockIdx.x] = t;
  }
}

__global__ __launch_bounds__(kNetNT) void k_norm_finish(double* __restrict__ scal, int np) {
  __shared__ double red[kNetNT];
  double t = 0.0;
  for (int i = threadIdx.x; i < np; i += kNetNT) t += scal[4 + i];
  red[threadIdx.x] = t;
  __syncthreads();
  for (int h = kNetNT / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) scal[1] = red[0];
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
  for (int t = 0; t < kNetMaxTensors; ++t) {
    if (act[t]) (void)hipFree(act[t]);
    if (grad[t]) (void)hipFree(grad[t]);
    act[t] = nullptr;
    grad[t] = nullptr;
  }
  if (outf) (void)hipFree(outf);
  if (inbuf) (void)hipFree(inbuf);
  outf = nullptr;
  inbuf = nullptr;
  D = H = W = 0;
  nvox = 0;
}

static std::atomic<unsigned long long> g_net_gen{0};

int Net::ensure_grid(int D_, int H_, int W_) {
  if (D_ == D && H_ == H && W_ == W && outf != nullptr) return kOk;
  release_grid();
  gen = ++g_net_gen;
  LD[0] = D_;
  LH[0] = H_;
  LW[0] = W_;
  for (int l = 1; l < kNetLevels; ++l) {
    LD[l] = (LD[l - 1] + 1) / 2;
    LH[l] = (LH[l - 1] + 1) / 2;
    LW[l] = (LW[l - 1] + 1) / 2;
  }
  for (int l = 0; l < kNetLevels; ++l) lvox[l] = (size_t)LD[l] * LH[l] * LW[l];
  const int last = ntensors - 1;
  for (int t = 1; t < ntensors; ++t) {
    const size_t bytes = std::max<size_t>(lvox[T[t].level], 1) * T[t].ch * sizeof(uint16_t);
    if (t != last) MSC_HIP_CHECK(hipMalloc(&act[t], bytes));
    MSC_HIP_CHECK(hipMalloc(&grad[t], bytes));
  }
  MSC_HIP_CHECK(hipMalloc(&outf, std::max<size_t>(lvox[0], 1) * kNetIOPad * sizeof(float)));
  MSC_HIP_CHECK(hipMalloc(&inbuf, std::max<size_t>(lvox[0], 1) * kNetIOPad * sizeof(uint16_t)));
  D = D_;
  H = H_;
  W = W_;
  nvox = lvox[0];
  return kOk;
}

namespace {

int pad16(int c) { return (c + 15) / 16 * 16; }

// Appends a layer (and its output tensor) to the graph; returns the output tensor id.
int add_layer(Net& net, int kind, int in, int cout, bool act, int skip = -1) {
  const int k = net.nlayers++;
  NetLayer& l = net.L[k];
  l.kind = kind;
  l.in = in;
  l.out = net.ntensors++;
  l.skip = skip;
  l.cin = in == 0 ? kNetIO : net.T[in].ch;
  l.cout = cout;
  l.pin = in == 0 ? kNetIOPad : net.T[in].ch;
  l.pout = cout == kNetIO ? kNetIOPad : pad16(cout);
  l.taps = kind == kConv3 ? 27 : 8;
  l.act = act;
  const int lvl = in == 0 ? 0 : net.T[in].level;
  NetTensor& t = net.T[l.out];
  t.level = kind == kDown2 ? lvl + 1 : (kind == kUp2 ? lvl - 1 : lvl);
  t.ch = l.pout;
  t.act = act;
  if (skip >= 0) net.T[skip].skip_user = k;
  return l.out;
}

int alloc_params(Net& net) {
  size_t off = 0, wp = 0, bp = 0;
  for (int k = 0; k < net.nlayers; ++k) {
    NetLayer& l = net.L[k];
    l.woff = off;
    off += (size_t)l.taps * l.cin * l.cout;
    l.boff = off;
    off += l.cout;
    l.wpad_off = wp;
    wp += (size_t)l.taps * l.pin * l.pout;
    l.bpad_off = bp;
    bp += l.pout;
  }
  net.nparams = off;
  net.nwpad = wp;
  net.nbpad = bp;
  net.gen = ++g_net_gen;
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
  MSC_HIP_CHECK(hipMalloc(&net.scal, kScalDoubles * sizeof(double)));
  return kOk;
}

}  // namespace

int net_init(Net& net, int hidden) {
  MSC_REQUIRE(hidden == 16 || hidden == 32 || hidden == 48, "hidden width must be 16, 32 or 48");
  net.arch = kArchLocal;
  net.hidden = hidden;
  net.nlayers = 0;
  net.ntensors = 1;
  int t = add_layer(net, kConv3, 0, hidden, true);
  t = add_layer(net, kConv3, t, hidden, true);
  t = add_layer(net, kConv3, t, hidden, true);
  add_layer(net, kConv3, t, kNetIO, false);
  return alloc_params(net);
}

int net_init_unet(Net& net, const int* c) {
  MSC_REQUIRE(c != nullptr, "widths must not be NULL");
  MSC_REQUIRE(c[0] == 16, "the full-resolution width c0 must be 16");
  for (int l = 1; l < kNetLevels; ++l)
    MSC_REQUIRE(c[l] == 16 || c[l] == 32 || c[l] == 48, "level widths c1..c3 must be 16, 32 or 48");
  net.arch = kArchUNet;
  net.hidden = c[0];
  for (int l = 0; l < kNetLevels; ++l) net.widths[l] = c[l];
  net.nlayers = 0;
  net.ntensors = 1;
  int x0 = add_layer(net, kConv3, 0, c[0], true);
  x0 = add_layer(net, kConv3, x0, c[0], true);
  int x1 = add_layer(net, kDown2, x0, c[1], true);
  x1 = add_layer(net, kConv3, x1, c[1], true);
  int x2 = add_layer(net, kDown2, x1, c[2], true);
  x2 = add_layer(net, kConv3, x2, c[2], true);
  int x3 = add_layer(net, kDown2, x2, c[3], true);
  x3 = add_layer(net, kConv3, x3, c[3], true);
  int y = add_layer(net, kUp2, x3, c[2], false, x2);
  y = add_layer(net, kConv3, y, c[2], true);
  y = add_layer(net, kUp2, y, c[1], false, x1);
  y = add_layer(net, kConv3, y, c[1], true);
  y = add_layer(net, kUp2, y, c[0], false, x0);
  y = add_layer(net, kConv3, y, c[0], true);
  add_layer(net, kConv3, y, kNetIO, false);
  return alloc_params(net);
}

int net_pack(Net& net, hipStream_t s) {
  for (int k = 0; k < net.nlayers; ++k) {
    const NetLayer& l = net.L[k];
    const size_t tot = (size_t)l.taps * l.pin * l.pout;
    hipLaunchKernelGGL(k_pack, dim3(grid_of(tot)), dim3(kNetNT), 0, s, net.params + l.woff,
                       net.params + l.boff, l.cin, l.cout, l.pin, l.pout, l.taps,
                       net.wpk + l.wpad_off, net.wtr + l.wpad_off, net.bpk + l.bpad_off);
  }
  MSC_LAUNCH_CHECK();
  net.packed = true;
  return kOk;
}

namespace {

// Planes per workgroup of the z-marching forward convolution: 4 stages 1.5
// input planes per output plane and leaves 512 workgroups on a 128x128x32
// grid (2 and 8 planes measured slower there).
constexpr int kConvZc = MSC_NET_ZC;

ConvArgs conv_args(const Net& net, int lvl, const uint16_t* in, const uint16_t* w,
                   const float* bias, const uint16_t* act, void* out, const int* done) {
  ConvArgs a{};
  a.in = in;
  a.w = w;
  a.bias = bias;
  a.act = act;
  a.out = out;
  a.done = done;
  a.D = net.LD[lvl];
  a.H = net.LH[lvl];
  a.W = net.LW[lvl];
  return a;
}

// Tiles per wave of the row-tiled kernels for a grid row of W voxels: 2 (rows of
// 32 voxels) when that leaves fewer idle voxels than rows of 64, else 4 (at
// W = 96 the 32-voxel rows took a hidden 16-channel layer from 21 to 14 us;
// at W = 128 the 64-voxel rows are faster).
int row_vt(int W) {
  const int p4 = (W + 63) / 64 * 64, p2 = (W + 31) / 32 * 32;
  return p2 < p4 ? 2 : 4;
}

dim3 conv_grid(const ConvArgs& a) {
  const int tw = 16 * row_vt(a.W);
  const int nbw = (a.W + tw - 1) / tw, nbh = (a.H + kTH - 1) / kTH;
  return dim3((unsigned)(a.D * nbh * nbw));
}

dim3 conv_grid_zm(const ConvArgs& a, int zc) {
  const int tw = 16 * row_vt(a.W);
  const int nbw = (a.W + tw - 1) / tw, nbh = (a.H + kTH - 1) / kTH;
  return dim3((unsigned)(((a.D + zc - 1) / zc) * nbh * nbw));
}

// Run-time width -> template instantiation (hidden widths 16, 32, 48, 64).
#define MSC_WIDTH_SWITCH(h, ...)   \
  switch (h) {                     \
    case 16: {                     \
      constexpr int HW = 16;       \
      __VA_ARGS__;                 \
    } break;                       \
    case 32: {                     \
      constexpr int HW = 32;       \
      __VA_ARGS__;                 \
    } break;                       \
    case 48: {                     \
      constexpr int HW = 48;       \
      __VA_ARGS__;                 \
    } break;                       \
    default: {                     \
      constexpr int HW = 64;       \
      __VA_ARGS__;                 \
    } break;                       \
  }

template <int NI, int NO, int EPI, int NS = NI, int SRC = 0>
void launch_zm(const ConvArgs& a, hipStream_t s) {
  if (row_vt(a.W) == 2)
    hipLaunchKernelGGL((k_conv_zm<NI, NO, EPI, NS, SRC, 2>), conv_grid_zm(a, kConvZc),
                       dim3(kNetNT), 0, s, a, kConvZc);
  else
    hipLaunchKernelGGL((k_conv_zm<NI, NO, EPI, NS, SRC, 4>), conv_grid_zm(a, kConvZc),
                       dim3(kNetNT), 0, s, a, kConvZc);
}

// Data gradient (mirrored taps, transposed weights) on the row-tiled kernel.
template <int NI, int NO, int EPI, int NS>
void launch_dgrad(const ConvArgs& a, hipStream_t s) {
  if (row_vt(a.W) == 2)
    hipLaunchKernelGGL((k_conv<NI, NO, EPI, true, NS, 2>), conv_grid(a), dim3(kNetNT), 0, s, a);
  else
    hipLaunchKernelGGL((k_conv<NI, NO, EPI, true, NS, 4>), conv_grid(a), dim3(kNetNT), 0, s, a);
}

// Forward of convolution layer k.  src: 0 = bf16 input grid, 1/2 = computed from
// a complex64/complex128 vector (first layer of an application); epi of the
// last layer: kEpiLast, kEpiApply64 or kEpiApply128.
void conv_forward(const Net& net, int k, ConvArgs a, int src, int epi, hipStream_t s) {
  const NetLayer& l = net.L[k];
  if (l.in == 0) {
    MSC_WIDTH_SWITCH(l.pout, {
      if (src == 0) launch_zm<kNetIOPad, HW, kEpiHidden, kIOStaged, 0>(a, s);
      else if (src == 1) launch_zm<kNetIOPad, HW, kEpiHidden, kIOStaged, 1>(a, s);
      else launch_zm<kNetIOPad, HW, kEpiHidden, kIOStaged, 2>(a, s);
    })
  } else if (l.cout == kNetIO) {
    MSC_WIDTH_SWITCH(l.pin, {
      if (epi == kEpiLast) launch_zm<HW, kNetIOPad, kEpiLast>(a, s);
      else if (epi == kEpiApply64) launch_zm<HW, kNetIOPad, kEpiApply64>(a, s);
      else launch_zm<HW, kNetIOPad, kEpiApply128>(a, s);
    })
  } else {
    MSC_WIDTH_SWITCH(l.pin, launch_zm<HW, HW, kEpiHidden>(a, s))
  }
}

DirArgs dir_args(const Net& net, int lvl, const uint16_t* in, const uint16_t* w,
                 const float* bias, const uint16_t* skip, const uint16_t* act, uint16_t* out,
                 const int* done) {
  DirArgs r{};
  r.in = in;
  r.w = w;
  r.bias = bias;
  r.skip = skip;
  r.act = act;
  r.out = out;
  r.done = done;
  r.D = net.LD[lvl];
  r.H = net.LH[lvl];
  r.W = net.LW[lvl];
  if (lvl > 0) {
    r.Df = net.LD[lvl - 1];
    r.Hf = net.LH[lvl - 1];
    r.Wf = net.LW[lvl - 1];
  }
  return r;
}


// k_conv_dir / k_down_mfma / k_up_mfma with run-time widths.
template <int EPI, bool MIRROR>
void launch_conv_dir(int pin, int pout, const DirArgs& a, hipStream_t s);

// One workgroup per (16-voxel tile of the coarse grid, 16-channel output block).
dim3 tile_grid(const DirArgs& a, int nout) {
  const size_t tiles = ((size_t)a.D * a.H * a.W + 15) / 16;
  return dim3((unsigned)(tiles * (nout / 16)));
}

template <int EPI>
void launch_down(int pin, int pout, const DirArgs& a, hipStream_t s) {
  MSC_WIDTH_SWITCH(pin, {
    constexpr int PI = HW;
    MSC_WIDTH_SWITCH(pout, hipLaunchKernelGGL((k_down_mfma<PI, HW, EPI>), tile_grid(a, HW),
                                              dim3(kNetNT), 0, s, a))
  })
}

template <int EPI, bool MIRROR>
void launch_conv_dir(int pin, int pout, const DirArgs& a, hipStream_t s) {
  MSC_WIDTH_SWITCH(pin, {
    constexpr int PI = HW;
    MSC_WIDTH_SWITCH(pout, hipLaunchKernelGGL((k_conv_dir<PI, HW, EPI, MIRROR>), tile_grid(a, HW),
                                              dim3(kNetNT), 0, s, a))
  })
}

template <int EPI>
void launch_up(int pin, int pout, const DirArgs& a, hipStream_t s) {
  MSC_WIDTH_SWITCH(pin, {
    constexpr int PI = HW;
    MSC_WIDTH_SWITCH(pout, hipLaunchKernelGGL((k_up_mfma<PI, HW, EPI>), tile_grid(a, HW),
                                              dim3(kNetNT), 0, s, a))
  })
}

// The whole forward pass.  Training (src 0): from the bf16 input `in`, the last
// layer writes outf (fp32).  Application (src 1/2, complex64/complex128): the
// first layer computes its input from r and the T0 parameters tp, the last
// writes z = r + N / s; src 3/4: the same from a gathered bf16 input `in`.
void forward_pass(Net& net, const uint16_t* in, int src, const void* r, const void* tp,
                  const unsigned int* tmax, void* z, hipStream_t s, const int* done) {
  const int last = net.nlayers - 1;
  for (int k = 0; k < net.nlayers; ++k) {
    const NetLayer& l = net.L[k];
    const uint16_t* w = net.wpk + l.wpad_off;
    const float* b = net.bpk + l.bpad_off;
    const uint16_t* x = l.in == 0 ? in : net.act[l.in];
    if (l.kind == kConv3 && net.T[l.out].level > 0) {
      launch_conv_dir<kDirHidden, false>(
          l.pin, l.pout,
          dir_args(net, net.T[l.out].level, x, w, b, nullptr, nullptr, net.act[l.out], done), s);
    } else if (l.kind == kConv3) {
      void* out = k == last ? (src == 0 ? (void*)net.outf : z) : (void*)net.act[l.out];
      ConvArgs a = conv_args(net, net.T[l.out].level, x, w, b, nullptr, out, done);
      if (l.in == 0 && src != 0) {
        a.r = r;
        a.tp = tp;
        a.tmax = tmax;
        a.scal = net.scal;
        a.scale = net.scale_arg;
      }
      if (k == last && src != 0) {
        a.r = r;
        a.scal = net.scal;
        a.scale = net.scale_arg;
      }
      const int epi =
          src == 0 ? kEpiLast : ((src == 1 || src == 3) ? kEpiApply64 : kEpiApply128);
      conv_forward(net, k, a, l.in == 0 && src <= 2 ? src : 0, epi, s);
    } else if (l.kind == kDown2) {
      launch_down<kDirHidden>(
          l.pin, l.pout,
          dir_args(net, net.T[l.out].level, x, w, b, nullptr, nullptr, net.act[l.out], done), s);
    } else {
      launch_up<kDirUp>(l.pin, l.pout,
                        dir_args(net, net.T[l.in].level, x, w, b,
                                 l.skip >= 0 ? net.act[l.skip] : nullptr, nullptr,
                                 net.act[l.out], done),
                        s);
    }
  }
}

template <int NI, int NO, int MODE>
void launch_wgrad(const uint16_t* in, const uint16_t* g, float* dW, int cin, int cout,
                  WgradGrid q, hipStream_t s) {
  const size_t nv = (size_t)q.D * q.H * q.W;
  const int vpb = 4096;
  const dim3 wg((unsigned)((nv + vpb - 1) / vpb), MODE == 0 ? 27 : 8);
  hipLaunchKernelGGL((k_conv_wgrad<NI, NO, MODE>), wg, dim3(kNetNT), 0, s, in, g, dW, cin, cout,
                     q, vpb);
}

template <int MODE>
void wgrad_dispatch(int pin, int pout, const uint16_t* in, const uint16_t* g, float* dW, int cin,
                    int cout, WgradGrid q, hipStream_t s) {
  // pin, pout in {16, 32, 48, 64}; (32, *) also covers the 18-channel input
  MSC_WIDTH_SWITCH(pin, {
    constexpr int PI = HW;
    MSC_WIDTH_SWITCH(pout, (launch_wgrad<PI, HW, MODE>(in, g, dW, cin, cout, q, s)))
  })
}

// Backward of the loss gradient held in grad[last tensor] ([v][kNetIOPad]);
// `in` is the network input.  Accumulates into net.grads.
void backward_pass(Net& net, const uint16_t* in, hipStream_t s) {
  const int last = net.nlayers - 1;
  float* Gr = net.grads;
  for (int k = last; k >= 0; --k) {
    const NetLayer& l = net.L[k];
    const int lo = net.T[l.out].level, li = l.in == 0 ? 0 : net.T[l.in].level;
    const uint16_t* g = net.grad[l.out];
    const uint16_t* x = l.in == 0 ? in : net.act[l.in];
    // weight and bias gradients
    if (l.kind == kConv3) {
      WgradGrid q{net.LD[lo], net.LH[lo], net.LW[lo], 0, 0, 0};
      wgrad_dispatch<0>(l.pin, l.pout, x, g, Gr + l.woff, l.cin, l.cout, q, s);
    } else {
      const int c = l.kind == kDown2 ? lo : li, f = c - 1;
      WgradGrid q{net.LD[c], net.LH[c], net.LW[c], net.LD[f], net.LH[f], net.LW[f]};
      if (l.kind == kDown2)
        wgrad_dispatch<1>(l.pin, l.pout, x, g, Gr + l.woff, l.cin, l.cout, q, s);
      else
        wgrad_dispatch<2>(l.pin, l.pout, x, g, Gr + l.woff, l.cin, l.cout, q, s);
    }
    MSC_WIDTH_SWITCH(l.pout, hipLaunchKernelGGL((k_bias_grad<HW>), dim3(kBiasGrid), dim3(kNetNT), 0,
                                                s, g, Gr + l.boff, l.cout, net.lvox[lo]))
    if (l.in == 0) continue;
    // data gradient into the layer's input tensor
    const uint16_t* wt = net.wtr + l.wpad_off;
    if (l.kind == kConv3 && li > 0) {
      const DirArgs a = dir_args(net, li, g, wt, nullptr, nullptr, net.act[l.in], net.grad[l.in],
                                 nullptr);
      if (net.T[l.in].act)
        launch_conv_dir<kDirBackward, true>(l.pout, l.pin, a, s);
      else
        launch_conv_dir<kDirBackLin, true>(l.pout, l.pin, a, s);
    } else if (l.kind == kConv3) {
      ConvArgs a = conv_args(net, li, g, wt, nullptr, net.act[l.in], net.grad[l.in], nullptr);
      const bool mask = net.T[l.in].act;
      if (k == last)
        MSC_WIDTH_SWITCH(l.pin, launch_dgrad<kNetIOPad, HW, kEpiBackward, kIOStaged>(a, s))
      else if (mask)
        MSC_WIDTH_SWITCH(l.pin, launch_dgrad<HW, HW, kEpiBackward, HW>(a, s))
      else
        MSC_WIDTH_SWITCH(l.pin, launch_dgrad<HW, HW, kEpiBackwardLin, HW>(a, s))
    } else if (l.kind == kDown2) {
      const int su = net.T[l.in].skip_user;
      const uint16_t* gskip = su >= 0 ? net.grad[net.L[su].out] : nullptr;
      launch_up<kDirDownBack>(l.pout, l.pin,
                              dir_args(net, lo, g, wt, nullptr, gskip, net.act[l.in],
                                       net.grad[l.in], nullptr),
                              s);
    } else {
      launch_down<kDirBackward>(l.pout, l.pin,
                                dir_args(net, li, g, wt, nullptr, nullptr, net.act[l.in],
                                         net.grad[l.in], nullptr),
                                s);
    }
  }
}

}  // namespace

int net_forward_out(Net& net, int D, int H, int W, const uint16_t* in, float* out,
                    hipStream_t s) {
  MSC_REQUIRE(net.packed, "network parameters are not set");
  int rc = net.ensure_grid(D, H, W);
  if (rc) return rc;
  forward_pass(net, in, 0, nullptr, nullptr, nullptr, nullptr, s, nullptr);
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
    forward_pass(net, xin, 0, nullptr, nullptr, nullptr, nullptr, s, nullptr);
    hipLaunchKernelGGL(k_loss_grad, dim3(grid_of(nv * kNetIOPad, 4096)), dim3(kNetNT), 0, s,
                       net.outf, tgt, net.grad[net.ntensors - 1], nv, gscale, net.loss);
    backward_pass(net, xin, s);
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
    const int np = grid_of((p.n + 3) / 4, kNormParts);
    hipLaunchKernelGGL((k_norm_part<R>), dim3(np), dim3(kNetNT), 0, s, r, p.n, scal);
    hipLaunchKernelGGL(k_norm_finish, dim3(1), dim3(kNetNT), 0, s, scal, np);
  }
  return sarg;
}

// (s / t_max) T0 r as bf16 [v][32] with the plan's cubic or general T.
template <typename R>
static void gather_input(const Plan& p, const cx<R>* r, const double* scal, float sarg,
                         uint16_t* out, hipStream_t s, const int* done) {
  const cx<R>* tp;
  const cx<R>* tg;
  if constexpr (sizeof(R) == 8) {
    tp = p.tp128;
    tg = p.tgen128;
  } else {
    tp = p.tp64;
    tg = p.tgen64;
  }
  if (p.general)
    hipLaunchKernelGGL((k_net_gather<R, true>), dim3(grid_of(p.nvox)), dim3(kNetNT), 0, s, r, tp,
                       p.tidx, tg, scal, sarg, p.tmax, out, p.nvox, done);
  else
    hipLaunchKernelGGL((k_net_gather<R, false>), dim3(grid_of(p.nvox)), dim3(kNetNT), 0, s, r, tp,
                       nullptr, nullptr, scal, sarg, p.tmax, out, p.nvox, done);
}

template <typename R>
int net_input(const Plan& p, const cx<R>* r, double scale, double* scal, uint16_t* out,
              hipStream_t s, const int* done) {
  MSC_REQUIRE(p.have_tparams, "T0 is not set on this plan");
  const float sarg = set_input_scale<R>(p, r, scale, scal, s);
  gather_input<R>(p, r, scal, sarg, out, s, done);
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
  if (p.general) {
    // a general T has no fused first-layer form: gather the input, then run
    gather_input<R>(p, r, net.scal, net.scale_arg, net.inbuf, s, done);
    forward_pass(net, net.inbuf, sizeof(R) == 4 ? 3 : 4, r, nullptr, p.tmax, z, s, done);
    MSC_LAUNCH_CHECK();
    return kOk;
  }
  const void* tp;
  if constexpr (sizeof(R) == 4) {
    tp = p.tp64;
  } else {
    tp = p.tp128;
  }
  forward_pass(net, nullptr, sizeof(R) == 4 ? 1 : 2, r, tp, p.tmax, z, s, done);
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
