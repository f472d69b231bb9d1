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
// D[4(l/16)+0..3][l%16].  Each wave owns 64 consecutive voxels and every
// output channel; the weight fragments of a tap are loaded once per tap and
// reused over the wave's four voxel tiles.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <new>

#include "msc_net.hpp"

namespace msc {

typedef short v4s __attribute__((ext_vector_type(4)));
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

enum ConvEpi : int { kEpiHidden = 0, kEpiLast = 1, kEpiBackward = 2 };

// EPI kEpiHidden: out bf16 = lrelu(acc + bias); kEpiLast: out fp32 = acc + bias;
// kEpiBackward: out bf16 = acc * lrelu'(act) (data gradient through the
// activation `act` of the layer input).  MIRROR negates the tap offsets (the
// data gradient is a convolution with the mirrored, transposed kernel).
template <int NI, int NO, int EPI, bool MIRROR>
__global__ __launch_bounds__(kNetNT) void k_conv(const uint16_t* __restrict__ in,
                                                 const uint16_t* __restrict__ w,
                                                 const float* __restrict__ bias,
                                                 const uint16_t* __restrict__ act,
                                                 void* __restrict__ out, int D, int H, int W,
                                                 const int* __restrict__ done) {
  constexpr int NOT = NO / 16, NIB = NI / 16;
  if (done != nullptr && *done) return;
  const int lane = threadIdx.x & 63;
  const int col = lane & 15, grp = lane >> 4;
  const size_t nvox = (size_t)D * H * W;
  const size_t v0 = ((size_t)blockIdx.x * (kNetNT / 64) + (threadIdx.x >> 6)) * (16 * kVT);
  if (v0 >= nvox) return;
  int vd[kVT], vh[kVT], vw[kVT];
  bool vok[kVT];
#pragma unroll
  for (int t = 0; t < kVT; ++t) {
    const size_t v = v0 + t * 16 + col;
    vok[t] = v < nvox;
    const size_t vv = vok[t] ? v : 0;
    vw[t] = (int)(vv % W);
    const size_t q = vv / W;
    vh[t] = (int)(q % H);
    vd[t] = (int)(q / H);
  }
  v4f acc[NOT][kVT];
#pragma unroll
  for (int o = 0; o < NOT; ++o)
#pragma unroll
    for (int t = 0; t < kVT; ++t) acc[o][t] = v4f{0.f, 0.f, 0.f, 0.f};
  const v4s zero4 = v4s{0, 0, 0, 0};
  for (int tap = 0; tap < 27; ++tap) {
    int dz = tap / 9 - 1, dx = (tap / 3) % 3 - 1, dy = tap % 3 - 1;
    if (MIRROR) {
      dz = -dz;
      dx = -dx;
      dy = -dy;
    }
    v4s a[NOT][NIB];
#pragma unroll
    for (int o = 0; o < NOT; ++o)
#pragma unroll
      for (int ib = 0; ib < NIB; ++ib)
        a[o][ib] = *reinterpret_cast<const v4s*>(w + ((size_t)tap * NO + o * 16 + col) * NI +
                                                 ib * 16 + grp * 4);
#pragma unroll
    for (int t = 0; t < kVT; ++t) {
      const int nd = vd[t] + dz, nh = vh[t] + dx, nw = vw[t] + dy;
      const bool ok = vok[t] && nd >= 0 && nd < D && nh >= 0 && nh < H && nw >= 0 && nw < W;
      const size_t nv = ok ? ((size_t)nd * H + nh) * W + nw : 0;
      const uint16_t* src = in + nv * NI + grp * 4;
#pragma unroll
      for (int ib = 0; ib < NIB; ++ib) {
        const v4s b = ok ? *reinterpret_cast<const v4s*>(src + ib * 16) : zero4;
#pragma unroll
        for (int o = 0; o < NOT; ++o)
          acc[o][t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[o][ib], b, acc[o][t], 0, 0, 0);
      }
    }
  }
#pragma unroll
  for (int t = 0; t < kVT; ++t) {
    if (!vok[t]) continue;
    const size_t v = v0 + t * 16 + col;
#pragma unroll
    for (int o = 0; o < NOT; ++o) {
      const int c0 = o * 16 + grp * 4;
      if constexpr (EPI == kEpiLast) {
        float4 r;
        r.x = acc[o][t][0] + bias[c0 + 0];
        r.y = acc[o][t][1] + bias[c0 + 1];
        r.z = acc[o][t][2] + bias[c0 + 2];
        r.w = acc[o][t][3] + bias[c0 + 3];
        *reinterpret_cast<float4*>(reinterpret_cast<float*>(out) + v * NO + c0) = r;
      } else {
        float r[4];
        if constexpr (EPI == kEpiHidden) {
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const float z = acc[o][t][i] + bias[c0 + i];
            r[i] = z > 0.f ? z : kNetSlope * z;
          }
        } else {
          const v4s av = *reinterpret_cast<const v4s*>(act + v * NO + c0);
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const float a = bf2f((uint16_t)av[i]);
            r[i] = a > 0.f ? acc[o][t][i] : kNetSlope * acc[o][t][i];
          }
        }
        v4s pk;
#pragma unroll
        for (int i = 0; i < 4; ++i) pk[i] = (short)f2bf(r[i]);
        *reinterpret_cast<v4s*>(reinterpret_cast<uint16_t*>(out) + v * NO + c0) = pk;
      }
    }
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
  for (size_t c0 = vbeg; c0 < vend; c0 += kWgVox) {
    for (int e = threadIdx.x; e < kWgVox * NO; e += kNetNT) {
      const int v = e / NO, c = e - v * NO;
      const size_t gv = c0 + v;
      sg[c * kWgLd + v] = gv < vend ? g[gv * NO + c] : (uint16_t)0;
    }
    for (int e = threadIdx.x; e < kWgVox * NI; e += kNetNT) {
      const int v = e / NI, c = e - v * NI;
      const size_t gv = c0 + v;
      uint16_t val = 0;
      if (gv < vend) {
        const int gw = (int)(gv % W);
        const size_t q = gv / W;
        const int nh = (int)(q % H) + dx, nd = (int)(q / H) + dz, nw = gw + dy;
        if (nd >= 0 && nd < D && nh >= 0 && nh < H && nw >= 0 && nw < W)
          val = in[(((size_t)nd * H + nh) * W + nw) * NI + c];
      }
      sx[c * kWgLd + v] = val;
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
template <int NO>
__global__ __launch_bounds__(kNetNT) void k_bias_grad(const uint16_t* __restrict__ g,
                                                      float* __restrict__ db, int cout,
                                                      size_t nvox) {
  __shared__ float s[NO];
  for (int c = threadIdx.x; c < NO; c += kNetNT) s[c] = 0.f;
  __syncthreads();
  const size_t tot = nvox * NO;
  for (size_t e = (size_t)blockIdx.x * kNetNT + threadIdx.x; e < tot;
       e += (size_t)gridDim.x * kNetNT)
    atomicAdd(&s[e % NO], bf2f(g[e]));
  __syncthreads();
  for (int c = threadIdx.x; c < cout; c += kNetNT) atomicAdd(&db[c], s[c]);
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
                                                       const unsigned int* __restrict__ tmax,
                                                       uint16_t* __restrict__ out, size_t nvox,
                                                       const int* __restrict__ done) {
  if (done != nullptr && *done) return;
  const float tm = __uint_as_float(*tmax);
  const float s0 = scale_of(scal);
  const float s = tm > 0.f ? s0 / tm : 0.f;
  for (size_t v = (size_t)blockIdx.x * kNetNT + threadIdx.x; v < nvox;
       v += (size_t)gridDim.x * kNetNT) {
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
    uint16_t pk[kNetIOPad];
#pragma unroll
    for (int c = 0; c < 9; ++c) {
      pk[2 * c] = f2bf(s * (float)y[c].x);
      pk[2 * c + 1] = f2bf(s * (float)y[c].y);
    }
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

// z = r + (out[.., 2c] + i out[.., 2c+1]) / s
template <typename R>
__global__ __launch_bounds__(kNetNT) void k_net_scatter(const cx<R>* __restrict__ r,
                                                        const float* __restrict__ out,
                                                        const double* __restrict__ scal,
                                                        cx<R>* __restrict__ z, size_t n,
                                                        const int* __restrict__ done) {
  if (done != nullptr && *done) return;
  const R inv = R(1) / R(scale_of(scal));
  for (size_t e = (size_t)blockIdx.x * kNetNT + threadIdx.x; e < n;
       e += (size_t)gridDim.x * kNetNT) {
    const size_t v = e / 9;
    const int c = (int)(e - v * 9);
    const float* o = out + v * kNetIOPad + 2 * c;
    const cx<R> a = r[e];
    z[e] = mk<R>(a.x + inv * R(o[0]), a.y + inv * R(o[1]));
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

static int conv_blocks(size_t nvox) {
  const size_t waves = (nvox + 16 * kVT - 1) / (16 * kVT);
  return (int)((waves + kNetNT / 64 - 1) / (kNetNT / 64));
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
  for (int l = 0; l < kNetLayers; ++l)
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

// The four-layer forward pass from act[0] (or `in`) to outf, per hidden width.
template <int HID>
struct NetOps {
  static void forward(Net& net, const uint16_t* in, hipStream_t s, const int* done) {
    const int B = conv_blocks(net.nvox);
    const int D = net.D, H = net.H, W = net.W;
    hipLaunchKernelGGL((k_conv<kNetIOPad, HID, kEpiHidden, false>), dim3(B), dim3(kNetNT), 0, s,
                       in, net.wpk + net.wpad_off[0], net.bpk + net.bpad_off[0], nullptr,
                       net.act[1], D, H, W, done);
    hipLaunchKernelGGL((k_conv<HID, HID, kEpiHidden, false>), dim3(B), dim3(kNetNT), 0, s,
                       net.act[1], net.wpk + net.wpad_off[1], net.bpk + net.bpad_off[1], nullptr,
                       net.act[2], D, H, W, done);
    hipLaunchKernelGGL((k_conv<HID, HID, kEpiHidden, false>), dim3(B), dim3(kNetNT), 0, s,
                       net.act[2], net.wpk + net.wpad_off[2], net.bpk + net.bpad_off[2], nullptr,
                       net.act[3], D, H, W, done);
    hipLaunchKernelGGL((k_conv<HID, kNetIOPad, kEpiLast, false>), dim3(B), dim3(kNetNT), 0, s,
                       net.act[3], net.wpk + net.wpad_off[3], net.bpk + net.bpad_off[3], nullptr,
                       net.outf, D, H, W, done);
  }

  // Backward of the loss gradient held in gbuf[0] ([v][kNetIOPad]); `in` is
  // the network input.  Accumulates into net.grads.
  static void backward(Net& net, const uint16_t* in, hipStream_t s) {
    const int B = conv_blocks(net.nvox);
    const int D = net.D, H = net.H, W = net.W;
    const size_t nv = net.nvox;
    const int vpb = 4096;
    const dim3 wg((unsigned)((nv + vpb - 1) / vpb), 27);
    const int gb = grid_of(nv * 64, 2048);
    uint16_t* g4 = net.gbuf[0];
    uint16_t* g3 = net.gbuf[1];
    uint16_t* g2 = net.gbuf[0];
    uint16_t* g1 = net.gbuf[1];
    float* G = net.grads;
    // layer 4 (HID -> IO)
    hipLaunchKernelGGL((k_conv_wgrad<HID, kNetIOPad>), wg, dim3(kNetNT), 0, s, net.act[3], g4,
                       G + net.woff[3], net.cin[3], net.cout[3], D, H, W, vpb);
    hipLaunchKernelGGL((k_bias_grad<kNetIOPad>), dim3(gb), dim3(kNetNT), 0, s, g4,
                       G + net.boff[3], net.cout[3], nv);
    hipLaunchKernelGGL((k_conv<kNetIOPad, HID, kEpiBackward, true>), dim3(B), dim3(kNetNT), 0, s,
                       g4, net.wtr + net.wpad_off[3], nullptr, net.act[3], g3, D, H, W, nullptr);
    // layer 3
    hipLaunchKernelGGL((k_conv_wgrad<HID, HID>), wg, dim3(kNetNT), 0, s, net.act[2], g3,
                       G + net.woff[2], net.cin[2], net.cout[2], D, H, W, vpb);
    hipLaunchKernelGGL((k_bias_grad<HID>), dim3(gb), dim3(kNetNT), 0, s, g3, G + net.boff[2],
                       net.cout[2], nv);
    hipLaunchKernelGGL((k_conv<HID, HID, kEpiBackward, true>), dim3(B), dim3(kNetNT), 0, s, g3,
                       net.wtr + net.wpad_off[2], nullptr, net.act[2], g2, D, H, W, nullptr);
    // layer 2
    hipLaunchKernelGGL((k_conv_wgrad<HID, HID>), wg, dim3(kNetNT), 0, s, net.act[1], g2,
                       G + net.woff[1], net.cin[1], net.cout[1], D, H, W, vpb);
    hipLaunchKernelGGL((k_bias_grad<HID>), dim3(gb), dim3(kNetNT), 0, s, g2, G + net.boff[1],
                       net.cout[1], nv);
    hipLaunchKernelGGL((k_conv<HID, HID, kEpiBackward, true>), dim3(B), dim3(kNetNT), 0, s, g2,
                       net.wtr + net.wpad_off[1], nullptr, net.act[1], g1, D, H, W, nullptr);
    // layer 1 (IO -> HID): weights and bias only
    hipLaunchKernelGGL((k_conv_wgrad<kNetIOPad, HID>), wg, dim3(kNetNT), 0, s, in, g1,
                       G + net.woff[0], net.cin[0], net.cout[0], D, H, W, vpb);
    hipLaunchKernelGGL((k_bias_grad<HID>), dim3(gb), dim3(kNetNT), 0, s, g1, G + net.boff[0],
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

template <typename R>
int net_input(const Plan& p, const cx<R>* r, double scale, double* scal, uint16_t* out,
              hipStream_t s, const int* done) {
  MSC_REQUIRE(p.have_tparams, "T0 is not set on this plan");
  if (scale > 0.0) {
    hipLaunchKernelGGL(k_set_scale, dim3(1), dim3(64), 0, s, scal, scale * scale, 1.0);
  } else {
    hipLaunchKernelGGL(k_set_scale, dim3(1), dim3(64), 0, s, scal, (double)p.n, 0.0);
    hipLaunchKernelGGL((k_norm_scale<R>), dim3(grid_of(p.n, 2048)), dim3(kNetNT), 0, s, r, p.n,
                       scal);
  }
  const cx<R>* tp;
  if constexpr (sizeof(R) == 8) {
    tp = p.tp128;
  } else {
    tp = p.tp64;
  }
  hipLaunchKernelGGL((k_net_gather<R>), dim3(grid_of(p.nvox)), dim3(kNetNT), 0, s, r, tp, scal,
                     p.tmax, out, p.nvox, done);
  MSC_LAUNCH_CHECK();
  return kOk;
}

template <typename R>
int net_apply(Net& net, const Plan& p, const cx<R>* r, cx<R>* z, double scale, hipStream_t s,
              const int* done) {
  MSC_REQUIRE(net.packed, "network parameters are not set");
  int rc = net.ensure_grid(p.Nz, p.M, p.M);
  if (rc) return rc;
  rc = net_input<R>(p, r, scale, net.scal, net.act[0], s, done);
  if (rc) return rc;
  by_width<Fwd>(net.hidden, net, net.act[0], s, done);
  hipLaunchKernelGGL((k_net_scatter<R>), dim3(grid_of(p.n)), dim3(kNetNT), 0, s, r, net.outf,
                     net.scal, z, p.n, done);
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
