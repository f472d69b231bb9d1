// FFT-accelerated Foldy-Lax operator y = (I - G0 T0) x on gfx950.
//
// Replaces _slab_matvec_gpu (cubic_scattering/slab_scattering_gpu.py:95-155)
// and _slab_matvec (cubic_scattering/slab_scattering.py:359-418).  The 3-D
// linear convolution of tau = T0 x with the spatial propagator is done on a
// power-of-two grid (Pz, Pxy, Pxy) >= (2N_z-1, 2M-1, 2M-1) with the kernel
// stored circularly (offset d at index d mod P), so the valid output window
// is [0, N_z) x [0, M) x [0, M).  Five LDS-staged passes, each one HBM
// round trip, with the zero padding never read or written:
//   1. y-forward  : T0 gather + zero pad + FFT along y      (N_z*M rows)
//   2. x-forward  : FFT along x of the M non-zero rows       (N_z planes)
//   3. z-multiply : FFT along z, 9x9 block multiply with the compressed
//                   table, inverse FFT along z, keep z < N_z (Pxy^2 columns)
//   4. x-inverse  : inverse FFT along x, keep x < M
//   5. y-inverse  : inverse FFT along y, keep y < M, y = x - acc (or the
//                   residual b - x + acc)
#include <algorithm>
#include <cstdint>

#include "msc_fft.hpp"
#include "msc_plan.hpp"

namespace msc {

constexpr int kNT = 256;  // threads per workgroup for every pass

// Element-wise staging loop over n items, thread-strided: the global loads of
// SB consecutive items of a thread are issued before any of them is consumed,
// so a thread waits for memory once per SB items (a plain load -> store loop
// waits once per item: the compiler does not pipeline a loop whose trip count
// depends on the thread index).
constexpr int kSB = 8;
template <typename T, typename Load, typename Use>
__device__ __forceinline__ void staged(int n, Load load, Use use) {
  for (int w0 = 0; w0 < n; w0 += kSB * kNT) {
    T v[kSB];
#pragma unroll
    for (int k = 0; k < kSB; ++k) {
      const int w = w0 + k * kNT + (int)threadIdx.x;
      if (w < n) v[k] = load(w);
    }
#pragma unroll
    for (int k = 0; k < kSB; ++k) {
      const int w = w0 + k * kNT + (int)threadIdx.x;
      if (w < n) use(w, v[k]);
    }
  }
}


template <typename R>
__device__ __forceinline__ cx<R> flip_if(cx<R> v, unsigned mask, int c) {
  return ((mask >> c) & 1u) ? mk<R>(-v.x, -v.y) : v;
}

// x<->y swap of the 9-vector: u_x<->u_y, e_xx<->e_yy, e_zy<->e_zx.
template <typename R>
__device__ __forceinline__ void swap_xy(cx<R>* t) {
  cx<R> q;
  q = t[1]; t[1] = t[2]; t[2] = q;
  q = t[4]; t[4] = t[5]; t[5] = q;
  q = t[7]; t[7] = t[8]; t[8] = q;
}

// o_p = K t_p for NP vectors sharing one canonical spectral propagator
// K = [[G, C], [H, S]], given by its 45 independent entries kp[e * stride]:
// G symmetric (6), C 3x6 row-major (18), the upper triangle of S' = S W^-1 (21),
// with H = W^-1 C^T, S = S' W and W = diag(1,1,1,1/2,1/2,1/2).  The entries are
// read one at a time and applied at once to every product.  t_p are already
// sign-flipped / swapped into the table's octant.
// Table loads kept in flight by the block multiply: every KG entries the
// accumulators are pinned, so later entries' loads cannot be hoisted above
// (the compiler otherwise issues all 45 at once: ~180 VGPRs in complex128).
#ifndef MSC_KMUL_GROUP
#define MSC_KMUL_GROUP 9
#endif
template <typename R, int NP>
__device__ __forceinline__ void kmul_pin(cx<R> (&o)[NP][9]) {
#pragma unroll
  for (int p = 0; p < NP; ++p)
#pragma unroll
    for (int i = 0; i < 9; ++i) asm volatile("" : "+v"(o[p][i].x), "+v"(o[p][i].y) : : "memory");
}

template <typename R, int NP>
__device__ __forceinline__ void kmul45_stream(const cx<R>* __restrict__ kp, int stride,
                                              const cx<R> (&t)[NP][9], cx<R> (&o)[NP][9]) {
  constexpr int KG = MSC_KMUL_GROUP;
#pragma unroll
  for (int p = 0; p < NP; ++p)
#pragma unroll
    for (int i = 0; i < 9; ++i) o[p][i] = czero<R>();
  // G (symmetric 3x3): entries (0,0) (0,1) (0,2) (1,1) (1,2) (2,2)
#pragma unroll
  for (int e = 0; e < 6; ++e) {
    const int i = e < 3 ? 0 : (e < 5 ? 1 : 2);
    const int j = e < 3 ? e : (e < 5 ? e - 2 : 2);
    const cx<R> kv = kp[e * stride];
#pragma unroll
    for (int p = 0; p < NP; ++p) {
      cfma<R>(o[p][i], kv, t[p][j]);
      if (i != j) cfma<R>(o[p][j], kv, t[p][i]);
    }
    if (KG > 0 && e % KG == KG - 1) kmul_pin<R, NP>(o);
  }
  // C (3x6) and H = W^-1 C^T: o_u += C t_e, o_e += (1 or 2) C^T t_u
#pragma unroll
  for (int e = 0; e < 18; ++e) {
    const int i = e / 6, al = e % 6;
    const cx<R> kv = kp[(6 + e) * stride];
#pragma unroll
    for (int p = 0; p < NP; ++p) {
      cfma<R>(o[p][i], kv, t[p][3 + al]);
      cfma<R>(o[p][3 + al], al < 3 ? kv : cscale<R>(kv, R(2)), t[p][i]);
    }
    if (KG > 0 && (6 + e) % KG == KG - 1) kmul_pin<R, NP>(o);
  }
  // S' (symmetric 6x6 upper triangle), S = S' W: o_e += S' (W t_e)
#pragma unroll
  for (int e = 0; e < 21; ++e) {
    int al = 0, rem = e;
#pragma unroll
    for (int k = 0; k < 6; ++k)
      if (rem >= 6 - k && al == k) {
        rem -= 6 - k;
        al = k + 1;
      }
    const int be = al + rem;
    const cx<R> kv = kp[(24 + e) * stride];
#pragma unroll
    for (int p = 0; p < NP; ++p) {
      const cx<R> twb = be < 3 ? t[p][3 + be] : cscale<R>(t[p][3 + be], R(0.5));
      cfma<R>(o[p][3 + al], kv, twb);
      if (be != al) {
        const cx<R> twa = al < 3 ? t[p][3 + al] : cscale<R>(t[p][3 + al], R(0.5));
        cfma<R>(o[p][3 + be], kv, twa);
      }
    }
    if (KG > 0 && (24 + e) % KG == KG - 1) kmul_pin<R, NP>(o);
  }
}

// Dispatch slot b runs on XCD b % 8: give each XCD a contiguous run of
// logical blocks, so the column segments that neighbouring blocks of the x
// passes share (rows of 9Q complex, not whole cache lines) sit in one L2
// (matvec 0.384-0.406 -> 0.365 ms in complex64, 0.699 -> 0.654 ms in
// complex128 at 128x128x32).
__device__ __forceinline__ int xcd_block(int b, int nblk) {
  if (nblk % 8 != 0) return b;
  return (b % 8) * (nblk / 8) + b / 8;
}

// ── pass 1: T0 gather + zero pad + forward FFT along y ─────────────────────
// GEN: voxel v applies the general 9x9 matrix tgen[tidx[v]] (row-major)
// instead of the 4-parameter cubic T0.
template <typename R, int P, int B, bool GEN>
__device__ __forceinline__ void fwd_y_body(unsigned bx_, unsigned gx_, const cx<R>* __restrict__ x,
                                              const cx<R>* __restrict__ tp,
                                              const int* __restrict__ tidx,
                                              const cx<R>* __restrict__ tgen,
                                              cx<R>* __restrict__ buf1, int M, int nrows,
                                              const cx<R>* __restrict__ tw,
                                              const int* __restrict__ done) {
  constexpr int NTR = 9 * B;
  constexpr int LD = lds_ld<NTR>::value;
  __shared__ cx<R> s[P * LD];
  if (done != nullptr && *done) return;
  const int row0 = bx_ * B;
  staged<cx<R>>(
      B * P * 9,
      [&](int w) {
        const int b = w / (P * 9);
        const int rem = w - b * P * 9;
        const int y = rem / 9;
        const int row = row0 + b;
        const bool ok = y < M && row < nrows;
        const cx<R> v = x[ok ? (size_t)row * M * 9 + rem : 0];
        return ok ? v : czero<R>();
      },
      [&](int w, cx<R> v) {
        const int b = w / (P * 9);
        const int rem = w - b * P * 9;
        const int y = rem / 9;
        const int c = rem - y * 9;
        s[y * LD + b * 9 + c] = v;
      });
  __syncthreads();
  if constexpr (GEN) {
    for (int w = threadIdx.x; w < B * M; w += kNT) {
      const int b = w / M;
      const int y = w - b * M;
      const int row = row0 + b;
      if (row >= nrows) continue;
      const cx<R>* T = tgen + (size_t)tidx[(size_t)row * M + y] * 81;
      cx<R>* v = s + y * LD + b * 9;
      cx<R> in[9];
#pragma unroll
      for (int c = 0; c < 9; ++c) in[c] = v[c];
#pragma unroll
      for (int i = 0; i < 9; ++i) {
        cx<R> acc = czero<R>();
#pragma unroll
        for (int j = 0; j < 9; ++j) acc = cadd<R>(acc, cmul<R>(T[i * 9 + j], in[j]));
        v[i] = acc;
      }
    }
  } else {
  struct T4 {
    cx<R> u, d, o, s;
  };
  staged<T4>(
      B * M,
      [&](int w) {
        const int b = w / M;
        const int y = w - b * M;
        const int row = min(row0 + b, nrows - 1);
        const cx<R>* t = tp + ((size_t)row * M + y) * 4;
        return T4{t[0], t[1], t[2], t[3]};
      },
      [&](int w, T4 tt) {
    const int b = w / M;
    const int y = w - b * M;
    if (row0 + b >= nrows) return;
    const cx<R> tu = tt.u, td = tt.d, to = tt.o, ts = tt.s;
    cx<R>* v = s + y * LD + b * 9;
    const cx<R> sum = cadd<R>(cadd<R>(v[3], v[4]), v[5]);
    const cx<R> dmo = csub<R>(td, to);
    const cx<R> tos = cmul<R>(to, sum);
#pragma unroll
    for (int c = 0; c < 3; ++c) v[c] = cmul<R>(tu, v[c]);
#pragma unroll
    for (int c = 3; c < 6; ++c) v[c] = cadd<R>(cmul<R>(dmo, v[c]), tos);
#pragma unroll
    for (int c = 6; c < 9; ++c) v[c] = cmul<R>(ts, v[c]);
      });
  }
  __syncthreads();
  lds_fft<R, P, false, NTR, LD, kNT>(s, tw);
  for (int w = threadIdx.x; w < B * P * 9; w += kNT) {
    const int b = w / (P * 9);
    const int rem = w - b * P * 9;
    const int y = rem / 9;
    const int c = rem - y * 9;
    const int row = row0 + b;
    if (row < nrows) buf1[(size_t)row * P * 9 + rem] = s[y * LD + b * 9 + c];
  }
}
template <typename R, int P, int B, bool GEN = false>
__global__ __launch_bounds__(kNT) void k_fwd_y(const cx<R>* __restrict__ x,
                                              const cx<R>* __restrict__ tp,
                                              const int* __restrict__ tidx,
                                              const cx<R>* __restrict__ tgen,
                                              cx<R>* __restrict__ buf1, int M, int nrows,
                                              const cx<R>* __restrict__ tw,
                                              const int* __restrict__ done) {
  fwd_y_body<R, P, B, GEN>(blockIdx.x, gridDim.x, x, tp, tidx, tgen, buf1, M, nrows, tw, done);
}


// ── pass 2: forward FFT along x of the M non-zero rows ─────────────────────
// A tile is NTR consecutive complex numbers of every x row of one z plane (a
// row holds P * 9: ky-major, component-minor); the NTR transforms along x are
// independent.  NTR = 9Q covers Q whole ky columns; NTR = 8, 16 or 32 covers
// whole 128-byte cache lines of each row and fills every lane of the LDS
// transform (NTR * P / radix is a multiple of the workgroup size).
template <typename R, int P, int NTR>
__device__ __forceinline__ void fwd_x_body(unsigned bx_, unsigned gx_, const cx<R>* __restrict__ buf1,
                                              cx<R>* __restrict__ buf2, int M,
                                              const cx<R>* __restrict__ tw,
                                              const int* __restrict__ done) {
  static_assert((P * 9) % NTR == 0, "x tile must divide the row");
  constexpr int LD = lds_ld<NTR>::value;
  __shared__ cx<R> s[P * LD];
  if (done != nullptr && *done) return;
  constexpr int nck = P * 9 / NTR;
  const int lb = xcd_block((int)bx_, (int)gx_);
  const int z = lb / nck;
  const int e0 = (lb - z * nck) * NTR;
  staged<cx<R>>(
      P * NTR,
      [&](int w) {
        const int xr = w / NTR;
        const int e = w - xr * NTR;
        const bool ok = xr < M;
        const cx<R> v = buf1[ok ? ((size_t)z * M + xr) * P * 9 + e0 + e : 0];
        return ok ? v : czero<R>();
      },
      [&](int w, cx<R> v) {
        const int xr = w / NTR;
        s[xr * LD + (w - xr * NTR)] = v;
      });
  __syncthreads();
  lds_fft<R, P, false, NTR, LD, kNT>(s, tw);
  for (int w = threadIdx.x; w < P * NTR; w += kNT) {
    const int kx = w / NTR;
    const int e = w - kx * NTR;
    buf2[((size_t)z * P + kx) * P * 9 + e0 + e] = s[kx * LD + e];
  }
}
template <typename R, int P, int NTR>
__global__ __launch_bounds__(kNT) void k_fwd_x(const cx<R>* __restrict__ buf1,
                                              cx<R>* __restrict__ buf2, int M,
                                              const cx<R>* __restrict__ tw,
                                              const int* __restrict__ done) {
  fwd_x_body<R, P, NTR>(blockIdx.x, gridDim.x, buf1, buf2, M, tw, done);
}


// ── pass 3: z FFT, spectral 9x9 multiply, inverse z FFT (in place) ─────────
// Block multiply of one z-pass tile held in LDS (Q columns ky0.. of row kx,
// PZ wavenumbers kz each), one (column, kz) per item: the table entries are
// streamed (each loaded and applied at once) rather than held in registers.
template <typename R, int PZ, int Q, int LD>
__device__ __forceinline__ void zmul_items(cx<R>* s, const cx<R>* __restrict__ table, int kx,
                                           int ky0, int Pxy) {
  constexpr int NC = PZ / 2 + 1;
  const int half = Pxy / 2;
  constexpr int HZ = PZ / 2;
  for (int w = threadIdx.x; w < PZ * Q; w += kNT) {
    const int q = w / PZ;
    const int kz = w - q * PZ;
    const int ky = ky0 + q;
    const int a = kx <= half ? kx : Pxy - kx;
    const int b = ky <= half ? ky : Pxy - ky;
    const int c = kz <= HZ ? kz : PZ - kz;
    const unsigned flip = (kz > HZ ? kFlipZ : 0u) ^ (kx > half ? kFlipX : 0u) ^
                          (ky > half ? kFlipY : 0u);
    const bool sw = a < b;
    const int u = sw ? b : a;
    const int vv = sw ? a : b;
    const cx<R>* kp = table + (size_t)((u * (u + 1)) / 2 + vv) * kNE * NC + c;
    cx<R>* sp = s + kz * LD + q * 9;
    cx<R> t[1][9], o[1][9];
#pragma unroll
    for (int i = 0; i < 9; ++i) t[0][i] = flip_if<R>(sp[i], flip, i);
    if (sw) swap_xy<R>(t[0]);
    kmul45_stream<R, 1>(kp, NC, t, o);
    if (sw) swap_xy<R>(o[0]);
#pragma unroll
    for (int i = 0; i < 9; ++i) sp[i] = flip_if<R>(o[0][i], flip, i);
  }
}
template <typename R, int PZ, int Q>
__device__ __forceinline__ void z_mult_body(unsigned bx_, unsigned gx_, cx<R>* __restrict__ buf2,
                                               const cx<R>* __restrict__ table, int Nz,
                                               int Pxy, const cx<R>* __restrict__ tw,
                                               const int* __restrict__ zmap,
                                               const int* __restrict__ done) {
  constexpr int NTR = 9 * Q;
  constexpr int LD = lds_ld<NTR>::value;
  __shared__ cx<R> s[PZ * LD];
  if (done != nullptr && *done) return;
  const int nck = Pxy / Q;
  const int tile = zmap[bx_];
  const int kx = tile / nck;
  const int ky0 = (tile - kx * nck) * Q;
  const size_t plane = (size_t)Pxy * Pxy * 9;
  const size_t col0 = ((size_t)kx * Pxy + ky0) * 9;
  for (int w = threadIdx.x; w < PZ * NTR; w += kNT) {
    const int z = w / NTR;
    const int e = w - z * NTR;
    const bool ok = z < Nz;
    const cx<R> v = buf2[ok ? (size_t)z * plane + col0 + e : col0];
    s[z * LD + e] = ok ? v : czero<R>();
  }
  __syncthreads();
  lds_fft<R, PZ, false, NTR, LD, kNT>(s, tw);
  zmul_items<R, PZ, Q, LD>(s, table, kx, ky0, Pxy);
  __syncthreads();
  lds_fft<R, PZ, true, NTR, LD, kNT>(s, tw);
  for (int w = threadIdx.x; w < Nz * NTR; w += kNT) {
    const int z = w / NTR;
    const int e = w - z * NTR;
    buf2[(size_t)z * plane + col0 + e] = s[z * LD + e];
  }
}

// Pass 3 with pruned z FFTs and no staging round trips through LDS: the
// forward transform's first pass reads its non-zero input rows (z < Nz)
// straight from HBM into registers, and the inverse transform's last pass
// writes the kept rows (z < Nz) straight back; the zero rows are neither
// read nor transformed, the rows z >= Nz of the result never formed.  Same
// tiles, tile order and block multiply as z_mult_body.
template <typename R, int PZ, int Q>
__device__ __forceinline__ void z_mult_pruned_body(unsigned bx_, cx<R>* __restrict__ buf2,
                                                   const cx<R>* __restrict__ table, int Nz,
                                                   int Pxy, const cx<R>* __restrict__ tw,
                                                   const int* __restrict__ zmap,
                                                   const int* __restrict__ done) {
  constexpr int NTR = 9 * Q;
  constexpr int LD = lds_ld<NTR>::value;
  __shared__ cx<R> s[PZ * LD];
  if (done != nullptr && *done) return;
  const int nck = Pxy / Q;
  const int tile = zmap[bx_];
  const int kx = tile / nck;
  const int ky0 = (tile - kx * nck) * Q;
  const size_t plane = (size_t)Pxy * Pxy * 9;
  cx<R>* col = buf2 + ((size_t)kx * Pxy + ky0) * 9;
  first_pass_pruned_global<R, PZ, NTR, LD, kNT>(col, plane, s, Nz);
  __syncthreads();
  fft_stages<R, PZ, pick_radix<PZ>::value, false, NTR, LD, kNT>(s, tw);
  zmul_items<R, PZ, Q, LD>(s, table, kx, ky0, Pxy);
  __syncthreads();
  fft_stages_but_last<R, PZ, 1, true, NTR, LD, kNT>(s, tw);
  last_pass_inv_store<R, PZ, NTR, LD, kNT>(
      s, tw, Nz, [&](int row, int c, cx<R> v) { col[(size_t)row * plane + c] = v; });
}
template <typename R, int PZ, int Q>
__global__ __launch_bounds__(kNT) void k_z_mult_pr(cx<R>* __restrict__ buf2,
                                                  const cx<R>* __restrict__ table, int Nz,
                                                  int Pxy, const cx<R>* __restrict__ tw,
                                                  const int* __restrict__ zmap,
                                                  const int* __restrict__ done) {
  z_mult_pruned_body<R, PZ, Q>(blockIdx.x, buf2, table, Nz, Pxy, tw, zmap, done);
}

template <typename R, int PZ, int Q>
__global__ __launch_bounds__(kNT) void k_z_mult(cx<R>* __restrict__ buf2,
                                               const cx<R>* __restrict__ table, int Nz,
                                               int Pxy, const cx<R>* __restrict__ tw,
                                               const int* __restrict__ zmap,
                                               const int* __restrict__ done) {
  z_mult_body<R, PZ, Q>(blockIdx.x, gridDim.x, buf2, table, Nz, Pxy, tw, zmap, done);
}


// Direct-to-LDS copy of the Nz input rows of one z-pass tile,
// dst[z * NTR + e] = src[z * plane + e] (z < Nz, e < NTR), in 16-byte chunks
// (global_load_lds_dwordx4): wave w issues chunks w*64 .. w*64+63, then every
// 4th group of 64; lane l of a group lands at the group's LDS base + 16 l.
// Nothing passes through VGPRs; the data is in LDS once the issuing wave's
// vmcnt drains.  Needs NTR * sizeof(cx<R>) a multiple of 16 and 16-byte
// aligned sources.
template <typename R, int NTR>
__device__ __forceinline__ void dma_rows(cx<R>* dst, const cx<R>* src, size_t plane, int Nz) {
  constexpr int EPC = 16 / (int)sizeof(cx<R>);  // complex per 16-byte chunk
  static_assert(NTR % EPC == 0, "a chunk must not straddle two rows");
  typedef __attribute__((address_space(3))) void* lds_ptr;
  typedef __attribute__((address_space(1))) void* gbl_ptr;
  const int nch = Nz * (NTR / EPC);
  const int lane = (int)threadIdx.x & 63;
  for (int c0 = ((int)threadIdx.x >> 6) * 64; c0 < nch; c0 += kNT) {
    const int c = c0 + lane;
    if (c < nch) {
      const int i = c * EPC;
      const int z = i / NTR;
      const int e = i - z * NTR;
      __builtin_amdgcn_global_load_lds((gbl_ptr)(src + (size_t)z * plane + e),
                                       (lds_ptr)(dst + (size_t)c0 * EPC), 16, 0, 0);
    }
  }
}

// Pass 3, persistent form: each workgroup walks the tiles blockIdx.x,
// blockIdx.x + gridDim.x, ... (gridDim.x a multiple of 8, so every tile keeps
// the XCD its zmap slot names).  The next tile's Nz input rows stream into
// a separate LDS buffer by direct-to-LDS loads while the current tile runs
// its FFTs and block multiply.  The z FFTs are pruned: the forward's first
// pass reads only the Nz non-zero rows (out of place, from the prefetch
// buffer), and the inverse's last pass writes back only the rows z < Nz.
template <typename R, int PZ, int Q>
__global__ __launch_bounds__(kNT) void k_z_mult_dma(cx<R>* __restrict__ buf2,
                                                   const cx<R>* __restrict__ table, int Nz,
                                                   int Pxy, const cx<R>* __restrict__ tw,
                                                   const int* __restrict__ zmap, int ntiles,
                                                   const int* __restrict__ done) {
  constexpr int NTR = 9 * Q;
  constexpr int LD = lds_ld<NTR>::value;
  constexpr int NC = PZ / 2 + 1;
  constexpr int HZ = PZ / 2;
  __shared__ cx<R> s[PZ * LD];
  __shared__ cx<R> in[HZ * NTR];
  if (done != nullptr && *done) return;
  int t = blockIdx.x;
  if (t >= ntiles) return;
  const int nck = Pxy / Q;
  const size_t plane = (size_t)Pxy * Pxy * 9;
  const int half = Pxy / 2;
  auto col_of = [&](int tile) {
    const int kx = tile / nck;
    return ((size_t)kx * Pxy + (tile - kx * nck) * Q) * 9;
  };
  dma_rows<R, NTR>(in, buf2 + col_of(zmap[t]), plane, Nz);
  for (; t < ntiles; t += gridDim.x) {
    const int tile = zmap[t];
    const int kx = tile / nck;
    const int ky0 = (tile - kx * nck) * Q;
    const size_t col0 = ((size_t)kx * Pxy + ky0) * 9;
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();  // this tile's rows are in `in`; the last tile's reads of s are done
    first_pass_pruned<R, PZ, NTR, LD, kNT>(in, s, Nz);
    __syncthreads();  // `in` is free
    if (t + (int)gridDim.x < ntiles)
      dma_rows<R, NTR>(in, buf2 + col_of(zmap[t + gridDim.x]), plane, Nz);
    fft_stages<R, PZ, pick_radix<PZ>::value, false, NTR, LD, kNT>(s, tw);
    for (int w = threadIdx.x; w < PZ * Q; w += kNT) {
      const int q = w / PZ;
      const int kz = w - q * PZ;
      const int ky = ky0 + q;
      const int a = kx <= half ? kx : Pxy - kx;
      const int b = ky <= half ? ky : Pxy - ky;
      const int c = kz <= HZ ? kz : PZ - kz;
      const unsigned flip = (kz > HZ ? kFlipZ : 0u) ^ (kx > half ? kFlipX : 0u) ^
                            (ky > half ? kFlipY : 0u);
      const bool sw = a < b;
      const int u = sw ? b : a;
      const int vv = sw ? a : b;
      const cx<R>* kp = table + (size_t)((u * (u + 1)) / 2 + vv) * kNE * NC + c;
      cx<R>* sp = s + kz * LD + q * 9;
      cx<R> tv[1][9], o[1][9];
#pragma unroll
      for (int i = 0; i < 9; ++i) tv[0][i] = flip_if<R>(sp[i], flip, i);
      if (sw) swap_xy<R>(tv[0]);
      kmul45_stream<R, 1>(kp, NC, tv, o);
      if (sw) swap_xy<R>(o[0]);
#pragma unroll
      for (int i = 0; i < 9; ++i) sp[i] = flip_if<R>(o[0][i], flip, i);
    }
    __syncthreads();
    fft_stages_but_last<R, PZ, 1, true, NTR, LD, kNT>(s, tw);
    // the last pass reads and writes the same rows of its own column, so it
    // runs in place without a barrier between its reads and writes
    last_pass_inv_store<R, PZ, NTR, LD, kNT>(
        s, tw, Nz, [&](int z, int e, cx<R> v) { s[z * LD + e] = v; });
    __syncthreads();
    for (int w = threadIdx.x; w < Nz * NTR; w += kNT) {
      const int z = w / NTR;
      const int e = w - z * NTR;
      buf2[(size_t)z * plane + col0 + e] = s[z * LD + e];
    }
  }
}

// ── pass 4: inverse FFT along x, keep x < M ────────────────────────────────
// Tiles as in pass 2.
template <typename R, int P, int NTR>
__device__ __forceinline__ void inv_x_body(unsigned bx_, unsigned gx_, const cx<R>* __restrict__ buf2,
                                              cx<R>* __restrict__ buf1, int M,
                                              const cx<R>* __restrict__ tw,
                                              const int* __restrict__ done) {
  static_assert((P * 9) % NTR == 0, "x tile must divide the row");
  constexpr int LD = lds_ld<NTR>::value;
  __shared__ cx<R> s[P * LD];
  if (done != nullptr && *done) return;
  constexpr int nck = P * 9 / NTR;
  const int lb = xcd_block((int)bx_, (int)gx_);
  const int z = lb / nck;
  const int e0 = (lb - z * nck) * NTR;
  staged<cx<R>>(
      P * NTR,
      [&](int w) {
        const int kx = w / NTR;
        return buf2[((size_t)z * P + kx) * P * 9 + e0 + (w - kx * NTR)];
      },
      [&](int w, cx<R> v) {
        const int kx = w / NTR;
        s[kx * LD + (w - kx * NTR)] = v;
      });
  __syncthreads();
  lds_fft<R, P, true, NTR, LD, kNT>(s, tw);
  for (int w = threadIdx.x; w < M * NTR; w += kNT) {
    const int xr = w / NTR;
    const int e = w - xr * NTR;
    buf1[((size_t)z * M + xr) * P * 9 + e0 + e] = s[xr * LD + e];
  }
}
template <typename R, int P, int NTR>
__global__ __launch_bounds__(kNT) void k_inv_x(const cx<R>* __restrict__ buf2,
                                              cx<R>* __restrict__ buf1, int M,
                                              const cx<R>* __restrict__ tw,
                                              const int* __restrict__ done) {
  inv_x_body<R, P, NTR>(blockIdx.x, gridDim.x, buf2, buf1, M, tw, done);
}


// ── pass 5: inverse FFT along y, keep y < M and form the output ────────────
// MODE 0: out = x - acc (the operator), 1: out = b - x + acc (residual),
// 2: out = acc (the scattered part G0 T0 x).
template <typename R, int P, int B, int MODE>
__device__ __forceinline__ void inv_y_body(unsigned bx_, unsigned gx_, const cx<R>* __restrict__ buf1,
                                              const cx<R>* __restrict__ x,
                                              const cx<R>* __restrict__ bvec,
                                              cx<R>* __restrict__ out, int M, int nrows,
                                              const cx<R>* __restrict__ tw,
                                              const int* __restrict__ done) {
  constexpr int NTR = 9 * B;
  constexpr int LD = lds_ld<NTR>::value;
  __shared__ cx<R> s[P * LD];
  if (done != nullptr && *done) return;
  const int row0 = bx_ * B;
  staged<cx<R>>(
      B * P * 9,
      [&](int w) {
        const int b = w / (P * 9);
        const int rem = w - b * P * 9;
        const int row = row0 + b;
        const bool ok = row < nrows;
        const cx<R> v = buf1[ok ? (size_t)row * P * 9 + rem : 0];
        return ok ? v : czero<R>();
      },
      [&](int w, cx<R> v) {
        const int b = w / (P * 9);
        const int rem = w - b * P * 9;
        const int y = rem / 9;
        s[y * LD + b * 9 + (rem - y * 9)] = v;
      });
  __syncthreads();
  lds_fft<R, P, true, NTR, LD, kNT>(s, tw);
  // the operator's own x (and b) terms: loaded in batches as above
  staged<cx<R>>(
      B * M * 9,
      [&](int w) {
        const int b = w / (M * 9);
        const int rem = w - b * M * 9;
        const int row = min(row0 + b, nrows - 1);
        const size_t g = (size_t)row * M * 9 + rem;
        if constexpr (MODE == 1) return csub<R>(bvec[g], x[g]);
        else if constexpr (MODE == 2) return czero<R>();
        else return x[g];
      },
      [&](int w, cx<R> xv) {
        const int b = w / (M * 9);
        const int rem = w - b * M * 9;
        const int y = rem / 9;
        const int row = row0 + b;
        if (row >= nrows) return;
        const size_t g = (size_t)row * M * 9 + rem;
        const cx<R> acc = s[y * LD + b * 9 + (rem - y * 9)];
        if constexpr (MODE == 2) out[g] = acc;
        else if constexpr (MODE == 1) out[g] = cadd<R>(xv, acc);
        else out[g] = csub<R>(xv, acc);
      });
}
template <typename R, int P, int B, int MODE>
__global__ __launch_bounds__(kNT) void k_inv_y(const cx<R>* __restrict__ buf1,
                                              const cx<R>* __restrict__ x,
                                              const cx<R>* __restrict__ bvec,
                                              cx<R>* __restrict__ out, int M, int nrows,
                                              const cx<R>* __restrict__ tw,
                                              const int* __restrict__ done) {
  inv_y_body<R, P, B, MODE>(blockIdx.x, gridDim.x, buf1, x, bvec, out, M, nrows, tw, done);
}


// ── volume-averaged nearest-neighbour correction (after pass 5) ──────────────
// acc_v = sum over the 26 offsets R of nn[R] tau(v - R), tau = T0 x, with
// periodic lateral wrap when asked; MODE 0: out -= acc, 1 and 2: out += acc
// (the operator, the residual and the scattered part).  The 26 x 81 blocks
// are staged in LDS.
template <typename R, int MODE>
__global__ __launch_bounds__(kNT) void k_nn_apply(const cx<R>* __restrict__ x,
                                                 const cx<R>* __restrict__ tp,
                                                 const cx<R>* __restrict__ nn,
                                                 cx<R>* __restrict__ out, int M, int Nz,
                                                 int periodic, const int* __restrict__ done) {
  __shared__ cx<R> sd[26 * 81];
  if (done != nullptr && *done) return;
  for (int i = threadIdx.x; i < 26 * 81; i += kNT) sd[i] = nn[i];
  __syncthreads();
  const size_t nvox = (size_t)Nz * M * M;
  for (size_t v = (size_t)blockIdx.x * kNT + threadIdx.x; v < nvox;
       v += (size_t)gridDim.x * kNT) {
    const int y = (int)(v % M);
    const size_t q = v / M;
    const int xr = (int)(q % M);
    const int z = (int)(q / M);
    cx<R> acc[9];
#pragma unroll
    for (int c = 0; c < 9; ++c) acc[c] = czero<R>();
    int slot = 0;
    for (int r = 0; r < 27; ++r) {
      if (r == 13) continue;
      const int dz = r / 9 - 1, dx = (r / 3) % 3 - 1, dy = r % 3 - 1;
      const cx<R>* K = sd + (slot++) * 81;
      const int uz = z - dz;
      int ux = xr - dx, uy = y - dy;
      if (uz < 0 || uz >= Nz) continue;
      if (periodic) {
        if (M == 1 && (dx != 0 || dy != 0)) continue;  // lateral images span |d| <= M-1
        ux = (ux + M) % M;
        uy = (uy + M) % M;
      } else if (ux < 0 || ux >= M || uy < 0 || uy >= M) {
        continue;
      }
      const size_t u = ((size_t)uz * M + ux) * M + uy;
      const cx<R>* t = tp + u * 4;
      const cx<R> tu = t[0], td = t[1], to = t[2], ts = t[3];
      cx<R> xv[9];
#pragma unroll
      for (int c = 0; c < 9; ++c) xv[c] = x[u * 9 + c];
      const cx<R> sum = cadd<R>(cadd<R>(xv[3], xv[4]), xv[5]);
      const cx<R> dmo = csub<R>(td, to);
      const cx<R> tos = cmul<R>(to, sum);
      cx<R> tau[9];
#pragma unroll
      for (int c = 0; c < 3; ++c) tau[c] = cmul<R>(tu, xv[c]);
#pragma unroll
      for (int c = 3; c < 6; ++c) tau[c] = cadd<R>(cmul<R>(dmo, xv[c]), tos);
#pragma unroll
      for (int c = 6; c < 9; ++c) tau[c] = cmul<R>(ts, xv[c]);
#pragma unroll
      for (int i = 0; i < 9; ++i)
#pragma unroll
        for (int j = 0; j < 9; ++j) cfma<R>(acc[i], K[i * 9 + j], tau[j]);
    }
#pragma unroll
    for (int c = 0; c < 9; ++c) {
      const cx<R> o = out[v * 9 + c];
      out[v * 9 + c] = MODE == 0 ? csub<R>(o, acc[c]) : cadd<R>(o, acc[c]);
    }
  }
}

// ── host dispatch ───────────────────────────────────────────────────────────

template <typename R> struct PlanBufs;
template <> struct PlanBufs<double> {
  static const double2* tgen(const Plan& p) { return p.tgen128; }
  static const double2* table(const Plan& p) { return p.table128; }
  static const double2* nn(const Plan& p) { return p.nn128; }
  static const double2* tp(const Plan& p) { return p.tp128; }
  static const double2* tw(const Plan& p) { return p.tw128; }
};
template <> struct PlanBufs<float> {
  static const float2* tgen(const Plan& p) { return p.tgen64; }
  static const float2* table(const Plan& p) { return p.table64; }
  static const float2* nn(const Plan& p) { return p.nn64; }
  static const float2* tp(const Plan& p) { return p.tp64; }
  static const float2* tw(const Plan& p) { return p.tw64; }
};

// Tiles: rows per workgroup of the y passes (B), the x-pass tile code (1, 2,
// 4 ky columns, or 8, 16, 32 complex of each row) and columns per workgroup
// of the z pass (Q).  Each plan starts from the defaults below
// (measured best at the 128x128x32 slab, P = 256, Pz = 64) and autotune()
// re-picks them per precision by timing every candidate in place.
constexpr int kLdsMax = 160 * 1024;
template <typename R, int N, int PER, int T>
constexpr bool lds_fits() {
  return N * (PER * T + 1) * (int)sizeof(cx<R>) <= kLdsMax;
}

int default_tile(int prec, int pass) {
  // pass: 0 y-forward B, 1 x-forward Q, 2 z Q, 3 x-inverse Q, 4 y-inverse B
  static const int t64[5] = {1, 16, 8, 16, 1};
  static const int t128[5] = {1, 16, 8, 8, 1};
  return prec == 1 ? t64[pass] : t128[pass];
}

template <typename R> constexpr int prec_of() { return sizeof(cx<R>) == 8 ? 1 : 0; }

template <typename R, int P, int B>
static void fwd_y_b(Plan& p, const cx<R>* x, cx<R>* b1, hipStream_t st, const int* done) {
  const int nrows = p.Nz * p.M;
  if (p.general)
    hipLaunchKernelGGL((k_fwd_y<R, P, B, true>), dim3((nrows + B - 1) / B), dim3(kNT), 0, st, x,
                       PlanBufs<R>::tp(p), p.tidx, PlanBufs<R>::tgen(p), b1, p.M, nrows,
                       PlanBufs<R>::tw(p), done);
  else
    hipLaunchKernelGGL((k_fwd_y<R, P, B>), dim3((nrows + B - 1) / B), dim3(kNT), 0, st, x,
                       PlanBufs<R>::tp(p), nullptr, nullptr, b1, p.M, nrows, PlanBufs<R>::tw(p),
                       done);
}
template <typename R, int P, int B, int MODE>
static void inv_y_b(Plan& p, const cx<R>* b1, const cx<R>* x, const cx<R>* bvec, cx<R>* out,
                    hipStream_t st, const int* done) {
  const int nrows = p.Nz * p.M;
  hipLaunchKernelGGL((k_inv_y<R, P, B, MODE>), dim3((nrows + B - 1) / B), dim3(kNT), 0, st, b1, x,
                     bvec, out, p.M, nrows, PlanBufs<R>::tw(p), done);
}

template <typename R, int P>
static int pass_fwd_y(Plan& p, const cx<R>* x, cx<R>* b1, hipStream_t st, const int* done) {
  switch (p.tile[prec_of<R>()][0]) {
    case 2:
      if constexpr (lds_fits<R, P, 9, 2>()) { fwd_y_b<R, P, 2>(p, x, b1, st, done); break; }
      [[fallthrough]];
    default:
      fwd_y_b<R, P, 1>(p, x, b1, st, done);
  }
  MSC_LAUNCH_CHECK();
  return kOk;
}

// x-pass tile codes: 1, 2, 4 = Q whole ky columns (9Q complex per row),
// 8, 16, 32 = that many complex per row (whole cache lines).
template <typename R, int P, int NTR>
constexpr bool x_tile_fits() {
  return P >= 2 && (P * 9) % NTR == 0 &&
         P * lds_ld<NTR>::value * (int)sizeof(cx<R>) <= kLdsMax;
}

template <typename R, int P, int NTR>
static bool launch_x(bool fwd, Plan& p, const cx<R>* src, cx<R>* dst, hipStream_t st,
                     const int* done) {
  if constexpr (x_tile_fits<R, P, NTR>()) {
    const dim3 g(p.Nz * (P * 9 / NTR));
    const cx<R>* tw = PlanBufs<R>::tw(p);
    if (fwd) hipLaunchKernelGGL((k_fwd_x<R, P, NTR>), g, dim3(kNT), 0, st, src, dst, p.M, tw, done);
    else hipLaunchKernelGGL((k_inv_x<R, P, NTR>), g, dim3(kNT), 0, st, src, dst, p.M, tw, done);
    return true;
  }
  return false;
}

// The plan's tile for the pass, else the next smaller one that fits.
template <typename R, int P>
static void pass_x(bool fwd, Plan& p, const cx<R>* src, cx<R>* dst, hipStream_t st,
                   const int* done) {
  const int code = p.tile[prec_of<R>()][fwd ? 1 : 3];
  bool ok = false;
  if (code >= 8) {
    if (code >= 32) ok = launch_x<R, P, 32>(fwd, p, src, dst, st, done);
    if (!ok && code >= 16) ok = launch_x<R, P, 16>(fwd, p, src, dst, st, done);
    if (!ok) ok = launch_x<R, P, 8>(fwd, p, src, dst, st, done);
  } else {
    if (code >= 4) ok = launch_x<R, P, 36>(fwd, p, src, dst, st, done);
    if (!ok && code >= 2) ok = launch_x<R, P, 18>(fwd, p, src, dst, st, done);
  }
  if (!ok) launch_x<R, P, 9>(fwd, p, src, dst, st, done);
}

template <typename R, int P>
static int pass_fwd_x(Plan& p, const cx<R>* b1, cx<R>* b2, hipStream_t st, const int* done) {
  pass_x<R, P>(true, p, b1, b2, st, done);
  MSC_LAUNCH_CHECK();
  return kOk;
}

template <typename R, int P>
static int pass_inv_x(Plan& p, const cx<R>* b2, cx<R>* b1, hipStream_t st, const int* done) {
  pass_x<R, P>(false, p, b2, b1, st, done);
  MSC_LAUNCH_CHECK();
  return kOk;
}

template <typename R, int P>
static int pass_inv_y(Plan& p, const cx<R>* b1, const cx<R>* x, const cx<R>* bvec, cx<R>* out,
                      int mode, hipStream_t st, const int* done) {
  const bool two = p.tile[prec_of<R>()][4] == 2 && lds_fits<R, P, 9, 2>();
  if (two) {
    if constexpr (lds_fits<R, P, 9, 2>()) {
      if (mode == 1) inv_y_b<R, P, 2, 1>(p, b1, x, bvec, out, st, done);
      else if (mode == 2) inv_y_b<R, P, 2, 2>(p, b1, x, bvec, out, st, done);
      else inv_y_b<R, P, 2, 0>(p, b1, x, bvec, out, st, done);
    }
  } else {
    if (mode == 1) inv_y_b<R, P, 1, 1>(p, b1, x, bvec, out, st, done);
    else if (mode == 2) inv_y_b<R, P, 1, 2>(p, b1, x, bvec, out, st, done);
    else inv_y_b<R, P, 1, 0>(p, b1, x, bvec, out, st, done);
  }
  MSC_LAUNCH_CHECK();
  return kOk;
}

int timing_begin(Plan& p, hipStream_t st, size_t* slot, int elem_bytes) {
  *slot = SIZE_MAX;
  if (!p.timing) return kOk;
  if (p.ev_used + 2 > p.ev_pool.size()) {
    for (int i = 0; i < 256; ++i) {
      hipEvent_t e;
      MSC_HIP_CHECK(hipEventCreate(&e));
      p.ev_pool.push_back(e);
    }
    p.ev_prec.resize(p.ev_pool.size() / 2);
  }
  p.ev_prec[p.ev_used / 2] = elem_bytes;
  MSC_HIP_CHECK(hipEventRecord(p.ev_pool[p.ev_used], st));
  *slot = p.ev_used;
  p.ev_used += 2;
  return kOk;
}

int timing_end(Plan& p, hipStream_t st, size_t slot) {
  if (slot == SIZE_MAX) return kOk;
  MSC_HIP_CHECK(hipEventRecord(p.ev_pool[slot + 1], st));
  return kOk;
}

template <typename R, int PZ, int Q>
static void z_q(Plan& p, cx<R>* b2, hipStream_t st, const int* done) {
  hipLaunchKernelGGL((k_z_mult<R, PZ, Q>), dim3(p.Pxy * (p.Pxy / Q)), dim3(kNT), 0, st, b2,
                     PlanBufs<R>::table(p), p.Nz, p.Pxy, PlanBufs<R>::tw(p),
                     p.zmap + zmap_offset(p.Pxy, Q), done);
}
template <typename R, int PZ, int Q>
static void z_q_pr(Plan& p, cx<R>* b2, hipStream_t st, const int* done) {
  hipLaunchKernelGGL((k_z_mult_pr<R, PZ, Q>), dim3(p.Pxy * (p.Pxy / Q)), dim3(kNT), 0, st, b2,
                     PlanBufs<R>::table(p), p.Nz, p.Pxy, PlanBufs<R>::tw(p),
                     p.zmap + zmap_offset(p.Pxy, Q), done);
}

// Persistent direct-to-LDS form: the z tile code is Q | kZDma.  The grid is
// the number of workgroups resident at once (occupancy from the runtime,
// times the CU count), a multiple of 8, and at most one per tile.
template <typename R, int PZ, int Q>
static void z_q_dma(Plan& p, cx<R>* b2, hipStream_t st, const int* done) {
  static const int per_cu = [] {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_z_mult_dma<R, PZ, Q>, kNT, 0) !=
            hipSuccess || n < 1)
      n = 1;
    return n;
  }();
  static const int ncu = [] {
    int dev = 0, n = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        n < 1)
      n = 8;
    return n;
  }();
  const int ntiles = p.Pxy * (p.Pxy / Q);
  int g = per_cu * ncu;
  g -= g % 8;
  g = std::max(8, std::min(g, ntiles));
  hipLaunchKernelGGL((k_z_mult_dma<R, PZ, Q>), dim3(g), dim3(kNT), 0, st, b2,
                     PlanBufs<R>::table(p), p.Nz, p.Pxy, PlanBufs<R>::tw(p),
                     p.zmap + zmap_offset(p.Pxy, Q), ntiles, done);
}

template <typename R, int PZ, int Q>
constexpr bool dma_fits() {
  return (9 * Q * (int)sizeof(cx<R>)) % 16 == 0 &&
         (PZ * lds_ld<9 * Q>::value + (PZ / 2) * 9 * Q) * (int)sizeof(cx<R>) <= kLdsMax;
}

template <typename R, int PZ>
static int launch_z(Plan& p, cx<R>* b2, hipStream_t st, const int* done) {
  size_t slot;
  int rc = timing_begin(p, st, &slot, (int)sizeof(cx<R>));
  if (rc) return rc;
  const int code = p.tile[prec_of<R>()][2];
  int q = code & (kZDma - 1);
  while (q > 1 && p.Pxy % q != 0) q >>= 1;
  if (code & kZPrune) {
    switch (q) {
      case 16:
      case 8:
        if constexpr (lds_fits<R, PZ, 9, 8>()) { z_q_pr<R, PZ, 8>(p, b2, st, done); break; }
        [[fallthrough]];
      case 4:
        if constexpr (lds_fits<R, PZ, 9, 4>()) { z_q_pr<R, PZ, 4>(p, b2, st, done); break; }
        [[fallthrough]];
      case 2:
        if constexpr (lds_fits<R, PZ, 9, 2>()) { z_q_pr<R, PZ, 2>(p, b2, st, done); break; }
        [[fallthrough]];
      default:
        z_q_pr<R, PZ, 1>(p, b2, st, done);
    }
    MSC_LAUNCH_CHECK();
    return timing_end(p, st, slot);
  }
  if (code & kZDma) {
    switch (q) {
      case 8:
        if constexpr (dma_fits<R, PZ, 8>()) { z_q_dma<R, PZ, 8>(p, b2, st, done); break; }
        [[fallthrough]];
      case 4:
        if constexpr (dma_fits<R, PZ, 4>()) { z_q_dma<R, PZ, 4>(p, b2, st, done); break; }
        [[fallthrough]];
      default:
        if constexpr (dma_fits<R, PZ, 2>()) {
          if (p.Pxy % 2 == 0) { z_q_dma<R, PZ, 2>(p, b2, st, done); break; }
        }
        z_q<R, PZ, 1>(p, b2, st, done);
    }
    MSC_LAUNCH_CHECK();
    return timing_end(p, st, slot);
  }
  switch (q) {
    case 16:
      if constexpr (lds_fits<R, PZ, 9, 16>()) { z_q<R, PZ, 16>(p, b2, st, done); break; }
      [[fallthrough]];
    case 8:
      if constexpr (lds_fits<R, PZ, 9, 8>()) { z_q<R, PZ, 8>(p, b2, st, done); break; }
      [[fallthrough]];
    case 4:
      if constexpr (lds_fits<R, PZ, 9, 4>()) { z_q<R, PZ, 4>(p, b2, st, done); break; }
      [[fallthrough]];
    case 2:
      if constexpr (lds_fits<R, PZ, 9, 2>()) { z_q<R, PZ, 2>(p, b2, st, done); break; }
      [[fallthrough]];
    default:
      z_q<R, PZ, 1>(p, b2, st, done);
  }
  MSC_LAUNCH_CHECK();
  return timing_end(p, st, slot);
}

template <typename R, int P>
static int launch_xy_fwd(Plan& p, const cx<R>* x, cx<R>* b1, cx<R>* b2, hipStream_t st,
                         const int* done) {
  int rc = pass_fwd_y<R, P>(p, x, b1, st, done);
  if (rc) return rc;
  return pass_fwd_x<R, P>(p, b1, b2, st, done);
}

template <typename R, int P>
static int launch_xy_inv(Plan& p, const cx<R>* x, const cx<R>* bvec, cx<R>* out, cx<R>* b1,
                         cx<R>* b2, int mode, hipStream_t st, const int* done) {
  int rc = pass_inv_x<R, P>(p, b2, b1, st, done);
  if (rc) return rc;
  return pass_inv_y<R, P>(p, b1, x, bvec, out, mode, st, done);
}

#define MSC_XY_CASES(F, ...)                 \
  switch (p.Pxy) {                           \
    case 2: return F<R, 2>(__VA_ARGS__);     \
    case 4: return F<R, 4>(__VA_ARGS__);     \
    case 8: return F<R, 8>(__VA_ARGS__);     \
    case 16: return F<R, 16>(__VA_ARGS__);   \
    case 32: return F<R, 32>(__VA_ARGS__);   \
    case 64: return F<R, 64>(__VA_ARGS__);   \
    case 128: return F<R, 128>(__VA_ARGS__); \
    case 256: return F<R, 256>(__VA_ARGS__); \
    case 512: return F<R, 512>(__VA_ARGS__); \
    case 1024: return F<R, 1024>(__VA_ARGS__); \
    default: break;                          \
  }

#define MSC_Z_CASES(F, ...)                  \
  switch (p.Pz) {                            \
    case 2: return F<R, 2>(__VA_ARGS__);     \
    case 4: return F<R, 4>(__VA_ARGS__);     \
    case 8: return F<R, 8>(__VA_ARGS__);     \
    case 16: return F<R, 16>(__VA_ARGS__);   \
    case 32: return F<R, 32>(__VA_ARGS__);   \
    case 64: return F<R, 64>(__VA_ARGS__);   \
    case 128: return F<R, 128>(__VA_ARGS__); \
    case 256: return F<R, 256>(__VA_ARGS__); \
    case 512: return F<R, 512>(__VA_ARGS__); \
    default: break;                          \
  }

template <typename R>
static int xy_fwd(Plan& p, const cx<R>* x, cx<R>* b1, cx<R>* b2, hipStream_t st, const int* done) {
  MSC_XY_CASES(launch_xy_fwd, p, x, b1, b2, st, done);
  set_error("unsupported lateral FFT length");
  return kUnsupported;
}
template <typename R>
static int xy_inv(Plan& p, const cx<R>* x, const cx<R>* bvec, cx<R>* out, cx<R>* b1, cx<R>* b2,
                  int mode, hipStream_t st, const int* done) {
  MSC_XY_CASES(launch_xy_inv, p, x, bvec, out, b1, b2, mode, st, done);
  set_error("unsupported lateral FFT length");
  return kUnsupported;
}
template <typename R>
static int z_pass(Plan& p, cx<R>* b2, hipStream_t st, const int* done) {
  MSC_Z_CASES(launch_z, p, b2, st, done);
  set_error("unsupported vertical FFT length");
  return kUnsupported;
}

template <typename R>
static int apply(Plan& p, const cx<R>* x, const cx<R>* bvec, cx<R>* out, int mode,
                 hipStream_t st, const int* done) {
  if (!p.have_table || !p.have_tparams) {
    set_error("plan needs build_g0 and set_tmatrix before the operator can be applied");
    return kNotReady;
  }
  cx<R>* b1 = reinterpret_cast<cx<R>*>(p.buf1);
  cx<R>* b2 = reinterpret_cast<cx<R>*>(p.buf2);
  int rc = xy_fwd<R>(p, x, b1, b2, st, done);
  if (rc) return rc;
  rc = z_pass<R>(p, b2, st, done);
  if (rc) return rc;
  rc = xy_inv<R>(p, x, bvec, out, b1, b2, mode, st, done);
  if (rc || !p.have_nn) return rc;
  if (p.general) {
    set_error("the volume-averaged correction needs the cubic T0, not a general per-voxel T");
    return kUnsupported;
  }
  const size_t nvox = (size_t)p.Nz * p.M * p.M;
  const int grid = (int)std::min<size_t>((nvox + kNT - 1) / kNT, 4096);
  const cx<R>* tp = PlanBufs<R>::tp(p);
  const cx<R>* nn = PlanBufs<R>::nn(p);
  const int per = p.periodic ? 1 : 0;
  if (mode == 0)
    hipLaunchKernelGGL((k_nn_apply<R, 0>), dim3(grid), dim3(kNT), 0, st, x, tp, nn, out, p.M,
                       p.Nz, per, done);
  else
    hipLaunchKernelGGL((k_nn_apply<R, 1>), dim3(grid), dim3(kNT), 0, st, x, tp, nn, out, p.M,
                       p.Nz, per, done);
  MSC_LAUNCH_CHECK();
  return kOk;
}

// ── batched operator: instance q = d[blockIdx.y] ──────────────────────────
template <typename R, int P, int B>
__global__ __launch_bounds__(kNT) void kb_fwd_y(const BPass<R>* __restrict__ d, int j,
                                               const cx<R>* __restrict__ tw) {
  const BPass<R>& q = d[blockIdx.y];
  const int nrows = q.Nz * q.M;
  if ((int)blockIdx.x * B >= nrows) return;
  fwd_y_body<R, P, B, false>(blockIdx.x, 0, q.in + (size_t)j * q.in_step, q.tp, nullptr,
                             nullptr, q.buf1, q.M, nrows, tw, q.done);
}

template <typename R, int P, int NTR>
__global__ __launch_bounds__(kNT) void kb_fwd_x(const BPass<R>* __restrict__ d,
                                               const cx<R>* __restrict__ tw) {
  const BPass<R>& q = d[blockIdx.y];
  const int nblk = q.Nz * (P * 9 / NTR);
  if ((int)blockIdx.x >= nblk) return;
  fwd_x_body<R, P, NTR>(blockIdx.x, nblk, q.buf1, q.buf2, q.M, tw, q.done);
}

template <typename R, int PZ, int Q>
__global__ __launch_bounds__(kNT) void kb_z_mult(const BPass<R>* __restrict__ d, int Pxy,
                                                const cx<R>* __restrict__ tw,
                                                const int* __restrict__ zmap) {
  const BPass<R>& q = d[blockIdx.y];
  z_mult_body<R, PZ, Q>(blockIdx.x, gridDim.x, q.buf2, q.table, q.Nz, Pxy, tw, zmap, q.done);
}

template <typename R, int PZ, int Q>
__global__ __launch_bounds__(kNT) void kb_z_mult_pr(const BPass<R>* __restrict__ d, int Pxy,
                                                   const cx<R>* __restrict__ tw,
                                                   const int* __restrict__ zmap) {
  const BPass<R>& q = d[blockIdx.y];
  z_mult_pruned_body<R, PZ, Q>(blockIdx.x, q.buf2, q.table, q.Nz, Pxy, tw, zmap, q.done);
}

template <typename R, int P, int NTR>
__global__ __launch_bounds__(kNT) void kb_inv_x(const BPass<R>* __restrict__ d,
                                               const cx<R>* __restrict__ tw) {
  const BPass<R>& q = d[blockIdx.y];
  const int nblk = q.Nz * (P * 9 / NTR);
  if ((int)blockIdx.x >= nblk) return;
  inv_x_body<R, P, NTR>(blockIdx.x, nblk, q.buf2, q.buf1, q.M, tw, q.done);
}

template <typename R, int P, int B, int MODE>
__global__ __launch_bounds__(kNT) void kb_inv_y(const BPass<R>* __restrict__ d, int j,
                                               const cx<R>* __restrict__ tw) {
  const BPass<R>& q = d[blockIdx.y];
  const int nrows = q.Nz * q.M;
  if ((int)blockIdx.x * B >= nrows) return;
  inv_y_body<R, P, B, MODE>(blockIdx.x, 0, q.buf1, q.in + (size_t)j * q.in_step, q.bvec,
                            q.out + (size_t)j * q.out_step, q.M, nrows, tw, q.done);
}

// blocks of a batched x pass: a multiple of 8, so that slot b of every
// instance runs on XCD b % 8 as in the single-instance launch
static unsigned round8(int v) { return (unsigned)((v + 7) / 8 * 8); }

template <typename R, int P, int NTR>
static bool batch_x_q(bool fwd, const BPass<R>* d, int nsys, int max_Nz, const cx<R>* tw,
                      hipStream_t st) {
  if constexpr (x_tile_fits<R, P, NTR>()) {
    const dim3 g(round8(max_Nz * (P * 9 / NTR)), nsys);
    if (fwd) hipLaunchKernelGGL((kb_fwd_x<R, P, NTR>), g, dim3(kNT), 0, st, d, tw);
    else hipLaunchKernelGGL((kb_inv_x<R, P, NTR>), g, dim3(kNT), 0, st, d, tw);
    return true;
  }
  return false;
}

template <typename R, int P>
static int batch_xy(const Plan& ref, const BPass<R>* d, int nsys, int max_M, int max_Nz, int mode,
                    int j, bool forward, hipStream_t st) {
  const int pr = prec_of<R>();
  const cx<R>* tw = PlanBufs<R>::tw(ref);
  const int rows = max_Nz * max_M;
  bool two = false;
  if constexpr (lds_fits<R, P, 9, 2>()) two = ref.tile[pr][forward ? 0 : 4] == 2;
  const dim3 gy(two ? (rows + 1) / 2 : rows, nsys);
  const int q = ref.tile[pr][forward ? 1 : 3];
  if (forward) {
    if constexpr (lds_fits<R, P, 9, 2>()) {
      if (two) hipLaunchKernelGGL((kb_fwd_y<R, P, 2>), gy, dim3(kNT), 0, st, d, j, tw);
    }
    if (!two) hipLaunchKernelGGL((kb_fwd_y<R, P, 1>), gy, dim3(kNT), 0, st, d, j, tw);
  }
  bool done = false;
  if (q >= 8) {
    if (q >= 32) done = batch_x_q<R, P, 32>(forward, d, nsys, max_Nz, tw, st);
    if (!done && q >= 16) done = batch_x_q<R, P, 16>(forward, d, nsys, max_Nz, tw, st);
    if (!done) done = batch_x_q<R, P, 8>(forward, d, nsys, max_Nz, tw, st);
  } else {
    if (q >= 4) done = batch_x_q<R, P, 36>(forward, d, nsys, max_Nz, tw, st);
    if (!done && q >= 2) done = batch_x_q<R, P, 18>(forward, d, nsys, max_Nz, tw, st);
  }
  if (!done) batch_x_q<R, P, 9>(forward, d, nsys, max_Nz, tw, st);
  if (!forward) {
    if constexpr (lds_fits<R, P, 9, 2>()) {
      if (two) {
        if (mode == 1) hipLaunchKernelGGL((kb_inv_y<R, P, 2, 1>), gy, dim3(kNT), 0, st, d, j, tw);
        else hipLaunchKernelGGL((kb_inv_y<R, P, 2, 2>), gy, dim3(kNT), 0, st, d, j, tw);
      }
    }
    if (!two) {
      if (mode == 1) hipLaunchKernelGGL((kb_inv_y<R, P, 1, 1>), gy, dim3(kNT), 0, st, d, j, tw);
      else hipLaunchKernelGGL((kb_inv_y<R, P, 1, 2>), gy, dim3(kNT), 0, st, d, j, tw);
    }
  }
  MSC_LAUNCH_CHECK();
  return kOk;
}

template <typename R, int PZ>
static int batch_z(const Plan& ref, const BPass<R>* d, int nsys, hipStream_t st) {
  const cx<R>* tw = PlanBufs<R>::tw(ref);
  const int code = ref.tile[prec_of<R>()][2];
  int q = code & (kZDma - 1);
  while (q > 1 && ref.Pxy % q != 0) q >>= 1;
  const int P = ref.Pxy;
  if (code & kZPrune) {
    switch (q) {
      case 16:
      case 8:
        if constexpr (lds_fits<R, PZ, 9, 8>()) {
          hipLaunchKernelGGL((kb_z_mult_pr<R, PZ, 8>), dim3(P * (P / 8), nsys), dim3(kNT), 0, st,
                             d, P, tw, ref.zmap + zmap_offset(P, 8));
          break;
        }
        [[fallthrough]];
      case 4:
        if constexpr (lds_fits<R, PZ, 9, 4>()) {
          if (P % 4 == 0) {
            hipLaunchKernelGGL((kb_z_mult_pr<R, PZ, 4>), dim3(P * (P / 4), nsys), dim3(kNT), 0,
                               st, d, P, tw, ref.zmap + zmap_offset(P, 4));
            break;
          }
        }
        [[fallthrough]];
      case 2:
        if constexpr (lds_fits<R, PZ, 9, 2>()) {
          if (P % 2 == 0) {
            hipLaunchKernelGGL((kb_z_mult_pr<R, PZ, 2>), dim3(P * (P / 2), nsys), dim3(kNT), 0,
                               st, d, P, tw, ref.zmap + zmap_offset(P, 2));
            break;
          }
        }
        [[fallthrough]];
      default:
        hipLaunchKernelGGL((kb_z_mult_pr<R, PZ, 1>), dim3(P * P, nsys), dim3(kNT), 0, st, d, P,
                           tw, ref.zmap);
    }
    MSC_LAUNCH_CHECK();
    return kOk;
  }
  switch (q) {
    case 16:
    case 8:
      if constexpr (lds_fits<R, PZ, 9, 8>()) {
        hipLaunchKernelGGL((kb_z_mult<R, PZ, 8>), dim3(P * (P / 8), nsys), dim3(kNT), 0, st, d, P,
                           tw, ref.zmap + zmap_offset(P, 8));
        break;
      }
      [[fallthrough]];
    case 4:
      if constexpr (lds_fits<R, PZ, 9, 4>()) {
        if (P % 4 == 0) {
          hipLaunchKernelGGL((kb_z_mult<R, PZ, 4>), dim3(P * (P / 4), nsys), dim3(kNT), 0, st, d,
                             P, tw, ref.zmap + zmap_offset(P, 4));
          break;
        }
      }
      [[fallthrough]];
    case 2:
      if constexpr (lds_fits<R, PZ, 9, 2>()) {
        if (P % 2 == 0) {
          hipLaunchKernelGGL((kb_z_mult<R, PZ, 2>), dim3(P * (P / 2), nsys), dim3(kNT), 0, st, d,
                             P, tw, ref.zmap + zmap_offset(P, 2));
          break;
        }
      }
      [[fallthrough]];
    default:
      hipLaunchKernelGGL((kb_z_mult<R, PZ, 1>), dim3(P * P, nsys), dim3(kNT), 0, st, d, P, tw,
                         ref.zmap);
  }
  MSC_LAUNCH_CHECK();
  return kOk;
}

template <typename R>
int apply_batch(const Plan& ref, const BPass<R>* d, int nsys, int max_M, int max_Nz, int mode,
                int j, hipStream_t st) {
  if (nsys < 1) return kOk;
  if (mode != 1 && mode != 2) {
    set_error("apply_batch: mode must be 1 (residual) or 2 (scattered)");
    return kInvalid;
  }
  const Plan& p = ref;
  int rc = kUnsupported;
  auto xy = [&](bool fwd) -> int {
    switch (p.Pxy) {
      case 2: return batch_xy<R, 2>(p, d, nsys, max_M, max_Nz, mode, j, fwd, st);
      case 4: return batch_xy<R, 4>(p, d, nsys, max_M, max_Nz, mode, j, fwd, st);
      case 8: return batch_xy<R, 8>(p, d, nsys, max_M, max_Nz, mode, j, fwd, st);
      case 16: return batch_xy<R, 16>(p, d, nsys, max_M, max_Nz, mode, j, fwd, st);
      case 32: return batch_xy<R, 32>(p, d, nsys, max_M, max_Nz, mode, j, fwd, st);
      case 64: return batch_xy<R, 64>(p, d, nsys, max_M, max_Nz, mode, j, fwd, st);
      case 128: return batch_xy<R, 128>(p, d, nsys, max_M, max_Nz, mode, j, fwd, st);
      case 256: return batch_xy<R, 256>(p, d, nsys, max_M, max_Nz, mode, j, fwd, st);
      case 512: return batch_xy<R, 512>(p, d, nsys, max_M, max_Nz, mode, j, fwd, st);
      case 1024: return batch_xy<R, 1024>(p, d, nsys, max_M, max_Nz, mode, j, fwd, st);
      default: set_error("unsupported lateral FFT length"); return kUnsupported;
    }
  };
  rc = xy(true);
  if (rc) return rc;
  switch (p.Pz) {
    case 2: rc = batch_z<R, 2>(p, d, nsys, st); break;
    case 4: rc = batch_z<R, 4>(p, d, nsys, st); break;
    case 8: rc = batch_z<R, 8>(p, d, nsys, st); break;
    case 16: rc = batch_z<R, 16>(p, d, nsys, st); break;
    case 32: rc = batch_z<R, 32>(p, d, nsys, st); break;
    case 64: rc = batch_z<R, 64>(p, d, nsys, st); break;
    case 128: rc = batch_z<R, 128>(p, d, nsys, st); break;
    case 256: rc = batch_z<R, 256>(p, d, nsys, st); break;
    case 512: rc = batch_z<R, 512>(p, d, nsys, st); break;
    default: set_error("unsupported vertical FFT length"); return kUnsupported;
  }
  if (rc) return rc;
  return xy(false);
}

template int apply_batch<double>(const Plan&, const BPass<double>*, int, int, int, int, int,
                                 hipStream_t);
template int apply_batch<float>(const Plan&, const BPass<float>*, int, int, int, int, int,
                                hipStream_t);

template <typename R>
int matvec(Plan& p, const cx<R>* x, cx<R>* y, hipStream_t st, const int* done) {
  return apply<R>(p, x, nullptr, y, 0, st, done);
}
template <typename R>
int residual(Plan& p, const cx<R>* b, const cx<R>* x, cx<R>* r, hipStream_t st, const int* done) {
  return apply<R>(p, x, b, r, 1, st, done);
}
template <typename R>
int scattered(Plan& p, const cx<R>* x, cx<R>* y, hipStream_t st, const int* done) {
  return apply<R>(p, x, nullptr, y, 2, st, done);
}

// Pick the tile of every pass for precision R by timing whole operator
// applications with each candidate (the other passes at their current tiles)
// on the plan's workspaces (inputs: x, outputs: out, both n complex).
template <typename R>
int autotune(Plan& p, const cx<R>* x, cx<R>* out, hipStream_t st) {
  if (!p.have_table || !p.have_tparams) {
    set_error("autotune needs build_g0 and set_tmatrix first");
    return kNotReady;
  }
  const int pr = prec_of<R>();
  const bool saved_timing = p.timing;
  p.timing = false;
  hipEvent_t e0, e1;
  MSC_HIP_CHECK(hipEventCreate(&e0));
  MSC_HIP_CHECK(hipEventCreate(&e1));
  static const int cand_y[] = {1, 2};
  static const int cand_x[] = {1, 2, 4, 8, 16, 32};
  static const int cand_z[] = {2, 4, 8, 2 | kZDma, 4 | kZDma, 8 | kZDma,
                               2 | kZPrune, 4 | kZPrune, 8 | kZPrune};
  static const int order[] = {2, 1, 3, 0, 4};
  int rc = apply<R>(p, x, nullptr, out, 2, st, nullptr);  // warm every buffer
  for (int pass : order) {
    if (rc) break;
    const int* cand = pass == 2 ? cand_z : (pass == 0 || pass == 4 ? cand_y : cand_x);
    const int nc = pass == 2 ? 9 : (pass == 0 || pass == 4 ? 2 : 6);
    int best = p.tile[pr][pass];
    float best_ms = 1e30f;
    for (int i = 0; i < nc && !rc; ++i) {
      p.tile[pr][pass] = cand[i];
      // whole operator applications: a pass repeated on its own keeps its
      // buffers and the table in the 256 MB last-level cache, which the
      // solve's other passes and Krylov vectors evict (the z pass measured
      // 150 us repeated alone against 193 us inside solves at 128x128x32)
      auto run = [&]() { return apply<R>(p, x, nullptr, out, 2, st, nullptr); };
      rc = run();
      if (rc) break;
      MSC_HIP_CHECK(hipEventRecord(e0, st));
      for (int rep = 0; rep < 4 && !rc; ++rep) rc = run();
      MSC_HIP_CHECK(hipEventRecord(e1, st));
      MSC_HIP_CHECK(hipEventSynchronize(e1));
      float ms = 0.f;
      MSC_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
      if (ms < best_ms) {
        best_ms = ms;
        best = cand[i];
      }
    }
    p.tile[pr][pass] = best;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  p.timing = saved_timing;
  if (rc) return rc;
  p.tuned[pr] = true;
  return kOk;
}

template int autotune<double>(Plan&, const double2*, double2*, hipStream_t);
template int autotune<float>(Plan&, const float2*, float2*, hipStream_t);

template int matvec<double>(Plan&, const double2*, double2*, hipStream_t, const int*);
template int matvec<float>(Plan&, const float2*, float2*, hipStream_t, const int*);
template int residual<double>(Plan&, const double2*, const double2*, double2*, hipStream_t,
                              const int*);
template int residual<float>(Plan&, const float2*, const float2*, float2*, hipStream_t,
                             const int*);
template int scattered<double>(Plan&, const double2*, double2*, hipStream_t, const int*);
template int scattered<float>(Plan&, const float2*, float2*, hipStream_t, const int*);

// ── generic strided batched FFT (table build, tests) ───────────────────────
// data [outer][N][inner] -> [outer][nout][inner] (the first nout frequencies
// kept, nout <= N; in place only when nout == N); a workgroup transforms NCOL
// consecutive inner indices of one outer slice.
template <typename R, int N, int NCOL, bool INV>
__global__ __launch_bounds__(kNT) void k_fft_axis(const cx<R>* __restrict__ in,
                                                 cx<R>* __restrict__ out, size_t inner, int nout,
                                                 const cx<R>* __restrict__ tw) {
  constexpr int LD = lds_ld<NCOL>::value;
  __shared__ cx<R> s[N * LD];
  const size_t nchunk = (inner + NCOL - 1) / NCOL;
  const size_t o = blockIdx.x / nchunk;
  const size_t i0 = (blockIdx.x - o * nchunk) * NCOL;
  const size_t base = o * (size_t)N * inner + i0;
  const size_t obase = o * (size_t)nout * inner + i0;
  for (int w = threadIdx.x; w < N * NCOL; w += kNT) {
    const int r = w / NCOL;
    const int e = w - r * NCOL;
    cx<R> v = czero<R>();
    if (i0 + e < inner) v = in[base + (size_t)r * inner + e];
    s[r * LD + e] = v;
  }
  __syncthreads();
  lds_fft<R, N, INV, NCOL, LD, kNT>(s, tw);
  for (int w = threadIdx.x; w < nout * NCOL; w += kNT) {
    const int r = w / NCOL;
    const int e = w - r * NCOL;
    if (i0 + e < inner) out[obase + (size_t)r * inner + e] = s[r * LD + e];
  }
}

template <typename R, int N>
static int launch_fft_axis(const cx<R>* in, cx<R>* out, int outer, size_t inner, int nout,
                           bool inv, const cx<R>* tw, hipStream_t st) {
  // 8 columns above 64 points: 4-5 workgroups per CU instead of 2 (LDS)
  constexpr int NCOL = (N <= 64) ? 16 : (N * 8 * sizeof(cx<R>) <= 80 * 1024 ? 8 : 4);
  const size_t nchunk = (inner + NCOL - 1) / NCOL;
  const size_t nblk = (size_t)outer * nchunk;
  MSC_REQUIRE(nblk < (1ull << 31), "FFT batch too large");
  if (inv) {
    hipLaunchKernelGGL((k_fft_axis<R, N, NCOL, true>), dim3((unsigned)nblk), dim3(kNT), 0, st, in,
                       out, inner, nout, tw);
  } else {
    hipLaunchKernelGGL((k_fft_axis<R, N, NCOL, false>), dim3((unsigned)nblk), dim3(kNT), 0, st, in,
                       out, inner, nout, tw);
  }
  MSC_LAUNCH_CHECK();
  return kOk;
}

template <typename R>
int fft_axis_pruned(const cx<R>* in, cx<R>* out, int outer, int N, size_t inner, int nout,
                    bool inverse, const cx<R>* tw, hipStream_t st) {
  MSC_REQUIRE(nout >= 1 && nout <= N, "fft_axis: nout must be in [1, N]");
  MSC_REQUIRE(nout == N || in != out, "fft_axis: a pruned transform cannot run in place");
  switch (N) {
    case 1: {
      if (in != out) {
        MSC_HIP_CHECK(hipMemcpyAsync(out, in, sizeof(cx<R>) * (size_t)outer * inner,
                                     hipMemcpyDeviceToDevice, st));
      }
      return kOk;
    }
    case 2: return launch_fft_axis<R, 2>(in, out, outer, inner, nout, inverse, tw, st);
    case 4: return launch_fft_axis<R, 4>(in, out, outer, inner, nout, inverse, tw, st);
    case 8: return launch_fft_axis<R, 8>(in, out, outer, inner, nout, inverse, tw, st);
    case 16: return launch_fft_axis<R, 16>(in, out, outer, inner, nout, inverse, tw, st);
    case 32: return launch_fft_axis<R, 32>(in, out, outer, inner, nout, inverse, tw, st);
    case 64: return launch_fft_axis<R, 64>(in, out, outer, inner, nout, inverse, tw, st);
    case 128: return launch_fft_axis<R, 128>(in, out, outer, inner, nout, inverse, tw, st);
    case 256: return launch_fft_axis<R, 256>(in, out, outer, inner, nout, inverse, tw, st);
    case 512: return launch_fft_axis<R, 512>(in, out, outer, inner, nout, inverse, tw, st);
    case 1024: return launch_fft_axis<R, 1024>(in, out, outer, inner, nout, inverse, tw, st);
    default: break;
  }
  set_error("fft_axis: length must be a power of two <= 1024");
  return kUnsupported;
}

template <typename R>
int fft_axis(const cx<R>* in, cx<R>* out, int outer, int N, size_t inner, bool inverse,
             const cx<R>* tw, hipStream_t st) {
  return fft_axis_pruned<R>(in, out, outer, N, inner, N, inverse, tw, st);
}

template int fft_axis<double>(const double2*, double2*, int, int, size_t, bool, const double2*,
                              hipStream_t);
template int fft_axis<float>(const float2*, float2*, int, int, size_t, bool, const float2*,
                             hipStream_t);
template int fft_axis_pruned<double>(const double2*, double2*, int, int, size_t, int, bool,
                                     const double2*, hipStream_t);

}  // namespace msc
