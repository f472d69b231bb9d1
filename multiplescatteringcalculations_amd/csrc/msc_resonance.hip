// Resonance-regime cube T-matrices: the internal Foldy-Lax system of a cube
// split into n^3 sub-cells, solved for a batch of material contrasts
// (msc_resonance_tmatrices, include/msc_slab.h).
//
// Replaces compute_resonance_tmatrix / _solve_coupled
// (cubic_scattering/resonance_tmatrix.py:650-850): for each contrast the
// 9N x 9N system (I - P (I x T_loc)) psi = psi_inc with 9 incident patterns
// (N = n^3 sub-cells, P the 9x9 Kupradze propagator between sub-cell centres,
// zero self block, T_loc the Rayleigh sub-cell T-matrix), then
// T_comp = sum_m T_loc psi_m.  On the device:
//   * k_res_offsets: P at the (2n-1)^3 distinct centre offsets (FP64), once;
//   * k_res_assemble: every system's augmented column-major matrix
//     [I - P T_loc | psi_inc] (T_loc in its 4-number form);
//   * k_res_lu: one workgroup per system, right-looking blocked Gaussian
//     elimination with partial pivoting (panel of NB columns and the U12 block
//     staged in LDS, trailing update streamed from HBM with lanes along the
//     contiguous rows), then back substitution of the 9 right-hand sides;
//   * k_res_tcomp: the 9x9 composite T-matrix of every system.
// The optional Neumann series of the reference (neumann_order > 0) runs as
// repeated products with B = I - A on the same assembled matrices.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/msc_slab.h"
#include "msc_common.hpp"
#include "msc_prop.hpp"
#include "msc_setup.hpp"

namespace msc {
namespace {

constexpr int kRT = 256;

__device__ __forceinline__ double2 zmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 zsub(double2 a, double2 b) {
  return make_double2(a.x - b.x, a.y - b.y);
}
__device__ __forceinline__ double2 zadd(double2 a, double2 b) {
  return make_double2(a.x + b.x, a.y + b.y);
}
// a -= b c
__device__ __forceinline__ void zfms(double2& a, double2 b, double2 c) {
  a.x = fma(-b.x, c.x, a.x);
  a.x = fma(b.y, c.y, a.x);
  a.y = fma(-b.x, c.y, a.y);
  a.y = fma(-b.y, c.x, a.y);
}

// 9x9 propagator blocks at offsets d * (i0, i1, i2), |i| < n, index
// ((i0 + n-1) (2n-1) + i1 + n-1) (2n-1) + i2 + n-1; zero block at the origin.
__global__ __launch_bounds__(64) void k_res_offsets(double2* __restrict__ P, int n, double d,
                                                   double kP, double kS, double pre) {
  const int no = 2 * n - 1;
  const int tot = no * no * no;
  for (int o = blockIdx.x * blockDim.x + threadIdx.x; o < tot; o += gridDim.x * blockDim.x) {
    const int i0 = o / (no * no) - (n - 1);
    const int i1 = (o / no) % no - (n - 1);
    const int i2 = o % no - (n - 1);
    double2 e45[45];
    for (int e = 0; e < 45; ++e) e45[e] = make_double2(0.0, 0.0);
    double2 K[81];
    if (i0 != 0 || i1 != 0 || i2 != 0) {
      add_prop45(i0 * d, i1 * d, i2 * d, kP, kS, pre, e45);
      expand45(e45, K);
    } else {
      for (int q = 0; q < 81; ++q) K[q] = make_double2(0.0, 0.0);
    }
    for (int q = 0; q < 81; ++q) P[(size_t)o * 81 + q] = K[q];
  }
}

// System b (blockIdx.y): A[c * n9 + r] for c < n9 is I - P T_loc, for
// c = n9 + q the incident pattern q.  T_loc = t_u I3 (+) [[t_d, t_o, t_o], ...]
// (+) t_s I3 (the sub-cell's Rayleigh T-matrix).
__global__ __launch_bounds__(kRT) void k_res_assemble(double2* __restrict__ Aall, size_t stride,
                                                     const double2* __restrict__ P,
                                                     const double2* __restrict__ tp,
                                                     const double2* __restrict__ psi, int n) {
  const int N = n * n * n, n9 = 9 * N, nc = n9 + 9, no = 2 * n - 1;
  double2* A = Aall + (size_t)blockIdx.y * stride;
  const double2* t = tp + (size_t)blockIdx.y * 4;
  const double2 tu = t[0], tdo = zsub(t[1], t[2]), to = t[2], ts = t[3];
  const size_t tot = (size_t)n9 * nc;
  for (size_t e = (size_t)blockIdx.x * kRT + threadIdx.x; e < tot; e += (size_t)gridDim.x * kRT) {
    const int c = (int)(e / n9), r = (int)(e % n9);
    double2 v;
    if (c >= n9) {
      v = psi[(size_t)(c - n9) * n9 + r];
    } else {
      const int m = r / 9, i = r % 9, s = c / 9, j = c % 9;
      if (m == s) {
        v = make_double2(i == j ? 1.0 : 0.0, 0.0);
      } else {
        const int d0 = m / (n * n) - s / (n * n), d1 = (m / n) % n - (s / n) % n,
                  d2 = m % n - s % n;
        const double2* K = P + ((size_t)((d0 + n - 1) * no + d1 + n - 1) * no + d2 + n - 1) * 81 +
                           i * 9;
        double2 pt;
        if (j < 3) {
          pt = zmul(K[j], tu);
        } else if (j < 6) {
          pt = zadd(zmul(K[j], tdo), zmul(zadd(zadd(K[3], K[4]), K[5]), to));
        } else {
          pt = zmul(K[j], ts);
        }
        v = make_double2(-pt.x, -pt.y);
      }
    }
    A[e] = v;
  }
}

// One workgroup (kLT threads) per system: [A | Y] (column-major, lda n, nc
// columns) is reduced in place to [U | Y'] by right-looking blocked
// elimination with partial pivoting, then the 9 solution columns overwrite Y.
// Panels of 16 columns (fewer for the larger systems, whose panels must still
// fit in LDS) are factorised in LDS (factor_panel); the back substitution
// keeps the right-hand sides in LDS when they fit (rhs_lds).  Two kernels
// share these pieces:
//   * k_res_lu_blocked (16-column panels): outer panels of K = 16 S columns.
//     Each 16-column sub-panel updates only the rest of its outer panel; the
//     columns right of the outer panel receive the K row exchanges, K rows of
//     U (forward substitution, column chunks through LDS) and one K-deep
//     trailing update on the matrix cores, so the trailing matrix crosses HBM
//     once per K columns rather than once per 16.
//   * k_res_lu (NB = 8, 4, 2, 1): one NB-column panel at a time, U12 in LDS,
//     the trailing update on the vector units.
constexpr int kLT = 512;
constexpr int kLW = kLT / 64;

// Pivot-search and pivot-growth state of one workgroup.
struct LuShared {
  double red_v[2][kLW];
  int red_i[2][kLW];
  int prow_of[16];        // the row chosen at each step of a panel
  int tpos[32], trow[32];  // positions holding another row, and that row
  int ntouch;
  double umin, umax;
  int sing;
};

__device__ __forceinline__ void lu_shared_init(LuShared& sh) {
  if (threadIdx.x == 0) {
    sh.umin = INFINITY;
    sh.umax = 0.0;
    sh.sing = 0;
  }
}

// One step of the wave arg-max (largest value, smallest index among equal
// values; an associative and commutative choice, so every lane ends with the
// same pair): combine with the lane the DPP control CTRL names.
template <int CTRL>
__device__ __forceinline__ void argmax_dpp(double& best, int& bi) {
  const int olo = __builtin_amdgcn_update_dpp(0, __double2loint(best), CTRL, 0xF, 0xF, true);
  const int ohi = __builtin_amdgcn_update_dpp(0, __double2hiint(best), CTRL, 0xF, 0xF, true);
  const int oi = __builtin_amdgcn_update_dpp(0, bi, CTRL, 0xF, 0xF, true);
  const double ob = __hiloint2double(ohi, olo);
  if (ob > best || (ob == best && oi < bi)) {
    best = ob;
    bi = oi;
  }
}

// The arg-max over the wave, in every lane: within each row of 16 lanes by
// DPP moves (quad permutations [1,0,3,2] and [2,3,0,1], half-row mirror, row
// mirror), then across the four rows by two shuffles.
__device__ __forceinline__ void wave_argmax(double& best, int& bi) {
  argmax_dpp<0xB1>(best, bi);
  argmax_dpp<0x4E>(best, bi);
  argmax_dpp<0x141>(best, bi);
  argmax_dpp<0x140>(best, bi);
  for (int off = 16; off < 64; off <<= 1) {
    const double ob = __shfl_xor(best, off);
    const int oi = __shfl_xor(bi, off);
    if (ob > best || (ob == best && oi < bi)) {
      best = ob;
      bi = oi;
    }
  }
}

// Factorises the panel of columns k0 .. k0 + nb - 1, rows k0 .. n - 1, in LDS
// (pan: column j at pan + j * n) with partial pivoting by implicit row
// exchanges: the rows stay in place and each thread marks its own rows once
// they have served as pivot rows (they leave the search and the updates).
// The maxima of the next column are found while the current one is applied,
// and the wave maxima alternate between two slots, so one barrier per column.
// The exchange sequence is then replayed on the positions it touches (at most
// 2 nb): piv[j] = the panel-local position exchanged with position j at step
// j, as an explicit-exchange elimination records it, and the panel goes back
// to A in that row order (unit L below the diagonal, U on and above it).
// L11[i * NB + j] = the panel's unit lower block (row i, column j).  Ends
// with a barrier.
template <int NB>
__device__ void factor_panel(double2* __restrict__ A, int n, int k0, int nb,
                             double2* __restrict__ pan, int* piv, double2* L11, LuShared& sh) {
  static_assert(NB <= 16 && 2 * NB <= 32, "the exchange map holds 2 NB positions");
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int m = n - k0;
  // 8 loads in flight per thread (indices clamped into the panel, the stores
  // guarded)
  const int tot = nb * m;
#pragma unroll 1
  for (int e0 = t; e0 < tot; e0 += 8 * kLT) {
    double2 v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int e = min(e0 + u * kLT, tot - 1), j = e / m, i = e - j * m;
      v[u] = A[(size_t)(k0 + j) * n + k0 + i];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int e = e0 + u * kLT;
      if (e < tot) {
        const int j = e / m, i = e - j * m;
        pan[j * n + i] = v[u];
      }
    }
  }
  __syncthreads();
  // this thread's rows are t + r kLT (r < 32: m <= 32 kLT); bit r of used:
  // that row has been a pivot row
  unsigned used = 0u;
  double best = -1.0;
  int bi = 0;
  for (int i = t; i < m; i += kLT) {
    const double2 v = pan[i];
    const double a = v.x * v.x + v.y * v.y;
    if (a > best) {
      best = a;
      bi = i;
    }
  }
  for (int j = 0; j < nb; ++j) {
    wave_argmax(best, bi);
    const int slot = j & 1;
    if (lane == 0) {
      sh.red_v[slot][w] = best;
      sh.red_i[slot][w] = bi;
    }
    __syncthreads();
    double b = sh.red_v[slot][0];
    int p = sh.red_i[slot][0];
    for (int q = 1; q < kLW; ++q)
      if (sh.red_v[slot][q] > b || (sh.red_v[slot][q] == b && sh.red_i[slot][q] < p)) {
        b = sh.red_v[slot][q];
        p = sh.red_i[slot][q];
      }
    if (p % kLT == t) used |= 1u << (p / kLT);
    const double2 dgn = pan[j * n + p];
    const double dd = dgn.x * dgn.x + dgn.y * dgn.y;
    if (t == 0) {
      sh.prow_of[j] = p;
      const double ad = sqrt(dd);
      sh.umin = fmin(sh.umin, ad);
      sh.umax = fmax(sh.umax, ad);
      if (!(dd > 0.0)) sh.sing = 1;
    }
    const double2 inv = dd > 0.0 ? make_double2(dgn.x / dd, -dgn.y / dd) : make_double2(0.0, 0.0);
    // scale column j of the rows not yet used, update their columns right of
    // it with the pivot row (row p, not written in this step), and find the
    // maxima of column j + 1 on the way
    best = -1.0;
    bi = 0;
    int r = 0;
    for (int i = t; i < m; i += kLT, ++r) {
      if ((used >> r) & 1u) continue;
      const double2 l = zmul(pan[j * n + i], inv);
      pan[j * n + i] = l;
      for (int jj = j + 1; jj < nb; ++jj) zfms(pan[jj * n + i], l, pan[jj * n + p]);
      if (j + 1 < nb) {
        const double2 v = pan[(j + 1) * n + i];
        const double a = v.x * v.x + v.y * v.y;
        if (a > best) {
          best = a;
          bi = i;
        }
      }
    }
  }
  __syncthreads();
  // the exchange sequence, by wave 0 with the map one entry per lane (lane k:
  // position pos_k holds row row_k, k < nt): position j takes the step's
  // pivot row p and the row it held goes where p was; nothing moves when p is
  // already at j (a pivot on the diagonal)
  if (w == 0) {
    int my_pos = -1, my_row = -1, nt = 0;
    for (int j = 0; j < nb; ++j) {
      const int p = sh.prow_of[j];
      const unsigned long long mp = __ballot(lane < nt && my_row == p);
      const int q = mp ? __shfl(my_pos, __ffsll((long long)mp) - 1) : p;
      if (lane == 0) piv[j] = q;
      if (q == j) continue;
      const unsigned long long mj = __ballot(lane < nt && my_pos == j);
      int kj = mj ? __ffsll((long long)mj) - 1 : -1;
      const int rj = mj ? __shfl(my_row, kj) : j;
      if (kj < 0) {
        kj = nt++;
        if (lane == kj) my_pos = j;
      }
      if (lane == kj) my_row = p;
      const unsigned long long mq = __ballot(lane < nt && my_pos == q);
      int kq = mq ? __ffsll((long long)mq) - 1 : -1;
      if (kq < 0) {
        kq = nt++;
        if (lane == kq) my_pos = q;
      }
      if (lane == kq) my_row = rj;
    }
    if (lane < nt) {
      sh.tpos[lane] = my_pos;
      sh.trow[lane] = my_row;
    }
    if (lane == 0) sh.ntouch = nt;
  }
  // rows to their own positions, then (after the barrier) the touched ones
  for (int e = t; e < nb * m; e += kLT) {
    const int j = e / m, i = e - j * m;
    A[(size_t)(k0 + j) * n + k0 + i] = pan[j * n + i];
  }
  __syncthreads();
  const int nt = sh.ntouch;
  for (int e = t; e < nt * nb; e += kLT) {
    const int k = e / nb, j = e - k * nb;
    A[(size_t)(k0 + j) * n + k0 + sh.tpos[k]] = pan[j * n + sh.trow[k]];
  }
  for (int e = t; e < nb * nb; e += kLT) {
    const int i = e / nb, j = e - i * nb;
    L11[i * NB + j] = pan[j * n + sh.prow_of[i]];
  }
  __syncthreads();
}

// Back substitution U X = Y for the 9 right-hand sides (columns n .. n + 8),
// in LDS (lds, 9 n complex) when rhs_lds, else in place; then the pivot
// growth ratio and the singular flag of the system.
__device__ void back_substitute(double2* __restrict__ A, int n, int rhs_lds, double2* lds,
                                double2* xs, const LuShared& sh, double* ratio, int* singular) {
  const int t = threadIdx.x;
  double2* Y = A + (size_t)n * n;
  double2* ys = rhs_lds ? lds : Y;
  if (rhs_lds) {
    for (int e = t; e < 9 * n; e += kLT) ys[e] = Y[e];
    __syncthreads();
  }
  for (int j = n - 1; j >= 0; --j) {
    if (t < 9) {
      const double2 dgn = A[(size_t)j * n + j];
      const double dd = dgn.x * dgn.x + dgn.y * dgn.y;
      const double2 inv = dd > 0.0 ? make_double2(dgn.x / dd, -dgn.y / dd) : make_double2(0.0, 0.0);
      const double2 x = zmul(ys[(size_t)t * n + j], inv);
      ys[(size_t)t * n + j] = x;
      xs[t] = x;
    }
    __syncthreads();
    const double2* uc = A + (size_t)j * n;
    for (int e = t; e < 9 * j; e += kLT) {
      const int q = e / j, i = e - q * j;
      zfms(ys[(size_t)q * n + i], uc[i], xs[q]);
    }
    __syncthreads();
  }
  if (rhs_lds)
    for (int e = t; e < 9 * n; e += kLT) Y[e] = ys[e];
  if (t == 0) {
    if (ratio) ratio[blockIdx.x] = sh.umin > 0.0 ? sh.umax / sh.umin : INFINITY;
    if (singular) singular[blockIdx.x] = sh.sing;
  }
}

// Back substitution U X = Y for the 9 right-hand sides (columns n .. n + 8)
// in LDS (ys, 9 n complex), 16 columns of U at a time from the bottom: the
// block's own rows by 9 threads (one per right-hand side, no barrier between
// its 16 steps), then the rows above it updated with the block's 16 columns of
// U (one thread per row, 16 independent loads) -- one round trip to memory
// per 16 columns.  Then the pivot growth ratio and the singular flag.
__device__ void back_substitute_blocked(double2* __restrict__ A, int n, double2* ys,
                                        const LuShared& sh, double* ratio, int* singular) {
  __shared__ double2 xb[9 * 16];
  __shared__ double2 ub[16 * 16];  // the diagonal block, column-major
  const int t = threadIdx.x;
  double2* Y = A + (size_t)n * n;
  for (int e = t; e < 9 * n; e += kLT) ys[e] = Y[e];
  __syncthreads();
#pragma unroll 1
  for (int j0 = (n - 1) / 16 * 16; j0 >= 0; j0 -= 16) {
    const int nbk = min(16, n - j0);
    if (t < 256) {
      const int r = t / 16, i = t % 16;
      if (r < nbk && i < nbk) ub[t] = A[(size_t)(j0 + r) * n + j0 + i];
    }
    __syncthreads();
    if (t < 9) {
      double2* y = ys + (size_t)t * n;
#pragma unroll 1
      for (int r = nbk - 1; r >= 0; --r) {
        const double2* uc = ub + r * 16;
        const double2 dgn = uc[r];
        const double dd = dgn.x * dgn.x + dgn.y * dgn.y;
        const double2 inv =
            dd > 0.0 ? make_double2(dgn.x / dd, -dgn.y / dd) : make_double2(0.0, 0.0);
        const double2 x = zmul(y[j0 + r], inv);
        y[j0 + r] = x;
        xb[t * 16 + r] = x;
        for (int i = 0; i < r; ++i) zfms(y[j0 + i], uc[i], x);
      }
      for (int r = nbk; r < 16; ++r) xb[t * 16 + r] = make_double2(0.0, 0.0);
    }
    __syncthreads();
#pragma unroll 1
    for (int i = t; i < j0; i += kLT) {
      double2 u[16];
#pragma unroll
      for (int r = 0; r < 16; ++r) u[r] = A[(size_t)(j0 + min(r, nbk - 1)) * n + i];
#pragma unroll 1
      for (int q = 0; q < 9; ++q) {
        double2 acc = ys[(size_t)q * n + i];
#pragma unroll
        for (int r = 0; r < 16; ++r) zfms(acc, u[r], xb[q * 16 + r]);  // zeros past nbk
        ys[(size_t)q * n + i] = acc;
      }
    }
    __syncthreads();
  }
  for (int e = t; e < 9 * n; e += kLT) Y[e] = ys[e];
  if (t == 0) {
    if (ratio) ratio[blockIdx.x] = sh.umin > 0.0 ? sh.umax / sh.umin : INFINITY;
    if (singular) singular[blockIdx.x] = sh.sing;
  }
}

template <int NB>
__global__ __launch_bounds__(kLT) void k_res_lu(double2* __restrict__ Aall, size_t stride, int n,
                                               int nc, int rhs_lds, double* __restrict__ ratio,
                                               int* __restrict__ singular) {
  extern __shared__ double2 lds[];
  __shared__ LuShared sh;
  __shared__ int piv[NB];
  __shared__ double2 L11[NB * NB];
  __shared__ double2 xs[9];
  double2* A = Aall + (size_t)blockIdx.x * stride;
  double2* pan = lds;  // panel: column j at pan + j * n (local rows)
  double2* U = lds;    // then U12: column c - c0 at U + (c - c0) * NB
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  lu_shared_init(sh);
  for (int k0 = 0; k0 < n; k0 += NB) {
    const int nb = min(NB, n - k0);
    factor_panel<NB>(A, n, k0, nb, pan, piv, L11, sh);
    // swaps and U12 = L11^-1 A12 for the columns right of the panel
    const int c0 = k0 + nb;
#pragma unroll 1
    for (int c = c0 + t; c < nc; c += kLT) {
      // keeps the L11 reads inside the loop (hoisting all NB^2 of them spills)
      asm volatile("" ::: "memory");
      double2* col = A + (size_t)c * n;
      for (int j = 0; j < nb; ++j) {
        const int p = k0 + piv[j];
        if (p != k0 + j) {
          const double2 a = col[k0 + j];
          col[k0 + j] = col[p];
          col[p] = a;
        }
      }
      double2 x[NB];
#pragma unroll
      for (int j = 0; j < NB; ++j) x[j] = j < nb ? col[k0 + j] : make_double2(0.0, 0.0);
#pragma unroll
      for (int i = 1; i < NB; ++i) {
#pragma unroll
        for (int j = 0; j < NB; ++j)
          if (j < i) zfms(x[i], L11[i * NB + j], x[j]);
        // x[i] is final here and the next row's L11 reads stay below: one
        // row of L11 live at a time (all NB^2 at once spills)
        asm volatile("" : "+v"(x[i].x), "+v"(x[i].y) : : "memory");
      }
#pragma unroll
      for (int j = 0; j < NB; ++j) {
        if (j < nb) col[k0 + j] = x[j];
        U[(size_t)(c - c0) * NB + j] = x[j];
      }
    }
    __syncthreads();
    // A22 -= L21 U12: lanes along rows (contiguous), waves along columns
#pragma unroll 1
    for (int ib = c0 + lane; ib < n; ib += 64) {
      double2 l[NB];
#pragma unroll
      for (int j = 0; j < NB; ++j)
        l[j] = j < nb ? A[(size_t)(k0 + j) * n + ib] : make_double2(0.0, 0.0);
      constexpr int CU = 4;  // columns in flight per thread
      int c = c0 + w;
#pragma unroll 1
      for (; c + (CU - 1) * kLW < nc; c += CU * kLW) {
        double2 v[CU];
#pragma unroll
        for (int u = 0; u < CU; ++u) v[u] = A[(size_t)(c + u * kLW) * n + ib];
#pragma unroll
        for (int u = 0; u < CU; ++u) {
          const double2* uu = U + (size_t)(c + u * kLW - c0) * NB;
#pragma unroll
          for (int j = 0; j < NB; ++j) zfms(v[u], l[j], uu[j]);
        }
#pragma unroll
        for (int u = 0; u < CU; ++u) A[(size_t)(c + u * kLW) * n + ib] = v[u];
      }
#pragma unroll 1
      for (; c < nc; c += kLW) {
        double2 v = A[(size_t)c * n + ib];
        const double2* uu = U + (size_t)(c - c0) * NB;
#pragma unroll
        for (int j = 0; j < NB; ++j) zfms(v, l[j], uu[j]);
        A[(size_t)c * n + ib] = v;
      }
    }
    __syncthreads();
  }
  back_substitute(A, n, rhs_lds, lds, xs, sh, ratio, singular);
}

typedef double msc_v4d __attribute__((ext_vector_type(4)));

// The row exchanges pv[0 .. kact) (absolute rows: row kb + j <-> pv[j], in
// order) on the columns cs .. ce - 1, one thread per column (the same column
// to thread mapping as solve_u_block, so no barrier between the two).
__device__ void exchange_rows(double2* __restrict__ A, int n, int kb, int kact, const int* pv,
                              int cs, int ce) {
#pragma unroll 1
  for (int c = cs + (int)threadIdx.x; c < ce; c += kLT) {
    double2* col = A + (size_t)c * n;
    for (int j = 0; j < kact; ++j) {
      const int p = pv[j];
      if (p != kb + j) {
        const double2 a = col[kb + j];
        col[kb + j] = col[p];
        col[p] = a;
      }
    }
  }
}

// Rows r0 .. r0 + nbb - 1 (nbb <= 16) of U for the columns cs .. ce - 1, one
// thread per column: forward substitution with the unit lower 16 x 16 block
// L (LDS, L[i * ldl + j]; its rows past nbb hold zeros, which keep x's rows
// past nbb at zero with no branch around a load) on those rows of A, written
// back to A and, row-major with leading dimension cw (column c at U + c - cs),
// to U's 16 rows (zeros past nbb).
__device__ void solve_u_block(double2* __restrict__ A, int n, int r0, int nbb, const double2* L,
                              int ldl, double2* U, int cw, int cs, int ce) {
#pragma unroll 1
  for (int c = cs + (int)threadIdx.x; c < ce; c += kLT) {
    double2* col = A + (size_t)c * n;
    double2* u = U + (c - cs);
    double2 x[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const double2 v = col[r0 + min(i, nbb - 1)];
      x[i] = i < nbb ? v : make_double2(0.0, 0.0);
    }
#pragma unroll
    for (int i = 1; i < 16; ++i) {
#pragma unroll
      for (int j = 0; j < 16; ++j)
        if (j < i) zfms(x[i], L[(size_t)i * ldl + j], x[j]);
      // x[i] is final here and the next row's L reads stay below: one row
      // of the block live at a time (all of it at once spills)
      asm volatile("" : "+v"(x[i].x), "+v"(x[i].y) : : "memory");
    }
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      if (i < nbb) col[r0 + i] = x[i];
      u[(size_t)i * cw] = x[i];
    }
  }
}

// A's rows r0 .. re - 1, columns cs .. ce - 1 -= L W on the matrix cores, with
// L = A's columns kb .. kb + kact - 1 of those rows (kact <= KM) and W = rows
// 0 .. kact - 1 of U (LDS, leading dimension cw, column c at U + c - cs).
// 16 x 16 complex tiles in row-major order, each wave taking one contiguous
// run of them: a wave keeps its tile row of L in registers while it moves
// along the columns and reloads it only when its run enters the next row.  Complex products as four real chains
// of v_mfma_f64_16x16x4f64 (Lr Ur, Li Ui, Lr Ui, Li Ur; independent
// accumulators, so the chains interleave without waiting on each other),
// combined at the end.  Lane l (i = l % 16, q = l / 16) feeds L[i][4 s + q]
// and W[4 s + q][i] at step s and receives rows q + 4 v (v < 4) of column i
// of the product (the result layout of the instruction on gfx950); entries
// past kact, re or ce are fed as zeros and not written.  C is loaded two
// tiles ahead (8 KB per wave in flight under the chains).  The tile loop is
// wave-uniform (the instruction uses every lane).
__device__ __forceinline__ void ctile_load(const double2* __restrict__ A, int n, int ce, int row0,
                                           int col, double2 (&cv)[4]) {
  // clamped into the tile grid's last row / column: entries past n or ce are
  // loaded but never stored
  const double2* cc = A + (size_t)min(col, ce - 1) * n;
#pragma unroll
  for (int v = 0; v < 4; ++v) cv[v] = cc[min(row0 + 4 * v, n - 1)];
}

template <int KM>
__device__ void trailing_mfma(double2* __restrict__ A, int n, int kb, int kact,
                              const double2* __restrict__ U, int cw, int r0, int re, int cs,
                              int ce) {
  constexpr int KS = KM / 4;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int i16 = lane & 15, q = lane >> 4;
  const int tr = (re - r0 + 15) / 16, tcn = (ce - cs + 15) / 16, total = tr * tcn;
  const int ns = (kact + 3) / 4;
  const int t_beg = (int)((long)total * w / kLW), t_end = (int)((long)total * (w + 1) / kLW);
  int cur = -1;
  double lr[KS], li[KS];
  double2 c1[4], c2[4];
  if (t_beg < t_end)
    ctile_load(A, n, ce, r0 + (t_beg / tcn) * 16 + q, cs + (t_beg % tcn) * 16 + i16, c1);
  if (t_beg + 1 < t_end) {
    const int tt = t_beg + 1;
    ctile_load(A, n, ce, r0 + (tt / tcn) * 16 + q, cs + (tt % tcn) * 16 + i16, c2);
  }
#pragma unroll 1
  for (int tile = t_beg; tile < t_end; ++tile) {
    const int ti = tile / tcn, tj = tile - ti * tcn;
    if (ti != cur) {
      cur = ti;
      const int rowL = r0 + ti * 16 + i16;
      const bool rin = rowL < re;
#pragma unroll
      for (int s = 0; s < KS; ++s) {
        const int k = 4 * s + q;
        const double2 l = A[(size_t)(kb + min(k, kact - 1)) * n + min(rowL, n - 1)];
        const bool in = rin && k < kact;
        lr[s] = in ? l.x : 0.0;
        li[s] = in ? l.y : 0.0;
      }
    }
    const int nt = tile + 2;
    double2 c3[4];
    if (nt < t_end) ctile_load(A, n, ce, r0 + (nt / tcn) * 16 + q, cs + (nt % tcn) * 16 + i16, c3);
    const int colU = tj * 16 + i16;
    msc_v4d rr = {0.0, 0.0, 0.0, 0.0}, ii = rr, ri = rr, ir = rr;
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      if (s < ns) {
        // colU < cw (ce - cs <= cw, a multiple of 16) and k < 4 ns: W's rows
        // up to the next multiple of 16, zero past kact
        const int k = 4 * s + q;
        const double2 u = U[(size_t)k * cw + colU];
        rr = __builtin_amdgcn_mfma_f64_16x16x4f64(lr[s], u.x, rr, 0, 0, 0);
        ii = __builtin_amdgcn_mfma_f64_16x16x4f64(li[s], u.y, ii, 0, 0, 0);
        ri = __builtin_amdgcn_mfma_f64_16x16x4f64(lr[s], u.y, ri, 0, 0, 0);
        ir = __builtin_amdgcn_mfma_f64_16x16x4f64(li[s], u.x, ir, 0, 0, 0);
      }
    }
    const int col = cs + tj * 16 + i16, row0 = r0 + ti * 16 + q;
    if (col < ce) {
      double2* cc = A + (size_t)col * n + row0;
#pragma unroll
      for (int v = 0; v < 4; ++v)
        if (row0 + 4 * v < re)
          cc[4 * v] = make_double2(c1[v].x - (rr[v] - ii[v]), c1[v].y - (ri[v] + ir[v]));
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      c1[v] = c2[v];
      c2[v] = c3[v];
    }
  }
}

// lds_entries: the dynamic LDS in complex entries, at least 16 n (the panel)
// and K * K + 16 K (the outer panel's L11 and a U chunk of 16 columns).
template <int S>
__global__ __launch_bounds__(kLT) void k_res_lu_blocked(double2* __restrict__ Aall, size_t stride,
                                                       int n, int nc, int lds_entries,
                                                       double* __restrict__ ratio,
                                                       int* __restrict__ singular) {
  constexpr int K = 16 * S;
  extern __shared__ double2 lds[];
  __shared__ LuShared sh;
  __shared__ int piv[16];
  __shared__ int pva[K];  // the outer panel's exchanges: row k0 + j <-> pva[j]
  __shared__ double2 L16[16 * 16];
  double2* A = Aall + (size_t)blockIdx.x * stride;
  const int t = threadIdx.x;
  const int cw = ((lds_entries - K * K) / K) & ~15;  // U chunk width of the far columns
  lu_shared_init(sh);
#pragma unroll 1
  for (int k0 = 0; k0 < n; k0 += K) {
    const int kact = min(K, n - k0);
#pragma unroll 1
    for (int p0 = k0; p0 < k0 + kact; p0 += 16) {
      const int nb = min(16, k0 + kact - p0);
      if (nb < 16 && t < 256) L16[t] = make_double2(0.0, 0.0);  // rows past nb stay zero
      factor_panel<16>(A, n, p0, nb, lds, piv, L16, sh);
      if (t < nb) pva[p0 - k0 + t] = p0 + piv[t];
      // this sub-panel's exchanges on the outer panel's earlier L columns
      for (int c = k0 + t; c < p0; c += kLT) {
        double2* col = A + (size_t)c * n;
        for (int j = 0; j < nb; ++j) {
          const int p = p0 + piv[j];
          if (p != p0 + j) {
            const double2 a = col[p0 + j];
            col[p0 + j] = col[p];
            col[p] = a;
          }
        }
      }
      __syncthreads();
      // the outer panel's later columns: 16 rows of U, then a 16-deep update
      const int cs = p0 + nb, ce = k0 + kact;
      if (cs < ce) {
        exchange_rows(A, n, p0, nb, pva + (p0 - k0), cs, ce);
        solve_u_block(A, n, p0, nb, L16, 16, lds, K, cs, ce);
        __syncthreads();
        trailing_mfma<16>(A, n, p0, nb, lds, K, cs, n, cs, ce);
        __syncthreads();
      }
    }
    // the columns right of the outer panel, in chunks of cw: kact rows of U,
    // then a kact-deep update of rows f0 .. n - 1
    const int f0 = k0 + kact;
    if (f0 < nc) {
      for (int e = t; e < K * K; e += kLT) {
        const int i = e / K, j = e - i * K;
        lds[e] = (i < kact && j < i) ? A[(size_t)(k0 + j) * n + k0 + i] : make_double2(0.0, 0.0);
      }
      __syncthreads();
#pragma unroll 1
      for (int cs = f0; cs < nc; cs += cw) {
        const int ce = min(cs + cw, nc);
        double2* U = lds + K * K;
        exchange_rows(A, n, k0, kact, pva, cs, ce);
        solve_u_block(A, n, k0, min(16, kact), lds, K, U, cw, cs, ce);
        // each further 16 rows of U: their rows of A less L U of the rows
        // above (matrix cores), then the 16 x 16 forward substitution
#pragma unroll 1
        for (int rb = 16; rb < kact; rb += 16) {
          __syncthreads();
          trailing_mfma<K>(A, n, k0, rb, U, cw, k0 + rb, k0 + min(rb + 16, kact), cs, ce);
          __syncthreads();
          solve_u_block(A, n, k0 + rb, min(16, kact - rb), lds + rb * K + rb, K, U + rb * cw,
                        cw, cs, ce);
        }
        __syncthreads();
        trailing_mfma<K>(A, n, k0, kact, U, cw, f0, n, cs, ce);
        __syncthreads();
      }
    }
  }
  back_substitute_blocked(A, n, lds, sh, ratio, singular);
}

// T_comp[b] (row-major 9x9) = T_loc sum_m X_m, X the 9 solution columns;
// psi_exc (optional) = X as (9N, 9) row-major.
__global__ __launch_bounds__(kRT) void k_res_tcomp(const double2* __restrict__ Aall, size_t stride,
                                                  const double2* __restrict__ tp, int n9,
                                                  double2* __restrict__ tcomp,
                                                  double2* __restrict__ psi_exc) {
  const int b = blockIdx.x;
  const double2* X = Aall + (size_t)b * stride + (size_t)n9 * n9;
  __shared__ double2 part[kRT / 64][81];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  // column sums s[k][q] = sum_m X[9m + k][q]: wave w, lane -> (k, q) pairs
  for (int kq = lane; kq < 81; kq += 64) {
    const int k = kq / 9, q = kq % 9;
    double2 s = make_double2(0.0, 0.0);
    for (int m = w; m < n9 / 9; m += kRT / 64) s = zadd(s, X[(size_t)q * n9 + 9 * m + k]);
    part[w][kq] = s;
  }
  if (psi_exc != nullptr)
    for (int e = t; e < n9 * 9; e += kRT) {
      const int r = e / 9, q = e % 9;
      psi_exc[(size_t)b * n9 * 9 + e] = X[(size_t)q * n9 + r];
    }
  __syncthreads();
  if (t < 81) {
    const int i = t / 9, q = t % 9;
    double2 s[9];
    for (int k = 0; k < 9; ++k) {
      s[k] = make_double2(0.0, 0.0);
      for (int ww = 0; ww < kRT / 64; ++ww) s[k] = zadd(s[k], part[ww][k * 9 + q]);
    }
    const double2* tl = tp + (size_t)b * 4;
    double2 v;
    if (i < 3) {
      v = zmul(tl[0], s[i]);
    } else if (i < 6) {
      v = zadd(zmul(zsub(tl[1], tl[2]), s[i]), zmul(tl[2], zadd(zadd(s[3], s[4]), s[5])));
    } else {
      v = zmul(tl[3], s[i]);
    }
    tcomp[(size_t)b * 81 + t] = v;
  }
}

// Neumann step: out = (I - A) in = in - A in over the n9 x n9 part, 9 columns;
// sum += out; norms[b] = (|out|^2, |sum|^2).
__global__ __launch_bounds__(kRT) void k_res_neumann(const double2* __restrict__ Aall,
                                                    size_t stride, int n9,
                                                    const double2* __restrict__ tin,
                                                    double2* __restrict__ tout,
                                                    double2* __restrict__ acc,
                                                    double* __restrict__ norms,
                                                    const int* __restrict__ active) {
  const int b = blockIdx.y;
  if (!active[b]) return;  // converged: its sum stays as it is
  const double2* A = Aall + (size_t)b * stride;
  const double2* x = tin + (size_t)b * n9 * 9;
  double2* y = tout + (size_t)b * n9 * 9;
  double2* s = acc + (size_t)b * n9 * 9;
  double nt = 0.0, ns = 0.0;
  for (int e = blockIdx.x * kRT + threadIdx.x; e < n9 * 9; e += gridDim.x * kRT) {
    const int q = e / n9, r = e - q * n9;
    double2 v = x[(size_t)q * n9 + r];
    for (int c = 0; c < n9; ++c) zfms(v, A[(size_t)c * n9 + r], x[(size_t)q * n9 + c]);
    y[e] = v;
    const double2 sv = zadd(s[e], v);
    s[e] = sv;
    nt += v.x * v.x + v.y * v.y;
    ns += sv.x * sv.x + sv.y * sv.y;
  }
  for (int off = 32; off > 0; off >>= 1) {
    nt += __shfl_xor(nt, off);
    ns += __shfl_xor(ns, off);
  }
  if ((threadIdx.x & 63) == 0) {
    atomicAdd(norms + 2 * b, nt);
    atomicAdd(norms + 2 * b + 1, ns);
  }
}

__global__ void k_zero_d(double* p, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0.0;
}

// The 9 incident patterns at the sub-cell centres (resonance_tmatrix.py:526-593),
// column-major (9N, 9): uniform displacement e_p (columns 0-2); linear
// displacement eps(e_al) (x - x_c) with uniform Voigt strain e_al (3-8);
// all times exp(i k khat.x).
std::vector<double2> incident_patterns(int n, double a, double k, const double* kh) {
  const int N = n * n * n, n9 = 9 * N;
  std::vector<double> coord(n);
  const double as = a / n;
  for (int i = 0; i < n; ++i)
    coord[i] = n == 1 ? 0.0 : (-a + as) + i * ((a - as) - (-a + as)) / (n - 1);
  if (n > 1) coord[n - 1] = a - as;
  std::vector<double> cx((size_t)N * 3);
  double mean[3] = {0.0, 0.0, 0.0};
  for (int m = 0; m < N; ++m) {
    const int i0 = m / (n * n), i1 = (m / n) % n, i2 = m % n;
    cx[3 * m] = coord[i0];
    cx[3 * m + 1] = coord[i1];
    cx[3 * m + 2] = coord[i2];
    for (int d = 0; d < 3; ++d) mean[d] += cx[3 * m + d];
  }
  for (int d = 0; d < 3; ++d) mean[d] /= N;
  const int VP[6][2] = {{0, 0}, {1, 1}, {2, 2}, {1, 2}, {0, 2}, {0, 1}};
  std::vector<double2> psi((size_t)n9 * 9, make_double2(0.0, 0.0));
  for (int m = 0; m < N; ++m) {
    const double ph = k * (cx[3 * m] * kh[0] + cx[3 * m + 1] * kh[1] + cx[3 * m + 2] * kh[2]);
    const double2 e = make_double2(std::cos(ph), std::sin(ph));
    const double dx[3] = {cx[3 * m] - mean[0], cx[3 * m + 1] - mean[1], cx[3 * m + 2] - mean[2]};
    for (int p = 0; p < 3; ++p) psi[(size_t)p * n9 + 9 * m + p] = e;
    for (int al = 0; al < 6; ++al) {
      const int p = VP[al][0], q = VP[al][1];
      double u[3] = {0.0, 0.0, 0.0};
      if (p == q) {
        u[p] = dx[p];
      } else {
        u[p] = 0.5 * dx[q];
        u[q] = 0.5 * dx[p];
      }
      double2* col = psi.data() + (size_t)(3 + al) * n9 + 9 * m;
      for (int i = 0; i < 3; ++i) col[i] = make_double2(u[i] * e.x, u[i] * e.y);
      col[3 + al] = e;
    }
  }
  return psi;
}

// The same 9 incident patterns as incident_patterns(), written on the device
// (column-major (9N, 9)): thread m owns rows 9m..9m+8 of every column, zeros
// included, so the output needs no clearing.  The sub-cell coordinates and
// their mean follow the host's arithmetic step by step; only cos/sin of the
// phase come from the device's math library.
__global__ __launch_bounds__(64) void k_res_incident(double2* __restrict__ psi, int n, double a,
                                                    double k, double kh0, double kh1,
                                                    double kh2) {
  const int N = n * n * n, n9 = 9 * N;
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= N) return;
  const double as = a / n;
  auto coord = [&](int i) {
    if (n == 1) return 0.0;
    if (i == n - 1) return a - as;
    return (-a + as) + i * ((a - as) - (-a + as)) / (n - 1);
  };
  double mean[3] = {0.0, 0.0, 0.0};
  for (int q = 0; q < N; ++q) {
    mean[0] += coord(q / (n * n));
    mean[1] += coord((q / n) % n);
    mean[2] += coord(q % n);
  }
  for (int d = 0; d < 3; ++d) mean[d] /= N;
  const double c0 = coord(m / (n * n)), c1 = coord((m / n) % n), c2 = coord(m % n);
  const double ph = k * (c0 * kh0 + c1 * kh1 + c2 * kh2);
  double sn, cs;
  sincos(ph, &sn, &cs);
  const double2 e = make_double2(cs, sn);
  const double dx[3] = {c0 - mean[0], c1 - mean[1], c2 - mean[2]};
  const int VP[6][2] = {{0, 0}, {1, 1}, {2, 2}, {1, 2}, {0, 2}, {0, 1}};
  for (int col = 0; col < 9; ++col) {
    double2 v[9];
    for (int i = 0; i < 9; ++i) v[i] = make_double2(0.0, 0.0);
    if (col < 3) {
      v[col] = e;
    } else {
      const int al = col - 3, p = VP[al][0], q = VP[al][1];
      double u[3] = {0.0, 0.0, 0.0};
      if (p == q) {
        u[p] = dx[p];
      } else {
        u[p] = 0.5 * dx[q];
        u[q] = 0.5 * dx[p];
      }
      for (int i = 0; i < 3; ++i) v[i] = make_double2(u[i] * e.x, u[i] * e.y);
      v[3 + al] = e;
    }
    double2* out = psi + (size_t)col * n9 + 9 * m;
    for (int i = 0; i < 9; ++i) out[i] = v[i];
  }
}

// Dynamic LDS of the elimination kernel for panel width nb.
size_t lu_lds(int nb, int n, int nc) { return (size_t)nb * std::max(n, nc) * sizeof(double2); }
constexpr size_t kLuLdsCap = 150 * 1024;  // leaves room for the static arrays (NB = 16)

template <int S>
int launch_lu_blocked(double2* A, size_t stride, int n, int nc, int count, double* ratio,
                      int* sing, hipStream_t st) {
  constexpr int K = 16 * S;
  const int entries = std::max(16 * n, K * K + 16 * K);
  const size_t lds = (size_t)entries * sizeof(double2);
  if (lds > kLuLdsCap) {
    set_error("resonance system too large for the blocked elimination");
    return kUnsupported;
  }
  MSC_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_res_lu_blocked<S>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(k_res_lu_blocked<S>, dim3(count), dim3(kLT), lds, st, A, stride, n, nc,
                     entries, ratio, sing);
  MSC_LAUNCH_CHECK();
  return kOk;
}

template <int NB>
int launch_lu(double2* A, size_t stride, int n, int nc, int count, double* ratio, int* sing,
              hipStream_t st) {
  // factor_panel keeps one bit per row of a thread: at most 32 rows each
  MSC_REQUIRE(n <= 32 * kLT, "resonance system too large for the elimination");
  if constexpr (NB == 16) {
    // outer panels of 32 columns (64 measured slower at n_sub = 4: its
    // 64-row forward substitutions run on fewer threads per column chunk)
    return launch_lu_blocked<2>(A, stride, n, nc, count, ratio, sing, st);
  } else {
    const size_t lds = lu_lds(NB, n, nc);
    const int rhs_lds = (size_t)9 * n * sizeof(double2) <= lds ? 1 : 0;
    MSC_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_res_lu<NB>),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(k_res_lu<NB>, dim3(count), dim3(kLT), lds, st, A, stride, n, nc, rhs_lds,
                       ratio, sing);
    MSC_LAUNCH_CHECK();
    return kOk;
  }
}

}  // namespace
}  // namespace msc

using namespace msc;

extern "C" int msc_resonance_tmatrices(double omega, double a, double alpha, double beta,
                                       double rho, int n_sub, const double* khat3, int wave_s,
                                       int count, const double* dlambda, const double* dmu,
                                       const double* drho, int neumann_order, double neumann_tol,
                                       void* tcomp, void* psi_exc, double* pivot_ratio,
                                       int* neumann_iters, void* stream) {
  set_error("");
  MSC_REQUIRE(omega > 0.0 && a > 0.0 && alpha > 0.0 && beta > 0.0 && rho > 0.0,
              "omega, a, alpha, beta, rho must be positive");
  MSC_REQUIRE(n_sub >= 1, "n_sub must be >= 1");
  MSC_REQUIRE(count >= 0, "count must be >= 0");
  MSC_REQUIRE(count == 0 || (dlambda && dmu && drho && tcomp), "null contrast or output pointer");
  MSC_REQUIRE(khat3 != nullptr, "khat must not be NULL");
  MSC_REQUIRE(wave_s == 0 || wave_s == 1, "wave_type must be 0 (P) or 1 (S)");
  MSC_REQUIRE(neumann_order >= 0, "neumann_order must be >= 0");
  const double kn = std::sqrt(khat3[0] * khat3[0] + khat3[1] * khat3[1] + khat3[2] * khat3[2]);
  MSC_REQUIRE(kn > 0.0, "khat must be non-zero");
  const int n = n_sub, N = n * n * n, n9 = 9 * N, nc = n9 + 9;
  int nb = 16;
  while (nb > 1 && lu_lds(nb, n9, nc) > kLuLdsCap) nb >>= 1;
  if (n_sub > 8 || lu_lds(nb, n9, nc) > kLuLdsCap) {
    set_error("n_sub too large for the resonance solver (at most 8)");
    return kUnsupported;
  }
  if (count == 0) return kOk;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const double as = a / n, kP = omega / alpha, kS = omega / beta;
  const double kh[3] = {khat3[0] / kn, khat3[1] / kn, khat3[2] / kn};
  const std::vector<double2> psi = incident_patterns(n, a, wave_s ? kS : kP, kh);
  double coef[8];
  int rc = msc_cube_coefficients(omega, as, alpha, beta, rho, coef);
  if (rc) return rc;
  const size_t stride = (size_t)n9 * nc;
  const size_t budget = (size_t)4 << 30;
  int chunk = (int)std::max<size_t>(
      1, std::min<size_t>(std::min<size_t>(count, 65535), budget / (stride * 16)));
  // whole rounds of one system per CU: the last round of a chunk with a few
  // systems left over would run the chip at a fraction of its width
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) == hipSuccess &&
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
      cus > 0 && chunk > cus)
    chunk = chunk / cus * cus;
  const int no = 2 * n - 1;
  double2 *dP = nullptr, *dpsi = nullptr, *dtp = nullptr, *dA = nullptr, *dt0 = nullptr,
          *dt1 = nullptr, *dacc = nullptr;
  double *dratio = nullptr, *dnorm = nullptr;
  int *dsing = nullptr, *dact = nullptr;
  auto cleanup = [&]() {
    for (void* p : {(void*)dP, (void*)dpsi, (void*)dtp, (void*)dA, (void*)dt0, (void*)dt1,
                    (void*)dacc, (void*)dratio, (void*)dnorm, (void*)dsing, (void*)dact})
      if (p) (void)hipFree(p);
  };
#define RES_CHECK(expr)                                                   \
  do {                                                                    \
    hipError_t _e = (expr);                                               \
    if (_e != hipSuccess) {                                               \
      set_error(std::string(#expr) + ": " + hipGetErrorString(_e));       \
      cleanup();                                                          \
      return kHip;                                                        \
    }                                                                     \
  } while (0)
  RES_CHECK(hipMalloc(&dP, (size_t)no * no * no * 81 * sizeof(double2)));
  RES_CHECK(hipMalloc(&dpsi, psi.size() * sizeof(double2)));
  RES_CHECK(hipMalloc(&dtp, (size_t)count * 4 * sizeof(double2)));
  RES_CHECK(hipMalloc(&dA, (size_t)chunk * stride * sizeof(double2)));
  RES_CHECK(hipMalloc(&dratio, (size_t)chunk * sizeof(double)));
  RES_CHECK(hipMalloc(&dsing, (size_t)chunk * sizeof(int)));
  if (neumann_order > 0) {
    RES_CHECK(hipMalloc(&dt0, (size_t)chunk * n9 * 9 * sizeof(double2)));
    RES_CHECK(hipMalloc(&dt1, (size_t)chunk * n9 * 9 * sizeof(double2)));
    RES_CHECK(hipMalloc(&dacc, (size_t)chunk * n9 * 9 * sizeof(double2)));
    RES_CHECK(hipMalloc(&dnorm, (size_t)chunk * 2 * sizeof(double)));
    RES_CHECK(hipMalloc(&dact, (size_t)chunk * sizeof(int)));
  }
  RES_CHECK(hipMemcpyAsync(dpsi, psi.data(), psi.size() * sizeof(double2), hipMemcpyHostToDevice,
                           st));
  hipLaunchKernelGGL(k_res_offsets, dim3((no * no * no + 63) / 64), dim3(64), 0, st, dP, n,
                     2.0 * as, kP, kS, 1.0 / (4.0 * M_PI * rho * omega * omega));
  RES_CHECK(hipGetLastError());
  rc = tparams_batch(dlambda, dmu, drho, (size_t)count, omega, coef, 8.0 * as * as * as, dtp, st);
  if (rc) {
    cleanup();
    return rc;
  }
  std::vector<double> hratio(chunk);
  std::vector<int> hsing(chunk);
  std::vector<double> hnorm(2 * (size_t)chunk);
  double2* out = static_cast<double2*>(tcomp);
  double2* pex = static_cast<double2*>(psi_exc);
  for (int b0 = 0; b0 < count; b0 += chunk) {
    const int cb = std::min(chunk, count - b0);
    const int gx = (int)std::min<size_t>((stride + kRT - 1) / kRT, 2048);
    hipLaunchKernelGGL(k_res_assemble, dim3(gx, cb), dim3(kRT), 0, st, dA, stride, dP,
                       dtp + (size_t)b0 * 4, dpsi, n);
    RES_CHECK(hipGetLastError());
    if (neumann_order == 0) {
      switch (nb) {
        case 16: rc = launch_lu<16>(dA, stride, n9, nc, cb, dratio, dsing, st); break;
        case 8: rc = launch_lu<8>(dA, stride, n9, nc, cb, dratio, dsing, st); break;
        case 4: rc = launch_lu<4>(dA, stride, n9, nc, cb, dratio, dsing, st); break;
        case 2: rc = launch_lu<2>(dA, stride, n9, nc, cb, dratio, dsing, st); break;
        default: rc = launch_lu<1>(dA, stride, n9, nc, cb, dratio, dsing, st); break;
      }
      if (rc) {
        cleanup();
        return rc;
      }
      RES_CHECK(hipMemcpyAsync(hratio.data(), dratio, cb * sizeof(double), hipMemcpyDeviceToHost,
                               st));
      RES_CHECK(hipMemcpyAsync(hsing.data(), dsing, cb * sizeof(int), hipMemcpyDeviceToHost, st));
      RES_CHECK(hipStreamSynchronize(st));
      for (int i = 0; i < cb; ++i) {
        if (hsing[i]) {
          set_error("the internal Foldy-Lax matrix is singular");
          cleanup();
          return kInvalid;
        }
        if (pivot_ratio) pivot_ratio[b0 + i] = hratio[i];
        if (neumann_iters) neumann_iters[b0 + i] = -1;
      }
    } else {
      // psi = sum_k B^k psi_inc: the incident columns are the first term
      for (int b = 0; b < cb; ++b) {
        RES_CHECK(hipMemcpyAsync(dt0 + (size_t)b * n9 * 9, dpsi, (size_t)n9 * 9 * sizeof(double2),
                                 hipMemcpyDeviceToDevice, st));
        RES_CHECK(hipMemcpyAsync(dacc + (size_t)b * n9 * 9, dpsi,
                                 (size_t)n9 * 9 * sizeof(double2), hipMemcpyDeviceToDevice, st));
      }
      std::vector<int> conv(cb, -1), act(cb, 1);
      double2 *tin = dt0, *tout = dt1;
      for (int k = 1; k <= neumann_order; ++k) {
        RES_CHECK(hipMemcpyAsync(dact, act.data(), cb * sizeof(int), hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_zero_d, dim3(1), dim3(256), 0, st, dnorm, 2 * cb);
        const int gxn = (n9 * 9 + kRT - 1) / kRT;
        hipLaunchKernelGGL(k_res_neumann, dim3(gxn, cb), dim3(kRT), 0, st, dA, stride, n9, tin,
                           tout, dacc, dnorm, dact);
        RES_CHECK(hipGetLastError());
        RES_CHECK(hipMemcpyAsync(hnorm.data(), dnorm, 2 * cb * sizeof(double),
                                 hipMemcpyDeviceToHost, st));
        RES_CHECK(hipStreamSynchronize(st));
        bool all = true;
        for (int i = 0; i < cb; ++i) {
          const double rel = std::sqrt(hnorm[2 * i]) / (std::sqrt(hnorm[2 * i + 1]) + 1e-300);
          if (conv[i] < 0 && act[i] && rel < neumann_tol) conv[i] = k;
          if (conv[i] >= 0) act[i] = 0;
          all = all && conv[i] >= 0;
        }
        std::swap(tin, tout);
        if (all) break;
      }
      // the Neumann sums replace the solution columns
      for (int b = 0; b < cb; ++b)
        RES_CHECK(hipMemcpyAsync(dA + (size_t)b * stride + (size_t)n9 * n9,
                                 dacc + (size_t)b * n9 * 9, (size_t)n9 * 9 * sizeof(double2),
                                 hipMemcpyDeviceToDevice, st));
      for (int i = 0; i < cb; ++i) {
        if (pivot_ratio) pivot_ratio[b0 + i] = NAN;
        if (neumann_iters) neumann_iters[b0 + i] = conv[i];
      }
    }
    hipLaunchKernelGGL(k_res_tcomp, dim3(cb), dim3(kRT), 0, st, dA, stride, dtp + (size_t)b0 * 4,
                       n9, out + (size_t)b0 * 81, pex ? pex + (size_t)b0 * n9 * 9 : nullptr);
    RES_CHECK(hipGetLastError());
  }
  RES_CHECK(hipStreamSynchronize(st));
#undef RES_CHECK
  cleanup();
  return kOk;
}

extern "C" int msc_resonance_system(double omega, double a, double alpha, double beta,
                                    double rho, int n_sub, double dlambda, double dmu,
                                    double drho, void* A_out, void* stream) {
  set_error("");
  MSC_REQUIRE(omega > 0.0 && a > 0.0 && alpha > 0.0 && beta > 0.0 && rho > 0.0,
              "omega, a, alpha, beta, rho must be positive");
  MSC_REQUIRE(n_sub >= 1 && n_sub <= 8, "n_sub must be in 1..8");
  MSC_REQUIRE(A_out != nullptr, "A_out must not be NULL");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int n = n_sub, N = n * n * n, n9 = 9 * N, nc = n9 + 9, no = 2 * n - 1;
  const double as = a / n, kP = omega / alpha, kS = omega / beta;
  double coef[8];
  int rc = msc_cube_coefficients(omega, as, alpha, beta, rho, coef);
  if (rc) return rc;
  const size_t stride = (size_t)n9 * nc;
  double2 *dP = nullptr, *dtp = nullptr, *dA = nullptr, *dpsi = nullptr;
  double* dc = nullptr;
  auto cleanup = [&]() {
    for (void* p : {(void*)dP, (void*)dtp, (void*)dA, (void*)dpsi, (void*)dc})
      if (p) (void)hipFree(p);
  };
#define SYS_CHECK(expr)                                                   \
  do {                                                                    \
    hipError_t _e = (expr);                                               \
    if (_e != hipSuccess) {                                               \
      set_error(std::string(#expr) + ": " + hipGetErrorString(_e));       \
      cleanup();                                                          \
      return kHip;                                                        \
    }                                                                     \
  } while (0)
  SYS_CHECK(hipMalloc(&dP, (size_t)no * no * no * 81 * sizeof(double2)));
  SYS_CHECK(hipMalloc(&dtp, 4 * sizeof(double2)));
  SYS_CHECK(hipMalloc(&dA, stride * sizeof(double2)));
  SYS_CHECK(hipMalloc(&dpsi, (size_t)n9 * 9 * sizeof(double2)));
  SYS_CHECK(hipMalloc(&dc, 3 * sizeof(double)));
  const double hc[3] = {dlambda, dmu, drho};
  SYS_CHECK(hipMemcpyAsync(dc, hc, sizeof(hc), hipMemcpyHostToDevice, st));
  // the right-hand-side columns are not part of A: zeros
  SYS_CHECK(hipMemsetAsync(dpsi, 0, (size_t)n9 * 9 * sizeof(double2), st));
  hipLaunchKernelGGL(k_res_offsets, dim3((no * no * no + 63) / 64), dim3(64), 0, st, dP, n,
                     2.0 * as, kP, kS, 1.0 / (4.0 * M_PI * rho * omega * omega));
  SYS_CHECK(hipGetLastError());
  rc = tparams_batch(dc, dc + 1, dc + 2, 1, omega, coef, 8.0 * as * as * as, dtp, st);
  if (rc) {
    cleanup();
    return rc;
  }
  const int gx = (int)std::min<size_t>((stride + kRT - 1) / kRT, 2048);
  hipLaunchKernelGGL(k_res_assemble, dim3(gx, 1), dim3(kRT), 0, st, dA, stride, dP, dtp, dpsi, n);
  SYS_CHECK(hipGetLastError());
  SYS_CHECK(hipMemcpyAsync(A_out, dA, (size_t)n9 * n9 * sizeof(double2), hipMemcpyDeviceToDevice,
                           st));
  SYS_CHECK(hipStreamSynchronize(st));
#undef SYS_CHECK
  cleanup();
  return kOk;
}

// ─────────────────────────────────────────────────────────────────────────────
// No-allocation variant (msc_resonance_tmatrices_noalloc).
//
// Its device workspace is `scratch`, a file-scope device array of
// kScratchBytes that every GPU the library runs on holds until the process
// ends.  Host code never passes `scratch` itself (its host-side name is only a
// handle): on the first call on a device, with that device current, the
// entry point asks the runtime for the array's address on that device
// (hipGetSymbolAddress) and keeps it per device; kernels, copies and fills
// take that address.  Calls that overlap on one GPU share the array and race;
// every call writes what it reads before reading it, so nothing of an earlier
// call's contents is used.
// ─────────────────────────────────────────────────────────────────────────────

namespace {
constexpr size_t kScratchBytes = size_t(64) << 20;
constexpr int kNoAllocMaxChunk = 64;
constexpr int kMaxDevices = 64;
}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

std::mutex g_scratch_mutex;
void* g_scratch_addr[kMaxDevices] = {};

// Device address of `scratch` on the current device (looked up once per device).
int scratch_address(void** out) {
  int dev = -1;
  MSC_HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) {
    msc::set_error("device ordinal out of range for the no-allocation scratch");
    return msc::kUnsupported;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (g_scratch_addr[dev] == nullptr) {
    void* p = nullptr;
    MSC_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (p == nullptr) {
      msc::set_error("hipGetSymbolAddress returned no address for the scratch array");
      return msc::kHip;
    }
    g_scratch_addr[dev] = p;
  }
  *out = g_scratch_addr[dev];
  return msc::kOk;
}

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

// Byte offsets of the workspace arrays inside the scratch for `chunk` systems.
struct ScratchLayout {
  size_t P, psi, tp, ratio, sing, A, t0, t1, acc, norm, act, total;
};

ScratchLayout scratch_layout(int n, int chunk, bool neumann) {
  const size_t N = (size_t)n * n * n, n9 = 9 * N, nc = n9 + 9, no = 2 * n - 1;
  const size_t c = (size_t)chunk;
  ScratchLayout L{};
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t o = off;
    off += align256(bytes);
    return o;
  };
  L.P = take(no * no * no * 81 * sizeof(double2));
  L.psi = take(n9 * 9 * sizeof(double2));
  L.tp = take(c * 4 * sizeof(double2));
  L.ratio = take(c * sizeof(double));
  L.sing = take(c * sizeof(int));
  L.A = take(c * n9 * nc * sizeof(double2));
  if (neumann) {
    L.t0 = take(c * n9 * 9 * sizeof(double2));
    L.t1 = take(c * n9 * 9 * sizeof(double2));
    L.acc = take(c * n9 * 9 * sizeof(double2));
    L.norm = take(c * 2 * sizeof(double));
    L.act = take(c * sizeof(int));
  }
  L.total = off;
  return L;
}

}  // namespace

extern "C" int msc_resonance_tmatrices_noalloc(double omega, double a, double alpha, double beta,
                                               double rho, int n_sub, const double* khat3,
                                               int wave_s, int count, const double* dlambda,
                                               const double* dmu, const double* drho,
                                               int neumann_order, double neumann_tol,
                                               void* tcomp, void* psi_exc, double* pivot_ratio,
                                               int* neumann_iters, void* stream) {
  set_error("");
  MSC_REQUIRE(omega > 0.0 && a > 0.0 && alpha > 0.0 && beta > 0.0 && rho > 0.0,
              "omega, a, alpha, beta, rho must be positive");
  MSC_REQUIRE(n_sub >= 1, "n_sub must be >= 1");
  MSC_REQUIRE(count >= 0, "count must be >= 0");
  MSC_REQUIRE(count == 0 || (dlambda && dmu && drho && tcomp), "null contrast or output pointer");
  MSC_REQUIRE(khat3 != nullptr, "khat must not be NULL");
  MSC_REQUIRE(wave_s == 0 || wave_s == 1, "wave_type must be 0 (P) or 1 (S)");
  MSC_REQUIRE(neumann_order >= 0, "neumann_order must be >= 0");
  const double kn = std::sqrt(khat3[0] * khat3[0] + khat3[1] * khat3[1] + khat3[2] * khat3[2]);
  MSC_REQUIRE(kn > 0.0, "khat must be non-zero");
  const int n = n_sub, N = n * n * n, n9 = 9 * N, nc = n9 + 9;
  int nb = 16;
  while (nb > 1 && lu_lds(nb, n9, nc) > kLuLdsCap) nb >>= 1;
  if (n_sub > 8 || lu_lds(nb, n9, nc) > kLuLdsCap) {
    set_error("n_sub too large for the resonance solver (at most 8)");
    return kUnsupported;
  }
  const bool neumann = neumann_order > 0;
  if (scratch_layout(n, 1, neumann).total > kScratchBytes) {
    set_error("n_sub = " + std::to_string(n_sub) +
              " needs more than the 64 MiB no-allocation scratch for one cube (at most n_sub = 6); "
              "use msc_resonance_tmatrices");
    return kUnsupported;
  }
  if (count == 0) return kOk;
  int chunk = std::min(count, kNoAllocMaxChunk);
  while (chunk > 1 && scratch_layout(n, chunk, neumann).total > kScratchBytes) --chunk;
  const ScratchLayout L = scratch_layout(n, chunk, neumann);
  void* base = nullptr;
  int rc = scratch_address(&base);
  if (rc) return rc;
  unsigned char* sb = static_cast<unsigned char*>(base);
  double2* dP = reinterpret_cast<double2*>(sb + L.P);
  double2* dpsi = reinterpret_cast<double2*>(sb + L.psi);
  double2* dtp = reinterpret_cast<double2*>(sb + L.tp);
  double* dratio = reinterpret_cast<double*>(sb + L.ratio);
  int* dsing = reinterpret_cast<int*>(sb + L.sing);
  double2* dA = reinterpret_cast<double2*>(sb + L.A);
  double2* dt0 = neumann ? reinterpret_cast<double2*>(sb + L.t0) : nullptr;
  double2* dt1 = neumann ? reinterpret_cast<double2*>(sb + L.t1) : nullptr;
  double2* dacc = neumann ? reinterpret_cast<double2*>(sb + L.acc) : nullptr;
  double* dnorm = neumann ? reinterpret_cast<double*>(sb + L.norm) : nullptr;
  int* dact = neumann ? reinterpret_cast<int*>(sb + L.act) : nullptr;

  hipStream_t st = static_cast<hipStream_t>(stream);
  const double as = a / n, kP = omega / alpha, kS = omega / beta;
  double coef[8];
  rc = msc_cube_coefficients(omega, as, alpha, beta, rho, coef);
  if (rc) return rc;
  const size_t stride = (size_t)n9 * nc;
  const int no = 2 * n - 1;
  hipLaunchKernelGGL(k_res_incident, dim3((N + 63) / 64), dim3(64), 0, st, dpsi, n, a,
                     wave_s ? kS : kP, khat3[0] / kn, khat3[1] / kn, khat3[2] / kn);
  MSC_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_res_offsets, dim3((no * no * no + 63) / 64), dim3(64), 0, st, dP, n,
                     2.0 * as, kP, kS, 1.0 / (4.0 * M_PI * rho * omega * omega));
  MSC_LAUNCH_CHECK();
  double hratio[kNoAllocMaxChunk], hnorm[2 * kNoAllocMaxChunk];
  int hsing[kNoAllocMaxChunk], conv[kNoAllocMaxChunk], act[kNoAllocMaxChunk];
  double2* out = static_cast<double2*>(tcomp);
  double2* pex = static_cast<double2*>(psi_exc);
  for (int b0 = 0; b0 < count; b0 += chunk) {
    const int cb = std::min(chunk, count - b0);
    rc = tparams_batch(dlambda + b0, dmu + b0, drho + b0, (size_t)cb, omega, coef,
                       8.0 * as * as * as, dtp, st);
    if (rc) return rc;
    const int gx = (int)std::min<size_t>((stride + kRT - 1) / kRT, 2048);
    hipLaunchKernelGGL(k_res_assemble, dim3(gx, cb), dim3(kRT), 0, st, dA, stride, dP, dtp, dpsi,
                       n);
    MSC_LAUNCH_CHECK();
    if (!neumann) {
      switch (nb) {
        case 16: rc = launch_lu<16>(dA, stride, n9, nc, cb, dratio, dsing, st); break;
        case 8: rc = launch_lu<8>(dA, stride, n9, nc, cb, dratio, dsing, st); break;
        case 4: rc = launch_lu<4>(dA, stride, n9, nc, cb, dratio, dsing, st); break;
        case 2: rc = launch_lu<2>(dA, stride, n9, nc, cb, dratio, dsing, st); break;
        default: rc = launch_lu<1>(dA, stride, n9, nc, cb, dratio, dsing, st); break;
      }
      if (rc) return rc;
      MSC_HIP_CHECK(hipMemcpyAsync(hratio, dratio, cb * sizeof(double), hipMemcpyDeviceToHost, st));
      MSC_HIP_CHECK(hipMemcpyAsync(hsing, dsing, cb * sizeof(int), hipMemcpyDeviceToHost, st));
      MSC_HIP_CHECK(hipStreamSynchronize(st));
      for (int i = 0; i < cb; ++i) {
        if (hsing[i]) {
          set_error("the internal Foldy-Lax matrix is singular");
          return kInvalid;
        }
        if (pivot_ratio) pivot_ratio[b0 + i] = hratio[i];
        if (neumann_iters) neumann_iters[b0 + i] = -1;
      }
    } else {
      for (int b = 0; b < cb; ++b) {
        MSC_HIP_CHECK(hipMemcpyAsync(dt0 + (size_t)b * n9 * 9, dpsi,
                                     (size_t)n9 * 9 * sizeof(double2), hipMemcpyDeviceToDevice, st));
        MSC_HIP_CHECK(hipMemcpyAsync(dacc + (size_t)b * n9 * 9, dpsi,
                                     (size_t)n9 * 9 * sizeof(double2), hipMemcpyDeviceToDevice, st));
      }
      for (int i = 0; i < cb; ++i) {
        conv[i] = -1;
        act[i] = 1;
      }
      double2 *tin = dt0, *tout = dt1;
      for (int k = 1; k <= neumann_order; ++k) {
        MSC_HIP_CHECK(hipMemcpyAsync(dact, act, cb * sizeof(int), hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_zero_d, dim3(1), dim3(256), 0, st, dnorm, 2 * cb);
        const int gxn = (n9 * 9 + kRT - 1) / kRT;
        hipLaunchKernelGGL(k_res_neumann, dim3(gxn, cb), dim3(kRT), 0, st, dA, stride, n9, tin,
                           tout, dacc, dnorm, dact);
        MSC_LAUNCH_CHECK();
        MSC_HIP_CHECK(hipMemcpyAsync(hnorm, dnorm, 2 * cb * sizeof(double), hipMemcpyDeviceToHost,
                                     st));
        MSC_HIP_CHECK(hipStreamSynchronize(st));
        bool all = true;
        for (int i = 0; i < cb; ++i) {
          const double rel = std::sqrt(hnorm[2 * i]) / (std::sqrt(hnorm[2 * i + 1]) + 1e-300);
          if (conv[i] < 0 && act[i] && rel < neumann_tol) conv[i] = k;
          if (conv[i] >= 0) act[i] = 0;
          all = all && conv[i] >= 0;
        }
        std::swap(tin, tout);
        if (all) break;
      }
      for (int b = 0; b < cb; ++b)
        MSC_HIP_CHECK(hipMemcpyAsync(dA + (size_t)b * stride + (size_t)n9 * n9,
                                     dacc + (size_t)b * n9 * 9, (size_t)n9 * 9 * sizeof(double2),
                                     hipMemcpyDeviceToDevice, st));
      for (int i = 0; i < cb; ++i) {
        if (pivot_ratio) pivot_ratio[b0 + i] = NAN;
        if (neumann_iters) neumann_iters[b0 + i] = conv[i];
      }
    }
    hipLaunchKernelGGL(k_res_tcomp, dim3(cb), dim3(kRT), 0, st, dA, stride, dtp, n9,
                       out + (size_t)b0 * 81, pex ? pex + (size_t)b0 * n9 * 9 : nullptr);
    MSC_LAUNCH_CHECK();
  }
  MSC_HIP_CHECK(hipStreamSynchronize(st));
  return kOk;
}
