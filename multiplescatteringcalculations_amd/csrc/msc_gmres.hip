// Device-resident restarted GMRES / GMRES-IR for the slab operator (gfx950).
//
// Replaces torch_gmres (cubic_scattering/torch_gmres.py:83-205) and the SciPy
// gmres call of compute_slab_scattering (slab_scattering.py:532-552).
//
// * Arnoldi with classical Gram-Schmidt as block GEMVs (one pass of dots
//   h = V^H w, one pass of updates w -= V h), repeated once when the norm
//   drops by more than 1/sqrt(2) ("twice is enough"), all reductions in FP64
//   through per-workgroup partials reduced in a fixed order (deterministic).
// * Hessenberg, Givens rotations (real c, complex s) and back-substitution in
//   FP64 on the device; a device flag `done` turns the remaining launches of
//   a cycle into no-ops, so a restart cycle needs no host synchronisation.
// * The true FP64 residual b - A x judges every cycle (as in
//   torch_gmres.py:202-203); the mixed form runs complex64 cycles inside an
//   FP64 iterative-refinement loop.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "msc_gmres.hpp"

namespace msc {

constexpr int kRT = 256;     // threads per workgroup of the vector kernels
constexpr int kChunk = 4096;  // vector elements per workgroup

__device__ __forceinline__ double2 wave_sum(double2 v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    v.x += __shfl_down(v.x, off, 64);
    v.y += __shfl_down(v.y, off, 64);
  }
  return v;
}
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// Sum over the workgroup; valid in thread 0.
__device__ __forceinline__ double2 block_sum(double2 v, double2* red) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double2 t = make_double2(0.0, 0.0);
  if (threadIdx.x == 0) {
    for (int i = 0; i < kRT / 64; ++i) {
      t.x += red[i].x;
      t.y += red[i].y;
    }
  }
  __syncthreads();
  return t;
}

__device__ __forceinline__ bool skip_step(const KState* st, int pass) {
  return st->done != 0 || (pass == 1 && st->reorth == 0);
}

// partial[blk][v] = sum_{t in chunk} conj(V_v[t]) w[t], v < nv;
// partial[blk][nv] = sum |w|^2 (stored in .x).
template <typename R>
__device__ __forceinline__ void dots_body(unsigned bx_, unsigned gx_, const cx<R>* __restrict__ V, size_t ldv, int nv,
                                              const cx<R>* __restrict__ w, size_t n,
                                              double2* __restrict__ part, const KState* st,
                                              int pass) {
  __shared__ double2 red[kRT / 64];
  if (skip_step(st, pass)) return;
  const size_t beg = (size_t)bx_ * kChunk;
  const size_t end = beg + kChunk < n ? beg + kChunk : n;
  const int ldp = nv + 1;
  double2 wn = make_double2(0.0, 0.0);
  for (size_t t = beg + threadIdx.x; t < end; t += kRT) {
    const double2 a = to_c128(w[t]);
    wn.x += a.x * a.x + a.y * a.y;
  }
  wn = block_sum(wn, red);
  if (threadIdx.x == 0) part[(size_t)bx_ * ldp + nv] = wn;
  constexpr int VB = 8;
  for (int v0 = 0; v0 < nv; v0 += VB) {
    double2 acc[VB];
#pragma unroll
    for (int k = 0; k < VB; ++k) acc[k] = make_double2(0.0, 0.0);
    for (size_t t = beg + threadIdx.x; t < end; t += kRT) {
      const double2 a = to_c128(w[t]);
#pragma unroll
      for (int k = 0; k < VB; ++k) {
        if (v0 + k < nv) {
          const double2 b = to_c128(V[(size_t)(v0 + k) * ldv + t]);
          // conj(b) * a
          acc[k].x += b.x * a.x + b.y * a.y;
          acc[k].y += b.x * a.y - b.y * a.x;
        }
      }
    }
#pragma unroll
    for (int k = 0; k < VB; ++k) {
      if (v0 + k < nv) {
        const double2 s = block_sum(acc[k], red);
        if (threadIdx.x == 0) part[(size_t)bx_ * ldp + v0 + k] = s;
      }
    }
  }
}
template <typename R>
__global__ __launch_bounds__(kRT) void k_dots(const cx<R>* __restrict__ V, size_t ldv, int nv,
                                              const cx<R>* __restrict__ w, size_t n,
                                              double2* __restrict__ part, const KState* st,
                                              int pass) {
  dots_body<R>(blockIdx.x, gridDim.x, V, ldv, nv, w, n, part, st, pass);
}


__device__ __forceinline__ double cabs2(double2 a) { return a.x * a.x + a.y * a.y; }
__device__ __forceinline__ double2 cm(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }

// Close Arnoldi step j with ||w|| = hn: form column j of the Hessenberg matrix
// of the iterated operator, apply the previous Givens rotations, build the new
// one (real c, complex s) and rotate g.  With `shifted`, the basis was built
// from B = G0 T0 and the operator is A = I - B, so H_A = [I; 0] - H_B.
__device__ void finish_step(KState* st, double2* __restrict__ H, int ldh,
                            double2* __restrict__ cs, double2* __restrict__ g,
                            double* __restrict__ hnext, int m, double hn, int shifted) {
  const int j = st->j;
  double2* col = H + (size_t)j * ldh;
  hnext[j] = hn;
  if (shifted) {
    for (int i = 0; i < j; ++i) col[i] = make_double2(-col[i].x, -col[i].y);
    col[j] = make_double2(1.0 - col[j].x, -col[j].y);
    col[j + 1] = make_double2(-hn, 0.0);
  } else {
    col[j + 1] = make_double2(hn, 0.0);
  }
  for (int i = 0; i < j; ++i) {  // [c s; -conj(s) c]
    const double c = cs[2 * i].x;
    const double2 s = cs[2 * i + 1];
    const double2 a = col[i], b = col[i + 1];
    const double2 sb = cm(s, b);
    const double2 csa = cm(cconj(s), a);
    col[i] = make_double2(c * a.x + sb.x, c * a.y + sb.y);
    col[i + 1] = make_double2(c * b.x - csa.x, c * b.y - csa.y);
  }
  const double2 a = col[j], b = col[j + 1];
  const double na = sqrt(cabs2(a)), nb = sqrt(cabs2(b));
  const double d = sqrt(na * na + nb * nb);
  double c;
  double2 s;
  if (d == 0.0) {
    c = 1.0;
    s = make_double2(0.0, 0.0);
  } else if (na == 0.0) {
    c = 0.0;
    s = make_double2(b.x / nb, -b.y / nb);  // conj(b)/|b|
  } else {
    c = na / d;  // s = (a/|a|) conj(b) / d
    const double2 u = make_double2(a.x / na, a.y / na);
    s = cm(u, cconj(b));
    s.x /= d;
    s.y /= d;
  }
  cs[2 * j] = make_double2(c, 0.0);
  cs[2 * j + 1] = s;
  const double2 sb = cm(s, b);
  col[j] = make_double2(c * a.x + sb.x, c * a.y + sb.y);
  col[j + 1] = make_double2(0.0, 0.0);
  const double2 gj = g[j];
  const double2 sg = cm(cconj(s), gj);
  g[j] = make_double2(c * gj.x, c * gj.y);
  g[j + 1] = make_double2(-sg.x, -sg.y);
  const double res = sqrt(cabs2(g[j + 1]));
  st->resid_est = res;
  st->scale = hn < 1e-30 ? 0.0 : 1.0 / hn;
  st->j = j + 1;
  if (res <= st->tol_abs || hn < 1e-30 || j + 1 >= m) {
    st->done = 1;
    if (st->done_host != nullptr) {
      *reinterpret_cast<volatile int*>(st->done_host) = 1;
      __threadfence_system();
    }
  }
}

// Reduce the dot partials (fixed order) into hcur and add them to H[:, j].
// Pass 0 also decides, from ||w||^2 - |h|^2 (exact for an orthonormal basis),
// whether a second Gram-Schmidt pass is needed; if not it closes the step with
// that norm, and k_update writes the normalised vector directly.
__device__ __forceinline__ void reduce_dots_body(unsigned bx_, unsigned gx_, const double2* __restrict__ part, int nblk,
                                                     int nv, KState* st, double2* __restrict__ H,
                                                     int ldh, double2* __restrict__ hcur,
                                                     double2* __restrict__ cs,
                                                     double2* __restrict__ g,
                                                     double* __restrict__ hnext, int m, int pass,
                                                     int orth_mode, int shifted) {
  __shared__ double wnb;
  if (skip_step(st, pass)) return;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int ldp = nv + 1;
  const int j = st->j;
  for (int v = wid; v <= nv; v += kRT / 64) {
    double2 s = make_double2(0.0, 0.0);
    for (int b = lane; b < nblk; b += 64) {
      const double2 p = part[(size_t)b * ldp + v];
      s.x += p.x;
      s.y += p.y;
    }
    s = wave_sum(s);
    if (lane == 0) {
      if (v < nv) {
        hcur[v] = s;
        double2& h = H[(size_t)j * ldh + v];
        if (pass == 0) {
          h = s;
        } else {
          h.x += s.x;
          h.y += s.y;
        }
      } else {
        wnb = s.x;
      }
    }
  }
  __syncthreads();
  if (threadIdx.x != 0 || pass != 0) return;
  double hh = 0.0;
  for (int v = 0; v < nv; ++v) hh += cabs2(hcur[v]);
  const double est = wnb - hh;
  const bool again = orth_mode == kOrthCGS2 ||
                     (orth_mode == kOrthAdaptive && !(est > 0.5 * wnb));
  if (again) {
    st->reorth = 1;
    return;
  }
  st->reorth = 0;
  finish_step(st, H, ldh, cs, g, hnext, m, sqrt(est > 0.0 ? est : 0.0), shifted);
}
__global__ __launch_bounds__(kRT) void k_reduce_dots(const double2* __restrict__ part, int nblk,
                                                     int nv, KState* st, double2* __restrict__ H,
                                                     int ldh, double2* __restrict__ hcur,
                                                     double2* __restrict__ cs,
                                                     double2* __restrict__ g,
                                                     double* __restrict__ hnext, int m, int pass,
                                                     int orth_mode, int shifted) {
  reduce_dots_body(blockIdx.x, gridDim.x, part, nblk, nv, st, H, ldh, hcur, cs, g, hnext, m, pass, orth_mode, shifted);
}


// w -= sum_v hcur[v] V_v.  Without a second pass pending (pass 0, reorth == 0)
// the result is scaled to unit norm in the same sweep; with one pending the
// unscaled vector is kept.  Pass 1 writes per-workgroup ||w||^2 partials.
template <typename R>
__device__ __forceinline__ void update_body(unsigned bx_, unsigned gx_, const cx<R>* __restrict__ V, size_t ldv, int nv,
                                                cx<R>* __restrict__ w, size_t n,
                                                const double2* __restrict__ hcur,
                                                double* __restrict__ part, const KState* st,
                                                int pass) {
  __shared__ double2 red[kRT / 64];
  if (pass == 0 ? st->done != 0 : skip_step(st, pass)) return;
  const bool finalise = pass == 0 && st->reorth == 0;
  const double sc = finalise ? st->scale : 1.0;
  const size_t beg = (size_t)bx_ * kChunk;
  const size_t end = beg + kChunk < n ? beg + kChunk : n;
  double2 nrm = make_double2(0.0, 0.0);
  for (size_t t = beg + threadIdx.x; t < end; t += kRT) {
    double2 a = to_c128(w[t]);
    for (int v = 0; v < nv; ++v) {
      const double2 h = hcur[v];
      const double2 b = to_c128(V[(size_t)v * ldv + t]);
      a.x -= h.x * b.x - h.y * b.y;
      a.y -= h.x * b.y + h.y * b.x;
    }
    const cx<R> o = from_c128<R>(make_double2(a.x * sc, a.y * sc));
    w[t] = o;
    const double2 oo = to_c128(o);
    nrm.x += oo.x * oo.x + oo.y * oo.y;
  }
  if (pass == 1) {
    nrm = block_sum(nrm, red);
    if (threadIdx.x == 0) part[bx_] = nrm.x;
  }
}
template <typename R>
__global__ __launch_bounds__(kRT) void k_update(const cx<R>* __restrict__ V, size_t ldv, int nv,
                                                cx<R>* __restrict__ w, size_t n,
                                                const double2* __restrict__ hcur,
                                                double* __restrict__ part, const KState* st,
                                                int pass) {
  update_body<R>(blockIdx.x, gridDim.x, V, ldv, nv, w, n, hcur, part, st, pass);
}


__device__ double reduce_scalar_partials(const double* part, int nblk, double* red) {
  // one workgroup, fixed order
  double s = 0.0;
  for (int b = threadIdx.x; b < nblk; b += kRT) s += part[b];
  s = wave_sum(s);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) red[wid] = s;
  __syncthreads();
  double t = 0.0;
  for (int i = 0; i < kRT / 64; ++i) t += red[i];
  __syncthreads();
  return t;
}

// After a second Gram-Schmidt pass: exact norm, then close the step.
__device__ __forceinline__ void norm_givens_body(unsigned bx_, unsigned gx_, const double* __restrict__ part, int nblk,
                                                     KState* st, double2* __restrict__ H, int ldh,
                                                     double2* __restrict__ cs,
                                                     double2* __restrict__ g,
                                                     double* __restrict__ hnext, int m,
                                                     int shifted) {
  __shared__ double red[kRT / 64];
  if (skip_step(st, 1)) return;
  const double wn = reduce_scalar_partials(part, nblk, red);
  if (threadIdx.x != 0) return;
  finish_step(st, H, ldh, cs, g, hnext, m, sqrt(wn), shifted);
}
__global__ __launch_bounds__(kRT) void k_norm_givens(const double* __restrict__ part, int nblk,
                                                     KState* st, double2* __restrict__ H, int ldh,
                                                     double2* __restrict__ cs,
                                                     double2* __restrict__ g,
                                                     double* __restrict__ hnext, int m,
                                                     int shifted) {
  norm_givens_body(blockIdx.x, gridDim.x, part, nblk, st, H, ldh, cs, g, hnext, m, shifted);
}


// V[j+1] = w / ||w|| after a second pass (zero on breakdown).
template <typename R>
__device__ __forceinline__ void scale_next_body(unsigned bx_, unsigned gx_, cx<R>* __restrict__ w, size_t n,
                                                    const KState* st) {
  if (skip_step(st, 1)) return;
  const double inv = st->scale;
  for (size_t t = (size_t)bx_ * kRT + threadIdx.x; t < n; t += (size_t)gx_ * kRT) {
    const double2 a = to_c128(w[t]);
    w[t] = from_c128<R>(make_double2(a.x * inv, a.y * inv));
  }
}
template <typename R>
__global__ __launch_bounds__(kRT) void k_scale_next(cx<R>* __restrict__ w, size_t n,
                                                    const KState* st) {
  scale_next_body<R>(blockIdx.x, gridDim.x, w, n, st);
}


// Back-substitution R y = g on the leading j x j block (one workgroup).
__device__ __forceinline__ void backsub_body(unsigned bx_, unsigned gx_, const KState* st, const double2* __restrict__ H,
                                                 int ldh, const double2* __restrict__ g,
                                                 double2* __restrict__ y, double2* __restrict__ tmp) {
  const int j = st->j;
  for (int i = threadIdx.x; i < j; i += kRT) tmp[i] = g[i];
  __syncthreads();
  for (int i = j - 1; i >= 0; --i) {
    if (threadIdx.x == 0) {
      const double2 r = H[(size_t)i * ldh + i];
      const double2 t = tmp[i];
      const double den = cabs2(r);
      // t / r
      y[i] = den == 0.0 ? make_double2(0.0, 0.0)
                        : make_double2((t.x * r.x + t.y * r.y) / den, (t.y * r.x - t.x * r.y) / den);
    }
    __syncthreads();
    const double2 yi = y[i];
    for (int k = threadIdx.x; k < i; k += kRT) {
      const double2 h = H[(size_t)i * ldh + k];
      tmp[k].x -= h.x * yi.x - h.y * yi.y;
      tmp[k].y -= h.x * yi.y + h.y * yi.x;
    }
    __syncthreads();
  }
}
__global__ __launch_bounds__(kRT) void k_backsub(const KState* st, const double2* __restrict__ H,
                                                 int ldh, const double2* __restrict__ g,
                                                 double2* __restrict__ y, double2* __restrict__ tmp) {
  backsub_body(blockIdx.x, gridDim.x, st, H, ldh, g, y, tmp);
}


// x += scale * sum_{v < j} y_v V_v  (x complex128 or the basis precision).
template <typename R, typename RX>
__device__ __forceinline__ void axpy_basis_body(unsigned bx_, unsigned gx_, const cx<R>* __restrict__ V, size_t ldv,
                                                    const double2* __restrict__ y,
                                                    const KState* st, cx<RX>* __restrict__ x,
                                                    size_t n, const double* __restrict__ scale) {
  const int j = st->j;
  const double sc = scale != nullptr ? *scale : 1.0;
  for (size_t t = (size_t)bx_ * kRT + threadIdx.x; t < n; t += (size_t)gx_ * kRT) {
    double2 acc = make_double2(0.0, 0.0);
    for (int v = 0; v < j; ++v) {
      const double2 b = to_c128(V[(size_t)v * ldv + t]);
      const double2 c = y[v];
      acc.x += c.x * b.x - c.y * b.y;
      acc.y += c.x * b.y + c.y * b.x;
    }
    double2 xv = to_c128(x[t]);
    xv.x += sc * acc.x;
    xv.y += sc * acc.y;
    x[t] = from_c128<RX>(xv);
  }
}
template <typename R, typename RX>
__global__ __launch_bounds__(kRT) void k_axpy_basis(const cx<R>* __restrict__ V, size_t ldv,
                                                    const double2* __restrict__ y,
                                                    const KState* st, cx<RX>* __restrict__ x,
                                                    size_t n, const double* __restrict__ scale) {
  axpy_basis_body<R, RX>(blockIdx.x, gridDim.x, V, ldv, y, st, x, n, scale);
}


// partial[blk] = sum |r|^2 over the block's chunk.
template <typename R>
__device__ __forceinline__ void norm2_part_body(unsigned bx_, unsigned gx_, const cx<R>* __restrict__ r, size_t n,
                                                    double* __restrict__ part) {
  __shared__ double2 red[kRT / 64];
  const size_t beg = (size_t)bx_ * kChunk;
  const size_t end = beg + kChunk < n ? beg + kChunk : n;
  double2 s = make_double2(0.0, 0.0);
  for (size_t t = beg + threadIdx.x; t < end; t += kRT) {
    const double2 a = to_c128(r[t]);
    s.x += a.x * a.x + a.y * a.y;
  }
  s = block_sum(s, red);
  if (threadIdx.x == 0) part[bx_] = s.x;
}
template <typename R>
__global__ __launch_bounds__(kRT) void k_norm2_part(const cx<R>* __restrict__ r, size_t n,
                                                    double* __restrict__ part) {
  norm2_part_body<R>(blockIdx.x, gridDim.x, r, n, part);
}


// *out = sqrt(sum partials)
__device__ __forceinline__ void norm_finish_body(unsigned bx_, unsigned gx_, const double* __restrict__ part, int nblk,
                                                     double* __restrict__ out) {
  __shared__ double red[kRT / 64];
  const double s = reduce_scalar_partials(part, nblk, red);
  if (threadIdx.x == 0) *out = sqrt(s);
}
__global__ __launch_bounds__(kRT) void k_norm_finish(const double* __restrict__ part, int nblk,
                                                     double* __restrict__ out) {
  norm_finish_body(blockIdx.x, gridDim.x, part, nblk, out);
}


// Start a cycle: V0 = r / ||r|| (r in precision RS, V0 in precision R),
// g = (||r|| or 1, 0, ...), reset the step state.
template <typename RS, typename R>
__device__ __forceinline__ void cycle_init_body(unsigned bx_, unsigned gx_, const cx<RS>* __restrict__ r, size_t n,
                                                    cx<R>* __restrict__ V0, KState* st,
                                                    double2* __restrict__ g, int m,
                                                    double tol_abs, int unit_rhs) {
  const double rn = st->rnorm;
  const double inv = rn > 0.0 ? 1.0 / rn : 0.0;
  for (size_t t = (size_t)bx_ * kRT + threadIdx.x; t < n; t += (size_t)gx_ * kRT) {
    const double2 a = to_c128(r[t]);
    V0[t] = from_c128<R>(make_double2(a.x * inv, a.y * inv));
  }
  if (bx_ == 0) {
    for (int i = threadIdx.x; i <= m; i += kRT) g[i] = make_double2(0.0, 0.0);
    __syncthreads();
    if (threadIdx.x == 0) {
      g[0] = make_double2(unit_rhs ? 1.0 : rn, 0.0);
      st->j = 0;
      st->reorth = 0;
      st->done = rn > 0.0 ? 0 : 1;
      if (st->done && st->done_host != nullptr) {
        *reinterpret_cast<volatile int*>(st->done_host) = 1;
        __threadfence_system();
      }
      st->tol_abs = tol_abs;
      st->resid_est = unit_rhs ? 1.0 : rn;
    }
  }
}
template <typename RS, typename R>
__global__ __launch_bounds__(kRT) void k_cycle_init(const cx<RS>* __restrict__ r, size_t n,
                                                    cx<R>* __restrict__ V0, KState* st,
                                                    double2* __restrict__ g, int m,
                                                    double tol_abs, int unit_rhs) {
  cycle_init_body<RS, R>(blockIdx.x, gridDim.x, r, n, V0, st, g, m, tol_abs, unit_rhs);
}


__device__ __forceinline__ void narrow_body(unsigned bx_, unsigned gx_, const double2* __restrict__ in,
                                                float2* __restrict__ out, size_t n) {
  for (size_t t = (size_t)bx_ * kRT + threadIdx.x; t < n; t += (size_t)gx_ * kRT)
    out[t] = make_float2((float)in[t].x, (float)in[t].y);
}
__global__ __launch_bounds__(kRT) void k_narrow(const double2* __restrict__ in,
                                                float2* __restrict__ out, size_t n) {
  narrow_body(blockIdx.x, gridDim.x, in, out, n);
}


// ── batched forms: system q = d[blockIdx.y], its own chunks / grid stride ────
template <typename R>
__global__ __launch_bounds__(kRT) void kb_dots(const BKry* __restrict__ d, int nv, int pass) {
  const BKry& q = d[blockIdx.y];
  if ((int)blockIdx.x >= q.nblk) return;
  const cx<R>* V = reinterpret_cast<const cx<R>*>(q.V);
  dots_body<R>(blockIdx.x, q.nblk, V, q.ldv, nv, V + (size_t)nv * q.ldv, q.n, q.part, q.st, pass);
}

__global__ __launch_bounds__(kRT) void kb_reduce_dots(const BKry* __restrict__ d, int nv, int pass,
                                                      int orth_mode) {
  const BKry& q = d[blockIdx.y];
  reduce_dots_body(0, 1, q.part, q.nblk, nv, q.st, q.H, q.ldh, q.hcur, q.cs, q.g, q.hnext, q.m,
                   pass, orth_mode, q.shifted);
}

template <typename R>
__global__ __launch_bounds__(kRT) void kb_update(const BKry* __restrict__ d, int nv, int pass) {
  const BKry& q = d[blockIdx.y];
  if ((int)blockIdx.x >= q.nblk) return;
  cx<R>* V = reinterpret_cast<cx<R>*>(q.V);
  update_body<R>(blockIdx.x, q.nblk, V, q.ldv, nv, V + (size_t)nv * q.ldv, q.n, q.hcur,
                 reinterpret_cast<double*>(q.part), q.st, pass);
}

__global__ __launch_bounds__(kRT) void kb_norm_givens(const BKry* __restrict__ d) {
  const BKry& q = d[blockIdx.y];
  norm_givens_body(0, 1, reinterpret_cast<const double*>(q.part), q.nblk, q.st, q.H, q.ldh, q.cs,
                   q.g, q.hnext, q.m, q.shifted);
}

template <typename R>
__global__ __launch_bounds__(kRT) void kb_scale_next(const BKry* __restrict__ d, int nv) {
  const BKry& q = d[blockIdx.y];
  scale_next_body<R>(blockIdx.x, gridDim.x, reinterpret_cast<cx<R>*>(q.V) + (size_t)nv * q.ldv,
                     q.n, q.st);
}

template <typename R>
__global__ __launch_bounds__(kRT) void kb_norm2_part(const BKry* __restrict__ d) {
  const BKry& q = d[blockIdx.y];
  if ((int)blockIdx.x >= q.nblk) return;
  norm2_part_body<R>(blockIdx.x, q.nblk, reinterpret_cast<const cx<R>*>(q.r), q.n,
                     reinterpret_cast<double*>(q.part));
}

// ||r|| into the system's KState::rnorm
__global__ __launch_bounds__(kRT) void kb_norm_finish(const BKry* __restrict__ d) {
  const BKry& q = d[blockIdx.y];
  norm_finish_body(0, 1, reinterpret_cast<const double*>(q.part), q.nblk, &q.st->rnorm);
}

template <typename RS, typename R>
__global__ __launch_bounds__(kRT) void kb_cycle_init(const BKry* __restrict__ d, int unit_rhs) {
  const BKry& q = d[blockIdx.y];
  cycle_init_body<RS, R>(blockIdx.x, gridDim.x, reinterpret_cast<const cx<RS>*>(q.r), q.n,
                         reinterpret_cast<cx<R>*>(q.V), q.st, q.g, q.m, q.tol_abs, unit_rhs);
}

__global__ __launch_bounds__(kRT) void kb_backsub(const BKry* __restrict__ d) {
  const BKry& q = d[blockIdx.y];
  backsub_body(0, 1, q.st, q.H, q.ldh, q.g, q.y, q.tmp);
}

// x += ||r|| sum_v y_v V_v (the complex64 cycle of iterative refinement, or
// the FP64 cycle with a unit right-hand side)
template <typename R>
__global__ __launch_bounds__(kRT) void kb_axpy_basis(const BKry* __restrict__ d) {
  const BKry& q = d[blockIdx.y];
  axpy_basis_body<R, double>(blockIdx.x, gridDim.x, reinterpret_cast<const cx<R>*>(q.V), q.ldv,
                             q.y, q.st, q.x, q.n, &q.st->rnorm);
}

__global__ __launch_bounds__(kRT) void kb_narrow(const BKry* __restrict__ d) {
  const BKry& q = d[blockIdx.y];
  narrow_body(blockIdx.x, gridDim.x, q.src, q.dst, q.n);
}

__global__ void kb_gather(const BKry* __restrict__ d, int nsys, double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nsys) return;
  out[2 * i] = d[i].st->rnorm;
  out[2 * i + 1] = (double)d[i].st->j;
}

__global__ void kb_deactivate(const BKry* __restrict__ d, int nsys, const int* __restrict__ active) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nsys && active[i] == 0) d[i].st->rnorm = 0.0;
}

// ── host side ───────────────────────────────────────────────────────────────

static int grid_for(size_t n) {
  const size_t b = (n + kRT - 1) / kRT;
  return (int)std::min<size_t>(b, 8192);
}

Krylov::~Krylov() { release(); }

void Krylov::drop_graphs() {
  for (hipGraphExec_t e : graphs)
    if (e) (void)hipGraphExecDestroy(e);
  graphs.clear();
  graph_orth = -1;
  graph_net = 0;
}

void Krylov::release() {
  drop_graphs();
  if (done_flag) (void)hipHostFree(done_flag);
  done_flag = nullptr;
  if (V && owns_V) (void)hipFree(V);
  owns_V = false;
  if (Z) (void)hipFree(Z);
  Z = nullptr;
  if (st) (void)hipFree(st);
  if (H) (void)hipFree(H);
  V = nullptr;
  st = nullptr;
  H = nullptr;
  if (host) (void)hipHostFree(host);
  host = nullptr;
}

int Krylov::ensure_state(size_t n_, int m_, int elem_bytes_) {
  if (st != nullptr && n_ <= cap_n && m_ == m && elem_bytes_ == elem_bytes) {
    if (n_ != n) {  // a smaller problem in the same buffers: re-derive the sizes
      drop_graphs();
      n = n_;
      nblk = (int)((n + kChunk - 1) / kChunk);
    }
    return kOk;
  }
  release();
  n = n_;
  cap_n = n_;
  m = m_;
  elem_bytes = elem_bytes_;
  ldv = (n + 63) & ~size_t(63);  // 64-element aligned rows
  nblk = (int)((n + kChunk - 1) / kChunk);
  ldh = m + 1;
  MSC_HIP_CHECK(hipMalloc(&st, sizeof(KState)));
  MSC_HIP_CHECK(hipMemset(st, 0, sizeof(KState)));
  // H, cs, g, hcur, y, tmp, 4 scalars, per-workgroup partials, hnext
  const size_t nH = (size_t)ldh * m;
  const size_t nsmall = nH + 2 * (size_t)m + (m + 1) + (m + 1) + m + m + 2;
  const size_t npart = (size_t)nblk * (m + 2);
  MSC_HIP_CHECK(
      hipMalloc(&H, (nsmall + npart) * sizeof(double2) + (size_t)(m + 1) * sizeof(double)));
  cs = H + nH;
  g = cs + 2 * m;
  hcur = g + (m + 1);
  y = hcur + (m + 1);
  tmp = y + m;
  scal = reinterpret_cast<double*>(tmp + m);
  part = reinterpret_cast<double2*>(scal) + 2;
  hnext = reinterpret_cast<double*>(part + npart);
  MSC_HIP_CHECK(hipHostMalloc(&host, sizeof(KState), hipHostMallocDefault));
  MSC_HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&done_flag), sizeof(int), hipHostMallocMapped));
  *done_flag = 0;
  int* dview = nullptr;
  MSC_HIP_CHECK(hipHostGetDevicePointer(reinterpret_cast<void**>(&dview), done_flag, 0));
  MSC_HIP_CHECK(hipMemcpy(&st->done_host, &dview, sizeof(int*), hipMemcpyHostToDevice));
  return kOk;
}

int Krylov::ensure(size_t n_, int m_, int elem_bytes_) {
  if (V != nullptr && owns_V && n_ <= cap_n && m_ == m && elem_bytes_ == elem_bytes)
    return ensure_state(n_, m_, elem_bytes_);
  release();
  int rc = ensure_state(n_, m_, elem_bytes_);
  if (rc) return rc;
  drop_graphs();
  if (V != nullptr && owns_V) (void)hipFree(V);
  V = nullptr;
  if (Z != nullptr) (void)hipFree(Z);
  Z = nullptr;
  MSC_HIP_CHECK(hipMalloc(&V, (size_t)(m + 1) * ldv * elem_bytes));
  owns_V = true;
  return kOk;
}

int Krylov::ensure_z() {
  if (Z != nullptr) return kOk;
  MSC_REQUIRE(st != nullptr, "Krylov workspace not allocated");
  MSC_HIP_CHECK(hipMalloc(&Z, (size_t)m * ldv * elem_bytes));
  return kOk;
}

template <typename R>
int krylov_step(Krylov& k, int j, int orth_mode, hipStream_t s) {
  const cx<R>* V = reinterpret_cast<const cx<R>*>(k.V);
  cx<R>* w = reinterpret_cast<cx<R>*>(k.V) + (size_t)(j + 1) * k.ldv;
  const int nv = j + 1;
  const int sh = k.shifted ? 1 : 0;
  double* dpart = reinterpret_cast<double*>(k.part);
  for (int pass = 0; pass < 2; ++pass) {
    hipLaunchKernelGGL((k_dots<R>), dim3(k.nblk), dim3(kRT), 0, s, V, k.ldv, nv, w, k.n, k.part,
                       k.st, pass);
    hipLaunchKernelGGL(k_reduce_dots, dim3(1), dim3(kRT), 0, s, k.part, k.nblk, nv, k.st, k.H,
                       k.ldh, k.hcur, k.cs, k.g, k.hnext, k.m, pass, orth_mode, sh);
    hipLaunchKernelGGL((k_update<R>), dim3(k.nblk), dim3(kRT), 0, s, V, k.ldv, nv, w, k.n,
                       k.hcur, dpart, k.st, pass);
    if (orth_mode == kOrthCGS1) break;
  }
  if (orth_mode != kOrthCGS1) {
    hipLaunchKernelGGL(k_norm_givens, dim3(1), dim3(kRT), 0, s, dpart, k.nblk, k.st, k.H, k.ldh,
                       k.cs, k.g, k.hnext, k.m, sh);
    hipLaunchKernelGGL((k_scale_next<R>), dim3(grid_for(k.n)), dim3(kRT), 0, s, w, k.n, k.st);
  }
  MSC_LAUNCH_CHECK();
  return kOk;
}

template <typename R>
int krylov_norm(Krylov& k, const cx<R>* r, double* out_dev, hipStream_t s) {
  double* part = reinterpret_cast<double*>(k.part);
  hipLaunchKernelGGL((k_norm2_part<R>), dim3(k.nblk), dim3(kRT), 0, s, r, k.n, part);
  hipLaunchKernelGGL(k_norm_finish, dim3(1), dim3(kRT), 0, s, part, k.nblk, out_dev);
  MSC_LAUNCH_CHECK();
  return kOk;
}

template <typename RS, typename R>
int krylov_cycle_init(Krylov& k, const cx<RS>* r, double tol_abs, bool unit_rhs, hipStream_t s) {
  hipLaunchKernelGGL((k_cycle_init<RS, R>), dim3(grid_for(k.n)), dim3(kRT), 0, s, r, k.n,
                     reinterpret_cast<cx<R>*>(k.V), k.st, k.g, k.m, tol_abs, unit_rhs ? 1 : 0);
  MSC_LAUNCH_CHECK();
  return kOk;
}

template <typename R, typename RX>
int krylov_finish(Krylov& k, cx<RX>* x, const double* scale_dev, hipStream_t s, bool flexible) {
  MSC_REQUIRE(!flexible || k.Z != nullptr, "flexible GMRES without a Z basis");
  hipLaunchKernelGGL(k_backsub, dim3(1), dim3(kRT), 0, s, k.st, k.H, k.ldh, k.g, k.y, k.tmp);
  hipLaunchKernelGGL((k_axpy_basis<R, RX>), dim3(grid_for(k.n)), dim3(kRT), 0, s,
                     reinterpret_cast<const cx<R>*>(flexible ? k.Z : k.V), k.ldv, k.y, k.st, x,
                     k.n, scale_dev);
  MSC_LAUNCH_CHECK();
  return kOk;
}

int krylov_narrow(const double2* in, float2* out, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_narrow, dim3(grid_for(n)), dim3(kRT), 0, s, in, out, n);
  MSC_LAUNCH_CHECK();
  return kOk;
}

int krylov_read_state(Krylov& k, KState* out, hipStream_t s) {
  MSC_HIP_CHECK(hipMemcpyAsync(k.host, k.st, sizeof(KState), hipMemcpyDeviceToHost, s));
  MSC_HIP_CHECK(hipStreamSynchronize(s));
  *out = *reinterpret_cast<KState*>(k.host);
  return kOk;
}

BKry bkry_of(const Krylov& k) {
  BKry q{};
  q.V = k.V;
  q.ldv = k.ldv;
  q.n = k.n;
  q.nblk = k.nblk;
  q.m = k.m;
  q.ldh = k.ldh;
  q.shifted = k.shifted ? 1 : 0;
  q.st = k.st;
  q.H = k.H;
  q.cs = k.cs;
  q.g = k.g;
  q.hcur = k.hcur;
  q.y = k.y;
  q.tmp = k.tmp;
  q.part = k.part;
  q.hnext = k.hnext;
  return q;
}

static dim3 bgrid(size_t gx, int nsys) { return dim3((unsigned)std::max<size_t>(gx, 1), nsys); }

template <typename R>
int bkrylov_step(const BKry* d, int nsys, size_t max_n, int max_nblk, int j, int orth_mode,
                 hipStream_t s) {
  const int nv = j + 1;
  for (int pass = 0; pass < 2; ++pass) {
    hipLaunchKernelGGL((kb_dots<R>), bgrid(max_nblk, nsys), dim3(kRT), 0, s, d, nv, pass);
    hipLaunchKernelGGL(kb_reduce_dots, bgrid(1, nsys), dim3(kRT), 0, s, d, nv, pass, orth_mode);
    hipLaunchKernelGGL((kb_update<R>), bgrid(max_nblk, nsys), dim3(kRT), 0, s, d, nv, pass);
    if (orth_mode == kOrthCGS1) break;
  }
  if (orth_mode != kOrthCGS1) {
    hipLaunchKernelGGL(kb_norm_givens, bgrid(1, nsys), dim3(kRT), 0, s, d);
    hipLaunchKernelGGL((kb_scale_next<R>), bgrid(grid_for(max_n), nsys), dim3(kRT), 0, s, d, nv);
  }
  MSC_LAUNCH_CHECK();
  return kOk;
}

template <typename R>
int bkrylov_norm(const BKry* d, int nsys, int max_nblk, hipStream_t s) {
  hipLaunchKernelGGL((kb_norm2_part<R>), bgrid(max_nblk, nsys), dim3(kRT), 0, s, d);
  hipLaunchKernelGGL(kb_norm_finish, bgrid(1, nsys), dim3(kRT), 0, s, d);
  MSC_LAUNCH_CHECK();
  return kOk;
}

template <typename RS, typename R>
int bkrylov_cycle_init(const BKry* d, int nsys, size_t max_n, bool unit_rhs, hipStream_t s) {
  hipLaunchKernelGGL((kb_cycle_init<RS, R>), bgrid(grid_for(max_n), nsys), dim3(kRT), 0, s, d,
                     unit_rhs ? 1 : 0);
  MSC_LAUNCH_CHECK();
  return kOk;
}

template <typename R>
int bkrylov_finish(const BKry* d, int nsys, size_t max_n, hipStream_t s) {
  hipLaunchKernelGGL(kb_backsub, bgrid(1, nsys), dim3(kRT), 0, s, d);
  hipLaunchKernelGGL((kb_axpy_basis<R>), bgrid(grid_for(max_n), nsys), dim3(kRT), 0, s, d);
  MSC_LAUNCH_CHECK();
  return kOk;
}

int bkrylov_narrow(const BKry* d, int nsys, size_t max_n, hipStream_t s) {
  hipLaunchKernelGGL(kb_narrow, bgrid(grid_for(max_n), nsys), dim3(kRT), 0, s, d);
  MSC_LAUNCH_CHECK();
  return kOk;
}

int bkrylov_gather(const BKry* d, int nsys, double* out, hipStream_t s) {
  hipLaunchKernelGGL(kb_gather, dim3((nsys + 63) / 64), dim3(64), 0, s, d, nsys, out);
  MSC_LAUNCH_CHECK();
  return kOk;
}

int bkrylov_deactivate(const BKry* d, int nsys, const int* active, hipStream_t s) {
  hipLaunchKernelGGL(kb_deactivate, dim3((nsys + 63) / 64), dim3(64), 0, s, d, nsys, active);
  MSC_LAUNCH_CHECK();
  return kOk;
}

template int bkrylov_step<double>(const BKry*, int, size_t, int, int, int, hipStream_t);
template int bkrylov_step<float>(const BKry*, int, size_t, int, int, int, hipStream_t);
template int bkrylov_norm<double>(const BKry*, int, int, hipStream_t);
template int bkrylov_norm<float>(const BKry*, int, int, hipStream_t);
template int bkrylov_cycle_init<double, float>(const BKry*, int, size_t, bool, hipStream_t);
template int bkrylov_cycle_init<float, float>(const BKry*, int, size_t, bool, hipStream_t);
template int bkrylov_cycle_init<double, double>(const BKry*, int, size_t, bool, hipStream_t);
template int bkrylov_finish<float>(const BKry*, int, size_t, hipStream_t);
template int bkrylov_finish<double>(const BKry*, int, size_t, hipStream_t);

template int krylov_step<double>(Krylov&, int, int, hipStream_t);
template int krylov_step<float>(Krylov&, int, int, hipStream_t);
template int krylov_norm<double>(Krylov&, const double2*, double*, hipStream_t);
template int krylov_norm<float>(Krylov&, const float2*, double*, hipStream_t);
template int krylov_cycle_init<double, double>(Krylov&, const double2*, double, bool, hipStream_t);
template int krylov_cycle_init<double, float>(Krylov&, const double2*, double, bool, hipStream_t);
template int krylov_cycle_init<float, float>(Krylov&, const float2*, double, bool, hipStream_t);
template int krylov_finish<double, double>(Krylov&, double2*, const double*, hipStream_t, bool);
template int krylov_finish<float, double>(Krylov&, double2*, const double*, hipStream_t, bool);
template int krylov_finish<float, float>(Krylov&, float2*, const double*, hipStream_t, bool);

}  // namespace msc
