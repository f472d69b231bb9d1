// Operator assembly on the device: spectral G0 table, per-voxel T0, incident
// field and the reductions behind the R_PP / far-field outputs (gfx950).
//
// * G0 table: replaces _build_slab_kernels (cubic_scattering/slab_scattering.py:
//   267-351) and _build_slab_kernels_gpu (slab_scattering_gpu.py:34-92).  The
//   9x9 Kupradze propagator [[G, C], [H, S]] (resonance_tmatrix.py:162-428) is
//   evaluated in FP64 at every lattice offset of the padded circular grid,
//   its 45 independent entries are 3-D FFT'd, and the octant that the lattice
//   reflections and the x<->y swap leave independent is kept (scaled by 1/P).
// * T0: the closed-form Rayleigh cube T-matrix reduced to 4 complex numbers
//   per voxel (effective_contrasts.py:360-451, resonance_tmatrix.py:461-487).
// * psi0: plane-wave incident field (slab_scattering.py:426-478).
#include <algorithm>
#include <cmath>

#include "msc_fft.hpp"
#include "msc_plan.hpp"
#include "msc_prop.hpp"
#include "msc_setup.hpp"

namespace msc {

namespace {

// Spatial kernel on the circular grid (Pz, Pxy, Pxy): offset d stored at
// d mod P.  Aperiodic: |dz| < Nz, |dx|,|dy| < M.  Periodic (fold without Bloch
// phase, slab_scattering.py:337-346): the lateral kernel becomes the circular
// one, circ[e] = sum of K[d] over |d| < M, d = e mod M, i.e. the offsets
// {j, j - M} for j >= 1 summed per axis.  With Pxy = M (M a power of two)
// circ is stored on the M-point grid; with Pxy >= 2M - 1 (any M) the linear
// kernel K_lin[d] = circ[d mod M], |d| < M, is stored instead, whose linear
// convolution over the M x M window equals the circular one.  The self term
// is zero.  Each thread
// evaluates one point of the octant 0 <= iz <= Pz/2, 0 <= ix, iy <= Pxy/2 into
// work[iz][ix][iy][45]: the reflections are not stored (the FFT passes read
// them from the stored half, k_fft_mirror).  Entry (i, j) of the block changes
// sign under a reflection when exactly one of components i, j flips.
__device__ __forceinline__ unsigned entry_pair(int e) {
  // component pair (i, j) of entry e in add_prop45's layout, as a bit mask
  if (e < 6) {
    const int i = e < 3 ? 0 : (e < 5 ? 1 : 2);
    const int j = e < 3 ? e : (e < 5 ? e - 2 : 2);
    return (1u << i) ^ (1u << j);
  }
  if (e < 24) return (1u << ((e - 6) / 6)) ^ (1u << (3 + (e - 6) % 6));
  int rem = e - 24, al = 0;
  while (rem >= 6 - al) {
    rem -= 6 - al;
    ++al;
  }
  return (1u << (3 + al)) ^ (1u << (3 + al + rem));
}

__global__ __launch_bounds__(128) void k_g0_spatial(double2* __restrict__ work, int Pz, int Pxy,
                                                    int Nz, int M, int periodic, double d,
                                                    double kP, double kS, double pre) {
  const int hz = Pz / 2, hxy = Pxy / 2;
  const size_t total = (size_t)(hz + 1) * (hxy + 1) * (hxy + 1);
  for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (size_t)gridDim.x * blockDim.x) {
    const int iy = (int)(idx % (hxy + 1));
    const int ix = (int)((idx / (hxy + 1)) % (hxy + 1));
    const int iz = (int)(idx / ((size_t)(hxy + 1) * (hxy + 1)));
    double2 acc[kNE];
#pragma unroll
    for (int e = 0; e < kNE; ++e) acc[e] = zc();
    const int dz = iz;
    if (dz < Nz) {
      int xs[2], ys[2], nx = 0, ny = 0;
      if (periodic) {
        if (ix < M) {
          xs[nx++] = ix;
          if (ix >= 1) xs[nx++] = ix - M;
        }
        if (iy < M) {
          ys[ny++] = iy;
          if (iy >= 1) ys[ny++] = iy - M;
        }
      } else {
        if (ix < M) xs[nx++] = ix;
        if (iy < M) ys[ny++] = iy;
      }
      for (int a = 0; a < nx; ++a) {
        for (int b = 0; b < ny; ++b) {
          if (dz == 0 && xs[a] == 0 && ys[b] == 0) continue;
          add_prop45(dz * d, xs[a] * d, ys[b] * d, kP, kS, pre, acc);
        }
      }
    }
    double2* o = work + idx * kNE;
#pragma unroll
    for (int e = 0; e < kNE; ++e) o[e] = acc[e];
  }
}

// Transform along N (forward) of data [outer][N][inner] of which only rows
// 0..N/2 are stored (in: [outer][N/2 + 1][inner]): row N - n is row n with
// entry e = (inner index mod 45) negated when the reflection FLIP changes its
// sign.  The first nout frequencies are kept: out [outer][nout][inner].
constexpr int kFT = 256;  // threads per workgroup

template <int N, int NCOL, unsigned FLIP>
__global__ __launch_bounds__(kFT) void k_fft_mirror(const double2* __restrict__ in,
                                                    double2* __restrict__ out, size_t inner,
                                                    int nout, const double2* __restrict__ tw) {
  constexpr int LD = lds_ld<NCOL>::value;
  constexpr int NH = N / 2 + 1;
  __shared__ double2 s[N * LD];
  const size_t nchunk = (inner + NCOL - 1) / NCOL;
  const size_t o = blockIdx.x / nchunk;
  const size_t i0 = (blockIdx.x - o * nchunk) * NCOL;
  const size_t base = o * (size_t)NH * inner + i0;
  const size_t obase = o * (size_t)nout * inner + i0;
  for (int w = threadIdx.x; w < N * NCOL; w += kFT) {
    const int r = w / NCOL;
    const int e = w - r * NCOL;
    double2 v = make_double2(0.0, 0.0);
    if (i0 + e < inner) {
      const int src = r < NH ? r : N - r;
      v = in[base + (size_t)src * inner + e];
      if (r >= NH && (__popc(FLIP & entry_pair((int)((i0 + e) % kNE))) & 1))
        v = make_double2(-v.x, -v.y);
    }
    s[r * LD + e] = v;
  }
  __syncthreads();
  lds_fft<double, N, false, NCOL, LD, kFT>(s, tw);
  for (int w = threadIdx.x; w < nout * NCOL; w += kFT) {
    const int r = w / NCOL;
    const int e = w - r * NCOL;
    if (i0 + e < inner) out[obase + (size_t)r * inner + e] = s[r * LD + e];
  }
}

template <int N, unsigned FLIP>
int launch_fft_mirror(const double2* in, double2* out, int outer, size_t inner, int nout,
                      const double2* tw, hipStream_t st) {
  constexpr int NCOL = (N <= 64) ? 16 : (N * 8 * sizeof(double2) <= 80 * 1024 ? 8 : 4);
  const size_t nblk = (size_t)outer * ((inner + NCOL - 1) / NCOL);
  MSC_REQUIRE(nblk < (1ull << 31), "FFT batch too large");
  hipLaunchKernelGGL((k_fft_mirror<N, NCOL, FLIP>), dim3((unsigned)nblk), dim3(kFT), 0, st, in,
                     out, inner, nout, tw);
  MSC_LAUNCH_CHECK();
  return kOk;
}

template <unsigned FLIP>
int fft_mirror(const double2* in, double2* out, int outer, int N, size_t inner, int nout,
               const double2* tw, hipStream_t st) {
  MSC_REQUIRE(nout >= 1 && nout <= N && in != out, "fft_mirror: bad arguments");
  switch (N) {
    case 2: return launch_fft_mirror<2, FLIP>(in, out, outer, inner, nout, tw, st);
    case 4: return launch_fft_mirror<4, FLIP>(in, out, outer, inner, nout, tw, st);
    case 8: return launch_fft_mirror<8, FLIP>(in, out, outer, inner, nout, tw, st);
    case 16: return launch_fft_mirror<16, FLIP>(in, out, outer, inner, nout, tw, st);
    case 32: return launch_fft_mirror<32, FLIP>(in, out, outer, inner, nout, tw, st);
    case 64: return launch_fft_mirror<64, FLIP>(in, out, outer, inner, nout, tw, st);
    case 128: return launch_fft_mirror<128, FLIP>(in, out, outer, inner, nout, tw, st);
    case 256: return launch_fft_mirror<256, FLIP>(in, out, outer, inner, nout, tw, st);
    case 512: return launch_fft_mirror<512, FLIP>(in, out, outer, inner, nout, tw, st);
    case 1024: return launch_fft_mirror<1024, FLIP>(in, out, outer, inner, nout, tw, st);
    default: break;
  }
  set_error("fft_mirror: length must be a power of two in [2, 1024]");
  return kUnsupported;
}

__device__ __forceinline__ void uv_split(int uv, int& u, int& v) {
  int t = (int)((sqrt(8.0 * uv + 1.0) - 1.0) * 0.5);
  while ((t + 1) * (t + 2) / 2 <= uv) ++t;
  while (t * (t + 1) / 2 > uv) --t;
  u = t;
  v = uv - t * (t + 1) / 2;
}

// table[(uv*45 + e)*NC + c] = K^(kz=c, kx=u, ky=v)[e] / P from the pruned
// spectrum spec[c][u][v][45] (c < NC, u, v < H = Pxy/2 + 1)
__global__ __launch_bounds__(256) void k_table_compress(const double2* __restrict__ spec, int H,
                                                        int NC, int NUV, double invP,
                                                        double2* __restrict__ t128,
                                                        float2* __restrict__ t64) {
  const size_t total = (size_t)NUV * kNE * NC;
  for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (size_t)gridDim.x * blockDim.x) {
    const int c = (int)(idx % NC);
    const int e = (int)((idx / NC) % kNE);
    const int uv = (int)(idx / ((size_t)NC * kNE));
    int u, v;
    uv_split(uv, u, v);
    const double2 s = spec[(((size_t)c * H + u) * H + v) * kNE + e];
    const double2 o = make_double2(s.x * invP, s.y * invP);
    t128[idx] = o;
    t64[idx] = make_float2((float)o.x, (float)o.y);
  }
}

// Per-voxel T0 parameters from (Dlambda, Dmu, Drho) and the frequency
// constants (Gamma0, A, B, C):
//   T1 = Dl (A + 4B + C) + 2 B Dm,  T2 = (A + B) Dm,  T3 = 2 C Dm
//   (the contraction of effective_contrasts.py:365-402 written out),
//   amplifications (405-420), effective contrasts (423-451), V = (2a)^3:
//   t_u = w^2 V Drho*, t_d = V(Dl* + 2 Dmu*_diag), t_o = V Dl*, t_s = 2 V Dmu*_off.
__global__ __launch_bounds__(256) void k_tparams(const double* __restrict__ Dl,
                                                 const double* __restrict__ Dm,
                                                 const double* __restrict__ Dr, size_t nvox,
                                                 double2 G0, double2 A, double2 B, double2 C,
                                                 double omega, double V,
                                                 double2* __restrict__ t128,
                                                 float2* __restrict__ t64) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nvox;
       i += (size_t)gridDim.x * blockDim.x) {
    const double dl = Dl[i], dm = Dm[i], dr = Dr[i];
    const double2 AB = cad(A, B);
    const double2 sum1 = cad(cad(A, csc(B, 4.0)), C);
    const double2 T1 = cad(csc(sum1, dl), csc(B, 2.0 * dm));
    const double2 T2 = csc(AB, dm);
    const double2 T3 = csc(C, 2.0 * dm);
    auto inv1m = [](double2 z) {  // 1 / (1 - z)
      const double2 q = make_double2(1.0 - z.x, -z.y);
      const double den = q.x * q.x + q.y * q.y;
      return make_double2(q.x / den, -q.y / den);
    };
    const double2 amp_u = inv1m(csc(G0, omega * omega * dr));
    const double2 amp_t = inv1m(cad(cad(csc(T1, 3.0), csc(T2, 2.0)), T3));
    const double2 amp_o = inv1m(csc(T2, 2.0));
    const double2 amp_d = inv1m(cad(csc(T2, 2.0), T3));
    const double2 drs = csc(amp_u, dr);
    const double2 dmo = csc(amp_o, dm);
    const double2 dmd = csc(amp_d, dm);
    const double2 dls = cad(csc(amp_t, dl + 2.0 / 3.0 * dm), csc(amp_d, -2.0 / 3.0 * dm));
    const double2 tu = csc(drs, omega * omega * V);
    const double2 td = csc(cad(dls, csc(dmd, 2.0)), V);
    const double2 to = csc(dls, V);
    const double2 ts = csc(dmo, 2.0 * V);
    double2* o = t128 + i * 4;
    o[0] = tu;
    o[1] = td;
    o[2] = to;
    o[3] = ts;
    if (t64 == nullptr) continue;
    float2* f = t64 + i * 4;
    f[0] = make_float2((float)tu.x, (float)tu.y);
    f[1] = make_float2((float)td.x, (float)td.y);
    f[2] = make_float2((float)to.x, (float)to.y);
    f[3] = make_float2((float)ts.x, (float)ts.y);
  }
}

__global__ __launch_bounds__(256) void k_narrow4(const double2* __restrict__ t128, size_t n,
                                                 float2* __restrict__ t64) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    t64[i] = make_float2((float)t128[i].x, (float)t128[i].y);
  }
}

// psi0[l,i,j,:] = inc[:] exp(i k khat . x_c), x_c = ((l+1/2)d, (i-(M-1)/2)d, (j-(M-1)/2)d)
struct Inc9 {
  double2 v[9];
};

__global__ __launch_bounds__(256) void k_incident(double2* __restrict__ psi0, int M, int Nz,
                                                  double d, double kz, double kx, double ky,
                                                  Inc9 inc) {
  const size_t nvox = (size_t)Nz * M * M;
  for (size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x; v < nvox;
       v += (size_t)gridDim.x * blockDim.x) {
    const int j = (int)(v % M);
    const int i = (int)((v / M) % M);
    const int l = (int)(v / ((size_t)M * M));
    const double z = (l + 0.5) * d;
    const double x = (i - (M - 1) / 2.0) * d;
    const double y = (j - (M - 1) / 2.0) * d;
    double s, c;
    sincos(kz * z + kx * x + ky * y, &s, &c);
    const double2 ph = make_double2(c, s);
#pragma unroll
    for (int k = 0; k < 9; ++k) psi0[v * 9 + k] = cmd(inc.v[k], ph);
  }
}

__device__ __forceinline__ void tapply9(const double2* t, const double2* p, double2* tau) {
  const double2 sum = cad(cad(p[3], p[4]), p[5]);
  const double2 dmo = make_double2(t[1].x - t[2].x, t[1].y - t[2].y);
  const double2 tos = cmd(t[2], sum);
  for (int c = 0; c < 3; ++c) tau[c] = cmd(t[0], p[c]);
  for (int c = 3; c < 6; ++c) tau[c] = cad(cmd(dmo, p[c]), tos);
  for (int c = 6; c < 9; ++c) tau[c] = cmd(t[3], p[c]);
}

// Sums of tau = T0 psi: per layer (Nz x 9) and the phased sums
// sum_v exp(-i k r_hat . x_v) tau_v for k in {kP, kS} (2 x 9).  Workgroup
// (c, l) sums chunk c of layer l's voxels into part[(l * C + c) * 27 + k];
// k_source_reduce adds the C chunks of each layer in a fixed order, so the
// sums are reproducible bit for bit.
__global__ __launch_bounds__(256) void k_source_sums(const double2* __restrict__ psi,
                                                     const double2* __restrict__ tp,
                                                     const int* __restrict__ tidx,
                                                     const double2* __restrict__ tgen, int M,
                                                     double d, double kP, double kS, double rz,
                                                     double rx, double ry,
                                                     double2* __restrict__ part) {
  __shared__ double2 red[27][4];
  const int c = blockIdx.x, C = gridDim.x, l = blockIdx.y;
  const int per = (M * M + C - 1) / C;
  const int v0 = c * per, v1 = min(M * M, v0 + per);
  double2 s[27];
  for (int k = 0; k < 27; ++k) s[k] = zc();
  const size_t base = (size_t)l * M * M;
  for (int v = v0 + (int)threadIdx.x; v < v1; v += blockDim.x) {
    const size_t g = base + v;
    double2 tau[9];
    if (tidx != nullptr) {
      const double2* T = tgen + (size_t)tidx[g] * 81;
      for (int i = 0; i < 9; ++i) {
        double2 acc = zc();
        for (int j = 0; j < 9; ++j) acc = cad(acc, cmd(T[i * 9 + j], psi[g * 9 + j]));
        tau[i] = acc;
      }
    } else {
      tapply9(tp + g * 4, psi + g * 9, tau);
    }
    const int i = v / M, j = v - (v / M) * M;
    const double z = (l + 0.5) * d;
    const double x = (i - (M - 1) / 2.0) * d;
    const double y = (j - (M - 1) / 2.0) * d;
    const double rdot = rz * z + rx * x + ry * y;
    double sp, cp, ss, cs;
    sincos(-kP * rdot, &sp, &cp);
    sincos(-kS * rdot, &ss, &cs);
    const double2 eP = make_double2(cp, sp), eS = make_double2(cs, ss);
    for (int k = 0; k < 9; ++k) {
      s[k] = cad(s[k], tau[k]);
      s[9 + k] = cad(s[9 + k], cmd(eP, tau[k]));
      s[18 + k] = cad(s[18 + k], cmd(eS, tau[k]));
    }
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int k = 0; k < 27; ++k) {
    double2 v = s[k];
    for (int off = 32; off > 0; off >>= 1) {
      v.x += __shfl_down(v.x, off, 64);
      v.y += __shfl_down(v.y, off, 64);
    }
    if (lane == 0) red[k][wid] = v;
  }
  __syncthreads();
  if (threadIdx.x < 27) {
    double2 t = zc();
    for (int w = 0; w < (int)(blockDim.x / 64); ++w) t = cad(t, red[threadIdx.x][w]);
    part[((size_t)l * C + c) * 27 + threadIdx.x] = t;
  }
}

__global__ __launch_bounds__(64) void k_source_reduce(const double2* __restrict__ part, int C,
                                                      int Nz, double2* __restrict__ layer_sum,
                                                      double2* __restrict__ phased) {
  const int l = blockIdx.x, k = threadIdx.x;
  if (l >= Nz || k >= 27) return;
  double2 t = zc();
  for (int c = 0; c < C; ++c) t = cad(t, part[((size_t)l * C + c) * 27 + k]);
  if (k < 9) layer_sum[(size_t)l * 9 + k] = t;
  else phased[(size_t)l * 18 + k - 9] = t;
}

int grid1d(size_t n, int bs) { return (int)std::min<size_t>((n + bs - 1) / bs, 65536); }

}  // namespace

// nn[slot] = VA[slot] - (point propagator at R d), both as full 9x9 blocks; the
// point block is expanded from the table's 45 independent entries
// (G symmetric, H = W^-1 C^T, S = S' W with S' symmetric, W = diag(1,1,1,1/2,1/2,1/2)).
__global__ __launch_bounds__(32) void k_nn_delta(const double* __restrict__ va, double d,
                                                 double kP, double kS, double pre,
                                                 double2* __restrict__ nn128,
                                                 float2* __restrict__ nn64) {
  const int slot = threadIdx.x;
  if (slot >= 26) return;
  const int r = slot < 13 ? slot : slot + 1;  // skip the centre of the 3x3x3 cube
  const int dz = r / 9 - 1, dx = (r / 3) % 3 - 1, dy = r % 3 - 1;
  double2 e45[45];
  for (int e = 0; e < 45; ++e) e45[e] = make_double2(0.0, 0.0);
  add_prop45(dz * d, dx * d, dy * d, kP, kS, pre, e45);
  double2 K[81];
  expand45(e45, K);
  for (int q = 0; q < 81; ++q) {
    const double2 v = make_double2(va[slot * 81 + q] - K[q].x, -K[q].y);
    nn128[slot * 81 + q] = v;
    nn64[slot * 81 + q] = make_float2((float)v.x, (float)v.y);
  }
}

// ── host entry points ────────────────────────────────────────────────────────

int ensure_build_work(Plan& p) {
  const int H = p.Pxy / 2 + 1;
  const size_t bytes = (size_t)p.NC * H * H * kNE * sizeof(double2);
  if (p.work == nullptr) MSC_HIP_CHECK(hipMalloc(&p.work, bytes));
  if (p.work2 == nullptr) MSC_HIP_CHECK(hipMalloc(&p.work2, bytes));
  return kOk;
}

int build_table(Plan& p, double omega, double alpha, double beta, double rho, hipStream_t st) {
  const size_t P = (size_t)p.Pz * p.Pxy * p.Pxy;
  const int P2 = p.Pxy, H = p.Pxy / 2 + 1, NC = p.NC;  // NC = Pz/2 + 1
  const int rc0 = ensure_build_work(p);
  if (rc0) return rc0;
  double2* work = reinterpret_cast<double2*>(p.work);
  double2* work2 = reinterpret_cast<double2*>(p.work2);
  const double kP = omega / alpha, kS = omega / beta;
  const double pre = 1.0 / (4.0 * M_PI * rho * omega * omega);
  const size_t octant = (size_t)NC * H * H;
  hipLaunchKernelGGL(k_g0_spatial, dim3(grid1d(octant, 128)), dim3(128), 0, st, work, p.Pz,
                     p.Pxy, p.Nz, p.M, p.periodic ? 1 : 0, 2.0 * p.a, kP, kS, pre);
  int rc = kOk;
  // Only the stored half of each axis is transformed or kept (the table holds
  // k_y, k_x <= Pxy/2 and k_z <= Pz/2):
  //   y: work [NC*H][H of Pxy rows][45] -> work2 [NC*H][H][45]   (y mirror)
  //   x: work2 [NC][H of Pxy rows][H*45] -> work [NC][H][H*45]   (x mirror)
  //   z: work [1][NC of Pz rows][H*H*45] -> work2 [NC][H*H*45]   (z mirror)
  rc = fft_mirror<kFlipY>(work, work2, NC * H, P2, kNE, H, p.tw128, st);
  if (!rc) rc = fft_mirror<kFlipX>(work2, work, NC, P2, (size_t)H * kNE, H, p.tw128, st);
  if (!rc) rc = fft_mirror<kFlipZ>(work, work2, 1, p.Pz, (size_t)H * H * kNE, NC, p.tw128, st);
  if (!rc) {
    const size_t tot = (size_t)p.NUV * kNE * p.NC;
    hipLaunchKernelGGL(k_table_compress, dim3(grid1d(tot, 256)), dim3(256), 0, st, work2, H,
                       p.NC, p.NUV, 1.0 / (double)P, p.table128, p.table64);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
      set_error(std::string("table compress launch: ") + hipGetErrorString(e));
      rc = kHip;
    }
  }
  hipError_t e1 = hipStreamSynchronize(st);
  if (rc) return rc;
  if (e1 != hipSuccess) {
    set_error(std::string("G0 table build: ") + hipGetErrorString(e1));
    return kHip;
  }
  p.omega = omega;
  p.alpha = alpha;
  p.beta = beta;
  p.rho = rho;
  p.have_table = true;
  return kOk;
}

// *tmax = max |t| over the 4 parameters of every voxel, as float bits
// (non-negative floats order like their bit patterns, so atomicMax works).
__global__ __launch_bounds__(256) void k_tmax_clear(unsigned int* tmax) {
  if (threadIdx.x == 0) *tmax = 0u;
}

// One atomic per workgroup (the waves' maxima meet in LDS first): a few
// hundred atomics on one address instead of one per wave.
__global__ __launch_bounds__(256) void k_tmax(const double2* __restrict__ tp, size_t n,
                                              unsigned int* __restrict__ tmax) {
  __shared__ float wm[4];
  float m = 0.f;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
    m = fmaxf(m, (float)sqrt(tp[i].x * tp[i].x + tp[i].y * tp[i].y));
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_down(m, off, 64));
  if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0)
    atomicMax(tmax, __float_as_uint(fmaxf(fmaxf(wm[0], wm[1]), fmaxf(wm[2], wm[3]))));
}

static int update_tmax(Plan& p, hipStream_t st) {
  hipLaunchKernelGGL(k_tmax_clear, dim3(1), dim3(64), 0, st, p.tmax);
  hipLaunchKernelGGL(k_tmax, dim3(std::min(grid1d(p.nvox * 4, 256), 512)), dim3(256), 0, st,
                     p.tp128, p.nvox * 4, p.tmax);
  MSC_LAUNCH_CHECK();
  return kOk;
}

int set_nn_correction(Plan& p, bool volume_averaged, int n_orders, hipStream_t st) {
  p.have_nn = false;
  if (!volume_averaged) return kOk;
  MSC_REQUIRE(p.have_table, "the G0 table must be built first");
  MSC_REQUIRE(n_orders >= 0 && n_orders <= 3, "n_orders must be in 0..3");
  double host[26 * 81];
  va_blocks(p.alpha, p.beta, p.rho, p.omega, n_orders, host);
  if (p.nn128 == nullptr) MSC_HIP_CHECK(hipMalloc(&p.nn128, 26 * 81 * sizeof(double2)));
  if (p.nn64 == nullptr) MSC_HIP_CHECK(hipMalloc(&p.nn64, 26 * 81 * sizeof(float2)));
  if (p.nn_stage == nullptr) MSC_HIP_CHECK(hipMalloc(&p.nn_stage, 26 * 81 * sizeof(double)));
  MSC_HIP_CHECK(hipMemcpyAsync(p.nn_stage, host, sizeof(host), hipMemcpyHostToDevice, st));
  const double kP = p.omega / p.alpha, kS = p.omega / p.beta;
  const double pre = 1.0 / (4.0 * M_PI * p.rho * p.omega * p.omega);
  hipLaunchKernelGGL(k_nn_delta, dim3(1), dim3(32), 0, st, p.nn_stage, 2.0 * p.a, kP, kS, pre,
                     p.nn128, p.nn64);
  MSC_LAUNCH_CHECK();
  MSC_HIP_CHECK(hipStreamSynchronize(st));  // `host` is a stack buffer
  p.have_nn = true;
  return kOk;
}

int set_tparams_from_contrasts(Plan& p, const double* Dl, const double* Dm, const double* Dr,
                               double omega, const double* coef8, hipStream_t st) {
  const double2 G0 = make_double2(coef8[0], coef8[1]);
  const double2 A = make_double2(coef8[2], coef8[3]);
  const double2 B = make_double2(coef8[4], coef8[5]);
  const double2 C = make_double2(coef8[6], coef8[7]);
  const double V = 8.0 * p.a * p.a * p.a;
  hipLaunchKernelGGL(k_tparams, dim3(grid1d(p.nvox, 256)), dim3(256), 0, st, Dl, Dm, Dr, p.nvox,
                     G0, A, B, C, omega, V, p.tp128, p.tp64);
  MSC_LAUNCH_CHECK();
  p.have_tparams = true;
  p.general = false;
  return update_tmax(p, st);
}

__global__ __launch_bounds__(256) void k_check_index(const int* __restrict__ idx, size_t n,
                                                     int count, int* __restrict__ bad) {
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
    const int v = idx[i];
    if (v < 0 || v >= count) atomicOr(bad, 1);
  }
}

__global__ __launch_bounds__(256) void k_narrow_n(const double2* __restrict__ in, size_t n,
                                                  float2* __restrict__ out) {
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
    out[i] = make_float2((float)in[i].x, (float)in[i].y);
}

int set_tgeneral(Plan& p, const int* tidx, const double2* table, int count, hipStream_t st) {
  if (p.flag == nullptr) MSC_HIP_CHECK(hipMalloc(&p.flag, sizeof(int)));
  if (p.tidx == nullptr) {
    const size_t capvox = (size_t)p.capNz * p.capM * p.capM;
    MSC_HIP_CHECK(hipMalloc(&p.tidx, capvox * sizeof(int)));
  }
  // indices outside the table would read past it: check them first
  MSC_HIP_CHECK(hipMemsetAsync(p.flag, 0, sizeof(int), st));
  hipLaunchKernelGGL(k_check_index, dim3(std::min(grid1d(p.nvox, 256), 1024)), dim3(256), 0, st,
                     tidx, p.nvox, count, p.flag);
  MSC_LAUNCH_CHECK();
  int bad = 0;
  MSC_HIP_CHECK(hipMemcpyAsync(&bad, p.flag, sizeof(int), hipMemcpyDeviceToHost, st));
  MSC_HIP_CHECK(hipStreamSynchronize(st));
  MSC_REQUIRE(bad == 0, "T index out of range [0, count)");
  if (count > p.gen_cap) {
    if (p.tgen128) (void)hipFree(p.tgen128);
    if (p.tgen64) (void)hipFree(p.tgen64);
    p.tgen128 = nullptr;
    p.tgen64 = nullptr;
    p.gen_cap = 0;
    MSC_HIP_CHECK(hipMalloc(&p.tgen128, (size_t)count * 81 * sizeof(double2)));
    MSC_HIP_CHECK(hipMalloc(&p.tgen64, (size_t)count * 81 * sizeof(float2)));
    p.gen_cap = count;
  }
  MSC_HIP_CHECK(hipMemcpyAsync(p.tidx, tidx, p.nvox * sizeof(int), hipMemcpyDeviceToDevice, st));
  MSC_HIP_CHECK(hipMemcpyAsync(p.tgen128, table, (size_t)count * 81 * sizeof(double2),
                               hipMemcpyDeviceToDevice, st));
  hipLaunchKernelGGL(k_narrow_n, dim3(std::min(grid1d((size_t)count * 81, 256), 1024)), dim3(256),
                     0, st, p.tgen128, (size_t)count * 81, p.tgen64);
  MSC_LAUNCH_CHECK();
  p.ngen = count;
  p.general = true;
  p.have_tparams = true;
  hipLaunchKernelGGL(k_tmax_clear, dim3(1), dim3(64), 0, st, p.tmax);
  hipLaunchKernelGGL(k_tmax, dim3(std::min(grid1d((size_t)count * 81, 256), 512)), dim3(256), 0,
                     st, p.tgen128, (size_t)count * 81, p.tmax);
  MSC_LAUNCH_CHECK();
  return kOk;
}

int tparams_batch(const double* Dl, const double* Dm, const double* Dr, size_t count,
                  double omega, const double* coef8, double V, double2* t128, hipStream_t st) {
  const double2 G0 = make_double2(coef8[0], coef8[1]);
  const double2 A = make_double2(coef8[2], coef8[3]);
  const double2 B = make_double2(coef8[4], coef8[5]);
  const double2 C = make_double2(coef8[6], coef8[7]);
  hipLaunchKernelGGL(k_tparams, dim3(grid1d(count, 256)), dim3(256), 0, st, Dl, Dm, Dr, count, G0,
                     A, B, C, omega, V, t128, static_cast<float2*>(nullptr));
  MSC_LAUNCH_CHECK();
  return kOk;
}

int set_tparams_direct(Plan& p, const double2* tp, hipStream_t st) {
  MSC_HIP_CHECK(hipMemcpyAsync(p.tp128, tp, p.nvox * 4 * sizeof(double2), hipMemcpyDeviceToDevice,
                               st));
  hipLaunchKernelGGL(k_narrow4, dim3(grid1d(p.nvox * 4, 256)), dim3(256), 0, st, p.tp128,
                     p.nvox * 4, p.tp64);
  MSC_LAUNCH_CHECK();
  p.have_tparams = true;
  p.general = false;
  return update_tmax(p, st);
}

int incident_field(Plan& p, double k_mag, const double* khat3, const double* inc18, double2* psi0,
                   hipStream_t st) {
  Inc9 inc;
  for (int k = 0; k < 9; ++k) inc.v[k] = make_double2(inc18[2 * k], inc18[2 * k + 1]);
  hipLaunchKernelGGL(k_incident, dim3(grid1d(p.nvox, 256)), dim3(256), 0, st, psi0, p.M, p.Nz,
                     2.0 * p.a, k_mag * khat3[0], k_mag * khat3[1], k_mag * khat3[2], inc);
  MSC_LAUNCH_CHECK();
  return kOk;
}

int source_sums(Plan& p, const double2* psi, double kP, double kS, const double* rhat3,
                double2* layer_sum, double2* phased, hipStream_t st) {
  // about 1024 workgroups in all, at least 256 voxels each
  const int vox = p.M * p.M;
  const int C = std::max(1, std::min((1024 + p.Nz - 1) / p.Nz, (vox + 255) / 256));
  const size_t need = (size_t)p.Nz * C * 27;
  if (need > p.sums_cap) {
    if (p.sums_part) (void)hipFree(p.sums_part);
    p.sums_part = nullptr;
    p.sums_cap = 0;
    MSC_HIP_CHECK(hipMalloc(&p.sums_part, need * sizeof(double2)));
    p.sums_cap = need;
  }
  hipLaunchKernelGGL(k_source_sums, dim3(C, p.Nz), dim3(256), 0, st, psi, p.tp128,
                     p.general ? p.tidx : nullptr, p.tgen128, p.M, 2.0 * p.a, kP, kS, rhat3[0],
                     rhat3[1], rhat3[2], p.sums_part);
  hipLaunchKernelGGL(k_source_reduce, dim3(p.Nz), dim3(64), 0, st, p.sums_part, C, p.Nz,
                     layer_sum, phased);
  MSC_LAUNCH_CHECK();
  return kOk;
}

}  // namespace msc
