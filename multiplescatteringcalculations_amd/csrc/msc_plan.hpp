// Plan object of the slab Foldy-Lax library: geometry, padded FFT grid, the
// compressed spectral G0 table, per-voxel T0 parameters and all workspaces.
//
// HBM layout (n_vox = N_z*M*M voxels, component index innermost):
//   field vectors      (N_z, M, M, 9)            complex, C-contiguous
//   T0 parameters      (N_z, M, M, 4)            complex: t_u, t_d, t_o, t_s
//   buf1 (y-spectra)   (N_z, M, Pxy, 9)          complex
//   buf2 (xy-spectra)  (N_z, Pxy, Pxy, 9)        complex
//   G0 table           (NUV, 45, NC)             complex, NC = Pz/2+1,
//                      NUV = (Pxy/2+1)(Pxy/2+2)/2 canonical (u >= v) columns
// The table holds the 45 independent entries of the 9x9 spectral propagator
// on the octant 0 <= kz <= Pz/2, 0 <= v <= u <= Pxy/2, pre-scaled by 1/P.
#pragma once

#include <vector>

#include "msc_common.hpp"

namespace msc {

constexpr int kZDma = 256;    // z tile code flag: persistent direct-to-LDS pass 3
constexpr int kZPrune = 512;  // z tile code flag: pass 3 with pruned z FFTs, HBM-direct
constexpr int kNE = 45;  // independent entries of the 9x9 spectral propagator

// Component sign flips of the lattice reflections (z, x, y) acting on the
// 9-vector (u_z, u_x, u_y, e_zz, e_xx, e_yy, e_xy, e_zy, e_zx):
// a component flips when it carries an odd number of the reflected index.
// The propagator block obeys P(R d) = S P(d) S (S the flip signs), in space
// and in k.
constexpr unsigned kFlipZ = (1u << 0) | (1u << 7) | (1u << 8);
constexpr unsigned kFlipX = (1u << 1) | (1u << 6) | (1u << 8);
constexpr unsigned kFlipY = (1u << 2) | (1u << 6) | (1u << 7);

struct Krylov;  // GMRES workspace (msc_gmres.hip)

struct Plan {
  int M = 0, Nz = 0;
  double a = 0.0;
  bool periodic = false;
  int Pz = 0, Pxy = 0;
  size_t nvox = 0, n = 0;  // voxels, complex unknowns (9 per voxel)
  // the largest slab the buffers hold (msc_plan_reshape stays within it)
  int capM = 0, capNz = 0;
  size_t cap_n = 0;
  int NC = 0, NUV = 0;

  double omega = 0.0, alpha = 0.0, beta = 0.0, rho = 0.0;
  // tiles per precision (0 complex128, 1 complex64) and pass (y-forward B,
  // x-forward Q, z Q, x-inverse Q, y-inverse B); see autotune().  A z code
  // Q | kZDma selects the persistent direct-to-LDS form of pass 3.
  int tile[2][5] = {{1, 1, 8, 1, 1}, {1, 2, 8, 2, 1}};
  bool tuned[2] = {false, false};
  bool have_table = false, have_tparams = false;

  double2* table128 = nullptr;
  float2* table64 = nullptr;
  double2* tp128 = nullptr;
  float2* tp64 = nullptr;
  // max over voxels of |t_u|, |t_d|, |t_o|, |t_s| (device, float bits; set with T0),
  // or over the entries of the general T-matrices
  unsigned int* tmax = nullptr;
  // general per-voxel T (msc_set_tgeneral): voxel v applies the 9x9 matrix
  // tgen[tidx[v]] (row-major, tau = T psi) instead of the 4-parameter cubic T0
  bool general = false;
  int ngen = 0, gen_cap = 0;
  int* tidx = nullptr;         // capvox
  double2* tgen128 = nullptr;  // gen_cap x 81
  float2* tgen64 = nullptr;
  int* flag = nullptr;         // device scratch flag (index validation)
  void* buf1 = nullptr;  // (Nz, M, Pxy, 9) complex128 bytes
  void* buf2 = nullptr;  // (Nz, Pxy, Pxy, 9) complex128 bytes
  // G0 table build: two octant workspaces (NC, Pxy/2 + 1, Pxy/2 + 1, 45)
  // complex128, kept for rebuilds; plans of one grid whose builds run on one
  // stream may share them (owns_work false on all but one)
  void* work = nullptr;
  void* work2 = nullptr;
  bool owns_work = true;
  // volume-averaged nearest-neighbour correction: for the 26 offsets R in
  // {-1,0,1}^3 \ 0 (lexicographic (dz, dx, dy)), the 9x9 difference between the
  // volume-averaged block and the point propagator the table holds, row-major
  bool have_nn = false;
  double2* nn128 = nullptr;  // 26 x 81
  float2* nn64 = nullptr;
  double* nn_stage = nullptr;  // 26 x 81 real volume-averaged blocks (device copy)
  // z-pass tile order per column tile width Q = 1, 2, 4, 8, 16 (concatenated,
  // zmap_offset): dispatch slot -> tile kx * (Pxy / Q) + chunk, grouping the
  // rows kx and Pxy - kx (which read the same table entries) on one XCD, with
  // the mirrored ky chunks next to each other in time.
  int* zmap = nullptr;
  // per-workgroup partial source sums (source_sums), allocated on first use
  double2* sums_part = nullptr;
  size_t sums_cap = 0;
  double2* tw128 = nullptr;
  float2* tw64 = nullptr;

  int device = 0;

  // HIP-event timing of the block-multiply pass (pass 3), one pair per launch
  bool timing = false;
  std::vector<hipEvent_t> ev_pool;
  std::vector<int> ev_prec;  // 8 (complex64) or 16 (complex128) per event pair
  size_t ev_used = 0;
};

// Event pair around one pass-3 launch when p.timing is set.
int timing_begin(Plan& p, hipStream_t st, size_t* slot, int elem_bytes);
int timing_end(Plan& p, hipStream_t st, size_t slot);

// Matvec y = (I - G0 T0) x in precision R on the caller's stream.  `done`
// (device int, may be null) turns every launch into a no-op when set.
template <typename R>
int matvec(Plan& p, const cx<R>* x, cx<R>* y, hipStream_t st, const int* done);

// Residual r = b - (I - G0 T0) x (fused into the last FFT pass).
template <typename R>
int residual(Plan& p, const cx<R>* b, const cx<R>* x, cx<R>* r, hipStream_t st, const int* done);

// Scattered part y = G0 T0 x = x - (I - G0 T0) x (the operator GMRES iterates on).
template <typename R>
int scattered(Plan& p, const cx<R>* x, cx<R>* y, hipStream_t st, const int* done);

// Time every tile candidate of every pass for precision R in place (x, out:
// n-element scratch vectors) and keep the fastest; sets p.tuned[prec].
template <typename R>
int autotune(Plan& p, const cx<R>* x, cx<R>* out, hipStream_t st);

// Generic strided batched FFT used by the table build:
// data viewed as [outer][N][inner], transform along N for every inner index.
template <typename R>
int fft_axis(const cx<R>* in, cx<R>* out, int outer, int N, size_t inner, bool inverse,
             const cx<R>* tw, hipStream_t st);
// The same keeping only the first nout frequencies: out is [outer][nout][inner]
// (not in place unless nout == N).
template <typename R>
int fft_axis_pruned(const cx<R>* in, cx<R>* out, int outer, int N, size_t inner, int nout,
                    bool inverse, const cx<R>* tw, hipStream_t st);

int build_table(Plan& p, double omega, double alpha, double beta, double rho, hipStream_t st);
// volume_averaged: set the nearest-neighbour correction for the plan's table
// (after build_table, same frequency and medium); otherwise clear it.
int set_nn_correction(Plan& p, bool volume_averaged, int n_orders, hipStream_t st);
// One instance of a batched operator launch (device copy): its T0, FFT
// buffers and G0 table, the input in + j * in_step and output out + j *
// out_step (j: the Arnoldi step of the launch, 0 for a residual), the
// residual's right-hand side and the done flag of its Krylov cycle.
template <typename R>
struct BPass {
  const cx<R>* tp;
  cx<R>* buf1;
  cx<R>* buf2;
  const cx<R>* table;
  const cx<R>* in;
  cx<R>* out;
  const cx<R>* bvec;
  size_t in_step, out_step;
  const int* done;
  int M, Nz;
};
// The operator on nsys instances of `ref`'s padded grid (Pz, Pxy, periodic)
// in one launch per pass, with `ref`'s tiles (the persistent z form maps to
// its one-tile-per-workgroup width): mode 1 out = bvec - in + acc (the
// residual), 2 out = acc (G0 T0 in).  Cubic T0 only, no near-neighbour
// correction.  max_M / max_Nz: the largest instance.
template <typename R>
int apply_batch(const Plan& ref, const BPass<R>* d, int nsys, int max_M, int max_Nz, int mode,
                int j, hipStream_t st);

// Offset of the tile order for width Q (a power of two <= 16) in Plan::zmap.
inline size_t zmap_offset(int Pxy, int Q) {
  size_t off = 0;
  for (int q = 1; q < Q; q <<= 1) off += (size_t)Pxy * (Pxy / q);
  return off;
}
int build_zmap(Plan& p);

}  // namespace msc

struct msc_plan;

namespace msc {
// Allocate a plan's GMRES workspaces for its capacity (restart m; the
// complex64 basis and vectors too when `mixed`), so that solves of any slab
// the plan can be re-shaped to reuse them.
int plan_reserve(msc_plan* pl, int m, bool mixed);
// Let `dst` use the G0-build workspaces of `src` (same padded grid; their
// builds must run on one stream, one after another).
int plan_share_build_work(msc_plan* dst, msc_plan* src);
}  // namespace msc
