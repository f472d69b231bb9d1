// GMRES workspace and building blocks (device-resident Arnoldi, FP64 state).
#pragma once

#include <vector>

#include "msc_common.hpp"

namespace msc {

enum OrthMode : int { kOrthAdaptive = 0, kOrthCGS1 = 1, kOrthCGS2 = 2 };

// Device-resident step state of one GMRES cycle (FP64).
struct KState {
  int done;          // set when the cycle must stop (converged, breakdown, full)
  int j;             // Arnoldi steps completed in the cycle
  int reorth;        // second Gram-Schmidt pass requested for the current step
  int pad;
  double rnorm;      // ||r|| of the cycle's starting residual
  double tol_abs;    // stop when |g_{j+1}| <= tol_abs
  double wn_before;  // ||w||^2 before the first projection
  double resid_est;  // |g_{j+1}| after the last step
  double scale;      // 1 / ||w|| of the last closed step (0 on breakdown)
  int* done_host;    // device view of a pinned host flag mirroring `done`
};

struct Krylov {
  size_t n = 0, ldv = 0;
  size_t cap_n = 0;           // n the buffers were allocated for (n <= cap_n)
  int m = 0, ldh = 0, nblk = 0, elem_bytes = 0;
  void* V = nullptr;          // (m+1) x ldv basis, complex64 or complex128
  KState* st = nullptr;
  double2* H = nullptr;       // (m+1) x m Hessenberg, column-major
  double2* cs = nullptr;      // rotations: cs[2i] = (c, 0), cs[2i+1] = s
  double2* g = nullptr;       // rotated right-hand side (m+1)
  double2* hcur = nullptr;    // projection coefficients of the current pass
  double2* y = nullptr;       // least-squares solution
  double2* tmp = nullptr;
  double* scal = nullptr;     // 4 device scalars (norms, scale factors)
  double2* part = nullptr;    // per-workgroup partials
  double* hnext = nullptr;    // ||w|| of each step (unrotated H[j+1, j])
  void* host = nullptr;       // pinned staging of KState
  int* done_flag = nullptr;   // pinned, device-mapped mirror of KState::done
  bool owns_V = false;
  void* Z = nullptr;          // flexible GMRES: m x ldv preconditioned vectors z_j = P(v_j)
  bool shifted = false;  // basis built from B = G0 T0, operator A = I - B
  // instantiated HIP graphs of whole Arnoldi steps (operator + orthogonalisation),
  // indexed by step; valid while the buffers and `graph_orth` are unchanged
  std::vector<hipGraphExec_t> graphs;
  int graph_orth = -1;
  unsigned long long graph_net = 0;  // generation of the preconditioner the graphs hold (0: none)
  void drop_graphs();

  ~Krylov();
  void release();
  // state only (caller-owned basis) / state plus an owned basis; a smaller n
  // than allocated reuses the buffers (graphs are dropped)
  int ensure_state(size_t n, int m, int elem_bytes);
  int ensure(size_t n, int m, int elem_bytes);
  // allocate Z (flexible GMRES) for the current n, m, precision
  int ensure_z();
};

// One system of a batched launch (device copy, filled by the host from its
// Krylov): the batched kernels take an array of these, system = blockIdx.y,
// and every system runs the single-system kernel body on its own workspace.
// `r` is the vector a norm or a cycle start reads (precision of the call),
// `x` the FP64 solution a cycle's update adds to, `src` / `dst` the
// complex128 -> complex64 narrowing pair; `tol_abs` the system's cycle
// tolerance.
struct BKry {
  void* V;
  size_t ldv, n;
  int nblk, m, ldh, shifted;
  KState* st;
  double2 *H, *cs, *g, *hcur, *y, *tmp, *part;
  double* hnext;
  const void* r;
  double2* x;
  const double2* src;
  float2* dst;
  double tol_abs;
};
BKry bkry_of(const Krylov& k);
// Batched counterparts of the calls below over `nsys` systems (max_n / max_nblk:
// the largest vector length / chunk count among them).
template <typename R>
int bkrylov_step(const BKry* d, int nsys, size_t max_n, int max_nblk, int j, int orth_mode,
                 hipStream_t s);
template <typename R>
int bkrylov_norm(const BKry* d, int nsys, int max_nblk, hipStream_t s);
template <typename RS, typename R>
int bkrylov_cycle_init(const BKry* d, int nsys, size_t max_n, bool unit_rhs, hipStream_t s);
template <typename R>
int bkrylov_finish(const BKry* d, int nsys, size_t max_n, hipStream_t s);
int bkrylov_narrow(const BKry* d, int nsys, size_t max_n, hipStream_t s);
// out[2 i] = ||r|| of system i's cycle (KState::rnorm), out[2 i + 1] = its
// completed steps (KState::j)
int bkrylov_gather(const BKry* d, int nsys, double* out, hipStream_t s);
// systems with active[i] == 0 start their next cycle done (rnorm = 0)
int bkrylov_deactivate(const BKry* d, int nsys, const int* active, hipStream_t s);

template <typename R>
int krylov_step(Krylov& k, int j, int orth_mode, hipStream_t s);
template <typename R>
int krylov_norm(Krylov& k, const cx<R>* r, double* out_dev, hipStream_t s);
template <typename RS, typename R>
int krylov_cycle_init(Krylov& k, const cx<RS>* r, double tol_abs, bool unit_rhs, hipStream_t s);
// x += (scale) * sum_{v<j} y_v B_v with B = V (GMRES) or Z (flexible GMRES)
template <typename R, typename RX>
int krylov_finish(Krylov& k, cx<RX>* x, const double* scale_dev, hipStream_t s,
                  bool flexible = false);
int krylov_read_state(Krylov& k, KState* out, hipStream_t s);
int krylov_narrow(const double2* in, float2* out, size_t n, hipStream_t s);

}  // namespace msc
