// Batched sweep of independent slab solves (msc_sweep, include/msc_slab.h).
//
// Replaces the sequential per-frequency loop of compute_ocean_bottom_reflection
// (cubic_scattering/ocean_bottom.py:290-307), where every frequency rebuilds
// the host T0 / G0 and runs one solve.  Here `workers` host threads pull
// instances from a shared counter; each thread owns a non-blocking HIP stream
// and a cache of plans keyed by padded FFT grid (P_z, P_xy, periodic).  A plan
// is created once at the largest slab its grid holds and re-shaped to each
// instance (msc_plan_reshape), so buffers are allocated once per grid and
// worker: no allocation (and no allocation-driven device synchronisation) in
// the steady state.  The G0 table is rebuilt only when the frequency, medium,
// cube size or slab change; tiles are autotuned once per plan (per grid).
// Independent solves on the workers' streams overlap on the device, which
// fills the GPU when the slabs are small.  Arnoldi steps run as individual
// launches (no HIP-graph capture, which must not overlap the other workers'
// allocations); the overlap of the workers hides the launch cost.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/msc_slab.h"

namespace {

int next_pow2(int v) {
  int p = 1;
  while (p < v) p <<= 1;
  return p;
}

bool is_pow2(int v) { return v > 0 && (v & (v - 1)) == 0; }

struct CachedPlan {
  int pz = 0, pxy = 0, periodic = 0;
  msc_plan* plan = nullptr;
  void* b = nullptr;  // psi0 workspace (capacity of the plan, complex128)
  // table key (valid while the plan's slab is unchanged)
  bool have_table = false;
  int M = 0, Nz = 0, va = 0, n_orders = 0;
  double a = 0, omega = 0, alpha = 0, beta = 0, rho = 0;
};

struct Shared {
  msc_instance* inst = nullptr;
  int count = 0;
  const msc_gmres_opts* opts = nullptr;
  int device = 0;
  std::atomic<int> next{0};
};

void record(msc_instance& I, int rc, const char* what) {
  I.status = rc;
  const char* msg = msc_last_error();
  std::snprintf(I.error, sizeof(I.error), "%s: %s", what, msg ? msg : "");
}

int run_one(msc_instance& I, std::vector<CachedPlan>& cache, const msc_gmres_opts& o,
            hipStream_t st) {
  const int pz = std::max(2, next_pow2(2 * I.N_z - 1));
  const bool circ = I.periodic && is_pow2(I.M) && I.M >= 2;  // periodic on the M-point grid
  const int pxy = circ ? I.M : std::max(2, next_pow2(2 * I.M - 1));
  CachedPlan* cp = nullptr;
  for (CachedPlan& c : cache)
    if (c.pz == pz && c.pxy == pxy && c.periodic == I.periodic) cp = &c;
  if (cp == nullptr) {
    CachedPlan c;
    c.pz = pz;
    c.pxy = pxy;
    c.periodic = I.periodic;
    // the largest slab of this grid, so that any instance of the grid fits
    const int capM = circ ? I.M : pxy / 2;
    const int capNz = pz / 2;
    int rc = msc_plan_create(&c.plan, std::max(capM, 1), std::max(capNz, 1), I.a, I.periodic,
                             pz, pxy);
    if (!rc) rc = msc_plan_use_graphs(c.plan, 0);  // no capture beside other workers
    if (rc) {
      record(I, rc, "plan");
      return rc;
    }
    const size_t capn = (size_t)std::max(capNz, 1) * capM * capM * 9;
    if (hipMalloc(&c.b, capn * 16) != hipSuccess) {
      msc_plan_destroy(c.plan);
      I.status = MSC_ERR_HIP;
      std::snprintf(I.error, sizeof(I.error), "plan: out of device memory");
      return MSC_ERR_HIP;
    }
    c.M = capM;
    c.Nz = capNz;
    c.a = I.a;
    cache.push_back(c);
    cp = &cache.back();
  }
  int rc = msc_plan_reshape(cp->plan, I.M, I.N_z, I.a);
  if (rc) {
    record(I, rc, "reshape");
    return rc;
  }
  const bool same_table = cp->have_table && cp->M == I.M && cp->Nz == I.N_z && cp->a == I.a &&
                          cp->omega == I.omega && cp->alpha == I.alpha &&
                          cp->beta == I.beta && cp->rho == I.rho &&
                          cp->va == I.volume_averaged &&
                          (!I.volume_averaged || cp->n_orders == I.n_orders);
  if (!same_table) {
    rc = msc_build_g0(cp->plan, I.omega, I.alpha, I.beta, I.rho, I.volume_averaged, I.n_orders,
                      st);
    if (rc) {
      cp->have_table = false;
      record(I, rc, "G0 table");
      return rc;
    }
    cp->have_table = true;
    cp->va = I.volume_averaged;
    cp->n_orders = I.n_orders;
    cp->omega = I.omega;
    cp->alpha = I.alpha;
    cp->beta = I.beta;
    cp->rho = I.rho;
  }
  cp->M = I.M;
  cp->Nz = I.N_z;
  cp->a = I.a;
  double coef[8];
  rc = msc_cube_coefficients(I.omega, I.a, I.alpha, I.beta, I.rho, coef);
  if (!rc) rc = msc_set_tmatrix(cp->plan, I.dlambda, I.dmu, I.drho, I.omega, coef, st);
  if (rc) {
    record(I, rc, "T0");
    return rc;
  }
  rc = msc_incident_field(cp->plan, I.k_mag, I.khat, I.inc18, cp->b, st);
  if (rc) {
    record(I, rc, "incident field");
    return rc;
  }
  if (I.psi0 != nullptr &&
      hipMemcpyAsync(I.psi0, cp->b, (size_t)I.N_z * I.M * I.M * 9 * 16, hipMemcpyDeviceToDevice,
                     st) != hipSuccess) {
    I.status = MSC_ERR_HIP;
    std::snprintf(I.error, sizeof(I.error), "copy of the incident field failed");
    return MSC_ERR_HIP;
  }
  rc = msc_gmres(cp->plan, cp->b, I.psi, &o, &I.stats, st);
  if (rc) {
    record(I, rc, "gmres");
    return rc;
  }
  if (I.layer_sum != nullptr && I.phased != nullptr) {
    const double rhat[3] = {-I.khat[0], -I.khat[1], -I.khat[2]};
    rc = msc_source_sums(cp->plan, I.psi, I.omega / I.alpha, I.omega / I.beta, rhat,
                         I.layer_sum, I.phased, st);
    if (rc) {
      record(I, rc, "source sums");
      return rc;
    }
  }
  if (hipStreamSynchronize(st) != hipSuccess) {
    I.status = MSC_ERR_HIP;
    std::snprintf(I.error, sizeof(I.error), "stream synchronise failed");
    return MSC_ERR_HIP;
  }
  I.status = MSC_OK;
  I.error[0] = '\0';
  return MSC_OK;
}

void worker(Shared* sh) {
  if (hipSetDevice(sh->device) != hipSuccess) return;
  hipStream_t st = nullptr;
  if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
    for (int i = sh->next.fetch_add(1); i < sh->count; i = sh->next.fetch_add(1)) {
      sh->inst[i].status = MSC_ERR_HIP;
      std::snprintf(sh->inst[i].error, sizeof(sh->inst[i].error), "stream creation failed");
    }
    return;
  }
  std::vector<CachedPlan> cache;
  cache.reserve(16);
  for (int i = sh->next.fetch_add(1); i < sh->count; i = sh->next.fetch_add(1))
    (void)run_one(sh->inst[i], cache, *sh->opts, st);
  (void)hipStreamSynchronize(st);
  for (CachedPlan& c : cache) {
    if (c.b) (void)hipFree(c.b);
    msc_plan_destroy(c.plan);
  }
  (void)hipStreamDestroy(st);
}

}  // namespace

extern "C" int msc_sweep(msc_instance* instances, int count, const msc_gmres_opts* opts,
                         int workers) {
  if (instances == nullptr || opts == nullptr || count < 0 || workers < 1 ||
      opts->precond != nullptr) {
    return MSC_ERR_INVALID;
  }
  for (int i = 0; i < count; ++i) {
    msc_instance& I = instances[i];
    I.status = MSC_ERR_NOT_READY;
    std::snprintf(I.error, sizeof(I.error), "not run");
    std::memset(&I.stats, 0, sizeof(I.stats));
    const bool ok = I.M >= 1 && I.N_z >= 1 && I.a > 0.0 && I.omega > 0.0 && I.alpha > 0.0 &&
                    I.beta > 0.0 && I.rho > 0.0 && I.dlambda && I.dmu && I.drho && I.psi &&
                    (I.periodic == 0 || I.periodic == 1) &&
                    (I.volume_averaged == 0 ||
                     (I.volume_averaged == 1 && I.n_orders >= 0 && I.n_orders <= 3));
    if (!ok) {
      I.status = MSC_ERR_INVALID;
      std::snprintf(I.error, sizeof(I.error), "invalid instance %d", i);
      return MSC_ERR_INVALID;
    }
  }
  Shared sh;
  sh.inst = instances;
  sh.count = count;
  sh.opts = opts;
  if (hipGetDevice(&sh.device) != hipSuccess) return MSC_ERR_HIP;
  const int nw = std::max(1, std::min(workers, count));
  std::vector<std::thread> threads;
  threads.reserve(nw);
  for (int w = 0; w < nw; ++w) threads.emplace_back(worker, &sh);
  for (std::thread& t : threads) t.join();
  for (int i = 0; i < count; ++i)
    if (instances[i].status != MSC_OK) return instances[i].status;
  return MSC_OK;
}
