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
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/msc_slab.h"
#include "msc_plan.hpp"

namespace {

int next_pow2(int v) {
  int p = 1;
  while (p < v) p <<= 1;
  return p;
}

bool is_pow2(int v) { return v > 0 && (v & (v - 1)) == 0; }

struct CachedPlan {
  int pz = 0, pxy = 0, periodic = 0;
  bool circ = false;  // periodic on the M-point grid (the plan cannot change M)
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
    if (c.pz == pz && c.pxy == pxy && c.periodic == I.periodic && c.circ == circ &&
        (!circ || c.M == I.M))
      cp = &c;
  if (cp == nullptr) {
    CachedPlan c;
    c.pz = pz;
    c.pxy = pxy;
    c.periodic = I.periodic;
    c.circ = circ;
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

int validate(msc_instance* instances, int count) {
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
  return MSC_OK;
}

}  // namespace

extern "C" int msc_sweep(msc_instance* instances, int count, const msc_gmres_opts* opts,
                         int workers) {
  if (instances == nullptr || opts == nullptr || count < 0 || workers < 1 ||
      opts->precond != nullptr) {
    return MSC_ERR_INVALID;
  }
  int rc0 = validate(instances, count);
  if (rc0) return rc0;
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

// ─────────────────────────────────────────────────────────────────────────────
// Batched sweep (msc_sweep_batched): the instances of one padded FFT grid
// (P_z, P_xy, periodic) are solved together by msc_gmres_batched, so every
// operator pass and Arnoldi kernel of a GMRES step is one launch for a whole
// batch, whatever the slab sizes.  A batch holds enough slabs to fill the
// device (about kBatchPoints padded grid points per pass), fewer if device
// memory is short.  Per grid a pool of plans at the grid's largest slab is
// re-shaped per instance (GMRES workspaces reserved at that capacity, one
// G0-build workspace shared by the pool); pools are kept between sweeps in a
// bounded process-wide cache (40% of the device memory; msc_sweep_release
// frees them), so a repeated
// sweep allocates nothing.  The set-up of the next batch (G0 table, T0,
// incident field) runs on a helper thread and stream while the current batch
// solves: a grid with more than one batch has a pool of two halves, the
// batches alternating between them.  Instances with the volume-averaged
// correction, which the batched operator does not apply, are solved one at a
// time at the end.
// ─────────────────────────────────────────────────────────────────────────────
namespace {

constexpr size_t kBatchPoints = size_t(1) << 23;

struct Slot {
  msc_plan* plan = nullptr;
  void* b = nullptr;  // right-hand side (psi0), complex128, plan capacity
  bool have_table = false;
  int M = 0, Nz = 0;
  double a = 0, omega = 0, alpha = 0, beta = 0, rho = 0;
};

// circ: periodic on the M-point grid (every slab of the group has M = pxy);
// otherwise the zero-padded grid, periodic or not
struct GridKey {
  int pz, pxy, periodic;
  bool circ;
  bool operator==(const GridKey& o) const {
    return pz == o.pz && pxy == o.pxy && periodic == o.periodic && circ == o.circ;
  }
};

GridKey grid_of(const msc_instance& I) {
  const int pz = std::max(2, next_pow2(2 * I.N_z - 1));
  const bool circ = I.periodic && is_pow2(I.M) && I.M >= 2;
  const int pxy = circ ? I.M : std::max(2, next_pow2(2 * I.M - 1));
  return {pz, pxy, I.periodic, circ};
}

// Device bytes of one pool plan at the grid's capacity (buffers, tables, T0,
// right-hand side, GMRES basis and vectors).
size_t plan_bytes(const GridKey& g, int capM, int capNz, int m, bool mixed) {
  const size_t capn = (size_t)capNz * capM * capM * 9;
  const size_t ldv = (capn + 63) & ~size_t(63);
  const size_t h = g.pxy / 2;
  const size_t nuv = (h + 1) * (h + 2) / 2, nc = g.pz / 2 + 1;
  size_t b = (size_t)capNz * capM * g.pxy * 9 * 16 + (size_t)capNz * g.pxy * g.pxy * 9 * 16 +
             nuv * 45 * nc * 24 + (size_t)capNz * capM * capM * 4 * 24 + capn * 16 * 2;
  b += mixed ? (size_t)(m + 1) * ldv * 8 + capn * 16 : (size_t)(m + 1) * ldv * 16;
  return b + (size_t)(8 << 20);
}

struct Pool {
  int device = 0;
  GridKey g{};
  int capM = 0, capNz = 0;
  std::vector<Slot> slots;
  size_t bytes = 0;  // estimate, for the cache bound
};

void destroy_slots(std::vector<Slot>& slots) {
  // the shared build workspace belongs to slot 0, which goes last
  for (size_t i = slots.size(); i-- > 0;) {
    if (slots[i].b) (void)hipFree(slots[i].b);
    msc_plan_destroy(slots[i].plan);
  }
  slots.clear();
}

std::mutex g_pool_mu;
std::vector<Pool> g_pools;  // idle pools, at most one per (device, grid)

size_t cache_budget() {
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return 0;
  return (size_t)(0.4 * (double)total_b);
}

// A pool on grid g holding slabs up to capM x capNz with at least `want`
// slots: the cached one when it is large enough, grown or created otherwise.
int acquire_pool(int dev, const GridKey& g, int capM, int capNz, double a, int m, bool mixed,
                 int want, Pool& out) {
  std::vector<Pool> drop;
  {
    std::lock_guard<std::mutex> lock(g_pool_mu);
    for (size_t i = 0; i < g_pools.size(); ++i) {
      Pool& c = g_pools[i];
      if (c.device != dev || !(c.g == g)) continue;
      if (c.capM >= capM && c.capNz >= capNz) out = std::move(c);
      else drop.push_back(std::move(c));
      g_pools.erase(g_pools.begin() + (long)i);
      break;
    }
  }
  for (Pool& d : drop) destroy_slots(d.slots);
  if (out.slots.empty()) {
    out.device = dev;
    out.g = g;
    out.capM = capM;
    out.capNz = capNz;
  }
  const size_t per = plan_bytes(g, out.capM, out.capNz, m, mixed);
  while ((int)out.slots.size() < want) {
    Slot sl;
    int rc = msc_plan_create(&sl.plan, out.capM, out.capNz, a, g.periodic, g.pz, g.pxy);
    if (!rc && !out.slots.empty()) rc = msc::plan_share_build_work(sl.plan, out.slots[0].plan);
    if (!rc && hipMalloc(&sl.b, (size_t)out.capNz * out.capM * out.capM * 9 * 16) != hipSuccess) {
      sl.b = nullptr;
      rc = MSC_ERR_HIP;
    }
    if (rc) {
      if (sl.b) (void)hipFree(sl.b);
      msc_plan_destroy(sl.plan);
      if (out.slots.empty()) return rc;
      break;  // a smaller pool
    }
    out.slots.push_back(sl);
  }
  for (Slot& sl : out.slots) {
    const int rc = msc::plan_reserve(sl.plan, m, mixed);
    if (rc) return rc;
  }
  out.bytes = per * out.slots.size();
  return MSC_OK;
}

// Back to the cache (tables marked stale: every sweep builds its own G0),
// evicting the oldest idle pools beyond the budget.
void release_pool(Pool&& p) {
  if (p.slots.empty()) return;
  for (Slot& sl : p.slots) sl.have_table = false;
  std::vector<Pool> drop;
  {
    std::lock_guard<std::mutex> lock(g_pool_mu);
    g_pools.push_back(std::move(p));
    const size_t budget = cache_budget();
    size_t total = 0;
    for (const Pool& c : g_pools) total += c.bytes;
    while (total > budget && !g_pools.empty()) {
      total -= g_pools.front().bytes;
      drop.push_back(std::move(g_pools.front()));
      g_pools.erase(g_pools.begin());
    }
  }
  for (Pool& d : drop) destroy_slots(d.slots);
}

// G0 (when the slot's table does not fit the instance), T0 and the incident
// field of instance I in slot s, on stream st.
int setup_instance(msc_instance& I, Slot& s, hipStream_t st) {
  int rc = msc_plan_reshape(s.plan, I.M, I.N_z, I.a);
  if (rc) {
    record(I, rc, "reshape");
    s.have_table = false;
    return rc;
  }
  const bool same = s.have_table && s.M == I.M && s.Nz == I.N_z && s.a == I.a &&
                    s.omega == I.omega && s.alpha == I.alpha && s.beta == I.beta &&
                    s.rho == I.rho;
  if (!same) {
    rc = msc_build_g0(s.plan, I.omega, I.alpha, I.beta, I.rho, 0, 0, st);
    if (rc) {
      s.have_table = false;
      record(I, rc, "G0 table");
      return rc;
    }
    s.have_table = true;
    s.M = I.M;
    s.Nz = I.N_z;
    s.a = I.a;
    s.omega = I.omega;
    s.alpha = I.alpha;
    s.beta = I.beta;
    s.rho = I.rho;
  }
  double coef[8];
  rc = msc_cube_coefficients(I.omega, I.a, I.alpha, I.beta, I.rho, coef);
  if (!rc) rc = msc_set_tmatrix(s.plan, I.dlambda, I.dmu, I.drho, I.omega, coef, st);
  if (rc) {
    record(I, rc, "T0");
    return rc;
  }
  rc = msc_incident_field(s.plan, I.k_mag, I.khat, I.inc18, s.b, st);
  if (rc) {
    record(I, rc, "incident field");
    return rc;
  }
  if (I.psi0 != nullptr &&
      hipMemcpyAsync(I.psi0, s.b, (size_t)I.N_z * I.M * I.M * 9 * 16, hipMemcpyDeviceToDevice,
                     st) != hipSuccess) {
    I.status = MSC_ERR_HIP;
    std::snprintf(I.error, sizeof(I.error), "copy of the incident field failed");
    return MSC_ERR_HIP;
  }
  return MSC_OK;
}

struct Group {
  GridKey g{};
  std::vector<int> idx;
  int capM = 1, capNz = 1, chunk = 1, nchunks = 1;
  Pool pool;
  bool have_pool = false;
};

struct Job {
  int group;
  size_t c0, c1;  // instances idx[c0..c1) of the group
  int half;       // pool slots [half * chunk, (half + 1) * chunk)
  std::vector<int> ready;  // instances whose set-up succeeded
};

}  // namespace

extern "C" int msc_sweep_batched(msc_instance* instances, int count, const msc_gmres_opts* opts,
                                 int max_batch, void* stream) {
  if (instances == nullptr || opts == nullptr || count < 0 || max_batch < 1 ||
      opts->precond != nullptr) {
    return MSC_ERR_INVALID;
  }
  int rc = validate(instances, count);
  if (rc) return rc;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return MSC_ERR_HIP;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const msc_gmres_opts& o = *opts;
  const int m = std::max(1, std::min(o.restart, std::max(o.max_iter, 1)));
  const bool mixed = o.mode == MSC_SOLVE_MIXED;

  std::vector<Group> groups;
  std::vector<int> single;
  for (int i = 0; i < count; ++i) {
    if (instances[i].volume_averaged) {
      single.push_back(i);
      continue;
    }
    const GridKey g = grid_of(instances[i]);
    size_t k = 0;
    while (k < groups.size() && !(groups[k].g == g)) ++k;
    if (k == groups.size()) {
      groups.emplace_back();
      groups[k].g = g;
    }
    groups[k].idx.push_back(i);
  }
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = (size_t)16 << 30;
  std::vector<Job> jobs;
  for (size_t k = 0; k < groups.size(); ++k) {
    Group& G = groups[k];
    for (int i : G.idx) {
      G.capM = std::max(G.capM, instances[i].M);
      G.capNz = std::max(G.capNz, instances[i].N_z);
    }
    if (G.g.circ) G.capM = G.g.pxy;
    const size_t per = plan_bytes(G.g, G.capM, G.capNz, m, mixed);
    const size_t pts = (size_t)G.g.pz * G.g.pxy * G.g.pxy;
    const size_t want = std::max<size_t>(1, (kBatchPoints + pts - 1) / pts);
    // two halves when the group needs more than one batch: 0.3 of the free
    // memory each at most
    const size_t fit = std::max<size_t>(1, (size_t)(0.3 * (double)free_b) / per);
    const size_t cap = std::min<size_t>({fit, want, (size_t)max_batch, G.idx.size()});
    // batches of equal size (17 slabs in 9 + 8, not 16 + 1)
    G.nchunks = (int)((G.idx.size() + cap - 1) / cap);
    G.chunk = (int)((G.idx.size() + G.nchunks - 1) / G.nchunks);
    for (int c = 0; c < G.nchunks; ++c)
      jobs.push_back(Job{(int)k, (size_t)c * G.chunk,
                         std::min(G.idx.size(), (size_t)(c + 1) * G.chunk), c % 2, {}});
  }

  // set-up of job j on stream sst (helper thread); returns a status for the
  // whole job (instance failures are recorded per instance)
  auto setup_job = [&](size_t j, hipStream_t sst) -> int {
    Job& J = jobs[j];
    Group& G = groups[J.group];
    if (!G.have_pool) {
      const int want = G.chunk * (G.nchunks > 1 ? 2 : 1);
      const int r = acquire_pool(dev, G.g, G.capM, G.capNz, instances[G.idx[0]].a, m, mixed,
                                 want, G.pool);
      if (r) {
        for (int i : G.idx) record(instances[i], r, "batched sweep plan");
        destroy_slots(G.pool.slots);
        return r;
      }
      G.have_pool = true;
      // a pool that came out smaller: fewer instances per batch (the jobs
      // of this group were cut at G.chunk; re-cut is not needed when the
      // pool holds one half at least)
      if ((int)G.pool.slots.size() < want) {
        for (int i : G.idx) record(instances[i], MSC_ERR_HIP, "batched sweep plan");
        return MSC_ERR_HIP;
      }
    }
    J.ready.clear();
    for (size_t k = J.c0; k < J.c1; ++k) {
      msc_instance& I = instances[G.idx[k]];
      Slot& sl = G.pool.slots[(size_t)J.half * G.chunk + (k - J.c0)];
      if (setup_instance(I, sl, sst) == MSC_OK) J.ready.push_back((int)k);
    }
    return MSC_OK;
  };

  hipStream_t sst = nullptr;
  if (hipStreamCreateWithFlags(&sst, hipStreamNonBlocking) != hipSuccess) return MSC_ERR_HIP;
  hipEvent_t ev_set = nullptr, ev_done[2] = {nullptr, nullptr};
  bool ok_ev = hipEventCreateWithFlags(&ev_set, hipEventDisableTiming) == hipSuccess &&
               hipEventCreateWithFlags(&ev_done[0], hipEventDisableTiming) == hipSuccess &&
               hipEventCreateWithFlags(&ev_done[1], hipEventDisableTiming) == hipSuccess;
  int first = ok_ev ? MSC_OK : MSC_ERR_HIP;
  std::vector<msc_plan*> plans;
  std::vector<const void*> bs;
  std::vector<void*> xs;
  std::vector<msc_gmres_stats> stats;
  if (ok_ev && !jobs.empty()) {
    int set_rc = setup_job(0, sst);
    for (size_t j = 0; j < jobs.size(); ++j) {
      if (hipEventRecord(ev_set, sst) != hipSuccess) {
        first = first ? first : MSC_ERR_HIP;
        break;
      }
      Job& J = jobs[j];
      Group& G = groups[J.group];
      // the next job's set-up on the helper thread, after the solve that last
      // used its half of the pool has finished on the main stream
      std::thread helper;
      int next_rc = MSC_OK;
      if (j + 1 < jobs.size()) {
        helper = std::thread([&, j]() {
          if (hipSetDevice(dev) != hipSuccess) {
            next_rc = MSC_ERR_HIP;
            return;
          }
          if (j >= 1 && jobs[j + 1].group == jobs[j - 1].group &&
              hipStreamWaitEvent(sst, ev_done[jobs[j - 1].half], 0) != hipSuccess) {
            next_rc = MSC_ERR_HIP;
            return;
          }
          next_rc = setup_job(j + 1, sst);
        });
      }
      if (set_rc == MSC_OK && !J.ready.empty() &&
          hipStreamWaitEvent(st, ev_set, 0) == hipSuccess) {
        plans.clear();
        bs.clear();
        xs.clear();
        for (int k : J.ready) {
          Slot& sl = G.pool.slots[(size_t)J.half * G.chunk + (k - J.c0)];
          plans.push_back(sl.plan);
          bs.push_back(sl.b);
          xs.push_back(instances[G.idx[k]].psi);
        }
        stats.assign(plans.size(), msc_gmres_stats{});
        rc = msc_gmres_batched(plans.data(), bs.data(), xs.data(), (int)plans.size(), &o,
                               stats.data(), st);
        if (rc) {
          for (int k : J.ready) record(instances[G.idx[k]], rc, "batched gmres");
          first = first ? first : rc;
        } else {
          for (size_t q = 0; q < J.ready.size(); ++q) {
            msc_instance& I = instances[G.idx[J.ready[q]]];
            I.stats = stats[q];
            if (I.layer_sum != nullptr && I.phased != nullptr) {
              const double rhat[3] = {-I.khat[0], -I.khat[1], -I.khat[2]};
              const int r2 = msc_source_sums(plans[q], I.psi, I.omega / I.alpha,
                                             I.omega / I.beta, rhat, I.layer_sum, I.phased, st);
              if (r2) {
                record(I, r2, "source sums");
                first = first ? first : r2;
                continue;
              }
            }
            I.status = MSC_OK;
            I.error[0] = '\0';
          }
        }
      } else if (set_rc) {
        first = first ? first : set_rc;
      }
      if (hipEventRecord(ev_done[J.half], st) != hipSuccess) first = first ? first : MSC_ERR_HIP;
      if (helper.joinable()) helper.join();
      set_rc = next_rc;
      // the group's last batch: its pool goes back to the cache once the
      // main stream is past it
      if (j + 1 == jobs.size() || jobs[j + 1].group != J.group) {
        if (hipStreamSynchronize(st) != hipSuccess) first = first ? first : MSC_ERR_HIP;
        if (G.have_pool) release_pool(std::move(G.pool));
        G.have_pool = false;
      }
    }
  }
  (void)hipStreamSynchronize(sst);
  if (hipStreamSynchronize(st) != hipSuccess && !first) first = MSC_ERR_HIP;
  for (Group& G : groups)
    if (G.have_pool) release_pool(std::move(G.pool));
  if (ev_set) (void)hipEventDestroy(ev_set);
  for (hipEvent_t e : ev_done)
    if (e) (void)hipEventDestroy(e);
  (void)hipStreamDestroy(sst);
  if (!single.empty()) {
    std::vector<CachedPlan> cache;
    for (int i : single) {
      rc = run_one(instances[i], cache, o, st);
      if (rc && !first) first = rc;
    }
    for (CachedPlan& c : cache) {
      if (c.b) (void)hipFree(c.b);
      msc_plan_destroy(c.plan);
    }
  }
  for (int i = 0; i < count; ++i)
    if (instances[i].status != MSC_OK) return instances[i].status;
  return first;
}

extern "C" int msc_sweep_release(void) {
  std::vector<Pool> drop;
  {
    std::lock_guard<std::mutex> lock(g_pool_mu);
    drop.swap(g_pools);
  }
  for (Pool& d : drop) destroy_slots(d.slots);
  return MSC_OK;
}
