// C ABI of libmsc_slab.so (declared in include/msc_slab.h): plan lifecycle,
// host-side frequency constants, solver drivers (FP64 GMRES(m) and GMRES-IR).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/msc_precond.h"
#include "msc_gmres.hpp"
#include "msc_net.hpp"
#include "msc_plan.hpp"
#include "msc_setup.hpp"

namespace msc {

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }
const char* get_error() { return g_err.c_str(); }

// Twiddles tw[m] = exp(-2 pi i m / kTwMax), one copy per device.
struct Twiddles {
  double2* t128 = nullptr;
  float2* t64 = nullptr;
};
static std::deque<Twiddles> g_tw;  // stable element addresses
static std::mutex g_tw_mutex;        // plans may be created from several host threads

static int twiddles(int dev, Twiddles** out) {
  std::lock_guard<std::mutex> lock(g_tw_mutex);
  while ((int)g_tw.size() <= dev) g_tw.emplace_back();
  Twiddles& t = g_tw[dev];
  if (t.t128 == nullptr) {
    constexpr int N = 1024;  // kTwMax
    std::vector<double2> h128(N);
    std::vector<float2> h64(N);
    for (int m = 0; m < N; ++m) {
      const long double ang = -2.0L * 3.141592653589793238462643383279502884L * m / N;
      h128[m] = make_double2((double)cosl(ang), (double)sinl(ang));
      h64[m] = make_float2((float)cosl(ang), (float)sinl(ang));
    }
    MSC_HIP_CHECK(hipMalloc(&t.t128, N * sizeof(double2)));
    MSC_HIP_CHECK(hipMalloc(&t.t64, N * sizeof(float2)));
    MSC_HIP_CHECK(hipMemcpy(t.t128, h128.data(), N * sizeof(double2), hipMemcpyHostToDevice));
    MSC_HIP_CHECK(hipMemcpy(t.t64, h64.data(), N * sizeof(float2), hipMemcpyHostToDevice));
  }
  *out = &t;
  return kOk;
}

static int next_pow2(int v) {
  int p = 1;
  while (p < v) p <<= 1;
  return p;
}
static bool is_pow2(int v) { return v > 0 && (v & (v - 1)) == 0; }

// ── frequency constants of the Rayleigh cube T-matrix (host, FP64) ──────────
// Restates effective_contrasts.py:109-357: static Kelvin/Eshelby parts plus
// the smooth radiation parts from the Taylor coefficients phi_n, psi_n and the
// exact cube moments S0, S1, S2, S11 (N_TAYLOR = 8 terms).
struct cd {
  double re, im;
};
static cd cmul_h(cd a, cd b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
static cd cadd_h(cd a, cd b) { return {a.re + b.re, a.im + b.im}; }
static cd cscl_h(cd a, double s) { return {a.re * s, a.im * s}; }
static cd cpow_i(double k, int m) {  // (i k)^m
  double mag = std::pow(k, m);
  switch (m & 3) {
    case 0: return {mag, 0.0};
    case 1: return {0.0, mag};
    case 2: return {-mag, 0.0};
    default: return {0.0, -mag};
  }
}
static double fact(int n) {
  double f = 1.0;
  for (int i = 2; i <= n; ++i) f *= i;
  return f;
}

static void cube_coefficients(double omega, double a, double alpha, double beta, double rho,
                              cd out[4]) {
  constexpr int NT = 8;
  const double pi = M_PI;
  const double den = 8.0 * pi * rho * alpha * alpha * beta * beta;
  const double a0 = (alpha * alpha + beta * beta) / den;
  const double b0 = (alpha * alpha - beta * beta) / den;
  // g0 = int_{[-1,1]^3} 1/|x| (Pell form of the closed expression)
  const double g0 = (4.0 / 3.0) * std::log(70226.0 + 40545.0 * std::sqrt(3.0)) - 2.0 * pi;
  cd phi[NT], psi[NT];
  for (int n = 0; n < NT; ++n) {
    phi[n] = cscl_h(cpow_i(omega / beta, 2 * n + 1), 1.0 / (fact(2 * n + 1) * 4.0 * pi * rho * beta * beta));
    cd pa = cscl_h(cpow_i(omega / alpha, 2 * n + 3), 1.0 / (alpha * alpha));
    cd pb = cscl_h(cpow_i(omega / beta, 2 * n + 3), -1.0 / (beta * beta));
    psi[n] = cscl_h(cadd_h(pa, pb), 1.0 / (fact(2 * n + 3) * 4.0 * pi * rho));
  }
  // moments of [-a, a]^3: mu_k = 2 a^(2k+1)/(2k+1), trinomial expansion
  double mu[NT + 3];
  for (int k = 0; k < NT + 3; ++k) mu[k] = 2.0 * std::pow(a, 2 * k + 1) / (2 * k + 1);
  double S0[NT] = {0}, S1[NT] = {0}, S2[NT] = {0}, S11[NT] = {0};
  for (int m = 0; m < NT; ++m) {
    for (int p = 0; p <= m; ++p) {
      for (int q = 0; q <= m - p; ++q) {
        const int r = m - p - q;
        const double c = fact(m) / (fact(p) * fact(q) * fact(r));
        S0[m] += c * mu[p] * mu[q] * mu[r];
        S1[m] += c * mu[p + 1] * mu[q] * mu[r];
        S2[m] += c * mu[p + 2] * mu[q] * mu[r];
        S11[m] += c * mu[p + 1] * mu[q + 1] * mu[r];
      }
    }
  }
  cd G0 = {a * a * (a0 + b0 / 3.0) * g0, 0.0};
  for (int n = 0; n < NT; ++n) {
    G0 = cadd_h(G0, cscl_h(phi[n], S0[n]));
    G0 = cadd_h(G0, cscl_h(psi[n], S1[n]));
  }
  const double j1 = 2.0 * pi / 3.0;
  const double j2 = -2.0 * (std::sqrt(3.0) - pi) / 9.0;
  const double k1 = 2.0 * (2.0 * std::sqrt(3.0) + pi) / 9.0;
  cd A = {2.0 * (-a0 * j1 - 3.0 * b0 * j2), 0.0};
  cd B = {2.0 * b0 * (j1 - 3.0 * j2), 0.0};
  cd C = {6.0 * b0 * (3.0 * j2 - k1), 0.0};
  for (int n = 0; n < NT; ++n) {
    B = cadd_h(B, cscl_h(psi[n], S0[n]));
    if (n >= 1) {
      A = cadd_h(A, cscl_h(phi[n], 2.0 * n * S0[n - 1]));
      A = cadd_h(A, cscl_h(psi[n], 2.0 * n * S1[n - 1]));
      B = cadd_h(B, cscl_h(psi[n], 4.0 * n * S1[n - 1]));
    }
    if (n >= 2) {
      const double f = 4.0 * n * (n - 1);
      A = cadd_h(A, cscl_h(phi[n], f * S1[n - 2]));
      A = cadd_h(A, cscl_h(psi[n], f * S11[n - 2]));
      B = cadd_h(B, cscl_h(psi[n], f * S11[n - 2]));
      C = cadd_h(C, cscl_h(psi[n], f * (S2[n - 2] - 3.0 * S11[n - 2])));
    }
  }
  out[0] = G0;
  out[1] = A;
  out[2] = B;
  out[3] = C;
}

}  // namespace msc

using namespace msc;

struct msc_net {
  Net n;
};

struct msc_plan {
  Plan p;
  Net* net = nullptr;  // learned preconditioner of the running solve (flexible GMRES)
  Krylov k128;
  Krylov k64;
  double2* r128 = nullptr;  // residual buffer
  float2* b64 = nullptr;    // right-hand side narrowed to complex64
  float2* r64 = nullptr;    // complex64 residual
  hipStream_t cap = nullptr;  // private stream for HIP graph capture
  bool use_graphs = false;  // direct launches measured as fast (msc_plan_use_graphs)
  double* scal = nullptr;     // device scalars of msc_net_input
};

struct msc_arnoldi {
  Krylov k;
  int precision = 0;
};

namespace {

hipStream_t S(void* s) { return reinterpret_cast<hipStream_t>(s); }

bool autotune_enabled() {
  const char* v = std::getenv("MSC_AUTOTUNE");
  return v == nullptr || std::strcmp(v, "0") != 0;
}

// Tiles tuned once per (device, FFT grid, precision) in the process: every
// later plan on that grid takes them (the kernels take any slab size on the
// grid), so a sweep's workers and new plans skip the timing runs.
std::mutex g_tile_mu;
std::map<std::array<int, 4>, std::array<int, 5>> g_tiles;

std::array<int, 4> tile_key(const Plan& p, int pr) { return {p.device, p.Pz, p.Pxy, pr}; }

bool lookup_tiles(Plan& p, int pr) {
  std::lock_guard<std::mutex> lock(g_tile_mu);
  const auto it = g_tiles.find(tile_key(p, pr));
  if (it == g_tiles.end()) return false;
  for (int i = 0; i < 5; ++i) p.tile[pr][i] = it->second[i];
  p.tuned[pr] = true;
  return true;
}

void store_tiles(const Plan& p, int pr) {
  std::lock_guard<std::mutex> lock(g_tile_mu);
  std::array<int, 5> t{};
  for (int i = 0; i < 5; ++i) t[i] = p.tile[pr][i];
  g_tiles[tile_key(p, pr)] = t;
}

int alloc_plan_buffers(Plan& p) {
  const size_t b1 = (size_t)p.capNz * p.capM * p.Pxy * 9;
  const size_t b2 = (size_t)p.capNz * p.Pxy * p.Pxy * 9;
  const size_t tab = (size_t)p.NUV * kNE * p.NC;
  MSC_HIP_CHECK(hipMalloc(&p.buf1, b1 * sizeof(double2)));
  MSC_HIP_CHECK(hipMalloc(&p.buf2, b2 * sizeof(double2)));
  MSC_HIP_CHECK(hipMalloc(&p.table128, tab * sizeof(double2)));
  MSC_HIP_CHECK(hipMalloc(&p.table64, tab * sizeof(float2)));
  const size_t capvox = (size_t)p.capNz * p.capM * p.capM;
  MSC_HIP_CHECK(hipMalloc(&p.tp128, capvox * 4 * sizeof(double2)));
  MSC_HIP_CHECK(hipMalloc(&p.tp64, capvox * 4 * sizeof(float2)));
  MSC_HIP_CHECK(hipMalloc(&p.tmax, sizeof(unsigned int)));
  return build_zmap(p);
}

void free_plan_buffers(Plan& p) {
  if (p.buf1) (void)hipFree(p.buf1);
  if (p.buf2) (void)hipFree(p.buf2);
  if (p.table128) (void)hipFree(p.table128);
  if (p.table64) (void)hipFree(p.table64);
  if (p.tp128) (void)hipFree(p.tp128);
  if (p.tp64) (void)hipFree(p.tp64);
  if (p.tmax) (void)hipFree(p.tmax);
  if (p.work && p.owns_work) (void)hipFree(p.work);
  if (p.work2 && p.owns_work) (void)hipFree(p.work2);
  if (p.nn128) (void)hipFree(p.nn128);
  if (p.nn64) (void)hipFree(p.nn64);
  if (p.nn_stage) (void)hipFree(p.nn_stage);
  if (p.zmap) (void)hipFree(p.zmap);
  if (p.sums_part) (void)hipFree(p.sums_part);
  p.sums_part = nullptr;
  p.sums_cap = 0;
  if (p.tidx) (void)hipFree(p.tidx);
  if (p.tgen128) (void)hipFree(p.tgen128);
  if (p.tgen64) (void)hipFree(p.tgen64);
  if (p.flag) (void)hipFree(p.flag);
  p.tidx = nullptr;
  p.tgen128 = nullptr;
  p.tgen64 = nullptr;
  p.flag = nullptr;
  p.zmap = nullptr;
  p.tmax = nullptr;
  p.work = nullptr;
  p.work2 = nullptr;
  p.nn128 = nullptr;
  p.nn64 = nullptr;
  p.nn_stage = nullptr;
  p.buf1 = p.buf2 = nullptr;
  p.table128 = nullptr;
  p.table64 = nullptr;
  p.tp128 = nullptr;
  p.tp64 = nullptr;
}

}  // namespace

// Tile order of the z pass: workgroups go to the XCDs in round-robin order
// (slot b on XCD b % 8), and each XCD has its own L2.  The rows kx and
// Pxy - kx fold onto the same table entries, so they form one group; group g
// (0 <= g <= Pxy/2) is placed on XCD g % 8, its tiles in chunk order with the
// two rows adjacent, so every table entry is fetched by at most the two XCDs
// of its folded (kx, ky) pair.  For Pxy >= 16 the XCDs get equal shares.
int msc::build_zmap(Plan& p) {
  // A column (kx, ky) reads the table slice of its folded pair
  // (max(a, b), min(a, b)), a = min(kx, P - kx), b = min(ky, P - ky): eight
  // columns (+-kx, +-ky and the swap) share a slice.  A tile is Q consecutive
  // ky at one kx (chunk c = ky / Q).  Key the tile by A = a / Q and by the
  // chunk's folded index C = min(c, nck - 1 - c): the swapped columns of tile
  // (A, C) lie in tiles (C, A), so every tile of the unordered pair {A, C}
  // goes to XCD (A + C) mod 8, back to back (+-kx, mirrored chunks and the
  // swapped block together), and a slice is fetched into one L2 once.
  constexpr int kXcd = 8;
  const int P = p.Pxy;
  std::vector<int> all;
  for (int Q = 1; Q <= 16; Q <<= 1) {
    const int nck = P / Q;
    std::vector<std::vector<int>> per(kXcd);
    if (nck >= 1 && P % Q == 0) {
      const int na = (P / 2) / Q + 1;  // groups of folded kx
      const int nc = (nck + 1) / 2;    // folded chunk indices
      std::vector<std::vector<int>> kx_of(na), ch_of(nc);
      for (int kx = 0; kx < P; ++kx) kx_of[std::min(kx, P - kx) / Q].push_back(kx);
      for (int c = 0; c < nck; ++c) ch_of[std::min(c, nck - 1 - c)].push_back(c);
      const int nkey = std::max(na, nc);
      for (int hi = 0; hi < nkey; ++hi) {
        for (int lo = 0; lo <= hi; ++lo) {
          std::vector<int>& L = per[(hi + lo) % kXcd];
          for (int o = 0; o < (lo == hi ? 1 : 2); ++o) {
            const int A = o == 0 ? hi : lo, C = o == 0 ? lo : hi;
            if (A >= na || C >= nc) continue;
            for (int kx : kx_of[A])
              for (int c : ch_of[C]) L.push_back(kx * nck + c);
          }
        }
      }
    }
    const size_t before = all.size();
    size_t longest = 0;
    for (auto& L : per) longest = std::max(longest, L.size());
    for (size_t s = 0; s < longest; ++s)
      for (int x = 0; x < kXcd; ++x)
        if (s < per[x].size()) all.push_back(per[x][s]);
    // every tile exactly once
    const size_t ntile = (size_t)P * std::max(nck, 0);
    std::vector<char> seen(ntile, 0);
    bool ok = all.size() - before == ntile;
    for (size_t i = before; ok && i < all.size(); ++i) {
      ok = all[i] >= 0 && (size_t)all[i] < ntile && !seen[all[i]];
      if (ok) seen[all[i]] = 1;
    }
    if (!ok) {
      set_error("internal: z-pass tile order is not a permutation");
      return kHip;
    }
  }
  MSC_HIP_CHECK(hipMalloc(&p.zmap, all.size() * sizeof(int)));
  MSC_HIP_CHECK(hipMemcpy(p.zmap, all.data(), all.size() * sizeof(int), hipMemcpyHostToDevice));
  return kOk;
}

namespace {

int check_opts(const msc_gmres_opts* o) {
  MSC_REQUIRE(o != nullptr, "opts must not be NULL");
  MSC_REQUIRE(o->restart >= 1, "restart must be >= 1");
  MSC_REQUIRE(o->max_iter >= 0, "max_iter must be >= 0");
  MSC_REQUIRE(o->tol > 0.0 && std::isfinite(o->tol), "tol must be positive and finite");
  MSC_REQUIRE(o->x0 >= 0 && o->x0 <= 2, "x0 must be 0 (zero), 1 (born) or 2 (given)");
  MSC_REQUIRE(o->mode == MSC_SOLVE_FP64 || o->mode == MSC_SOLVE_MIXED, "unknown solver mode");
  MSC_REQUIRE(o->orth >= 0 && o->orth <= 2, "orth must be 0, 1 or 2");
  return kOk;
}

}  // namespace

namespace {

// Launch the work of Arnoldi step j (operator and orthogonalisation) on stream
// s.  The basis is built from w with A (step vector) = v_j - w, so that w does
// not carry the identity part of A = I - B along v_j (whose removal by the
// orthogonalisation cancels and asks for a second Gram-Schmidt pass):
//   plain GMRES:    w = B v_j
//   flexible GMRES: z_j = P(v_j) (v_j has unit norm, so the network's input
//                   scale is sqrt(n)), w = v_j - A z_j (the fused residual).
template <typename R>
int step_launches(msc_plan* pl, Krylov& k, int j, int orth, hipStream_t s) {
  cx<R>* V = reinterpret_cast<cx<R>*>(k.V);
  const cx<R>* vj = V + (size_t)j * k.ldv;
  cx<R>* w = V + (size_t)(j + 1) * k.ldv;
  int rc;
  if (pl->net != nullptr) {
    cx<R>* zj = reinterpret_cast<cx<R>*>(k.Z) + (size_t)j * k.ldv;
    rc = net_apply<R>(*pl->net, pl->p, vj, zj, std::sqrt((double)pl->p.n), s, &k.st->done);
    if (!rc) rc = residual<R>(pl->p, vj, zj, w, s, &k.st->done);
  } else {
    rc = scattered<R>(pl->p, vj, w, s, &k.st->done);
  }
  if (rc) return rc;
  return krylov_step<R>(k, j, orth, s);
}

// The instantiated graph of step j (captured once per Krylov buffer set).
template <typename R>
int step_graph(msc_plan* pl, Krylov& k, int j, int orth, hipGraphExec_t* out) {
  const unsigned long long key = pl->net != nullptr ? pl->net->gen : 0ull;
  if (k.graph_orth != orth || k.graph_net != key) {
    k.drop_graphs();
    k.graph_orth = orth;
    k.graph_net = key;
  }
  if ((int)k.graphs.size() <= j) k.graphs.resize(j + 1, nullptr);
  if (k.graphs[j] == nullptr) {
    if (pl->cap == nullptr) MSC_HIP_CHECK(hipStreamCreateWithFlags(&pl->cap, hipStreamNonBlocking));
    MSC_HIP_CHECK(hipStreamBeginCapture(pl->cap, hipStreamCaptureModeRelaxed));
    const int rc = step_launches<R>(pl, k, j, orth, pl->cap);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(pl->cap, &graph);
    if (rc) {
      if (graph) (void)hipGraphDestroy(graph);
      return rc;
    }
    if (e != hipSuccess) {
      set_error(std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
      return kHip;
    }
    hipGraphExec_t exec = nullptr;
    const hipError_t e2 = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e2 != hipSuccess) {
      set_error(std::string("hipGraphInstantiate: ") + hipGetErrorString(e2));
      return kHip;
    }
    k.graphs[j] = exec;
  }
  *out = k.graphs[j];
  return kOk;
}

// One cycle's Arnoldi steps.  The host keeps at most two steps queued ahead
// of the device and stops launching once the device has raised `done` (read
// through a pinned, device-mapped flag, so no stream synchronisation is
// needed inside the cycle).  Steps already queued when `done` rises are
// no-ops.  Pass-3 event timing (msc_timing_enable) disables the graphs so the
// events bracket the kernel itself.
template <typename R>
int run_cycle(msc_plan* pl, Krylov& k, int steps, int orth, hipStream_t st, int* taken) {
  const bool graphs = pl->use_graphs && !pl->p.timing;
  volatile int* flag = k.done_flag;
  hipEvent_t ev[2];
  MSC_HIP_CHECK(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
  MSC_HIP_CHECK(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
  int rc = kOk;
  for (int j = 0; j < steps && rc == kOk; ++j) {
    if (j >= 2) {
      const hipError_t e = hipEventSynchronize(ev[j & 1]);
      if (e != hipSuccess) {
        set_error(std::string("hipEventSynchronize: ") + hipGetErrorString(e));
        rc = kHip;
        break;
      }
      if (*flag) break;
    }
    if (graphs) {
      hipGraphExec_t exec = nullptr;
      rc = step_graph<R>(pl, k, j, orth, &exec);
      if (rc) break;
      const hipError_t e = hipGraphLaunch(exec, st);
      if (e != hipSuccess) {
        set_error(std::string("hipGraphLaunch: ") + hipGetErrorString(e));
        rc = kHip;
        break;
      }
    } else {
      rc = step_launches<R>(pl, k, j, orth, st);
      if (rc) break;
    }
    const hipError_t e = hipEventRecord(ev[j & 1], st);
    if (e != hipSuccess) {
      set_error(std::string("hipEventRecord: ") + hipGetErrorString(e));
      rc = kHip;
    }
  }
  (void)hipEventDestroy(ev[0]);
  (void)hipEventDestroy(ev[1]);
  if (rc) return rc;
  KState hs{};
  rc = krylov_read_state(k, &hs, st);
  if (rc) return rc;
  *taken = hs.j;
  return kOk;
}

int solve(msc_plan* pl, const double2* b, double2* x, const msc_gmres_opts& o,
          msc_gmres_stats* s, hipStream_t st) {
  Plan& p = pl->p;
  const size_t n = p.n;
  msc_gmres_stats out{};
  const int m = std::max(1, std::min(o.restart, std::max(o.max_iter, 1)));
  int rc = pl->k128.ensure(n, m, sizeof(double2));
  if (rc) return rc;
  Krylov& k = pl->k128;
  Net* net = o.precond != nullptr ? &o.precond->n : nullptr;
  if (net != nullptr) {
    MSC_REQUIRE(net->packed, "preconditioner parameters are not set");
    MSC_REQUIRE(net->device == p.device, "preconditioner and plan are on different devices");
    rc = net->ensure_grid(p.Nz, p.M, p.M);
    if (rc) return rc;
    rc = k.ensure_z();
    if (rc) return rc;
  }
  // the network of the running solve; cleared on every exit
  struct NetScope {
    msc_plan* pl;
    ~NetScope() { pl->net = nullptr; }
  } scope{pl};
  pl->net = net;
  // both forms build the basis from w with A (step vector) = v_j - w
  k.shifted = true;
  bool mixed = o.mode == MSC_SOLVE_MIXED;
  if (mixed) {
    rc = pl->k64.ensure(n, m, sizeof(float2));
    if (rc) return rc;
    pl->k64.shifted = true;
    if (net != nullptr) {
      rc = pl->k64.ensure_z();
      if (rc) return rc;
    }
    if (pl->b64 == nullptr) MSC_HIP_CHECK(hipMalloc(&pl->b64, p.cap_n * sizeof(float2)));
    if (pl->r64 == nullptr) MSC_HIP_CHECK(hipMalloc(&pl->r64, p.cap_n * sizeof(float2)));
  }
  if (pl->r128 == nullptr) MSC_HIP_CHECK(hipMalloc(&pl->r128, p.cap_n * sizeof(double2)));
  double2* r = pl->r128;
  // tile autotuning once per plan and precision (MSC_AUTOTUNE=0 keeps the defaults)
  if (autotune_enabled()) {
    if (!p.tuned[0]) {
      if (!lookup_tiles(p, 0)) {
        rc = autotune<double>(p, b, r, st);
        if (rc) return rc;
        store_tiles(p, 0);
      }
      k.drop_graphs();
    }
    if (mixed && !p.tuned[1]) {
      if (!lookup_tiles(p, 1)) {
        rc = krylov_narrow(b, pl->b64, n, st);
        if (!rc) rc = autotune<float>(p, pl->b64, pl->r64, st);
        if (rc) return rc;
        store_tiles(p, 1);
      }
      pl->k64.drop_graphs();
    }
  }

  // ||b||
  rc = krylov_norm<double>(k, b, &k.scal[0], st);
  if (rc) return rc;
  double bnorm = 0.0;
  MSC_HIP_CHECK(hipMemcpyAsync(&bnorm, &k.scal[0], sizeof(double), hipMemcpyDeviceToHost, st));
  MSC_HIP_CHECK(hipStreamSynchronize(st));
  out.bnorm = bnorm;
  if (o.x0 == MSC_X0_ZERO) {
    MSC_HIP_CHECK(hipMemsetAsync(x, 0, n * sizeof(double2), st));
  } else if (o.x0 == MSC_X0_BORN) {
    MSC_HIP_CHECK(hipMemcpyAsync(x, b, n * sizeof(double2), hipMemcpyDeviceToDevice, st));
  }
  if (bnorm < 1e-30) {  // zero right-hand side (torch_gmres.py:115-116)
    out.converged = 1;
    out.rel_residual = 0.0;
    MSC_HIP_CHECK(hipStreamSynchronize(st));
    *s = out;
    return kOk;
  }
  double prev_rel = INFINITY;
  const double eta_min = o.inner_tol > 0.0 ? o.inner_tol : 1e-4;
  // Mixed form: the first residual needs no FP64 operator.  From x0 = 0 it is
  // b itself; from the Born guess x0 = b it is b - A b = G0 T0 b, taken with
  // the complex64 operator (its rounding only limits what the first complex64
  // cycle could reach anyway).  Every later residual is an FP64 one, and
  // convergence is only ever declared on an FP64 residual.
  bool cheap = mixed && o.x0 != MSC_X0_GIVEN;
  for (;;) {
    Krylov& kk = mixed ? pl->k64 : k;
    const bool from64 = cheap && o.x0 == MSC_X0_BORN;
    if (cheap) {
      if (o.x0 == MSC_X0_ZERO) {
        MSC_HIP_CHECK(hipMemcpyAsync(&kk.st->rnorm, &k.scal[0], sizeof(double),
                                     hipMemcpyDeviceToDevice, st));
      } else {
        rc = krylov_narrow(b, pl->b64, n, st);
        if (!rc) rc = scattered<float>(p, pl->b64, pl->r64, st, nullptr);
        if (!rc) rc = krylov_norm<float>(kk, pl->r64, &kk.st->rnorm, st);
        if (rc) return rc;
        ++out.matvecs_c64;
      }
    } else {
      rc = residual<double>(p, b, x, r, st, nullptr);
      if (rc) return rc;
      ++out.matvecs_c128;
      rc = krylov_norm<double>(kk, r, &kk.st->rnorm, st);
      if (rc) return rc;
    }
    double rnorm = 0.0;
    MSC_HIP_CHECK(hipMemcpyAsync(&rnorm, &kk.st->rnorm, sizeof(double), hipMemcpyDeviceToHost, st));
    MSC_HIP_CHECK(hipStreamSynchronize(st));
    const double rel = rnorm / bnorm;
    if (cheap) {
      cheap = false;
      if (!(rel > o.tol)) continue;  // confirm with an FP64 residual
    } else {
      out.rel_residual = rel;
      if (!(rel > o.tol) || out.iterations >= o.max_iter) break;
      if (out.cycles > 0 && rel > 0.9 * prev_rel) {
        if (pl->net != nullptr) {
          // the learned preconditioner stalled: plain GMRES cycles take over
          pl->net = nullptr;
          k.shifted = true;
          pl->k64.shifted = true;
          out.fell_back |= 2;
        } else if (mixed) {
          // a complex64 cycle failed to make progress: finish with FP64 cycles
          mixed = false;
          out.fell_back |= 1;
        }
        if (!mixed) {
          rc = krylov_norm<double>(k, r, &k.st->rnorm, st);
          if (rc) return rc;
        }
      }
    }
    prev_rel = rel;
    const int steps = std::min(m, o.max_iter - out.iterations);
    int taken = 0;
    if (mixed) {
      // inner tolerance: no tighter than needed to reach tol, no looser than eta
      const double need = 0.3 * o.tol / rel;
      const double eta = std::min(std::max(eta_min, need), 0.5);
      Krylov& k6 = pl->k64;
      *k6.done_flag = 0;
      if (from64) {
        rc = krylov_cycle_init<float, float>(k6, pl->r64, eta, true, st);
      } else {
        rc = krylov_cycle_init<double, float>(k6, o.x0 == MSC_X0_ZERO && out.cycles == 0 ? b : r,
                                              eta, true, st);
      }
      if (rc) return rc;
      rc = run_cycle<float>(pl, k6, steps, o.orth, st, &taken);
      if (rc) return rc;
      rc = krylov_finish<float, double>(k6, x, &k6.st->rnorm, st, pl->net != nullptr);
      if (rc) return rc;
      out.matvecs_c64 += taken;
      if (pl->net != nullptr) out.precond_applies += taken;
    } else {
      *k.done_flag = 0;
      rc = krylov_cycle_init<double, double>(k, r, o.tol * bnorm, false, st);
      if (rc) return rc;
      rc = run_cycle<double>(pl, k, steps, o.orth, st, &taken);
      if (rc) return rc;
      rc = krylov_finish<double, double>(k, x, nullptr, st, pl->net != nullptr);
      if (rc) return rc;
      out.matvecs_c128 += taken;
      if (pl->net != nullptr) out.precond_applies += taken;
    }
    out.iterations += taken;
    ++out.cycles;
    if (taken == 0) {  // breakdown at the first step: nothing more to gain
      rc = residual<double>(p, b, x, r, st, nullptr);
      if (rc) return rc;
      ++out.matvecs_c128;
      rc = krylov_norm<double>(k, r, &k.scal[0], st);
      if (rc) return rc;
      MSC_HIP_CHECK(hipMemcpyAsync(&rnorm, &k.scal[0], sizeof(double), hipMemcpyDeviceToHost, st));
      MSC_HIP_CHECK(hipStreamSynchronize(st));
      out.rel_residual = rnorm / bnorm;
      break;
    }
  }
  out.converged = out.rel_residual <= o.tol ? 1 : 0;
  *s = out;
  return kOk;
}

// Batched GMRES over `ns` plans that share one padded FFT grid, with the
// cycle structure of solve(): every operator pass and every Arnoldi kernel is
// one launch for all systems (system = blockIdx.y, its own chunks, workspace
// and done flag), and a system that has converged, or whose cycle is over,
// turns its part of each launch into a no-op.  The host decides per system
// after each cycle from one gathered copy of the residual norms.  The
// complex64 cycles of MIXED use a unit right-hand side (as solve() does); the
// FP64 cycles do too here, with tolerance tol ||b|| / ||r||, which is the same
// stopping rule.  A system whose complex64 cycle stalls leaves the batch and is
// finished by solve() in FP64 from its current iterate.  No learned
// preconditioner, general T or near-neighbour correction.
int solve_batched(msc_plan* const* pls, const double2* const* b, double2* const* x, int ns,
                  const msc_gmres_opts& o, msc_gmres_stats* stats, hipStream_t st) {
  MSC_REQUIRE(ns >= 1, "batched solve needs at least one system");
  MSC_REQUIRE(o.precond == nullptr, "the batched solve takes no preconditioner");
  const Plan& p0 = pls[0]->p;
  size_t max_n = 0;
  int max_nblk = 0, max_M = 0, max_Nz = 0;
  for (int i = 0; i < ns; ++i) {
    const Plan& p = pls[i]->p;
    MSC_REQUIRE(p.Pz == p0.Pz && p.Pxy == p0.Pxy && p.periodic == p0.periodic &&
                    p.device == p0.device,
                "batched plans must share one padded grid and device");
    MSC_REQUIRE(p.have_table && p.have_tparams,
                "every plan needs build_g0 and set_tmatrix before the batched solve");
    MSC_REQUIRE(!p.general && !p.have_nn,
                "the batched solve takes the cubic T0 without the near-neighbour correction");
    MSC_REQUIRE(b[i] != nullptr && x[i] != nullptr, "b and x must not be NULL");
    max_n = std::max(max_n, p.n);
    max_M = std::max(max_M, p.M);
    max_Nz = std::max(max_Nz, p.Nz);
  }
  const int m = std::max(1, std::min(o.restart, std::max(o.max_iter, 1)));
  const bool mixed = o.mode == MSC_SOLVE_MIXED;
  int rc = kOk;
  for (int i = 0; i < ns; ++i) {
    msc_plan* pl = pls[i];
    const size_t n = pl->p.n;
    rc = mixed ? pl->k128.ensure_state(n, m, sizeof(double2)) : pl->k128.ensure(n, m, sizeof(double2));
    if (rc) return rc;
    pl->k128.shifted = true;
    if (mixed) {
      rc = pl->k64.ensure(n, m, sizeof(float2));
      if (rc) return rc;
      pl->k64.shifted = true;
      if (pl->b64 == nullptr) MSC_HIP_CHECK(hipMalloc(&pl->b64, pl->p.cap_n * sizeof(float2)));
      if (pl->r64 == nullptr) MSC_HIP_CHECK(hipMalloc(&pl->r64, pl->p.cap_n * sizeof(float2)));
    }
    if (pl->r128 == nullptr) MSC_HIP_CHECK(hipMalloc(&pl->r128, pl->p.cap_n * sizeof(double2)));
    max_nblk = std::max(max_nblk, pl->k128.nblk);
  }
  // the batched passes use the first plan's tiles (tuned once per grid)
  Plan& q0 = pls[0]->p;
  if (autotune_enabled()) {
    if (!q0.tuned[0] && !lookup_tiles(q0, 0)) {
      rc = autotune<double>(q0, b[0], pls[0]->r128, st);
      if (rc) return rc;
      store_tiles(q0, 0);
    }
    if (mixed && !q0.tuned[1] && !lookup_tiles(q0, 1)) {
      rc = krylov_narrow(b[0], pls[0]->b64, q0.n, st);
      if (!rc) rc = autotune<float>(q0, pls[0]->b64, pls[0]->r64, st);
      if (rc) return rc;
      store_tiles(q0, 1);
    }
  }

  // host descriptors: Krylov of FP64 norms / cycles, Krylov of complex64
  // cycles, and the operator launches they need
  std::vector<BKry> h128(ns), h64(ns);
  std::vector<BPass<double>> hres(ns), hstep128(ns);
  std::vector<BPass<float>> hcheap(ns), hstep64(ns);
  const size_t bytes_kry = sizeof(BKry) * ns, bytes_p128 = sizeof(BPass<double>) * ns,
               bytes_p64 = sizeof(BPass<float>) * ns;
  const size_t off_k64 = bytes_kry, off_res = 2 * bytes_kry, off_s128 = off_res + bytes_p128,
               off_cheap = off_s128 + bytes_p128, off_s64 = off_cheap + bytes_p64,
               off_act = off_s64 + bytes_p64, off_fin = off_act + sizeof(int) * ns,
               off_gat = (off_fin + sizeof(int) * ns + 15) / 16 * 16,
               total = off_gat + sizeof(double) * 2 * ns;
  unsigned char* dev = nullptr;
  unsigned char* hst = nullptr;  // pinned mirror of dev: one copy per upload
  hipEvent_t ev_up = nullptr;    // the last upload's copy (hst is reused after it)
  MSC_HIP_CHECK(hipMalloc(&dev, total));
  struct Free {
    unsigned char* p;
    unsigned char*& h;
    hipEvent_t& e;
    ~Free() {
      if (p) (void)hipFree(p);
      if (e) {
        (void)hipEventSynchronize(e);
        (void)hipEventDestroy(e);
      }
      if (h) (void)hipHostFree(h);
    }
  } free_dev{dev, hst, ev_up};
  MSC_HIP_CHECK(hipHostMalloc(&hst, total, hipHostMallocDefault));
  std::memset(hst, 0, total);
  MSC_HIP_CHECK(hipEventCreateWithFlags(&ev_up, hipEventDisableTiming));
  BKry* d128 = reinterpret_cast<BKry*>(dev);
  BKry* d64 = reinterpret_cast<BKry*>(dev + off_k64);
  BPass<double>* dres = reinterpret_cast<BPass<double>*>(dev + off_res);
  BPass<double>* ds128 = reinterpret_cast<BPass<double>*>(dev + off_s128);
  BPass<float>* dcheap = reinterpret_cast<BPass<float>*>(dev + off_cheap);
  BPass<float>* ds64 = reinterpret_cast<BPass<float>*>(dev + off_s64);
  int* dact = reinterpret_cast<int*>(dev + off_act);
  int* dfin = reinterpret_cast<int*>(dev + off_fin);
  double* dgat = reinterpret_cast<double*>(dev + off_gat);
  for (int i = 0; i < ns; ++i) {
    msc_plan* pl = pls[i];
    const Plan& p = pl->p;
    h128[i] = bkry_of(pl->k128);
    h128[i].r = b[i];
    h128[i].x = x[i];
    if (mixed) {
      h64[i] = bkry_of(pl->k64);
      h64[i].x = x[i];
      h64[i].src = b[i];
      h64[i].dst = pl->b64;
    }
    auto common = [&](auto& d, const auto* tp, const auto* table) {
      d.tp = tp;
      d.buf1 = reinterpret_cast<decltype(d.buf1)>(p.buf1);
      d.buf2 = reinterpret_cast<decltype(d.buf2)>(p.buf2);
      d.table = table;
      d.M = p.M;
      d.Nz = p.Nz;
      d.in_step = d.out_step = 0;
      d.bvec = nullptr;
      d.done = nullptr;
    };
    common(hres[i], p.tp128, p.table128);
    hres[i].in = x[i];
    hres[i].out = pl->r128;
    hres[i].bvec = b[i];
    hres[i].done = dfin + i;
    common(hstep128[i], p.tp128, p.table128);
    if (!mixed) {
      double2* V = reinterpret_cast<double2*>(pl->k128.V);
      hstep128[i].in = V;
      hstep128[i].out = V + pl->k128.ldv;
      hstep128[i].in_step = hstep128[i].out_step = pl->k128.ldv;
      hstep128[i].done = &pl->k128.st->done;
    }
    common(hcheap[i], p.tp64, p.table64);
    common(hstep64[i], p.tp64, p.table64);
    if (mixed) {
      hcheap[i].in = pl->b64;
      hcheap[i].out = pl->r64;
      float2* V = reinterpret_cast<float2*>(pl->k64.V);
      hstep64[i].in = V;
      hstep64[i].out = V + pl->k64.ldv;
      hstep64[i].in_step = hstep64[i].out_step = pl->k64.ldv;
      hstep64[i].done = &pl->k64.st->done;
    }
  }
  std::vector<int> act(ns, 1), fin(ns, 0), stalled(ns, 0), broke(ns, 0);
  std::vector<double> bnorm(ns), prev(ns, INFINITY);
  // the host descriptors go to the device as one copy from the pinned mirror
  // (per-array copies from pageable memory cost tens of microseconds each)
  auto upload = [&]() -> int {
    MSC_HIP_CHECK(hipEventSynchronize(ev_up));
    std::memcpy(hst, h128.data(), bytes_kry);
    if (mixed) std::memcpy(hst + off_k64, h64.data(), bytes_kry);
    std::memcpy(hst + off_res, hres.data(), bytes_p128);
    std::memcpy(hst + off_s128, hstep128.data(), bytes_p128);
    std::memcpy(hst + off_cheap, hcheap.data(), bytes_p64);
    std::memcpy(hst + off_s64, hstep64.data(), bytes_p64);
    std::memcpy(hst + off_act, act.data(), sizeof(int) * ns);
    std::memcpy(hst + off_fin, fin.data(), sizeof(int) * ns);
    MSC_HIP_CHECK(hipMemcpyAsync(dev, hst, off_gat, hipMemcpyHostToDevice, st));
    MSC_HIP_CHECK(hipEventRecord(ev_up, st));
    return kOk;
  };
  // the gathered norms land in the pinned mirror's last 2 ns doubles
  const double* gat = reinterpret_cast<const double*>(hst + off_gat);
  auto gather = [&](const BKry* d) -> int {
    int r2 = bkrylov_gather(d, ns, dgat, st);
    if (r2) return r2;
    MSC_HIP_CHECK(hipMemcpyAsync(hst + off_gat, dgat, sizeof(double) * 2 * ns,
                                 hipMemcpyDeviceToHost, st));
    MSC_HIP_CHECK(hipStreamSynchronize(st));
    return kOk;
  };
  rc = upload();
  if (rc) return rc;
  // ||b_i||
  rc = bkrylov_norm<double>(d128, ns, max_nblk, st);
  if (!rc) rc = gather(d128);
  if (rc) return rc;
  std::vector<msc_gmres_stats> out(ns);
  for (int i = 0; i < ns; ++i) {
    out[i] = msc_gmres_stats{};
    bnorm[i] = gat[2 * i];
    out[i].bnorm = bnorm[i];
    const size_t bytes = pls[i]->p.n * sizeof(double2);
    if (o.x0 == MSC_X0_ZERO) MSC_HIP_CHECK(hipMemsetAsync(x[i], 0, bytes, st));
    else if (o.x0 == MSC_X0_BORN)
      MSC_HIP_CHECK(hipMemcpyAsync(x[i], b[i], bytes, hipMemcpyDeviceToDevice, st));
    if (bnorm[i] < 1e-30) {  // zero right-hand side
      out[i].converged = 1;
      fin[i] = 1;
      act[i] = 0;
    }
  }
  const double eta_min = o.inner_tol > 0.0 ? o.inner_tol : 1e-4;
  const BKry* dcyc = mixed ? d64 : d128;
  std::vector<BKry>& hcyc = mixed ? h64 : h128;
  bool cheap = mixed && o.x0 != MSC_X0_GIVEN;
  for (;;) {
    const bool from64 = cheap && o.x0 == MSC_X0_BORN;
    // residual norms of the systems still running, into the cycle Krylov
    for (int i = 0; i < ns; ++i) hcyc[i].r = from64 ? (const void*)pls[i]->r64
                                          : (cheap ? (const void*)b[i] : (const void*)pls[i]->r128);
    rc = upload();
    if (rc) return rc;
    if (from64) {
      rc = bkrylov_narrow(d64, ns, max_n, st);
      if (!rc) rc = apply_batch<float>(q0, dcheap, ns, max_M, max_Nz, 2, 0, st);
      if (!rc) rc = bkrylov_norm<float>(d64, ns, max_nblk, st);
      for (int i = 0; i < ns; ++i) out[i].matvecs_c64 += fin[i] ? 0 : 1;
    } else if (cheap) {
      rc = bkrylov_norm<double>(dcyc, ns, max_nblk, st);  // x0 = 0: r = b
    } else {
      rc = apply_batch<double>(q0, dres, ns, max_M, max_Nz, 1, 0, st);
      if (!rc) rc = bkrylov_norm<double>(dcyc, ns, max_nblk, st);
      for (int i = 0; i < ns; ++i) out[i].matvecs_c128 += fin[i] ? 0 : 1;
    }
    if (!rc) rc = gather(dcyc);
    if (rc) return rc;
    int running = 0;
    for (int i = 0; i < ns; ++i) {
      act[i] = 0;
      if (fin[i]) continue;
      const double rnorm = gat[2 * i];
      const double rel = rnorm / bnorm[i];
      if (cheap) {
        if (!(rel > o.tol)) continue;  // confirmed by an FP64 residual next round
      } else {
        out[i].rel_residual = rel;
        if (!(rel > o.tol) || out[i].iterations >= o.max_iter || broke[i]) {
          fin[i] = 1;
          continue;
        }
        if (mixed && out[i].cycles > 0 && rel > 0.9 * prev[i]) {
          stalled[i] = 1;  // finished below in FP64
          fin[i] = 1;
          continue;
        }
      }
      prev[i] = rel;
      act[i] = 1;
      ++running;
      if (mixed) {
        const double need = 0.3 * o.tol / rel;
        hcyc[i].tol_abs = std::min(std::max(eta_min, need), 0.5);
      } else {
        hcyc[i].tol_abs = o.tol * bnorm[i] / rnorm;
      }
    }
    const bool was_cheap = cheap;
    cheap = false;
    if (running == 0) {
      if (was_cheap) continue;
      break;
    }
    int steps = 0;
    for (int i = 0; i < ns; ++i)
      if (act[i]) steps = std::max(steps, std::min(m, o.max_iter - out[i].iterations));
    rc = upload();
    if (rc) return rc;
    for (int i = 0; i < ns; ++i) {
      Krylov& kk = mixed ? pls[i]->k64 : pls[i]->k128;
      *kk.done_flag = act[i] ? 0 : 1;
    }
    rc = bkrylov_deactivate(dcyc, ns, dact, st);
    if (!rc) {
      if (from64) rc = bkrylov_cycle_init<float, float>(dcyc, ns, max_n, true, st);
      else if (mixed) rc = bkrylov_cycle_init<double, float>(dcyc, ns, max_n, true, st);
      else rc = bkrylov_cycle_init<double, double>(dcyc, ns, max_n, true, st);
    }
    if (rc) return rc;
    // the cycle: at most two steps queued ahead of the device
    hipEvent_t ev[2];
    MSC_HIP_CHECK(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
    MSC_HIP_CHECK(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
    for (int j = 0; j < steps && rc == kOk; ++j) {
      if (j >= 2) {
        if (hipEventSynchronize(ev[j & 1]) != hipSuccess) {
          set_error("hipEventSynchronize failed in the batched cycle");
          rc = kHip;
          break;
        }
        bool all = true;
        for (int i = 0; i < ns && all; ++i) {
          Krylov& kk = mixed ? pls[i]->k64 : pls[i]->k128;
          all = *reinterpret_cast<volatile int*>(kk.done_flag) != 0;
        }
        if (all) break;
      }
      if (mixed) {
        rc = apply_batch<float>(q0, ds64, ns, max_M, max_Nz, 2, j, st);
        if (!rc) rc = bkrylov_step<float>(d64, ns, max_n, max_nblk, j, o.orth, st);
      } else {
        rc = apply_batch<double>(q0, ds128, ns, max_M, max_Nz, 2, j, st);
        if (!rc) rc = bkrylov_step<double>(d128, ns, max_n, max_nblk, j, o.orth, st);
      }
      if (!rc && hipEventRecord(ev[j & 1], st) != hipSuccess) {
        set_error("hipEventRecord failed in the batched cycle");
        rc = kHip;
      }
    }
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    if (rc) return rc;
    rc = mixed ? bkrylov_finish<float>(d64, ns, max_n, st) : bkrylov_finish<double>(d128, ns, max_n, st);
    if (!rc) rc = gather(dcyc);
    if (rc) return rc;
    for (int i = 0; i < ns; ++i) {
      if (!act[i]) continue;
      const int taken = (int)gat[2 * i + 1];
      out[i].iterations += taken;
      ++out[i].cycles;
      if (mixed) out[i].matvecs_c64 += taken;
      else out[i].matvecs_c128 += taken;
      if (taken == 0) broke[i] = 1;  // breakdown at the first step: stop after the residual
    }
  }
  // systems whose complex64 cycle stalled: FP64 cycles from the current iterate
  for (int i = 0; i < ns; ++i) {
    if (stalled[i]) {
      msc_gmres_opts o2 = o;
      o2.mode = MSC_SOLVE_FP64;
      o2.x0 = MSC_X0_GIVEN;
      o2.max_iter = std::max(1, o.max_iter - out[i].iterations);
      msc_gmres_stats s2{};
      rc = solve(pls[i], b[i], x[i], o2, &s2, st);
      if (rc) return rc;
      out[i].iterations += s2.iterations;
      out[i].cycles += s2.cycles;
      out[i].matvecs_c128 += s2.matvecs_c128;
      out[i].rel_residual = s2.rel_residual;
      out[i].fell_back |= 1;
    }
    out[i].converged = out[i].rel_residual <= o.tol ? 1 : 0;
    stats[i] = out[i];
  }
  MSC_HIP_CHECK(hipStreamSynchronize(st));
  return kOk;
}

}  // namespace

int msc::plan_reserve(msc_plan* pl, int m, bool mixed) {
  MSC_REQUIRE(pl != nullptr && m >= 1, "plan_reserve: NULL plan or restart < 1");
  const size_t cap = pl->p.cap_n;
  int rc = mixed ? pl->k128.ensure_state(cap, m, sizeof(double2))
                 : pl->k128.ensure(cap, m, sizeof(double2));
  if (rc) return rc;
  if (mixed) {
    rc = pl->k64.ensure(cap, m, sizeof(float2));
    if (rc) return rc;
    if (pl->b64 == nullptr) MSC_HIP_CHECK(hipMalloc(&pl->b64, cap * sizeof(float2)));
    if (pl->r64 == nullptr) MSC_HIP_CHECK(hipMalloc(&pl->r64, cap * sizeof(float2)));
  }
  if (pl->r128 == nullptr) MSC_HIP_CHECK(hipMalloc(&pl->r128, cap * sizeof(double2)));
  return kOk;
}

int msc::plan_share_build_work(msc_plan* dst, msc_plan* src) {
  MSC_REQUIRE(dst != nullptr && src != nullptr, "plan_share_build_work: NULL plan");
  Plan& d = dst->p;
  Plan& s = src->p;
  MSC_REQUIRE(d.Pz == s.Pz && d.Pxy == s.Pxy && d.device == s.device,
              "plans sharing a G0-build workspace need one padded grid and device");
  if (dst == src) return kOk;
  int rc = ensure_build_work(s);
  if (rc) return rc;
  if (d.owns_work) {
    if (d.work) (void)hipFree(d.work);
    if (d.work2) (void)hipFree(d.work2);
  }
  d.work = s.work;
  d.work2 = s.work2;
  d.owns_work = false;
  return kOk;
}

// ── exported C ABI ───────────────────────────────────────────────────────────

extern "C" {

const char* msc_last_error(void) { return get_error(); }
const char* msc_version(void) { return "msc_slab 0.1 (gfx950)"; }

int msc_plan_create(msc_plan** plan, int M, int N_z, double a, int periodic, int pz, int pxy) {
  MSC_REQUIRE(plan != nullptr, "plan must not be NULL");
  *plan = nullptr;
  MSC_REQUIRE(M >= 1, "M must be >= 1");
  MSC_REQUIRE(N_z >= 1, "N_z must be >= 1");
  MSC_REQUIRE(a > 0.0 && std::isfinite(a), "a must be > 0");
  if (pz == 0) pz = std::max(2, next_pow2(2 * N_z - 1));
  if (pxy == 0)
    pxy = periodic && is_pow2(M) && M >= 2 ? M : std::max(2, next_pow2(2 * M - 1));
  MSC_REQUIRE(is_pow2(pz) && pz >= 2 * N_z - 1 && pz >= 2 && pz <= 512,
              "pz must be a power of two in [max(2, 2N_z-1), 512]");
  if (periodic && pxy == M) {
    // circular lateral grid: the folded kernel on the M-point grid
    if (!(is_pow2(M) && M >= 2 && M <= 1024)) {
      set_error("periodic mode on the M-point grid needs M to be a power of two in [2, 1024]");
      return kUnsupported;
    }
  } else {
    // aperiodic, or periodic on a zero-padded grid with the folded kernel (any M)
    MSC_REQUIRE(is_pow2(pxy) && pxy >= 2 * M - 1 && pxy >= 2 && pxy <= 1024,
                "pxy must be a power of two in [max(2, 2M-1), 1024]");
  }
  msc_plan* pl = new (std::nothrow) msc_plan();
  MSC_REQUIRE(pl != nullptr, "out of host memory");
  Plan& p = pl->p;
  p.M = M;
  p.Nz = N_z;
  p.a = a;
  p.periodic = periodic != 0;
  p.Pz = pz;
  p.Pxy = pxy;
  p.nvox = (size_t)N_z * M * M;
  p.n = p.nvox * 9;
  p.capM = M;
  p.capNz = N_z;
  p.cap_n = p.n;
  p.NC = pz / 2 + 1;
  const int h = pxy / 2;
  p.NUV = (h + 1) * (h + 2) / 2;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) {
    delete pl;
    set_error(std::string("hipGetDevice: ") + hipGetErrorString(e));
    return kHip;
  }
  p.device = dev;
  Twiddles* tw = nullptr;
  int rc = twiddles(dev, &tw);
  if (rc == kOk) rc = alloc_plan_buffers(p);
  if (rc) {
    free_plan_buffers(p);
    delete pl;
    return rc;
  }
  p.tw128 = tw->t128;
  p.tw64 = tw->t64;
  *plan = pl;
  return kOk;
}

int msc_plan_reshape(msc_plan* plan, int M, int N_z, double a) {
  MSC_REQUIRE(plan != nullptr, "plan must not be NULL");
  Plan& p = plan->p;
  MSC_REQUIRE(M >= 1 && N_z >= 1, "M and N_z must be >= 1");
  MSC_REQUIRE(a > 0.0 && std::isfinite(a), "a must be > 0");
  MSC_REQUIRE(M <= p.capM && N_z <= p.capNz,
              "the slab exceeds the plan's capacity (create the plan at the largest size)");
  MSC_REQUIRE(2 * N_z - 1 <= p.Pz, "N_z does not fit the plan's vertical FFT length");
  if (p.periodic && p.Pxy == p.M) {
    MSC_REQUIRE(M == p.M, "a periodic plan on the M-point grid cannot change M");
  } else {
    MSC_REQUIRE(2 * M - 1 <= p.Pxy, "M does not fit the plan's lateral FFT length");
  }
  const bool same = M == p.M && N_z == p.Nz && a == p.a;
  if (same) return kOk;
  p.have_table = false;  // the kernel window (and for a new a, the lattice) changes
  p.have_nn = false;
  p.have_tparams = false;
  p.M = M;
  p.Nz = N_z;
  p.a = a;
  p.nvox = (size_t)N_z * M * M;
  p.n = p.nvox * 9;
  plan->k128.drop_graphs();
  plan->k64.drop_graphs();
  return kOk;
}

int msc_plan_use_graphs(msc_plan* plan, int enable) {
  MSC_REQUIRE(plan != nullptr, "plan must not be NULL");
  plan->use_graphs = enable != 0;
  if (!plan->use_graphs) {
    plan->k128.drop_graphs();
    plan->k64.drop_graphs();
  }
  return kOk;
}

int msc_plan_copy_g0(msc_plan* dst, const msc_plan* src, void* stream) {
  MSC_REQUIRE(dst != nullptr && src != nullptr, "NULL argument");
  MSC_REQUIRE(dst != src, "dst and src must be different plans");
  const Plan& s = src->p;
  Plan& d = dst->p;
  if (!s.have_table) {
    set_error("the source plan has no G0 table");
    return kNotReady;
  }
  MSC_REQUIRE(s.M == d.M && s.Nz == d.Nz && s.a == d.a && s.periodic == d.periodic &&
                  s.Pz == d.Pz && s.Pxy == d.Pxy && s.NUV == d.NUV && s.NC == d.NC,
              "plans differ in slab size, lattice, periodicity or FFT grid");
  MSC_REQUIRE(d.device == s.device, "plans are on different devices");
  hipStream_t st = S(stream);
  const size_t tab = (size_t)s.NUV * kNE * s.NC;
  MSC_HIP_CHECK(hipMemcpyAsync(d.table128, s.table128, tab * sizeof(double2),
                               hipMemcpyDeviceToDevice, st));
  MSC_HIP_CHECK(hipMemcpyAsync(d.table64, s.table64, tab * sizeof(float2),
                               hipMemcpyDeviceToDevice, st));
  if (s.have_nn) {
    if (d.nn128 == nullptr) MSC_HIP_CHECK(hipMalloc(&d.nn128, 26 * 81 * sizeof(double2)));
    if (d.nn64 == nullptr) MSC_HIP_CHECK(hipMalloc(&d.nn64, 26 * 81 * sizeof(float2)));
    MSC_HIP_CHECK(hipMemcpyAsync(d.nn128, s.nn128, 26 * 81 * sizeof(double2),
                                 hipMemcpyDeviceToDevice, st));
    MSC_HIP_CHECK(hipMemcpyAsync(d.nn64, s.nn64, 26 * 81 * sizeof(float2),
                                 hipMemcpyDeviceToDevice, st));
  }
  d.have_nn = s.have_nn;
  d.omega = s.omega;
  d.alpha = s.alpha;
  d.beta = s.beta;
  d.rho = s.rho;
  d.have_table = true;
  for (int pr = 0; pr < 2; ++pr) {
    for (int k = 0; k < 5; ++k) d.tile[pr][k] = s.tile[pr][k];
    d.tuned[pr] = s.tuned[pr];
  }
  dst->k128.drop_graphs();
  dst->k64.drop_graphs();
  return kOk;
}

int msc_plan_destroy(msc_plan* plan) {
  if (plan == nullptr) return kOk;
  (void)hipDeviceSynchronize();
  free_plan_buffers(plan->p);
  if (plan->r128) (void)hipFree(plan->r128);
  if (plan->b64) (void)hipFree(plan->b64);
  if (plan->cap) (void)hipStreamDestroy(plan->cap);
  if (plan->r64) (void)hipFree(plan->r64);
  if (plan->scal) (void)hipFree(plan->scal);
  for (hipEvent_t e : plan->p.ev_pool) (void)hipEventDestroy(e);
  delete plan;
  return kOk;
}

int msc_plan_info(const msc_plan* plan, long long* out6) {
  MSC_REQUIRE(plan != nullptr && out6 != nullptr, "plan and out must not be NULL");
  const Plan& p = plan->p;
  out6[0] = p.M;
  out6[1] = p.Nz;
  out6[2] = p.Pz;
  out6[3] = p.Pxy;
  out6[4] = p.NUV;
  out6[5] = p.NC;
  return kOk;
}

int msc_plan_autotune(msc_plan* plan, int precision, void* stream) {
  MSC_REQUIRE(plan != nullptr, "plan must not be NULL");
  MSC_REQUIRE(precision == MSC_C128 || precision == MSC_C64, "precision must be 0 or 1");
  Plan& p = plan->p;
  void* xb = nullptr;
  void* ob = nullptr;
  const size_t eb = precision == MSC_C128 ? sizeof(double2) : sizeof(float2);
  MSC_HIP_CHECK(hipMalloc(&xb, p.n * eb));
  MSC_HIP_CHECK(hipMalloc(&ob, p.n * eb));
  MSC_HIP_CHECK(hipMemsetAsync(xb, 0, p.n * eb, S(stream)));
  const int rc = precision == MSC_C128
                     ? autotune<double>(p, reinterpret_cast<const double2*>(xb),
                                        reinterpret_cast<double2*>(ob), S(stream))
                     : autotune<float>(p, reinterpret_cast<const float2*>(xb),
                                       reinterpret_cast<float2*>(ob), S(stream));
  (void)hipStreamSynchronize(S(stream));
  (void)hipFree(xb);
  (void)hipFree(ob);
  plan->k128.drop_graphs();
  plan->k64.drop_graphs();
  if (!rc) store_tiles(plan->p, precision == MSC_C128 ? 0 : 1);
  return rc;
}

int msc_plan_set_tiles(msc_plan* plan, const int* tiles10) {
  MSC_REQUIRE(plan != nullptr && tiles10 != nullptr, "plan and tiles must not be NULL");
  for (int i = 0; i < 10; ++i) {
    const int v = tiles10[i];
    const int pass = i % 5;
    const bool ok = (pass == 0 || pass == 4) ? (v == 1 || v == 2)
                    : (pass == 2) ? (v == 1 || v == 2 || v == 4 || v == 8 || v == 16 ||
                                     v == (2 | kZDma) || v == (4 | kZDma) || v == (8 | kZDma) ||
                                     v == (2 | kZPrune) || v == (4 | kZPrune) ||
                                     v == (8 | kZPrune))
                                  : (v == 1 || v == 2 || v == 4 || v == 8 || v == 16 || v == 32);
    MSC_REQUIRE(ok, "tile out of range (y passes 1-2, x passes 1/2/4/8/16/32, z pass 1/2/4/8/16, "
                    "2/4/8 + 256 for the persistent z pass or 2/4/8 + 512 for the pruned one)");
  }
  for (int pr = 0; pr < 2; ++pr) {
    for (int i = 0; i < 5; ++i) plan->p.tile[pr][i] = tiles10[pr * 5 + i];
    plan->p.tuned[pr] = true;
  }
  plan->k128.drop_graphs();
  plan->k64.drop_graphs();
  return kOk;
}

int msc_plan_tiles(const msc_plan* plan, int* out10) {
  MSC_REQUIRE(plan != nullptr && out10 != nullptr, "plan and out must not be NULL");
  for (int pr = 0; pr < 2; ++pr)
    for (int i = 0; i < 5; ++i) out10[pr * 5 + i] = plan->p.tile[pr][i];
  return kOk;
}

int msc_cube_coefficients(double omega, double a, double alpha, double beta, double rho,
                          double* out8) {
  MSC_REQUIRE(out8 != nullptr, "out must not be NULL");
  MSC_REQUIRE(omega > 0 && a > 0 && alpha > 0 && beta > 0 && rho > 0,
              "omega, a, alpha, beta and rho must be positive");
  cd c[4];
  cube_coefficients(omega, a, alpha, beta, rho, c);
  for (int i = 0; i < 4; ++i) {
    out8[2 * i] = c[i].re;
    out8[2 * i + 1] = c[i].im;
  }
  return kOk;
}

int msc_build_g0(msc_plan* plan, double omega, double alpha, double beta, double rho,
                 int volume_averaged, int n_orders, void* stream) {
  MSC_REQUIRE(plan != nullptr, "plan must not be NULL");
  MSC_REQUIRE(omega > 0 && alpha > 0 && beta > 0 && rho > 0,
              "omega, alpha, beta and rho must be positive");
  MSC_REQUIRE(!volume_averaged || (n_orders >= 0 && n_orders <= 3), "n_orders must be in 0..3");
  if (volume_averaged && plan->p.general) {
    set_error("the volume-averaged correction cannot be combined with a general per-voxel T");
    return kUnsupported;
  }
  const bool had_nn = plan->p.have_nn;
  int rc = build_table(plan->p, omega, alpha, beta, rho, S(stream));
  if (!rc) rc = set_nn_correction(plan->p, volume_averaged != 0, n_orders, S(stream));
  // captured Arnoldi steps hold the near-neighbour launch only when it was on
  if (plan->p.have_nn != had_nn) {
    plan->k128.drop_graphs();
    plan->k64.drop_graphs();
  }
  return rc;
}

int msc_set_tmatrix(msc_plan* plan, const double* dlambda, const double* dmu, const double* drho,
                    double omega, const double* coef8, void* stream) {
  MSC_REQUIRE(plan != nullptr && dlambda && dmu && drho && coef8, "NULL argument");
  if (plan->p.general) {  // back to the cubic T0: captured steps hold the other pass 1
    plan->k128.drop_graphs();
    plan->k64.drop_graphs();
  }
  return set_tparams_from_contrasts(plan->p, dlambda, dmu, drho, omega, coef8, S(stream));
}

int msc_set_tparams(msc_plan* plan, const void* tparams, void* stream) {
  MSC_REQUIRE(plan != nullptr && tparams != nullptr, "NULL argument");
  if (plan->p.general) {
    plan->k128.drop_graphs();
    plan->k64.drop_graphs();
  }
  return set_tparams_direct(plan->p, reinterpret_cast<const double2*>(tparams), S(stream));
}

int msc_set_tgeneral(msc_plan* plan, const int* tidx, const void* table, int count,
                     void* stream) {
  MSC_REQUIRE(plan != nullptr && tidx != nullptr && table != nullptr, "NULL argument");
  MSC_REQUIRE(count >= 1, "count must be >= 1");
  if (plan->p.have_nn) {
    set_error("a general per-voxel T cannot be combined with the volume-averaged correction");
    return kUnsupported;
  }
  // captured Arnoldi steps hold the pass-1 variant and the table pointers
  plan->k128.drop_graphs();
  plan->k64.drop_graphs();
  return set_tgeneral(plan->p, tidx, reinterpret_cast<const double2*>(table), count, S(stream));
}

int msc_get_tparams(const msc_plan* plan, void* out, void* stream) {
  MSC_REQUIRE(plan != nullptr && out != nullptr, "NULL argument");
  if (!plan->p.have_tparams || plan->p.general) {
    set_error("T0 parameters not set");
    return kNotReady;
  }
  MSC_HIP_CHECK(hipMemcpyAsync(out, plan->p.tp128, plan->p.nvox * 4 * sizeof(double2),
                               hipMemcpyDeviceToDevice, S(stream)));
  return kOk;
}

int msc_incident_field(msc_plan* plan, double k_mag, const double* khat3, const double* inc18,
                       void* psi0, void* stream) {
  MSC_REQUIRE(plan != nullptr && khat3 && inc18 && psi0, "NULL argument");
  return incident_field(plan->p, k_mag, khat3, inc18, reinterpret_cast<double2*>(psi0), S(stream));
}

int msc_matvec(msc_plan* plan, const void* x, void* y, int precision, void* stream) {
  MSC_REQUIRE(plan != nullptr && x != nullptr && y != nullptr, "NULL argument");
  MSC_REQUIRE(x != y, "x and y must not alias");
  if (precision == MSC_C128)
    return matvec<double>(plan->p, reinterpret_cast<const double2*>(x),
                          reinterpret_cast<double2*>(y), S(stream), nullptr);
  if (precision == MSC_C64)
    return matvec<float>(plan->p, reinterpret_cast<const float2*>(x),
                         reinterpret_cast<float2*>(y), S(stream), nullptr);
  MSC_REQUIRE(false, "precision must be 0 (complex128) or 1 (complex64)");
}

int msc_residual(msc_plan* plan, const void* b, const void* x, void* r, int precision,
                 void* stream) {
  MSC_REQUIRE(plan != nullptr && b && x && r, "NULL argument");
  MSC_REQUIRE(x != r, "x and r must not alias");
  if (precision == MSC_C128)
    return residual<double>(plan->p, reinterpret_cast<const double2*>(b),
                            reinterpret_cast<const double2*>(x), reinterpret_cast<double2*>(r),
                            S(stream), nullptr);
  if (precision == MSC_C64)
    return residual<float>(plan->p, reinterpret_cast<const float2*>(b),
                           reinterpret_cast<const float2*>(x), reinterpret_cast<float2*>(r),
                           S(stream), nullptr);
  MSC_REQUIRE(false, "precision must be 0 (complex128) or 1 (complex64)");
}

int msc_gmres(msc_plan* plan, const void* b, void* x, const msc_gmres_opts* opts,
              msc_gmres_stats* stats, void* stream) {
  MSC_REQUIRE(plan != nullptr && b != nullptr && x != nullptr && stats != nullptr,
              "NULL argument");
  MSC_REQUIRE(b != x, "b and x must not alias");
  int rc = check_opts(opts);
  if (rc) return rc;
  if (!plan->p.have_table || !plan->p.have_tparams) {
    set_error("plan needs msc_build_g0 and msc_set_tmatrix before msc_gmres");
    return kNotReady;
  }
  return solve(plan, reinterpret_cast<const double2*>(b), reinterpret_cast<double2*>(x), *opts,
               stats, S(stream));
}

int msc_gmres_batched(msc_plan* const* plans, const void* const* b, void* const* x, int count,
                      const msc_gmres_opts* opts, msc_gmres_stats* stats, void* stream) {
  MSC_REQUIRE(plans != nullptr && b != nullptr && x != nullptr && stats != nullptr && count >= 1,
              "NULL argument or empty batch");
  int rc = check_opts(opts);
  if (rc) return rc;
  MSC_REQUIRE(opts->precond == nullptr, "msc_gmres_batched takes no preconditioner");
  for (int i = 0; i < count; ++i) {
    MSC_REQUIRE(plans[i] != nullptr, "NULL plan in the batch");
    MSC_REQUIRE(b[i] != x[i], "b and x must not alias");
    for (int k = 0; k < i; ++k)
      MSC_REQUIRE(plans[k] != plans[i], "a plan appears twice in the batch");
  }
  return solve_batched(plans, reinterpret_cast<const double2* const*>(b),
                       reinterpret_cast<double2* const*>(x), count, *opts, stats, S(stream));
}

int msc_source_sums(msc_plan* plan, const void* psi, double kP, double kS, const double* rhat3,
                    void* layer_sum, void* phased, void* stream) {
  MSC_REQUIRE(plan != nullptr && psi && rhat3 && layer_sum && phased, "NULL argument");
  if (!plan->p.have_tparams) {
    set_error("T0 parameters not set");
    return kNotReady;
  }
  return source_sums(plan->p, reinterpret_cast<const double2*>(psi), kP, kS, rhat3,
                     reinterpret_cast<double2*>(layer_sum), reinterpret_cast<double2*>(phased),
                     S(stream));
}

int msc_fft_axis(const void* in, void* out, int outer, int n, long long inner, int inverse,
                 int precision, void* stream) {
  MSC_REQUIRE(in != nullptr && out != nullptr, "NULL argument");
  MSC_REQUIRE(outer >= 1 && inner >= 1 && n >= 1, "sizes must be positive");
  int dev = 0;
  MSC_HIP_CHECK(hipGetDevice(&dev));
  Twiddles* tw = nullptr;
  int rc = twiddles(dev, &tw);
  if (rc) return rc;
  if (precision == MSC_C128)
    return fft_axis<double>(reinterpret_cast<const double2*>(in), reinterpret_cast<double2*>(out),
                            outer, n, (size_t)inner, inverse != 0, tw->t128, S(stream));
  if (precision == MSC_C64)
    return fft_axis<float>(reinterpret_cast<const float2*>(in), reinterpret_cast<float2*>(out),
                           outer, n, (size_t)inner, inverse != 0, tw->t64, S(stream));
  MSC_REQUIRE(false, "precision must be 0 (complex128) or 1 (complex64)");
}

int msc_arnoldi_create(msc_arnoldi** ctx, long long n, int m, int precision) {
  MSC_REQUIRE(ctx != nullptr, "ctx must not be NULL");
  *ctx = nullptr;
  MSC_REQUIRE(n >= 1 && m >= 1, "n and m must be >= 1");
  MSC_REQUIRE(precision == MSC_C128 || precision == MSC_C64, "precision must be 0 or 1");
  msc_arnoldi* c = new (std::nothrow) msc_arnoldi();
  MSC_REQUIRE(c != nullptr, "out of host memory");
  c->precision = precision;
  const int rc = c->k.ensure_state((size_t)n, m, precision == MSC_C128 ? 16 : 8);
  if (rc) {
    delete c;
    return rc;
  }
  *ctx = c;
  return kOk;
}

int msc_arnoldi_destroy(msc_arnoldi* ctx) {
  if (ctx == nullptr) return kOk;
  (void)hipDeviceSynchronize();
  delete ctx;
  return kOk;
}

int msc_arnoldi_start(msc_arnoldi* ctx, void* V, long long ldv, const void* r, double tol_abs,
                      void* stream) {
  MSC_REQUIRE(ctx != nullptr && V != nullptr && r != nullptr, "NULL argument");
  MSC_REQUIRE(ldv >= (long long)ctx->k.n, "ldv must be >= n");
  Krylov& k = ctx->k;
  k.V = V;
  k.ldv = (size_t)ldv;
  int rc;
  if (ctx->precision == MSC_C128) {
    rc = krylov_norm<double>(k, reinterpret_cast<const double2*>(r), &k.st->rnorm, S(stream));
    if (!rc) rc = krylov_cycle_init<double, double>(k, reinterpret_cast<const double2*>(r), tol_abs,
                                                    false, S(stream));
  } else {
    rc = krylov_norm<float>(k, reinterpret_cast<const float2*>(r), &k.st->rnorm, S(stream));
    if (!rc) rc = krylov_cycle_init<float, float>(k, reinterpret_cast<const float2*>(r), tol_abs,
                                                  false, S(stream));
  }
  return rc;
}

int msc_arnoldi_step(msc_arnoldi* ctx, void* V, long long ldv, int j, int orth, void* stream) {
  MSC_REQUIRE(ctx != nullptr && V != nullptr, "NULL argument");
  MSC_REQUIRE(j >= 0 && j < ctx->k.m, "step index out of range");
  MSC_REQUIRE(orth >= 0 && orth <= 2, "orth must be 0, 1 or 2");
  Krylov& k = ctx->k;
  k.V = V;
  k.ldv = (size_t)ldv;
  return ctx->precision == MSC_C128 ? krylov_step<double>(k, j, orth, S(stream))
                                    : krylov_step<float>(k, j, orth, S(stream));
}

int msc_arnoldi_state(msc_arnoldi* ctx, double* out4, void* stream) {
  MSC_REQUIRE(ctx != nullptr && out4 != nullptr, "NULL argument");
  KState s{};
  const int rc = krylov_read_state(ctx->k, &s, S(stream));
  if (rc) return rc;
  out4[0] = s.j;
  out4[1] = s.done;
  out4[2] = s.resid_est;
  out4[3] = s.rnorm;
  return kOk;
}

int msc_arnoldi_solution(msc_arnoldi* ctx, const void* V, long long ldv, void* x, void* stream) {
  MSC_REQUIRE(ctx != nullptr && V != nullptr && x != nullptr, "NULL argument");
  Krylov& k = ctx->k;
  k.V = const_cast<void*>(V);
  k.ldv = (size_t)ldv;
  if (ctx->precision == MSC_C128)
    return krylov_finish<double, double>(k, reinterpret_cast<double2*>(x), nullptr, S(stream));
  return krylov_finish<float, float>(k, reinterpret_cast<float2*>(x), nullptr, S(stream));
}

int msc_timing_enable(msc_plan* plan, int enable) {
  MSC_REQUIRE(plan != nullptr, "plan must not be NULL");
  plan->p.timing = enable != 0;
  plan->p.ev_used = 0;
  return kOk;
}

int msc_timing_read(msc_plan* plan, double* out4) {
  MSC_REQUIRE(plan != nullptr && out4 != nullptr, "NULL argument");
  double ms64 = 0.0, ms128 = 0.0;
  size_t n64 = 0, n128 = 0;
  const Plan& p = plan->p;
  for (size_t i = 0; i + 1 < p.ev_used; i += 2) {
    MSC_HIP_CHECK(hipEventSynchronize(p.ev_pool[i + 1]));
    float ms = 0.0f;
    MSC_HIP_CHECK(hipEventElapsedTime(&ms, p.ev_pool[i], p.ev_pool[i + 1]));
    if (p.ev_prec[i / 2] == 8) {
      ms64 += ms;
      ++n64;
    } else {
      ms128 += ms;
      ++n128;
    }
  }
  out4[0] = (double)n64;
  out4[1] = ms64;
  out4[2] = (double)n128;
  out4[3] = ms128;
  return kOk;
}

int msc_timing_durations(msc_plan* plan, int elem_bytes, double* out, int cap, int* n) {
  MSC_REQUIRE(plan != nullptr && n != nullptr && (out != nullptr || cap == 0) && cap >= 0,
              "NULL argument");
  MSC_REQUIRE(elem_bytes == 8 || elem_bytes == 16, "elem_bytes must be 8 or 16");
  const Plan& p = plan->p;
  int k = 0;
  for (size_t i = 0; i + 1 < p.ev_used; i += 2) {
    if (p.ev_prec[i / 2] != elem_bytes) continue;
    MSC_HIP_CHECK(hipEventSynchronize(p.ev_pool[i + 1]));
    float ms = 0.0f;
    MSC_HIP_CHECK(hipEventElapsedTime(&ms, p.ev_pool[i], p.ev_pool[i + 1]));
    if (k < cap) out[k] = ms;
    ++k;
  }
  *n = k;
  return kOk;
}

// ── learned preconditioner (msc_precond.h) ──────────────────────────────────

int msc_net_create(msc_net** net, int hidden) {
  MSC_REQUIRE(net != nullptr, "net must not be NULL");
  *net = nullptr;
  msc_net* h = new (std::nothrow) msc_net();
  MSC_REQUIRE(h != nullptr, "out of host memory");
  int rc = net_init(h->n, hidden);
  if (!rc) rc = net_pack(h->n, nullptr);
  if (!rc) {
    const hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) {
      set_error(std::string("hipDeviceSynchronize: ") + hipGetErrorString(e));
      rc = kHip;
    }
  }
  if (rc) {
    delete h;
    return rc;
  }
  *net = h;
  return kOk;
}

int msc_net_create_unet(msc_net** net, const int* c4) {
  MSC_REQUIRE(net != nullptr && c4 != nullptr, "NULL argument");
  *net = nullptr;
  msc_net* h = new (std::nothrow) msc_net();
  MSC_REQUIRE(h != nullptr, "out of host memory");
  int rc = net_init_unet(h->n, c4);
  if (!rc) rc = net_pack(h->n, nullptr);
  if (!rc) {
    const hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) {
      set_error(std::string("hipDeviceSynchronize: ") + hipGetErrorString(e));
      rc = kHip;
    }
  }
  if (rc) {
    delete h;
    return rc;
  }
  *net = h;
  return kOk;
}

int msc_net_destroy(msc_net* net) {
  if (net == nullptr) return kOk;
  (void)hipDeviceSynchronize();
  delete net;
  return kOk;
}

int msc_net_num_params(const msc_net* net, long long* n) {
  MSC_REQUIRE(net != nullptr && n != nullptr, "NULL argument");
  *n = (long long)net->n.nparams;
  return kOk;
}

int msc_net_set_params(msc_net* net, const float* params, void* stream) {
  MSC_REQUIRE(net != nullptr && params != nullptr, "NULL argument");
  MSC_HIP_CHECK(hipMemcpyAsync(net->n.params, params, net->n.nparams * sizeof(float),
                               hipMemcpyDeviceToDevice, S(stream)));
  return net_pack(net->n, S(stream));
}

int msc_net_get_params(const msc_net* net, float* params, void* stream) {
  MSC_REQUIRE(net != nullptr && params != nullptr, "NULL argument");
  MSC_HIP_CHECK(hipMemcpyAsync(params, net->n.params, net->n.nparams * sizeof(float),
                               hipMemcpyDeviceToDevice, S(stream)));
  return kOk;
}

int msc_net_input(msc_plan* plan, const void* r, int precision, double scale, void* out_bf16,
                  void* stream) {
  MSC_REQUIRE(plan != nullptr && r != nullptr && out_bf16 != nullptr, "NULL argument");
  if (!plan->p.have_tparams) {
    set_error("plan needs msc_set_tmatrix before msc_net_input");
    return kNotReady;
  }
  if (plan->scal == nullptr) MSC_HIP_CHECK(hipMalloc(&plan->scal, kScalDoubles * sizeof(double)));
  uint16_t* out = reinterpret_cast<uint16_t*>(out_bf16);
  if (precision == MSC_C128)
    return net_input<double>(plan->p, reinterpret_cast<const double2*>(r), scale, plan->scal,
                             out, S(stream));
  if (precision == MSC_C64)
    return net_input<float>(plan->p, reinterpret_cast<const float2*>(r), scale, plan->scal, out,
                            S(stream));
  MSC_REQUIRE(false, "precision must be 0 (complex128) or 1 (complex64)");
}

int msc_net_forward(msc_net* net, int D, int H, int W, const void* in_bf16, float* out,
                    void* stream) {
  MSC_REQUIRE(net != nullptr && in_bf16 != nullptr && out != nullptr, "NULL argument");
  MSC_REQUIRE(D >= 1 && H >= 1 && W >= 1, "grid dimensions must be >= 1");
  return net_forward_out(net->n, D, H, W, reinterpret_cast<const uint16_t*>(in_bf16), out,
                         S(stream));
}

int msc_net_gradients(msc_net* net, int D, int H, int W, int batch, const void* in_bf16,
                      const float* target, float* grads, double* loss, void* stream) {
  MSC_REQUIRE(net != nullptr && in_bf16 != nullptr && target != nullptr && grads != nullptr,
              "NULL argument");
  MSC_REQUIRE(D >= 1 && H >= 1 && W >= 1, "grid dimensions must be >= 1");
  int rc = net_gradients(net->n, D, H, W, batch, reinterpret_cast<const uint16_t*>(in_bf16),
                         target, loss, S(stream));
  if (rc) return rc;
  MSC_HIP_CHECK(hipMemcpyAsync(grads, net->n.grads, net->n.nparams * sizeof(float),
                               hipMemcpyDeviceToDevice, S(stream)));
  return kOk;
}

int msc_net_train_step(msc_net* net, int D, int H, int W, int batch, const void* in_bf16,
                       const float* target, double lr, double* loss, void* stream) {
  MSC_REQUIRE(net != nullptr && in_bf16 != nullptr && target != nullptr, "NULL argument");
  MSC_REQUIRE(D >= 1 && H >= 1 && W >= 1, "grid dimensions must be >= 1");
  MSC_REQUIRE(lr > 0.0 && std::isfinite(lr), "lr must be positive and finite");
  int rc = net_gradients(net->n, D, H, W, batch, reinterpret_cast<const uint16_t*>(in_bf16),
                         target, loss, S(stream));
  if (rc) return rc;
  return net_adam(net->n, lr, S(stream));
}

int msc_gmres_basis(msc_plan* plan, int precision, int j, void* out, void* stream) {
  MSC_REQUIRE(plan != nullptr && out != nullptr, "NULL argument");
  MSC_REQUIRE(precision == MSC_C128 || precision == MSC_C64,
              "precision must be 0 (complex128) or 1 (complex64)");
  const Krylov& k = precision == MSC_C128 ? plan->k128 : plan->k64;
  if (k.V == nullptr || !k.owns_V) {
    set_error("no GMRES basis of this precision: run msc_gmres first");
    return kNotReady;
  }
  MSC_REQUIRE(j >= 0 && j <= k.m, "j must be in [0, restart]");
  const size_t eb = (size_t)k.elem_bytes;
  MSC_HIP_CHECK(hipMemcpyAsync(out, reinterpret_cast<const char*>(k.V) + (size_t)j * k.ldv * eb,
                               k.n * eb, hipMemcpyDeviceToDevice, S(stream)));
  return kOk;
}

int msc_net_apply(msc_net* net, msc_plan* plan, const void* r, void* z, int precision,
                  void* stream) {
  MSC_REQUIRE(net != nullptr && plan != nullptr && r != nullptr && z != nullptr,
              "NULL argument");
  MSC_REQUIRE(r != z, "r and z must not alias");
  if (!plan->p.have_tparams) {
    set_error("plan needs msc_set_tmatrix before msc_net_apply");
    return kNotReady;
  }
  MSC_REQUIRE(net->n.device == plan->p.device, "network and plan are on different devices");
  if (precision == MSC_C128)
    return net_apply<double>(net->n, plan->p, reinterpret_cast<const double2*>(r),
                             reinterpret_cast<double2*>(z), -1.0, S(stream));
  if (precision == MSC_C64)
    return net_apply<float>(net->n, plan->p, reinterpret_cast<const float2*>(r),
                            reinterpret_cast<float2*>(z), -1.0, S(stream));
  MSC_REQUIRE(false, "precision must be 0 (complex128) or 1 (complex64)");
}

}  // extern "C"
