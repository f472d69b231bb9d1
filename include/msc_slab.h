/*
 * msc_slab.h -- C ABI of the MI355X (gfx950) slab Foldy-Lax library
 * (libmsc_slab.so).
 *
 * The library solves (I - G0 T0) psi = psi0 for an M x M x N_z slab of cubic
 * voxels with 9 complex unknowns each (3 displacement + 6 Voigt strain,
 * order (z, x, y) / (zz, xx, yy, xy, zy, zx)).  All arrays are caller-owned
 * device pointers (interleaved complex, C-contiguous (N_z, M, M, 9) unless
 * stated) and all work is ordered on the caller's HIP stream (`stream` is a
 * hipStream_t, NULL = the default stream).  Every entry point returns an
 * msc_status; msc_last_error() describes the last failure of the calling
 * thread.  One host thread per plan.
 *
 * The reference exposes a Python API only; each function below cites the
 * reference interface (path:line in tmnestor/MultipleScatteringCalculations)
 * it replaces.
 */
#ifndef MSC_SLAB_H
#define MSC_SLAB_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum msc_status {
  MSC_OK = 0,
  MSC_ERR_INVALID = 1,     /* bad argument          -> ValueError   */
  MSC_ERR_UNSUPPORTED = 2, /* valid but not handled -> ValueError   */
  MSC_ERR_HIP = 3,         /* HIP runtime failure   -> RuntimeError */
  MSC_ERR_NOT_READY = 4    /* plan not initialised  -> RuntimeError */
} msc_status;

typedef enum msc_precision {
  MSC_C128 = 0, /* complex128 throughout                                   */
  MSC_C64 = 1   /* complex64 operator (FFTs, block multiply, basis)        */
} msc_precision;

typedef enum msc_solver_mode {
  MSC_SOLVE_FP64 = 0,  /* plain GMRES(m) in complex128                      */
  MSC_SOLVE_MIXED = 1  /* complex64 GMRES cycles inside FP64 iterative      */
                       /* refinement, FP64 true residual judges each cycle  */
} msc_solver_mode;

typedef enum msc_x0 {
  MSC_X0_ZERO = 0, /* x0 = 0                                                */
  MSC_X0_BORN = 1, /* x0 = b (the reference default, "born")                */
  MSC_X0_GIVEN = 2 /* x holds the initial guess on entry                    */
} msc_x0;

typedef struct msc_plan msc_plan;
typedef struct msc_arnoldi msc_arnoldi;
typedef struct msc_net msc_net; /* learned preconditioner (msc_precond.h)     */

typedef struct msc_gmres_opts {
  int restart;       /* Krylov dimension per cycle (>= 1)                    */
  int max_iter;      /* cap on Arnoldi steps over all cycles                 */
  double tol;        /* relative FP64 residual ||b - A x|| / ||b||          */
  int x0;            /* msc_x0                                               */
  int mode;          /* msc_solver_mode                                      */
  double inner_tol;  /* MIXED: relative reduction asked of each c64 cycle    */
  int orth;          /* 0 adaptive CGS (reorthogonalise when needed),       */
                     /* 1 CGS once, 2 CGS twice                              */
  msc_net* precond;  /* NULL: GMRES; else flexible GMRES with this learned   */
                     /* right preconditioner (z_j = P(v_j), x = x0 + Z y)    */
} msc_gmres_opts;

typedef struct msc_gmres_stats {
  int iterations;      /* Arnoldi steps (all cycles, all precisions)         */
  int cycles;          /* restart / refinement cycles                        */
  int matvecs_c128;    /* complex128 operator applications                   */
  int matvecs_c64;     /* complex64 operator applications                    */
  int converged;       /* 1 if rel_residual <= tol                           */
  int fell_back;       /* bit 0: MIXED switched to FP64 cycles; bit 1: the   */
                       /* learned preconditioner stalled and plain GMRES    */
                       /* cycles finished the solve                          */
  int precond_applies; /* learned-preconditioner applications                */
  double rel_residual; /* final true FP64 ||b - A x|| / ||b||               */
  double bnorm;        /* ||b||                                              */
} msc_gmres_stats;

/* One independent solve of a sweep (msc_sweep). */
typedef struct msc_instance {
  int M, N_z;               /* slab of M x M x N_z voxels                      */
  int periodic;             /* 1: laterally periodic (any M)                    */
  int volume_averaged;      /* 1: inter-voxel blocks at the 26 neighbours      */
  int n_orders;             /* their frequency orders, 0..3 (volume_averaged)  */
  double a;                 /* cube half-width                                 */
  double omega;             /* angular frequency                               */
  double alpha, beta, rho;  /* background medium                               */
  double k_mag;             /* incident wavenumber (omega / alpha for P)       */
  double khat[3];           /* incident direction                              */
  double inc18[18];         /* incident (polarisation, Voigt strain), re/im    */
  const double* dlambda;    /* device (N_z, M, M) float64 contrasts            */
  const double* dmu;
  const double* drho;
  void* psi;                /* out: device (N_z, M, M, 9) complex128           */
  void* psi0;               /* out (may be NULL): the incident field, same     */
  void* layer_sum;          /* out (may be NULL): device (N_z, 9) complex128   */
  void* phased;             /* out (with layer_sum): device (N_z, 2, 9): sums  */
                            /* of tau = T0 psi as msc_source_sums with kP, kS  */
                            /* = omega/alpha, omega/beta and r_hat = -khat     */
  msc_gmres_stats stats;    /* out                                             */
  int status;               /* out: msc_status of this instance                */
  char error[160];          /* out: message when status != MSC_OK              */
} msc_instance;

/*
 * Resonance-regime T-matrices of `count` cubes of half-width a, each split
 * into n_sub^3 sub-cells (n_sub 1..8): the internal 9N x 9N Foldy-Lax system
 * (N = n_sub^3) with 9 incident patterns, solved by Gaussian elimination with
 * partial pivoting on the device, and the composite 9x9 T-matrix
 * T_comp = sum_m T_loc psi_m.  Replaces compute_resonance_tmatrix /
 * _solve_coupled (cubic_scattering/resonance_tmatrix.py:650-850) and
 * _build_incident_field_coupled (526-593) for a batch of contrasts.
 *   khat3, wave_s: incidence direction (normalised here) and 0 = P, 1 = S
 *   dlambda, dmu, drho: device float64 arrays of `count` contrasts
 *   neumann_order > 0: truncated Neumann series instead of the direct solve,
 *     stopping per cube when |term| / |sum| < neumann_tol
 *   tcomp: device out, count x 9 x 9 complex128 (row-major)
 *   psi_exc: device out or NULL, count x 9N x 9 complex128 (row-major)
 *   pivot_ratio: host out or NULL, max |u_ii| / min |u_ii| (NaN for Neumann)
 *   neumann_iters: host out or NULL, converged order, -1 if none / direct
 * Synchronises `stream` before returning.  A singular system is
 * MSC_ERR_INVALID.
 */
int msc_resonance_tmatrices(double omega, double a, double alpha, double beta, double rho,
                            int n_sub, const double* khat3, int wave_s, int count,
                            const double* dlambda, const double* dmu, const double* drho,
                            int neumann_order, double neumann_tol, void* tcomp, void* psi_exc,
                            double* pivot_ratio, int* neumann_iters, void* stream);

/*
 * The internal Foldy-Lax matrix A = I - P (I x T_loc) (9N x 9N, N = n_sub^3,
 * complex128, column-major: A[c * 9N + r]) of one cube with contrasts
 * (dlambda, dmu, drho) (host values), written to the device buffer A_out:
 * the matrix whose 2-norm condition number _solve_coupled
 * (cubic_scattering/resonance_tmatrix.py:770-795) reports and warns on.
 * Synchronises `stream` before returning.
 */
int msc_resonance_system(double omega, double a, double alpha, double beta, double rho,
                         int n_sub, double dlambda, double dmu, double drho, void* A_out,
                         void* stream);

/*
 * msc_resonance_tmatrices without any device allocation.  No counterpart in
 * the reference: this entry point is for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Same arguments, results and errors as
 * msc_resonance_tmatrices; the workspace is a 64 MiB device array inside the
 * library that every GPU it runs on holds until the process ends, so n_sub
 * is limited to 6 (larger n_sub is MSC_ERR_UNSUPPORTED; any count is taken,
 * in chunks of at most 64 cubes).  Calls no hipMalloc/hipFree.  All calls on
 * one GPU share that array: calls that overlap on one GPU (other threads,
 * other streams) race and must be serialised by the caller.  Synchronises
 * `stream` before returning.
 */
int msc_resonance_tmatrices_noalloc(double omega, double a, double alpha, double beta, double rho,
                                    int n_sub, const double* khat3, int wave_s, int count,
                                    const double* dlambda, const double* dmu, const double* drho,
                                    int neumann_order, double neumann_tol, void* tcomp,
                                    void* psi_exc, double* pivot_ratio, int* neumann_iters,
                                    void* stream);

/* Last error message of the calling thread ("" if none). */
const char* msc_last_error(void);

/* Library version string. */
const char* msc_version(void);

/*
 * Plan for an M x M x N_z slab of cubes of half-width a (side d = 2a).
 * pz, pxy: padded FFT lengths (powers of two, pz <= 512 and pxy <= 1024, so
 * N_z <= 256 and M <= 512); 0 picks the smallest power of two >= 2N_z-1 (z)
 * and >= 2M-1 (x, y), or pxy = M when periodic and M is a power of two.  Periodic mode (circular lateral convolution, any M) runs on
 * the M-point lateral grid when pxy == M (M a power of two), otherwise on the
 * padded grid (pxy >= 2M-1) with the periodically folded kernel.
 * Replaces the geometry handling of _build_slab_kernels_gpu /
 * _slab_matvec_gpu (cubic_scattering/slab_scattering_gpu.py:34-155) and
 * SlabGeometry (cubic_scattering/slab_scattering.py:40-99).
 */
int msc_plan_create(msc_plan** plan, int M, int N_z, double a, int periodic, int pz, int pxy);
int msc_plan_destroy(msc_plan* plan);

/*
 * Re-target a plan to a smaller slab (M <= the M it was created with, N_z
 * likewise, same FFT grid; a may change) without reallocating.  The G0 table
 * and T0 must be set again.  No work on the plan may be in flight.  No
 * reference counterpart: it lets msc_sweep reuse one plan per FFT grid across
 * the instances of a sweep.
 */
int msc_plan_reshape(msc_plan* plan, int M, int N_z, double a);

/*
 * Copy the G0 table (with its volume-averaged near-neighbour correction) and
 * the autotuned tiles of `src` into `dst`, a plan of the same slab size,
 * lattice, periodicity and FFT grid on the same device, in order on `stream`:
 * the same state as msc_build_g0 on `dst` at src's frequency and medium, at
 * the cost of a device copy.  `dst`'s T0 is unchanged.  No reference
 * counterpart: it gives the concurrent solves of one sphere (one plan per
 * host thread) a single table build.
 */
int msc_plan_copy_g0(msc_plan* dst, const msc_plan* src, void* stream);

/*
 * Run each Arnoldi step of msc_gmres as one instantiated HIP graph (enable = 1)
 * or as individual launches (0, the default: as fast on MI355X, with no capture
 * on a plan's first solve).  Graph capture must not overlap
 * another host thread's allocation or synchronous copy calls, so callers that
 * drive several plans from several threads turn graphs off (msc_sweep does).
 * No reference counterpart.
 */
int msc_plan_use_graphs(msc_plan* plan, int enable);

/*
 * Solve `count` independent slabs (different frequencies, incidences, slab
 * sizes and contrasts) with `workers` host threads.  Replaces the sequential
 * per-frequency loop of compute_ocean_bottom_reflection
 * (cubic_scattering/ocean_bottom.py:290-307).  Each worker owns a HIP stream
 * and one plan per FFT grid (created at the largest slab of that grid and
 * re-shaped per instance), rebuilds G0 only when (omega, medium, a, slab)
 * change and autotunes tiles once per grid; solves on different workers
 * overlap on the device.  Blocks until all instances are done.  Every
 * instance records its own status and stats; the return value is the first
 * failing instance's status.  opts->precond must be NULL.
 */
int msc_sweep(msc_instance* instances, int count, const msc_gmres_opts* opts, int workers);
/*
 * msc_sweep with the instances of one padded FFT grid solved as batches
 * (msc_gmres_batched): every operator pass and Arnoldi kernel of a GMRES step
 * is one launch for a batch of slabs of that grid (as many as fill the device,
 * at most `max_batch`), whatever their sizes, frequencies and incidences.
 * Replaces the same per-frequency loop of compute_ocean_bottom_reflection
 * (cubic_scattering/ocean_bottom.py:290-307).  Per grid a pool of plans at
 * the grid's largest slab is re-shaped per instance; pools stay in a bounded
 * process-wide cache (40% of the device memory) for later sweeps until
 * msc_sweep_release.  Solves and source sums run on `stream`; the next batch's
 * set-up (G0, T0, incident field) runs meanwhile on a helper thread and
 * stream of the library's own.  Instances with the volume-averaged correction
 * are solved one at a time.  Blocks until all instances are done; statuses
 * and return value as msc_sweep.  opts->precond must be NULL.
 */
int msc_sweep_batched(msc_instance* instances, int count, const msc_gmres_opts* opts,
                      int max_batch, void* stream);
/* Free the plan pools msc_sweep_batched keeps between calls.  No reference
 * counterpart. */
int msc_sweep_release(void);
/* out[0..5] = M, N_z, Pz, Pxy, NUV (table columns), NC (table depth). */
int msc_plan_info(const msc_plan* plan, long long* out6);

/*
 * Tile autotuning: time every tile candidate of the five FFT passes in whole
 * operator applications
 * for one precision and keep the fastest (needs msc_build_g0 and
 * msc_set_tmatrix first).  msc_gmres tunes each precision it uses once per
 * plan unless the environment sets MSC_AUTOTUNE=0.  tiles: out10 = for
 * complex128 then complex64, (y-forward rows, x-forward tile, z columns,
 * x-inverse tile, y-inverse rows) per workgroup; an x tile of 1, 2 or 4 is
 * that many whole ky columns (9 complex each) of every x row, 8, 16 or 32 is
 * that many consecutive complex of every row (whole cache lines).  No
 * counterpart in the reference (its FFTs are library calls).
 */
int msc_plan_autotune(msc_plan* plan, int precision, void* stream);
int msc_plan_tiles(const msc_plan* plan, int* out10);
/* Set the tiles explicitly (same layout as msc_plan_tiles; marks both
 * precisions tuned).  A tile larger than the LDS allows for the plan's FFT
 * lengths falls back to the next smaller one.  A z tile of Q + 256 (Q = 2, 4,
 * 8) selects the persistent z pass: resident workgroups walk the tiles and
 * prefetch the next tile's rows into LDS with direct-to-LDS loads.  A z tile
 * of Q + 512 selects the pruned z pass: the forward z FFT's first pass reads
 * the N_z non-zero rows straight from HBM and the inverse's last pass writes
 * the N_z kept rows straight back. */
int msc_plan_set_tiles(msc_plan* plan, const int* tiles10);

/*
 * Frequency constants of the Rayleigh cube T-matrix on the host:
 * out8 = (Gamma0, A^c, B^c, C^c) as (re, im) pairs.
 * Replaces _compute_Gamma0_analytical and _compute_ABC_polynomial
 * (cubic_scattering/effective_contrasts.py:112-152, 278-357).
 */
int msc_cube_coefficients(double omega, double a, double alpha, double beta, double rho,
                          double* out8);

/*
 * Spectral G0 table on the device (FP64 evaluation, stored complex128 and
 * complex64).  Replaces _build_slab_kernels (slab_scattering.py:267-351) and
 * _build_slab_kernels_gpu (slab_scattering_gpu.py:34-92).  volume_averaged
 * != 0 replaces the point propagator at the 26 nearest-neighbour offsets by
 * the inter-voxel block of inter_voxel_propagator_9x9 (slab_scattering.py:
 * 316-329; inter_voxel_propagator.py:1295) with n_orders in 0..3 frequency orders;
 * it is applied as a 26-block stencil after the spectral passes, so every
 * operator entry point (matvec, residual, scattered, GMRES) carries it.
 */
int msc_build_g0(msc_plan* plan, double omega, double alpha, double beta, double rho,
                 int volume_averaged, int n_orders, void* stream);

/*
 * Per-voxel T0 from contrast arrays (device float64, (N_z, M, M) each) and
 * the frequency constants of msc_cube_coefficients.  Replaces
 * compute_slab_tmatrices (slab_scattering.py:169-205).
 */
int msc_set_tmatrix(msc_plan* plan, const double* dlambda, const double* dmu, const double* drho,
                    double omega, const double* coef8, void* stream);

/*
 * Per-voxel T0 given directly as 4 complex128 parameters (t_u, t_d, t_o, t_s)
 * per voxel: T0 = blockdiag(t_u I3, [[t_d, t_o, t_o], [t_o, t_d, t_o],
 * [t_o, t_o, t_d]], t_s I3).  get copies them out.
 */
int msc_set_tparams(msc_plan* plan, const void* tparams_c128, void* stream);
int msc_get_tparams(const msc_plan* plan, void* out_c128, void* stream);

/*
 * A general 9x9 T-matrix per voxel (for instance the resonance-regime
 * composite T-matrices of msc_resonance_tmatrices): voxel v applies
 * table[tidx[v]] (row-major, tau = T psi) in the operator, the source sums
 * and the learned preconditioner's input.  tidx: device int32 array of the
 * slab's N_z*M*M voxels with values in [0, count) (checked: out of range is
 * MSC_ERR_INVALID); table: device count x 81 complex128.  Both are copied.
 * The reference's slab operator takes the cubic Rayleigh T0 only
 * (slab_scattering.py:169-205, 359-418, resonance_tmatrix.py:650-850 for the
 * composite T); this extends it to the resonance regime.  Not combinable with
 * the volume-averaged correction (MSC_ERR_UNSUPPORTED).  msc_set_tmatrix or
 * msc_set_tparams switches back to the cubic T0.  Synchronises `stream`.
 */
int msc_set_tgeneral(msc_plan* plan, const int* tidx, const void* table_c128, int count,
                     void* stream);

/*
 * Incident plane wave psi0 = inc9 * exp(i k khat . x_c) (complex128 out).
 * inc18 = 9 complex (re, im) on the host: polarisation and Voigt strain.
 * Replaces _build_slab_incident_field (slab_scattering.py:426-478).
 */
int msc_incident_field(msc_plan* plan, double k_mag, const double* khat3, const double* inc18,
                       void* psi0_c128, void* stream);

/*
 * y = (I - G0 T0) x in the given precision (x, y complex128 or complex64).
 * Replaces _slab_matvec_gpu (slab_scattering_gpu.py:95-155) and
 * _slab_matvec (slab_scattering.py:359-418).
 */
int msc_matvec(msc_plan* plan, const void* x, void* y, int precision, void* stream);

/* r = b - (I - G0 T0) x, fused (complex128 or complex64). */
int msc_residual(msc_plan* plan, const void* b, const void* x, void* r, int precision,
                 void* stream);

/*
 * GMRES solve of (I - G0 T0) x = b (b, x complex128).  Replaces torch_gmres
 * (cubic_scattering/torch_gmres.py:83-205) as called by
 * compute_slab_scattering_gpu (slab_scattering_gpu.py:229-241) and the SciPy
 * gmres call of compute_slab_scattering (slab_scattering.py:532-552).
 * Synchronises `stream` once per cycle.  Non-convergence is not an error
 * (stats->converged = 0), as the reference only warns.
 */
int msc_gmres(msc_plan* plan, const void* b, void* x, const msc_gmres_opts* opts,
              msc_gmres_stats* stats, void* stream);
/* msc_gmres on `count` plans of one padded FFT grid (same P_z, P_xy and
 * periodic flag; slab sizes may differ) as one batch: every operator pass and
 * every Arnoldi kernel is a single launch for all systems, and each system
 * keeps its own Krylov space, tolerance, restarts and statistics (stats[i]).
 * Replaces the per-frequency solve loop of ocean_bottom.py:290-307 (one
 * msc_gmres per frequency); cubic T0 only (no msc_set_tgeneral, no
 * volume-averaged correction), no preconditioner.  A system whose complex64
 * cycles stall is finished in FP64 cycles, as msc_gmres does. */
int msc_gmres_batched(msc_plan* const* plans, const void* const* b, void* const* x, int count,
                      const msc_gmres_opts* opts, msc_gmres_stats* stats, void* stream);

/*
 * Sums of the sources tau = T0 psi (psi complex128): per layer
 * layer_sum (N_z, 9) and per-layer phased sums phased (N_z, 2, 9) of
 * exp(-i k r_hat . x_v) tau_v for k = kP, kS.  Device outputs, complex128.
 * Back end of slab_rpp_periodic (slab_scattering.py:672-735) and
 * slab_reflected_field (slab_scattering.py:589-669).
 */
int msc_source_sums(msc_plan* plan, const void* psi, double kP, double kS, const double* rhat3,
                    void* layer_sum, void* phased, void* stream);

/*
 * Batched FFT along one axis of data viewed as [outer][n][inner] (complex,
 * in-place allowed), forward unscaled / inverse unscaled.
 */
int msc_fft_axis(const void* in, void* out, int outer, int n, long long inner, int inverse,
                 int precision, void* stream);

/*
 * Arnoldi building blocks over a caller-owned basis V ((m+1) rows of ldv
 * complex, precision of the context) for a caller-supplied operator -- the
 * engine of the generic torch_gmres (torch_gmres.py:83-205).
 *   start : V[0] = r / ||r||, g = (||r||, 0, ...); the cycle stops when
 *           |g_{j+1}| <= tol_abs.
 *   step  : the caller has written A V[j] into V[j+1]; orthogonalise,
 *           update the Givens QR and normalise V[j+1].
 *   state : out4 = (steps j, done flag, |g_{j+1}|, ||r||) after a sync.
 *   solution: x += sum_{i<j} y_i V[i] (x in the context precision).
 */
int msc_arnoldi_create(msc_arnoldi** ctx, long long n, int m, int precision);
int msc_arnoldi_destroy(msc_arnoldi* ctx);
int msc_arnoldi_start(msc_arnoldi* ctx, void* V, long long ldv, const void* r, double tol_abs,
                      void* stream);
int msc_arnoldi_step(msc_arnoldi* ctx, void* V, long long ldv, int j, int orth, void* stream);
int msc_arnoldi_state(msc_arnoldi* ctx, double* out4, void* stream);
int msc_arnoldi_solution(msc_arnoldi* ctx, const void* V, long long ldv, void* x, void* stream);

/*
 * Kernel timing with HIP events recorded around every launch of the
 * block-multiply pass (pass 3) on the launch stream: enable resets the
 * counters; read returns (complex64 launches, their total ms, complex128
 * launches, their total ms).
 */
int msc_timing_enable(msc_plan* plan, int enable);
int msc_timing_read(msc_plan* plan, double* out4);
/* The individual pass-3 launch durations (ms) of one precision (8: complex64,
 * 16: complex128) recorded since timing was enabled, in launch order: at most
 * cap values to out, their number to *n.  Launches that a raised done flag
 * turned into no-ops are among them (a few microseconds each).  No reference
 * counterpart. */
int msc_timing_durations(msc_plan* plan, int elem_bytes, double* out, int cap, int* n);

#ifdef __cplusplus
}
#endif

#endif /* MSC_SLAB_H */
