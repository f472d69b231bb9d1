// Device-side operator assembly entry points (implemented in msc_setup.hip).
#pragma once

#include "msc_plan.hpp"

namespace msc {

// The G0 build's two octant workspaces (NC x H x H x 45 complex128 each,
// H = Pxy/2 + 1), allocated once and kept for rebuilds.
int ensure_build_work(Plan& p);
int set_tparams_from_contrasts(Plan& p, const double* Dl, const double* Dm, const double* Dr,
                               double omega, const double* coef8, hipStream_t st);
int set_tparams_direct(Plan& p, const double2* tp, hipStream_t st);
// General per-voxel T: tidx (device, nvox indices in [0, count)), table (device,
// count x 81 complex128, row-major).  Validates the indices.
int set_tgeneral(Plan& p, const int* tidx, const double2* table, int count, hipStream_t st);
// (t_u, t_d, t_o, t_s) of `count` cubes of volume V from contrast arrays
// (device), with the frequency constants of msc_cube_coefficients.
int tparams_batch(const double* Dl, const double* Dm, const double* Dr, size_t count,
                  double omega, const double* coef8, double V, double2* t128, hipStream_t st);
int incident_field(Plan& p, double k_mag, const double* khat3, const double* inc18, double2* psi0,
                   hipStream_t st);
// The 26 volume-averaged nearest-neighbour 9x9 blocks (row-major, real) for
// R = (dz, dx, dy) in {-1,0,1}^3 \ 0, lexicographic (msc_va.hip).
void va_blocks(double alpha, double beta, double rho, double omega, int n_orders,
               double* out26x81);
int source_sums(Plan& p, const double2* psi, double kP, double kS, const double* rhat3,
                double2* layer_sum, double2* phased, hipStream_t st);

}  // namespace msc
