/*
 * msc_precond.h -- learned 3-D convolutional preconditioner of the slab
 * Foldy-Lax solve (libmsc_slab.so, gfx950).
 *
 * The network maps the induced source of a residual r to a correction:
 *     z = P(r) = r + N((s / t_max) T0 r) / s,   s = sqrt(n) / ||r||
 * (n = 9 N_z M^2, t_max = max over the slab of the 4 T0 parameters' moduli).
 * N is either four 3x3x3 zero-padded convolutions 18 -> h -> h -> h -> 18 over
 * the (N_z, M, M) voxel grid with leaky-ReLU (slope 0.1) between them
 * (msc_net_create), or a four-level network (msc_net_create_unet) whose
 * levels 1-3 are 2x, 4x, 8x coarser grids reached by 2x2x2 stride-2
 * convolutions and left by 2x2x2 stride-2 transposed convolutions added to
 * the finer level's activations.  Its 18 channels are the real and imaginary
 * parts of the 9 components, interleaved (channel 2c = Re, 2c+1 = Im of
 * component c).  Convolutions run as BF16
 * MFMA with FP32 accumulation; activations are BF16.  It is used as the right
 * preconditioner of flexible GMRES (msc_gmres_opts.precond), and is trained
 * on the GPU by msc_net_train_step from residual -> correction pairs.
 *
 * The reference has no learned preconditioner: it solves with SciPy gmres
 * (slab_scattering.py:532-552) or torch_gmres (torch_gmres.py:83-205), and
 * these entry points extend that solve.
 *
 * Parameters are one flat float32 vector, layer by layer in forward order:
 * weights [out][in][k][k][k] (k = 3, or 2 for the stride-2 layers; the
 * kernel offset (kz, kx, ky) within a 2x2x2 block is 4 kz + 2 kx + ky)
 * followed by the bias [out].
 * Training inputs are BF16 [sample][voxel][32] (msc_net_input; channels
 * 18..31 zero), targets float32 [sample][voxel][18].
 */
#ifndef MSC_PRECOND_H
#define MSC_PRECOND_H

#include "msc_slab.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Create a network of hidden width 16, 32 or 48 (parameters zero:
 * P(r) = r until parameters are set) on the current device. */
int msc_net_create(msc_net** net, int hidden);

/* Create the four-level network with level widths c4[0..3] (c4[0] = 16,
 * c4[1..3] in {16, 32, 48}; parameters zero) on the current device.  Layers:
 *   x0 = conv3(conv3(in)),  x1 = conv3(down(x0)),  x2 = conv3(down(x1)),
 *   x3 = conv3(down(x2)),   y2 = conv3(up(x3) + x2),  y1 = conv3(up(y2) + x1),
 *   out = conv3(conv3(up(y1) + x0)) -> 18 channels
 * with leaky-ReLU after every conv3 and down layer except the last conv3;
 * up layers are linear.  Level l has ceil(D / 2^l) x ceil(H / 2^l) x
 * ceil(W / 2^l) voxels.  (16, 16, 32, 48) has 197,938 parameters. */
int msc_net_create_unet(msc_net** net, const int* c4);
int msc_net_destroy(msc_net* net);

/* Number of float32 parameters (local network: 27*(18h + 2h^2 + 18h) + 3h + 18). */
int msc_net_num_params(const msc_net* net, long long* n);

/* Copy the flat float32 parameter vector in / out (device pointers). */
int msc_net_set_params(msc_net* net, const float* params, void* stream);
int msc_net_get_params(const msc_net* net, float* params, void* stream);

/* Network input (s / t_max) T0 r for the plan's grid as BF16 [voxel][32]
 * (r complex128 or complex64 by `precision`); scale <= 0 selects
 * s = sqrt(n) / ||r||, else s = scale. */
int msc_net_input(msc_plan* plan, const void* r, int precision, double scale, void* out_bf16,
                  void* stream);

/* Forward pass of a [D][H][W][32] BF16 input: out float32 [voxel][18]. */
int msc_net_forward(msc_net* net, int D, int H, int W, const void* in_bf16, float* out,
                    void* stream);

/* Gradient of loss = mean over samples, voxels and the 18 channels of
 * (N(in) - target)^2 for `batch` contiguous samples: grads float32 (flat
 * parameter layout, device), loss on the host (may be NULL). */
int msc_net_gradients(msc_net* net, int D, int H, int W, int batch, const void* in_bf16,
                      const float* target, float* grads, double* loss, void* stream);

/* One Adam step (beta 0.9 / 0.999, eps 1e-8) on that loss; loss before the
 * step on the host (may be NULL). */
int msc_net_train_step(msc_net* net, int D, int H, int W, int batch, const void* in_bf16,
                       const float* target, double lr, double* loss, void* stream);

/* Copy vector j (0 <= j <= restart) of the orthonormal Krylov basis of the
 * plan's last GMRES cycle in `precision` (plain GMRES builds it from
 * G0 T0 and psi0: the same space as (I - G0 T0) and psi0).  Training pairs
 * are harvested from it: input (I - G0 T0) v_j, target v_j. */
int msc_gmres_basis(msc_plan* plan, int precision, int j, void* out, void* stream);

/* Apply the preconditioner on the plan's grid: z = r + N((s / t_max) T0 r) / s
 * with s = sqrt(n) / ||r|| (r, z complex128 or complex64 by `precision`). */
int msc_net_apply(msc_net* net, msc_plan* plan, const void* r, void* z, int precision,
                  void* stream);

#ifdef __cplusplus
}
#endif

#endif /* MSC_PRECOND_H */
