/*
 * Building blocks of the learned pressure preconditioner (MI355X
 * implementation): 3x3x3 convolutions over channels-last bfloat16 volumes on
 * the gfx950 matrix cores, with FP32 accumulation.  No counterpart exists in
 * the reference; these entry points serve the preconditioner's training and
 * its tests.  All pointers are device pointers; every call is ordered on
 * `stream` (a hipStream_t, NULL = the default stream) and asynchronous.
 *
 * Volume layout: element (b, z, y, x, c) of a batch of volumes with C
 * channels is at ((((b * D + z) * H + y) * W + x) * C + c), bfloat16.
 *
 * Convolution modes, output voxel o, tap t in {0,1,2}^3, input voxel i:
 *   SAME (stride 1):            i = o + t - 1                 (out dims = in dims)
 *   DOWN (stride 2):            i = 2 o + t - 1               (out dims = ceil(in / 2))
 *   UP   (transposed, stride 2): o = 2 i + t - 1, i.e. i = (o + 1 - t) / 2 when
 *                               even                          (out dims = 2 in)
 * Input voxels outside the volume read as zero.  y[o][n] = bias[n] +
 * sum_{t, c} W[t][c][n] * x[i(o, t)][c].  Weights W are FP32 [27][cin][cout]
 * with t = (tz * 3 + ty) * 3 + tx; the kernels read them packed to bfloat16
 * by smoke_nn_pack_weights.  cin must be a multiple of 8 and cout of 16
 * (pad with zero channels).
 */
#ifndef VISUAL_SIMULATION_OF_SMOKE_3D_NN_H
#define VISUAL_SIMULATION_OF_SMOKE_3D_NN_H

#include <stdint.h>

#include "visual-simulation-of-smoke-3d.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef enum SmokeNnConvMode {
    SMOKE_NN_CONV_SAME = 0,
    SMOKE_NN_CONV_DOWN = 1,
    SMOKE_NN_CONV_UP = 2,
} SmokeNnConvMode;

typedef struct SmokeNnConvDesc {
    int32_t batch;
    int32_t in[3];    /* W, H, D of the input volume */
    int32_t out[3];   /* W, H, D of the output volume (must match the mode) */
    int32_t cin;
    int32_t cout;
    int32_t mode;     /* SmokeNnConvMode */
} SmokeNnConvDesc;

/* Bytes of one packed weight array (forward or data-gradient layout). */
SMOKE_SIMULATION_API uint64_t smoke_nn_packed_weight_bytes(const SmokeNnConvDesc* desc);

/* FP32 weights [27][cin][cout] -> bfloat16 forward layout and data-gradient
   layout (either output may be NULL). */
SMOKE_SIMULATION_API SmokeSimulationResult smoke_nn_pack_weights(const SmokeNnConvDesc* desc, const float* weights, uint16_t* forward_packed,
                                                                 uint16_t* backward_packed, void* stream);

/* y = act(bias + conv(x) + skip): bias (FP32, cout) and skip (bfloat16, the
   output's shape) may be NULL; relu != 0 applies max(., 0).  Writes the
   bfloat16 output y and/or, when y_f32 is not NULL, the FP32 values of the
   first `f32_channels` channels as a [voxels][f32_channels] array. */
SMOKE_SIMULATION_API SmokeSimulationResult smoke_nn_conv_forward(const SmokeNnConvDesc* desc, const uint16_t* x, const uint16_t* forward_packed,
                                                                 const float* bias, const uint16_t* skip, int32_t relu, uint16_t* y, float* y_f32,
                                                                 int32_t f32_channels, void* stream);

/* Gradient with respect to the input: dx = (conv^T(dy) + add) * [mask > 0],
   where add (bfloat16, the input's shape) and mask (bfloat16, the input's
   shape: the forward activation whose ReLU the gradient passes back through)
   may be NULL. */
SMOKE_SIMULATION_API SmokeSimulationResult smoke_nn_conv_backward_data(const SmokeNnConvDesc* desc, const uint16_t* dy, const uint16_t* backward_packed,
                                                                       const uint16_t* add, const uint16_t* mask, uint16_t* dx, void* stream);

/* Accumulates dW[27][cin][cout] += sum_o x[i(o,t)][c] dy[o][n] and, when db is
   not NULL, db[n] += sum_o dy[o][n] (FP32, atomically: the summation order
   is not fixed). */
SMOKE_SIMULATION_API SmokeSimulationResult smoke_nn_conv_backward_weights(const SmokeNnConvDesc* desc, const uint16_t* x, const uint16_t* dy, float* dw,
                                                                          float* db, void* stream);

/* Kernel choice for the convolutions above, process-wide.  AUTO (the
   default) runs the LDS-staged kernels on volumes with many output tiles and
   the gather kernels (split-K for small volumes) elsewhere; GATHER forces
   the gather kernels; STAGED runs the staged kernels wherever their shape
   table covers (mode, cin, cout), at any volume size.  Forward and data
   gradients give the same bits with either choice; the weight gradient's
   atomic sums do not have a fixed order.  For tests and measurements. */
typedef enum SmokeNnConvKernels {
    SMOKE_NN_CONV_KERNELS_AUTO = 0,
    SMOKE_NN_CONV_KERNELS_GATHER = 1,
    SMOKE_NN_CONV_KERNELS_STAGED = 2,
} SmokeNnConvKernels;

SMOKE_SIMULATION_API SmokeSimulationResult smoke_nn_set_conv_kernels(int32_t choice);

/* The learned preconditioner: a 3-level convolutional encoder-decoder with
   3x3x3 kernels and 16-32-64 channels (stride-2 down, transposed stride-2 up,
   ReLU, skip additions), mapping a residual r to z ~ A^-1 r.
     input (8 channels: r / s, identity-row mask, 6 zero channels)
     c0a SAME 8->16, c0b SAME 16->16 = e0;  d1 DOWN 16->32, c1 SAME 32->32 = e1;
     d2 DOWN 32->64, c2 SAME 64->64 = e2;   u1 = UP(e2) 64->32 + e1, c1u SAME 32->32;
     u0 = UP(c1u) 32->16 + e0, c0u SAME 16->16;  out SAME 16->1 (no ReLU)
   every layer but `out` followed by ReLU; s is the RMS of r over the volume,
   z = s * out, and z = 0 on identity rows (solid cells and the pressure
   anchor), so M(a r) = a M(r).  Parameters: per layer W [27][cin][cout] then
   bias [cout], FP32 masters (the output layer is stored with 16 channels of
   which only channel 0 is used), bfloat16 copies for the matrix cores.
   Training minimises mean_b ||A z_b - r_b||^2 / ||r_b||^2 over a batch of
   volumes with unit-RMS r, A the pressure operator restricted to the volume
   (flags as in the training-pair files; neighbours outside the volume read as
   zero), with Adam (beta 0.9, 0.999, eps 1e-8). */
typedef struct SmokeNnPreconditioner SmokeNnPreconditioner;

SMOKE_SIMULATION_API SmokeSimulationResult smoke_nn_preconditioner_create(uint64_t seed, SmokeNnPreconditioner** out);
SMOKE_SIMULATION_API SmokeSimulationResult smoke_nn_preconditioner_destroy(SmokeNnPreconditioner* net);
SMOKE_SIMULATION_API int64_t smoke_nn_preconditioner_parameter_count(const SmokeNnPreconditioner* net);
/* host arrays of parameter_count floats */
SMOKE_SIMULATION_API SmokeSimulationResult smoke_nn_preconditioner_get_parameters(SmokeNnPreconditioner* net, float* out);
SMOKE_SIMULATION_API SmokeSimulationResult smoke_nn_preconditioner_set_parameters(SmokeNnPreconditioner* net, const float* values);
SMOKE_SIMULATION_API SmokeSimulationResult smoke_nn_preconditioner_save(SmokeNnPreconditioner* net, const char* path);
SMOKE_SIMULATION_API SmokeSimulationResult smoke_nn_preconditioner_load(SmokeNnPreconditioner* net, const char* path);

/* One training evaluation on `batch` volumes of dims (W, H, D): r FP32 and
   flags uint8, [batch][D][H][W] device arrays.  Writes the loss to *loss;
   copies the parameter gradients to `host_gradients` (parameter_count floats)
   when not NULL; applies one Adam step with learning rate `lr` when lr > 0.
   Synchronous. */
SMOKE_SIMULATION_API SmokeSimulationResult smoke_nn_preconditioner_train_step(SmokeNnPreconditioner* net, const float* r, const uint8_t* flags,
                                                                              int32_t batch, const int32_t dims[3], float lr, double* loss,
                                                                              float* host_gradients, void* stream);

/* Training loss of later train_step calls:
     SMOKE_NN_LOSS_RESIDUAL (default): mean_b ||A z_b - r_b||^2 / ||r_b||^2;
     SMOKE_NN_LOSS_ENERGY: mean_b (z_b^T A z_b / 2 - z_b^T r_b) / ||r_b||^2, whose
       minimiser is also z = A^-1 r but which weighs the error e = z - A^-1 r
       by the energy norm e^T A e (what conjugate gradients minimise) instead
       of by |A e|^2, so smooth (small-eigenvalue) error is not discounted --
       the loss for a coarse-grid solver.  Its value is negative at the optimum. */
typedef enum SmokeNnLoss {
    SMOKE_NN_LOSS_RESIDUAL = 0,
    SMOKE_NN_LOSS_ENERGY = 1,
} SmokeNnLoss;

SMOKE_SIMULATION_API SmokeSimulationResult smoke_nn_preconditioner_set_loss(SmokeNnPreconditioner* net, uint32_t loss);

/* z = M(r) for one volume of dims (W, H, D) (FP32 device arrays, flags uint8). */
SMOKE_SIMULATION_API SmokeSimulationResult smoke_nn_preconditioner_apply(SmokeNnPreconditioner* net, const float* r, const uint8_t* flags,
                                                                         const int32_t dims[3], float* z, void* stream);

/* No counterpart in the reference.  For callers that cannot allocate device
   memory (code under HIP graph capture, tight serving loops, harnesses that
   forbid hipMalloc): the same z = M(r) as smoke_nn_preconditioner_apply, with
   the same kernels and the same bits, but the inference's workspace
   (activations, normalised residual, output, sums) lives in a static 16 MiB
   device array of the library instead of buffers the network allocates on
   its first call for a volume size; it calls no hipMalloc / hipFree.  The
   network's weights were packed at create or by an earlier call.  Limit: the
   workspace of a W x H x D volume is about 188 bytes per voxel (plus 256-byte
   alignment per buffer), so volumes up to about 89,000 voxels fit (44^3,
   64 x 48 x 28); a larger volume returns BACKEND_FAILURE and prints the
   reason to stderr.  Every GPU holds its own array until the process ends;
   calls on one GPU share it, so two calls that overlap on one GPU (two
   streams, or two host threads) race and must be ordered by the caller. */
SMOKE_SIMULATION_API SmokeSimulationResult smoke_nn_preconditioner_apply_static(SmokeNnPreconditioner* net, const float* r, const uint8_t* flags,
                                                                                const int32_t dims[3], float* z, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* VISUAL_SIMULATION_OF_SMOKE_3D_NN_H */
