// Internal C++ interface of the bfloat16 convolutions (nn.hip) and of the
// learned preconditioner (precond.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "visual-simulation-of-smoke-3d-nn.h"

namespace smoke {
namespace nn {

uint64_t packed_bytes(const SmokeNnConvDesc& d);
void pack(const SmokeNnConvDesc& d, const float* w, uint16_t* fwd, uint16_t* bwd, hipStream_t s);
void forward(const SmokeNnConvDesc& d, const uint16_t* x, const uint16_t* w, const float* bias, const uint16_t* skip, int relu, uint16_t* y, float* yf,
             int f32_channels, hipStream_t s);
void backward_data(const SmokeNnConvDesc& d, const uint16_t* dy, const uint16_t* wb, const uint16_t* add, const uint16_t* mask, uint16_t* dx,
                   hipStream_t s);
void backward_weights(const SmokeNnConvDesc& d, const uint16_t* x, const uint16_t* dy, float* dw, float* db, hipStream_t s);

}  // namespace nn
}  // namespace smoke

// The learned preconditioner as the pressure solve uses it: z = M(r) on one
// volume, on the caller's stream.
struct SmokeNnPreconditioner;
namespace smoke {
void preconditioner_apply(SmokeNnPreconditioner& net, const float* r, const uint8_t* flags, const int dims[3], float* z, hipStream_t s);

// The multilevel preconditioner's coarse-grid solve with the network: the
// deflation of v = 1 (defl[0] = v^T A v on entry; writes defl[1] = c and
// the deflated right-hand side rd), z' = M(rd), `sweeps` damped-Jacobi
// sweeps on A z' = rd from z', and z = z' + (c - Av.z' / v.Av) v, in two
// single-block launches around the network's convolutions.  For volumes of
// at most kCoarseSolveMaxCells cells.
constexpr int64_t kCoarseSolveMaxCells = 8192;
constexpr int kCoarseSolveMaxSweeps = 64;
// the coarse level's operator as the sweeps see it: the flag byte's
// neighbour bits (a set bit at the domain edge is a periodic neighbour) plus
// the OUTFLOW domain faces (x-, x+, y-, y+, z-, z+) on the diagonal; the
// damping and the sweep count
struct CoarseSweeps {
    int outflow[6];
    float omega;
    int count;
};
void preconditioner_coarse_solve(SmokeNnPreconditioner& net, const float* r, const float* av, const uint8_t* flags, double* defl, const int dims[3],
                                 const CoarseSweeps& sweeps, float* rd, float* z, hipStream_t s);
}  // namespace smoke
