// Matrix-free 7-point pressure solve.
//
// Replaces the host CSR build + cuSPARSE SpMV + cuBLAS dot/axpy loop of
// backend-cuda.cu:1688-1824.  The operator is applied from one flag byte per
// cell (fluid neighbours + identity row) and the cell coordinates (OUTFLOW
// faces add to the diagonal), which reproduces the CSR values of
// B:1343-1412 exactly: the anchor keeps its -1 coupling here, which is exact
// because the anchor entry of every iterate is 0.
//
// One PCG iteration is two fused passes (34 bytes of compulsory HBM traffic
// per cell):
//   pass A: p <- z + beta p (z = r / diag on the fly), Ap, partial p.Ap
//           reads r, p_old, flags; writes p_new                   (13 B/cell)
//   pass B: Ap recomputed from p_new, e += alpha p, r -= alpha Ap,
//           partial r.z and r.r
//           reads p_new, flags, e, r; writes e, r                  (21 B/cell)
// Both march along z with the current plane (plus an x/y halo ring) in LDS and
// the z neighbours in registers.  Reductions are FP64: wave64 shuffles, LDS
// across waves, per-block partials, and the last block to finish sums the
// partials in block order (deterministic) and updates alpha / beta on the
// device, so fixed-iteration solves never return to the host.
#include "context.hpp"

#include <algorithm>
#include <vector>

namespace smoke {
namespace {

constexpr uint8_t kIdentity = 0x40;

struct PcgArgs {
    int n[3];
    int periodic[3];
    int outflow[6];
    unsigned skip;        // off-diagonal directions ignored (periodic axes with n <= 2)
    int zchunk;
    int tiles[3];         // tile grid (x tiles, y tiles, z chunks)
    const uint8_t* flags;
    PcgScalars* s;
    double* partials;
};

__device__ __forceinline__ float diag_of(const PcgArgs& a, uint8_t f, int x, int y, int z) {
    if (f & kIdentity) return 1.0f;
    int d = __popc(unsigned(f & 0x3f));
    d += (x == 0 && a.outflow[0]) + (x == a.n[0] - 1 && a.outflow[1]);
    d += (y == 0 && a.outflow[2]) + (y == a.n[1] - 1 && a.outflow[3]);
    d += (z == 0 && a.outflow[4]) + (z == a.n[2] - 1 && a.outflow[5]);
    return d > 0 ? float(d) : 1.0f;
}

// Resolves a possibly out-of-range cell: periodic axes wrap, otherwise the
// cell does not exist (its value is never used by the stencil).
__device__ __forceinline__ bool resolve(const PcgArgs& a, int& x, int& y, int& z) {
    if (x < 0 || x >= a.n[0]) {
        if (!a.periodic[0]) return false;
        x = wrapi(x, a.n[0]);
    }
    if (y < 0 || y >= a.n[1]) {
        if (!a.periodic[1]) return false;
        y = wrapi(y, a.n[1]);
    }
    if (z < 0 || z >= a.n[2]) {
        if (!a.periodic[2]) return false;
        z = wrapi(z, a.n[2]);
    }
    return true;
}

__device__ __forceinline__ int64_t lin(const PcgArgs& a, int x, int y, int z) { return (int64_t(z) * a.n[1] + y) * a.n[0] + x; }

// Tile index of this block; consecutive tiles are kept on one XCD (blocks are
// dealt round-robin over the 8 XCDs) so that neighbouring tiles share halo
// rows through the same L2.
__device__ __forceinline__ int tile_id(int nblocks) {
    const int b = blockIdx.x;
    if ((nblocks & 7) != 0) return b;
    return (b & 7) * (nblocks >> 3) + (b >> 3);
}

template <int NV, int NWAVES>
__device__ __forceinline__ void block_sum(double (&v)[NV], double (&red)[NV][NWAVES]) {
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        double x = v[q];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
        v[q] = x;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
#pragma unroll
        for (int q = 0; q < NV; ++q) red[q][wave] = v[q];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NV; ++q) {
            double s = 0.0;
            for (int w = 0; w < NWAVES; ++w) s += red[q][w];
            v[q] = s;
        }
    }
}

// Writes this block's partials, and returns true in the one block that
// finishes last, with `tot` holding the sums over all blocks in block order.
template <int NV, int NWAVES>
__device__ bool last_block_total(double (&v)[NV], double (&tot)[NV], double* partials, unsigned* ticket, int nblocks, int bid) {
    __shared__ double red[NV][NWAVES];
    __shared__ int is_last;
    block_sum<NV, NWAVES>(v, red);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NV; ++q) partials[q * nblocks + bid] = v[q];
        __threadfence();
        const unsigned t = atomicAdd(ticket, 1u);
        is_last = t == unsigned(nblocks - 1);
    }
    __syncthreads();
    if (!is_last) return false;
    __threadfence();
    double acc[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        double s = 0.0;
        for (int b = threadIdx.x; b < nblocks; b += NWAVES * 64)
            s += __hip_atomic_load(&partials[q * nblocks + b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        acc[q] = s;
    }
    __syncthreads();
    block_sum<NV, NWAVES>(acc, red);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NV; ++q) tot[q] = acc[q];
        *ticket = 0u;
    }
    return threadIdx.x == 0;
}

__device__ __forceinline__ double guarded(double num, double den, bool fixed) {
    const double lim = fixed ? 1.0e-20 : 1.0e-300;
    return fabs(den) > lim ? num / den : 0.0;
}

// One z-marching pass over a TX x TY x zchunk tile.
template <int TX, int TY, bool PASS_A, bool JACOBI>
__global__ __launch_bounds__(TX* TY) void pcg_pass_kernel(PcgArgs a, float* __restrict__ r, const float* __restrict__ p_in,
                                                           float* __restrict__ p_out, float* __restrict__ e) {
    PcgScalars* __restrict__ S = a.s;
    if (S->done) return;
    constexpr int NWAVES = TX * TY / 64;
    const int nblocks = a.tiles[0] * a.tiles[1] * a.tiles[2];
    const int tile = tile_id(nblocks);
    const int bx = tile % a.tiles[0];
    const int by = (tile / a.tiles[0]) % a.tiles[1];
    const int bz = tile / (a.tiles[0] * a.tiles[1]);
    const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
    const int x0 = bx * TX, y0 = by * TY;
    const int x = x0 + tx, y = y0 + ty;
    const int z0 = bz * a.zchunk;
    const int z1 = min(z0 + a.zchunk, a.n[2]);
    const bool owner = x < a.n[0] && y < a.n[1];
    const bool first = S->first != 0;
    const float beta = first ? 0.0f : float(S->beta);
    const float alpha = float(S->alpha);

    // Operand of the stencil at a cell: pass A builds p_new = z + beta p_old,
    // pass B reads p_new.
    auto operand = [&](int cx, int cy, int cz) -> float {
        if (!resolve(a, cx, cy, cz)) return 0.0f;
        const int64_t i = lin(a, cx, cy, cz);
        if constexpr (PASS_A) {
            const float rv = r[i];
            float zv = rv;
            if constexpr (JACOBI) zv = rv / diag_of(a, a.flags[i], cx, cy, cz);
            return first ? zv : zv + beta * p_in[i];
        } else {
            return p_in[i];
        }
    };

    // halo ring slot of this thread (x halo first, then y halo)
    const int q = threadIdx.x;
    int ring_x = 0, ring_y = 0, ring_lx = 0, ring_ly = 0;
    const bool ring = q < 2 * TY + 2 * TX;
    if (q < 2 * TY) {
        const int side = q / TY, row = q % TY;
        ring_x = side ? x0 + TX : x0 - 1;
        ring_y = y0 + row;
        ring_lx = side ? TX + 1 : 0;
        ring_ly = row + 1;
    } else if (ring) {
        const int qq = q - 2 * TY, side = qq / TX, col = qq % TX;
        ring_x = x0 + col;
        ring_y = side ? y0 + TY : y0 - 1;
        ring_lx = col + 1;
        ring_ly = side ? TY + 1 : 0;
    }

    __shared__ float plane[TY + 2][TX + 2];
    float vm = operand(x, y, z0 - 1);
    float vc = operand(x, y, z0);
    float rc = ring ? operand(ring_x, ring_y, z0) : 0.0f;
    double acc0 = 0.0, acc1 = 0.0;

    for (int z = z0; z < z1; ++z) {
        const float vp = operand(x, y, z + 1);
        const float rp = (ring && z + 1 < z1) ? operand(ring_x, ring_y, z + 1) : 0.0f;
        plane[ty + 1][tx + 1] = vc;
        if (ring) plane[ring_ly][ring_lx] = rc;
        __syncthreads();
        if (owner) {
            const int64_t i = lin(a, x, y, z);
            const uint8_t f = a.flags[i];
            float ap;
            float d = 1.0f;
            if (f & kIdentity) {
                ap = vc;
            } else {
                d = diag_of(a, f, x, y, z);
                const unsigned m = unsigned(f) & ~a.skip;
                float s = d * vc;
                s -= (m & 1u) ? plane[ty + 1][tx] : 0.0f;
                s -= (m & 2u) ? plane[ty + 1][tx + 2] : 0.0f;
                s -= (m & 4u) ? plane[ty][tx + 1] : 0.0f;
                s -= (m & 8u) ? plane[ty + 2][tx + 1] : 0.0f;
                s -= (m & 16u) ? vm : 0.0f;
                s -= (m & 32u) ? vp : 0.0f;
                ap = s;
            }
            if constexpr (PASS_A) {
                p_out[i] = vc;
                acc0 += double(vc) * double(ap);
            } else {
                e[i] = (first ? 0.0f : e[i]) + alpha * vc;
                const float rn = r[i] - alpha * ap;
                r[i] = rn;
                const float zn = JACOBI ? rn / d : rn;
                acc0 += double(rn) * double(zn);
                acc1 += double(rn) * double(rn);
            }
        }
        __syncthreads();
        vm = vc;
        vc = vp;
        rc = rp;
    }

    if constexpr (PASS_A) {
        double v[1] = {acc0}, tot[1];
        if (last_block_total<1, NWAVES>(v, tot, a.partials, &S->ticket[0], nblocks, tile)) {
            S->alpha = guarded(S->rz, tot[0], S->fixed_mode != 0);
        }
    } else {
        double v[2] = {acc0, acc1}, tot[2];
        if (last_block_total<2, NWAVES>(v, tot, a.partials, &S->ticket[1], nblocks, tile)) {
            const double rz_new = tot[0];
            S->beta = guarded(rz_new, S->rz, S->fixed_mode != 0);
            S->rz = rz_new;
            S->rr = tot[1];
            S->first = 0;
            const int it = S->iteration + 1;
            S->iteration = it;
            if (!S->fixed_mode && (tot[1] <= S->stop2 || it >= S->max_iterations)) S->done = 1;
        }
    }
}

// r = b - A x' with x' = x_old + e_add (either may be absent) evaluated in
// FP64; optionally stores x' (FP64) and starts a fresh inner solve.
__global__ __launch_bounds__(256) void residual_kernel(PcgArgs a, const float* __restrict__ b, const double* __restrict__ x_old,
                                                       const float* __restrict__ e_add, double* __restrict__ x_new, float* __restrict__ r,
                                                       int compute_bb, int reset_iterations) {
    PcgScalars* __restrict__ S = a.s;
    const int nblocks = gridDim.x * gridDim.y * gridDim.z;
    const int bid = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    // 256 threads = 64 (x) x 4 (y) cells of one z plane
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int z = blockIdx.z;
    double rr = 0.0, rz = 0.0, bb = 0.0;
    auto xv = [&](int cx, int cy, int cz) -> double {
        if (!resolve(a, cx, cy, cz)) return 0.0;
        const int64_t i = lin(a, cx, cy, cz);
        double s = x_old ? x_old[i] : 0.0;
        if (e_add) s += double(e_add[i]);
        return s;
    };
    if (x < a.n[0] && y < a.n[1]) {
        const int64_t i = lin(a, x, y, z);
        const uint8_t f = a.flags[i];
        const double xc = xv(x, y, z);
        double ax;
        float d = 1.0f;
        if (f & kIdentity) {
            ax = xc;
        } else {
            d = diag_of(a, f, x, y, z);
            const unsigned m = unsigned(f) & ~a.skip;
            ax = double(d) * xc;
            if (m & 1u) ax -= xv(x - 1, y, z);
            if (m & 2u) ax -= xv(x + 1, y, z);
            if (m & 4u) ax -= xv(x, y - 1, z);
            if (m & 8u) ax -= xv(x, y + 1, z);
            if (m & 16u) ax -= xv(x, y, z - 1);
            if (m & 32u) ax -= xv(x, y, z + 1);
        }
        const double bv = double(b[i]);
        const double r64 = bv - ax;
        if (x_new) x_new[i] = xc;
        const float rf = float(r64);
        r[i] = rf;
        const float zf = S->jacobi ? rf / d : rf;
        rr = r64 * r64;
        rz = double(rf) * double(zf);
        bb = bv * bv;
    }
    double v[3] = {rr, rz, bb}, tot[3];
    if (last_block_total<3, 4>(v, tot, a.partials, &S->ticket[2], nblocks, bid)) {
        if (compute_bb) S->bb = tot[2];
        S->rr_true = tot[0];
        S->rz = tot[1];
        S->rr = tot[0];
        S->beta = 0.0;
        S->first = 1;
        if (reset_iterations) S->iteration = 0;
        const double bnorm = sqrt(S->bb);
        const double target = S->tolerance * bnorm;
        const double inner = fmax(0.5 * target, S->inner_reduction * sqrt(tot[0]));
        S->stop2 = S->fixed_mode ? -1.0 : inner * inner;
        S->done = S->fixed_mode ? 0 : (sqrt(tot[0]) <= target ? 1 : 0);
    }
}

__global__ void pressure_from_x64_kernel(const double* __restrict__ x, float* __restrict__ p, int64_t n) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = float(x[i]);
}

__global__ void copy_kernel(const float4* __restrict__ src, float4* __restrict__ dst, int64_t n) {
    const int64_t stride = int64_t(gridDim.x) * blockDim.x;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

// b = -(h^2/dt) div(u*) + sum of OUTFLOW face pressures, 0 on identity rows (B:1304-1328).
__global__ void pressure_rhs_kernel(Grid g, const float* __restrict__ us, const float* __restrict__ vs, const float* __restrict__ ws,
                                    const uint8_t* __restrict__ flags, const float* __restrict__ dt_ptr, float* __restrict__ rhs) {
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int z = blockIdx.z;
    if (x >= g.n[0] || y >= g.n[1]) return;
    const int64_t i = cidx(g, x, y, z);
    if (flags[i] & kIdentity) {
        rhs[i] = 0.0f;
        return;
    }
    const float h = g.h, dt = *dt_ptr;
    const float div = (us[fidx(g, 0, x + 1, y, z)] - us[fidx(g, 0, x, y, z)] + vs[fidx(g, 1, x, y + 1, z)] - vs[fidx(g, 1, x, y, z)] +
                       ws[fidx(g, 2, x, y, z + 1)] - ws[fidx(g, 2, x, y, z)]) /
                      h;
    float bsum = 0.0f;
    if (x == 0 && g.flow_type[0] == kOutflow) bsum += g.flow_pressure[0];
    if (x == g.n[0] - 1 && g.flow_type[1] == kOutflow) bsum += g.flow_pressure[1];
    if (y == 0 && g.flow_type[2] == kOutflow) bsum += g.flow_pressure[2];
    if (y == g.n[1] - 1 && g.flow_type[3] == kOutflow) bsum += g.flow_pressure[3];
    if (z == 0 && g.flow_type[4] == kOutflow) bsum += g.flow_pressure[4];
    if (z == g.n[2] - 1 && g.flow_type[5] == kOutflow) bsum += g.flow_pressure[5];
    rhs[i] = -(h * h / dt) * div + bsum;
}

PcgArgs make_args(const Context& c) {
    PcgArgs a{};
    for (int k = 0; k < 3; ++k) {
        a.n[k] = c.grid.n[k];
        a.periodic[k] = c.grid.flow_periodic[k];
    }
    for (int f = 0; f < 6; ++f) a.outflow[f] = c.grid.flow_type[f] == kOutflow;
    unsigned skip = 0;
    for (int k = 0; k < 3; ++k) {
        if (!a.periodic[k]) continue;
        if (a.n[k] == 1) skip |= 3u << (2 * k);        // both neighbours are the cell itself
        else if (a.n[k] == 2) skip |= 2u << (2 * k);   // both neighbours are the same cell: one -1
    }
    a.skip = skip;
    a.zchunk = c.zchunk;
    a.tiles[0] = (c.grid.n[0] + c.tile_x - 1) / c.tile_x;
    a.tiles[1] = (c.grid.n[1] + c.tile_y - 1) / c.tile_y;
    a.tiles[2] = (c.grid.n[2] + c.zchunk - 1) / c.zchunk;
    a.flags = c.stencil_flags;
    a.s = c.scalars;
    a.partials = c.partials;
    return a;
}

template <int TX, int TY>
void launch_pass(Context& c, bool pass_a) {
    const PcgArgs a = make_args(c);
    const unsigned blocks = unsigned(a.tiles[0] * a.tiles[1] * a.tiles[2]);
    const float* p_old = c.pcg_p[c.p_current];
    float* p_new = c.pcg_p[c.p_current ^ 1];
    const bool jac = c.options.pressure_mode == SMOKE_SIMULATION_PRESSURE_TOLERANCE && c.options.preconditioner == SMOKE_SIMULATION_PRECONDITIONER_JACOBI;
    float* e = c.options.pressure_mode == SMOKE_SIMULATION_PRESSURE_TOLERANCE ? c.pcg_e : c.pressure;
    if (pass_a) {
        if (jac) pcg_pass_kernel<TX, TY, true, true><<<blocks, TX * TY, 0, c.stream>>>(a, c.pcg_r, p_old, p_new, e);
        else pcg_pass_kernel<TX, TY, true, false><<<blocks, TX * TY, 0, c.stream>>>(a, c.pcg_r, p_old, p_new, e);
    } else {
        if (jac) pcg_pass_kernel<TX, TY, false, true><<<blocks, TX * TY, 0, c.stream>>>(a, c.pcg_r, p_new, nullptr, e);
        else pcg_pass_kernel<TX, TY, false, false><<<blocks, TX * TY, 0, c.stream>>>(a, c.pcg_r, p_new, nullptr, e);
    }
}

void pass(Context& c, bool pass_a) {
    if (c.tile_y == 8) launch_pass<64, 8>(c, pass_a);
    else launch_pass<64, 4>(c, pass_a);
    hip_check(hipGetLastError(), "pcg pass");
}

void launch_residual(Context& c, const double* x_old, const float* e_add, double* x_new, bool bb, bool reset) {
    const PcgArgs a = make_args(c);
    const dim3 grid((c.grid.n[0] + 63) / 64, (c.grid.n[1] + 3) / 4, c.grid.n[2]);
    residual_kernel<<<grid, 256, 0, c.stream>>>(a, c.pressure_rhs, x_old, e_add, x_new, c.pcg_r, bb ? 1 : 0, reset ? 1 : 0);
    hip_check(hipGetLastError(), "residual_kernel");
}

}  // namespace

void configure_pcg(Context& c) {
    const int64_t plane_tiles = int64_t((c.grid.n[0] + 63) / 64) * ((c.grid.n[1] + 7) / 8);
    c.tile_x = 64;
    c.tile_y = c.grid.n[1] >= 8 ? 8 : 4;
    // enough tiles to fill 256 CUs several times over, z chunks of 4..32 planes
    int zc = int(std::max<int64_t>(4, std::min<int64_t>(32, plane_tiles * c.grid.n[2] / 2048)));
    c.zchunk = std::min(zc, c.grid.n[2]);
    const int64_t tiles = int64_t((c.grid.n[0] + c.tile_x - 1) / c.tile_x) * ((c.grid.n[1] + c.tile_y - 1) / c.tile_y) *
                          ((c.grid.n[2] + c.zchunk - 1) / c.zchunk);
    const int64_t residual_blocks = int64_t((c.grid.n[0] + 63) / 64) * ((c.grid.n[1] + 3) / 4) * c.grid.n[2];
    c.pcg_blocks = int(tiles);
    c.partial_capacity = size_t(std::max<int64_t>(2 * tiles, 3 * residual_blocks));
}

void enqueue_pressure_rhs(Context& c) {
    const dim3 grid((c.grid.n[0] + 63) / 64, (c.grid.n[1] + 3) / 4, c.grid.n[2]);
    pressure_rhs_kernel<<<grid, 256, 0, c.stream>>>(c.grid, c.face_star[0], c.face_star[1], c.face_star[2], c.stencil_flags, c.dt_device,
                                                    c.pressure_rhs);
    hip_check(hipGetLastError(), "pressure_rhs_kernel");
}

void enqueue_pcg_init(Context& c, bool use_x64) {
    launch_residual(c, use_x64 ? c.x64[c.x64_current] : nullptr, nullptr, nullptr, true, true);
}

void enqueue_pcg_iteration(Context& c) {
    pass(c, true);
    pass(c, false);
    c.p_current ^= 1;
}

void enqueue_pcg_refine(Context& c) {
    launch_residual(c, c.x64[c.x64_current], c.pcg_e, c.x64[c.x64_current ^ 1], false, false);
    c.x64_current ^= 1;
}

void enqueue_pressure_from_x64(Context& c) {
    const int64_t n = c.grid.cells;
    pressure_from_x64_kernel<<<unsigned((n + 255) / 256), 256, 0, c.stream>>>(c.x64[c.x64_current], c.pressure, n);
    hip_check(hipGetLastError(), "pressure_from_x64_kernel");
}

// Reference semantics (B:1800-1823): p = 0, K plain CG iterations, no test.
void enqueue_fixed_pcg(Context& c) {
    enqueue_pcg_init(c, false);
    for (int k = 0; k < c.config.pressure_iterations; ++k) enqueue_pcg_iteration(c);
}

// FP64 ||b - A p||^2 of the FP32 pressure (statistics of fixed mode).
void enqueue_true_residual(Context& c) {
    // x_old absent, e_add = pressure: x' = p
    launch_residual(c, nullptr, c.pressure, nullptr, true, false);
}

void run_tolerance_pcg(Context& c) {
    const auto& o = c.options;
    const int64_t n = c.grid.cells;
    if (!o.warm_start) {
        hip_check(hipMemsetAsync(c.x64[c.x64_current], 0, size_t(n) * sizeof(double), c.stream), "zero x64");
    }
    enqueue_pcg_init(c, true);
    const int chunk = std::max(1, o.check_interval);
    int passes = 0;
    PcgScalars host{};
    auto read_scalars = [&]() {
        hip_check(hipMemcpyAsync(&host, c.scalars, sizeof(PcgScalars), hipMemcpyDeviceToHost, c.stream), "read pcg scalars");
        hip_check(hipStreamSynchronize(c.stream), "pcg sync");
    };
    read_scalars();
    if (host.bb == 0.0) {
        // b = 0: the solution is x = 0 exactly
        hip_check(hipMemsetAsync(c.x64[c.x64_current], 0, size_t(n) * sizeof(double), c.stream), "zero x64");
        host.rr_true = 0.0;
        host.done = 1;
    }
    int observed = host.iteration;
    while (!host.done) {
        // inner solve: chunks of `chunk` iterations, at most two in flight; in
        // timing mode the first iteration of every chunk is bracketed by events
        int slot = 0;
        auto enqueue_chunk = [&](int s) {
            for (int k = 0; k < chunk; ++k) {
                if (c.timing && k == 0) {
                    hip_check(hipEventRecord(c.timing_events[s][0], c.stream), "event");
                    pass(c, true);
                    hip_check(hipEventRecord(c.timing_events[s][1], c.stream), "event");
                    pass(c, false);
                    hip_check(hipEventRecord(c.timing_events[s][2], c.stream), "event");
                    c.p_current ^= 1;
                } else {
                    enqueue_pcg_iteration(c);
                }
            }
            hip_check(hipMemcpyAsync(&c.host_status[s], &c.scalars->iteration, sizeof(PcgStatus), hipMemcpyDeviceToHost, c.stream), "status copy");
            hip_check(hipEventRecord(c.status_event[s], c.stream), "status event");
        };
        enqueue_chunk(0);
        enqueue_chunk(1);
        for (;;) {
            hip_check(hipEventSynchronize(c.status_event[slot]), "status wait");
            const PcgStatus st = c.host_status[slot];
            if (c.timing && st.iteration > observed) {
                float ma = 0.f, mb = 0.f;
                hip_check(hipEventElapsedTime(&ma, c.timing_events[slot][0], c.timing_events[slot][1]), "elapsed");
                hip_check(hipEventElapsedTime(&mb, c.timing_events[slot][1], c.timing_events[slot][2]), "elapsed");
                c.timing_totals.pcg_pass_a_ms += ma;
                c.timing_totals.pcg_pass_b_ms += mb;
                c.timing_totals.pcg_pass_a_launches += 1;
                c.timing_totals.pcg_pass_b_launches += 1;
            }
            if (st.iteration > observed) c.timing_totals.pcg_iterations += uint64_t(st.iteration - observed);
            observed = st.iteration > observed ? st.iteration : observed;
            if (st.done) break;
            enqueue_chunk(slot);
            slot ^= 1;
        }
        // FP64 refinement: x += e, true residual
        enqueue_pcg_refine(c);
        ++passes;
        read_scalars();
        if (host.done) break;
        if (passes > o.max_refinements || host.iteration >= o.max_iterations) break;
    }
    enqueue_pressure_from_x64(c);
    c.stats.iterations = host.iteration;
    c.stats.refinements = passes > 0 ? passes - 1 : 0;
    c.stats.rhs_norm = std::sqrt(host.bb);
    c.stats.relative_residual = host.bb > 0 ? std::sqrt(host.rr_true / host.bb) : 0.0;
    c.stats.converged = host.bb > 0 ? (std::sqrt(host.rr_true) <= o.relative_tolerance * std::sqrt(host.bb)) : 1;
}

double measure_copy(uint64_t bytes, int reps, double* ms_per_copy) {
    const int64_t n4 = int64_t(bytes / 16);
    float4 *a = nullptr, *b = nullptr;
    hip_check(hipMalloc(&a, size_t(n4) * 16), "copy src");
    hipError_t e = hipMalloc(&b, size_t(n4) * 16);
    if (e != hipSuccess) {
        hipFree(a);
        hip_check(e, "copy dst");
    }
    hipStream_t s;
    hip_check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "copy stream");
    hipEvent_t t0, t1;
    hip_check(hipEventCreate(&t0), "event");
    hip_check(hipEventCreate(&t1), "event");
    const unsigned blocks = 256u * 8u;
    copy_kernel<<<blocks, 256, 0, s>>>(a, b, n4);   // warm-up
    hip_check(hipEventRecord(t0, s), "event");
    for (int r = 0; r < reps; ++r) copy_kernel<<<blocks, 256, 0, s>>>(r & 1 ? b : a, r & 1 ? a : b, n4);
    hip_check(hipEventRecord(t1, s), "event");
    hip_check(hipEventSynchronize(t1), "event sync");
    float ms = 0.f;
    hipEventElapsedTime(&ms, t0, t1);
    hipEventDestroy(t0);
    hipEventDestroy(t1);
    hipStreamDestroy(s);
    hipFree(a);
    hipFree(b);
    const double per = double(ms) / std::max(1, reps);
    if (ms_per_copy) *ms_per_copy = per;
    return per > 0 ? 2.0 * double(n4) * 16.0 / (per * 1e-3) / 1e9 : 0.0;
}

}  // namespace smoke
