// Device-side building blocks of the smoke step: grid descriptor, MAC
// indexing, boundary (ghost) rules, linear / monotone-cubic samplers and the
// RK2 backtrace.  Semantics follow backend-cuda.cu ("B") line for line where
// cited; the implementation is organised for gfx950 (flat descriptor arrays
// indexed by axis, an all-in-range fast path for the 64-tap cubic gathers).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "visual-simulation-of-smoke-3d.h"

namespace smoke {

// Workgroups of a launch are dispatched round-robin over the 8 XCDs, each
// with its own L2.  A kernel over a grid of tiles (numbered x fastest, then y,
// then z, as a 3-D launch numbers its blocks) launched as xcd_blocks(tiles)
// one-dimensional blocks takes its tile from xcd_tile: XCD k walks the
// contiguous tiles [k ceil(T/8), (k+1) ceil(T/8)), so the halo rows a tile
// reads from its neighbours were mostly brought into the same L2 by them.
// Blocks past the last tile return at once (the whole block: the test
// depends on blockIdx only).
constexpr unsigned kXcds = 8;

inline unsigned xcd_blocks(const dim3& tiles) { return (tiles.x * tiles.y * tiles.z + kXcds - 1) / kXcds * kXcds; }

// Block index of a kernel body: the hardware's in an ordinary launch, a
// grid's own numbering inside a batched launch over many grids (batched.hpp).
struct VBlock {
    uint3 idx;   // block index in the grid's own launch
    uint3 dim;   // the grid's own launch dimensions
};

__device__ __forceinline__ VBlock hw_block() {
    VBlock b;
    b.idx = make_uint3(blockIdx.x, blockIdx.y, blockIdx.z);
    b.dim = make_uint3(gridDim.x, gridDim.y, gridDim.z);
    return b;
}

__device__ __forceinline__ bool xcd_tile_of(unsigned block, const uint3& tiles, uint3& tile, unsigned& id) {
    const unsigned total = tiles.x * tiles.y * tiles.z;
    const unsigned per = (total + kXcds - 1) / kXcds;
    id = (block % kXcds) * per + block / kXcds;
    if (id >= total) return false;
    tile.x = id % tiles.x;
    const unsigned rest = id / tiles.x;
    tile.y = rest % tiles.y;
    tile.z = rest / tiles.y;
    return true;
}

__device__ __forceinline__ bool xcd_tile(const uint3& tiles, uint3& tile, unsigned& id) { return xcd_tile_of(blockIdx.x, tiles, tile, id); }
__device__ __forceinline__ bool xcd_tile(const VBlock& vb, const uint3& tiles, uint3& tile, unsigned& id) {
    return xcd_tile_of(vb.idx.x, tiles, tile, id);
}

// Boundary kinds (values of the ABI enums).
enum : uint32_t { kNoSlip = 0, kFreeSlip = 1, kOutflow = 2, kPeriodic = 3 };
enum : uint32_t { kFixedValue = 0, kZeroFlux = 1, kScalarPeriodic = 2 };

struct ScalarBC {
    uint32_t type[6];   // x-, x+, y-, y+, z-, z+
    float value[6];
    int periodic[3];
};

// Flat, axis-indexed copy of SmokeSimulationConfig passed by value to kernels.
struct Grid {
    int n[3];
    float h;
    uint32_t flow_type[6];
    float flow_velocity[6][3];
    float flow_pressure[6];
    int flow_periodic[3];
    ScalarBC density;
    ScalarBC temperature;
    int cubic;
    int64_t cells;
    int64_t faces[3];
};

__host__ inline Grid make_grid(const SmokeSimulationConfig& c) {
    Grid g{};
    g.n[0] = c.nx;
    g.n[1] = c.ny;
    g.n[2] = c.nz;
    g.h = c.cell_size;
    const SmokeSimulationFlowBoundaryFaceDesc* ff[6] = {&c.flow_boundary.x_minus, &c.flow_boundary.x_plus, &c.flow_boundary.y_minus,
                                                        &c.flow_boundary.y_plus, &c.flow_boundary.z_minus, &c.flow_boundary.z_plus};
    for (int f = 0; f < 6; ++f) {
        g.flow_type[f] = ff[f]->type;
        g.flow_velocity[f][0] = ff[f]->velocity_x;
        g.flow_velocity[f][1] = ff[f]->velocity_y;
        g.flow_velocity[f][2] = ff[f]->velocity_z;
        g.flow_pressure[f] = ff[f]->pressure;
    }
    for (int a = 0; a < 3; ++a) g.flow_periodic[a] = g.flow_type[2 * a] == kPeriodic && g.flow_type[2 * a + 1] == kPeriodic;
    auto scal = [](const SmokeSimulationScalarBoundaryConfig& s) {
        ScalarBC b{};
        const SmokeSimulationScalarBoundaryFaceDesc* sf[6] = {&s.x_minus, &s.x_plus, &s.y_minus, &s.y_plus, &s.z_minus, &s.z_plus};
        for (int f = 0; f < 6; ++f) {
            b.type[f] = sf[f]->type;
            b.value[f] = sf[f]->value;
        }
        for (int a = 0; a < 3; ++a) b.periodic[a] = b.type[2 * a] == kScalarPeriodic && b.type[2 * a + 1] == kScalarPeriodic;
        return b;
    };
    g.density = scal(c.density_boundary);
    g.temperature = scal(c.temperature_boundary);
    g.cubic = c.scalar_advection_mode == SMOKE_SIMULATION_SCALAR_ADVECTION_MONOTONIC_CUBIC;
    g.cells = int64_t(c.nx) * c.ny * c.nz;
    for (int k = 0; k < 3; ++k) g.faces[k] = int64_t(c.nx + (k == 0)) * (c.ny + (k == 1)) * (c.nz + (k == 2));
    return g;
}

// B:121-135 -- x-fastest SoA cells and staggered faces.
__device__ __forceinline__ int64_t cidx(const Grid& g, int x, int y, int z) { return (int64_t(z) * g.n[1] + y) * g.n[0] + x; }
__device__ __forceinline__ int64_t fidx(const Grid& g, int c, int i, int j, int k) {
    const int d0 = g.n[0] + (c == 0), d1 = g.n[1] + (c == 1);
    return (int64_t(k) * d1 + j) * d0 + i;
}

// B:137-142
__device__ __forceinline__ int wrapi(int v, int n) {
    v %= n;
    return v < 0 ? v + n : v;
}
__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// B:148-153: wrap flow-periodic axes, then the bounds test.
__device__ __forceinline__ bool resolve_flow(const Grid& g, int c[3]) {
#pragma unroll
    for (int a = 0; a < 3; ++a)
        if (g.flow_periodic[a]) c[a] = wrapi(c[a], g.n[a]);
    return c[0] >= 0 && c[0] < g.n[0] && c[1] >= 0 && c[1] < g.n[1] && c[2] >= 0 && c[2] < g.n[2];
}

__device__ __forceinline__ bool resolve_scalar(const Grid& g, const ScalarBC& bc, int c[3]) {
#pragma unroll
    for (int a = 0; a < 3; ++a)
        if (bc.periodic[a]) c[a] = wrapi(c[a], g.n[a]);
    return c[0] >= 0 && c[0] < g.n[0] && c[1] >= 0 && c[1] < g.n[1] && c[2] >= 0 && c[2] < g.n[2];
}

// Kernels take the Grid by value (it lives in the kernel-argument segment and
// is read with scalar loads).  A runtime index into one of its arrays would
// make the compiler copy the whole struct to per-thread scratch memory, so
// per-face entries are picked with selects: `a` is a compile-time index after
// unrolling, `side` and `comp` only choose between values.
template <typename T>
__device__ __forceinline__ T face_entry(const T (&v)[6], int a, int side) {
    return side ? v[2 * a + 1] : v[2 * a];
}

__device__ __forceinline__ float pick3(const float (&v)[3], int comp) { return comp == 0 ? v[0] : (comp == 1 ? v[1] : v[2]); }

__device__ __forceinline__ int pick3(const int (&v)[3], int comp) { return comp == 0 ? v[0] : (comp == 1 ? v[1] : v[2]); }

__device__ __forceinline__ float face_velocity(const Grid& g, int a, int side, int comp) {
    return side ? pick3(g.flow_velocity[2 * a + 1], comp) : pick3(g.flow_velocity[2 * a], comp);
}

// B:162-166 -- out of domain counts as solid.
__device__ __forceinline__ bool occupied(const Grid& g, const uint8_t* occ, int x, int y, int z) {
    int c[3] = {x, y, z};
    if (!resolve_flow(g, c)) return true;
    return occ[cidx(g, c[0], c[1], c[2])] != 0;
}

// B:168-200 -- scalar ghost value.
__device__ inline float load_scalar(const Grid& g, const ScalarBC& bc, const float* __restrict__ f, int x, int y, int z) {
    int c[3] = {x, y, z};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (c[a] >= 0 && c[a] < g.n[a]) continue;
        const int side = c[a] < 0 ? 0 : 1;
        if (bc.periodic[a]) c[a] = wrapi(c[a], g.n[a]);
        else if (face_entry(bc.type, a, side) == kZeroFlux) c[a] = side ? g.n[a] - 1 : 0;
        else return face_entry(bc.value, a, side);
    }
    return f[cidx(g, c[0], c[1], c[2])];
}

// B:202-225 -- wrap or clamp.
__device__ inline float load_flow_cell(const Grid& g, const float* __restrict__ f, int x, int y, int z) {
    int c[3] = {x, y, z};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (c[a] >= 0 && c[a] < g.n[a]) continue;
        c[a] = g.flow_periodic[a] ? wrapi(c[a], g.n[a]) : (c[a] < 0 ? 0 : g.n[a] - 1);
    }
    return f[cidx(g, c[0], c[1], c[2])];
}

// Cell-centred velocity component `comp` (face average, B:859-868) at an
// in-domain cell.
__device__ __forceinline__ float center_velocity(const Grid& g, const float* __restrict__ f, int comp, int x, int y, int z) {
    int hi[3] = {x, y, z};
    hi[comp] += 1;
    return 0.5f * (f[fidx(g, comp, x, y, z)] + f[fidx(g, comp, hi[0], hi[1], hi[2])]);
}

// B:227-274 -- ghost rule for the cell-centred velocity, evaluated from the
// face array directly.
__device__ inline float load_center_component(const Grid& g, const float* __restrict__ f, int comp, int x, int y, int z) {
    int c[3] = {x, y, z};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (c[a] >= 0 && c[a] < g.n[a]) continue;
        const int side = c[a] < 0 ? 0 : 1;
        if (g.flow_periodic[a]) {
            c[a] = wrapi(c[a], g.n[a]);
            continue;
        }
        int cl[3];
#pragma unroll
        for (int b = 0; b < 3; ++b) cl[b] = b == a ? (side ? g.n[a] - 1 : 0) : clampi(c[b], 0, g.n[b] - 1);
        const float interior = center_velocity(g, f, comp, cl[0], cl[1], cl[2]);
        const uint32_t t = face_entry(g.flow_type, a, side);
        if (t == kOutflow) return interior;
        if (t == kFreeSlip && comp != a) return interior;
        return 2.0f * face_velocity(g, a, side, comp) - interior;
    }
    return center_velocity(g, f, comp, c[0], c[1], c[2]);
}

// B:276-376 -- MAC face ghost value for component `comp`.
__device__ inline float load_face(const Grid& g, const float* __restrict__ f, int comp, int i, int j, int k) {
    int c[3] = {i, j, k};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int hi = g.n[a] - (a == comp ? 0 : 1);
        if (c[a] >= 0 && c[a] <= hi) continue;
        const int side = c[a] < 0 ? 0 : 1;
        if (g.flow_periodic[a]) {
            c[a] = wrapi(c[a], g.n[a]);
            continue;
        }
        int cl[3];
#pragma unroll
        for (int b = 0; b < 3; ++b) cl[b] = b == a ? (side ? hi : 0) : clampi(c[b], 0, g.n[b] - (b == comp ? 0 : 1));
        const float interior = f[fidx(g, comp, cl[0], cl[1], cl[2])];
        const uint32_t t = face_entry(g.flow_type, a, side);
        if (a == comp) {
            if (t == kOutflow) return interior;
        } else if (t == kOutflow || t == kFreeSlip) {
            return interior;
        }
        return 2.0f * face_velocity(g, a, side, comp) - interior;
    }
    return f[fidx(g, comp, c[0], c[1], c[2])];
}

// B:378-392: the monotone cubic of p0..p3 at t in [0, 1] -- the Hermite
// basis h00 p1 + h10 m1 + h01 p2 + h11 m2 (evaluated as p1 + h01 (p2 - p1) +
// h10 m1 + h11 m2, h00 = 1 - h01) with slopes m1 = (p2 - p0) / 2,
// m2 = (p3 - p1) / 2, each zeroed when its sign differs from that of
// delta = p2 - p1 (or is 0), and both zeroed when |delta| < 1e-6.
// The basis of a fraction t is computed once (Hermite) and shared by every
// row sampled at that fraction; the slopes' factor 1/2 is carried by the
// basis (exact: a power of two), so a row costs its three differences, the
// limiter and four multiply-adds.  The limiter is one clamp per slope: the
// slope clamped to the interval between 0 and +-inf on delta's side (to 0
// when |delta| < 1e-6) -- the same slopes as the sign tests, except a slope
// whose product with delta underflows to 0, which the sign tests would zero
// and the clamp keeps (its contribution is then below 2^-126 |h10|).
struct Hermite {
    float h00, h10, h01, h11;   // h10 and h11 times 1/2
};

__device__ __forceinline__ Hermite hermite(float t) {
    const float t2 = t * t;
    const float t3 = t2 * t;
    Hermite w;
    // every multiply-add written out, so that the rounding does not depend
    // on how the compiler contracts the expression where it is inlined
    w.h00 = fmaf(2.0f, t3, fmaf(-3.0f, t2, 1.0f));
    w.h10 = 0.5f * fmaf(-2.0f, t2, t3 + t);
    w.h01 = fmaf(-2.0f, t3, 3.0f * t2);
    w.h11 = 0.5f * (t3 - t2);
    return w;
}

__device__ __forceinline__ float monotonic_cubic(float p0, float p1, float p2, float p3, const Hermite& w) {
    const float delta = p2 - p1;
    const float d1 = p2 - p0, d2 = p3 - p1;   // twice the slopes
    const float side = fabsf(delta) < 1.0e-6f ? 0.0f : copysignf(__builtin_inff(), delta);
    const float s1 = __builtin_amdgcn_fmed3f(d1, 0.0f, side);
    const float s2 = __builtin_amdgcn_fmed3f(d2, 0.0f, side);
    return fmaf(w.h11, s2, fmaf(w.h10, s1, fmaf(w.h01, delta, p1)));
}

__device__ __forceinline__ float monotonic_cubic(float p0, float p1, float p2, float p3, float t) {
    return monotonic_cubic(p0, p1, p2, p3, hermite(t));
}

// fmodf(p, extent) then + extent when negative (the periodic wrap of the
// reference's samplers and hit test), with one subtraction or addition when
// |p| < 2 extent, which gives the same value: for extent <= p < 2 extent,
// p - extent is exact (Sterbenz) and equals fmodf's exact result; for
// -extent < p < 0, fmodf returns p itself.
__device__ __forceinline__ float wrap_position(float p, float extent) {
    if (p >= 0.0f && p < extent) return p;
    if (p >= extent && p < 2.0f * extent) return p - extent;
    if (p < 0.0f && p > -extent) return p + extent;
    p = fmodf(p, extent);
    return p < 0.0f ? p + extent : p;
}

// Sampling position -> base index and fraction on each axis.  Positions are
// in cell units (world position / h: the face of component C with index i
// lies at i on axis C and at j + 1/2 on the others).  `comp` is the staggered
// axis of a face component, -1 for cell-centred data; `periodic` selects the
// fmod wrap (B:394-410 and the other samplers' preambles, there in world
// units: the wrap extent is n cells).
struct Stencil {
    int b[3];
    float t[3];
};

// WRAP = false compiles the periodic wrap out (for callers that know no axis
// of `periodic` is set).
template <bool WRAP = true>
__device__ __forceinline__ Stencil locate(const Grid& g, const int periodic[3], int comp, float x, float y, float z) {
    float p[3] = {x, y, z};
    Stencil s;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (WRAP && periodic[a]) {
            p[a] = wrap_position(p[a], float(g.n[a]));
        }
        // Beyond two cells outside the domain every tap is the same ghost
        // value, so clamping there changes no result; it keeps b +- 2 (and a
        // non-finite position) inside int range for the bounds tests below.
        const float gc = __builtin_amdgcn_fmed3f(a == comp ? p[a] : p[a] - 0.5f, -4.0f, float(g.n[a]) + 4.0f);
        s.b[a] = int(floorf(gc));
        s.t[a] = gc - float(s.b[a]);
    }
    return s;
}

// Generic tap loader for a face component or a scalar.
template <int COMP>
__device__ __forceinline__ float tap(const Grid& g, const ScalarBC& bc, const float* __restrict__ f, int i, int j, int k) {
    if constexpr (COMP < 0) return load_scalar(g, bc, f, i, j, k);
    else return load_face(g, f, COMP, i, j, k);
}

template <int COMP>
__device__ __forceinline__ bool stencil_inside(const Grid& g, const Stencil& s, int lo, int hi) {
    bool ok = true;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int top = g.n[a] - ((a == COMP) ? 0 : 1);
        ok = ok && (s.b[a] + lo >= 0) && (s.b[a] + hi <= top);
    }
    return ok;
}

// Trilinear combination of the 8 taps c[(dz << 2) | (dy << 1) | dx], lerp
// order x, y, z (B:394-440), each lerp one written-out multiply-add.
__device__ __forceinline__ float lerp1(float a, float b, float t) { return fmaf(b - a, t, a); }

__device__ __forceinline__ float trilinear(const float (&c)[8], const Stencil& s) {
    const float c00 = lerp1(c[0], c[1], s.t[0]);
    const float c10 = lerp1(c[2], c[3], s.t[0]);
    const float c01 = lerp1(c[4], c[5], s.t[0]);
    const float c11 = lerp1(c[6], c[7], s.t[0]);
    const float c0 = lerp1(c00, c10, s.t[1]);
    const float c1 = lerp1(c01, c11, s.t[1]);
    return lerp1(c0, c1, s.t[2]);
}

// B:394-440 / 486-530 etc. -- trilinear, lerp order x, y, z.
template <int COMP>
__device__ inline float sample_linear(const Grid& g, const ScalarBC& bc, const float* __restrict__ f, const Stencil& s) {
    float c[8];
    if (stencil_inside<COMP>(g, s, 0, 1)) {
        const int d0 = g.n[0] + (COMP == 0), d1 = g.n[1] + (COMP == 1);
        const int64_t base = (int64_t(s.b[2]) * d1 + s.b[1]) * d0 + s.b[0];
        const int64_t sy = d0, sz = int64_t(d0) * d1;
        c[0] = f[base];
        c[1] = f[base + 1];
        c[2] = f[base + sy];
        c[3] = f[base + sy + 1];
        c[4] = f[base + sz];
        c[5] = f[base + sz + 1];
        c[6] = f[base + sz + sy];
        c[7] = f[base + sz + sy + 1];
    } else {
#pragma unroll
        for (int q = 0; q < 8; ++q) c[q] = tap<COMP>(g, bc, f, s.b[0] + (q & 1), s.b[1] + ((q >> 1) & 1), s.b[2] + (q >> 2));
    }
    return trilinear(c, s);
}

// B:442-484 / 532-573 etc. -- 4x4x4 taps, monotone cubic along x, then y, then z.
template <int COMP>
__device__ inline float sample_cubic(const Grid& g, const ScalarBC& bc, const float* __restrict__ f, const Stencil& s) {
    float zs[4];
    const Hermite wx = hermite(s.t[0]), wy = hermite(s.t[1]), wz = hermite(s.t[2]);
    if (stencil_inside<COMP>(g, s, -1, 2)) {
        const int d0 = g.n[0] + (COMP == 0), d1 = g.n[1] + (COMP == 1);
        const int64_t sy = d0, sz = int64_t(d0) * d1;
        const int64_t base = (int64_t(s.b[2] - 1) * d1 + (s.b[1] - 1)) * d0 + (s.b[0] - 1);
#pragma unroll
        for (int dz = 0; dz < 4; ++dz) {
            float ys[4];
#pragma unroll
            for (int dy = 0; dy < 4; ++dy) {
                const float* row = f + base + dz * sz + dy * sy;
                ys[dy] = monotonic_cubic(row[0], row[1], row[2], row[3], wx);
            }
            zs[dz] = monotonic_cubic(ys[0], ys[1], ys[2], ys[3], wy);
        }
    } else {
        for (int dz = 0; dz < 4; ++dz) {
            float ys[4];
            for (int dy = 0; dy < 4; ++dy) {
                const int yy = s.b[1] + dy - 1, zz = s.b[2] + dz - 1;
                const float p0 = tap<COMP>(g, bc, f, s.b[0] - 1, yy, zz);
                const float p1 = tap<COMP>(g, bc, f, s.b[0], yy, zz);
                const float p2 = tap<COMP>(g, bc, f, s.b[0] + 1, yy, zz);
                const float p3 = tap<COMP>(g, bc, f, s.b[0] + 2, yy, zz);
                ys[dy] = monotonic_cubic(p0, p1, p2, p3, wx);
            }
            zs[dz] = monotonic_cubic(ys[0], ys[1], ys[2], ys[3], wy);
        }
    }
    return monotonic_cubic(zs[0], zs[1], zs[2], zs[3], wz);
}

template <int COMP>
__device__ __forceinline__ float sample_face(const Grid& g, const float* __restrict__ f, float x, float y, float z, bool cubic) {
    const Stencil s = locate(g, g.flow_periodic, COMP, x, y, z);
    return cubic ? sample_cubic<COMP>(g, g.density, f, s) : sample_linear<COMP>(g, g.density, f, s);
}

__device__ __forceinline__ float sample_scalar(const Grid& g, const ScalarBC& bc, const float* __restrict__ f, float x, float y, float z, bool cubic) {
    const Stencil s = locate(g, bc.periodic, -1, x, y, z);
    return cubic ? sample_cubic<-1>(g, bc, f, s) : sample_linear<-1>(g, bc, f, s);
}

// Two cell-centred fields sampled at one point in one pass over the cubic
// stencil (temperature and density share the trace): each field's value is
// computed exactly as by sample_scalar; the interior fast path loads both
// fields' rows together and shares the index arithmetic.
__device__ inline void sample_scalar2(const Grid& g, const ScalarBC& bca, const float* __restrict__ fa, const ScalarBC& bcb,
                                      const float* __restrict__ fb, float x, float y, float z, bool cubic, float& va, float& vb) {
    // the stencil depends on the fields' periodic axes only (the same for
    // both in every configuration the benchmark and tests use)
    const bool same = bca.periodic[0] == bcb.periodic[0] && bca.periodic[1] == bcb.periodic[1] && bca.periodic[2] == bcb.periodic[2];
    const Stencil sa = locate(g, bca.periodic, -1, x, y, z);
    if (!cubic || !same || !stencil_inside<-1>(g, sa, -1, 2)) {
        va = sample_scalar(g, bca, fa, x, y, z, cubic);
        vb = sample_scalar(g, bcb, fb, x, y, z, cubic);
        return;
    }
    const int64_t sy = g.n[0], sz = int64_t(g.n[0]) * g.n[1];
    const int64_t base = (int64_t(sa.b[2] - 1) * g.n[1] + (sa.b[1] - 1)) * g.n[0] + (sa.b[0] - 1);
    const Hermite wx = hermite(sa.t[0]), wy = hermite(sa.t[1]), wz = hermite(sa.t[2]);
    float za[4], zb[4];
#pragma unroll
    for (int dz = 0; dz < 4; ++dz) {
        float ya[4], yb[4];
#pragma unroll
        for (int dy = 0; dy < 4; ++dy) {
            const int64_t o = base + dz * sz + dy * sy;
            const float a0 = fa[o], a1 = fa[o + 1], a2 = fa[o + 2], a3 = fa[o + 3];
            const float b0 = fb[o], b1 = fb[o + 1], b2 = fb[o + 2], b3 = fb[o + 3];
            ya[dy] = monotonic_cubic(a0, a1, a2, a3, wx);
            yb[dy] = monotonic_cubic(b0, b1, b2, b3, wx);
        }
        za[dz] = monotonic_cubic(ya[0], ya[1], ya[2], ya[3], wy);
        zb[dz] = monotonic_cubic(yb[0], yb[1], yb[2], yb[3], wy);
    }
    va = monotonic_cubic(za[0], za[1], za[2], za[3], wz);
    vb = monotonic_cubic(zb[0], zb[1], zb[2], zb[3], wz);
}

// Hit test of B:785-794 / 819-828, position in cell units.
__device__ __forceinline__ bool hits_solid(const Grid& g, const uint8_t* __restrict__ occ, float qx, float qy, float qz) {
    float q[3] = {qx, qy, qz};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (g.flow_periodic[a]) q[a] = wrap_position(q[a], float(g.n[a]));
    }
#pragma unroll
    for (int a = 0; a < 3; ++a)
        if (q[a] < 0.0f || q[a] > float(g.n[a])) return true;
    int c[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        c[a] = int(floorf(q[a]));
        if (c[a] == g.n[a]) c[a] = g.n[a] - 1;
    }
    if (c[0] < 0 || c[0] >= g.n[0] || c[1] < 0 || c[1] >= g.n[1] || c[2] < 0 || c[2] >= g.n[2]) return true;
    return occ[cidx(g, c[0], c[1], c[2])] != 0;
}

// The linear velocity samples at the start of a trace.  A trace starts at a
// cell centre (scalars) or at a face (velocity), where every sample fraction
// is 0 or 1/2: the trilinear form then reduces to one or two half-way lerps of
// taps next to the start, bit for bit (a lerp at fraction 0 returns its first
// value; only a zero's sign can differ, and the positions built from the
// samples cannot), so these read 2 (cell centre) or 4 taps instead of 8.
// `tap(comp, i, j, k)` returns a face value with the ghost rules.
__device__ __forceinline__ float half_lerp(float a, float b) { return lerp1(a, b, 0.5f); }

// component D at the start of a trace from face (i, j, k) of component C
// (D != C): fraction 1/2 on axis C (taps idx_C - 1, idx_C), 1/2 on axis D
// (taps idx_D, idx_D + 1), 0 on the third axis; lerps in x, y, z order
template <int C, int D, class Tap>
__device__ __forceinline__ float face_start_component(const Tap& tap, int i, int j, int k) {
    constexpr int A1 = C < D ? C : D, A2 = C < D ? D : C;
    auto at = [&](int o1, int o2) {
        int c[3] = {i, j, k};
        c[A1] += o1;
        c[A2] += o2;
        return tap(D, c[0], c[1], c[2]);
    };
    constexpr int lo1 = A1 == C ? -1 : 0, lo2 = A2 == C ? -1 : 0;   // lower tap on each axis
    return half_lerp(half_lerp(at(lo1, lo2), at(lo1 + 1, lo2)), half_lerp(at(lo1, lo2 + 1), at(lo1 + 1, lo2 + 1)));
}

// (v0x, v0y, v0z) at the start of a trace from face (i, j, k) of component C,
// or (C = -1) from the centre of cell (i, j, k)
template <int C, class Tap>
__device__ __forceinline__ float3 start_velocity(const Tap& tap, int i, int j, int k) {
    if constexpr (C < 0) {
        return make_float3(half_lerp(tap(0, i, j, k), tap(0, i + 1, j, k)), half_lerp(tap(1, i, j, k), tap(1, i, j + 1, k)),
                           half_lerp(tap(2, i, j, k), tap(2, i, j, k + 1)));
    } else {
        float v[3];
        v[C] = tap(C, i, j, k);
        if constexpr (C != 0) v[0] = face_start_component<C, 0>(tap, i, j, k);
        if constexpr (C != 1) v[1] = face_start_component<C, 1>(tap, i, j, k);
        if constexpr (C != 2) v[2] = face_start_component<C, 2>(tap, i, j, k);
        return make_float3(v[0], v[1], v[2]);
    }
}

// B:756-838 -- RK2 midpoint backtrace with linear velocity samples, bisection
// clipping against walls and solids; the result is not wrapped.  In cell
// units: a velocity v moves dt v / h cells, so `s` = dt / h scales the
// samples (the reference divides each world position by h instead).  This
// is the rest of the trace from its start (sx, sy, sz) and start velocity v0.
__device__ inline float3 trace_rk2_from(const Grid& g, const float* __restrict__ u, const float* __restrict__ v, const float* __restrict__ w,
                                        const uint8_t* __restrict__ occ, float s, float sx, float sy, float sz, float3 v0) {
    const float hs = -0.5f * s;
    const float mx = fmaf(hs, v0.x, sx), my = fmaf(hs, v0.y, sy), mz = fmaf(hs, v0.z, sz);
    const float v1x = sample_face<0>(g, u, mx, my, mz, false);
    const float v1y = sample_face<1>(g, v, mx, my, mz, false);
    const float v1z = sample_face<2>(g, w, mx, my, mz, false);
    const float tx = fmaf(-s, v1x, sx), ty = fmaf(-s, v1y, sy), tz = fmaf(-s, v1z, sz);
    if (!hits_solid(g, occ, tx, ty, tz)) return make_float3(tx, ty, tz);
    float lo = 0.0f, hi = 1.0f;
    for (int it = 0; it < 10; ++it) {
        const float mt = 0.5f * (lo + hi);
        if (hits_solid(g, occ, fmaf(tx - sx, mt, sx), fmaf(ty - sy, mt, sy), fmaf(tz - sz, mt, sz))) hi = mt;
        else lo = mt;
    }
    return make_float3(fmaf(tx - sx, lo, sx), fmaf(ty - sy, lo, sy), fmaf(tz - sz, lo, sz));
}

// RK2 trace from face (i, j, k) of component C (C = -1: from the centre of
// cell (i, j, k)), ghost-rule taps from the global arrays
template <int C>
__device__ inline float3 trace_rk2_at(const Grid& g, const float* __restrict__ u, const float* __restrict__ v, const float* __restrict__ w,
                                      const uint8_t* __restrict__ occ, float s, int i, int j, int k) {
    auto tap = [&](int comp, int a, int b, int c) { return load_face(g, comp == 0 ? u : (comp == 1 ? v : w), comp, a, b, c); };
    const float3 v0 = start_velocity<C>(tap, i, j, k);
    const float sx = C == 0 ? float(i) : float(i) + 0.5f;
    const float sy = C == 1 ? float(j) : float(j) + 0.5f;
    const float sz = C == 2 ? float(k) : float(k) + 0.5f;
    return trace_rk2_from(g, u, v, w, occ, s, sx, sy, sz, v0);
}

// B:1007-1012 with the face-neighbour logic of B:282-304: mean solid velocity
// of the solid cells beside face (i,j,k) of component `comp`.
__device__ inline bool solid_face_override(const Grid& g, const uint8_t* __restrict__ occ, const float* __restrict__ solid_vel, int comp, int i, int j,
                                           int k, float& value) {
    int left[3] = {i, j, k};
    int right[3] = {i, j, k};
    left[comp] -= 1;
    const bool hl = resolve_flow(g, left);
    const bool hr = resolve_flow(g, right);
    const int64_t il = hl ? cidx(g, left[0], left[1], left[2]) : 0;
    const int64_t ir = hr ? cidx(g, right[0], right[1], right[2]) : 0;
    const bool lo = hl && occ[il] != 0;
    const bool ro = hr && occ[ir] != 0;
    if (!lo && !ro) return false;
    float sum = 0.0f, weight = 0.0f;
    if (lo) {
        sum += solid_vel[il];
        weight += 1.0f;
    }
    if (ro) {
        sum += solid_vel[ir];
        weight += 1.0f;
    }
    value = sum / weight;
    return true;
}

// Face (i,j,k) of component comp is a domain face handled by the boundary
// pass: a non-periodic end face, or the duplicate end face n of a periodic axis.
__device__ __forceinline__ bool boundary_face(const Grid& g, int comp, int idx) {
    if (idx == 0) return !pick3(g.flow_periodic, comp);
    return idx == pick3(g.n, comp);
}

}  // namespace smoke
