// Multilevel preconditioner of the pressure solve: one V-cycle of a
// geometric hierarchy with damped-Jacobi smoothing, whose coarsest level is
// solved by the learned network (or by Jacobi sweeps).  No counterpart in the
// reference, whose solve is unpreconditioned CG (backend-cuda.cu:1788-1824);
// every level applies the reference's 7-point matrix (B:1343-1412)
// rediscretised on that level's occupancy, from the same flag byte the PCG
// kernels use.
//
// Hierarchy: level l+1 has ceil(n/2) cells per axis; a coarse cell is solid
// when every existing child is solid, and the cell holding the finer level's
// anchor is the coarse anchor (an identity row), so each level is a
// nonsingular pinned Poisson problem like the pressure system itself.
// Transfers are piecewise constant: r_(l+1) = 1/2 (sum of the children's
// residuals) -- the 2^3 average times the (2h/h)^2 = 4 of the rediscretised
// operator -- and the prolongation adds the parent's correction to each fluid
// child.
//
// Kernels are one thread per cell (per coarse cell for the restriction)
// reading their neighbours through the cache.  The first sweep from zero
// (x = omega D^-1 r) is never stored: the next kernel recomputes it at every
// tap from r, so two pre-sweeps cost one pass, and the prolongation is folded
// into the first post-sweep the same way.
#include "context.hpp"
#include "nn.hpp"

#include <algorithm>
#include <cstring>

namespace smoke {
namespace {

constexpr uint8_t kIdentity = 0x40;

struct LevelArgs {
    int n[3];
    int periodic[3];
    int outflow[6];
    int any_outflow;
    const uint8_t* flags;
    float omega;
};

__device__ __forceinline__ int outflow_faces(const LevelArgs& a, int x, int y, int z) {
    if (!a.any_outflow) return 0;
    return (x == 0 && a.outflow[0]) + (x == a.n[0] - 1 && a.outflow[1]) + (y == 0 && a.outflow[2]) + (y == a.n[1] - 1 && a.outflow[3]) +
           (z == 0 && a.outflow[4]) + (z == a.n[2] - 1 && a.outflow[5]);
}

// diagonal of a fluid (non-identity) row: fluid neighbours + OUTFLOW faces, at least 1
__device__ __forceinline__ float row_diag(const LevelArgs& a, unsigned f, int x, int y, int z) {
    const int d = __popc(f & 0x3fu) + outflow_faces(a, x, y, z);
    return float(d > 1 ? d : 1);
}

__device__ __forceinline__ int64_t lin(const LevelArgs& a, int x, int y, int z) { return (int64_t(z) * a.n[1] + y) * a.n[0] + x; }

// neighbour in direction d (0..5: x-, x+, y-, y+, z-, z+); only called for
// set flag bits, so it exists (periodic axes wrap by one period)
__device__ __forceinline__ void neighbour(const LevelArgs& a, int d, int& x, int& y, int& z) {
    const int ax = d >> 1, s = (d & 1) ? 1 : -1;
    int* c[3] = {&x, &y, &z};
    int v = *c[ax] + s;
    if (v < 0) v += a.n[ax];
    else if (v >= a.n[ax]) v -= a.n[ax];
    *c[ax] = v;
}

// What the sweep and restriction kernels take as the current iterate x:
enum XIn {
    kZero = 0,        // x = 0
    kFromR = 1,       // x = omega D^-1 r (one sweep from zero, recomputed per tap)
    kBuffer = 2,      // x = buffer
    kBufferPlusC = 3, // x = buffer + coarse correction of the parent cell
    kFromRPlusC = 4,  // x = omega D^-1 r + coarse correction of the parent cell
};

struct Coarse {
    const float* e;   // correction on the next level
    int n0, n1;       // its x and y extent
};

template <int XIN>
__device__ __forceinline__ float x_at(const LevelArgs& a, const float* __restrict__ r, const float* __restrict__ xb, const Coarse& cc, int x, int y,
                                      int z, int64_t i, unsigned f) {
    if (XIN == kZero || (f & kIdentity)) return 0.0f;
    float v = 0.0f;
    if constexpr (XIN == kFromR || XIN == kFromRPlusC) v = a.omega * r[i] / row_diag(a, f, x, y, z);
    if constexpr (XIN == kBuffer || XIN == kBufferPlusC) v = xb[i];
    if constexpr (XIN == kBufferPlusC || XIN == kFromRPlusC) v += cc.e[(int64_t(z >> 1) * cc.n1 + (y >> 1)) * cc.n0 + (x >> 1)];
    return v;
}

// (r - A x)_i of fluid row i with flag byte f
template <int XIN>
__device__ __forceinline__ float residual_at(const LevelArgs& a, const float* __restrict__ r, const float* __restrict__ xb, const Coarse& cc, int x,
                                             int y, int z, int64_t i, unsigned f, float& xi, float& diag) {
    diag = row_diag(a, f, x, y, z);
    xi = x_at<XIN>(a, r, xb, cc, x, y, z, i, f);
    float ax = diag * xi;
    if constexpr (XIN != kZero) {
#pragma unroll
        for (int d = 0; d < 6; ++d) {
            if (!(f & (1u << d))) continue;
            int xx = x, yy = y, zz = z;
            neighbour(a, d, xx, yy, zz);
            const int64_t j = lin(a, xx, yy, zz);
            ax -= x_at<XIN>(a, r, xb, cc, xx, yy, zz, j, a.flags[j]);
        }
    }
    return r[i] - ax;
}

// out = x + omega D^-1 (r - A x) on fluid rows, 0 on identity rows
template <int XIN>
__global__ __launch_bounds__(256) void mg_sweep_kernel(LevelArgs a, const float* __restrict__ r, const float* __restrict__ xb, Coarse cc,
                                                       float* __restrict__ out) {
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int z = blockIdx.z;
    if (x >= a.n[0] || y >= a.n[1]) return;
    const int64_t i = lin(a, x, y, z);
    const unsigned f = a.flags[i];
    if (f & kIdentity) {
        out[i] = 0.0f;
        return;
    }
    float xi, diag;
    const float res = residual_at<XIN>(a, r, xb, cc, x, y, z, i, f, xi, diag);
    out[i] = xi + a.omega * res / diag;
}

// r_c = 1/2 * sum over the children of (r - A x); 0 on coarse identity rows
template <int XIN>
__global__ __launch_bounds__(256) void mg_restrict_kernel(LevelArgs a, const float* __restrict__ r, const float* __restrict__ xb, int nc0, int nc1,
                                                          int nc2, const uint8_t* __restrict__ cflags, float* __restrict__ rc) {
    const int X = blockIdx.x * 64 + (threadIdx.x & 63);
    const int Y = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int Z = blockIdx.z;
    if (X >= nc0 || Y >= nc1 || Z >= nc2) return;
    const int64_t ci = (int64_t(Z) * nc1 + Y) * nc0 + X;
    if (cflags[ci] & kIdentity) {
        rc[ci] = 0.0f;
        return;
    }
    const Coarse none{nullptr, 0, 0};
    float s = 0.0f;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const int x = 2 * X + (c & 1), y = 2 * Y + ((c >> 1) & 1), z = 2 * Z + (c >> 2);
        if (x >= a.n[0] || y >= a.n[1] || z >= a.n[2]) continue;
        const int64_t i = lin(a, x, y, z);
        const unsigned f = a.flags[i];
        if (f & kIdentity) continue;
        float xi, diag;
        s += residual_at<XIN>(a, r, xb, none, x, y, z, i, f, xi, diag);
    }
    rc[ci] = 0.5f * s;
}

// coarse occupancy: solid when every existing child is solid
__global__ void mg_coarsen_kernel(int f0, int f1, int f2, const uint8_t* __restrict__ occf, int c0, int c1, int c2, uint8_t* __restrict__ occc) {
    const int X = blockIdx.x * 64 + (threadIdx.x & 63);
    const int Y = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int Z = blockIdx.z;
    if (X >= c0 || Y >= c1 || Z >= c2) return;
    bool solid = true;
    for (int c = 0; c < 8; ++c) {
        const int x = 2 * X + (c & 1), y = 2 * Y + ((c >> 1) & 1), z = 2 * Z + (c >> 2);
        if (x >= f0 || y >= f1 || z >= f2) continue;
        solid = solid && occf[(int64_t(z) * f1 + y) * f0 + x] != 0;
    }
    occc[(int64_t(Z) * c1 + Y) * c0 + X] = solid ? 1 : 0;
}

// coarse anchor = parent of the fine anchor (fine "no fluid cell" = cells -> coarse cells)
__global__ void mg_anchor_kernel(const int* __restrict__ af, int f0, int f1, int f2, int c0, int c1, int c2, int* __restrict__ ac) {
    const int a = *af;
    if (a < 0 || int64_t(a) >= int64_t(f0) * f1 * f2) {
        *ac = c0 * c1 * c2;
        return;
    }
    const int x = a % f0, y = (a / f0) % f1, z = a / (f0 * f1);
    *ac = ((z >> 1) * c1 + (y >> 1)) * c0 + (x >> 1);
}

// flag byte of a level (same bits as the pressure grid's): fluid neighbours
// after periodic wrap, identity for solid cells and the anchor
__global__ void mg_flags_kernel(LevelArgs a, const uint8_t* __restrict__ occ, const int* __restrict__ anchor, uint8_t* __restrict__ flags) {
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int z = blockIdx.z;
    if (x >= a.n[0] || y >= a.n[1]) return;
    const int64_t i = lin(a, x, y, z);
    if (occ[i] != 0 || i == int64_t(*anchor)) {
        flags[i] = kIdentity;
        return;
    }
    unsigned m = 0;
    const int base[3] = {x, y, z};
    for (int d = 0; d < 6; ++d) {
        const int ax = d >> 1;
        int c[3] = {base[0], base[1], base[2]};
        c[ax] += (d & 1) ? 1 : -1;
        if (c[ax] < 0 || c[ax] >= a.n[ax]) {
            if (!a.periodic[ax]) continue;
            c[ax] = c[ax] < 0 ? c[ax] + a.n[ax] : c[ax] - a.n[ax];
        }
        if (occ[lin(a, c[0], c[1], c[2])] == 0) m |= 1u << d;
    }
    flags[i] = uint8_t(m);
}

dim3 cell_grid(const int n[3]) { return dim3(unsigned((n[0] + 63) / 64), unsigned((n[1] + 3) / 4), unsigned(n[2])); }

LevelArgs level_args(const Context& c, const MgLevel& L) {
    LevelArgs a{};
    for (int k = 0; k < 3; ++k) {
        a.n[k] = L.n[k];
        a.periodic[k] = c.grid.flow_periodic[k];
    }
    for (int f = 0; f < 6; ++f) a.outflow[f] = c.grid.flow_type[f] == kOutflow;
    a.any_outflow = a.outflow[0] | a.outflow[1] | a.outflow[2] | a.outflow[3] | a.outflow[4] | a.outflow[5];
    a.flags = L.flags;
    a.omega = c.mg.options.omega;
    return a;
}

template <class T>
void alloc_level(T*& p, int64_t count) {
    void* q = nullptr;
    hip_check(hipMalloc(&q, size_t(std::max<int64_t>(count, 1)) * sizeof(T)), "multilevel buffer");
    p = static_cast<T*>(q);
}

template <class T>
void free_level(T*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

void sweep(Context& c, const MgLevel& L, int xin, const float* r, const float* xb, Coarse cc, float* out) {
    const LevelArgs a = level_args(c, L);
    const dim3 g = cell_grid(L.n);
    switch (xin) {
        case kZero: mg_sweep_kernel<kZero><<<g, 256, 0, c.stream>>>(a, r, xb, cc, out); break;
        case kFromR: mg_sweep_kernel<kFromR><<<g, 256, 0, c.stream>>>(a, r, xb, cc, out); break;
        case kBuffer: mg_sweep_kernel<kBuffer><<<g, 256, 0, c.stream>>>(a, r, xb, cc, out); break;
        case kBufferPlusC: mg_sweep_kernel<kBufferPlusC><<<g, 256, 0, c.stream>>>(a, r, xb, cc, out); break;
        default: mg_sweep_kernel<kFromRPlusC><<<g, 256, 0, c.stream>>>(a, r, xb, cc, out); break;
    }
    hip_check(hipGetLastError(), "mg_sweep_kernel");
}

// `count` sweeps from zero on level L, ending in `out`; scratch ta / tb
void sweeps_from_zero(Context& c, const MgLevel& L, int count, const float* r, float* out) {
    const Coarse none{nullptr, 0, 0};
    if (count <= 1) {
        sweep(c, L, kZero, r, nullptr, none, out);
        return;
    }
    // the first two sweeps are one pass; the rest ping-pong through ta / tb
    float* bufs[2] = {L.ta, L.tb};
    int next = 0;
    const float* cur = nullptr;
    for (int k = 2; k <= count; ++k) {
        float* dst = k == count ? out : bufs[next];
        if (k == 2) sweep(c, L, kFromR, r, nullptr, none, dst);
        else sweep(c, L, kBuffer, r, cur, none, dst);
        cur = dst;
        next ^= 1;
    }
}

void vcycle(Context& c, int l) {
    Multilevel& m = c.mg;
    MgLevel& L = m.level[l];
    const Coarse none{nullptr, 0, 0};
    const auto& o = m.options;
    if (l == m.built - 1) {
        if (o.coarse_solver == SMOKE_SIMULATION_COARSE_NETWORK) {
            if (!c.learned) throw HipError("multilevel preconditioner: no network attached");
            preconditioner_apply(*c.learned, L.r, L.flags, L.n, L.z, c.stream);
            // optional sweeps after the network's correction
            float* bufs[2] = {L.ta, L.tb};
            const float* cur = L.z;
            for (int k = 1; k <= o.coarse_sweeps; ++k) {
                float* dst = bufs[(k - 1) & 1];
                sweep(c, L, kBuffer, L.r, cur, none, dst);
                cur = dst;
            }
            if (cur != L.z) hip_check(hipMemcpyAsync(L.z, cur, size_t(L.cells) * sizeof(float), hipMemcpyDeviceToDevice, c.stream), "coarse copy");
        } else {
            sweeps_from_zero(c, L, std::max(1, o.coarse_sweeps), L.r, L.z);
        }
        return;
    }
    MgLevel& C = m.level[l + 1];
    const LevelArgs a = level_args(c, L);
    // pre-smoothing: x on the fly (one sweep) or in ta
    const float* xb = nullptr;
    if (o.pre_sweeps >= 2) {
        float* bufs[2] = {L.ta, L.tb};
        for (int k = 2; k <= o.pre_sweeps; ++k) {
            float* dst = bufs[k & 1];
            if (k == 2) sweep(c, L, kFromR, L.r, nullptr, none, dst);
            else sweep(c, L, kBuffer, L.r, xb, none, dst);
            xb = dst;
        }
    }
    const dim3 gc = cell_grid(C.n);
    if (xb) mg_restrict_kernel<kBuffer><<<gc, 256, 0, c.stream>>>(a, L.r, xb, C.n[0], C.n[1], C.n[2], C.flags, C.r);
    else mg_restrict_kernel<kFromR><<<gc, 256, 0, c.stream>>>(a, L.r, nullptr, C.n[0], C.n[1], C.n[2], C.flags, C.r);
    hip_check(hipGetLastError(), "mg_restrict_kernel");
    if (l + 1 == m.built - 1 && m.capture && m.captured < m.capacity) {
        hip_check(hipMemcpyAsync(m.capture + size_t(m.captured) * size_t(C.cells), C.r, size_t(C.cells) * sizeof(float), hipMemcpyDeviceToDevice,
                                 c.stream),
                  "multilevel capture");
        ++m.captured;
    }
    vcycle(c, l + 1);
    // post-smoothing: the first sweep adds the prolonged correction on the fly
    const Coarse cc{C.z, C.n[0], C.n[1]};
    float* other = (xb == L.ta) ? L.tb : L.ta;   // the buffer not holding x
    float* first_out = o.post_sweeps == 1 ? L.z : other;
    sweep(c, L, xb ? kBufferPlusC : kFromRPlusC, L.r, xb, cc, first_out);
    const float* cur = first_out;
    float* bufs[2] = {L.ta, L.tb};
    for (int k = 2; k <= o.post_sweeps; ++k) {
        float* dst = k == o.post_sweeps ? L.z : (cur == bufs[0] ? bufs[1] : bufs[0]);
        sweep(c, L, kBuffer, L.r, cur, none, dst);
        cur = dst;
    }
}

}  // namespace

SmokeSimulationMultilevelOptions default_multilevel_options() {
    SmokeSimulationMultilevelOptions o{};
    o.levels = 4;
    o.pre_sweeps = 2;
    o.post_sweeps = 2;
    o.omega = 2.0f / 3.0f;
    o.coarse_solver = SMOKE_SIMULATION_COARSE_NETWORK;
    o.coarse_sweeps = 0;
    return o;
}

bool multilevel_options_valid(const Context& c, const SmokeSimulationMultilevelOptions& o) {
    if (o.levels < 2 || o.levels > kMaxLevels || o.pre_sweeps < 1 || o.post_sweeps < 1 || !(o.omega > 0.0f) || !(o.omega <= 1.0f)) return false;
    if (o.coarse_solver > 1u || o.coarse_sweeps < 0 || (o.coarse_solver == SMOKE_SIMULATION_COARSE_JACOBI && o.coarse_sweeps < 1)) return false;
    if (o.pre_sweeps > 64 || o.post_sweeps > 64 || o.coarse_sweeps > 4096) return false;
    for (int k = 0; k < 3; ++k) {
        // the network's two stride-2 levels need at least 4 cells per axis
        int m = c.grid.n[k];
        for (int l = 1; l < o.levels; ++l) m = (m + 1) / 2;
        if (o.coarse_solver == SMOKE_SIMULATION_COARSE_NETWORK && m < 4) return false;
        if (!c.grid.flow_periodic[k]) continue;
        int n = c.grid.n[k];
        for (int l = 1; l < o.levels; ++l) {
            if (n % 2 != 0) return false;
            n /= 2;
        }
        if (n < 3) return false;   // the flag byte cannot express the n <= 2 periodic couplings
    }
    return true;
}

void multilevel_release(Context& c) {
    Multilevel& m = c.mg;
    for (int l = 0; l < kMaxLevels; ++l) {
        MgLevel& L = m.level[l];
        if (l > 0) {
            free_level(L.occ);
            free_level(L.flags);
            free_level(L.anchor);
            free_level(L.r);
            free_level(L.z);
        }
        free_level(L.ta);
        free_level(L.tb);
        L = MgLevel{};
    }
    m.allocated = 0;
    m.built = 0;
}

void multilevel_build(Context& c) {
    Multilevel& m = c.mg;
    const int levels = m.options.levels;
    if (m.built == levels) return;
    if (m.allocated != levels) {
        multilevel_release(c);
        int n[3] = {c.grid.n[0], c.grid.n[1], c.grid.n[2]};
        for (int l = 0; l < levels; ++l) {
            MgLevel& L = m.level[l];
            for (int k = 0; k < 3; ++k) L.n[k] = n[k];
            L.cells = int64_t(n[0]) * n[1] * n[2];
            if (l > 0) {
                alloc_level(L.occ, L.cells);
                alloc_level(L.flags, L.cells);
                alloc_level(L.anchor, 1);
                alloc_level(L.r, L.cells);
                alloc_level(L.z, L.cells);
            }
            alloc_level(L.ta, L.cells);
            alloc_level(L.tb, L.cells);
            for (int k = 0; k < 3; ++k) n[k] = (n[k] + 1) / 2;
        }
        m.allocated = levels;
    }
    MgLevel& L0 = m.level[0];
    L0.occ = c.occupancy;
    L0.flags = c.stencil_flags;
    L0.anchor = c.anchor;
    for (int l = 1; l < levels; ++l) {
        const MgLevel& F = m.level[l - 1];
        MgLevel& L = m.level[l];
        const dim3 g = cell_grid(L.n);
        mg_coarsen_kernel<<<g, 256, 0, c.stream>>>(F.n[0], F.n[1], F.n[2], F.occ, L.n[0], L.n[1], L.n[2], L.occ);
        mg_anchor_kernel<<<1, 1, 0, c.stream>>>(F.anchor, F.n[0], F.n[1], F.n[2], L.n[0], L.n[1], L.n[2], L.anchor);
        mg_flags_kernel<<<g, 256, 0, c.stream>>>(level_args(c, L), L.occ, L.anchor, L.flags);
        hip_check(hipGetLastError(), "multilevel hierarchy");
    }
    m.built = levels;
}

void multilevel_apply(Context& c, const float* r, float* z) {
    multilevel_build(c);
    MgLevel& L0 = c.mg.level[0];
    L0.r = const_cast<float*>(r);
    L0.z = z;
    vcycle(c, 0);
    L0.r = nullptr;
    L0.z = nullptr;
}

}  // namespace smoke
