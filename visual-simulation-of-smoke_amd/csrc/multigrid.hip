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
// The coarsest level is solved inside an exact deflation of v = 1 on its
// fluid rows: z = c v + P^T S(P r), c = v.r / v.Av, P = I - A v v^T / v.Av,
// with S the network (or Jacobi sweeps).  The identity
// A^-1 = Q + P^T A^-1 P (Q = v v^T / v.Av) makes this exact for an exact S;
// it keeps the near-null mode of the anchored operator -- nearly constant,
// with an eigenvalue far below the rest -- away from S, which a network with a
// local receptive field cannot resolve (it needs the global sum of r).
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
    int sy, sz;          // strides of y and z (n0, n0 n1)
    const uint8_t* flags;
    float omega;
};

__device__ __forceinline__ int outflow_faces(const LevelArgs& a, int x, int y, int z) {
    if (!a.any_outflow) return 0;
    return (x == 0 && a.outflow[0]) + (x == a.n[0] - 1 && a.outflow[1]) + (y == 0 && a.outflow[2]) + (y == a.n[1] - 1 && a.outflow[3]) +
           (z == 0 && a.outflow[4]) + (z == a.n[2] - 1 && a.outflow[5]);
}

// diagonal of a fluid (non-identity) row: fluid neighbours + OUTFLOW faces, at least 1
__device__ __forceinline__ int row_diag(const LevelArgs& a, unsigned f, int x, int y, int z) {
    const int d = __popc(f & 0x3fu) + outflow_faces(a, x, y, z);
    return d > 1 ? d : 1;
}

// 1 / d for the small integer diagonal (the hardware reciprocal, within one
// ulp; the preconditioner does not need the correctly rounded quotient)
__device__ __forceinline__ float recip(int d) { return __builtin_amdgcn_rcpf(float(d)); }

__device__ __forceinline__ int lin(const LevelArgs& a, int x, int y, int z) { return z * a.sz + y * a.sy + x; }

// Neighbour D (0..5: x-, x+, y-, y+, z-, z+) of cell i at (x, y, z): its index
// and coordinates.  Only used for set flag bits, so the neighbour exists; at
// the domain edge it is the periodic image.
template <int D>
__device__ __forceinline__ int neighbour(const LevelArgs& a, int i, int& x, int& y, int& z) {
    if constexpr (D == 0) {
        if (x > 0) return --x, i - 1;
        x = a.n[0] - 1;
        return i + a.n[0] - 1;
    } else if constexpr (D == 1) {
        if (x < a.n[0] - 1) return ++x, i + 1;
        x = 0;
        return i - (a.n[0] - 1);
    } else if constexpr (D == 2) {
        if (y > 0) return --y, i - a.sy;
        y = a.n[1] - 1;
        return i + (a.n[1] - 1) * a.sy;
    } else if constexpr (D == 3) {
        if (y < a.n[1] - 1) return ++y, i + a.sy;
        y = 0;
        return i - (a.n[1] - 1) * a.sy;
    } else if constexpr (D == 4) {
        if (z > 0) return --z, i - a.sz;
        z = a.n[2] - 1;
        return i + (a.n[2] - 1) * a.sz;
    } else {
        if (z < a.n[2] - 1) return ++z, i + a.sz;
        z = 0;
        return i - (a.n[2] - 1) * a.sz;
    }
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
    int n0;           // its x extent (y stride)
    int s2;           // its z stride
};

__device__ __forceinline__ int parent(const Coarse& cc, int x, int y, int z) { return (z >> 1) * cc.s2 + (y >> 1) * cc.n0 + (x >> 1); }

// x at cell i (flag byte f): identity rows hold 0
template <int XIN>
__device__ __forceinline__ float x_at(const LevelArgs& a, const float* __restrict__ r, const float* __restrict__ xb, const Coarse& cc, int x, int y,
                                      int z, int i, unsigned f) {
    if (XIN == kZero) return 0.0f;
    if constexpr (XIN == kBuffer) return xb[i];   // already 0 on identity rows
    if (f & kIdentity) return 0.0f;
    float v = 0.0f;
    if constexpr (XIN == kFromR || XIN == kFromRPlusC) v = a.omega * r[i] * recip(row_diag(a, f, x, y, z));
    if constexpr (XIN == kBufferPlusC) v = xb[i];
    if constexpr (XIN == kBufferPlusC || XIN == kFromRPlusC) v += cc.e[parent(cc, x, y, z)];
    return v;
}

template <int XIN, int D>
__device__ __forceinline__ float tap(const LevelArgs& a, const float* __restrict__ r, const float* __restrict__ xb, const Coarse& cc, int x, int y,
                                     int z, int i, unsigned f) {
    if (!(f & (1u << D))) return 0.0f;
    int xx = x, yy = y, zz = z;
    const int j = neighbour<D>(a, i, xx, yy, zz);
    const unsigned fj = (XIN == kBuffer) ? 0u : unsigned(a.flags[j]);
    return x_at<XIN>(a, r, xb, cc, xx, yy, zz, j, fj);
}

// (r - A x)_i of fluid row i with flag byte f; also returns x_i and 1 / diag_i
template <int XIN>
__device__ __forceinline__ float residual_at(const LevelArgs& a, const float* __restrict__ r, const float* __restrict__ xb, const Coarse& cc, int x,
                                             int y, int z, int i, unsigned f, float& xi, float& inv) {
    const int diag = row_diag(a, f, x, y, z);
    inv = recip(diag);
    if constexpr (XIN == kZero) {
        xi = 0.0f;
        return r[i];
    } else {
        float ri = r[i];
        if constexpr (XIN == kFromR || XIN == kFromRPlusC) xi = a.omega * ri * inv;
        else xi = xb[i];
        if constexpr (XIN == kBufferPlusC || XIN == kFromRPlusC) xi += cc.e[parent(cc, x, y, z)];
        const float s = tap<XIN, 0>(a, r, xb, cc, x, y, z, i, f) + tap<XIN, 1>(a, r, xb, cc, x, y, z, i, f) +
                        tap<XIN, 2>(a, r, xb, cc, x, y, z, i, f) + tap<XIN, 3>(a, r, xb, cc, x, y, z, i, f) +
                        tap<XIN, 4>(a, r, xb, cc, x, y, z, i, f) + tap<XIN, 5>(a, r, xb, cc, x, y, z, i, f);
        return ri - (float(diag) * xi - s);
    }
}

// out = x + omega D^-1 (r - A x) on fluid rows, 0 on identity rows
template <int XIN>
__global__ __launch_bounds__(256) void mg_sweep_kernel(LevelArgs a, const float* __restrict__ r, const float* __restrict__ xb, Coarse cc,
                                                       float* __restrict__ out) {
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int z = blockIdx.z;
    if (x >= a.n[0] || y >= a.n[1]) return;
    const int i = lin(a, x, y, z);
    const unsigned f = a.flags[i];
    if (f & kIdentity) {
        out[i] = 0.0f;
        return;
    }
    float xi, inv;
    const float res = residual_at<XIN>(a, r, xb, cc, x, y, z, i, f, xi, inv);
    out[i] = xi + a.omega * res * inv;
}

// x = omega D^-1 r (one sweep from zero) on fluid rows, 0 on identity rows
__global__ __launch_bounds__(256) void mg_scale_kernel(LevelArgs a, const float* __restrict__ r, float* __restrict__ out) {
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int z = blockIdx.z;
    if (x >= a.n[0] || y >= a.n[1]) return;
    const int i = lin(a, x, y, z);
    const unsigned f = a.flags[i];
    out[i] = (f & kIdentity) ? 0.0f : a.omega * r[i] * recip(row_diag(a, f, x, y, z));
}

// x += correction of the parent cell, on fluid rows (identity rows stay 0)
__global__ __launch_bounds__(256) void mg_prolong_kernel(LevelArgs a, Coarse cc, float* __restrict__ xb) {
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int z = blockIdx.z;
    if (x >= a.n[0] || y >= a.n[1]) return;
    const int i = lin(a, x, y, z);
    if (!(a.flags[i] & kIdentity)) xb[i] += cc.e[parent(cc, x, y, z)];
}

// r_c = 1/2 * sum over the children of (r - A x); 0 on coarse identity rows.
// Threads own fine cells: a 64 x 4 block covers two z planes (each thread a
// column of two cells); the x pair is summed with a lane shuffle, the y pair
// through LDS, and the thread of the even (x, y) child writes.
template <int XIN>
__global__ __launch_bounds__(256) void mg_restrict_kernel(LevelArgs a, const float* __restrict__ r, const float* __restrict__ xb, int nc0, int nc1,
                                                          const uint8_t* __restrict__ cflags, float* __restrict__ rc) {
    __shared__ float rows[4][64];
    const int lane = threadIdx.x & 63, row = threadIdx.x >> 6;
    const int x = blockIdx.x * 64 + lane;
    const int y = blockIdx.y * 4 + row;
    const int Z = blockIdx.z;
    const Coarse none{nullptr, 0, 0};
    float s = 0.0f;
    if (x < a.n[0] && y < a.n[1]) {
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const int z = 2 * Z + k;
            if (z >= a.n[2]) break;
            const int i = lin(a, x, y, z);
            const unsigned f = a.flags[i];
            if (f & kIdentity) continue;
            float xi, inv;
            s += residual_at<XIN>(a, r, xb, none, x, y, z, i, f, xi, inv);
        }
    }
    s += __shfl_xor(s, 1, 64);
    rows[row][lane] = s;
    __syncthreads();
    if ((row & 1) || (lane & 1)) return;
    const int X = x >> 1, Y = y >> 1;
    if (X >= nc0 || Y >= nc1) return;
    const int ci = (Z * nc1 + Y) * nc0 + X;
    rc[ci] = (cflags[ci] & kIdentity) ? 0.0f : 0.5f * (s + rows[row + 1][lane]);
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

constexpr int kDeflThreads = 1024;

// one block: sum over the block's threads in FP64
__device__ double block_sum(double v, double* red) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < kDeflThreads / 64; ++w) t += red[w];
        red[kDeflThreads / 64] = t;
    }
    __syncthreads();
    return red[kDeflThreads / 64];
}

// A v for v = 1 on the fluid rows (the anchor's fluid neighbours and the
// OUTFLOW faces carry it) and v^T A v (one block)
__global__ __launch_bounds__(kDeflThreads) void mg_deflation_setup_kernel(LevelArgs a, float* __restrict__ av, double* __restrict__ defl) {
    __shared__ double red[kDeflThreads / 64 + 1];
    const int64_t cells = int64_t(a.n[0]) * a.n[1] * a.n[2];
    double s = 0.0;
    for (int64_t i = threadIdx.x; i < cells; i += kDeflThreads) {
        const unsigned f = a.flags[i];
        float v = 0.0f;
        if (!(f & kIdentity)) {
            const int x = int(i % a.n[0]), y = int((i / a.n[0]) % a.n[1]), z = int(i / (int64_t(a.n[0]) * a.n[1]));
            int k = outflow_faces(a, x, y, z);
            for (int d = 0; d < 6; ++d) {
                if (!(f & (1u << d))) continue;
                int c[3] = {x, y, z};
                const int ax = d >> 1;
                c[ax] += (d & 1) ? 1 : -1;
                if (c[ax] < 0) c[ax] += a.n[ax];
                else if (c[ax] >= a.n[ax]) c[ax] -= a.n[ax];
                k += (a.flags[lin(a, c[0], c[1], c[2])] & kIdentity) ? 1 : 0;
            }
            v = float(k);
        }
        av[i] = v;
        s += v;
    }
    const double t = block_sum(s, red);
    if (threadIdx.x == 0) defl[0] = t;
}

// sum over i of f(i) by one block, eight independent loads in flight per thread
template <class F>
__device__ __forceinline__ double block_sum_of(int64_t cells, F f, double* red) {
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int64_t i = threadIdx.x;
    for (; i + 7 * kDeflThreads < cells; i += 8 * kDeflThreads) {
#pragma unroll
        for (int u = 0; u < 8; ++u) acc[u] += f(i + u * kDeflThreads);
    }
    for (; i < cells; i += kDeflThreads) acc[0] += f(i);
    return block_sum(((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7])), red);
}

// c = v.r / v.Av, rd = r - c A v (one block)
__global__ __launch_bounds__(kDeflThreads) void mg_deflate_rhs_kernel(int64_t cells, const uint8_t* __restrict__ flags, const float* __restrict__ r,
                                                                       const float* __restrict__ av, double* __restrict__ defl, float* __restrict__ rd) {
    __shared__ double red[kDeflThreads / 64 + 1];
    const double t = block_sum_of(cells, [&](int64_t i) { return (flags[i] & kIdentity) ? 0.0 : double(r[i]); }, red);
    const double vav = defl[0];
    const double c = vav > 0.0 ? t / vav : 0.0;
    if (threadIdx.x == 0) defl[1] = c;
    const float cf = float(c);
#pragma unroll 8
    for (int64_t i = threadIdx.x; i < cells; i += kDeflThreads) rd[i] = (flags[i] & kIdentity) ? 0.0f : r[i] - cf * av[i];
}

// z = z' + (c - Av.z' / v.Av) v (one block, in place)
__global__ __launch_bounds__(kDeflThreads) void mg_deflate_solution_kernel(int64_t cells, const uint8_t* __restrict__ flags,
                                                                            const float* __restrict__ av, const double* __restrict__ defl,
                                                                            float* __restrict__ z) {
    __shared__ double red[kDeflThreads / 64 + 1];
    const double t = block_sum_of(cells, [&](int64_t i) { return double(av[i]) * double(z[i]); }, red);
    const double vav = defl[0];
    const float coef = vav > 0.0 ? float(defl[1] - t / vav) : 0.0f;
#pragma unroll 8
    for (int64_t i = threadIdx.x; i < cells; i += kDeflThreads)
        if (!(flags[i] & kIdentity)) z[i] += coef;
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
    a.sy = L.n[0];
    a.sz = L.n[0] * L.n[1];
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
        // deflated right-hand side (what the coarse solver sees, and what the
        // capture stores as training data)
        mg_deflate_rhs_kernel<<<1, kDeflThreads, 0, c.stream>>>(L.cells, L.flags, L.r, L.av, L.defl, L.rd);
        hip_check(hipGetLastError(), "mg_deflate_rhs_kernel");
        if (m.capture && m.captured < m.capacity) {
            hip_check(hipMemcpyAsync(m.capture + size_t(m.captured) * size_t(L.cells), L.rd, size_t(L.cells) * sizeof(float),
                                     hipMemcpyDeviceToDevice, c.stream),
                      "multilevel capture");
            ++m.captured;
        }
        if (o.coarse_solver == SMOKE_SIMULATION_COARSE_NETWORK) {
            if (!c.learned) throw HipError("multilevel preconditioner: no network attached");
            preconditioner_apply(*c.learned, L.rd, L.flags, L.n, L.z, c.stream);
            // optional sweeps after the network's correction
            float* bufs[2] = {L.ta, L.tb};
            const float* cur = L.z;
            for (int k = 1; k <= o.coarse_sweeps; ++k) {
                float* dst = bufs[(k - 1) & 1];
                sweep(c, L, kBuffer, L.rd, cur, none, dst);
                cur = dst;
            }
            if (cur != L.z) hip_check(hipMemcpyAsync(L.z, cur, size_t(L.cells) * sizeof(float), hipMemcpyDeviceToDevice, c.stream), "coarse copy");
        } else {
            sweeps_from_zero(c, L, std::max(1, o.coarse_sweeps), L.rd, L.z);
        }
        mg_deflate_solution_kernel<<<1, kDeflThreads, 0, c.stream>>>(L.cells, L.flags, L.av, L.defl, L.z);
        hip_check(hipGetLastError(), "mg_deflate_solution_kernel");
        return;
    }
    MgLevel& C = m.level[l + 1];
    const LevelArgs a = level_args(c, L);
    const dim3 g = cell_grid(L.n);
    const Coarse cc{C.z, C.n[0], C.n[0] * C.n[1]};
    float* bufs[2] = {L.ta, L.tb};
    // Large levels run every sweep from a stored iterate: the first sweep
    // and the prolongation are element-wise passes, so no stencil tap needs
    // its neighbour's flag byte (identity rows hold 0 in every stored
    // iterate).  Small levels are bound by launches instead and recompute
    // the first sweep and the prolongation at every tap.
    const bool split = L.cells >= (int64_t(1) << 20);
    float* xb = nullptr;   // x in a buffer; nullptr: x = omega D^-1 r on the fly
    if (split) {
        mg_scale_kernel<<<g, 256, 0, c.stream>>>(a, L.r, L.ta);
        hip_check(hipGetLastError(), "mg_scale_kernel");
        xb = L.ta;
        for (int k = 2; k <= o.pre_sweeps; ++k) {
            float* dst = xb == L.ta ? L.tb : L.ta;
            sweep(c, L, kBuffer, L.r, xb, none, dst);
            xb = dst;
        }
    } else {
        for (int k = 2; k <= o.pre_sweeps; ++k) {
            float* dst = bufs[k & 1];
            if (k == 2) sweep(c, L, kFromR, L.r, nullptr, none, dst);
            else sweep(c, L, kBuffer, L.r, xb, none, dst);
            xb = dst;
        }
    }
    // fine-cell threads, two z planes per block
    const dim3 gc(unsigned((L.n[0] + 63) / 64), unsigned((L.n[1] + 3) / 4), unsigned((L.n[2] + 1) / 2));
    if (xb) mg_restrict_kernel<kBuffer><<<gc, 256, 0, c.stream>>>(a, L.r, xb, C.n[0], C.n[1], C.flags, C.r);
    else mg_restrict_kernel<kFromR><<<gc, 256, 0, c.stream>>>(a, L.r, nullptr, C.n[0], C.n[1], C.flags, C.r);
    hip_check(hipGetLastError(), "mg_restrict_kernel");
    vcycle(c, l + 1);
    // post-smoothing from x + P z_(l+1)
    const float* cur;
    int done_sweeps;
    if (split) {
        mg_prolong_kernel<<<g, 256, 0, c.stream>>>(a, cc, xb);
        hip_check(hipGetLastError(), "mg_prolong_kernel");
        cur = xb;
        done_sweeps = 0;
    } else {
        float* other = (xb == L.ta) ? L.tb : L.ta;   // the buffer not holding x
        float* first_out = o.post_sweeps == 1 ? L.z : other;
        sweep(c, L, xb ? kBufferPlusC : kFromRPlusC, L.r, xb, cc, first_out);
        cur = first_out;
        done_sweeps = 1;
    }
    for (int k = done_sweeps + 1; k <= o.post_sweeps; ++k) {
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
        free_level(L.av);
        free_level(L.rd);
        free_level(L.defl);
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
            if (l == levels - 1) {
                alloc_level(L.av, L.cells);
                alloc_level(L.rd, L.cells);
                alloc_level(L.defl, 2);
            }
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
    MgLevel& B = m.level[levels - 1];
    mg_deflation_setup_kernel<<<1, kDeflThreads, 0, c.stream>>>(level_args(c, B), B.av, B.defl);
    hip_check(hipGetLastError(), "mg_deflation_setup_kernel");
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
