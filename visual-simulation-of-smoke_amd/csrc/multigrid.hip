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

__global__ void mg_copy_kernel(const float* __restrict__ src, float* __restrict__ dst, int n) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

// ---- float4 row kernels for large levels with nx % 4 == 0: a thread owns
// four consecutive x cells of one row; the rows at y +- 1 and z +- 1 are
// float4 loads, the cells at x - 1 and x + 4 scalar loads.  A neighbour
// outside a non-periodic domain has no flag bit, so its (clamped, in-bounds)
// load is never used.
struct Row4 {
    int i0;                 // index of the first cell
    int x0, y, z;
    unsigned f[4];          // flag bytes
};

__device__ __forceinline__ Row4 row4_at(const LevelArgs& a, int x0, int y, int z) {
    Row4 q;
    q.x0 = x0;
    q.y = y;
    q.z = z;
    q.i0 = lin(a, x0, y, z);
    const unsigned w = *reinterpret_cast<const unsigned*>(a.flags + q.i0);
#pragma unroll
    for (int k = 0; k < 4; ++k) q.f[k] = (w >> (8 * k)) & 0xffu;
    return q;
}

__device__ __forceinline__ float4 ld4(const float* __restrict__ p, int i) { return *reinterpret_cast<const float4*>(p + i); }
__device__ __forceinline__ float comp(const float4& v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : (k == 2 ? v.z : v.w)); }

// (r - A x) of the four cells (0 where identity) and their 1 / diag; x from a
// buffer that holds 0 on identity rows
__device__ __forceinline__ void row4_residual(const LevelArgs& a, const float* __restrict__ r, const float* __restrict__ xb, const Row4& q,
                                              float res[4], float xi[4], float inv[4]) {
    const int nx = a.n[0], ny = a.n[1], nz = a.n[2];
    const int row = q.i0 - q.x0;
    const int ym = q.y > 0 ? q.y - 1 : (a.periodic[1] ? ny - 1 : q.y);
    const int yp = q.y < ny - 1 ? q.y + 1 : (a.periodic[1] ? 0 : q.y);
    const int zm = q.z > 0 ? q.z - 1 : (a.periodic[2] ? nz - 1 : q.z);
    const int zp = q.z < nz - 1 ? q.z + 1 : (a.periodic[2] ? 0 : q.z);
    const int im1 = q.x0 > 0 ? q.i0 - 1 : (a.periodic[0] ? row + nx - 1 : q.i0);
    const int ip4 = q.x0 + 4 < nx ? q.i0 + 4 : (a.periodic[0] ? row : q.i0);
    const float4 rv = ld4(r, q.i0);
    const float4 xv = ld4(xb, q.i0);
    const float4 xym = ld4(xb, q.i0 + (ym - q.y) * a.sy);
    const float4 xyp = ld4(xb, q.i0 + (yp - q.y) * a.sy);
    const float4 xzm = ld4(xb, q.i0 + (zm - q.z) * a.sz);
    const float4 xzp = ld4(xb, q.i0 + (zp - q.z) * a.sz);
    const float xm1 = xb[im1], xp4 = xb[ip4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const unsigned f = q.f[k];
        const float xk = comp(xv, k);
        const float left = k == 0 ? xm1 : comp(xv, k - 1);
        const float right = k == 3 ? xp4 : comp(xv, k + 1);
        float s = 0.0f;
        s += (f & 1u) ? left : 0.0f;
        s += (f & 2u) ? right : 0.0f;
        s += (f & 4u) ? comp(xym, k) : 0.0f;
        s += (f & 8u) ? comp(xyp, k) : 0.0f;
        s += (f & 16u) ? comp(xzm, k) : 0.0f;
        s += (f & 32u) ? comp(xzp, k) : 0.0f;
        const int d = row_diag(a, f, q.x0 + k, q.y, q.z);
        inv[k] = recip(d);
        xi[k] = xk;
        res[k] = (f & kIdentity) ? 0.0f : comp(rv, k) - (float(d) * xk - s);
    }
}

// DOT: also the block's partial sum of r . out (FP64), written to
// partials[block] -- the r.z of flexible PCG when `out` is its z
//
// LPR lanes per row: a block of 256 threads covers 4 LPR x cells of 256 / LPR
// rows (64: 256 x 4; 32 for levels at most 128 cells wide, whose rows would
// leave half of 64 lanes idle: 128 x 8).
template <bool DOT, int LPR>
__global__ __launch_bounds__(256) void mg_sweep4_kernel(LevelArgs a, uint3 tiles, const float* __restrict__ r, const float* __restrict__ xb,
                                                        float* __restrict__ out, double* __restrict__ partials) {
    uint3 tile;
    unsigned id;
    if (!xcd_tile(tiles, tile, id)) return;
    const int x0 = (int(tile.x) * LPR + int(threadIdx.x % LPR)) * 4;
    const int y = int(tile.y) * (256 / LPR) + int(threadIdx.x / LPR);
    const int z = int(tile.z);
    const bool inside = x0 < a.n[0] && y < a.n[1];
    double dot = 0.0;
    if (inside) {
        const Row4 q = row4_at(a, x0, y, z);
        float res[4], xi[4], inv[4];
        row4_residual(a, r, xb, q, res, xi, inv);
        float o[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) o[k] = (q.f[k] & kIdentity) ? 0.0f : xi[k] + a.omega * res[k] * inv[k];
        *reinterpret_cast<float4*>(out + q.i0) = make_float4(o[0], o[1], o[2], o[3]);
        if constexpr (DOT) {
            const float4 rv = ld4(r, q.i0);
            dot = (double(rv.x) * double(o[0]) + double(rv.y) * double(o[1])) + (double(rv.z) * double(o[2]) + double(rv.w) * double(o[3]));
        }
    }
    if constexpr (DOT) {
        __shared__ double red[4];
        for (int off = 32; off > 0; off >>= 1) dot += __shfl_down(dot, off, 64);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = dot;
        __syncthreads();
        if (threadIdx.x == 0) partials[id] = (red[0] + red[1]) + (red[2] + red[3]);
    }
}

template <int LPR>
__global__ __launch_bounds__(256) void mg_scale4_kernel(LevelArgs a, const float* __restrict__ r, float* __restrict__ out) {
    const int x0 = (int(blockIdx.x) * LPR + int(threadIdx.x % LPR)) * 4;
    const int y = int(blockIdx.y) * (256 / LPR) + int(threadIdx.x / LPR);
    const int z = blockIdx.z;
    if (x0 >= a.n[0] || y >= a.n[1]) return;
    const Row4 q = row4_at(a, x0, y, z);
    const float4 rv = ld4(r, q.i0);
    float o[4];
#pragma unroll
    for (int k = 0; k < 4; ++k)
        o[k] = (q.f[k] & kIdentity) ? 0.0f : a.omega * comp(rv, k) * recip(row_diag(a, q.f[k], x0 + k, y, z));
    *reinterpret_cast<float4*>(out + q.i0) = make_float4(o[0], o[1], o[2], o[3]);
}

// x += correction of the parent cell on fluid rows; the four cells have the
// parents X = x0 / 2 and X + 1 (x0 % 4 == 0, so the pair is one aligned float2)
template <int LPR>
__global__ __launch_bounds__(256) void mg_prolong4_kernel(LevelArgs a, Coarse cc, float* __restrict__ xb) {
    const int x0 = (int(blockIdx.x) * LPR + int(threadIdx.x % LPR)) * 4;
    const int y = int(blockIdx.y) * (256 / LPR) + int(threadIdx.x / LPR);
    const int z = blockIdx.z;
    if (x0 >= a.n[0] || y >= a.n[1]) return;
    const Row4 q = row4_at(a, x0, y, z);
    const float2 e = *reinterpret_cast<const float2*>(cc.e + parent(cc, x0, y, z));
    float4 v = ld4(xb, q.i0);
    if (!(q.f[0] & kIdentity)) v.x += e.x;
    if (!(q.f[1] & kIdentity)) v.y += e.x;
    if (!(q.f[2] & kIdentity)) v.z += e.y;
    if (!(q.f[3] & kIdentity)) v.w += e.y;
    *reinterpret_cast<float4*>(xb + q.i0) = v;
}

// restriction with float4 rows: a block covers 4 LPR x cells, 256 / LPR rows
// and two z planes; each thread sums its 2 x 4 cells into two coarse x cells,
// the y pair is summed through LDS
template <int LPR>
__global__ __launch_bounds__(256) void mg_restrict4_kernel(LevelArgs a, uint3 tiles, const float* __restrict__ r, const float* __restrict__ xb, int nc0,
                                                           int nc1, const uint8_t* __restrict__ cflags, float* __restrict__ rc) {
    __shared__ float2 rows[256 / LPR][LPR];
    uint3 tile;
    unsigned id;
    if (!xcd_tile(tiles, tile, id)) return;
    const int lane = int(threadIdx.x % LPR), rowi = int(threadIdx.x / LPR);
    const int x0 = (int(tile.x) * LPR + lane) * 4;
    const int y = int(tile.y) * (256 / LPR) + rowi;
    const int Z = int(tile.z);
    float s0 = 0.0f, s1 = 0.0f;
    if (x0 < a.n[0] && y < a.n[1]) {
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const int z = 2 * Z + k;
            if (z >= a.n[2]) break;
            const Row4 q = row4_at(a, x0, y, z);
            float res[4], xi[4], inv[4];
            row4_residual(a, r, xb, q, res, xi, inv);
            s0 += res[0] + res[1];
            s1 += res[2] + res[3];
        }
    }
    rows[rowi][lane] = make_float2(s0, s1);
    __syncthreads();
    if (rowi & 1) return;
    const int X = x0 >> 1, Y = y >> 1;
    if (X >= nc0 || Y >= nc1) return;
    const float2 o = rows[rowi + 1][lane];
    const int ci = (Z * nc1 + Y) * nc0 + X;
    // X is even and nc0 = nx / 2 is even: the pair is one aligned float2
    const unsigned cf = *reinterpret_cast<const unsigned short*>(cflags + ci);
    const float v0 = (cf & kIdentity) ? 0.0f : 0.5f * (s0 + o.x);
    const float v1 = ((cf >> 8) & kIdentity) ? 0.0f : 0.5f * (s1 + o.y);
    *reinterpret_cast<float2*>(rc + ci) = make_float2(v0, v1);
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

// The deflation's two global sums (v.r before the coarse solve, Av.z' after
// it) are per-block partial sums followed by an element-wise kernel whose
// every block adds the partials in the same order (deterministic).
constexpr int kDeflBlocks = 64;

__device__ __forceinline__ double wave_block_sum256(double v, double* red) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    return (red[0] + red[1]) + (red[2] + red[3]);
}

// MODE 0: partial sums of r over fluid rows; MODE 1: of Av . z
template <int MODE>
__global__ __launch_bounds__(256) void mg_defl_sum_kernel(int cells, const uint8_t* __restrict__ flags, const float* __restrict__ v,
                                                          const float* __restrict__ av, double* __restrict__ partials) {
    __shared__ double red[4];
    double s = 0.0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < cells; i += kDeflBlocks * 256) {
        if constexpr (MODE == 0) s += (flags[i] & kIdentity) ? 0.0 : double(v[i]);
        else s += double(av[i]) * double(v[i]);
    }
    const double t = wave_block_sum256(s, red);
    if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

__device__ __forceinline__ double sum_partials(const double* __restrict__ partials) {
    double t = 0.0;
#pragma unroll 8
    for (int b = 0; b < kDeflBlocks; ++b) t += partials[b];
    return t;
}

// rd = r - c A v with c = v.r / v.Av (block 0 stores c)
__global__ __launch_bounds__(256) void mg_deflate_rhs_kernel(int cells, const uint8_t* __restrict__ flags, const float* __restrict__ r,
                                                             const float* __restrict__ av, double* __restrict__ defl,
                                                             const double* __restrict__ partials, float* __restrict__ rd) {
    const double vav = defl[0];
    const double c = vav > 0.0 ? sum_partials(partials) / vav : 0.0;
    if (blockIdx.x == 0 && threadIdx.x == 0) defl[1] = c;
    const float cf = float(c);
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < cells) rd[i] = (flags[i] & kIdentity) ? 0.0f : fmaf(-cf, av[i], r[i]);
}

// z = z' + (c - Av.z' / v.Av) v
__global__ __launch_bounds__(256) void mg_deflate_solution_kernel(int cells, const uint8_t* __restrict__ flags, const double* __restrict__ defl,
                                                                  const double* __restrict__ partials, float* __restrict__ z) {
    const double vav = defl[0];
    const float coef = vav > 0.0 ? float(defl[1] - sum_partials(partials) / vav) : 0.0f;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < cells && !(flags[i] & kIdentity)) z[i] += coef;
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

// Launches of the float4-row kernels (levels over 2^20 cells with nx % 8 ==
// 0): 32 lanes per row when a row has at most 128 cells, else 64.
struct Vec4Grid {
    int lpr;
    dim3 g;    // one z plane per block
    dim3 gc;   // two z planes per block (restriction)
};

Vec4Grid vec4_grid(const MgLevel& L) {
    Vec4Grid v;
    v.lpr = L.n[0] / 4 <= 32 ? 32 : 64;
    const unsigned gx = unsigned((L.n[0] / 4 + v.lpr - 1) / v.lpr), gy = unsigned((L.n[1] + 256 / v.lpr - 1) / (256 / v.lpr));
    v.g = dim3(gx, gy, unsigned(L.n[2]));
    v.gc = dim3(gx, gy, unsigned((L.n[2] + 1) / 2));
    return v;
}

void sweep4(Context& c, const Vec4Grid& v, const LevelArgs& a, const float* r, const float* xb, float* out, double* partials) {
    const uint3 t = make_uint3(v.g.x, v.g.y, v.g.z);
    const unsigned blocks = xcd_blocks(v.g);
    if (v.lpr == 32) {
        if (partials) mg_sweep4_kernel<true, 32><<<blocks, 256, 0, c.stream>>>(a, t, r, xb, out, partials);
        else mg_sweep4_kernel<false, 32><<<blocks, 256, 0, c.stream>>>(a, t, r, xb, out, nullptr);
    } else {
        if (partials) mg_sweep4_kernel<true, 64><<<blocks, 256, 0, c.stream>>>(a, t, r, xb, out, partials);
        else mg_sweep4_kernel<false, 64><<<blocks, 256, 0, c.stream>>>(a, t, r, xb, out, nullptr);
    }
    hip_check(hipGetLastError(), "mg_sweep4_kernel");
}

void scale4(Context& c, const Vec4Grid& v, const LevelArgs& a, const float* r, float* out) {
    if (v.lpr == 32) mg_scale4_kernel<32><<<v.g, 256, 0, c.stream>>>(a, r, out);
    else mg_scale4_kernel<64><<<v.g, 256, 0, c.stream>>>(a, r, out);
    hip_check(hipGetLastError(), "mg_scale4_kernel");
}

void prolong4(Context& c, const Vec4Grid& v, const LevelArgs& a, Coarse cc, float* xb) {
    if (v.lpr == 32) mg_prolong4_kernel<32><<<v.g, 256, 0, c.stream>>>(a, cc, xb);
    else mg_prolong4_kernel<64><<<v.g, 256, 0, c.stream>>>(a, cc, xb);
    hip_check(hipGetLastError(), "mg_prolong4_kernel");
}

void restrict4(Context& c, const Vec4Grid& v, const LevelArgs& a, const float* r, const float* xb, const MgLevel& C) {
    const uint3 t = make_uint3(v.gc.x, v.gc.y, v.gc.z);
    const unsigned blocks = xcd_blocks(v.gc);
    if (v.lpr == 32) mg_restrict4_kernel<32><<<blocks, 256, 0, c.stream>>>(a, t, r, xb, C.n[0], C.n[1], C.flags, C.r);
    else mg_restrict4_kernel<64><<<blocks, 256, 0, c.stream>>>(a, t, r, xb, C.n[0], C.n[1], C.flags, C.r);
    hip_check(hipGetLastError(), "mg_restrict4_kernel");
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

void fold_timing_pair(Context& c, int k) {
    Multilevel& m = c.mg;
    if (!m.timing_pending[k]) return;
    hip_check(hipEventSynchronize(m.timing_events[k][1]), "timing event");
    float ms = 0.0f;
    hip_check(hipEventElapsedTime(&ms, m.timing_events[k][0], m.timing_events[k][1]), "elapsed");
    c.timing_totals.mg_sweep_ms += ms;
    c.timing_totals.mg_sweeps += uint64_t(m.timing_count[k]);
    m.timing_pending[k] = false;
}

// the event pair around the next `count` back-to-back timed launches (created
// on first use)
hipEvent_t* next_timing_pair(Context& c, int count) {
    Multilevel& m = c.mg;
    const int k = m.timing_next;
    m.timing_next = (k + 1) % Multilevel::kTimingRing;
    fold_timing_pair(c, k);
    if (!m.timing_events[k][0]) {
        hip_check(hipEventCreate(&m.timing_events[k][0]), "event");
        hip_check(hipEventCreate(&m.timing_events[k][1]), "event");
    }
    m.timing_pending[k] = true;
    m.timing_count[k] = count;
    return m.timing_events[k];
}

void vcycle(Context& c, int l) {
    Multilevel& m = c.mg;
    MgLevel& L = m.level[l];
    const Coarse none{nullptr, 0, 0};
    const auto& o = m.options;
    if (l == m.built - 1) {
        const bool capturing = m.capture && m.captured < m.capacity;
        if (o.coarse_solver == SMOKE_SIMULATION_COARSE_NETWORK && o.coarse_sweeps <= kCoarseSolveMaxSweeps && !capturing &&
            L.cells <= kCoarseSolveMaxCells) {
            // deflation, network, sweeps and deflated correction in few launches
            if (!c.learned) throw HipError("multilevel preconditioner: no network attached");
            const LevelArgs a = level_args(c, L);
            CoarseSweeps cs{};
            for (int f = 0; f < 6; ++f) cs.outflow[f] = a.outflow[f];
            cs.omega = a.omega;
            cs.count = o.coarse_sweeps;
            preconditioner_coarse_solve(*c.learned, L.r, L.av, L.flags, L.defl, L.n, cs, L.rd, L.z, c.stream);
            return;
        }
        // deflated right-hand side (what the coarse solver sees, and what the
        // capture stores as training data)
        const int cells = int(L.cells);
        const unsigned eb = unsigned((cells + 255) / 256);
        mg_defl_sum_kernel<0><<<kDeflBlocks, 256, 0, c.stream>>>(cells, L.flags, L.r, nullptr, L.defl + 2);
        mg_deflate_rhs_kernel<<<eb, 256, 0, c.stream>>>(cells, L.flags, L.r, L.av, L.defl, L.defl + 2, L.rd);
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
            if (cur != L.z) {
                mg_copy_kernel<<<eb, 256, 0, c.stream>>>(cur, L.z, cells);
                hip_check(hipGetLastError(), "mg_copy_kernel");
            }
        } else {
            sweeps_from_zero(c, L, std::max(1, o.coarse_sweeps), L.rd, L.z);
        }
        mg_defl_sum_kernel<1><<<kDeflBlocks, 256, 0, c.stream>>>(cells, L.flags, L.z, L.av, L.defl + 2);
        mg_deflate_solution_kernel<<<eb, 256, 0, c.stream>>>(cells, L.flags, L.defl, L.defl + 2, L.z);
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
    // float4 rows: x extent a multiple of 4 (and of 8 when the coarse x pair of
    // the restriction must be a float2-aligned pair: nc0 even)
    const bool vec = split && L.n[0] % 8 == 0;
    const Vec4Grid v4 = vec4_grid(L);
    float* xb = nullptr;   // x in a buffer; nullptr: x = omega D^-1 r on the fly
    if (split) {
        if (vec) scale4(c, v4, a, L.r, L.ta);
        else mg_scale_kernel<<<g, 256, 0, c.stream>>>(a, L.r, L.ta);
        hip_check(hipGetLastError(), "mg_scale_kernel");
        xb = L.ta;
        // the pre-sweeps after the first run back to back: one event pair
        hipEvent_t* ev = (c.timing && l == 0 && o.pre_sweeps >= 2) ? next_timing_pair(c, o.pre_sweeps - 1) : nullptr;
        if (ev) hip_check(hipEventRecord(ev[0], c.stream), "event");
        for (int k = 2; k <= o.pre_sweeps; ++k) {
            float* dst = xb == L.ta ? L.tb : L.ta;
            if (vec) sweep4(c, v4, a, L.r, xb, dst, nullptr);
            else sweep(c, L, kBuffer, L.r, xb, none, dst);
            xb = dst;
        }
        if (ev) hip_check(hipEventRecord(ev[1], c.stream), "event");
    } else {
        hipEvent_t* ev = (c.timing && l == 0 && o.pre_sweeps >= 2) ? next_timing_pair(c, o.pre_sweeps - 1) : nullptr;
        if (ev) hip_check(hipEventRecord(ev[0], c.stream), "event");
        for (int k = 2; k <= o.pre_sweeps; ++k) {
            float* dst = bufs[k & 1];
            if (k == 2) sweep(c, L, kFromR, L.r, nullptr, none, dst);
            else sweep(c, L, kBuffer, L.r, xb, none, dst);
            xb = dst;
        }
        if (ev) hip_check(hipEventRecord(ev[1], c.stream), "event");
    }
    // fine-cell threads, two z planes per block
    const dim3 gc(unsigned((L.n[0] + 63) / 64), unsigned((L.n[1] + 3) / 4), unsigned((L.n[2] + 1) / 2));
    if (vec) restrict4(c, v4, a, L.r, xb, C);
    else if (xb) mg_restrict_kernel<kBuffer><<<gc, 256, 0, c.stream>>>(a, L.r, xb, C.n[0], C.n[1], C.flags, C.r);
    else mg_restrict_kernel<kFromR><<<gc, 256, 0, c.stream>>>(a, L.r, nullptr, C.n[0], C.n[1], C.flags, C.r);
    hip_check(hipGetLastError(), "mg_restrict_kernel");
    vcycle(c, l + 1);
    // post-smoothing from x + P z_(l+1)
    const float* cur;
    int done_sweeps;
    if (split) {
        if (vec) prolong4(c, v4, a, cc, xb);
        else mg_prolong_kernel<<<g, 256, 0, c.stream>>>(a, cc, xb);
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
    // the remaining post-sweeps run back to back: one event pair
    const int post_timed = o.post_sweeps - done_sweeps;
    hipEvent_t* ev = (c.timing && l == 0 && post_timed > 0) ? next_timing_pair(c, post_timed) : nullptr;
    if (ev) hip_check(hipEventRecord(ev[0], c.stream), "event");
    for (int k = done_sweeps + 1; k <= o.post_sweeps; ++k) {
        float* dst = k == o.post_sweeps ? L.z : (cur == bufs[0] ? bufs[1] : bufs[0]);
        const bool last0 = l == 0 && k == o.post_sweeps;
        const int64_t blocks4 = int64_t(v4.g.x) * v4.g.y * v4.g.z;
        if (vec && last0 && m.dot_partials && blocks4 <= m.dot_capacity) {
            sweep4(c, v4, a, L.r, cur, dst, m.dot_partials);
            m.dot_blocks = int(blocks4);
        } else if (vec) {
            sweep4(c, v4, a, L.r, cur, dst, nullptr);
        } else {
            sweep(c, L, kBuffer, L.r, cur, none, dst);
        }
        hip_check(hipGetLastError(), "mg_sweep_kernel");
        cur = dst;
    }
    if (ev) hip_check(hipEventRecord(ev[1], c.stream), "event");
}

}  // namespace

SmokeSimulationMultilevelOptions default_multilevel_options() {
    SmokeSimulationMultilevelOptions o{};
    o.levels = 5;
    o.pre_sweeps = 4;
    o.post_sweeps = 4;
    // the damping that minimises the smoothing factor of damped Jacobi on
    // the 7-point Laplacian in 3D: max |1 - omega (1 - c)| over the
    // oscillatory modes, c in [-1, 2/3], is 5/7 at omega = 6/7 (7/9 at 2/3)
    o.omega = 6.0f / 7.0f;
    o.coarse_solver = SMOKE_SIMULATION_COARSE_NETWORK;
    o.coarse_sweeps = 8;
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

void multilevel_drain_timing(Context& c) {
    for (int k = 0; k < Multilevel::kTimingRing; ++k) fold_timing_pair(c, k);
}

void multilevel_release(Context& c) {
    Multilevel& m = c.mg;
    for (auto& pair : m.timing_events)
        for (hipEvent_t& e : pair)
            if (e) (void)hipEventDestroy(e), e = nullptr;
    for (bool& p : m.timing_pending) p = false;
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
                alloc_level(L.defl, 2 + kDeflBlocks);   // v.Av, c, partial sums
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

int multilevel_apply(Context& c, const float* r, float* z, double* dot_partials, size_t dot_capacity) {
    multilevel_build(c);
    Multilevel& m = c.mg;
    MgLevel& L0 = m.level[0];
    L0.r = const_cast<float*>(r);
    L0.z = z;
    m.dot_partials = dot_partials;
    m.dot_capacity = int64_t(dot_capacity);
    m.dot_blocks = 0;
    vcycle(c, 0);
    L0.r = nullptr;
    L0.z = nullptr;
    m.dot_partials = nullptr;
    return m.dot_blocks;
}

}  // namespace smoke

namespace smoke {

SmokeSimulationMultilevelTiming multilevel_timing(Context& c, int reps) {
    multilevel_build(c);
    Multilevel& m = c.mg;
    if (m.built < 2) throw HipError("no hierarchy");
    if (!c.pcg_z) throw HipError("multilevel timing needs the preconditioner buffers");
    MgLevel& L = m.level[0];
    const MgLevel& C = m.level[1];
    const LevelArgs a = level_args(c, L);
    const bool vec = L.cells >= (int64_t(1) << 20) && L.n[0] % 8 == 0;
    const dim3 g = cell_grid(L.n);
    const dim3 gc(unsigned((L.n[0] + 63) / 64), unsigned((L.n[1] + 3) / 4), unsigned((L.n[2] + 1) / 2));
    const Vec4Grid v4 = vec4_grid(L);
    const Coarse none{nullptr, 0, 0};
    const Coarse cc{C.z, C.n[0], C.n[0] * C.n[1]};
    const float* r = c.pcg_r;
    hipEvent_t t0, t1;
    hip_check(hipEventCreate(&t0), "event");
    hip_check(hipEventCreate(&t1), "event");
    auto time = [&](auto&& launch) {
        launch();   // warm-up
        hip_check(hipEventRecord(t0, c.stream), "event");
        for (int k = 0; k < reps; ++k) launch();
        hip_check(hipEventRecord(t1, c.stream), "event");
        hip_check(hipEventSynchronize(t1), "event sync");
        float ms = 0.0f;
        hip_check(hipEventElapsedTime(&ms, t0, t1), "elapsed");
        hip_check(hipGetLastError(), "multilevel timing");
        return 1e3f * ms / float(reps);
    };
    SmokeSimulationMultilevelTiming out{};
    out.vector_rows = vec ? 1 : 0;
    out.repetitions = reps;
    out.scale_us = time([&] {
        if (vec) scale4(c, v4, a, r, L.ta);
        else mg_scale_kernel<<<g, 256, 0, c.stream>>>(a, r, L.ta);
    });
    out.sweep_us = time([&] {
        if (vec) sweep4(c, v4, a, r, L.ta, L.tb, nullptr);
        else mg_sweep_kernel<kBuffer><<<g, 256, 0, c.stream>>>(a, r, L.ta, none, L.tb);
    });
    out.restrict_us = time([&] {
        if (vec) restrict4(c, v4, a, r, L.tb, C);
        else mg_restrict_kernel<kBuffer><<<gc, 256, 0, c.stream>>>(a, r, L.tb, C.n[0], C.n[1], C.flags, C.r);
    });
    hip_check(hipMemsetAsync(C.z, 0, size_t(C.cells) * sizeof(float), c.stream), "zero correction");
    out.prolong_us = time([&] {
        if (vec) prolong4(c, v4, a, cc, L.tb);
        else mg_prolong_kernel<<<g, 256, 0, c.stream>>>(a, cc, L.tb);
    });
    float* saved = m.capture;
    m.capture = nullptr;   // timing must not add training samples
    out.vcycle_us = time([&] { multilevel_apply(c, r, c.pcg_z); });
    m.capture = saved;
    hipEventDestroy(t0);
    hipEventDestroy(t1);
    return out;
}

}  // namespace smoke
