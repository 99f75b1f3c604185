// Matrix-free 7-point pressure solve.
//
// Replaces the host CSR build + cuSPARSE SpMV + cuBLAS dot/axpy loop of
// backend-cuda.cu:1688-1824.  The operator is applied from one flag byte per
// cell (fluid neighbours + identity row) and the cell coordinates (OUTFLOW
// faces add to the diagonal), which reproduces the CSR values of
// B:1343-1412 exactly: the anchor keeps its -1 coupling here, which is exact
// because the anchor entry of every iterate is 0.
//
// One PCG iteration is three passes (39 bytes of compulsory HBM traffic per
// cell):
//   pass 1: p <- z + beta p in place (z = r / diag on the fly)
//           reads r, p, flags; writes p                            (13 B/cell)
//   pass 2: Ap from p, partial p.Ap                        reads p, flags (5 B/cell)
//   pass 3: Ap recomputed, e += alpha p, r -= alpha Ap, partial r.z and r.r
//                      reads p, flags, e, r; writes e, r           (21 B/cell)
// Pass 1 is a flat float4 stream.  Passes 2 and 3 give each thread four
// consecutive x cells (one float4) on two z planes and read every neighbour
// straight through the cache: with no LDS staging and no barriers all loads
// of a thread are independent and in flight together, which keeps the HBM
// pipe full where a z-marching LDS pass stalls on every plane.  Reductions
// are FP64: wave64 shuffles, LDS across waves, per-block partials, and a
// one-block finalize kernel after passes 2 and 3 sums the partials in block
// order (deterministic) and updates alpha / beta on the device, so
// fixed-iteration solves never return to the host.
#include "context.hpp"
#include "nn.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

namespace smoke {
namespace {

constexpr uint8_t kIdentity = 0x40;
constexpr int kResidualPlanes = 8;   // z planes per block of the FP64 residual kernel

// Division by a divisor d >= 1 known at launch, without a divide
// instruction: q = (t + ((n - t) >> s1)) >> s2 with t = mulhi(n, m), exact for
// every 32-bit n (Granlund and Montgomery, "Division by invariant integers
// using multiplication", 1994, in the form that keeps m at 32 bits by
// splitting the shift).  For a uniform n it runs on the scalar unit.
struct Divisor {
    unsigned m;
    int s1, s2;
};

Divisor make_divisor(unsigned d) {
    int l = 0;   // ceil(log2 d)
    while ((uint64_t(1) << l) < d) ++l;
    Divisor v;
    v.m = unsigned(((uint64_t(1) << 32) * ((uint64_t(1) << l) - d)) / d + 1);
    v.s1 = l > 0 ? 1 : 0;
    v.s2 = l > 0 ? l - 1 : 0;
    return v;
}

__device__ __forceinline__ unsigned divide(unsigned n, const Divisor& v) {
    const unsigned t = __umulhi(n, v.m);
    return (t + ((n - t) >> v.s1)) >> v.s2;
}

struct PcgArgs {
    int n[3];
    int periodic[3];
    int outflow[6];
    unsigned skip;        // off-diagonal directions ignored (periodic axes with n <= 2)
    int any_outflow;
    int tiles[3];         // tile grid of a stencil pass (x tiles, y tiles, z runs)
    Divisor tile_div[2];  // division by tiles[2] and by tiles[1]
    const uint8_t* flags;
    PcgScalars* s;
    double* partials;
};

// OUTFLOW faces of cell (x, y, z) (each adds 1 to the diagonal of a fluid row).
__device__ __forceinline__ int outflow_faces(const PcgArgs& a, int x, int y, int z) {
    if (!a.any_outflow) return 0;
    return (x == 0 && a.outflow[0]) + (x == a.n[0] - 1 && a.outflow[1]) + (y == 0 && a.outflow[2]) + (y == a.n[1] - 1 && a.outflow[3]) +
           (z == 0 && a.outflow[4]) + (z == a.n[2] - 1 && a.outflow[5]);
}

// Diagonal of the row with flag byte f: fluid neighbours plus OUTFLOW faces,
// at least 1.  An identity row's byte is exactly kIdentity (no neighbour
// bits), so its bit count is already its diagonal 1 and `extra` is ignored.
__device__ __forceinline__ float diag_from(unsigned f, int extra) {
    const int d = __popc(f) + ((f & kIdentity) ? 0 : extra);
    return float(d > 1 ? d : 1);
}

__device__ __forceinline__ float diag_of(const PcgArgs& a, uint8_t f, int x, int y, int z) { return diag_from(f, outflow_faces(a, x, y, z)); }

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

// Tile index of block b of a launch of nblocks; consecutive tiles are kept on
// one XCD (blocks are dealt round-robin over the 8 XCDs) so that neighbouring
// tiles share halo rows through the same L2.
__device__ __forceinline__ int tile_id(int b, int nblocks) {
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

// Writes this block's sums to partials[q * nblocks + bid].  The partials are
// summed by a one-block finalize kernel launched next on the same stream: a
// kernel boundary makes them visible across the XCDs' L2 caches for the
// price of one launch, where a last-block-done ticket would need an
// agent-scope release fence (an L2 writeback) in every block.
template <int NV, int NWAVES>
__device__ __forceinline__ void store_block_partials(double (&v)[NV], double* partials, int nblocks, int bid) {
    __shared__ double red[NV][NWAVES];
    block_sum<NV, NWAVES>(v, red);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NV; ++q) partials[q * nblocks + bid] = v[q];
    }
}

// Sums the partials of nblocks blocks in a fixed order (deterministic); the
// totals are valid in thread 0 of the finalize block.  Each thread keeps four
// independent loads in flight: the partials were just written from every
// XCD, so they come from memory and the sum is latency-bound.
constexpr int kFinalizeThreads = 1024;

template <int NV>
__device__ __forceinline__ void sum_partials(const double* __restrict__ partials, int nblocks, double (&tot)[NV]) {
    __shared__ double red[NV][kFinalizeThreads / 64];
    constexpr int T = kFinalizeThreads;
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        const double* __restrict__ row = partials + int64_t(q) * nblocks;
        double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
        int b = threadIdx.x;
        for (; b + 3 * T < nblocks; b += 4 * T) {
            s0 += row[b];
            s1 += row[b + T];
            s2 += row[b + 2 * T];
            s3 += row[b + 3 * T];
        }
        for (; b < nblocks; b += T) s0 += row[b];
        tot[q] = (s0 + s1) + (s2 + s3);
    }
    block_sum<NV, kFinalizeThreads / 64>(tot, red);
}

__device__ __forceinline__ double guarded(double num, double den, bool fixed) {
    const double lim = fixed ? 1.0e-20 : 1.0e-300;
    return fabs(den) > lim ? num / den : 0.0;
}

// Row (or column, or plane) to load the neighbour at coordinate v in [-1, n]
// from: wrapped by one period on a periodic axis; on a closed axis some row
// that exists, because a neighbour outside the domain has no flag bit and
// its value is never used.  No integer division.
__device__ __forceinline__ int neighbour_row(int v, int n, int periodic) {
    if (v < 0) return periodic ? v + n : 0;
    if (v >= n) return periodic ? v - n : n - 1;
    return v;
}

// v when bit `pos` of w is set, else +0.0f: the bit sign-extended to a mask
// (one bit-field extract) ANDed with the bits of v, the value of the select
// `(w >> pos) & 1 ? v : 0.0f` in two instructions.
__device__ __forceinline__ float if_bit(unsigned w, int pos, float v) {
    return __int_as_float(__float_as_int(v) & __builtin_amdgcn_sbfe(int(w), pos, 1));
}

// Tile of this block for the stencil passes: z fastest, consecutive tiles
// kept on one XCD (blocks are dealt round-robin over the 8 XCDs), so that
// the neighbours a tile reads through the cache were just loaded by the
// tiles beside it in the same L2.
__device__ __forceinline__ void stencil_tile(const PcgArgs& a, int block, int& bx, int& by, int& bz) {
    const int nb = a.tiles[0] * a.tiles[1] * a.tiles[2];
    const unsigned b = unsigned(tile_id(block, nb));
    const unsigned t = divide(b, a.tile_div[0]);
    const unsigned u = divide(t, a.tile_div[1]);
    bz = int(b - t * unsigned(a.tiles[2]));
    by = int(t - u * unsigned(a.tiles[1]));
    bx = int(u);
}

// Pass 1 of an iteration: p <- z + beta p in place, z = r / diag (Jacobi)
// or r; a fresh Krylov space (first) takes p = z.  Elementwise, 4 cells per
// thread: reads r, p, flags, writes p (13 B/cell).
// (the bodies of the pass kernels take their block index: the batched
// launches over many contexts run them with each context's own numbering)
template <bool JACOBI>
__device__ __forceinline__ void update_p_body(const PcgArgs& a, const float* __restrict__ r, float* __restrict__ p, int block) {
    const PcgScalars* __restrict__ S = a.s;
    if (S->done) return;
    const bool first = S->first != 0;
    const float beta = first ? 0.0f : float(S->beta);
    const int64_t cells = int64_t(a.n[0]) * a.n[1] * a.n[2];
    const int64_t base = (int64_t(block) * 256 + threadIdx.x) * 4;
    if (base >= cells) return;
    auto zval = [&](float rv, uint8_t f, int64_t i) {
        if constexpr (!JACOBI) return rv;
        int x = 0, y = 0, z = 0;
        if (a.any_outflow) {
            x = int(i % a.n[0]);
            y = int((i / a.n[0]) % a.n[1]);
            z = int(i / (int64_t(a.n[0]) * a.n[1]));
        }
        return rv / diag_of(a, f, x, y, z);
    };
    if (base + 4 <= cells) {
        const float4 rv = *reinterpret_cast<const float4*>(r + base);
        const float4 pv = *reinterpret_cast<const float4*>(p + base);
        const uchar4 f = *reinterpret_cast<const uchar4*>(a.flags + base);
        float4 o;
        o.x = zval(rv.x, f.x, base) + beta * (first ? 0.0f : pv.x);
        o.y = zval(rv.y, f.y, base + 1) + beta * (first ? 0.0f : pv.y);
        o.z = zval(rv.z, f.z, base + 2) + beta * (first ? 0.0f : pv.z);
        o.w = zval(rv.w, f.w, base + 3) + beta * (first ? 0.0f : pv.w);
        *reinterpret_cast<float4*>(p + base) = o;
    } else {
        for (int64_t i = base; i < cells; ++i) p[i] = zval(r[i], a.flags[i], i) + beta * (first ? 0.0f : p[i]);
    }
}

template <bool JACOBI>
__global__ __launch_bounds__(256) void pcg_update_p_kernel(PcgArgs a, const float* __restrict__ r, float* __restrict__ p) {
    update_p_body<JACOBI>(a, r, p, blockIdx.x);
}

// Stencil passes on the stored direction p, VEC cells per thread along x
// (float4 when nx % 4 == 0), ZR planes per thread, neighbours read through
// the cache (no LDS, no barriers: every load of a thread is independent):
//   MODE 0 (pass 2): partial p.Ap per block.  Reads p, flags (5 B/cell).
//   MODE 1 (pass 3): e += alpha p, r -= alpha Ap, partials r.z and r.r per
//                    block.  Reads p, flags, e, r; writes e, r (21 B/cell).
//   MODE 2 (flexible PCG update): as MODE 1, with the partial r.z taken
//                    against the stored previous preconditioned residual z
//                    (r_(k+1) . z_k, for the Polak-Ribiere beta).
// Addressing is 32-bit (fewer than 2^31 cells): the z planes of a block are
// uniform, and each lane resolves its row, its y neighbours and its two x
// edge cells once; every load is then one add from those offsets.
//
// Without skipped directions (every periodic axis longer than 2 cells) the
// operator needs no per-cell neighbour masks.  p is exactly +0 on identity
// rows: the right-hand side and the FP64 iterate are 0 there, so r, z
// (Jacobi or learned) and every direction stay +0, and subtracting the p of
// a solid neighbour subtracts +0 -- the same value, in the same order, as
// leaving the neighbour out.  A neighbour outside the domain is read from a
// zero plane stored after p (elements cells .. cells + nx ny - 1), chosen
// per lane (y rows, x edge cells) or per block (z planes).  Only the r
// update masks the identity rows themselves.  With skipped directions each
// neighbour is masked by its flag bit instead.
template <int VEC>
struct Row {
    float v[VEC];
};

template <int VEC>
__device__ __forceinline__ Row<VEC> load_row(const float* __restrict__ f, unsigned i) {
    Row<VEC> o;
    if constexpr (VEC == 4) {
        const float4 q = *reinterpret_cast<const float4*>(f + i);
        o.v[0] = q.x;
        o.v[1] = q.y;
        o.v[2] = q.z;
        o.v[3] = q.w;
    } else {
        o.v[0] = f[i];
    }
    return o;
}

template <int VEC>
__device__ __forceinline__ void store_row(float* __restrict__ f, unsigned i, const Row<VEC>& o) {
    if constexpr (VEC == 4) *reinterpret_cast<float4*>(f + i) = make_float4(o.v[0], o.v[1], o.v[2], o.v[3]);
    else f[i] = o.v[0];
}

template <int TXV, int TY, int ZR, int VEC, int MODE, bool JACOBI>
__device__ __forceinline__ void stencil_body(const PcgArgs& a, const float* __restrict__ p, float* __restrict__ e, float* __restrict__ r,
                                             const float* __restrict__ zold, int block) {
    PcgScalars* __restrict__ S = a.s;
    if (S->done) return;
    constexpr int NWAVES = TXV * TY / 64;
    const int nblocks = a.tiles[0] * a.tiles[1] * a.tiles[2];
    int bx, by, bz;
    stencil_tile(a, block, bx, by, bz);
    const int nx = a.n[0], ny = a.n[1], nz = a.n[2];
    const int xv = (bx * TXV + int(threadIdx.x) % TXV) * VEC;
    const int y = by * TY + int(threadIdx.x) / TXV;
    const int z0 = bz * ZR;
    const bool owner = xv < nx && y < ny;
    // a padding thread of a partial tile reads the cells of (0, 0): in bounds,
    // never used
    const int xs = owner ? xv : 0;
    const int ys = owner ? y : 0;
    const bool first = S->first != 0;
    const float alpha = float(S->alpha);
    const uint8_t* __restrict__ flags = a.flags;

    const unsigned plane = unsigned(nx) * unsigned(ny);
    const unsigned zero = plane * unsigned(nz);   // first element of the zero plane after p
    const unsigned row = unsigned(ys) * unsigned(nx);
    const unsigned cen = row + unsigned(xs);
    const unsigned ym = unsigned(neighbour_row(ys - 1, ny, a.periodic[1])) * unsigned(nx) + unsigned(xs);
    const unsigned yp = unsigned(neighbour_row(ys + 1, ny, a.periodic[1])) * unsigned(nx) + unsigned(xs);
    const unsigned xl = row + unsigned(neighbour_row(xs - 1, nx, a.periodic[0]));
    const unsigned xr = row + unsigned(neighbour_row(xs + VEC, nx, a.periodic[0]));

    Row<VEC> col[ZR + 2];
    double acc0 = 0.0, acc1 = 0.0;

    // Everything a plane reads besides the column of p, loaded for all ZR
    // planes before any arithmetic so that a thread's loads are in flight
    // together.
    struct Plane {
        Row<VEC> rm, rp, ev, rv, zv;
        float lft, rgt;
        unsigned w, i;
    };
    auto load_rest = [&](Plane& q, unsigned i, unsigned ym_at, unsigned yp_at, unsigned xl_at, unsigned xr_at) {
        q.i = i;
        q.rm = load_row<VEC>(p, ym_at);
        q.rp = load_row<VEC>(p, yp_at);
        q.lft = p[xl_at];
        q.rgt = p[xr_at];
        if constexpr (VEC == 4) q.w = *reinterpret_cast<const unsigned*>(flags + i);
        else q.w = flags[i];
        if constexpr (MODE >= 1) {
            q.ev = load_row<VEC>(e, i);
            q.rv = load_row<VEC>(r, i);
        }
        if constexpr (MODE == 2) q.zv = load_row<VEC>(zold, i);
    };

    // Plane j of the tile from its VEC products ap (+0 on identity rows) and
    // diagonals d.
    auto finish = [&](int j, const Plane& q, const float (&ap)[VEC], const float (&d)[VEC]) {
        const int z = z0 + j;
        const bool active = owner && z < nz;
        Row<VEC> en, rn;
        double s0 = 0.0, s1 = 0.0;
#pragma unroll
        for (int k = 0; k < VEC; ++k) {
            const float c = col[j + 1].v[k];
            if constexpr (MODE == 0) {
                s0 += double(c) * double(ap[k]);
            } else {
                en.v[k] = (first ? 0.0f : q.ev.v[k]) + alpha * c;
                rn.v[k] = q.rv.v[k] - alpha * ap[k];
                const float zk = MODE == 2 ? q.zv.v[k] : (JACOBI ? rn.v[k] / d[k] : rn.v[k]);
                s0 += double(rn.v[k]) * double(zk);
                s1 += double(rn.v[k]) * double(rn.v[k]);
            }
        }
        if constexpr (MODE >= 1) {
            // e + alpha p taken before the branch: otherwise the compiler sinks
            // the e load into it, after the stencil's loads have all returned
#pragma unroll
            for (int k = 0; k < VEC; ++k) asm volatile("" ::"v"(en.v[k]));
            if (active) {
                store_row<VEC>(e, q.i, en);
                store_row<VEC>(r, q.i, rn);
            }
        }
        // a product, not a select: the loads stay unconditional (a padding
        // thread's cells are in bounds and finite, so keep = 0 adds +0)
        const double keep = active ? 1.0 : 0.0;
        acc0 += s0 * keep;
        acc1 += s1 * keep;
    };

    // No skipped direction.  OUT: OUTFLOW faces add to the diagonals.
    auto unmasked = [&](auto out_tag) {
        constexpr bool OUT = decltype(out_tag)::value;
        const bool hxm = xs > 0 || a.periodic[0], hxp = xs + VEC < nx || a.periodic[0];
        const bool hym = ys > 0 || a.periodic[1], hyp = ys < ny - 1 || a.periodic[1];
        unsigned zoff[ZR + 2];   // planes z0 - 1 .. z0 + ZR, the zero plane where none exists
#pragma unroll
        for (int j = 0; j < ZR + 2; ++j) {
            const int v = z0 - 1 + j;
            unsigned off = zero;
            if (v >= 0 && v < nz) off = unsigned(v) * plane;
            else if (a.periodic[2] && v == -1) off = unsigned(nz - 1) * plane;
            else if (a.periodic[2] && v == nz) off = 0;
            zoff[j] = off;
        }
        Plane q[ZR];
#pragma unroll
        for (int j = 0; j < ZR + 2; ++j) col[j] = load_row<VEC>(p, zoff[j] + cen);
#pragma unroll
        for (int j = 0; j < ZR; ++j) {
            const unsigned zp = zoff[j + 1];
            // flags, e, r of a padding plane: the last plane (in bounds, unused)
            load_rest(q[j], unsigned(min(z0 + j, nz - 1)) * plane + cen, hym ? zp + ym : zero + unsigned(xs), hyp ? zp + yp : zero + unsigned(xs),
                      hxm ? zp + xl : zero, hxp ? zp + xr : zero);
        }
#pragma unroll
        for (int j = 0; j < ZR; ++j) {
            float ap[VEC], d[VEC];
#pragma unroll
            for (int k = 0; k < VEC; ++k) {
                const unsigned f = (q[j].w >> (8 * k)) & 0xffu;
                if constexpr (OUT) d[k] = diag_from(f, outflow_faces(a, xs + k, ys, z0 + j));
                else d[k] = diag_from(f, 0);
                const float c = col[j + 1].v[k];
                const float xm = k == 0 ? q[j].lft : col[j + 1].v[k - 1 < 0 ? 0 : k - 1];
                const float xp = k == VEC - 1 ? q[j].rgt : col[j + 1].v[k + 1 < VEC ? k + 1 : VEC - 1];
                float v = d[k] * c;
                v -= xm;
                v -= xp;
                v -= q[j].rm.v[k];
                v -= q[j].rp.v[k];
                v -= col[j].v[k];
                v -= col[j + 2].v[k];
                // identity row: A p = p = +0 (the p.Ap term c * v is +0 already)
                if constexpr (MODE >= 1) v = __int_as_float(__float_as_int(v) & ~__builtin_amdgcn_sbfe(int(q[j].w), 8 * k + 6, 1));
                ap[k] = v;
            }
            finish(j, q[j], ap, d);
        }
    };

    // Skipped directions: every neighbour masked by its flag bit.
    auto masked = [&]() {
        // neighbours that do not exist read some cell that does (never used)
        const unsigned skip_mask = ~(a.skip * (VEC == 4 ? 0x01010101u : 1u));
        unsigned zoff[ZR + 2];
#pragma unroll
        for (int j = 0; j < ZR + 2; ++j) zoff[j] = unsigned(min(neighbour_row(z0 - 1 + j, nz, a.periodic[2]), nz - 1)) * plane;
        Plane q[ZR];
#pragma unroll
        for (int j = 0; j < ZR + 2; ++j) col[j] = load_row<VEC>(p, zoff[j] + cen);
#pragma unroll
        for (int j = 0; j < ZR; ++j) {
            const unsigned zp = zoff[j + 1];
            load_rest(q[j], zp + cen, zp + ym, zp + yp, zp + xl, zp + xr);
        }
#pragma unroll
        for (int j = 0; j < ZR; ++j) {
            const unsigned w = q[j].w;
            const unsigned wm = w & skip_mask;   // the diagonal counts the skipped bits: it uses w
            float ap[VEC], d[VEC];
#pragma unroll
            for (int k = 0; k < VEC; ++k) {
                d[k] = diag_from((w >> (8 * k)) & 0xffu, outflow_faces(a, xs + k, ys, z0 + j));
                const float c = col[j + 1].v[k];
                const float xm = k == 0 ? q[j].lft : col[j + 1].v[k - 1 < 0 ? 0 : k - 1];
                const float xp = k == VEC - 1 ? q[j].rgt : col[j + 1].v[k + 1 < VEC ? k + 1 : VEC - 1];
                float v = d[k] * c;   // identity rows: no neighbour bits, d = 1
                v -= if_bit(wm, 8 * k + 0, xm);
                v -= if_bit(wm, 8 * k + 1, xp);
                v -= if_bit(wm, 8 * k + 2, q[j].rm.v[k]);
                v -= if_bit(wm, 8 * k + 3, q[j].rp.v[k]);
                v -= if_bit(wm, 8 * k + 4, col[j].v[k]);
                v -= if_bit(wm, 8 * k + 5, col[j + 2].v[k]);
                ap[k] = v;
            }
            finish(j, q[j], ap, d);
        }
    };

    if (a.skip != 0) masked();
    else if (a.any_outflow) unmasked(std::true_type{});
    else unmasked(std::false_type{});

    if constexpr (MODE == 0) {
        double v[1] = {acc0};
        store_block_partials<1, NWAVES>(v, a.partials, nblocks, block);
    } else {
        double v[2] = {acc0, acc1};
        store_block_partials<2, NWAVES>(v, a.partials, nblocks, block);
    }
}

template <int TXV, int TY, int ZR, int VEC, int MODE, bool JACOBI>
__global__ __launch_bounds__(TXV* TY) void pcg_stencil_kernel(PcgArgs a, const float* __restrict__ p, float* __restrict__ e, float* __restrict__ r,
                                                              const float* __restrict__ zold) {
    stencil_body<TXV, TY, ZR, VEC, MODE, JACOBI>(a, p, e, r, zold, blockIdx.x);
}

// alpha = r.z / p.Ap from the partials of pass 2.
__device__ __forceinline__ void alpha_body(const PcgArgs& a, int nblocks) {
    PcgScalars* __restrict__ S = a.s;
    if (S->done) return;
    double tot[1];
    sum_partials<1>(a.partials, nblocks, tot);
    if (threadIdx.x == 0) S->alpha = guarded(S->rz, tot[0], S->fixed_mode != 0);
}

// beta = r.z_new / r.z_old, iteration count and stop test from the partials
// of pass 3.
__device__ __forceinline__ void beta_body(const PcgArgs& a, int nblocks) {
    PcgScalars* __restrict__ S = a.s;
    if (S->done) return;
    double tot[2];
    sum_partials<2>(a.partials, nblocks, tot);
    if (threadIdx.x == 0) {
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

__global__ __launch_bounds__(kFinalizeThreads) void pcg_alpha_kernel(PcgArgs a, int nblocks) { alpha_body(a, nblocks); }
__global__ __launch_bounds__(kFinalizeThreads) void pcg_beta_kernel(PcgArgs a, int nblocks) { beta_body(a, nblocks); }

// ---- the pass bodies as functors of batched launches over many contexts
// (batched.hpp): each context's blocks run with its own numbering
template <bool JACOBI>
struct update_p_k {
    __device__ __forceinline__ static void run(const VBlock& vb, PcgArgs a, const float* r, float* p) { update_p_body<JACOBI>(a, r, p, int(vb.idx.x)); }
};

template <int TXV, int TY, int ZR, int VEC, int MODE, bool JACOBI>
struct stencil_k {
    __device__ __forceinline__ static void run(const VBlock& vb, PcgArgs a, const float* p, float* e, float* r) {
        stencil_body<TXV, TY, ZR, VEC, MODE, JACOBI>(a, p, e, r, nullptr, int(vb.idx.x));
    }
};

struct alpha_k {
    __device__ __forceinline__ static void run(const VBlock&, PcgArgs a, int nblocks) { alpha_body(a, nblocks); }
};

struct beta_k {
    __device__ __forceinline__ static void run(const VBlock&, PcgArgs a, int nblocks) { beta_body(a, nblocks); }
};

// the scalar block of one context into the batch's array (one thread)
struct gather_scalars_k {
    __device__ __forceinline__ static void run(const VBlock&, const PcgScalars* src, PcgScalars* dst) {
        if (threadIdx.x == 0) *dst = *src;
    }
};

struct zero64_k {
    __device__ __forceinline__ static void run(const VBlock& vb, double* x, int64_t n) {
        const int64_t i = int64_t(vb.idx.x) * 256 + threadIdx.x;
        if (i < n) x[i] = 0.0;
    }
};

struct copy64_k {
    __device__ __forceinline__ static void run(const VBlock& vb, const double* src, double* dst, int64_t n) {
        const int64_t i = int64_t(vb.idx.x) * 256 + threadIdx.x;
        if (i < n) dst[i] = src[i];
    }
};

// ---- flexible PCG (learned preconditioner): z is a stored vector
// p <- z (first) or z + beta p
__global__ __launch_bounds__(256) void fpcg_direction_kernel(PcgArgs a, const float* __restrict__ z, float* __restrict__ p) {
    const PcgScalars* __restrict__ S = a.s;
    if (S->done) return;
    const bool first = S->first != 0;
    const float beta = first ? 0.0f : float(S->beta);
    const int64_t cells = int64_t(a.n[0]) * a.n[1] * a.n[2];
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < cells; i += int64_t(gridDim.x) * 256)
        p[i] = first ? z[i] : z[i] + beta * p[i];
}

// partial r.z per block
__global__ __launch_bounds__(256) void fpcg_dot_kernel(PcgArgs a, const float* __restrict__ r, const float* __restrict__ z) {
    const int64_t cells = int64_t(a.n[0]) * a.n[1] * a.n[2];
    double s = 0.0;
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < cells; i += int64_t(gridDim.x) * 256) s += double(r[i]) * double(z[i]);
    double v[1] = {s};
    store_block_partials<1, 4>(v, a.partials, int(gridDim.x), blockIdx.x);
}

// after the update pass: r_(k+1).z_k, ||r||^2, iteration count, stop test
__global__ __launch_bounds__(kFinalizeThreads) void fpcg_update_finalize_kernel(PcgArgs a, int nblocks) {
    PcgScalars* __restrict__ S = a.s;
    if (S->done) return;
    double tot[2];
    sum_partials<2>(a.partials, nblocks, tot);
    if (threadIdx.x == 0) {
        S->rz_old_new = tot[0];
        S->rr = tot[1];
        S->first = 0;
        const int it = S->iteration + 1;
        S->iteration = it;
        if (tot[1] <= S->stop2 || it >= S->max_iterations) S->done = 1;
    }
}

// after z = M(r): r.z; beta = (r.z - r.z_old) / rz_old (Polak-Ribiere), or a
// fresh start (first) that only records r.z
__global__ __launch_bounds__(kFinalizeThreads) void fpcg_beta_kernel(PcgArgs a, int nblocks, int start) {
    PcgScalars* __restrict__ S = a.s;
    if (S->done) return;
    double tot[1];
    sum_partials<1>(a.partials, nblocks, tot);
    if (threadIdx.x == 0) {
        if (start) {
            S->first = 1;
            S->beta = 0.0;
        } else {
            S->beta = guarded(tot[0] - S->rz_old_new, S->rz, false);
        }
        S->rz = tot[0];
    }
}

// r = b - A x' with x' = x_old + e_add (either may be absent) evaluated in
// FP64; optionally stores x' (FP64) and starts a fresh inner solve.
__device__ __forceinline__ void residual_body(const VBlock& vb, PcgArgs a, uint3 tiles, const float* __restrict__ b, const double* __restrict__ x_old,
                                                       const float* __restrict__ e_add, double* __restrict__ x_new, float* __restrict__ r,
                                                       const PcgScalars* __restrict__ S) {
    // tiles of 64 (x) x 4 (y) columns by kResidualPlanes z planes, in the
    // XCD-aware order (neighbouring tiles' rows are read through one L2)
    uint3 tile;
    unsigned id;
    if (!xcd_tile(vb, tiles, tile, id)) return;
    const int nblocks = int(tiles.x * tiles.y * tiles.z);
    const int bid = int(id);
    const int x = int(tile.x) * 64 + (threadIdx.x & 63);
    const int y = int(tile.y) * 4 + (threadIdx.x >> 6);
    double rr = 0.0, rz = 0.0, bb = 0.0;
    auto xv = [&](int cx, int cy, int cz) -> double {
        if (!resolve(a, cx, cy, cz)) return 0.0;
        const int64_t i = lin(a, cx, cy, cz);
        double s = x_old ? x_old[i] : 0.0;
        if (e_add) s += double(e_add[i]);
        return s;
    };
    const int zb = int(tile.z) * kResidualPlanes;
    const int ze = min(zb + kResidualPlanes, a.n[2]);
    for (int z = zb; z < ze && x < a.n[0] && y < a.n[1]; ++z) {
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
        rr += r64 * r64;
        rz += double(rf) * double(zf);
        bb += bv * bv;
    }
    double v[3] = {rr, rz, bb};
    store_block_partials<3, 4>(v, a.partials, nblocks, bid);
}

__global__ __launch_bounds__(256) void residual_kernel(PcgArgs a, uint3 tiles, const float* __restrict__ b, const double* __restrict__ x_old,
                                                       const float* __restrict__ e_add, double* __restrict__ x_new, float* __restrict__ r,
                                                       const PcgScalars* __restrict__ S) { residual_body(hw_block(), a, tiles, b, x_old, e_add, x_new, r, S); }

struct residual_k {
    template <class... A>
    __device__ __forceinline__ static void run(const VBlock& vb, A... a) { residual_body(vb, a...); }
};

// The same pass with float4 rows (nx % 4 == 0): a thread owns 4 consecutive
// x cells of a row and marches kResidualPlanes z planes, keeping the rows at
// z - 1, z and z + 1 of x' in registers; the rows at y +- 1 are vector loads,
// the cells at x - 1 and x + 4 scalar loads.  Every cell subtracts its
// neighbours in the order of residual_kernel, so r and x' are the same; only
// the grouping of the FP64 partial sums differs.
__device__ __forceinline__ void residual4_body(const VBlock& vb, PcgArgs a, uint3 tiles, const float* __restrict__ b, const double* __restrict__ x_old,
                                                        const float* __restrict__ e_add, double* __restrict__ x_new, float* __restrict__ r,
                                                        const PcgScalars* __restrict__ S) {
    uint3 tile;
    unsigned id;
    if (!xcd_tile(vb, tiles, tile, id)) return;
    const int nblocks = int(tiles.x * tiles.y * tiles.z);
    const int nx = a.n[0], ny = a.n[1], nz = a.n[2];
    const int x0 = (int(tile.x) * 64 + (threadIdx.x & 63)) * 4;
    const int y = int(tile.y) * 4 + (threadIdx.x >> 6);
    double rr = 0.0, rz = 0.0, bb = 0.0;
    // x' = x_old + e_add of the 4 cells of row (yy, zz) (an in-domain row)
    auto row = [&](int yy, int zz, double (&v)[4]) {
        const int64_t i = lin(a, x0, yy, zz);
        if (x_old) {
            const double2 p = *reinterpret_cast<const double2*>(x_old + i);
            const double2 q = *reinterpret_cast<const double2*>(x_old + i + 2);
            v[0] = p.x, v[1] = p.y, v[2] = q.x, v[3] = q.y;
        } else {
            v[0] = v[1] = v[2] = v[3] = 0.0;
        }
        if (e_add) {
            const float4 e = *reinterpret_cast<const float4*>(e_add + i);
            v[0] += double(e.x), v[1] += double(e.y), v[2] += double(e.z), v[3] += double(e.w);
        }
    };
    auto cell = [&](int xx, int yy, int zz) -> double {
        const int64_t i = lin(a, xx, yy, zz);
        double s = x_old ? x_old[i] : 0.0;
        if (e_add) s += double(e_add[i]);
        return s;
    };
    const int zb = int(tile.z) * kResidualPlanes;
    const int ze = min(zb + kResidualPlanes, nz);
    if (x0 < nx && y < ny && zb < ze) {
        // rows of a neighbour outside a closed axis are some existing row:
        // no flag bit selects them
        const int ym = neighbour_row(y - 1, ny, a.periodic[1]), yp = neighbour_row(y + 1, ny, a.periodic[1]);
        const int xm = neighbour_row(x0 - 1, nx, a.periodic[0]), xp = neighbour_row(x0 + 4, nx, a.periodic[0]);
        double zm[4], c[4], zp[4];
        row(y, neighbour_row(zb - 1, nz, a.periodic[2]), zm);
        row(y, zb, c);
        for (int z = zb; z < ze; ++z) {
            row(y, neighbour_row(z + 1, nz, a.periodic[2]), zp);
            double ym4[4], yp4[4];
            row(ym, z, ym4);
            row(yp, z, yp4);
            const double left = cell(xm, y, z), right = cell(xp, y, z);
            const int64_t i = lin(a, x0, y, z);
            const uchar4 f4 = *reinterpret_cast<const uchar4*>(a.flags + i);
            const float4 b4 = *reinterpret_cast<const float4*>(b + i);
            const uint8_t fs[4] = {f4.x, f4.y, f4.z, f4.w};
            const float bs[4] = {b4.x, b4.y, b4.z, b4.w};
            float rs[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const uint8_t f = fs[k];
                const double xc = c[k];
                double ax;
                float d = 1.0f;
                if (f & kIdentity) {
                    ax = xc;
                } else {
                    d = diag_of(a, f, x0 + k, y, z);
                    const unsigned m = unsigned(f) & ~a.skip;
                    ax = double(d) * xc;
                    if (m & 1u) ax -= k == 0 ? left : c[k - 1];
                    if (m & 2u) ax -= k == 3 ? right : c[k + 1];
                    if (m & 4u) ax -= ym4[k];
                    if (m & 8u) ax -= yp4[k];
                    if (m & 16u) ax -= zm[k];
                    if (m & 32u) ax -= zp[k];
                }
                const double bv = double(bs[k]);
                const double r64 = bv - ax;
                const float rf = float(r64);
                rs[k] = rf;
                const float zf = S->jacobi ? rf / d : rf;
                rr += r64 * r64;
                rz += double(rf) * double(zf);
                bb += bv * bv;
            }
            *reinterpret_cast<float4*>(r + i) = make_float4(rs[0], rs[1], rs[2], rs[3]);
            if (x_new) {
                *reinterpret_cast<double2*>(x_new + i) = make_double2(c[0], c[1]);
                *reinterpret_cast<double2*>(x_new + i + 2) = make_double2(c[2], c[3]);
            }
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                zm[k] = c[k];
                c[k] = zp[k];
            }
        }
    }
    double v[3] = {rr, rz, bb};
    store_block_partials<3, 4>(v, a.partials, nblocks, int(id));
}

__global__ __launch_bounds__(256) void residual4_kernel(PcgArgs a, uint3 tiles, const float* __restrict__ b, const double* __restrict__ x_old,
                                                        const float* __restrict__ e_add, double* __restrict__ x_new, float* __restrict__ r,
                                                        const PcgScalars* __restrict__ S) { residual4_body(hw_block(), a, tiles, b, x_old, e_add, x_new, r, S); }

struct residual4_k {
    template <class... A>
    __device__ __forceinline__ static void run(const VBlock& vb, A... a) { residual4_body(vb, a...); }
};

// Totals of the residual kernel: true ||r||^2, r.z, ||b||^2; starts a fresh
// inner solve (p = z at the next iteration) with its stop threshold.
__device__ __forceinline__ void residual_finalize_body(const VBlock& vb, PcgArgs a, int nblocks, int compute_bb, int reset_iterations) {
    PcgScalars* __restrict__ S = a.s;
    double tot[3];
    sum_partials<3>(a.partials, nblocks, tot);
    if (threadIdx.x == 0) {
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

__global__ __launch_bounds__(kFinalizeThreads) void residual_finalize_kernel(PcgArgs a, int nblocks, int compute_bb, int reset_iterations) { residual_finalize_body(hw_block(), a, nblocks, compute_bb, reset_iterations); }

struct residual_finalize_k {
    template <class... A>
    __device__ __forceinline__ static void run(const VBlock& vb, A... a) { residual_finalize_body(vb, a...); }
};

// Warm start extrapolated from the two previous solutions: g = 2 x_n - x_{n-1}.
__device__ __forceinline__ void extrapolate_body(const VBlock& vb, const double* __restrict__ x, const double* __restrict__ prev, double* __restrict__ g, int64_t n) {
    const int64_t i = int64_t(vb.idx.x) * blockDim.x + threadIdx.x;
    if (i < n) g[i] = 2.0 * x[i] - prev[i];
}

__global__ void extrapolate_kernel(const double* __restrict__ x, const double* __restrict__ prev, double* __restrict__ g, int64_t n) { extrapolate_body(hw_block(), x, prev, g, n); }

struct extrapolate_k {
    template <class... A>
    __device__ __forceinline__ static void run(const VBlock& vb, A... a) { extrapolate_body(vb, a...); }
};

__device__ __forceinline__ void pressure_from_x64_body(const VBlock& vb, const double* __restrict__ x, float* __restrict__ p, int64_t n) {
    const int64_t i = int64_t(vb.idx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = float(x[i]);
}

__global__ void pressure_from_x64_kernel(const double* __restrict__ x, float* __restrict__ p, int64_t n) { pressure_from_x64_body(hw_block(), x, p, n); }

struct pressure_from_x64_k {
    template <class... A>
    __device__ __forceinline__ static void run(const VBlock& vb, A... a) { pressure_from_x64_body(vb, a...); }
};

__global__ void copy_kernel(const float4* __restrict__ src, float4* __restrict__ dst, int64_t n) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

// b = -(h^2/dt) div(u*) + sum of OUTFLOW face pressures, 0 on identity rows (B:1304-1328).
__device__ __forceinline__ void pressure_rhs_body(const VBlock& vb, Grid g, const float* __restrict__ us, const float* __restrict__ vs, const float* __restrict__ ws,
                                    const uint8_t* __restrict__ flags, const float* __restrict__ dt_ptr, float* __restrict__ rhs) {
    const int x = vb.idx.x * 64 + (threadIdx.x & 63);
    const int y = vb.idx.y * 4 + (threadIdx.x >> 6);
    const int z = vb.idx.z;
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

__global__ void pressure_rhs_kernel(Grid g, const float* __restrict__ us, const float* __restrict__ vs, const float* __restrict__ ws,
                                    const uint8_t* __restrict__ flags, const float* __restrict__ dt_ptr, float* __restrict__ rhs) { pressure_rhs_body(hw_block(), g, us, vs, ws, flags, dt_ptr, rhs); }

struct pressure_rhs_k {
    template <class... A>
    __device__ __forceinline__ static void run(const VBlock& vb, A... a) { pressure_rhs_body(vb, a...); }
};

PcgArgs make_args(const Context& c) {
    PcgArgs a{};
    for (int k = 0; k < 3; ++k) {
        a.n[k] = c.grid.n[k];
        a.periodic[k] = c.grid.flow_periodic[k];
    }
    for (int f = 0; f < 6; ++f) a.outflow[f] = c.grid.flow_type[f] == kOutflow;
    a.any_outflow = a.outflow[0] | a.outflow[1] | a.outflow[2] | a.outflow[3] | a.outflow[4] | a.outflow[5];
    unsigned skip = 0;
    for (int k = 0; k < 3; ++k) {
        if (!a.periodic[k]) continue;
        if (a.n[k] == 1) skip |= 3u << (2 * k);        // both neighbours are the cell itself
        else if (a.n[k] == 2) skip |= 2u << (2 * k);   // both neighbours are the same cell: one -1
    }
    a.skip = skip;
    a.flags = c.stencil_flags;
    a.s = c.scalars;
    a.partials = c.partials;
    return a;
}

// Candidate launch shapes of the stencil passes: threads along x (each owning
// VEC consecutive cells: float4 rows need nx % 4 == 0), threads along y, z
// planes per tile.  Each grid class has its own candidates; the table
// default (measured on MI355X at 256^3) is 64 x 4 x 2 for pass 2 and
// 64 x 8 x 2 for pass 3 with float4 rows, 16 x 16 x 2 on narrow grids, scalar
// 64 x 4 x 2 when nx % 4 != 0.  Every candidate tile covers at least 64 x 4 x 2
// cells, which bounds the partials buffer.
struct StencilShape {
    int txv, ty, zr, vec;
};

constexpr StencilShape kStencilShapes[] = {
    {64, 4, 2, 4}, {64, 8, 2, 4}, {32, 8, 2, 4}, {64, 4, 4, 4},   // 0-3: float4 rows, nx >= 128
    {16, 16, 2, 4}, {16, 8, 2, 4},                                 // 4-5: float4 rows, narrow
    {64, 4, 2, 1}, {64, 8, 2, 1}, {32, 8, 2, 1},                   // 6-8: scalar rows
};
constexpr int kStencilShapeCount = int(sizeof(kStencilShapes) / sizeof(kStencilShapes[0]));

// [first, last) candidates of this grid's class
void stencil_candidates(const Context& c, int& first, int& last) {
    const bool vec = c.grid.n[0] % 4 == 0;
    if (vec && c.grid.n[0] >= 128) first = 0, last = 4;
    else if (vec) first = 4, last = 6;
    else first = 6, last = 9;
}

int64_t stencil_tiles(const Context& c, const StencilShape& sh) {
    return int64_t((c.grid.n[0] + sh.txv * sh.vec - 1) / (sh.txv * sh.vec)) * ((c.grid.n[1] + sh.ty - 1) / sh.ty) *
           ((c.grid.n[2] + sh.zr - 1) / sh.zr);
}

template <int TXV, int TY, int ZR, int VEC, bool JACOBI>
void launch_stencils(Context& c, PcgArgs a, int mode) {
    a.tiles[0] = (c.grid.n[0] + TXV * VEC - 1) / (TXV * VEC);
    a.tiles[1] = (c.grid.n[1] + TY - 1) / TY;
    a.tiles[2] = (c.grid.n[2] + ZR - 1) / ZR;
    a.tile_div[0] = make_divisor(unsigned(a.tiles[2]));
    a.tile_div[1] = make_divisor(unsigned(a.tiles[1]));
    const unsigned blocks = unsigned(a.tiles[0] * a.tiles[1] * a.tiles[2]);
    float* e = c.options.pressure_mode == SMOKE_SIMULATION_PRESSURE_TOLERANCE ? c.pcg_e : c.pressure;
    if (mode == 0) pcg_stencil_kernel<TXV, TY, ZR, VEC, 0, JACOBI><<<blocks, TXV * TY, 0, c.stream>>>(a, c.pcg_p, e, c.pcg_r, nullptr);
    else if (mode == 1) pcg_stencil_kernel<TXV, TY, ZR, VEC, 1, JACOBI><<<blocks, TXV * TY, 0, c.stream>>>(a, c.pcg_p, e, c.pcg_r, nullptr);
    else pcg_stencil_kernel<TXV, TY, ZR, VEC, 2, false><<<blocks, TXV * TY, 0, c.stream>>>(a, c.pcg_p, e, c.pcg_r, c.pcg_z);
    hip_check(hipGetLastError(), "pcg_stencil_kernel");
    c.pcg_stencil_blocks = int(blocks);
}

template <bool JACOBI>
void stencil_pass_shape(Context& c, int mode, int shape) {
    const PcgArgs a = make_args(c);
    switch (shape) {
        case 0: launch_stencils<64, 4, 2, 4, JACOBI>(c, a, mode); break;
        case 1: launch_stencils<64, 8, 2, 4, JACOBI>(c, a, mode); break;
        case 2: launch_stencils<32, 8, 2, 4, JACOBI>(c, a, mode); break;
        case 3: launch_stencils<64, 4, 4, 4, JACOBI>(c, a, mode); break;
        case 4: launch_stencils<16, 16, 2, 4, JACOBI>(c, a, mode); break;
        case 5: launch_stencils<16, 8, 2, 4, JACOBI>(c, a, mode); break;
        case 6: launch_stencils<64, 4, 2, 1, JACOBI>(c, a, mode); break;
        case 7: launch_stencils<64, 8, 2, 1, JACOBI>(c, a, mode); break;
        case 8: launch_stencils<32, 8, 2, 1, JACOBI>(c, a, mode); break;
        default: throw HipError("invalid stencil shape");
    }
}

template <bool JACOBI>
void stencil_pass(Context& c, int mode) {
    stencil_pass_shape<JACOBI>(c, mode, c.pcg_shape[mode == 0 ? 0 : 1]);
}

// Jacobi kernels: the Jacobi preconditioner, and the learned one's fallback
bool jacobi_on(const Context& c) {
    return c.options.pressure_mode == SMOKE_SIMULATION_PRESSURE_TOLERANCE && c.options.preconditioner != SMOKE_SIMULATION_PRECONDITIONER_NONE;
}

// pass 1: p <- z + beta p
void pass_update_p(Context& c) {
    const PcgArgs a = make_args(c);
    const unsigned blocks = unsigned((c.grid.cells + 1023) / 1024);
    if (jacobi_on(c)) pcg_update_p_kernel<true><<<blocks, 256, 0, c.stream>>>(a, c.pcg_r, c.pcg_p);
    else pcg_update_p_kernel<false><<<blocks, 256, 0, c.stream>>>(a, c.pcg_r, c.pcg_p);
    hip_check(hipGetLastError(), "pcg_update_p_kernel");
}

// passes 2 (mode 0: p.Ap) and 3 (mode 1: e, r update, r.z, r.r)
void pass_stencil(Context& c, int mode) {
    if (jacobi_on(c)) stencil_pass<true>(c, mode);
    else stencil_pass<false>(c, mode);
}

// alpha after pass 2 (mode 0), beta and the stop test after pass 3 (mode 1)
void pass_finalize(Context& c, int mode) {
    const PcgArgs a = make_args(c);
    if (mode == 0) pcg_alpha_kernel<<<1, kFinalizeThreads, 0, c.stream>>>(a, c.pcg_stencil_blocks);
    else pcg_beta_kernel<<<1, kFinalizeThreads, 0, c.stream>>>(a, c.pcg_stencil_blocks);
    hip_check(hipGetLastError(), "pcg finalize kernel");
}

// ---- flexible PCG with the learned preconditioner
// z = M(r)
void learned_apply(Context& c) {
    c.rz_fused_blocks = 0;
    if (c.options.preconditioner == SMOKE_SIMULATION_PRECONDITIONER_MULTILEVEL) {
        c.rz_fused_blocks = multilevel_apply(c, c.pcg_r, c.pcg_z, c.partials, c.partial_capacity);
        return;
    }
    const int dims[3] = {c.grid.n[0], c.grid.n[1], c.grid.n[2]};
    preconditioner_apply(*c.learned, c.pcg_r, c.stencil_flags, dims, c.pcg_z, c.stream);
}

// r.z, then beta (or a fresh Krylov start)
void learned_rz(Context& c, bool start) {
    const PcgArgs a = make_args(c);
    if (c.rz_fused_blocks > 0) {   // the preconditioner's last pass wrote the partials
        fpcg_beta_kernel<<<1, kFinalizeThreads, 0, c.stream>>>(a, c.rz_fused_blocks, start ? 1 : 0);
        hip_check(hipGetLastError(), "flexible pcg beta");
        return;
    }
    const unsigned blocks = unsigned(std::min<size_t>(1024, c.partial_capacity));
    fpcg_dot_kernel<<<blocks, 256, 0, c.stream>>>(a, c.pcg_r, c.pcg_z);
    fpcg_beta_kernel<<<1, kFinalizeThreads, 0, c.stream>>>(a, int(blocks), start ? 1 : 0);
    hip_check(hipGetLastError(), "flexible pcg r.z");
}

// p = z + beta p; alpha = r.z / p.Ap; e += alpha p, r -= alpha Ap with r_(k+1).z_k and ||r||^2
void learned_iteration(Context& c) {
    const PcgArgs a = make_args(c);
    fpcg_direction_kernel<<<unsigned(std::min<int64_t>((c.grid.cells + 255) / 256, 8192)), 256, 0, c.stream>>>(a, c.pcg_z, c.pcg_p);
    hip_check(hipGetLastError(), "fpcg_direction_kernel");
    pass_stencil(c, 0);
    pass_finalize(c, 0);
    pass_stencil(c, 2);
    fpcg_update_finalize_kernel<<<1, kFinalizeThreads, 0, c.stream>>>(a, c.pcg_stencil_blocks);
    hip_check(hipGetLastError(), "fpcg_update_finalize_kernel");
}

void launch_residual(Context& c, const double* x_old, const float* e_add, double* x_new, bool bb, bool reset) {
    const PcgArgs a = make_args(c);
    const bool rows4 = c.grid.n[0] % 4 == 0;
    const dim3 grid((c.grid.n[0] + (rows4 ? 255 : 63)) / (rows4 ? 256 : 64), (c.grid.n[1] + 3) / 4,
                    (c.grid.n[2] + kResidualPlanes - 1) / kResidualPlanes);
    if (rows4)
        residual4_kernel<<<xcd_blocks(grid), 256, 0, c.stream>>>(a, make_uint3(grid.x, grid.y, grid.z), c.pressure_rhs, x_old, e_add, x_new,
                                                                  c.pcg_r, c.scalars);
    else
        residual_kernel<<<xcd_blocks(grid), 256, 0, c.stream>>>(a, make_uint3(grid.x, grid.y, grid.z), c.pressure_rhs, x_old, e_add, x_new,
                                                                 c.pcg_r, c.scalars);
    hip_check(hipGetLastError(), "residual_kernel");
    residual_finalize_kernel<<<1, kFinalizeThreads, 0, c.stream>>>(a, int(grid.x * grid.y * grid.z), bb ? 1 : 0, reset ? 1 : 0);
    hip_check(hipGetLastError(), "residual_finalize_kernel");
}

}  // namespace

void configure_pcg(Context& c) {
    // upper bound of the stencil-pass block count over all candidate shapes
    int64_t tiles_max = 0;
    for (const StencilShape& sh : kStencilShapes) tiles_max = std::max(tiles_max, stencil_tiles(c, sh));
    const int64_t residual_blocks =
        int64_t((c.grid.n[0] + 63) / 64) * ((c.grid.n[1] + 3) / 4) * ((c.grid.n[2] + kResidualPlanes - 1) / kResidualPlanes);
    c.partial_capacity = size_t(std::max<int64_t>(2 * tiles_max, 3 * residual_blocks));
    c.pcg_stencil_blocks = int(tiles_max);
}

void table_pcg_shapes(Context& c) {
    int first, last;
    stencil_candidates(c, first, last);
    c.pcg_shape[0] = first;                                 // 64x4x2 / 16x16x2 / scalar 64x4x2
    c.pcg_shape[1] = first == 0 ? 1 : first;                // 64x8x2 for the update pass on wide grids
    // a float4 tile row of 64 threads covers 256 x cells: below 256 (or not a
    // multiple of it) a 32-thread row (128 cells) leaves fewer threads idle
    if (first == 0 && c.grid.n[0] % 256 != 0) c.pcg_shape[0] = c.pcg_shape[1] = 2;   // 32x8x2
}

void describe_pcg_shape(int index, SmokeSimulationLaunchShape& out) {
    if (index < 0 || index >= kStencilShapeCount) throw HipError("invalid stencil shape");
    const StencilShape& sh = kStencilShapes[index];
    out = {sh.txv, sh.ty, sh.zr, sh.vec};
}

int find_pcg_shape(const Context& c, const SmokeSimulationLaunchShape& shape) {
    int first, last;
    stencil_candidates(c, first, last);
    for (int i = first; i < last; ++i) {
        const StencilShape& sh = kStencilShapes[i];
        if (sh.txv == shape.threads_x && sh.ty == shape.threads_y && sh.zr == shape.threads_z && sh.vec == shape.cells_per_thread) return i;
    }
    throw HipError("PCG launch shape is not a candidate for this grid");
}

// Times the candidate tiles of passes 2 and 3 on this grid and keeps the
// fastest.  The scalars are set so that every launch does its full work and
// leaves the vectors unchanged: not done, no stop test, alpha = 0 (e and r
// are updated by zero) and not the first iteration.  The caller restores the
// solver constants afterwards.
int tune_pcg_shapes(Context& c, SmokeSimulationLaunchReport& report) {
    PcgScalars s{};
    s.fixed_mode = 1;
    s.max_iterations = 1 << 30;
    s.rz = 1.0;
    hip_check(hipStreamSynchronize(c.stream), "sync");
    hip_check(hipMemcpy(c.scalars, &s, sizeof(s), hipMemcpyHostToDevice), "tuning scalars");
    hipEvent_t t0, t1;
    hip_check(hipEventCreate(&t0), "event");
    hip_check(hipEventCreate(&t1), "event");
    constexpr int kReps = 5;
    int first, last, timed = 0;
    stencil_candidates(c, first, last);
    const int table[2] = {c.pcg_shape[0], c.pcg_shape[1]};
    float* out_us[2] = {report.pcg_curvature_us, report.pcg_update_us};
    for (int mode = 0; mode < 2; ++mode) {
        float best_us = 0.0f;
        int best = table[mode];
        for (int shape = first; shape < last; ++shape) {
            auto launch = [&] {
                if (jacobi_on(c)) stencil_pass_shape<true>(c, mode, shape);
                else stencil_pass_shape<false>(c, mode, shape);
            };
            launch();   // warm-up
            hip_check(hipEventRecord(t0, c.stream), "event");
            for (int r = 0; r < kReps; ++r) launch();
            hip_check(hipEventRecord(t1, c.stream), "event");
            hip_check(hipEventSynchronize(t1), "event sync");
            float ms = 0.0f;
            hip_check(hipEventElapsedTime(&ms, t0, t1), "elapsed");
            const float us = 1e3f * ms / kReps;
            ++timed;
            if (shape == table[mode]) out_us[mode][0] = us;
            if (shape == first || us < best_us) best_us = us, best = shape;
        }
        c.pcg_shape[mode] = best;
        out_us[mode][1] = best_us;
    }
    hipEventDestroy(t0);
    hipEventDestroy(t1);
    return timed;
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
    pass_update_p(c);
    pass_stencil(c, 0);
    pass_finalize(c, 0);
    pass_stencil(c, 1);
    pass_finalize(c, 1);
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
    } else if (o.warm_start == 2) {
        // linear extrapolation in time once two solutions exist: the buffer
        // roles rotate so that the history keeps the last solution
        if (c.solves_done >= 2) {
            double* cur = c.x64[c.x64_current];
            double* spare = c.x64[c.x64_current ^ 1];
            extrapolate_kernel<<<unsigned((n + 255) / 256), 256, 0, c.stream>>>(cur, c.x64_hist, spare, n);
            hip_check(hipGetLastError(), "extrapolate_kernel");
            c.x64[c.x64_current ^ 1] = c.x64_hist;
            c.x64_hist = cur;
            c.x64[c.x64_current] = spare;
        } else if (c.solves_done == 1) {
            hip_check(hipMemcpyAsync(c.x64_hist, c.x64[c.x64_current], size_t(n) * sizeof(double), hipMemcpyDeviceToDevice, c.stream),
                      "warm-start history");
        }
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
    // flexible PCG: the network on the pressure grid, or the multilevel cycle
    // (which fails loudly when its coarse solver needs a network and none is attached)
    const bool multilevel = o.preconditioner == SMOKE_SIMULATION_PRECONDITIONER_MULTILEVEL && c.pcg_z;
    bool learned = (o.preconditioner == SMOKE_SIMULATION_PRECONDITIONER_LEARNED && c.learned && c.pcg_z) || multilevel;
    int learned_iterations = 0;
    bool fell_back = false;
    // Inner solve with the learned (or multilevel) preconditioner: flexible
    // PCG, convergence checked after every iteration (an application of the
    // preconditioner costs more than the host round trip).  Returns false when the residual fell
    // by less than 10 % over 5 iterations or turned non-finite.
    // Up to two iterations stay in flight: iteration k's scalars are copied
    // to a pinned slot and signalled by an event, and the host judges
    // iteration k while iteration k + 1 runs.  Each iteration is enqueued
    // with the preconditioner application that prepares the next one (z =
    // M(r), r.z, beta), which costs most of an iteration's device time and
    // is wasted after the iteration that ends the solve.  So an iteration
    // predicted to end the solve (r.r falling by the last ratio again for
    // each iteration ahead comes within a factor 4 of the stop threshold, or
    // the iteration cap) is enqueued without its application and with
    // nothing after it; if the judgement then finds it did not end the
    // solve, its application is enqueued at that point (one host round trip
    // with the device idle).  Every iteration's device work is unchanged
    // (same iterations, same bits); only applications whose z is never used
    // are left out.
    auto learned_inner = [&]() -> bool {
        learned_apply(c);
        learned_rz(c, true);
        std::vector<double> rr{host.rr};
        struct Pending {
            int slot;
            bool applied;
        };
        Pending pending[2] = {};
        int in_flight = 0, next_slot = 0;
        auto enqueue = [&](bool apply) {
            const int s = next_slot;
            next_slot ^= 1;
            learned_iteration(c);
            hip_check(hipMemcpyAsync(&c.host_scalars[s], c.scalars, sizeof(PcgScalars), hipMemcpyDeviceToHost, c.stream), "scalars copy");
            hip_check(hipEventRecord(c.status_event[s], c.stream), "status event");
            if (apply) {
                learned_apply(c);
                learned_rz(c, false);
            }
            pending[in_flight++] = Pending{s, apply};
        };
        // is the iteration `ahead` iterations after the last judged one
        // predicted to reach the stop test (or the iteration cap)?
        auto predicted_end = [&](int ahead) {
            if (host.iteration + ahead >= host.max_iterations) return true;
            const size_t k = rr.size() - 1;
            // the last ratio of this solve, or the previous solve's mean one
            double ratio = c.fpcg_ratio;
            if (k >= 1) ratio = rr[k - 1] > 0.0 ? rr[k] / rr[k - 1] : 1.0;
            if (!(ratio > 0.0 && ratio < 1.0)) return false;
            double next = rr[k];
            for (int q = 0; q < ahead; ++q) next *= ratio;
            return next <= 4.0 * host.stop2;
        };
        // up to two in flight; nothing after an iteration enqueued without
        // its application or predicted to end the solve
        auto fill = [&]() {
            while (in_flight < 2) {
                if (in_flight > 0 && (!pending[in_flight - 1].applied || predicted_end(in_flight))) break;
                enqueue(!predicted_end(in_flight + 1));
            }
        };
        fill();
        for (;;) {
            const Pending judged = pending[0];
            pending[0] = pending[1];
            --in_flight;
            hip_check(hipEventSynchronize(c.status_event[judged.slot]), "status wait");
            host = c.host_scalars[judged.slot];
            ++learned_iterations;
            if (!std::isfinite(host.rr) || !std::isfinite(host.rz)) return false;
            rr.push_back(host.rr);
            if (host.done) {
                const size_t its = rr.size() - 1;
                if (its >= 1 && rr[0] > 0.0 && host.rr > 0.0) c.fpcg_ratio = std::pow(host.rr / rr[0], 1.0 / double(its));
                return true;
            }
            const size_t k = rr.size() - 1;
            if (!multilevel) {
                if (k >= 5 && rr[k] > 0.81 * rr[k - 5]) return false;
            } else if (k >= 20) {
                // The V-cycle is a convergent linear map, but ||r|| under it
                // is not monotone: the near-null mode of the anchored
                // operator (the pressure level) is built over the first
                // iterations while ||r|| grows.  Stagnation here means the
                // best residual of the last 10 iterations is not 10 % below
                // the best one before them.
                const double recent = *std::min_element(rr.end() - 10, rr.end());
                const double before = *std::min_element(rr.begin(), rr.end() - 10);
                if (recent > 0.81 * before) return false;
            }
            if (!judged.applied) {
                // predicted to end the solve, and it did not
                learned_apply(c);
                learned_rz(c, false);
            }
            fill();
        }
    };
    while (!host.done) {
        if (learned) {
            if (!learned_inner()) {
                // the rest of this solve runs Jacobi PCG; a non-finite
                // correction is dropped before the refinement adds it
                learned = false;
                fell_back = true;
                if (!std::isfinite(host.rr) || !std::isfinite(host.rz))
                    hip_check(hipMemsetAsync(c.pcg_e, 0, size_t(n) * sizeof(float), c.stream), "drop correction");
            }
            enqueue_pcg_refine(c);
            ++passes;
            read_scalars();
            if (host.done) break;
            if (passes > o.max_refinements + (fell_back ? 1 : 0) || host.iteration >= o.max_iterations) break;
            continue;
        }
        // inner solve: chunks of `chunk` iterations, at most two in flight; in
        // timing mode the first iteration of every chunk is bracketed by events
        int slot = 0;
        auto enqueue_chunk = [&](int s) {
            for (int k = 0; k < chunk; ++k) {
                if (c.timing && k == 0) {
                    hipEvent_t* ev = c.timing_events[s];
                    hip_check(hipEventRecord(ev[0], c.stream), "event");
                    pass_update_p(c);
                    hip_check(hipEventRecord(ev[1], c.stream), "event");
                    pass_stencil(c, 0);
                    hip_check(hipEventRecord(ev[2], c.stream), "event");
                    pass_finalize(c, 0);
                    hip_check(hipEventRecord(ev[3], c.stream), "event");
                    pass_stencil(c, 1);
                    hip_check(hipEventRecord(ev[4], c.stream), "event");
                    pass_finalize(c, 1);
                    hip_check(hipEventRecord(ev[5], c.stream), "event");
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
                float m[5] = {};
                for (int q = 0; q < 5; ++q)
                    hip_check(hipEventElapsedTime(&m[q], c.timing_events[slot][q], c.timing_events[slot][q + 1]), "elapsed");
                c.timing_totals.pcg_direction_ms += m[0];
                c.timing_totals.pcg_curvature_ms += m[1];
                c.timing_totals.pcg_update_ms += m[3];
                c.timing_totals.pcg_reduction_ms += m[2] + m[4];
                c.timing_totals.pcg_timed_iterations += 1;
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
    c.solves_done = o.warm_start == 2 ? std::min(c.solves_done + 1, 2) : 1;
    c.stats.iterations = host.iteration;
    c.stats.refinements = passes > 0 ? passes - 1 : 0;
    c.stats.rhs_norm = std::sqrt(host.bb);
    c.stats.relative_residual = host.bb > 0 ? std::sqrt(host.rr_true / host.bb) : 0.0;
    c.stats.converged = host.bb > 0 ? (std::sqrt(host.rr_true) <= o.relative_tolerance * std::sqrt(host.bb)) : 1;
    c.stats.learned_iterations = learned_iterations;
    c.stats.learned_fallback = fell_back ? 1 : 0;
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
    const unsigned blocks = unsigned((n4 + 255) / 256);
    copy_kernel<<<blocks, 256, 0, s>>>(a, b, n4);   // warm-up
    hip_check(hipEventRecord(t0, s), "event");
    for (int r = 0; r < reps; ++r) copy_kernel<<<blocks, 256, 0, s>>>(a, b, n4);
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

// ---- batched solves (batch.hip): the pressure stage of many contexts as
// one launch per pass.  Static launches (right-hand side, the iteration
// passes, the fixed-mode start, the scalar gather) are built once per batch
// plan; the launches that name a context's current FP64 solution buffers are
// built per solve phase.
namespace {

template <int TXV, int TY, int ZR, int VEC, int MODE, bool JACOBI>
void stencil_group(const std::vector<Context*>& P, const std::vector<unsigned>& members, TableArena& A, std::vector<BatchedLaunch>& out) {
    LaunchBuilder<stencil_k<TXV, TY, ZR, VEC, MODE, JACOBI>, TXV * TY, PcgArgs, const float*, float*, float*> L;
    for (unsigned p : members) {
        const Context& c = *P[p];
        PcgArgs a = make_args(c);
        a.tiles[0] = (c.grid.n[0] + TXV * VEC - 1) / (TXV * VEC);
        a.tiles[1] = (c.grid.n[1] + TY - 1) / TY;
        a.tiles[2] = (c.grid.n[2] + ZR - 1) / ZR;
        a.tile_div[0] = make_divisor(unsigned(a.tiles[2]));
        a.tile_div[1] = make_divisor(unsigned(a.tiles[1]));
        float* e = c.options.pressure_mode == SMOKE_SIMULATION_PRESSURE_TOLERANCE ? c.pcg_e : c.pressure;
        L.add(p, dim3(unsigned(a.tiles[0] * a.tiles[1] * a.tiles[2])), a, c.pcg_p, e, c.pcg_r);
    }
    if (!L.empty()) out.push_back(L.finish(A, dim3(TXV * TY)));
}

template <int MODE, bool JACOBI>
void stencil_groups(const std::vector<Context*>& P, TableArena& A, std::vector<BatchedLaunch>& out) {
    for (int shape = 0; shape < kStencilShapeCount; ++shape) {
        std::vector<unsigned> m;
        for (unsigned p = 0; p < P.size(); ++p)
            if (P[p]->pcg_shape[MODE] == shape && jacobi_on(*P[p]) == JACOBI) m.push_back(p);
        if (m.empty()) continue;
        switch (shape) {
            case 0: stencil_group<64, 4, 2, 4, MODE, JACOBI>(P, m, A, out); break;
            case 1: stencil_group<64, 8, 2, 4, MODE, JACOBI>(P, m, A, out); break;
            case 2: stencil_group<32, 8, 2, 4, MODE, JACOBI>(P, m, A, out); break;
            case 3: stencil_group<64, 4, 4, 4, MODE, JACOBI>(P, m, A, out); break;
            case 4: stencil_group<16, 16, 2, 4, MODE, JACOBI>(P, m, A, out); break;
            case 5: stencil_group<16, 8, 2, 4, MODE, JACOBI>(P, m, A, out); break;
            case 6: stencil_group<64, 4, 2, 1, MODE, JACOBI>(P, m, A, out); break;
            case 7: stencil_group<64, 8, 2, 1, MODE, JACOBI>(P, m, A, out); break;
            case 8: stencil_group<32, 8, 2, 1, MODE, JACOBI>(P, m, A, out); break;
            default: throw HipError("invalid stencil shape");
        }
    }
}

template <bool JACOBI>
void update_group(const std::vector<Context*>& P, TableArena& A, std::vector<BatchedLaunch>& out) {
    LaunchBuilder<update_p_k<JACOBI>, 256, PcgArgs, const float*, float*> L;
    for (unsigned p = 0; p < P.size(); ++p) {
        const Context& c = *P[p];
        if (jacobi_on(c) != JACOBI) continue;
        L.add(p, dim3(unsigned((c.grid.cells + 1023) / 1024)), make_args(c), c.pcg_r, c.pcg_p);
    }
    if (!L.empty()) out.push_back(L.finish(A, dim3(256)));
}

int stencil_blocks_of(const Context& c, int mode) {
    return int(stencil_tiles(c, kStencilShapes[c.pcg_shape[mode]]));
}

// r = b - A (x_old + e_add) in FP64 (x_new = that sum when given), then the
// totals; per context the residual kernel of its row width
void residual_launches(const std::vector<Context*>& P, const std::vector<const double*>& x_old, const std::vector<const float*>& e_add,
                       const std::vector<double*>& x_new, bool bb, bool reset, TableArena& A, std::vector<BatchedLaunch>& out) {
    LaunchBuilder<residual4_k, 256, PcgArgs, uint3, const float*, const double*, const float*, double*, float*, const PcgScalars*> R4;
    LaunchBuilder<residual_k, 256, PcgArgs, uint3, const float*, const double*, const float*, double*, float*, const PcgScalars*> R1;
    LaunchBuilder<residual_finalize_k, kFinalizeThreads, PcgArgs, int, int, int> F;
    for (unsigned p = 0; p < P.size(); ++p) {
        const Context& c = *P[p];
        const PcgArgs a = make_args(c);
        const bool rows4 = c.grid.n[0] % 4 == 0;
        const dim3 grid((c.grid.n[0] + (rows4 ? 255 : 63)) / (rows4 ? 256 : 64), (c.grid.n[1] + 3) / 4,
                        (c.grid.n[2] + kResidualPlanes - 1) / kResidualPlanes);
        const uint3 tiles = make_uint3(grid.x, grid.y, grid.z);
        if (rows4) R4.add(p, dim3(xcd_blocks(grid)), a, tiles, c.pressure_rhs, x_old[p], e_add[p], x_new[p], c.pcg_r, c.scalars);
        else R1.add(p, dim3(xcd_blocks(grid)), a, tiles, c.pressure_rhs, x_old[p], e_add[p], x_new[p], c.pcg_r, c.scalars);
        F.add(p, dim3(1), a, int(grid.x * grid.y * grid.z), bb ? 1 : 0, reset ? 1 : 0);
    }
    if (!R4.empty()) out.push_back(R4.finish(A, dim3(256)));
    if (!R1.empty()) out.push_back(R1.finish(A, dim3(256)));
    out.push_back(F.finish(A, dim3(kFinalizeThreads)));
}

}  // namespace

void build_batched_solve(const std::vector<Context*>& P, PcgScalars* gathered, TableArena& A, BatchedSolve& out) {
    out = BatchedSolve{};
    {
        LaunchBuilder<pressure_rhs_k, 256, Grid, const float*, const float*, const float*, const uint8_t*, const float*, float*> L;
        for (unsigned p = 0; p < P.size(); ++p) {
            const Context& c = *P[p];
            L.add(p, dim3((c.grid.n[0] + 63) / 64, (c.grid.n[1] + 3) / 4, c.grid.n[2]), c.grid, c.face_star[0], c.face_star[1], c.face_star[2],
                  c.stencil_flags, c.dt_device, c.pressure_rhs);
        }
        out.rhs.push_back(L.finish(A, dim3(256)));
    }
    // one iteration: p update, p.Ap, alpha, e / r update, beta
    update_group<false>(P, A, out.iteration);
    update_group<true>(P, A, out.iteration);
    stencil_groups<0, false>(P, A, out.iteration);
    stencil_groups<0, true>(P, A, out.iteration);
    {
        LaunchBuilder<alpha_k, kFinalizeThreads, PcgArgs, int> L;
        for (unsigned p = 0; p < P.size(); ++p) L.add(p, dim3(1), make_args(*P[p]), stencil_blocks_of(*P[p], 0));
        out.iteration.push_back(L.finish(A, dim3(kFinalizeThreads)));
    }
    stencil_groups<1, false>(P, A, out.iteration);
    stencil_groups<1, true>(P, A, out.iteration);
    {
        LaunchBuilder<beta_k, kFinalizeThreads, PcgArgs, int> L;
        for (unsigned p = 0; p < P.size(); ++p) L.add(p, dim3(1), make_args(*P[p]), stencil_blocks_of(*P[p], 1));
        out.iteration.push_back(L.finish(A, dim3(kFinalizeThreads)));
    }
    // fixed mode starts from p = 0: r = b
    {
        const size_t n = P.size();
        residual_launches(P, std::vector<const double*>(n, nullptr), std::vector<const float*>(n, nullptr), std::vector<double*>(n, nullptr), true,
                          true, A, out.fixed_start);
    }
    {
        LaunchBuilder<gather_scalars_k, 64, const PcgScalars*, PcgScalars*> L;
        for (unsigned p = 0; p < P.size(); ++p) L.add(p, dim3(1), P[p]->scalars, gathered + p);
        out.gather.push_back(L.finish(A, dim3(64)));
    }
}

void build_batched_warm_start(const std::vector<Context*>& P, TableArena& A, std::vector<BatchedLaunch>& out) {
    // the per-context logic of run_tolerance_pcg's start, buffer roles included
    LaunchBuilder<zero64_k, 256, double*, int64_t> Z;
    LaunchBuilder<extrapolate_k, 256, const double*, const double*, double*, int64_t> E;
    LaunchBuilder<copy64_k, 256, const double*, double*, int64_t> Cp;
    for (unsigned p = 0; p < P.size(); ++p) {
        Context& c = *P[p];
        const int64_t n = c.grid.cells;
        const dim3 grid(unsigned((n + 255) / 256));
        if (!c.options.warm_start) {
            Z.add(p, grid, c.x64[c.x64_current], n);
        } else if (c.options.warm_start == 2) {
            if (c.solves_done >= 2) {
                double* cur = c.x64[c.x64_current];
                double* spare = c.x64[c.x64_current ^ 1];
                E.add(p, grid, cur, c.x64_hist, spare, n);
                c.x64[c.x64_current ^ 1] = c.x64_hist;
                c.x64_hist = cur;
                c.x64[c.x64_current] = spare;
            } else if (c.solves_done == 1) {
                Cp.add(p, grid, c.x64[c.x64_current], c.x64_hist, n);
            }
        }
    }
    if (!Z.empty()) out.push_back(Z.finish(A, dim3(256)));
    if (!E.empty()) out.push_back(E.finish(A, dim3(256)));
    if (!Cp.empty()) out.push_back(Cp.finish(A, dim3(256)));
}

void build_batched_residual(const std::vector<Context*>& P, bool refine, TableArena& A, std::vector<BatchedLaunch>& out) {
    const size_t n = P.size();
    std::vector<const double*> xo(n);
    std::vector<const float*> e(n, nullptr);
    std::vector<double*> xn(n, nullptr);
    for (size_t p = 0; p < n; ++p) {
        xo[p] = P[p]->x64[P[p]->x64_current];
        if (refine) {
            e[p] = P[p]->pcg_e;
            xn[p] = P[p]->x64[P[p]->x64_current ^ 1];
        }
    }
    residual_launches(P, xo, e, xn, !refine, !refine, A, out);
}

void build_batched_zero_x64(const std::vector<Context*>& P, TableArena& A, std::vector<BatchedLaunch>& out) {
    LaunchBuilder<zero64_k, 256, double*, int64_t> Z;
    for (unsigned p = 0; p < P.size(); ++p) Z.add(p, dim3(unsigned((P[p]->grid.cells + 255) / 256)), P[p]->x64[P[p]->x64_current], P[p]->grid.cells);
    out.push_back(Z.finish(A, dim3(256)));
}

void build_batched_pressure_from_x64(const std::vector<Context*>& P, TableArena& A, std::vector<BatchedLaunch>& out) {
    LaunchBuilder<pressure_from_x64_k, 256, const double*, float*, int64_t> L;
    for (unsigned p = 0; p < P.size(); ++p) {
        const Context& c = *P[p];
        L.add(p, dim3(unsigned((c.grid.cells + 255) / 256)), c.x64[c.x64_current], c.pressure, c.grid.cells);
    }
    out.push_back(L.finish(A, dim3(256)));
}

}  // namespace smoke
