// Stage kernels of the smoke step other than the pressure solve.
//
// Face kernels never write the domain end faces (non-periodic faces 0 and n,
// and face n of a periodic axis); a per-line boundary pass writes them after
// the interior is final, which gives the race-free reading of the reference's
// enforce / project passes (B:1014-1189, B:1425-1600): OUTFLOW copies the
// finished neighbour face, walls take their prescribed normal velocity, a
// periodic face n copies face 0.
#include "batched.hpp"
#include "context.hpp"

namespace smoke {
namespace {

constexpr int kBlock = 256;

inline dim3 grid3(int nx, int ny, int nz, dim3 b) { return dim3((nx + b.x - 1) / b.x, (ny + b.y - 1) / b.y, (nz + b.z - 1) / b.z); }
inline unsigned grid1(int64_t n) { return unsigned((n + kBlock - 1) / kBlock); }
const dim3 kCellBlock(64, 4, 1);

// Candidate launches of the advection kernels: the blocks of the gather form
// (threads along x, y, z; every thread computes one output independently, so
// the block shape changes only the speed), then the LDS-staged form
// (kAdvectStaged).  Measured on MI355X at 256^3 (12-sphere plume state after
// 30 steps, cubic advection): the staged form takes 0.44, 0.45, 0.45 and
// 0.75 ms for u, v, w and the scalars; the best gather block 0.50 ms for u
// (16 x 4 x 4) and about twice the staged time for the others.
const dim3 kAdvectBlocks[] = {dim3(64, 4, 1), dim3(32, 8, 1), dim3(16, 16, 1), dim3(32, 4, 2), dim3(64, 2, 2), dim3(128, 2, 1),
                              dim3(8, 8, 4), dim3(8, 8, 8), dim3(16, 4, 4), dim3(8, 16, 2)};
constexpr int kAdvectBlockCount = int(sizeof(kAdvectBlocks) / sizeof(kAdvectBlocks[0]));
constexpr int kAdvectStaged = kAdvectBlockCount;
constexpr int kAdvectCandidates = kAdvectBlockCount + 1;
constexpr int kAdvectTable[4] = {kAdvectStaged, kAdvectStaged, kAdvectStaged, kAdvectStaged};

__global__ void fill_kernel(float* __restrict__ p, float value, int64_t n) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = value;
}

__global__ void fill_u8_kernel(uint8_t* __restrict__ p, uint8_t value, int64_t n) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = value;
}

// ------------------------------------------------------------ time step
// max |f| over one array: float4 loads (the arrays are separate, 256-byte
// aligned allocations), then the tail; the maximum does not depend on the order
__device__ __forceinline__ float abs_max(const float* __restrict__ f, int64_t n, int64_t first, int64_t stride, float m) {
    const int64_t n4 = n / 4;
    const float4* __restrict__ f4 = reinterpret_cast<const float4*>(f);
    for (int64_t i = first; i < n4; i += stride) {
        const float4 q = f4[i];
        m = fmaxf(m, fmaxf(fmaxf(fabsf(q.x), fabsf(q.y)), fmaxf(fabsf(q.z), fabsf(q.w))));
    }
    for (int64_t i = 4 * n4 + first; i < n; i += stride) m = fmaxf(m, fabsf(f[i]));
    return m;
}

__device__ __forceinline__ void max_speed_body(const VBlock& vb, const float* __restrict__ u, int64_t nu, const float* __restrict__ v, int64_t nv, const float* __restrict__ w,
                                 int64_t nw, unsigned* __restrict__ max_bits) {
    const int64_t stride = int64_t(vb.dim.x) * blockDim.x;
    const int64_t first = int64_t(vb.idx.x) * blockDim.x + threadIdx.x;
    float m = abs_max(u, nu, first, stride, 0.0f);
    m = abs_max(v, nv, first, stride, m);
    m = abs_max(w, nw, first, stride, m);
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_down(m, off));
    __shared__ float red[kBlock / 64];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 1; q < kBlock / 64; ++q) m = fmaxf(m, red[q]);
        red[0] = m;
        // non-negative floats order like their bit patterns
        atomicMax(max_bits, __float_as_uint(red[0]));
    }
}

__global__ void max_speed_kernel(const float* __restrict__ u, int64_t nu, const float* __restrict__ v, int64_t nv, const float* __restrict__ w,
                                 int64_t nw, unsigned* __restrict__ max_bits) { max_speed_body(hw_block(), u, nu, v, nv, w, nw, max_bits); }

struct max_speed_k {
    template <class... A>
    __device__ __forceinline__ static void run(const VBlock& vb, A... a) { max_speed_body(vb, a...); }
};

// dt = min(dt_max, cfl h / max|u|) (adaptive) or the configured dt.
__device__ __forceinline__ void time_step_body(const VBlock& vb, unsigned* __restrict__ max_bits, float* __restrict__ max_speed, float* __restrict__ dt, int adaptive, float cfl,
                                 float h, float dt_max, float dt_fixed) {
    const float m = __uint_as_float(*max_bits);
    *max_speed = m;
    *max_bits = 0u;
    if (!adaptive) {
        *dt = dt_fixed;
        return;
    }
    const float cand = m > 0.0f ? cfl * h / m : dt_max;
    *dt = fminf(dt_max, cand);
}

__global__ void time_step_kernel(unsigned* __restrict__ max_bits, float* __restrict__ max_speed, float* __restrict__ dt, int adaptive, float cfl,
                                 float h, float dt_max, float dt_fixed) { time_step_body(hw_block(), max_bits, max_speed, dt, adaptive, cfl, h, dt_max, dt_fixed); }

struct time_step_k {
    template <class... A>
    __device__ __forceinline__ static void run(const VBlock& vb, A... a) { time_step_body(vb, a...); }
};

// arrays of the grid (faces included) are indexable in 32 bits
__host__ __device__ __forceinline__ bool index32(const Grid& g) {
    return (int64_t(g.n[0]) + 1) * (int64_t(g.n[1]) + 1) * (int64_t(g.n[2]) + 1) < (int64_t(1) << 31);
}
inline bool index32_host(const Grid& g) { return index32(g); }

// distance between neighbouring cells along axis C in the cell arrays
template <int C>
__device__ __forceinline__ int cell_stride(const Grid& g) {
    return C == 0 ? 1 : (C == 1 ? g.n[0] : g.n[0] * g.n[1]);
}

// Tiles of kCellBlock (64 x 4 x 1 cells) over the grid, taken in the
// XCD-aware order (xcd_tile): a launch of xcd_blocks(cell_tiles(g)) blocks.
// Consecutive tiles run on one XCD, so a tile's z +- 1 neighbour planes were
// mostly brought into that XCD's L2 by the tiles before it.
inline uint3 cell_tiles(const Grid& g) { return make_uint3(unsigned((g.n[0] + 63) / 64), unsigned((g.n[1] + 3) / 4), unsigned(g.n[2])); }
inline unsigned cell_tile_blocks(const Grid& g) {
    const uint3 t = cell_tiles(g);
    return xcd_blocks(dim3(t.x, t.y, t.z));
}

__device__ __forceinline__ bool xcd_cell(const VBlock& vb, const uint3& tiles, int& x, int& y, int& z) {
    uint3 t;
    unsigned id;
    if (!xcd_tile(vb, tiles, t, id)) return false;
    x = int(t.x) * 64 + int(threadIdx.x);
    y = int(t.y) * 4 + int(threadIdx.y);
    z = int(t.z);
    return true;
}

// ------------------------------------------------------------ forces (stage 0)
// omega from the centred differences of the cell-centred velocity
// (B:870-900); the arguments are centre components at the six neighbours.
// Both the two-pass force stage and the fused one evaluate omega, |omega|
// and the cell force through these helpers, so they produce the same bits.
struct Omega {
    float x, y, z;
};

// half_inv_h = 0.5 / h: the centred difference (a - b) / (2 h) as one
// multiplication (exactly the quotient when h is a power of two)
__device__ __forceinline__ Omega omega_from(float w_yp, float w_ym, float v_zp, float v_zm, float u_zp, float u_zm, float w_xp, float w_xm,
                                            float v_xp, float v_xm, float u_yp, float u_ym, float half_inv_h) {
    const float dvz_dy = (w_yp - w_ym) * half_inv_h;
    const float dvy_dz = (v_zp - v_zm) * half_inv_h;
    const float dvx_dz = (u_zp - u_zm) * half_inv_h;
    const float dvz_dx = (w_xp - w_xm) * half_inv_h;
    const float dvy_dx = (v_xp - v_xm) * half_inv_h;
    const float dvx_dy = (u_yp - u_ym) * half_inv_h;
    return Omega{dvz_dy - dvy_dz, dvx_dz - dvz_dx, dvy_dx - dvx_dy};
}

// |omega| with its fused multiply-adds written out, so that the rounding
// does not depend on how the compiler contracts the sum of squares
__device__ __forceinline__ float omega_norm(const Omega& o) { return sqrtf(fmaf(o.z, o.z, fmaf(o.x, o.x, o.y * o.y))); }

// centre value of component comp at cell (xx, yy, zz): a domain cell's is the
// face average of center_velocity (32-bit indices when `fast`), the ghost
// rules only for cells outside the domain
__device__ __forceinline__ float centre_component(const Grid& g, bool fast, const float* __restrict__ f, int comp, int xx, int yy, int zz) {
    const int n0 = g.n[0], n1 = g.n[1], n2 = g.n[2];
    if (fast && xx >= 0 && yy >= 0 && zz >= 0 && xx < n0 && yy < n1 && zz < n2) {
        const int d0 = n0 + (comp == 0), d1 = n1 + (comp == 1);
        const int lo = (zz * d1 + yy) * d0 + xx;
        const int hi = lo + (comp == 0 ? 1 : (comp == 1 ? d0 : d0 * d1));
        return 0.5f * (f[lo] + f[hi]);
    }
    return load_center_component(g, f, comp, xx, yy, zz);
}

// omega at the fluid domain cell (x, y, z), every input read from the face arrays
__device__ __forceinline__ Omega omega_at(const Grid& g, bool fast, const float* __restrict__ u, const float* __restrict__ v,
                                          const float* __restrict__ w, int x, int y, int z) {
    return omega_from(centre_component(g, fast, w, 2, x, y + 1, z), centre_component(g, fast, w, 2, x, y - 1, z),
                      centre_component(g, fast, v, 1, x, y, z + 1), centre_component(g, fast, v, 1, x, y, z - 1),
                      centre_component(g, fast, u, 0, x, y, z + 1), centre_component(g, fast, u, 0, x, y, z - 1),
                      centre_component(g, fast, w, 2, x + 1, y, z), centre_component(g, fast, w, 2, x - 1, y, z),
                      centre_component(g, fast, v, 1, x + 1, y, z), centre_component(g, fast, v, 1, x - 1, y, z),
                      centre_component(g, fast, u, 0, x, y + 1, z), centre_component(g, fast, u, 0, x, y - 1, z), 0.5f / g.h);
}

// B:902-948 for a fluid cell: buoyancy (+y) and vorticity confinement added
// to the user force (sx, sy, sz); om_* are |omega| at the six neighbours
// (ghost rules applied), o is omega at the cell
__device__ __forceinline__ void fluid_cell_force(float& sx, float& sy, float& sz, float a, float b, float tamb, float eps, float h, float rho, float T,
                                                 float om_xp, float om_xm, float om_yp, float om_ym, float om_zp, float om_zm, const Omega& o) {
    sy += -a * rho + b * (T - tamb);
    const float half_inv_h = 0.5f / h;
    const float gx = (om_xp - om_xm) * half_inv_h;
    const float gy = (om_yp - om_ym) * half_inv_h;
    const float gz = (om_zp - om_zm) * half_inv_h;
    const float gm = sqrtf(gx * gx + gy * gy + gz * gz);
    if (!(gm < 1.0e-6f)) {
        const float inv = 1.0f / gm;
        const float nx = gx * inv, ny = gy * inv, nz = gz * inv;
        sx += eps * h * (ny * o.z - nz * o.y);
        sy += eps * h * (nz * o.x - nx * o.z);
        sz += eps * h * (nx * o.y - ny * o.x);
    }
}

// B:2001-2003: solid temperature, then omega and |omega| from the
// cell-centred velocity with the flow ghost rules (B:870-900).
__device__ __forceinline__ void vorticity_body(const VBlock& vb, Grid g, const float* __restrict__ u, const float* __restrict__ v, const float* __restrict__ w,
                                 const uint8_t* __restrict__ occ, float* __restrict__ ox, float* __restrict__ oy, float* __restrict__ oz,
                                 float* __restrict__ om, float* __restrict__ temperature, const float* __restrict__ solid_temperature, uint3 tiles) {
    int x, y, z;
    if (!xcd_cell(vb, tiles, x, y, z)) return;
    if (x >= g.n[0] || y >= g.n[1]) return;
    const int64_t i = cidx(g, x, y, z);
    if (occ[i] != 0) {
        if (temperature) temperature[i] = solid_temperature[i];
        ox[i] = oy[i] = oz[i] = om[i] = 0.0f;
        return;
    }
    const Omega o = omega_at(g, index32(g), u, v, w, x, y, z);
    ox[i] = o.x;
    oy[i] = o.y;
    oz[i] = o.z;
    om[i] = omega_norm(o);
}

__global__ void vorticity_kernel(Grid g, const float* __restrict__ u, const float* __restrict__ v, const float* __restrict__ w,
                                 const uint8_t* __restrict__ occ, float* __restrict__ ox, float* __restrict__ oy, float* __restrict__ oz,
                                 float* __restrict__ om, float* __restrict__ temperature, const float* __restrict__ solid_temperature, uint3 tiles) { vorticity_body(hw_block(), g, u, v, w, occ, ox, oy, oz, om, temperature, solid_temperature, tiles); }

struct vorticity_k {
    template <class... A>
    __device__ __forceinline__ static void run(const VBlock& vb, A... a) { vorticity_body(vb, a...); }
};

// B:902-948: user force, buoyancy (+y) and vorticity confinement in fluid cells.
__device__ __forceinline__ void cell_force_body(const VBlock& vb, Grid g, const float* __restrict__ fux, const float* __restrict__ fuy, const float* __restrict__ fuz,
                                  const float* __restrict__ rho, const float* __restrict__ T, const uint8_t* __restrict__ occ,
                                  const float* __restrict__ ox, const float* __restrict__ oy, const float* __restrict__ oz,
                                  const float* __restrict__ om, float a, float b, float tamb, float eps, float* __restrict__ fx,
                                  float* __restrict__ fy, float* __restrict__ fz, uint3 tiles) {
    int x, y, z;
    if (!xcd_cell(vb, tiles, x, y, z)) return;
    if (x >= g.n[0] || y >= g.n[1]) return;
    const int64_t i = cidx(g, x, y, z);
    float sx = fux[i], sy = fuy[i], sz = fuz[i];
    if (occ[i] == 0) {
        // |omega| at a neighbour: a plain load for domain cells (32-bit
        // indices), the ghost rule only for the neighbours outside the domain
        const bool fast = index32(g);
        auto om_at = [&](int xx, int yy, int zz) {
            if (fast && xx >= 0 && yy >= 0 && zz >= 0 && xx < g.n[0] && yy < g.n[1] && zz < g.n[2]) return om[(zz * g.n[1] + yy) * g.n[0] + xx];
            return load_flow_cell(g, om, xx, yy, zz);
        };
        fluid_cell_force(sx, sy, sz, a, b, tamb, eps, g.h, rho[i], T[i], om_at(x + 1, y, z), om_at(x - 1, y, z), om_at(x, y + 1, z),
                         om_at(x, y - 1, z), om_at(x, y, z + 1), om_at(x, y, z - 1), Omega{ox[i], oy[i], oz[i]});
    }
    fx[i] = sx;
    fy[i] = sy;
    fz[i] = sz;
}

__global__ void cell_force_kernel(Grid g, const float* __restrict__ fux, const float* __restrict__ fuy, const float* __restrict__ fuz,
                                  const float* __restrict__ rho, const float* __restrict__ T, const uint8_t* __restrict__ occ,
                                  const float* __restrict__ ox, const float* __restrict__ oy, const float* __restrict__ oz,
                                  const float* __restrict__ om, float a, float b, float tamb, float eps, float* __restrict__ fx,
                                  float* __restrict__ fy, float* __restrict__ fz, uint3 tiles) { cell_force_body(hw_block(), g, fux, fuy, fuz, rho, T, occ, ox, oy, oz, om, a, b, tamb, eps, fx, fy, fz, tiles); }

struct cell_force_k {
    template <class... A>
    __device__ __forceinline__ static void run(const VBlock& vb, A... a) { cell_force_body(vb, a...); }
};

// ---- fused force pass: omega, |omega| and the cell force in one kernel,
// with no omega arrays in HBM.  A block owns a 64 x 4 column of cells and
// marches through kFfPlanes z planes of it.  Per plane it stages the
// cell-centred velocity of the plane two ahead on a halo of 2 (ghost rules
// for cells outside the domain) in LDS, computes omega and |omega| of the
// plane one ahead on a halo of 1 from those, and then the cell force of the
// current plane, whose |omega| neighbours are all in LDS.  Rings of three
// planes hold both.  Same helpers, so the same bits as vorticity_kernel
// followed by cell_force_kernel; the omega views are recomputed lazily.
constexpr int kFfTx = 64, kFfTy = 4, kFfPlanes = 16;
constexpr int kFfCx = kFfTx + 4, kFfCy = kFfTy + 4, kFfCn = kFfCx * kFfCy;   // centre-velocity plane, halo 2
constexpr int kFfOx = kFfTx + 2, kFfOy = kFfTy + 2, kFfOn = kFfOx * kFfOy;   // omega plane, halo 1

inline uint3 force_tiles(const Grid& g) {
    return make_uint3(unsigned((g.n[0] + kFfTx - 1) / kFfTx), unsigned((g.n[1] + kFfTy - 1) / kFfTy), unsigned((g.n[2] + kFfPlanes - 1) / kFfPlanes));
}

template <bool I32>
__global__ __launch_bounds__(kFfTx * kFfTy) void force_fused_kernel(Grid g, const float* __restrict__ u, const float* __restrict__ v,
                                                                     const float* __restrict__ w, const uint8_t* __restrict__ occ,
                                                                     const float* __restrict__ fux, const float* __restrict__ fuy,
                                                                     const float* __restrict__ fuz, const float* __restrict__ rho,
                                                                     float* __restrict__ T, const float* __restrict__ solid_temperature, float a,
                                                                     float b, float tamb, float eps, float* __restrict__ fx, float* __restrict__ fy,
                                                                     float* __restrict__ fz, uint3 tiles) {
    using Index = typename std::conditional<I32, int, int64_t>::type;
    constexpr int kThreads = kFfTx * kFfTy;
    constexpr int kPer = (kFfCn + kThreads - 1) / kThreads;   // staged centre cells per thread
    __shared__ float cv[3][3][kFfCn];   // [plane ring][component][cell]
    __shared__ float om[3][4][kFfOn];   // [plane ring][wx, wy, wz, |w|][cell]
    uint3 tile;
    unsigned id;
    if (!xcd_tile(tiles, tile, id)) return;   // the whole block
    const int t = int(threadIdx.x);
    const int n0 = g.n[0], n1 = g.n[1], n2 = g.n[2];
    const int x0 = int(tile.x) * kFfTx, y0 = int(tile.y) * kFfTy, z0 = int(tile.z) * kFfPlanes;
    const int zend = min(z0 + kFfPlanes, n2);
    const float h = g.h;
    const Index d0u = n0 + 1, d1v = n1 + 1;
    const Index su = Index(d0u) * n1, sv = Index(n0) * d1v, sw = Index(n0) * n1;   // z strides of the face arrays
    auto ring = [](int z) { return (z + 3) % 3; };   // z >= -3
    // A staged position off a periodic axis stands for its periodic image
    // (the ghost rule of that axis is the wrap), so the halo values there are
    // the far side's.
    auto wrap = [&](int q, int k) { return g.flow_periodic[k] ? wrapi(q, g.n[k]) : q; };

    const float half_inv_h = 0.5f / h;

    // Centre velocity of plane zz on the halo of 2, in two halves.  fetch
    // issues the two face loads per component of every staged position into
    // a register set, with the coordinates clamped into the domain (always in
    // bounds), two planes before they are stored.  store averages them -- the
    // centre value of a domain position, and the interior value the ghost
    // rule of a position one cell off a non-periodic side starts from (the
    // only off-domain positions an omega of a domain cell or of a periodic
    // image reads; 0 further out).  A thread's staged positions are the same
    // in every plane: their x, y parts are worked out once.
    int qx[kPer], qy[kPer], off_u[kPer], off_vw[kPer];
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        const int i = min(t + k * kThreads, kFfCn - 1);
        qx[k] = wrap(x0 - 2 + i % kFfCx, 0);
        qy[k] = wrap(y0 - 2 + i / kFfCx, 1);
        const int xc = clampi(qx[k], 0, n0 - 1), yc = clampi(qy[k], 0, n1 - 1);
        off_u[k] = yc * (n0 + 1) + xc;
        off_vw[k] = yc * n0 + xc;
    }
    struct Fetch {
        float lo[3][kPer], hi[3][kPer];
    };
    auto fetch_centre = [&](int zz, Fetch& f) {
        const Index zc = clampi(wrap(zz, 2), 0, n2 - 1);
        const float* __restrict__ pu = u + zc * su;
        const float* __restrict__ pv = v + zc * sv;
        const float* __restrict__ pw = w + zc * sw;
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            f.lo[0][k] = pu[off_u[k]];
            f.hi[0][k] = pu[off_u[k] + 1];
            f.lo[1][k] = pv[off_vw[k]];
            f.hi[1][k] = pv[off_vw[k] + n0];
            f.lo[2][k] = pw[off_vw[k]];
            f.hi[2][k] = pw[off_vw[k] + sw];
        }
    };
    auto store_centre = [&](int zz, const Fetch& f) {
        const int s = ring(zz);
        const int zw = wrap(zz, 2);
        const bool offz = zw < 0 || zw >= n2;
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            const int i = t + k * kThreads;
            if (i >= kFfCn) break;
            float cc[3];
#pragma unroll
            for (int comp = 0; comp < 3; ++comp) cc[comp] = 0.5f * (f.lo[comp][k] + f.hi[comp][k]);
            const bool offx = qx[k] < 0 || qx[k] >= n0, offy = qy[k] < 0 || qy[k] >= n1;
            if (offx || offy || offz) {
                // load_center_component's rule at its first off-domain axis
                // (the others clamped, as fetched)
                const bool near = qx[k] >= -1 && qx[k] <= n0 && qy[k] >= -1 && qy[k] <= n1 && zw >= -1 && zw <= n2;
                const int ax = offx ? 0 : (offy ? 1 : 2);
                const int face = 2 * ax + ((offx ? qx[k] : (offy ? qy[k] : zw)) < 0 ? 0 : 1);
                // the face's kind and wall velocity read with constant indices
                // (a computed index into the kernel's Grid argument would
                // put the whole struct in scratch memory)
#pragma unroll
                for (int comp = 0; comp < 3; ++comp) {
                    float vv = cc[comp];
#pragma unroll
                    for (int f6 = 0; f6 < 6; ++f6) {
                        if (face != f6) continue;
                        const uint32_t type = g.flow_type[f6];
                        if (!(type == kOutflow || (type == kFreeSlip && comp != (f6 >> 1)))) vv = 2.0f * g.flow_velocity[f6][comp] - vv;
                    }
                    cc[comp] = near ? vv : 0.0f;
                }
            }
            cv[s][0][i] = cc[0];
            cv[s][1][i] = cc[1];
            cv[s][2][i] = cc[2];
        }
    };
    // this thread's omega positions: LDS index of the position in a centre
    // plane and the xy offset of its cell (-1 off the domain)
    constexpr int kOmPer = (kFfOn + kThreads - 1) / kThreads;
    int om_c[kOmPer], om_off[kOmPer];
#pragma unroll
    for (int k = 0; k < kOmPer; ++k) {
        const int i = t + k * kThreads;
        const int ox = i % kFfOx, oy = i / kFfOx;
        const int xw = wrap(x0 - 1 + ox, 0), yw = wrap(y0 - 1 + oy, 1);
        om_c[k] = (oy + 1) * kFfCx + ox + 1;
        om_off[k] = (i < kFfOn && xw >= 0 && xw < n0 && yw >= 0 && yw < n1) ? yw * n0 + xw : -1;
    }
    // occupancy of this thread's omega positions of plane zz (1 off the
    // domain or past the plane), fetched an iteration before stage_omega
    auto fetch_occ = [&](int zz, uint8_t (&o)[kOmPer]) {
        const int zw = wrap(zz, 2);
        const bool inz = zw >= 0 && zw < n2;
        const uint8_t* __restrict__ po = occ + Index(inz ? zw : 0) * sw;
#pragma unroll
        for (int k = 0; k < kOmPer; ++k) o[k] = (inz && om_off[k] >= 0) ? po[om_off[k]] : uint8_t(1);
    };
    // omega and |omega| of plane zz on the halo of 1, at the fluid domain
    // cells and periodic images (0 elsewhere, as vorticity_kernel writes in
    // solid cells); centre planes zz - 1, zz, zz + 1 staged
    auto stage_omega = [&](int zz, const uint8_t (&o)[kOmPer]) {
        const int s = ring(zz), sm = ring(zz - 1), sp = ring(zz + 1);
#pragma unroll
        for (int k = 0; k < kOmPer; ++k) {
            const int i = t + k * kThreads;
            if (i >= kFfOn) break;
            Omega w3{0.0f, 0.0f, 0.0f};
            float m = 0.0f;
            if (o[k] == 0) {
                const int c = om_c[k];
                w3 = omega_from(cv[s][2][c + kFfCx], cv[s][2][c - kFfCx], cv[sp][1][c], cv[sm][1][c], cv[sp][0][c], cv[sm][0][c], cv[s][2][c + 1],
                                cv[s][2][c - 1], cv[s][1][c + 1], cv[s][1][c - 1], cv[s][0][c + kFfCx], cv[s][0][c - kFfCx], half_inv_h);
                m = omega_norm(w3);
            }
            om[s][0][i] = w3.x;
            om[s][1][i] = w3.y;
            om[s][2][i] = w3.z;
            om[s][3][i] = m;
        }
    };

    // prologue: centre planes z0 - 2 .. z0 + 2, omega planes z0 - 1, z0;
    // plane z0 + 3 fetched, the occupancy of omega plane z0 + 1 fetched
    Fetch fa, fb;
    uint8_t occ_a[kOmPer], occ_b[kOmPer];
    fetch_centre(z0 - 2, fa);
    fetch_centre(z0 - 1, fb);
    fetch_occ(z0 - 1, occ_a);
    fetch_occ(z0, occ_b);
    store_centre(z0 - 2, fa);
    store_centre(z0 - 1, fb);
    fetch_centre(z0, fa);
    fetch_centre(z0 + 1, fb);
    store_centre(z0, fa);
    __syncthreads();
    stage_omega(z0 - 1, occ_a);
    __syncthreads();
    store_centre(z0 + 1, fb);   // over plane z0 - 2, last read above
    fetch_centre(z0 + 2, fa);
    fetch_occ(z0 + 1, occ_a);
    __syncthreads();
    stage_omega(z0, occ_b);
    __syncthreads();
    store_centre(z0 + 2, fa);   // over plane z0 - 1, last read above
    fetch_centre(z0 + 3, fa);

    const int tx = t % kFfTx, ty = t / kFfTx;
    const int x = x0 + tx, y = y0 + ty;
    const bool mine = x < n0 && y < n1;
    const int c = (ty + 1) * kFfOx + tx + 1;   // this thread's cell in an omega plane
    // neighbours off a non-periodic side of the domain
    const bool off_xm = !g.flow_periodic[0] && x == 0, off_xp = !g.flow_periodic[0] && x == n0 - 1;
    const bool off_ym = !g.flow_periodic[1] && y == 0, off_yp = !g.flow_periodic[1] && y == n1 - 1;
    // one output plane: centre planes z, z + 1, z + 2 and omega planes z - 1,
    // z staged; `cur` holds centre plane z + 3, `ocur` the occupancy of
    // omega plane z + 1; the next ones are fetched into `next` and `onext`
    auto plane = [&](int z, Fetch& cur, Fetch& next, uint8_t (&ocur)[kOmPer], uint8_t (&onext)[kOmPer]) {
        fetch_centre(z + 4, next);
        fetch_occ(z + 2, onext);
        const Index i = (Index(z) * n1 + (mine ? y : 0)) * n0 + (mine ? x : 0);
        const float ux = fux[i], uy = fuy[i], uz = fuz[i], ri = rho[i], Ti = T[i];
        const uint8_t oi = occ[i];
        __syncthreads();
        stage_omega(z + 1, ocur);   // over omega(z - 2), last read by plane z - 1
        __syncthreads();
        if (!mine) {
            store_centre(z + 3, cur);
            return;
        }
        {
            float sx = ux, sy = uy, sz = uz;
            if (oi == 0) {
                const int s = ring(z);
                // |omega| of a neighbour (load_flow_cell's rule): staged in
                // LDS for domain cells and periodic images; off a
                // non-periodic side the clamp gives the cell itself
                const float self = om[s][3][c];
                const bool off_zm = !g.flow_periodic[2] && z == 0, off_zp = !g.flow_periodic[2] && z == n2 - 1;
                fluid_cell_force(sx, sy, sz, a, b, tamb, eps, h, ri, Ti, off_xp ? self : om[s][3][c + 1], off_xm ? self : om[s][3][c - 1],
                                 off_yp ? self : om[s][3][c + kFfOx], off_ym ? self : om[s][3][c - kFfOx], off_zp ? self : om[ring(z + 1)][3][c],
                                 off_zm ? self : om[ring(z - 1)][3][c], Omega{om[s][0][c], om[s][1][c], om[s][2][c]});
            } else {
                T[i] = solid_temperature[i];
            }
            fx[i] = sx;
            fy[i] = sy;
            fz[i] = sz;
        }
        store_centre(z + 3, cur);   // over centre plane z, last read by stage_omega(z + 1)
    };
    // two planes per trip, so the register sets swap roles without copies
    for (int z = z0; z < zend; z += 2) {
        plane(z, fa, fb, occ_a, occ_b);
        if (z + 1 < zend) plane(z + 1, fb, fa, occ_b, occ_a);
    }
}

// B:950-1005 + solid part of B:1014-1168: u_face += dt * mean of the
// in-domain neighbour cell forces (no wrap), then the solid override.
template <int C>
__device__ __forceinline__ void face_force_body(const VBlock& vb, Grid g, float* __restrict__ f, const float* __restrict__ force, const uint8_t* __restrict__ occ,
                                  const float* __restrict__ solid_vel, const float* __restrict__ dt_ptr) {
    const int i = vb.idx.x * blockDim.x + threadIdx.x;
    const int j = vb.idx.y * blockDim.y + threadIdx.y;
    const int k = vb.idx.z * blockDim.z + threadIdx.z;
    if (i >= g.n[0] + (C == 0) || j >= g.n[1] + (C == 1) || k >= g.n[2] + (C == 2)) return;
    const int c[3] = {i, j, k};
    if (boundary_face(g, C, c[C])) return;
    if (c[C] > 0 && c[C] < g.n[C] && index32(g)) {
        // an inner face (0 < index < n): both neighbour cells are domain
        // cells; when both are fluid the face gets the mean of their forces
        const int ir = (k * g.n[1] + j) * g.n[0] + i, il = ir - cell_stride<C>(g);
        if (occ[il] == 0 && occ[ir] == 0) {
            const int fi32 = (k * (g.n[1] + (C == 1)) + j) * (g.n[0] + (C == 0)) + i;
            float sum = 0.0f;
            sum += force[il];
            sum += force[ir];
            f[fi32] += *dt_ptr * (sum / 2.0f);
            return;
        }
    }
    const int64_t fi = fidx(g, C, i, j, k);
    float value;
    if (solid_face_override(g, occ, solid_vel, C, i, j, k, value)) {
        f[fi] = value;
        return;
    }
    float sum = 0.0f, weight = 0.0f;
    if (c[C] > 0) {
        int l[3] = {i, j, k};
        l[C] -= 1;
        sum += force[cidx(g, l[0], l[1], l[2])];
        weight += 1.0f;
    }
    if (c[C] < g.n[C]) {
        sum += force[cidx(g, i, j, k)];
        weight += 1.0f;
    }
    f[fi] += *dt_ptr * (sum / weight);
}

template <int C>
__global__ void face_force_kernel(Grid g, float* __restrict__ f, const float* __restrict__ force, const uint8_t* __restrict__ occ,
                                  const float* __restrict__ solid_vel, const float* __restrict__ dt_ptr) { face_force_body<C>(hw_block(), g, f, force, occ, solid_vel, dt_ptr); }

template <int C>
struct face_force_k {
    template <class... A>
    __device__ __forceinline__ static void run(const VBlock& vb, A... a) { face_force_body<C>(vb, a...); }
};

// Domain end faces of component C, one thread per line along axis C; both
// ends are read before either is written.
template <int C>
__device__ __forceinline__ void boundary_faces_body(const VBlock& vb, Grid g, float* __restrict__ f) {
    constexpr int A1 = (C + 1) % 3, A2 = (C + 2) % 3;
    const int q1 = vb.idx.x * blockDim.x + threadIdx.x;
    const int q2 = vb.idx.y * blockDim.y + threadIdx.y;
    if (q1 >= g.n[A1] || q2 >= g.n[A2]) return;
    auto at = [&](int idx) {
        int c[3];
        c[C] = idx;
        c[A1] = q1;
        c[A2] = q2;
        return fidx(g, C, c[0], c[1], c[2]);
    };
    const int n = g.n[C];
    if (g.flow_periodic[C]) {
        f[at(n)] = f[at(0)];
        return;
    }
    const float inner_lo = f[at(1)];
    const float inner_hi = f[at(n - 1)];
    const uint32_t tl = g.flow_type[2 * C], th = g.flow_type[2 * C + 1];
    const float lo = tl == kOutflow ? inner_lo : g.flow_velocity[2 * C][C];
    const float hi = th == kOutflow ? inner_hi : g.flow_velocity[2 * C + 1][C];
    f[at(0)] = lo;
    f[at(n)] = hi;
}

template <int C>
__global__ void boundary_faces_kernel(Grid g, float* __restrict__ f) { boundary_faces_body<C>(hw_block(), g, f); }

template <int C>
struct boundary_faces_k {
    template <class... A>
    __device__ __forceinline__ static void run(const VBlock& vb, A... a) { boundary_faces_body<C>(vb, a...); }
};

// ------------------------------------------------------------ velocity advection (stage 1)
// B:1191-1219 + the enforce that follows (B:2025-2030): every interior face is
// traced and sampled unless a solid override replaces it.
// face (i,j,k) of component C traced back and sampled by the global path
template <int C>
__device__ __forceinline__ float advect_face_value(const Grid& g, const float* __restrict__ u, const float* __restrict__ v,
                                                   const float* __restrict__ w, const uint8_t* __restrict__ occ, float s, int i, int j, int k) {
    // positions in cell units, s = dt / h
    const float3 p = trace_rk2_at<C>(g, u, v, w, occ, s, i, j, k);
    const float* src = C == 0 ? u : (C == 1 ? v : w);
    return sample_face<C>(g, src, p.x, p.y, p.z, g.cubic != 0);
}

template <int C>
__device__ __forceinline__ void advect_face_body(const VBlock& vb, Grid g, const float* __restrict__ u, const float* __restrict__ v, const float* __restrict__ w,
                                   float* __restrict__ dst, const uint8_t* __restrict__ occ, const float* __restrict__ solid_vel,
                                   const float* __restrict__ dt_ptr) {
    const int i = vb.idx.x * blockDim.x + threadIdx.x;
    const int j = vb.idx.y * blockDim.y + threadIdx.y;
    const int k = vb.idx.z * blockDim.z + threadIdx.z;
    if (i >= g.n[0] + (C == 0) || j >= g.n[1] + (C == 1) || k >= g.n[2] + (C == 2)) return;
    const int c[3] = {i, j, k};
    if (boundary_face(g, C, c[C])) return;
    const int64_t fi = fidx(g, C, i, j, k);
    float value;
    if (solid_face_override(g, occ, solid_vel, C, i, j, k, value)) {
        dst[fi] = value;
        return;
    }
    dst[fi] = advect_face_value<C>(g, u, v, w, occ, *dt_ptr / g.h, i, j, k);
}

template <int C>
__global__ void advect_face_kernel(Grid g, const float* __restrict__ u, const float* __restrict__ v, const float* __restrict__ w,
                                   float* __restrict__ dst, const uint8_t* __restrict__ occ, const float* __restrict__ solid_vel,
                                   const float* __restrict__ dt_ptr) { advect_face_body<C>(hw_block(), g, u, v, w, dst, occ, solid_vel, dt_ptr); }

template <int C>
struct advect_face_k {
    template <class... A>
    __device__ __forceinline__ static void run(const VBlock& vb, A... a) { advect_face_body<C>(vb, a...); }
};

// ------------------------------------------------------------ projection (stage 3)
// B:1425-1600 + enforce + copy back (B:2032-2043): the projected faces are
// written straight into the state array.
template <int C>
__device__ __forceinline__ void project_body(const VBlock& vb, Grid g, const float* __restrict__ ustar, float* __restrict__ out, const float* __restrict__ p,
                               const uint8_t* __restrict__ occ, const float* __restrict__ solid_vel, const float* __restrict__ dt_ptr) {
    const int i = vb.idx.x * blockDim.x + threadIdx.x;
    const int j = vb.idx.y * blockDim.y + threadIdx.y;
    const int k = vb.idx.z * blockDim.z + threadIdx.z;
    if (i >= g.n[0] + (C == 0) || j >= g.n[1] + (C == 1) || k >= g.n[2] + (C == 2)) return;
    const int c[3] = {i, j, k};
    if (boundary_face(g, C, c[C])) return;
    if (c[C] > 0 && c[C] < g.n[C] && index32(g)) {
        // an inner face between two fluid cells
        const int ir = (k * g.n[1] + j) * g.n[0] + i, il = ir - cell_stride<C>(g);
        if (occ[il] == 0 && occ[ir] == 0) {
            const int fi32 = (k * (g.n[1] + (C == 1)) + j) * (g.n[0] + (C == 0)) + i;
            float result = ustar[fi32];
            result -= *dt_ptr * (p[ir] - p[il]) / g.h;
            out[fi32] = result;
            return;
        }
    }
    const int64_t fi = fidx(g, C, i, j, k);
    float value;
    if (solid_face_override(g, occ, solid_vel, C, i, j, k, value)) {
        out[fi] = value;
        return;
    }
    int l[3] = {i, j, k};
    int r[3] = {i, j, k};
    l[C] -= 1;
    const bool hl = resolve_flow(g, l), hr = resolve_flow(g, r);
    float result = ustar[fi];
    if (hl && hr) result -= *dt_ptr * (p[cidx(g, r[0], r[1], r[2])] - p[cidx(g, l[0], l[1], l[2])]) / g.h;
    out[fi] = result;
}

template <int C>
__global__ void project_kernel(Grid g, const float* __restrict__ ustar, float* __restrict__ out, const float* __restrict__ p,
                               const uint8_t* __restrict__ occ, const float* __restrict__ solid_vel, const float* __restrict__ dt_ptr) { project_body<C>(hw_block(), g, ustar, out, p, occ, solid_vel, dt_ptr); }

template <int C>
struct project_k {
    template <class... A>
    __device__ __forceinline__ static void run(const VBlock& vb, A... a) { project_body<C>(vb, a...); }
};

// ------------------------------------------------------------ scalars (stage 4)
// B:852-857: tmp = cur + dt * src for both scalars.
__device__ __forceinline__ void scalar_sources_body(const VBlock& vb, const float* __restrict__ T, const float* __restrict__ ST, const float* __restrict__ rho,
                                      const float* __restrict__ Srho, float* __restrict__ Ttmp, float* __restrict__ rtmp,
                                      const float* __restrict__ dt_ptr, int64_t n) {
    const int64_t i = int64_t(vb.idx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float dt = *dt_ptr;
    Ttmp[i] = T[i] + dt * ST[i];
    rtmp[i] = rho[i] + dt * Srho[i];
}

__global__ void scalar_sources_kernel(const float* __restrict__ T, const float* __restrict__ ST, const float* __restrict__ rho,
                                      const float* __restrict__ Srho, float* __restrict__ Ttmp, float* __restrict__ rtmp,
                                      const float* __restrict__ dt_ptr, int64_t n) { scalar_sources_body(hw_block(), T, ST, rho, Srho, Ttmp, rtmp, dt_ptr, n); }

struct scalar_sources_k {
    template <class... A>
    __device__ __forceinline__ static void run(const VBlock& vb, A... a) { scalar_sources_body(vb, a...); }
};

// B:1221-1241: one trace per fluid cell shared by temperature and density;
// solid cells take the solid temperature and a density filled afterwards.
__device__ __forceinline__ void advect_scalars_body(const VBlock& vb, Grid g, const float* __restrict__ u, const float* __restrict__ v, const float* __restrict__ w,
                                      const uint8_t* __restrict__ occ, const float* __restrict__ Ttmp, const float* __restrict__ rtmp,
                                      float* __restrict__ T, float* __restrict__ rho, const float* __restrict__ solid_T,
                                      const float* __restrict__ dt_ptr) {
    const int x = vb.idx.x * blockDim.x + threadIdx.x;
    const int y = vb.idx.y * blockDim.y + threadIdx.y;
    const int z = vb.idx.z * blockDim.z + threadIdx.z;
    if (x >= g.n[0] || y >= g.n[1] || z >= g.n[2]) return;
    const int64_t i = cidx(g, x, y, z);
    if (occ[i] != 0) {
        T[i] = solid_T[i];
        rho[i] = 0.0f;
        return;
    }
    const float3 p = trace_rk2_at<-1>(g, u, v, w, occ, *dt_ptr / g.h, x, y, z);
    float tv, rv;
    sample_scalar2(g, g.temperature, Ttmp, g.density, rtmp, p.x, p.y, p.z, g.cubic != 0, tv, rv);
    T[i] = tv;
    rho[i] = rv;
}

__global__ void advect_scalars_kernel(Grid g, const float* __restrict__ u, const float* __restrict__ v, const float* __restrict__ w,
                                      const uint8_t* __restrict__ occ, const float* __restrict__ Ttmp, const float* __restrict__ rtmp,
                                      float* __restrict__ T, float* __restrict__ rho, const float* __restrict__ solid_T,
                                      const float* __restrict__ dt_ptr) { advect_scalars_body(hw_block(), g, u, v, w, occ, Ttmp, rtmp, T, rho, solid_T, dt_ptr); }

struct advect_scalars_k {
    template <class... A>
    __device__ __forceinline__ static void run(const VBlock& vb, A... a) { advect_scalars_body(vb, a...); }
};

// B:1243-1290 as a gather through the nearest-fluid map built at occupancy update.
__device__ __forceinline__ void density_fill_body(const VBlock& vb, float* __restrict__ rho, const int64_t* __restrict__ pairs, const unsigned* __restrict__ count) {
    const unsigned n = *count;
    for (unsigned q = vb.idx.x * blockDim.x + threadIdx.x; q < n; q += vb.dim.x * blockDim.x) {
        const int64_t cell = pairs[2 * q], src = pairs[2 * q + 1];
        rho[cell] = src >= 0 ? rho[src] : 0.0f;
    }
}

__global__ void density_fill_kernel(float* __restrict__ rho, const int64_t* __restrict__ pairs, const unsigned* __restrict__ count) { density_fill_body(hw_block(), rho, pairs, count); }

struct density_fill_k {
    template <class... A>
    __device__ __forceinline__ static void run(const VBlock& vb, A... a) { density_fill_body(vb, a...); }
};

// ------------------------------------------------------------ occupancy-derived data
__global__ void occupancy_float_kernel(const uint8_t* __restrict__ occ, float* __restrict__ out, int64_t n) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) out[i] = float(occ[i]);
}

__global__ void anchor_reset_kernel(int* anchor, int cells, unsigned* solid_count) {
    *anchor = cells;
    *solid_count = 0u;
}

// B:1297-1302 (first fluid cell): block minimum first, one atomic per block
__global__ void anchor_kernel(const uint8_t* __restrict__ occ, int* __restrict__ anchor, int64_t n) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    int m = (i < n && occ[i] == 0) ? int(i) : 0x7fffffff;
    for (int off = 32; off > 0; off >>= 1) m = min(m, __shfl_down(m, off));
    __shared__ int red[kBlock / 64];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 1; q < kBlock / 64; ++q) m = min(m, red[q]);
        if (m != 0x7fffffff) atomicMin(anchor, m);
    }
}

// Per-cell stencil flags of the Poisson operator (the structure of
// B:1343-1412): bit d set when the neighbour in direction d (x-,x+,y-,y+,z-,z+)
// exists after periodic wrap and is fluid; bit 6 marks identity rows (solid
// cells and the anchor).
__global__ void stencil_flags_kernel(Grid g, const uint8_t* __restrict__ occ, const int* __restrict__ anchor, uint8_t* __restrict__ flags) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    const int z = blockIdx.z * blockDim.z + threadIdx.z;
    if (x >= g.n[0] || y >= g.n[1] || z >= g.n[2]) return;
    const int64_t i = cidx(g, x, y, z);
    if (occ[i] != 0 || i == int64_t(*anchor)) {
        flags[i] = 0x40;
        return;
    }
    uint8_t m = 0;
    const int base[3] = {x, y, z};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            int c[3] = {base[0], base[1], base[2]};
            c[a] += s ? 1 : -1;
            if (c[a] < 0 || c[a] >= g.n[a]) {
                if (!g.flow_periodic[a]) continue;
                c[a] = wrapi(c[a], g.n[a]);
            }
            if (occ[cidx(g, c[0], c[1], c[2])] == 0) m |= uint8_t(1u << (2 * a + s));
        }
    }
    flags[i] = m;
}

// B:1243-1290 run once per occupancy: nearest fluid cell of every solid cell.
__global__ void nearest_fluid_kernel(Grid g, const uint8_t* __restrict__ occ, int64_t* __restrict__ pairs, unsigned* __restrict__ count) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    const int z = blockIdx.z * blockDim.z + threadIdx.z;
    if (x >= g.n[0] || y >= g.n[1] || z >= g.n[2]) return;
    const int64_t i = cidx(g, x, y, z);
    if (occ[i] == 0) return;
    const int max_r = max(g.n[0], max(g.n[1], g.n[2]));
    int64_t best = -1;
    for (int r = 1; r <= max_r && best < 0; ++r) {
        int best_d2 = 0x7fffffff;
        for (int dz = -r; dz <= r; ++dz)
            for (int dy = -r; dy <= r; ++dy) {
                const bool face_yz = abs(dz) == r || abs(dy) == r;
                for (int dx = -r; dx <= r; dx += (face_yz ? 1 : 2 * r)) {
                    int c[3] = {x + dx, y + dy, z + dz};
                    if (!resolve_scalar(g, g.density, c)) continue;
                    const int64_t ni = cidx(g, c[0], c[1], c[2]);
                    if (occ[ni] != 0) continue;
                    const int d2 = dx * dx + dy * dy + dz * dz;
                    if (d2 < best_d2) {
                        best_d2 = d2;
                        best = ni;
                    }
                }
            }
    }
    const unsigned slot = atomicAdd(count, 1u);
    pairs[2 * slot] = i;
    pairs[2 * slot + 1] = best;
}

// Identity rows (solid cells, the anchor) hold exactly +0 in every FP64
// solution buffer, the warm-start history included: the PCG passes rely on
// p = 0 there, and an extrapolated start 2 x_n - x_(n-1) keeps the zeros only
// if both terms have them.
__global__ void zero_identity_rows_kernel(const uint8_t* __restrict__ flags, double* __restrict__ x0, double* __restrict__ x1,
                                          double* __restrict__ hist, int64_t n) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n && (flags[i] & 0x40)) {
        x0[i] = 0.0;
        x1[i] = 0.0;
        hist[i] = 0.0;
    }
}

// ------------------------------------------------------------ diagnostics (views)
// B:859-868, 1330-1341, 1602-1609.
__global__ void diagnostics_kernel(Grid g, const float* __restrict__ u, const float* __restrict__ v, const float* __restrict__ w,
                                   const uint8_t* __restrict__ occ, float* __restrict__ cx, float* __restrict__ cy, float* __restrict__ cz,
                                   float* __restrict__ vmag, float* __restrict__ div) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    const int z = blockIdx.z * blockDim.z + threadIdx.z;
    if (x >= g.n[0] || y >= g.n[1] || z >= g.n[2]) return;
    const int64_t i = cidx(g, x, y, z);
    const float u0 = u[fidx(g, 0, x, y, z)], u1 = u[fidx(g, 0, x + 1, y, z)];
    const float v0 = v[fidx(g, 1, x, y, z)], v1 = v[fidx(g, 1, x, y + 1, z)];
    const float w0 = w[fidx(g, 2, x, y, z)], w1 = w[fidx(g, 2, x, y, z + 1)];
    const float a = 0.5f * (u0 + u1), b = 0.5f * (v0 + v1), c = 0.5f * (w0 + w1);
    cx[i] = a;
    cy[i] = b;
    cz[i] = c;
    vmag[i] = sqrtf(a * a + b * b + c * c);
    div[i] = occ[i] != 0 ? 0.0f : (u1 - u0 + v1 - v0 + w1 - w0) / g.h;
}

// Spherical emitter of the scenario harness (same bits as the CPU oracle).
__device__ __forceinline__ uint32_t mix32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    x ^= x >> 16;
    return x;
}

__device__ __forceinline__ void plume_body(const VBlock& vb, Grid g, SmokeSimulationPlumeSource src, uint32_t step, float* __restrict__ rs, float* __restrict__ ts) {
    const int x = vb.idx.x * blockDim.x + threadIdx.x;
    const int y = vb.idx.y * blockDim.y + threadIdx.y;
    const int z = vb.idx.z * blockDim.z + threadIdx.z;
    if (x >= g.n[0] || y >= g.n[1] || z >= g.n[2]) return;
    const int64_t i = cidx(g, x, y, z);
    const float dx = (float(x) + 0.5f) - src.center_x;
    const float dy = (float(y) + 0.5f) - src.center_y;
    const float dz = (float(z) + 0.5f) - src.center_z;
    const float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, __fmul_rn(dx, dx)));
    if (r2 > __fmul_rn(src.radius, src.radius)) {
        rs[i] = 0.0f;
        ts[i] = 0.0f;
        return;
    }
    const uint32_t bits = mix32(uint32_t(i) ^ mix32(step ^ mix32(src.seed)));
    const float unit = __fsub_rn(__fmul_rn(__fmul_rn(float(bits >> 8), 1.0f / 16777216.0f), 2.0f), 1.0f);
    rs[i] = src.density_rate;
    ts[i] = __builtin_fmaf(src.temperature_noise, unit, src.temperature_rate);
}

__global__ void plume_kernel(Grid g, SmokeSimulationPlumeSource src, uint32_t step, float* __restrict__ rs, float* __restrict__ ts) { plume_body(hw_block(), g, src, step, rs, ts); }

struct plume_k {
    template <class... A>
    __device__ __forceinline__ static void run(const VBlock& vb, A... a) { plume_body(vb, a...); }
};

template <int C>
void launch_face_force(Context& c) {
    const Grid& g = c.grid;
    const dim3 grid = grid3(g.n[0] + (C == 0), g.n[1] + (C == 1), g.n[2] + (C == 2), kCellBlock);
    face_force_kernel<C><<<grid, kCellBlock, 0, c.stream>>>(g, c.face[C], c.cell_force[C], c.occupancy, c.solid_velocity[C], c.dt_device);
}

template <int C>
void launch_boundary(Context& c, float* f) {
    constexpr int A1 = (C + 1) % 3, A2 = (C + 2) % 3;
    const dim3 block(64, 4, 1);
    const dim3 grid = grid3(c.grid.n[A1], c.grid.n[A2], 1, block);
    boundary_faces_kernel<C><<<grid, block, 0, c.stream>>>(c.grid, f);
}

void launch_boundaries(Context& c, float* const f[3]) {
    launch_boundary<0>(c, f[0]);
    launch_boundary<1>(c, f[1]);
    launch_boundary<2>(c, f[2]);
}


// ---- LDS-staged scalar advection (the same results as advect_scalars_kernel).
// A block owns a 32 x 8 x 4 tile of cells (two per thread).  It stages, with the ghost
// rules applied once per staged index, the temperature and density sources
// over the tile plus a halo of 2 (every cubic tap of a backtrace shorter than
// one cell), the three face components over the boxes the two linear samples
// of such a trace touch, and the occupancy of the tile plus a halo of 1.  A
// cell whose every sample stencil and hit-test cell lies in its staged box is
// advected from LDS -- the same values in the same arithmetic order as the
// global samplers, which return f at in-domain indices and the ghost value
// elsewhere.  Any other cell (a trace longer than a cell, a periodic wrap, a
// hit cell outside the box) is appended to a list that a second kernel
// advects with the global path, so the common path carries none of the
// ghost-rule code and keeps its register budget.  Tiles whose boxes lie inside
// the arrays (all but the domain's edge tiles) stage their FP32 boxes by LDS
// DMA, one box row per wavefront pass (dma_rows); grids with no periodic axis
// run the kernels compiled without the wrap code (staged_wrap).  Measured at
// 256^3 against 32 x 8 x 2 tiles with one output per thread and against
// 32 x 8 x 4 tiles of 1024 threads, two outputs per thread was the fastest
// (with register staging: the scalars 0.87 against 0.95 ms, v 0.52 against
// 0.68 and 0.76 ms).  The DMA staging took the scalars from 0.87 to 0.73 ms
// and v and w from 0.52 to 0.45: staged through registers, each row's load
// waited for its memory round trip before the next was issued.  Edge tiles
// keep the per-entry ghost-rule staging: the same clamped-DMA-then-patch
// scheme for them raised the kernels' scalar register spills tenfold and
// made every tile slower.
constexpr int kSaX = 32, kSaY = 8, kSaZ = 4;
// threads of a staged block: each owns an (x, y) column of the tile and the
// z planes tz0, tz0 + kSaZStep, ... of it
constexpr int kSaThreads = 512;
constexpr int kSaWaves = kSaThreads / 64;
constexpr int kSaPlane = kSaX * kSaY;
constexpr int kSaZStep = kSaThreads / kSaPlane;
static_assert(kSaThreads % kSaPlane == 0 && kSaZ % kSaZStep == 0, "staged tile and block");
// every staged box is at most one wavefront wide (row-wise staging)
static_assert(kSaX + 4 <= 64, "staged box width");

struct LdsBox {
    int o[3];   // origin (global index of the first staged entry)
    int d[3];   // extent
};

__device__ __forceinline__ int box_size(const LdsBox& b) { return b.d[0] * b.d[1] * b.d[2]; }

template <int LO, int HI>
__device__ __forceinline__ bool in_box(const LdsBox& bx, const Stencil& s) {
    bool ok = true;
#pragma unroll
    for (int a = 0; a < 3; ++a) ok = ok && s.b[a] + LO >= bx.o[a] && s.b[a] + HI < bx.o[a] + bx.d[a];
    return ok;
}

// Offset of entry (x, y, z) in a box it lies in: non-negative offsets and
// extents below 2^24, so the products are full-rate 24-bit multiplies.
__device__ __forceinline__ int box_base(const LdsBox& bx, int x, int y, int z) {
    return int(__umul24(__umul24(unsigned(z - bx.o[2]), unsigned(bx.d[1])) + unsigned(y - bx.o[1]), unsigned(bx.d[0]))) + (x - bx.o[0]);
}

// True when every index of the box lies in [0, top[a]] on each axis a (n - 1
// for cells, n on a face component's own axis): its entries are then the
// array's values themselves and the ghost rules never apply.
__device__ __forceinline__ bool box_inside(const LdsBox& bx, int t0, int t1, int t2) {
    return bx.o[0] >= 0 && bx.o[1] >= 0 && bx.o[2] >= 0 && bx.o[0] + bx.d[0] - 1 <= t0 && bx.o[1] + bx.d[1] - 1 <= t1 &&
           bx.o[2] + bx.d[2] - 1 <= t2;
}

// Staging of a box of D0 x D1 x D2 entries, one box row (j, k) = (r % D1,
// r / D1) per wavefront pass: wavefront w takes rows w, w + kSaWaves, ...,
// lane x entry x of the row, so a row's indices are wavefront-uniform (scalar
// arithmetic) and each row is one contiguous load.  The FP32 rows go by LDS
// DMA (global_load_lds_dword: lane x's value is written to lds[r * D0 + x]
// without passing through registers), so a wavefront's row loads are all in
// flight together; the byte rows (occupancy) go through registers, all loaded
// before any is stored.  `wave` must be wavefront-uniform.
//
// dma_rows: a box inside an array of row pitch d0 and plane pitch d0 * d1,
// 32-bit indices (the caller checks that the array has fewer than 2^31
// entries).
template <int D0, int D1, int D2>
__device__ __forceinline__ void dma_rows(float* lds, const float* __restrict__ f, const LdsBox& bx, int d0, int d1, int wave, int lane) {
    static_assert(D0 <= 64, "one row per wavefront pass");
    for (int r = wave; r < D1 * D2; r += kSaWaves) {
        const int j = r % D1, k = r / D1;
        if (lane < D0) __builtin_amdgcn_global_load_lds(f + ((bx.o[2] + k) * d1 + bx.o[1] + j) * d0 + bx.o[0] + lane, lds + r * D0, 4, 0, 0);
    }
}

template <int D0, int D1, int D2>
struct ByteRows {
    static_assert(D0 <= 64, "one row per wavefront pass");
    static constexpr int kRows = D1 * D2;
    static constexpr int kPasses = (kRows + kSaWaves - 1) / kSaWaves;
    uint8_t v[kPasses];
    // value(i, j, k) at the box's entry (i, j, k) (global indices)
    template <class V>
    __device__ __forceinline__ void load(const LdsBox& bx, int wave, int lane, V value) {
#pragma unroll
        for (int p = 0; p < kPasses; ++p) {
            const int r = wave + p * kSaWaves;
            if (r < kRows && lane < D0) v[p] = value(bx.o[0] + lane, bx.o[1] + r % D1, bx.o[2] + r / D1);
        }
    }
    __device__ __forceinline__ void store(uint8_t* lds, int wave, int lane) const {
#pragma unroll
        for (int p = 0; p < kPasses; ++p) {
            const int r = wave + p * kSaWaves;
            if (r < kRows && lane < D0) lds[r * D0 + lane] = v[p];
        }
    }
};

// sample_linear on staged values; the stencil must lie in the box
__device__ __forceinline__ float staged_linear(const float* lds, const LdsBox& bx, const Stencil& s) {
    const int base = box_base(bx, s.b[0], s.b[1], s.b[2]);
    const int sy = bx.d[0], sz = bx.d[0] * bx.d[1];
    float c[8];
    c[0] = lds[base];
    c[1] = lds[base + 1];
    c[2] = lds[base + sy];
    c[3] = lds[base + sy + 1];
    c[4] = lds[base + sz];
    c[5] = lds[base + sz + 1];
    c[6] = lds[base + sz + sy];
    c[7] = lds[base + sz + sy + 1];
    return trilinear(c, s);
}

// sample_cubic of two staged fields at one stencil (which must lie in the box)
__device__ __forceinline__ void staged_cubic2(const float* la, const float* lb, const LdsBox& bx, const Stencil& s, float& va, float& vb) {
    const int sy = bx.d[0], sz = bx.d[0] * bx.d[1];
    const int base = box_base(bx, s.b[0] - 1, s.b[1] - 1, s.b[2] - 1);
    const Hermite wx = hermite(s.t[0]), wy = hermite(s.t[1]), wz = hermite(s.t[2]);
    float za[4], zb[4];
#pragma unroll
    for (int dz = 0; dz < 4; ++dz) {
        float ya[4], yb[4];
#pragma unroll
        for (int dy = 0; dy < 4; ++dy) {
            const int o = base + dz * sz + dy * sy;
            ya[dy] = monotonic_cubic(la[o], la[o + 1], la[o + 2], la[o + 3], wx);
            yb[dy] = monotonic_cubic(lb[o], lb[o + 1], lb[o + 2], lb[o + 3], wx);
        }
        za[dz] = monotonic_cubic(ya[0], ya[1], ya[2], ya[3], wy);
        zb[dz] = monotonic_cubic(yb[0], yb[1], yb[2], yb[3], wy);
    }
    va = monotonic_cubic(za[0], za[1], za[2], za[3], wz);
    vb = monotonic_cubic(zb[0], zb[1], zb[2], zb[3], wz);
}

// hits_solid (B:785-794) against the staged occupancy: 1 hit, 0 free,
// -1 when the hit cell lies outside the staged box (the caller falls back)
template <bool WRAP>
__device__ __forceinline__ int hit_staged(const Grid& g, const uint8_t* locc, const LdsBox& bx, float qx, float qy, float qz) {
    float q[3] = {qx, qy, qz};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (WRAP && g.flow_periodic[a]) q[a] = wrap_position(q[a], float(g.n[a]));
    }
#pragma unroll
    for (int a = 0; a < 3; ++a)
        if (q[a] < 0.0f || q[a] > float(g.n[a])) return 1;
    int c[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        c[a] = int(floorf(q[a]));
        if (c[a] == g.n[a]) c[a] = g.n[a] - 1;
    }
    if (c[0] < 0 || c[0] >= g.n[0] || c[1] < 0 || c[1] >= g.n[1] || c[2] < 0 || c[2] >= g.n[2]) return 1;
    bool inside = true;
#pragma unroll
    for (int a = 0; a < 3; ++a) inside = inside && c[a] >= bx.o[a] && c[a] < bx.o[a] + bx.d[a];
    if (!inside) return -1;
    return locc[box_base(bx, c[0], c[1], c[2])] != 0 ? 1 : 0;
}

// trace_rk2 (B:756-838) from face (i, j, k) of component C (C = -1: the
// centre of cell (i, j, k)) with linear velocity samples of the staged
// components and the staged hit test: false (nothing written) when a sample
// stencil or a hit-test cell lies outside its staged box.  The start samples
// read the taps next to the start (start_velocity), which every box holds.
// WRAP = false: no periodic flow axis (the wrap code compiled out).
template <int C, bool WRAP>
__device__ __forceinline__ bool staged_trace(const Grid& g, const float* lu, const LdsBox& bu, const float* lv, const LdsBox& bv, const float* lw,
                                             const LdsBox& bw, const uint8_t* locc, const LdsBox& bo, float s, int i, int j, int k, float& px,
                                             float& py, float& pz) {
    auto tap = [&](int comp, int a, int b, int c) {
        const float* l = comp == 0 ? lu : (comp == 1 ? lv : lw);
        const LdsBox& bx = comp == 0 ? bu : (comp == 1 ? bv : bw);
        return l[box_base(bx, a, b, c)];
    };
    const float3 v0 = start_velocity<C>(tap, i, j, k);
    const float sx = C == 0 ? float(i) : float(i) + 0.5f;
    const float sy = C == 1 ? float(j) : float(j) + 0.5f;
    const float sz = C == 2 ? float(k) : float(k) + 0.5f;
    const float hs = -0.5f * s;
    const float mx = fmaf(hs, v0.x, sx), my = fmaf(hs, v0.y, sy), mz = fmaf(hs, v0.z, sz);
    const Stencil m0 = locate<WRAP>(g, g.flow_periodic, 0, mx, my, mz);
    const Stencil m1 = locate<WRAP>(g, g.flow_periodic, 1, mx, my, mz);
    const Stencil m2 = locate<WRAP>(g, g.flow_periodic, 2, mx, my, mz);
    if (!(in_box<0, 1>(bu, m0) && in_box<0, 1>(bv, m1) && in_box<0, 1>(bw, m2))) return false;
    const float v1x = staged_linear(lu, bu, m0), v1y = staged_linear(lv, bv, m1), v1z = staged_linear(lw, bw, m2);
    const float tx = fmaf(-s, v1x, sx), ty = fmaf(-s, v1y, sy), tz = fmaf(-s, v1z, sz);
    px = tx;
    py = ty;
    pz = tz;
    const int hit = hit_staged<WRAP>(g, locc, bo, tx, ty, tz);
    if (hit < 0) return false;
    if (hit > 0) {
        float lo = 0.0f, hi = 1.0f;
#pragma nounroll
        for (int it = 0; it < 10; ++it) {
            const float mt = 0.5f * (lo + hi);
            const int hm = hit_staged<WRAP>(g, locc, bo, fmaf(tx - sx, mt, sx), fmaf(ty - sy, mt, sy), fmaf(tz - sz, mt, sz));
            if (hm < 0) return false;
            if (hm > 0) hi = mt;
            else lo = mt;
        }
        px = fmaf(tx - sx, lo, sx);
        py = fmaf(ty - sy, lo, sy);
        pz = fmaf(tz - sz, lo, sz);
    }
    return true;
}

// the global path for one cell (the body of advect_scalars_kernel)
__device__ __forceinline__ void advect_scalar_cell(const Grid& g, const float* __restrict__ u, const float* __restrict__ v,
                                                   const float* __restrict__ w, const uint8_t* __restrict__ occ, const float* __restrict__ Ttmp,
                                                   const float* __restrict__ rtmp, float* __restrict__ T, float* __restrict__ rho,
                                                   const float* __restrict__ solid_T, float s, int x, int y, int z) {
    const int64_t i = cidx(g, x, y, z);
    if (occ[i] != 0) {
        T[i] = solid_T[i];
        rho[i] = 0.0f;
        return;
    }
    const float3 p = trace_rk2_at<-1>(g, u, v, w, occ, s, x, y, z);
    float tv, rv;
    sample_scalar2(g, g.temperature, Ttmp, g.density, rtmp, p.x, p.y, p.z, g.cubic != 0, tv, rv);
    T[i] = tv;
    rho[i] = rv;
}

// WRAP = false: no periodic flow or scalar axis (the wrap code compiled out).
// Two phases through one LDS pool: the velocity and occupancy boxes serve
// the traces (each lane keeps its cells' sample stencils in registers),
// then the same LDS takes the T and rho boxes for the samples.  The pool is
// the larger of the two (27 KB, not the 56 KB of all five boxes at once),
// so three blocks fit a CU instead of two.
template <bool WRAP>
__device__ __forceinline__ void advect_scalars_staged_body(const VBlock& vb, Grid g, const float* __restrict__ u, const float* __restrict__ v,
                                                                           const float* __restrict__ w, const uint8_t* __restrict__ occ,
                                                                           const float* __restrict__ Ttmp, const float* __restrict__ rtmp,
                                                                           float* __restrict__ T, float* __restrict__ rho,
                                                                           const float* __restrict__ solid_T, const float* __restrict__ dt_ptr,
                                                                           int* __restrict__ fallback, unsigned* __restrict__ fallback_count,
                                                                           uint3 tiles) {
    constexpr int kS = (kSaX + 4) * (kSaY + 4) * (kSaZ + 4);
    constexpr int kU = (kSaX + 3) * (kSaY + 2) * (kSaZ + 2);
    constexpr int kV = (kSaX + 2) * (kSaY + 3) * (kSaZ + 2);
    constexpr int kW = (kSaX + 2) * (kSaY + 2) * (kSaZ + 3);
    constexpr int kO = (kSaX + 2) * (kSaY + 2) * (kSaZ + 2);
    constexpr int kPool = kU + kV + kW > 2 * kS ? kU + kV + kW : 2 * kS;
    constexpr int kCells = kSaZ / kSaZStep;   // cells per lane
    __shared__ float pool[kPool];
    __shared__ uint8_t so[kO];
    float* const su = pool;
    float* const sv = pool + kU;
    float* const sw = pool + kU + kV;
    float* const sT = pool;
    float* const sR = pool + kS;
    uint3 tile;
    unsigned id;
    if (!xcd_tile(vb, tiles, tile, id)) return;
    const int tid = threadIdx.x;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
    const int X0 = __builtin_amdgcn_readfirstlane(int(tile.x) * kSaX), Y0 = __builtin_amdgcn_readfirstlane(int(tile.y) * kSaY),
              Z0 = __builtin_amdgcn_readfirstlane(int(tile.z) * kSaZ);
    const LdsBox bs{{X0 - 2, Y0 - 2, Z0 - 2}, {kSaX + 4, kSaY + 4, kSaZ + 4}};
    const LdsBox bu{{X0 - 1, Y0 - 1, Z0 - 1}, {kSaX + 3, kSaY + 2, kSaZ + 2}};
    const LdsBox bv{{X0 - 1, Y0 - 1, Z0 - 1}, {kSaX + 2, kSaY + 3, kSaZ + 2}};
    const LdsBox bw{{X0 - 1, Y0 - 1, Z0 - 1}, {kSaX + 2, kSaY + 2, kSaZ + 3}};
    const LdsBox bo{{X0 - 1, Y0 - 1, Z0 - 1}, {kSaX + 2, kSaY + 2, kSaZ + 2}};
    // staging: every staged entry is what the global samplers' taps return;
    // a tile whose boxes lie inside the domain (most of them) stages the
    // arrays' values directly
    const int n0 = g.n[0], n1 = g.n[1], n2 = g.n[2];
    const bool inside = index32(g) && box_inside(bs, n0 - 1, n1 - 1, n2 - 1) && box_inside(bu, n0, n1 - 1, n2 - 1) &&
                        box_inside(bv, n0 - 1, n1, n2 - 1) && box_inside(bw, n0 - 1, n1 - 1, n2);
    if (inside) {
        ByteRows<kSaX + 2, kSaY + 2, kSaZ + 2> ro;
        ro.load(bo, wave, lane, [&](int i, int j, int k) { return occ[(k * n1 + j) * n0 + i]; });   // bo lies inside bs
        dma_rows<kSaX + 3, kSaY + 2, kSaZ + 2>(su, u, bu, n0 + 1, n1, wave, lane);
        dma_rows<kSaX + 2, kSaY + 3, kSaZ + 2>(sv, v, bv, n0, n1 + 1, wave, lane);
        dma_rows<kSaX + 2, kSaY + 2, kSaZ + 3>(sw, w, bw, n0, n1, wave, lane);
        ro.store(so, wave, lane);
    } else {
        // edge tiles (and grids too large for 32-bit indices): every entry with the ghost rules
        auto stage_face = [&](int comp, const float* __restrict__ f, float* lds, const LdsBox& bx) {
            const int n = box_size(bx);
            for (int q = tid; q < n; q += kSaThreads) {
                const int i = bx.o[0] + q % bx.d[0], j = bx.o[1] + (q / bx.d[0]) % bx.d[1], k = bx.o[2] + q / (bx.d[0] * bx.d[1]);
                lds[q] = load_face(g, f, comp, i, j, k);
            }
        };
        stage_face(0, u, su, bu);
        stage_face(1, v, sv, bv);
        stage_face(2, w, sw, bw);
        for (int q = tid; q < kO; q += kSaThreads) {
            const int i = bo.o[0] + q % bo.d[0], j = bo.o[1] + (q / bo.d[0]) % bo.d[1], k = bo.o[2] + q / (bo.d[0] * bo.d[1]);
            // entries outside the domain are never read (the hit test returns first)
            so[q] = (i >= 0 && i < g.n[0] && j >= 0 && j < g.n[1] && k >= 0 && k < g.n[2]) ? occ[cidx(g, i, j, k)] : uint8_t(1);
        }
    }
    __syncthreads();
    // phase 1: the traces; solid cells and fallbacks are finished here
    const int x = X0 + tid % kSaX, y = Y0 + (tid / kSaX) % kSaY;
    const bool column = x < g.n[0] && y < g.n[1];
    const float cps = *dt_ptr / g.h;   // cells moved per unit velocity and time step
    const bool cubic = g.cubic != 0;
    const bool same_periodic = g.temperature.periodic[0] == g.density.periodic[0] && g.temperature.periodic[1] == g.density.periodic[1] &&
                               g.temperature.periodic[2] == g.density.periodic[2];
    const int tz0 = tid / kSaPlane;
    Stencil st[kCells];
    bool sample[kCells];
#pragma unroll
    for (int c = 0; c < kCells; ++c) {
        sample[c] = false;
        const int z = Z0 + tz0 + c * kSaZStep;
        if (!column || z >= g.n[2]) continue;
        const int64_t i = cidx(g, x, y, z);
        if (so[box_base(bo, x, y, z)] != 0) {
            T[i] = solid_T[i];
            rho[i] = 0.0f;
            continue;
        }
        float px, py, pz;
        bool ok = same_periodic && staged_trace<-1, WRAP>(g, su, bu, sv, bv, sw, bw, so, bo, cps, x, y, z, px, py, pz);
        if (ok) {
            st[c] = locate<WRAP>(g, g.temperature.periodic, -1, px, py, pz);
            ok = cubic ? in_box<-1, 2>(bs, st[c]) : in_box<0, 1>(bs, st[c]);
        }
        if (!ok) {
            fallback[atomicAdd(fallback_count, 1u)] = int(i);
            continue;
        }
        sample[c] = true;
    }
    __syncthreads();   // the velocity boxes are read: the pool takes the scalars
    if (inside) {
        dma_rows<kSaX + 4, kSaY + 4, kSaZ + 4>(sT, Ttmp, bs, n0, n1, wave, lane);
        dma_rows<kSaX + 4, kSaY + 4, kSaZ + 4>(sR, rtmp, bs, n0, n1, wave, lane);
    } else {
        for (int q = tid; q < kS; q += kSaThreads) {
            const int i = bs.o[0] + q % bs.d[0], j = bs.o[1] + (q / bs.d[0]) % bs.d[1], k = bs.o[2] + q / (bs.d[0] * bs.d[1]);
            sT[q] = load_scalar(g, g.temperature, Ttmp, i, j, k);
            sR[q] = load_scalar(g, g.density, rtmp, i, j, k);
        }
    }
    __syncthreads();
    // phase 2: the samples
#pragma unroll
    for (int c = 0; c < kCells; ++c) {
        if (!sample[c]) continue;
        const int64_t i = cidx(g, x, y, Z0 + tz0 + c * kSaZStep);
        float tv, rv;
        if (cubic) {
            staged_cubic2(sT, sR, bs, st[c], tv, rv);
        } else {
            tv = staged_linear(sT, bs, st[c]);
            rv = staged_linear(sR, bs, st[c]);
        }
        T[i] = tv;
        rho[i] = rv;
    }
}

template <bool WRAP>
__global__ __launch_bounds__(kSaThreads) void advect_scalars_staged_kernel(Grid g, const float* __restrict__ u, const float* __restrict__ v,
                                                                           const float* __restrict__ w, const uint8_t* __restrict__ occ,
                                                                           const float* __restrict__ Ttmp, const float* __restrict__ rtmp,
                                                                           float* __restrict__ T, float* __restrict__ rho,
                                                                           const float* __restrict__ solid_T, const float* __restrict__ dt_ptr,
                                                                           int* __restrict__ fallback, unsigned* __restrict__ fallback_count,
                                                                           uint3 tiles) { advect_scalars_staged_body<WRAP>(hw_block(), g, u, v, w, occ, Ttmp, rtmp, T, rho, solid_T, dt_ptr, fallback, fallback_count, tiles); }

template <bool WRAP>
struct advect_scalars_staged_k {
    template <class... A>
    __device__ __forceinline__ static void run(const VBlock& vb, A... a) { advect_scalars_staged_body<WRAP>(vb, a...); }
};

// the cells the staged kernel could not advect from LDS, by the global path
__device__ __forceinline__ void advect_scalars_fallback_body(const VBlock& vb, Grid g, const float* __restrict__ u, const float* __restrict__ v,
                                                                      const float* __restrict__ w, const uint8_t* __restrict__ occ,
                                                                      const float* __restrict__ Ttmp, const float* __restrict__ rtmp,
                                                                      float* __restrict__ T, float* __restrict__ rho,
                                                                      const float* __restrict__ solid_T, const float* __restrict__ dt_ptr,
                                                                      const int* __restrict__ fallback, const unsigned* __restrict__ fallback_count) {
    const unsigned n = *fallback_count;
    const float cps = *dt_ptr / g.h;
    for (unsigned q = vb.idx.x * 256 + threadIdx.x; q < n; q += vb.dim.x * 256) {
        const int i = fallback[q];
        const int x = i % g.n[0], y = (i / g.n[0]) % g.n[1], z = i / (g.n[0] * g.n[1]);
        advect_scalar_cell(g, u, v, w, occ, Ttmp, rtmp, T, rho, solid_T, cps, x, y, z);
    }
}

__global__ __launch_bounds__(256) void advect_scalars_fallback_kernel(Grid g, const float* __restrict__ u, const float* __restrict__ v,
                                                                      const float* __restrict__ w, const uint8_t* __restrict__ occ,
                                                                      const float* __restrict__ Ttmp, const float* __restrict__ rtmp,
                                                                      float* __restrict__ T, float* __restrict__ rho,
                                                                      const float* __restrict__ solid_T, const float* __restrict__ dt_ptr,
                                                                      const int* __restrict__ fallback, const unsigned* __restrict__ fallback_count) { advect_scalars_fallback_body(hw_block(), g, u, v, w, occ, Ttmp, rtmp, T, rho, solid_T, dt_ptr, fallback, fallback_count); }

struct advect_scalars_fallback_k {
    template <class... A>
    __device__ __forceinline__ static void run(const VBlock& vb, A... a) { advect_scalars_fallback_body(vb, a...); }
};

__device__ __forceinline__ void zero_count_body(const VBlock& vb, unsigned* __restrict__ count) { *count = 0u; }

__global__ void zero_count_kernel(unsigned* __restrict__ count) { zero_count_body(hw_block(), count); }

struct zero_count_k {
    template <class... A>
    __device__ __forceinline__ static void run(const VBlock& vb, A... a) { zero_count_body(vb, a...); }
};

// sample_cubic of one staged field (the stencil must lie in the box)
__device__ __forceinline__ float staged_cubic(const float* lds, const LdsBox& bx, const Stencil& s) {
    const int sy = bx.d[0], sz = bx.d[0] * bx.d[1];
    const int base = box_base(bx, s.b[0] - 1, s.b[1] - 1, s.b[2] - 1);
    const Hermite wx = hermite(s.t[0]), wy = hermite(s.t[1]), wz = hermite(s.t[2]);
    float zs[4];
#pragma unroll
    for (int dz = 0; dz < 4; ++dz) {
        float ys[4];
#pragma unroll
        for (int dy = 0; dy < 4; ++dy) {
            const int o = base + dz * sz + dy * sy;
            ys[dy] = monotonic_cubic(lds[o], lds[o + 1], lds[o + 2], lds[o + 3], wx);
        }
        zs[dz] = monotonic_cubic(ys[0], ys[1], ys[2], ys[3], wy);
    }
    return monotonic_cubic(zs[0], zs[1], zs[2], zs[3], wz);
}

// ---- LDS-staged velocity advection (the same results as advect_face_kernel).
// A block owns a 32 x 8 x 4 tile of the faces of component C (face indices
// below n on every axis: face n is always a boundary face).  For a backtrace
// shorter than one cell it stages component C over the tile plus a halo of 2
// (the cubic taps of the foot, the linear taps of C's own velocity samples),
// the other two components over the tile plus a halo of 1 (their linear taps
// at the start and the midpoint) and the occupancy plus a halo of 1 (the hit
// test and the solid-face test).  Faces whose trace or foot stencil leaves the
// staged boxes go to the fallback list.
template <int C, bool WRAP>
__device__ __forceinline__ void advect_face_staged_body(const VBlock& vb, Grid g, const float* __restrict__ u, const float* __restrict__ v,
                                                                        const float* __restrict__ w, float* __restrict__ dst,
                                                                        const uint8_t* __restrict__ occ, const float* __restrict__ solid_vel,
                                                                        const float* __restrict__ dt_ptr, int* __restrict__ fallback,
                                                                        unsigned* __restrict__ fallback_count, uint3 tiles) {
    constexpr int kF = (kSaX + 4) * (kSaY + 4) * (kSaZ + 4);
    constexpr int kG = (kSaX + 3) * (kSaY + 3) * (kSaZ + 3);
    constexpr int kO = (kSaX + 2) * (kSaY + 2) * (kSaZ + 2);
    // the other two components, in increasing order
    constexpr int D0 = C == 0 ? 1 : 0, D1 = C == 2 ? 1 : 2;
    __shared__ float sf[kF], sg0[kG], sg1[kG];
    __shared__ uint8_t so[kO];
    uint3 tile;
    unsigned id;
    if (!xcd_tile(vb, tiles, tile, id)) return;
    const int tid = threadIdx.x;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
    const int X0 = __builtin_amdgcn_readfirstlane(int(tile.x) * kSaX), Y0 = __builtin_amdgcn_readfirstlane(int(tile.y) * kSaY),
              Z0 = __builtin_amdgcn_readfirstlane(int(tile.z) * kSaZ);
    const LdsBox bf{{X0 - 2, Y0 - 2, Z0 - 2}, {kSaX + 4, kSaY + 4, kSaZ + 4}};
    const LdsBox bg{{X0 - 1, Y0 - 1, Z0 - 1}, {kSaX + 3, kSaY + 3, kSaZ + 3}};
    const LdsBox bo{{X0 - 1, Y0 - 1, Z0 - 1}, {kSaX + 2, kSaY + 2, kSaZ + 2}};
    const float* fc = C == 0 ? u : (C == 1 ? v : w);
    const float* f0 = D0 == 0 ? u : v;
    const float* f1 = D1 == 1 ? v : w;
    // a tile whose boxes lie inside the arrays (most of them) stages the
    // arrays' values directly; the others go through the ghost rules
    const int n0 = g.n[0], n1 = g.n[1], n2 = g.n[2];
    auto inside_for = [&](int comp, const LdsBox& bx) {
        return box_inside(bx, n0 - 1 + (comp == 0), n1 - 1 + (comp == 1), n2 - 1 + (comp == 2));
    };
    const bool inside = index32(g) && inside_for(C, bf) && inside_for(D0, bg) && inside_for(D1, bg) && box_inside(bo, n0 - 1, n1 - 1, n2 - 1);
    if (inside) {
        ByteRows<kSaX + 2, kSaY + 2, kSaZ + 2> ro;
        ro.load(bo, wave, lane, [&](int i, int j, int k) { return occ[(k * n1 + j) * n0 + i]; });
        dma_rows<kSaX + 4, kSaY + 4, kSaZ + 4>(sf, fc, bf, n0 + (C == 0), n1 + (C == 1), wave, lane);
        dma_rows<kSaX + 3, kSaY + 3, kSaZ + 3>(sg0, f0, bg, n0 + (D0 == 0), n1 + (D0 == 1), wave, lane);
        dma_rows<kSaX + 3, kSaY + 3, kSaZ + 3>(sg1, f1, bg, n0 + (D1 == 0), n1 + (D1 == 1), wave, lane);
        ro.store(so, wave, lane);
    } else {
        // edge tiles (and grids too large for 32-bit indices): every entry with the ghost rules
        auto stage_face = [&](int comp, const float* __restrict__ f, float* lds, const LdsBox& bx) {
            const int n = box_size(bx);
            for (int q = tid; q < n; q += kSaThreads) {
                const int i = bx.o[0] + q % bx.d[0], j = bx.o[1] + (q / bx.d[0]) % bx.d[1], k = bx.o[2] + q / (bx.d[0] * bx.d[1]);
                lds[q] = load_face(g, f, comp, i, j, k);
            }
        };
        stage_face(C, fc, sf, bf);
        stage_face(D0, f0, sg0, bg);
        stage_face(D1, f1, sg1, bg);
        for (int q = tid; q < kO; q += kSaThreads) {
            const int i = bo.o[0] + q % bo.d[0], j = bo.o[1] + (q / bo.d[0]) % bo.d[1], k = bo.o[2] + q / (bo.d[0] * bo.d[1]);
            // entries outside the domain are never read (both tests check the domain first)
            so[q] = (i >= 0 && i < g.n[0] && j >= 0 && j < g.n[1] && k >= 0 && k < g.n[2]) ? occ[cidx(g, i, j, k)] : uint8_t(1);
        }
    }
    __syncthreads();
    const int i = X0 + tid % kSaX, j = Y0 + (tid / kSaX) % kSaY;
    if (i >= g.n[0] || j >= g.n[1]) return;
    // staged copies of u, v, w in component order
    const float* lu = C == 0 ? sf : sg0;
    const LdsBox& bu = C == 0 ? bf : bg;
    const float* lv = C == 1 ? sf : (C == 0 ? sg0 : sg1);
    const LdsBox& bv = C == 1 ? bf : bg;
    const float* lw = C == 2 ? sf : sg1;
    const LdsBox& bw = C == 2 ? bf : bg;
    const float cps = *dt_ptr / g.h;   // cells moved per unit velocity and time step
    const bool cubic = g.cubic != 0;
    for (int tz = tid / kSaPlane; tz < kSaZ; tz += kSaZStep) {
        const int k = Z0 + tz;
        if (k >= g.n[2]) break;
        const int c3[3] = {i, j, k};
        if (boundary_face(g, C, c3[C])) continue;
        const int64_t fi = fidx(g, C, i, j, k);
        // solid_face_override (B:1007-1012) returns false when both cells
        // beside the face lie in the domain and are fluid
        bool free_face = false;
        if (c3[C] >= 1 && c3[C] < g.n[C]) {
            int l[3] = {i, j, k};
            l[C] -= 1;
            free_face = so[box_base(bo, l[0], l[1], l[2])] == 0 && so[box_base(bo, i, j, k)] == 0;
        }
        if (!free_face) {
            float value;
            if (solid_face_override(g, occ, solid_vel, C, i, j, k, value)) {
                dst[fi] = value;
                continue;
            }
        }
        float px, py, pz;
        bool ok = staged_trace<C, WRAP>(g, lu, bu, lv, bv, lw, bw, so, bo, cps, i, j, k, px, py, pz);
        Stencil st{};
        if (ok) {
            st = locate<WRAP>(g, g.flow_periodic, C, px, py, pz);
            ok = cubic ? in_box<-1, 2>(bf, st) : in_box<0, 1>(bf, st);
        }
        if (!ok) {
            fallback[atomicAdd(fallback_count, 1u)] = int(fi);
            continue;
        }
        dst[fi] = cubic ? staged_cubic(sf, bf, st) : staged_linear(sf, bf, st);
    }
}

template <int C, bool WRAP>
__global__ __launch_bounds__(kSaThreads) void advect_face_staged_kernel(Grid g, const float* __restrict__ u, const float* __restrict__ v,
                                                                        const float* __restrict__ w, float* __restrict__ dst,
                                                                        const uint8_t* __restrict__ occ, const float* __restrict__ solid_vel,
                                                                        const float* __restrict__ dt_ptr, int* __restrict__ fallback,
                                                                        unsigned* __restrict__ fallback_count, uint3 tiles) { advect_face_staged_body<C, WRAP>(hw_block(), g, u, v, w, dst, occ, solid_vel, dt_ptr, fallback, fallback_count, tiles); }

template <int C, bool WRAP>
struct advect_face_staged_k {
    template <class... A>
    __device__ __forceinline__ static void run(const VBlock& vb, A... a) { advect_face_staged_body<C, WRAP>(vb, a...); }
};

// the faces the staged kernel could not advect from LDS, by the global path
template <int C>
__device__ __forceinline__ void advect_face_fallback_body(const VBlock& vb, Grid g, const float* __restrict__ u, const float* __restrict__ v,
                                                                   const float* __restrict__ w, float* __restrict__ dst, const uint8_t* __restrict__ occ,
                                                                   const float* __restrict__ dt_ptr, const int* __restrict__ fallback,
                                                                   const unsigned* __restrict__ fallback_count) {
    const unsigned n = *fallback_count;
    const float cps = *dt_ptr / g.h;
    const int d0 = g.n[0] + (C == 0), d1 = g.n[1] + (C == 1);
    for (unsigned q = vb.idx.x * 256 + threadIdx.x; q < n; q += vb.dim.x * 256) {
        const int fi = fallback[q];
        const int i = fi % d0, j = (fi / d0) % d1, k = fi / (d0 * d1);
        dst[fi] = advect_face_value<C>(g, u, v, w, occ, cps, i, j, k);
    }
}

template <int C>
__global__ __launch_bounds__(256) void advect_face_fallback_kernel(Grid g, const float* __restrict__ u, const float* __restrict__ v,
                                                                   const float* __restrict__ w, float* __restrict__ dst, const uint8_t* __restrict__ occ,
                                                                   const float* __restrict__ dt_ptr, const int* __restrict__ fallback,
                                                                   const unsigned* __restrict__ fallback_count) { advect_face_fallback_body<C>(hw_block(), g, u, v, w, dst, occ, dt_ptr, fallback, fallback_count); }

template <int C>
struct advect_face_fallback_k {
    template <class... A>
    __device__ __forceinline__ static void run(const VBlock& vb, A... a) { advect_face_fallback_body<C>(vb, a...); }
};

// The staged kernels come in two forms: with the periodic wrap, and (for grids
// with no periodic flow or scalar axis) without it.
bool staged_wrap(const Grid& g) {
    for (int a = 0; a < 3; ++a)
        if (g.flow_periodic[a] || g.temperature.periodic[a] || g.density.periodic[a]) return true;
    return false;
}

template <int C>
void launch_advect(Context& c, int block_index) {
    const Grid& g = c.grid;
    if (block_index == kAdvectStaged) {
        // tiles over face indices below n (at most `cells` faces reach the list)
        const dim3 tiles(unsigned((g.n[0] + kSaX - 1) / kSaX), unsigned((g.n[1] + kSaY - 1) / kSaY), unsigned((g.n[2] + kSaZ - 1) / kSaZ));
        zero_count_kernel<<<1, 1, 0, c.stream>>>(c.advect_fallback_count);
        auto* kernel = staged_wrap(g) ? advect_face_staged_kernel<C, true> : advect_face_staged_kernel<C, false>;
        kernel<<<xcd_blocks(tiles), kSaThreads, 0, c.stream>>>(g, c.face[0], c.face[1], c.face[2], c.face_star[C], c.occupancy, c.solid_velocity[C],
                                                               c.dt_device, c.advect_fallback, c.advect_fallback_count,
                                                               make_uint3(tiles.x, tiles.y, tiles.z));
        advect_face_fallback_kernel<C><<<1024, 256, 0, c.stream>>>(g, c.face[0], c.face[1], c.face[2], c.face_star[C], c.occupancy, c.dt_device,
                                                                   c.advect_fallback, c.advect_fallback_count);
        hip_check(hipGetLastError(), "staged velocity advection");
        return;
    }
    const dim3 block = kAdvectBlocks[block_index];
    const dim3 grid = grid3(g.n[0] + (C == 0), g.n[1] + (C == 1), g.n[2] + (C == 2), block);
    advect_face_kernel<C><<<grid, block, 0, c.stream>>>(g, c.face[0], c.face[1], c.face[2], c.face_star[C], c.occupancy, c.solid_velocity[C],
                                                        c.dt_device);
}

// T and rho advection from (Ttmp, rtmp) into (T, rho)
void launch_advect_scalars(Context& c, int block_index, float* T, float* rho) {
    const Grid& g = c.grid;
    if (block_index == kAdvectStaged) {
        const dim3 tiles(unsigned((g.n[0] + kSaX - 1) / kSaX), unsigned((g.n[1] + kSaY - 1) / kSaY), unsigned((g.n[2] + kSaZ - 1) / kSaZ));
        zero_count_kernel<<<1, 1, 0, c.stream>>>(c.advect_fallback_count);
        auto* kernel = staged_wrap(g) ? advect_scalars_staged_kernel<true> : advect_scalars_staged_kernel<false>;
        kernel<<<xcd_blocks(tiles), kSaThreads, 0, c.stream>>>(g, c.face[0], c.face[1], c.face[2], c.occupancy, c.temperature_tmp, c.density_tmp, T,
                                                               rho, c.solid_temperature, c.dt_device, c.advect_fallback, c.advect_fallback_count,
                                                               make_uint3(tiles.x, tiles.y, tiles.z));
        advect_scalars_fallback_kernel<<<1024, 256, 0, c.stream>>>(g, c.face[0], c.face[1], c.face[2], c.occupancy, c.temperature_tmp,
                                                                 c.density_tmp, T, rho, c.solid_temperature, c.dt_device, c.advect_fallback,
                                                                 c.advect_fallback_count);
        hip_check(hipGetLastError(), "staged scalar advection");
        return;
    }
    const dim3 block = kAdvectBlocks[block_index];
    advect_scalars_kernel<<<grid3(g.n[0], g.n[1], g.n[2], block), block, 0, c.stream>>>(g, c.face[0], c.face[1], c.face[2], c.occupancy,
                                                                                      c.temperature_tmp, c.density_tmp, T, rho,
                                                                                      c.solid_temperature, c.dt_device);
}

template <int C>
void launch_project(Context& c) {
    const Grid& g = c.grid;
    const dim3 grid = grid3(g.n[0] + (C == 0), g.n[1] + (C == 1), g.n[2] + (C == 2), kCellBlock);
    project_kernel<C><<<grid, kCellBlock, 0, c.stream>>>(g, c.face_star[C], c.face[C], c.pressure, c.occupancy, c.solid_velocity[C], c.dt_device);
}

dim3 cell_grid(const Grid& g) { return grid3(g.n[0], g.n[1], g.n[2], kCellBlock); }

}  // namespace

void enqueue_fill(Context& c, float* p, float value, int64_t n) {
    fill_kernel<<<grid1(n), kBlock, 0, c.stream>>>(p, value, n);
    hip_check(hipGetLastError(), "fill_kernel");
}

void enqueue_fill_u8(Context& c, uint8_t* p, uint8_t value, int64_t n) {
    fill_u8_kernel<<<grid1(n), kBlock, 0, c.stream>>>(p, value, n);
    hip_check(hipGetLastError(), "fill_u8_kernel");
}

void enqueue_occupancy_derived(Context& c) {
    c.anchor_host = -1;                           // the anchor is recomputed below
    c.solves_done = std::min(c.solves_done, 1);   // the warm-start history predates this occupancy
    c.mg.built = 0;                               // the multilevel hierarchy is rebuilt from the new occupancy
    const Grid& g = c.grid;
    occupancy_float_kernel<<<grid1(g.cells), kBlock, 0, c.stream>>>(c.occupancy, c.occupancy_float, g.cells);
    anchor_reset_kernel<<<1, 1, 0, c.stream>>>(c.anchor, int(g.cells), c.solid_count);
    anchor_kernel<<<grid1(g.cells), kBlock, 0, c.stream>>>(c.occupancy, c.anchor, g.cells);
    stencil_flags_kernel<<<cell_grid(g), kCellBlock, 0, c.stream>>>(g, c.occupancy, c.anchor, c.stencil_flags);
    nearest_fluid_kernel<<<cell_grid(g), kCellBlock, 0, c.stream>>>(g, c.occupancy, c.solid_pairs, c.solid_count);
    zero_identity_rows_kernel<<<grid1(g.cells), kBlock, 0, c.stream>>>(c.stencil_flags, c.x64[0], c.x64[1], c.x64_hist, g.cells);
    hip_check(hipGetLastError(), "occupancy-derived kernels");
    c.diagnostics_dirty = true;
}

void enqueue_time_step(Context& c) {
    const Grid& g = c.grid;
    if (c.options.adaptive_dt) {
        const int64_t total = g.faces[0] + g.faces[1] + g.faces[2];
        const unsigned blocks = unsigned(std::min<int64_t>(2048, (total + kBlock - 1) / kBlock));
        max_speed_kernel<<<blocks, kBlock, 0, c.stream>>>(c.face[0], g.faces[0], c.face[1], g.faces[1], c.face[2], g.faces[2], c.max_speed_bits);
    }
    time_step_kernel<<<1, 1, 0, c.stream>>>(c.max_speed_bits, c.max_speed, c.dt_device, c.options.adaptive_dt ? 1 : 0, c.options.cfl, g.h,
                                            c.options.dt_max, c.config.dt);
    hip_check(hipGetLastError(), "time_step_kernel");
    if (c.options.pressure_mode != SMOKE_SIMULATION_PRESSURE_FIXED_ITERATIONS && c.host_step) {
        // tolerance mode (never captured into a graph): dt and max |u_face|
        // for the solver stats, behind their own event
        hip_check(hipMemcpyAsync(&c.host_step[0], c.dt_device, sizeof(float), hipMemcpyDeviceToHost, c.stream), "step dt");
        hip_check(hipMemcpyAsync(&c.host_step[1], c.max_speed, sizeof(float), hipMemcpyDeviceToHost, c.stream), "step speed");
        hip_check(hipEventRecord(c.step_event, c.stream), "step event");
        c.step_info_pending = true;
    } else {
        c.step_info_pending = false;
    }
}

void enqueue_forces(Context& c) {
    const Grid& g = c.grid;
    if (c.fused_force) {
        const uint3 tiles = force_tiles(g);
        auto kernel = index32_host(g) ? force_fused_kernel<true> : force_fused_kernel<false>;
        kernel<<<xcd_blocks(dim3(tiles.x, tiles.y, tiles.z)), kFfTx * kFfTy, 0, c.stream>>>(
            g, c.face[0], c.face[1], c.face[2], c.occupancy, c.force_user[0], c.force_user[1], c.force_user[2], c.density, c.temperature,
            c.solid_temperature, c.config.buoyancy_density_factor, c.config.buoyancy_temperature_factor, c.config.ambient_temperature,
            c.config.vorticity_confinement, c.cell_force[0], c.cell_force[1], c.cell_force[2], tiles);
    } else {
        const uint3 tiles = cell_tiles(g);
        vorticity_kernel<<<cell_tile_blocks(g), kCellBlock, 0, c.stream>>>(g, c.face[0], c.face[1], c.face[2], c.occupancy, c.vorticity[0],
                                                                             c.vorticity[1], c.vorticity[2], c.vorticity_magnitude, c.temperature,
                                                                             c.solid_temperature, tiles);
        cell_force_kernel<<<cell_tile_blocks(g), kCellBlock, 0, c.stream>>>(
            g, c.force_user[0], c.force_user[1], c.force_user[2], c.density, c.temperature, c.occupancy, c.vorticity[0], c.vorticity[1],
            c.vorticity[2], c.vorticity_magnitude, c.config.buoyancy_density_factor, c.config.buoyancy_temperature_factor,
            c.config.ambient_temperature, c.config.vorticity_confinement, c.cell_force[0], c.cell_force[1], c.cell_force[2], tiles);
    }
    c.diagnostics_dirty = true;   // the velocity changes; the views are recomputed when read
    launch_face_force<0>(c);
    launch_face_force<1>(c);
    launch_face_force<2>(c);
    launch_boundaries(c, c.face);
    hip_check(hipGetLastError(), "force stage");
}

void enqueue_advect_velocity(Context& c) {
    launch_advect<0>(c, c.advect_block[0]);
    launch_advect<1>(c, c.advect_block[1]);
    launch_advect<2>(c, c.advect_block[2]);
    launch_boundaries(c, c.face_star);
    hip_check(hipGetLastError(), "velocity advection");
}

void enqueue_project(Context& c) {
    launch_project<0>(c);
    launch_project<1>(c);
    launch_project<2>(c);
    launch_boundaries(c, c.face);
    hip_check(hipGetLastError(), "projection");
    c.diagnostics_dirty = true;
}

void enqueue_scalars(Context& c) {
    const Grid& g = c.grid;
    scalar_sources_kernel<<<grid1(g.cells), kBlock, 0, c.stream>>>(c.temperature, c.temperature_source, c.density, c.density_source,
                                                                     c.temperature_tmp, c.density_tmp, c.dt_device, g.cells);
    launch_advect_scalars(c, c.advect_block[3], c.temperature, c.density);
    density_fill_kernel<<<256, kBlock, 0, c.stream>>>(c.density, c.solid_pairs, c.solid_count);
    hip_check(hipGetLastError(), "scalar stage");
    c.diagnostics_dirty = true;
}

void enqueue_diagnostics(Context& c) {
    const Grid& g = c.grid;
    vorticity_kernel<<<cell_tile_blocks(g), kCellBlock, 0, c.stream>>>(g, c.face[0], c.face[1], c.face[2], c.occupancy, c.vorticity[0],
                                                                                 c.vorticity[1], c.vorticity[2], c.vorticity_magnitude, nullptr,
                                                                                 nullptr, cell_tiles(g));
    diagnostics_kernel<<<cell_grid(g), kCellBlock, 0, c.stream>>>(g, c.face[0], c.face[1], c.face[2], c.occupancy, c.center_velocity[0],
                                                                   c.center_velocity[1], c.center_velocity[2], c.velocity_magnitude, c.divergence);
    hip_check(hipGetLastError(), "diagnostics");
    c.diagnostics_dirty = false;
}

void enqueue_plume_sources(Context& c, const SmokeSimulationPlumeSource& src, uint32_t step) {
    plume_kernel<<<cell_grid(c.grid), kCellBlock, 0, c.stream>>>(c.grid, src, step, c.density_source, c.temperature_source);
    hip_check(hipGetLastError(), "plume_kernel");
}

// ---- batched stages: one launch per kernel over the plumes of a batch
// (batched.hpp), with the launch shapes of the single-context path above.
namespace {

constexpr int kCellThreads = 256;   // kCellBlock

template <int C>
void batch_boundaries(const std::vector<Context*>& P, float* (Context::*faces)[3], TableArena& A, std::vector<BatchedLaunch>& out) {
    constexpr int A1 = (C + 1) % 3, A2 = (C + 2) % 3;
    const dim3 block(64, 4, 1);
    LaunchBuilder<boundary_faces_k<C>, 256, Grid, float*> L;
    for (unsigned p = 0; p < P.size(); ++p) {
        const Context& c = *P[p];
        L.add(p, grid3(c.grid.n[A1], c.grid.n[A2], 1, block), c.grid, (c.*faces)[C]);
    }
    if (!L.empty()) out.push_back(L.finish(A, block));
}

void batch_all_boundaries(const std::vector<Context*>& P, float* (Context::*faces)[3], TableArena& A, std::vector<BatchedLaunch>& out) {
    batch_boundaries<0>(P, faces, A, out);
    batch_boundaries<1>(P, faces, A, out);
    batch_boundaries<2>(P, faces, A, out);
}

// `blocks` of the fallback passes: their loops stride over the list, so any
// count gives the same results; a plume needs no more than its cells / 256
unsigned fallback_blocks(const Context& c) { return unsigned(std::min<int64_t>(1024, (c.grid.cells + 255) / 256)); }

template <int C>
void batch_advect_face(const std::vector<Context*>& P, TableArena& A, std::vector<BatchedLaunch>& out) {
    LaunchBuilder<zero_count_k, 64, unsigned*> Z;
    LaunchBuilder<advect_face_staged_k<C, true>, kSaThreads, Grid, const float*, const float*, const float*, float*, const uint8_t*, const float*,
                  const float*, int*, unsigned*, uint3>
        SW;
    LaunchBuilder<advect_face_staged_k<C, false>, kSaThreads, Grid, const float*, const float*, const float*, float*, const uint8_t*, const float*,
                  const float*, int*, unsigned*, uint3>
        S;
    LaunchBuilder<advect_face_fallback_k<C>, 256, Grid, const float*, const float*, const float*, float*, const uint8_t*, const float*, const int*,
                  const unsigned*>
        F;
    for (unsigned p = 0; p < P.size(); ++p) {
        const Context& c = *P[p];
        if (c.advect_block[C] != kAdvectStaged) continue;
        const Grid& g = c.grid;
        const dim3 tiles(unsigned((g.n[0] + kSaX - 1) / kSaX), unsigned((g.n[1] + kSaY - 1) / kSaY), unsigned((g.n[2] + kSaZ - 1) / kSaZ));
        Z.add(p, dim3(1), c.advect_fallback_count);
        auto add = [&](auto& L) {
            L.add(p, dim3(xcd_blocks(tiles)), g, c.face[0], c.face[1], c.face[2], c.face_star[C], c.occupancy, c.solid_velocity[C], c.dt_device,
                  c.advect_fallback, c.advect_fallback_count, make_uint3(tiles.x, tiles.y, tiles.z));
        };
        if (staged_wrap(g)) add(SW);
        else add(S);
        F.add(p, dim3(fallback_blocks(c)), g, c.face[0], c.face[1], c.face[2], c.face_star[C], c.occupancy, c.dt_device, c.advect_fallback,
              c.advect_fallback_count);
    }
    if (!Z.empty()) {
        out.push_back(Z.finish(A, dim3(1)));
        if (!SW.empty()) out.push_back(SW.finish(A, dim3(kSaThreads)));
        if (!S.empty()) out.push_back(S.finish(A, dim3(kSaThreads)));
        out.push_back(F.finish(A, dim3(256)));
    }
    for (int b = 0; b < kAdvectBlockCount; ++b) {
        const dim3 block = kAdvectBlocks[b];
        // the gather blocks go up to 512 threads (8 x 8 x 8)
        LaunchBuilder<advect_face_k<C>, 1024, Grid, const float*, const float*, const float*, float*, const uint8_t*, const float*, const float*> G;
        for (unsigned p = 0; p < P.size(); ++p) {
            const Context& c = *P[p];
            if (c.advect_block[C] != b) continue;
            const Grid& g = c.grid;
            G.add(p, grid3(g.n[0] + (C == 0), g.n[1] + (C == 1), g.n[2] + (C == 2), block), g, c.face[0], c.face[1], c.face[2], c.face_star[C],
                  c.occupancy, c.solid_velocity[C], c.dt_device);
        }
        if (!G.empty()) out.push_back(G.finish(A, block));
    }
}

void batch_advect_scalars(const std::vector<Context*>& P, TableArena& A, std::vector<BatchedLaunch>& out) {
    LaunchBuilder<zero_count_k, 64, unsigned*> Z;
    LaunchBuilder<advect_scalars_staged_k<true>, kSaThreads, Grid, const float*, const float*, const float*, const uint8_t*, const float*,
                  const float*, float*, float*, const float*, const float*, int*, unsigned*, uint3>
        SW;
    LaunchBuilder<advect_scalars_staged_k<false>, kSaThreads, Grid, const float*, const float*, const float*, const uint8_t*, const float*,
                  const float*, float*, float*, const float*, const float*, int*, unsigned*, uint3>
        S;
    LaunchBuilder<advect_scalars_fallback_k, 256, Grid, const float*, const float*, const float*, const uint8_t*, const float*, const float*, float*,
                  float*, const float*, const float*, const int*, const unsigned*>
        F;
    for (unsigned p = 0; p < P.size(); ++p) {
        const Context& c = *P[p];
        if (c.advect_block[3] != kAdvectStaged) continue;
        const Grid& g = c.grid;
        const dim3 tiles(unsigned((g.n[0] + kSaX - 1) / kSaX), unsigned((g.n[1] + kSaY - 1) / kSaY), unsigned((g.n[2] + kSaZ - 1) / kSaZ));
        Z.add(p, dim3(1), c.advect_fallback_count);
        auto add = [&](auto& L) {
            L.add(p, dim3(xcd_blocks(tiles)), g, c.face[0], c.face[1], c.face[2], c.occupancy, c.temperature_tmp, c.density_tmp, c.temperature,
                  c.density, c.solid_temperature, c.dt_device, c.advect_fallback, c.advect_fallback_count, make_uint3(tiles.x, tiles.y, tiles.z));
        };
        if (staged_wrap(g)) add(SW);
        else add(S);
        F.add(p, dim3(fallback_blocks(c)), g, c.face[0], c.face[1], c.face[2], c.occupancy, c.temperature_tmp, c.density_tmp, c.temperature, c.density,
              c.solid_temperature, c.dt_device, c.advect_fallback, c.advect_fallback_count);
    }
    if (!Z.empty()) {
        out.push_back(Z.finish(A, dim3(1)));
        if (!SW.empty()) out.push_back(SW.finish(A, dim3(kSaThreads)));
        if (!S.empty()) out.push_back(S.finish(A, dim3(kSaThreads)));
        out.push_back(F.finish(A, dim3(256)));
    }
    for (int b = 0; b < kAdvectBlockCount; ++b) {
        const dim3 block = kAdvectBlocks[b];
        LaunchBuilder<advect_scalars_k, 1024, Grid, const float*, const float*, const float*, const uint8_t*, const float*, const float*, float*, float*,
                      const float*, const float*>
            G;
        for (unsigned p = 0; p < P.size(); ++p) {
            const Context& c = *P[p];
            if (c.advect_block[3] != b) continue;
            const Grid& g = c.grid;
            G.add(p, grid3(g.n[0], g.n[1], g.n[2], block), g, c.face[0], c.face[1], c.face[2], c.occupancy, c.temperature_tmp, c.density_tmp,
                  c.temperature, c.density, c.solid_temperature, c.dt_device);
        }
        if (!G.empty()) out.push_back(G.finish(A, block));
    }
}

template <int C>
void batch_face_force(const std::vector<Context*>& P, TableArena& A, std::vector<BatchedLaunch>& out) {
    LaunchBuilder<face_force_k<C>, kCellThreads, Grid, float*, const float*, const uint8_t*, const float*, const float*> L;
    for (unsigned p = 0; p < P.size(); ++p) {
        const Context& c = *P[p];
        const Grid& g = c.grid;
        L.add(p, grid3(g.n[0] + (C == 0), g.n[1] + (C == 1), g.n[2] + (C == 2), kCellBlock), g, c.face[C], c.cell_force[C], c.occupancy,
              c.solid_velocity[C], c.dt_device);
    }
    out.push_back(L.finish(A, kCellBlock));
}

template <int C>
void batch_project(const std::vector<Context*>& P, TableArena& A, std::vector<BatchedLaunch>& out) {
    LaunchBuilder<project_k<C>, kCellThreads, Grid, const float*, float*, const float*, const uint8_t*, const float*, const float*> L;
    for (unsigned p = 0; p < P.size(); ++p) {
        const Context& c = *P[p];
        const Grid& g = c.grid;
        L.add(p, grid3(g.n[0] + (C == 0), g.n[1] + (C == 1), g.n[2] + (C == 2), kCellBlock), g, c.face_star[C], c.face[C], c.pressure, c.occupancy,
              c.solid_velocity[C], c.dt_device);
    }
    out.push_back(L.finish(A, kCellBlock));
}

}  // namespace

void build_batched_stages(const std::vector<Context*>& P, TableArena& A, BatchedStages& out) {
    out = BatchedStages{};
    // time step: the running max |u_face| of the adaptive plumes, then dt
    {
        LaunchBuilder<max_speed_k, kBlock, const float*, int64_t, const float*, int64_t, const float*, int64_t, unsigned*> M;
        LaunchBuilder<time_step_k, 64, unsigned*, float*, float*, int, float, float, float, float> T;
        for (unsigned p = 0; p < P.size(); ++p) {
            const Context& c = *P[p];
            const Grid& g = c.grid;
            if (c.options.adaptive_dt) {
                const int64_t total = g.faces[0] + g.faces[1] + g.faces[2];
                const unsigned blocks = unsigned(std::min<int64_t>(2048, (total + kBlock - 1) / kBlock));
                M.add(p, dim3(blocks), c.face[0], g.faces[0], c.face[1], g.faces[1], c.face[2], g.faces[2], c.max_speed_bits);
            }
            T.add(p, dim3(1), c.max_speed_bits, c.max_speed, c.dt_device, c.options.adaptive_dt ? 1 : 0, c.options.cfl, g.h, c.options.dt_max,
                  c.config.dt);
        }
        if (!M.empty()) out.time_step.push_back(M.finish(A, dim3(kBlock)));
        out.time_step.push_back(T.finish(A, dim3(1)));
    }
    // forces
    {
        LaunchBuilder<vorticity_k, kCellThreads, Grid, const float*, const float*, const float*, const uint8_t*, float*, float*, float*, float*, float*,
                      const float*, uint3>
            V;
        LaunchBuilder<cell_force_k, kCellThreads, Grid, const float*, const float*, const float*, const float*, const float*, const uint8_t*,
                      const float*, const float*, const float*, const float*, float, float, float, float, float*, float*, float*, uint3>
            F;
        for (unsigned p = 0; p < P.size(); ++p) {
            Context& c = *P[p];
            const Grid& g = c.grid;
            const uint3 tiles = cell_tiles(g);
            V.add(p, dim3(cell_tile_blocks(g)), g, c.face[0], c.face[1], c.face[2], c.occupancy, c.vorticity[0], c.vorticity[1], c.vorticity[2],
                  c.vorticity_magnitude, c.temperature, c.solid_temperature, tiles);
            F.add(p, dim3(cell_tile_blocks(g)), g, c.force_user[0], c.force_user[1], c.force_user[2], c.density, c.temperature, c.occupancy,
                  c.vorticity[0], c.vorticity[1], c.vorticity[2], c.vorticity_magnitude, c.config.buoyancy_density_factor,
                  c.config.buoyancy_temperature_factor, c.config.ambient_temperature, c.config.vorticity_confinement, c.cell_force[0], c.cell_force[1],
                  c.cell_force[2], tiles);
        }
        out.forces.push_back(V.finish(A, kCellBlock));
        out.forces.push_back(F.finish(A, kCellBlock));
        batch_face_force<0>(P, A, out.forces);
        batch_face_force<1>(P, A, out.forces);
        batch_face_force<2>(P, A, out.forces);
        batch_all_boundaries(P, &Context::face, A, out.forces);
    }
    // velocity advection
    batch_advect_face<0>(P, A, out.advect);
    batch_advect_face<1>(P, A, out.advect);
    batch_advect_face<2>(P, A, out.advect);
    batch_all_boundaries(P, &Context::face_star, A, out.advect);
    // projection
    batch_project<0>(P, A, out.project);
    batch_project<1>(P, A, out.project);
    batch_project<2>(P, A, out.project);
    batch_all_boundaries(P, &Context::face, A, out.project);
    // scalars
    {
        LaunchBuilder<scalar_sources_k, kBlock, const float*, const float*, const float*, const float*, float*, float*, const float*, int64_t> S;
        for (unsigned p = 0; p < P.size(); ++p) {
            const Context& c = *P[p];
            S.add(p, dim3(grid1(c.grid.cells)), c.temperature, c.temperature_source, c.density, c.density_source, c.temperature_tmp, c.density_tmp,
                  c.dt_device, c.grid.cells);
        }
        out.scalars.push_back(S.finish(A, dim3(kBlock)));
        batch_advect_scalars(P, A, out.scalars);
        LaunchBuilder<density_fill_k, kBlock, float*, const int64_t*, const unsigned*> D;
        for (unsigned p = 0; p < P.size(); ++p) {
            const Context& c = *P[p];
            D.add(p, dim3(256), c.density, c.solid_pairs, c.solid_count);
        }
        out.scalars.push_back(D.finish(A, dim3(kBlock)));
    }
}

void build_batched_sources(const std::vector<Context*>& P, const SmokeSimulationPlumeSource* src, uint32_t step, TableArena& A,
                           std::vector<BatchedLaunch>& out) {
    LaunchBuilder<plume_k, kCellThreads, Grid, SmokeSimulationPlumeSource, uint32_t, float*, float*> L;
    for (unsigned p = 0; p < P.size(); ++p) {
        const Context& c = *P[p];
        L.add(p, cell_grid(c.grid), c.grid, src[p], step, c.density_source, c.temperature_source);
    }
    out.push_back(L.finish(A, kCellBlock));
}

void table_advect_blocks(Context& c) {
    for (int k = 0; k < 4; ++k) c.advect_block[k] = kAdvectTable[k];
}

void describe_advect_block(int index, SmokeSimulationLaunchShape& out) {
    if (index < 0 || index >= kAdvectCandidates) throw HipError("invalid advection launch");
    if (index == kAdvectStaged) {
        // the staged form's block: kSaX x kSaY x kSaZStep threads, each with
        // kSaZ / kSaZStep outputs of the kSaX x kSaY x kSaZ tile
        out = {kSaX, kSaY, kSaZStep, kSaZ / kSaZStep};
        return;
    }
    const dim3 b = kAdvectBlocks[index];
    out = {int32_t(b.x), int32_t(b.y), int32_t(b.z), 1};
}

int find_advect_block(const SmokeSimulationLaunchShape& shape) {
    for (int b = 0; b < kAdvectCandidates; ++b) {
        SmokeSimulationLaunchShape d;
        describe_advect_block(b, d);
        if (d.threads_x == shape.threads_x && d.threads_y == shape.threads_y && d.threads_z == shape.threads_z &&
            d.cells_per_thread == shape.cells_per_thread)
            return b;
    }
    throw HipError("advection launch shape is not a candidate");
}

// Times every candidate block of the four advection kernels on the current
// state and keeps the fastest per kernel.  Velocity advection writes the
// intermediate u* (recomputed at the next step); the scalar kernel writes into
// the solver scratch vectors instead of the state.
int tune_advect_blocks(Context& c, SmokeSimulationLaunchReport& report) {
    hipEvent_t t0, t1;
    hip_check(hipEventCreate(&t0), "event");
    hip_check(hipEventCreate(&t1), "event");
    constexpr int kReps = 3;
    int timed = 0;
    for (int k = 0; k < 4; ++k) {
        float best_us = 0.0f;
        int best = c.advect_block[k];
        for (int b = 0; b < kAdvectCandidates; ++b) {
            auto launch = [&] {
                if (k == 0) launch_advect<0>(c, b);
                else if (k == 1) launch_advect<1>(c, b);
                else if (k == 2) launch_advect<2>(c, b);
                else launch_advect_scalars(c, b, c.pcg_r, c.pcg_p);
            };
            launch();   // warm-up
            hip_check(hipEventRecord(t0, c.stream), "event");
            for (int r = 0; r < kReps; ++r) launch();
            hip_check(hipEventRecord(t1, c.stream), "event");
            hip_check(hipEventSynchronize(t1), "event sync");
            hip_check(hipGetLastError(), "advection tuning");
            float ms = 0.0f;
            hip_check(hipEventElapsedTime(&ms, t0, t1), "elapsed");
            const float us = 1e3f * ms / kReps;
            ++timed;
            if (b == c.advect_block[k]) report.advect_us[k][0] = us;
            if (b == 0 || us < best_us) best_us = us, best = b;
        }
        c.advect_block[k] = best;
        report.advect_us[k][1] = best_us;
    }
    hipEventDestroy(t0);
    hipEventDestroy(t1);
    return timed;
}

}  // namespace smoke
