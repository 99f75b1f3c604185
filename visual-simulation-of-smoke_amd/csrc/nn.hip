// 3x3x3 convolutions over channels-last bfloat16 volumes on the gfx950 matrix
// cores (v_mfma_f32_16x16x32_bf16), FP32 accumulation: the building blocks of
// the learned pressure preconditioner.  Modes, layouts and the entry points'
// contracts are documented in visual-simulation-of-smoke-3d-nn.h.
//
// Every convolution is an implicit GEMM: output voxels are the M dimension,
// output channels N, and K = 27 * cin ordered (tap, channel), so the eight
// consecutive K values a lane feeds the matrix core are eight channels of one
// tap -- a single 16-byte load from the channels-last volume.  A wave owns 16
// consecutive x voxels of one output row and all output channels
// (cout / 16 accumulator tiles).  The data gradient is the same kernel run in
// the dual mode (SAME with flipped taps, DOWN <-> UP) on a second packing of
// the weights; the weight gradient runs the matrix core with K = voxels.
#include "context.hpp"
#include "nn.hpp"

#include <algorithm>
#include <atomic>

namespace smoke {
namespace nn {
namespace {

typedef short short8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float float4v __attribute__((ext_vector_type(4)));

struct ConvArgs {
    int batch;
    int in[3];
    int out[3];
    int cin, cout;
    int mode;
    int kpad;   // 27 * cin rounded up to the matrix-core K step (32)
};

constexpr int kStep = 32;

__host__ __device__ constexpr int padded_k(int c) { return (27 * c + kStep - 1) / kStep * kStep; }

__device__ __forceinline__ float bf16_to_f32(uint16_t h) { return __uint_as_float(uint32_t(h) << 16); }

// round to nearest even (NaN stays a quiet NaN)
__device__ __forceinline__ uint16_t f32_to_bf16(float v) {
    uint32_t u = __float_as_uint(v);
    if ((u & 0x7fffffffu) > 0x7f800000u) return uint16_t((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return uint16_t(u >> 16);
}

// Input coordinate feeding output coordinate o through tap t (one axis).
__device__ __forceinline__ bool tap_source(int mode, int o, int t, int n_in, int& i) {
    if (mode == SMOKE_NN_CONV_SAME) {
        i = o + t - 1;
    } else if (mode == SMOKE_NN_CONV_DOWN) {
        i = 2 * o + t - 1;
    } else {
        const int q = o + 1 - t;
        if (q & 1) return false;
        i = q >> 1;
    }
    return i >= 0 && i < n_in;
}

__device__ __forceinline__ bf16x8 load8(const uint16_t* __restrict__ p) {
    return __builtin_bit_cast(bf16x8, *reinterpret_cast<const short8*>(p));
}

__device__ __forceinline__ bf16x8 zero8() { return __builtin_bit_cast(bf16x8, short8{0, 0, 0, 0, 0, 0, 0, 0}); }

// y (or dx) for one 16-voxel x segment of one output row: per wave
// (SPLIT = 1), or per block with the K steps dealt round-robin over its four
// waves and summed through LDS (SPLIT = 4; small volumes, where one wave per
// tile leaves the chip idle and every wave waits out its chain of gathers).
// The next K step's loads are issued before the current step's matrix-core
// work.  Epilogue: + bias, + add, relu, * [mask > 0].
//
// ngroups > 1 (split-K blocks only): the output channels are dealt over
// ngroups consecutive blocks of each tile, 16 NT channels each, so a tiny
// layer (a few tiles) runs on ngroups times as many CUs; every channel's
// chain of matrix-core instructions is unchanged (same bits).
template <int NT, int SPLIT>
__global__ __launch_bounds__(SPLIT == 1 ? 256 : 64 * SPLIT) void conv_kernel(ConvArgs a, const uint16_t* __restrict__ x, const uint16_t* __restrict__ w,
                                                   const float* __restrict__ bias, const uint16_t* __restrict__ add,
                                                   const uint16_t* __restrict__ mask, int relu, uint16_t* __restrict__ y,
                                                   float* __restrict__ yf, int f32_channels, int ngroups) {
    const int lane = threadIdx.x & 63;
    const int wib = threadIdx.x >> 6;   // wave in block
    const int64_t wave = SPLIT == 1 ? int64_t(blockIdx.x) * 4 + wib : int64_t(blockIdx.x) / ngroups;
    const int cbase = SPLIT == 1 ? 0 : int(blockIdx.x % unsigned(ngroups)) * 16 * NT;   // first output channel
    w += int64_t(cbase) * a.kpad;
    const int xt = (a.out[0] + 15) / 16;
    const int64_t tiles = int64_t(a.batch) * a.out[2] * a.out[1] * xt;
    if (wave >= tiles) return;   // uniform per block when SPLIT > 1
    int64_t t = wave;
    const int xtile = int(t % xt);
    t /= xt;
    const int oy = int(t % a.out[1]);
    t /= a.out[1];
    const int oz = int(t % a.out[2]);
    const int b = int(t / a.out[2]);
    const int m = lane & 15, g = lane >> 4;
    const int ox = xtile * 16 + m;   // the voxel whose A row this lane feeds

    auto load_a = [&](int s) -> bf16x8 {
        const int kb = s * kStep + g * 8;
        const int tap = kb / a.cin;
        const int c0 = kb - tap * a.cin;
        if (tap < 27 && ox < a.out[0]) {
            const int tx = tap % 3, ty = (tap / 3) % 3, tz = tap / 9;
            int ix, iy, iz;
            if (tap_source(a.mode, ox, tx, a.in[0], ix) && tap_source(a.mode, oy, ty, a.in[1], iy) &&
                tap_source(a.mode, oz, tz, a.in[2], iz)) {
                const int64_t idx = (((int64_t(b) * a.in[2] + iz) * a.in[1] + iy) * a.in[0] + ix) * a.cin + c0;
                return load8(x + idx);
            }
        }
        return zero8();
    };

    float4v acc[NT];
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) acc[nt] = float4v{0.0f, 0.0f, 0.0f, 0.0f};
    const int ksteps = a.kpad / kStep;
    const int first = SPLIT == 1 ? 0 : wib;
    if constexpr (SPLIT > 1) {
        // a wave's few K steps (up to 4 per round): every load issued before
        // the round's matrix-core work, which keeps the wave's K order
        constexpr int ROUND = NT >= 3 ? 2 : 4;   // within the 128 VGPRs of a 1024-thread block
        for (int s0 = first; s0 < ksteps; s0 += ROUND * SPLIT) {
            bf16x8 ar[ROUND], br[ROUND][NT];
#pragma unroll
            for (int j = 0; j < ROUND; ++j) {
                const int s = s0 + j * SPLIT;
                ar[j] = zero8();
                if (s < ksteps) {
                    ar[j] = load_a(s);
#pragma unroll
                    for (int nt = 0; nt < NT; ++nt) br[j][nt] = load8(w + int64_t(nt * 16 + m) * a.kpad + s * kStep + g * 8);
                }
            }
#pragma unroll
            for (int j = 0; j < ROUND; ++j) {
                if (s0 + j * SPLIT >= ksteps) break;
#pragma unroll
                for (int nt = 0; nt < NT; ++nt) acc[nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ar[j], br[j][nt], acc[nt], 0, 0, 0);
            }
        }
    } else if (first < ksteps) {
        bf16x8 av = load_a(first);
        bf16x8 bv[NT];
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) bv[nt] = load8(w + int64_t(nt * 16 + m) * a.kpad + first * kStep + g * 8);
        for (int s = first; s < ksteps; s += SPLIT) {
            const int sn = s + SPLIT;
            bf16x8 an = zero8();
            bf16x8 bn[NT];
            if (sn < ksteps) {
                an = load_a(sn);
#pragma unroll
                for (int nt = 0; nt < NT; ++nt) bn[nt] = load8(w + int64_t(nt * 16 + m) * a.kpad + sn * kStep + g * 8);
            }
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) acc[nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv[nt], acc[nt], 0, 0, 0);
            av = an;
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) bv[nt] = bn[nt];
        }
    }
    if constexpr (SPLIT > 1) {
        // waves 1..SPLIT-1 hand their partial sums to wave 0, which adds them
        // in wave order (deterministic)
        __shared__ float4v part[SPLIT - 1][NT][64];
        if (wib > 0) {
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) part[wib - 1][nt][lane] = acc[nt];
        }
        __syncthreads();
        if (wib > 0) return;
        // wave order outermost, not unrolled: NT partials live at a time
#pragma unroll 1
        for (int q = 0; q < SPLIT - 1; ++q) {
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) acc[nt] = acc[nt] + part[q][nt][lane];
        }
    }
    // lane holds D[4 g + i][m]: voxel xtile * 16 + 4 g + i, channel cbase + nt * 16 + m
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
        const int n = cbase + nt * 16 + m;
        const float bn = bias ? bias[n] : 0.0f;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int vx = xtile * 16 + 4 * g + i;
            if (vx >= a.out[0]) continue;
            const int64_t vox = ((int64_t(b) * a.out[2] + oz) * a.out[1] + oy) * a.out[0] + vx;
            const int64_t e = vox * a.cout + n;
            float v = acc[nt][i] + bn;
            if (add) v += bf16_to_f32(add[e]);
            if (relu) v = fmaxf(v, 0.0f);
            if (mask && !(bf16_to_f32(mask[e]) > 0.0f)) v = 0.0f;
            if (y) y[e] = f32_to_bf16(v);
            if (yf && n < f32_channels) yf[vox * f32_channels + n] = v;
        }
    }
}

// ---- LDS-staged convolution (volumes with many tiles) ----
//
// A block of four waves computes a 16 MT-voxel x segment of four
// consecutive output rows (wave w: row oy0 + w) of one z plane.  The input
// patch that these outputs read through any of the 27 taps is copied into
// LDS once, zero outside the volume, so the taps' gathers become LDS reads
// with no bounds tests; each wave keeps MT x tiles times NT channel tiles of
// accumulators, so one weight fragment (read from global memory) feeds MT
// matrix-core instructions and one activation fragment NT.  Per output tile
// the chain of matrix-core instructions is the gather kernel's -- the same
// operands in the same K order from zero -- so both kernels produce the same
// bits.

// Voxel stride of a staged patch in 16-byte units: channels / 8, made odd so
// that eight consecutive voxels fall into eight different groups of banks.
__host__ __device__ constexpr int patch_stride(int c) { return (c / 8) | 1; }

// First input coordinate of the patch of output origin o0 along one axis.
template <int MODE>
__device__ __forceinline__ int patch_lo(int o0) {
    if constexpr (MODE == SMOKE_NN_CONV_SAME) return o0 - 1;
    else if constexpr (MODE == SMOKE_NN_CONV_DOWN) return 2 * o0 - 1;
    else return (o0 - 1) >> 1;   // floor((o0 - 1) / 2)
}

template <int MODE, int MT>
struct PatchShape {
    static constexpr int TX = 16 * MT;   // output voxels along x per block
    static constexpr int EX = MODE == SMOKE_NN_CONV_SAME ? TX + 2 : (MODE == SMOKE_NN_CONV_DOWN ? 2 * TX + 1 : TX / 2 + 2);
    static constexpr int EY = MODE == SMOKE_NN_CONV_SAME ? 6 : (MODE == SMOKE_NN_CONV_DOWN ? 9 : 4);
    static constexpr int EZ = MODE == SMOKE_NN_CONV_UP ? 2 : 3;
    // DOWN reads every other patch voxel along x: the even ones are stored
    // first, so that the 16 lanes of a fragment read consecutive slots
    static constexpr int HALF = (EX + 1) / 2;
    __device__ static constexpr int slot(int px) { return MODE == SMOKE_NN_CONV_DOWN ? (px & 1) * HALF + (px >> 1) : px; }
};

template <int MODE, int CIN, int NT, int MT>
__global__ __launch_bounds__(256) void conv_staged_kernel(ConvArgs a, const uint16_t* __restrict__ x, const uint16_t* __restrict__ w,
                                                          const float* __restrict__ bias, const uint16_t* __restrict__ add,
                                                          const uint16_t* __restrict__ mask, int relu, uint16_t* __restrict__ y,
                                                          float* __restrict__ yf, int f32_channels) {
    using P = PatchShape<MODE, MT>;
    constexpr int VS = patch_stride(CIN);
    constexpr int C8 = CIN / 8;
    constexpr int KPAD = padded_k(CIN);
    constexpr int KSTEPS = KPAD / kStep;
    __shared__ short8 patch[P::EX * P::EY * P::EZ * VS];

    const int xt = (a.out[0] + P::TX - 1) / P::TX;
    const int yt = (a.out[1] + 3) / 4;
    unsigned t = blockIdx.x;
    const int bx = int(t % unsigned(xt));
    t /= unsigned(xt);
    const int by = int(t % unsigned(yt));
    t /= unsigned(yt);
    const int oz = int(t % unsigned(a.out[2]));
    const int b = int(t / unsigned(a.out[2]));
    const int ox0 = bx * P::TX, oy0 = by * 4;
    const int lx = patch_lo<MODE>(ox0), ly = patch_lo<MODE>(oy0), lz = patch_lo<MODE>(oz);

    // the patch: 16-byte loads of 8 channels, consecutive threads along
    // channels then x (coalesced in the channels-last volume)
    // (all of a thread's loads are issued before its stores)
    constexpr int ELEMS = P::EX * P::EY * P::EZ * C8;
    constexpr int ITERS = (ELEMS + 255) / 256;
    short8 val[ITERS];
    int dst[ITERS];
#pragma unroll
    for (int k = 0; k < ITERS; ++k) {
        const int i = int(threadIdx.x) + 256 * k;
        const int c8 = i % C8;
        const int v = i / C8;
        const int px = v % P::EX;
        const int py = (v / P::EX) % P::EY;
        const int pz = v / (P::EX * P::EY);
        const int gx = lx + px, gy = ly + py, gz = lz + pz;
        val[k] = short8{0, 0, 0, 0, 0, 0, 0, 0};
        dst[k] = i < ELEMS ? ((pz * P::EY + py) * P::EX + P::slot(px)) * VS + c8 : -1;
        if (i < ELEMS && gx >= 0 && gx < a.in[0] && gy >= 0 && gy < a.in[1] && gz >= 0 && gz < a.in[2])
            val[k] = *reinterpret_cast<const short8*>(x + ((((int64_t(b) * a.in[2] + gz) * a.in[1] + gy) * a.in[0] + gx) * CIN + c8 * 8));
    }
#pragma unroll
    for (int k = 0; k < ITERS; ++k)
        if (dst[k] >= 0) patch[dst[k]] = val[k];
    __syncthreads();

    const int lane = int(threadIdx.x) & 63, wv = int(threadIdx.x) >> 6;
    const int m = lane & 15, g = lane >> 4;
    const int oy = oy0 + wv;

    // A fragment of K step s for x tile mt: eight channels of one tap of the
    // voxel this lane feeds (s and mt are compile-time after unrolling)
    auto load_a = [&](int s, int mt) -> bf16x8 {
        const int kb = s * kStep + g * 8;
        const int tap = kb / CIN;
        const int c8 = (kb % CIN) / 8;
        if (tap >= 27) return zero8();
        const int tx = tap % 3, ty = (tap / 3) % 3, tz = tap / 9;
        const int mx = 16 * mt + m;
        int px, py, pz;
        if constexpr (MODE == SMOKE_NN_CONV_SAME) {
            px = mx + tx;
            py = wv + ty;
            pz = tz;
        } else if constexpr (MODE == SMOKE_NN_CONV_DOWN) {
            px = 2 * mx + tx;
            py = 2 * wv + ty;
            pz = tz;
        } else {
            // ox0 and oy0 are even: the parity of o + 1 - t is that of the
            // in-block offset's
            const int qx = mx + 1 - tx, qy = wv + 1 - ty, qz = oz + 1 - tz;
            if ((qx | qy | qz) & 1) return zero8();
            px = (qx >> 1) + 1;
            py = (qy >> 1) + 1;
            pz = (qz >> 1) - lz;
        }
        return __builtin_bit_cast(bf16x8, patch[((pz * P::EY + py) * P::EX + P::slot(px)) * VS + c8]);
    };
    auto load_b = [&](int s, int nt) -> bf16x8 { return load8(w + int64_t(nt * 16 + m) * KPAD + s * kStep + g * 8); };

    float4v acc[MT][NT];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = float4v{0.0f, 0.0f, 0.0f, 0.0f};
    bf16x8 av[MT], bv[NT];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) av[mt] = load_a(0, mt);
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) bv[nt] = load_b(0, nt);
#pragma unroll
    for (int s = 0; s < KSTEPS; ++s) {
        bf16x8 an[MT], bn[NT];
        if (s + 1 < KSTEPS) {
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) bn[nt] = load_b(s + 1, nt);
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) an[mt] = load_a(s + 1, mt);
        }
#pragma unroll
        for (int mt = 0; mt < MT; ++mt)
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[mt], bv[nt], acc[mt][nt], 0, 0, 0);
        if (s + 1 < KSTEPS) {
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) av[mt] = an[mt];
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) bv[nt] = bn[nt];
        }
    }
    if (oy >= a.out[1]) return;
    // lane holds D[4 g + i][m]: voxel ox0 + 16 mt + 4 g + i, channel nt * 16 + m
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
        const int n = nt * 16 + m;
        const float bn = bias ? bias[n] : 0.0f;
#pragma unroll
        for (int mt = 0; mt < MT; ++mt)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int vx = ox0 + 16 * mt + 4 * g + i;
                if (vx >= a.out[0]) continue;
                const int64_t vox = ((int64_t(b) * a.out[2] + oz) * a.out[1] + oy) * a.out[0] + vx;
                const int64_t e = vox * (NT * 16) + n;
                float v = acc[mt][nt][i] + bn;
                if (add) v += bf16_to_f32(add[e]);
                if (relu) v = fmaxf(v, 0.0f);
                if (mask && !(bf16_to_f32(mask[e]) > 0.0f)) v = 0.0f;
                if (y) y[e] = f32_to_bf16(v);
                if (yf && n < f32_channels) yf[vox * f32_channels + n] = v;
            }
    }
}

// ---- LDS-staged weight gradient ----
//
// dW[t][c][n] = sum over output voxels o of x[i(o, t)][c] dy[o][n] on the
// matrix cores with M = input channels, N = output channels and K = 32
// output voxels.  A lane's eight K values are eight voxels of one channel,
// which the channels-last volumes hold 2 * channels bytes apart; a block
// therefore stages a region of RY output rows x XB voxels of one z plane
// transposed into LDS: dy as [n][row][x], and x as [tx][c][input row][x]
// holding, for every output voxel and each of the three x taps, the input
// value that tap reads (zero where it reads outside the volume or, for UP,
// nothing).  Every fragment is then one aligned 16-byte LDS read.  A wave
// owns TW taps times all channel tiles (its accumulators) and reuses each dy
// fragment across its taps; blocks loop over regions, accumulating, and add
// their sums to dW with FP32 atomics at the end.
template <int MODE, int RY>
struct RowShape {
    static constexpr int EY = MODE == SMOKE_NN_CONV_SAME ? RY + 2 : (MODE == SMOKE_NN_CONV_DOWN ? 2 * RY + 1 : RY / 2 + 2);
    static constexpr int EZ = MODE == SMOKE_NN_CONV_UP ? 2 : 3;
    static constexpr int IR = EY * EZ;   // input rows a region reads
};

// odd number of 16-byte units (bank groups), at least u
__host__ __device__ constexpr int odd_units(int u) { return u | 1; }

// the input coordinate of output coordinate o through tap t along x
template <int MODE>
__device__ __forceinline__ bool x_source(int o, int t, int n_in, int& i) {
    if constexpr (MODE == SMOKE_NN_CONV_SAME) i = o + t - 1;
    else if constexpr (MODE == SMOKE_NN_CONV_DOWN) i = 2 * o + t - 1;
    else {
        const int q = o + 1 - t;
        if (q & 1) return false;
        i = q >> 1;
    }
    return i >= 0 && i < n_in;
}

// 8 voxels x 8 channels of bfloat16 -> 8 channels x 8 voxels
__device__ __forceinline__ void transpose8(const short8 (&v)[8], short8 (&u)[8]) {
#pragma unroll
    for (int c = 0; c < 8; ++c)
#pragma unroll
        for (int j = 0; j < 8; ++j) u[c][j] = v[j][c];
}

template <int MODE, int CIN, int COUT, int RY, int XB, int TW>
__global__ __launch_bounds__(256) void wgrad_staged_kernel(ConvArgs a, const uint16_t* __restrict__ x, const uint16_t* __restrict__ dy,
                                                           float* __restrict__ dw, int regions) {
    using R = RowShape<MODE, RY>;
    constexpr int CT = (CIN + 15) / 16, NT = COUT / 16;
    constexpr int XU = XB / 8;                      // 16-byte units per staged row
    constexpr int CS = odd_units(R::IR * XU);       // channel stride of the x stage
    constexpr int NS = odd_units(RY * XU);          // channel stride of the dy stage
    constexpr int KS = RY * XB / kStep;             // K steps per region
    constexpr int PASSES = (27 + 4 * TW - 1) / (4 * TW);
    static_assert(XB % 8 == 0 && (RY * XB) % kStep == 0, "region must hold whole K steps");
    __shared__ short8 xs[3 * CIN * CS];
    __shared__ short8 ds[COUT * NS];

    const int lane = int(threadIdx.x) & 63, wv = int(threadIdx.x) >> 6;
    const int m = lane & 15, g = lane >> 4;
    const int xr = (a.out[0] + XB - 1) / XB;
    const int yr = (a.out[1] + RY - 1) / RY;

    for (int pass = 0; pass < PASSES; ++pass) {
        const int tap0 = pass * 4 * TW + wv * TW;
        float4v acc[TW][CT][NT];
#pragma unroll
        for (int j = 0; j < TW; ++j)
#pragma unroll
            for (int ct = 0; ct < CT; ++ct)
#pragma unroll
                for (int nt = 0; nt < NT; ++nt) acc[j][ct][nt] = float4v{0.0f, 0.0f, 0.0f, 0.0f};

        for (int reg = int(blockIdx.x); reg < regions; reg += int(gridDim.x)) {
            int q = reg;
            const int rx = q % xr;
            q /= xr;
            const int ry = q % yr;
            q /= yr;
            const int oz = q % a.out[2];
            const int b = q / a.out[2];
            const int x0 = rx * XB, oy0 = ry * RY;
            const int ly = patch_lo<MODE>(oy0), lz = patch_lo<MODE>(oz);

            __syncthreads();   // the previous region's reads are done
            // x: unit = (tx, channel octet, input row, voxel octet)
            constexpr int XUNITS = 3 * (CIN / 8) * R::IR * XU;
#pragma unroll
            for (int k = 0; k < (XUNITS + 255) / 256; ++k) {
                const int u = int(threadIdx.x) + 256 * k;
                if (u >= XUNITS) break;
                int r = u;
                const int xo8 = r % XU;
                r /= XU;
                const int ir = r % R::IR;
                r /= R::IR;
                const int c8 = r % (CIN / 8);
                const int tx = r / (CIN / 8);
                const int iy = ly + ir % R::EY, iz = lz + ir / R::EY;
                const bool row_ok = iy >= 0 && iy < a.in[1] && iz >= 0 && iz < a.in[2];
                const int64_t row = (int64_t(b) * a.in[2] + iz) * a.in[1] + iy;
                short8 v[8], t8[8];
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    int ix;
                    v[j] = short8{0, 0, 0, 0, 0, 0, 0, 0};
                    if (row_ok && x_source<MODE>(x0 + xo8 * 8 + j, tx, a.in[0], ix))
                        v[j] = *reinterpret_cast<const short8*>(x + (row * a.in[0] + ix) * CIN + c8 * 8);
                }
                transpose8(v, t8);
#pragma unroll
                for (int c = 0; c < 8; ++c) xs[(tx * CIN + c8 * 8 + c) * CS + ir * XU + xo8] = t8[c];
            }
            // dy: unit = (channel octet, row, voxel octet); zero outside the volume
            constexpr int DUNITS = (COUT / 8) * RY * XU;
#pragma unroll
            for (int k = 0; k < (DUNITS + 255) / 256; ++k) {
                const int u = int(threadIdx.x) + 256 * k;
                if (u >= DUNITS) break;
                const int xo8 = u % XU;
                const int r = (u / XU) % RY;
                const int n8 = u / (XU * RY);
                const int oy = oy0 + r;
                short8 v[8], t8[8];
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const int ox = x0 + xo8 * 8 + j;
                    v[j] = short8{0, 0, 0, 0, 0, 0, 0, 0};
                    if (oy < a.out[1] && ox < a.out[0])
                        v[j] = *reinterpret_cast<const short8*>(dy + (((int64_t(b) * a.out[2] + oz) * a.out[1] + oy) * a.out[0] + ox) * COUT + n8 * 8);
                }
                transpose8(v, t8);
#pragma unroll
                for (int c = 0; c < 8; ++c) ds[(n8 * 8 + c) * NS + r * XU + xo8] = t8[c];
            }
            __syncthreads();

            for (int s = 0; s < KS; ++s) {
                // this lane's eight K values: voxels kv .. kv + 7 of the region
                const int kv = s * kStep + g * 8;
                const int r = kv / XB, xo8 = (kv % XB) / 8;
                bf16x8 bv[NT];
#pragma unroll
                for (int nt = 0; nt < NT; ++nt) bv[nt] = __builtin_bit_cast(bf16x8, ds[(nt * 16 + m) * NS + r * XU + xo8]);
#pragma unroll
                for (int j = 0; j < TW; ++j) {
                    const int tap = tap0 + j;
                    if (tap >= 27) break;
                    const int tx = tap % 3, ty = (tap / 3) % 3, tz = tap / 9;
                    int ir;
                    bool ok = true;
                    if constexpr (MODE == SMOKE_NN_CONV_SAME) ir = tz * R::EY + r + ty;
                    else if constexpr (MODE == SMOKE_NN_CONV_DOWN) ir = tz * R::EY + 2 * r + ty;
                    else {
                        // oy0 is even: the parity of oy + 1 - ty is r + 1 - ty's
                        const int qy = r + 1 - ty, qz = oz + 1 - tz;
                        ok = ((qy | qz) & 1) == 0;
                        ir = ((qz >> 1) - lz) * R::EY + (qy >> 1) + 1;
                    }
                    bf16x8 av[CT];
#pragma unroll
                    for (int ct = 0; ct < CT; ++ct) {
                        const int c = ct * 16 + m;
                        av[ct] = (ok && c < CIN) ? __builtin_bit_cast(bf16x8, xs[(tx * CIN + c) * CS + ir * XU + xo8]) : zero8();
                    }
#pragma unroll
                    for (int ct = 0; ct < CT; ++ct)
#pragma unroll
                        for (int nt = 0; nt < NT; ++nt) acc[j][ct][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[ct], bv[nt], acc[j][ct][nt], 0, 0, 0);
                }
            }
        }
        // lane holds D[4 g + i][m]: input channel ct * 16 + 4 g + i, output channel nt * 16 + m
#pragma unroll
        for (int j = 0; j < TW; ++j) {
            const int tap = tap0 + j;
            if (tap >= 27) break;
#pragma unroll
            for (int ct = 0; ct < CT; ++ct)
#pragma unroll
                for (int nt = 0; nt < NT; ++nt)
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        const int ci = ct * 16 + 4 * g + i;
                        const float v = acc[j][ct][nt][i];
                        if (ci < CIN && v != 0.0f) atomicAdd(&dw[(int64_t(tap) * CIN + ci) * COUT + nt * 16 + m], v);
                    }
        }
    }
}

// dW[t][c][n] += sum over output voxels o of x[i(o, t)][c] * dy[o][n].  One
// wave: tap t, a 16-channel tile of c, a 16-channel tile of n and `rows`
// consecutive output rows; K runs over the voxels of a row, 32 at a time.
__global__ __launch_bounds__(256) void wgrad_kernel(ConvArgs a, const uint16_t* __restrict__ x, const uint16_t* __restrict__ dy,
                                                    float* __restrict__ dw, int rows) {
    const int lane = threadIdx.x & 63;
    const int64_t wave = int64_t(blockIdx.x) * 4 + (threadIdx.x >> 6);
    const int ct_n = (a.cin + 15) / 16, nt_n = a.cout / 16;
    const int64_t total_rows = int64_t(a.batch) * a.out[2] * a.out[1];
    const int64_t groups = (total_rows + rows - 1) / rows;
    const int64_t waves = groups * 27 * ct_n * nt_n;
    if (wave >= waves) return;
    int64_t q = wave;
    const int nt = int(q % nt_n);
    q /= nt_n;
    const int ct = int(q % ct_n);
    q /= ct_n;
    const int tap = int(q % 27);
    const int64_t group = q / 27;
    const int tx = tap % 3, ty = (tap / 3) % 3, tz = tap / 9;
    const int m = lane & 15, g = lane >> 4;
    const int c = ct * 16 + m;   // this lane's A row (input channel)
    const int n = nt * 16 + m;   // this lane's B column (output channel)
    float4v acc = float4v{0.0f, 0.0f, 0.0f, 0.0f};
    const int64_t r0 = group * rows, r1 = min(total_rows, r0 + rows);
    for (int64_t r = r0; r < r1; ++r) {
        const int oy = int(r % a.out[1]);
        const int oz = int((r / a.out[1]) % a.out[2]);
        const int b = int(r / (int64_t(a.out[1]) * a.out[2]));
        int iy, iz;
        const bool row_ok = bool(int(tap_source(a.mode, oy, ty, a.in[1], iy)) & int(tap_source(a.mode, oz, tz, a.in[2], iz)));   // both taps resolved
        const int64_t in_row = (int64_t(b) * a.in[2] + iz) * a.in[1] + iy;
        const int64_t out_row = r;   // (b * D + z) * H + y
        for (int x0 = 0; x0 < a.out[0]; x0 += kStep) {
            short8 av = short8{0, 0, 0, 0, 0, 0, 0, 0}, bv = short8{0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int ox = x0 + g * 8 + i;
                if (ox >= a.out[0]) continue;
                bv[i] = short(dy[(out_row * a.out[0] + ox) * a.cout + n]);
                int ix;
                if (row_ok && c < a.cin && tap_source(a.mode, ox, tx, a.in[0], ix)) av[i] = short(x[(in_row * a.in[0] + ix) * a.cin + c]);
            }
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, av), __builtin_bit_cast(bf16x8, bv), acc, 0, 0, 0);
        }
    }
    // lane holds D[4 g + i][m]: input channel ct * 16 + 4 g + i, output channel n
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int ci = ct * 16 + 4 * g + i;
        if (ci < a.cin && acc[i] != 0.0f) atomicAdd(&dw[(int64_t(tap) * a.cin + ci) * a.cout + n], acc[i]);
    }
}

// db[n] += sum over voxels of dy[.][n].  A thread reads 8 channels of a voxel
// per 16-byte load (cout / 8 threads per voxel), the grid strides over the
// volume, and each block folds its threads' sums per channel through LDS and
// adds one value per channel: kBiasBlocks atomics per address.
constexpr int kBiasBlocks = 512;

__global__ __launch_bounds__(256) void bias_grad_kernel(const uint16_t* __restrict__ dy, int64_t voxels, int cout, float* __restrict__ db) {
    __shared__ float part[256][8];
    const int groups = cout / 8;                 // threads per voxel
    const int per_block = 256 / groups;          // voxels per block and pass
    const int n8 = int(threadIdx.x) % groups;
    const int slot = int(threadIdx.x) / groups;
    float s[8] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    if (slot < per_block) {
        for (int64_t v = int64_t(blockIdx.x) * per_block + slot; v < voxels; v += int64_t(gridDim.x) * per_block) {
            const short8 h = *reinterpret_cast<const short8*>(dy + v * cout + n8 * 8);
#pragma unroll
            for (int j = 0; j < 8; ++j) s[j] += bf16_to_f32(uint16_t(h[j]));
        }
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) part[threadIdx.x][j] = s[j];
    __syncthreads();
    if (int(threadIdx.x) < cout) {
        const int g = int(threadIdx.x) / 8, j = int(threadIdx.x) % 8;
        float t = 0.0f;
        for (int k = 0; k < per_block; ++k) t += part[k * groups + g][j];
        atomicAdd(&db[threadIdx.x], t);
    }
}

// forward packing [cout][kpad] with k = t * cin + c; backward packing
// [cin][kpad'] with k' = t' * cout + n, t' = 26 - t for SAME (flipped taps).
__global__ void pack_kernel(ConvArgs a, const float* __restrict__ w, uint16_t* __restrict__ fwd, uint16_t* __restrict__ bwd) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int kf = a.kpad, kb = padded_k(a.cout);
    if (fwd && i < int64_t(a.cout) * kf) {
        const int n = int(i / kf), k = int(i % kf);
        const int t = k / a.cin, c = k % a.cin;
        fwd[i] = t < 27 ? f32_to_bf16(w[(int64_t(t) * a.cin + c) * a.cout + n]) : uint16_t(0);
    }
    if (bwd && i < int64_t(a.cin) * kb) {
        const int c = int(i / kb), k = int(i % kb);
        const int tp = k / a.cout, n = k % a.cout;
        uint16_t v = 0;
        if (tp < 27) {
            const int t = a.mode == SMOKE_NN_CONV_SAME ? 26 - tp : tp;
            v = f32_to_bf16(w[(int64_t(t) * a.cin + c) * a.cout + n]);
        }
        bwd[i] = v;
    }
}

int expected_out(int mode, int in) { return mode == SMOKE_NN_CONV_SAME ? in : (mode == SMOKE_NN_CONV_DOWN ? (in + 1) / 2 : 2 * in); }

ConvArgs make(const SmokeNnConvDesc& d) {
    if (d.batch <= 0 || d.cin <= 0 || d.cout <= 0 || d.cin % 8 != 0 || d.cout % 16 != 0 || d.cout > 64 || d.cin > 64)
        throw HipError("conv: channels must be cin % 8 == 0, cout % 16 == 0, at most 64");
    if (d.mode < 0 || d.mode > 2) throw HipError("conv: bad mode");
    ConvArgs a{};
    a.batch = d.batch;
    for (int k = 0; k < 3; ++k) {
        a.in[k] = d.in[k];
        a.out[k] = d.out[k];
        if (d.in[k] <= 0 || d.out[k] <= 0) throw HipError("conv: empty volume");
        // UP may produce an odd fine size that DOWN maps back to this coarse size
        const bool ok = d.mode == SMOKE_NN_CONV_UP ? (d.out[k] == 2 * d.in[k] || d.out[k] == 2 * d.in[k] - 1) : d.out[k] == expected_out(d.mode, d.in[k]);
        if (!ok) throw HipError("conv: output size does not match the mode");
    }
    a.cin = d.cin;
    a.cout = d.cout;
    a.mode = d.mode;
    a.kpad = padded_k(d.cin);
    return a;
}

// the dual convolution computing the data gradient
ConvArgs dual(const ConvArgs& a) {
    ConvArgs d = a;
    for (int k = 0; k < 3; ++k) {
        d.in[k] = a.out[k];
        d.out[k] = a.in[k];
    }
    d.cin = a.cout;
    d.cout = a.cin;
    d.mode = a.mode == SMOKE_NN_CONV_SAME ? SMOKE_NN_CONV_SAME : (a.mode == SMOKE_NN_CONV_DOWN ? SMOKE_NN_CONV_UP : SMOKE_NN_CONV_DOWN);
    d.kpad = padded_k(a.cout);
    return d;
}

// Volumes with fewer tiles than kSplitTiles deal each tile's K steps over
// the waves of a block: 4 waves, 8 below kSplit8Tiles, 16 below
// kSplit16Tiles.  The coarse levels of the network (4^3 to 16^3: 16 to 256
// tiles) would otherwise leave most CUs idle while each wave works through
// its chain of dependent gathers.
constexpr int64_t kSplitTiles = 8192;
constexpr int64_t kSplit8Tiles = 512;
constexpr int64_t kSplit16Tiles = 128;

template <int SPLIT>
void launch_conv_split(const ConvArgs& a, unsigned blocks, const uint16_t* x, const uint16_t* w, const float* bias, const uint16_t* add,
                       const uint16_t* mask, int relu, uint16_t* y, float* yf, int f32_channels, hipStream_t s) {
    constexpr unsigned threads = SPLIT == 1 ? 256 : 64 * SPLIT;
    if (a.cout % 16 != 0 || a.cout < 16 || a.cout > 64) throw HipError("conv: cout must be 16..64");
    if constexpr (SPLIT == 16) {
        // the fewest tiles: one 16-channel group per block
        const int groups = a.cout / 16;
        conv_kernel<1, SPLIT><<<blocks * unsigned(groups), threads, 0, s>>>(a, x, w, bias, add, mask, relu, y, yf, f32_channels, groups);
        return;
    }
    switch (a.cout / 16) {
        case 1: conv_kernel<1, SPLIT><<<blocks, threads, 0, s>>>(a, x, w, bias, add, mask, relu, y, yf, f32_channels, 1); break;
        case 2: conv_kernel<2, SPLIT><<<blocks, threads, 0, s>>>(a, x, w, bias, add, mask, relu, y, yf, f32_channels, 1); break;
        case 3: conv_kernel<3, SPLIT><<<blocks, threads, 0, s>>>(a, x, w, bias, add, mask, relu, y, yf, f32_channels, 1); break;
        default: conv_kernel<4, SPLIT><<<blocks, threads, 0, s>>>(a, x, w, bias, add, mask, relu, y, yf, f32_channels, 1); break;
    }
}

// Which kernels run (smoke_nn_set_conv_kernels): automatic, the gather
// kernels only, or the staged kernels wherever their shape table applies.
std::atomic<int> g_conv_kernels{SMOKE_NN_CONV_KERNELS_AUTO};

template <int MODE, int CIN, int NT, int MT>
void launch_staged(const ConvArgs& a, const uint16_t* x, const uint16_t* w, const float* bias, const uint16_t* add, const uint16_t* mask, int relu,
                   uint16_t* y, float* yf, int f32_channels, hipStream_t s) {
    using P = PatchShape<MODE, MT>;
    const int64_t blocks = int64_t(a.batch) * a.out[2] * ((a.out[1] + 3) / 4) * ((a.out[0] + P::TX - 1) / P::TX);
    if (blocks > int64_t(0x7fffffff)) throw HipError("conv: volume too large");
    conv_staged_kernel<MODE, CIN, NT, MT><<<unsigned(blocks), 256, 0, s>>>(a, x, w, bias, add, mask, relu, y, yf, f32_channels);
}

// x tiles per block: the table's MT, fewer for narrow volumes
template <int MODE, int CIN, int NT, int MTMAX>
void staged_mt(const ConvArgs& a, const uint16_t* x, const uint16_t* w, const float* bias, const uint16_t* add, const uint16_t* mask, int relu,
               uint16_t* y, float* yf, int f32_channels, hipStream_t s) {
    const int xtiles = (a.out[0] + 15) / 16;
    if constexpr (MTMAX >= 4) {
        if (xtiles >= 4) return launch_staged<MODE, CIN, NT, 4>(a, x, w, bias, add, mask, relu, y, yf, f32_channels, s);
    }
    if constexpr (MTMAX >= 2) {
        if (xtiles >= 2) return launch_staged<MODE, CIN, NT, 2>(a, x, w, bias, add, mask, relu, y, yf, f32_channels, s);
    }
    launch_staged<MODE, CIN, NT, 1>(a, x, w, bias, add, mask, relu, y, yf, f32_channels, s);
}

// The staged kernel's shapes: every (mode, cin, cout) of the network's
// layers and of their data gradients, with the largest MT whose patch fits
// about 64 KB of LDS.  Returns false for any other shape.
bool launch_conv_staged(const ConvArgs& a, const uint16_t* x, const uint16_t* w, const float* bias, const uint16_t* add, const uint16_t* mask,
                        int relu, uint16_t* y, float* yf, int f32_channels, hipStream_t s) {
    const int nt = a.cout / 16;
#define SMOKE_STAGED(MODE, CIN, NT, MT)                                                              \
    if (a.mode == MODE && a.cin == CIN && nt == NT) {                                                \
        staged_mt<MODE, CIN, NT, MT>(a, x, w, bias, add, mask, relu, y, yf, f32_channels, s);        \
        return true;                                                                                 \
    }
    SMOKE_STAGED(SMOKE_NN_CONV_SAME, 8, 1, 4)
    SMOKE_STAGED(SMOKE_NN_CONV_SAME, 16, 1, 4)
    SMOKE_STAGED(SMOKE_NN_CONV_SAME, 32, 2, 2)
    SMOKE_STAGED(SMOKE_NN_CONV_SAME, 64, 4, 1)
    SMOKE_STAGED(SMOKE_NN_CONV_DOWN, 16, 2, 1)
    SMOKE_STAGED(SMOKE_NN_CONV_DOWN, 32, 4, 1)
    SMOKE_STAGED(SMOKE_NN_CONV_UP, 32, 1, 4)
    SMOKE_STAGED(SMOKE_NN_CONV_UP, 64, 2, 4)
#undef SMOKE_STAGED
    return false;
}

void launch_conv(const ConvArgs& a, const uint16_t* x, const uint16_t* w, const float* bias, const uint16_t* add, const uint16_t* mask, int relu,
                 uint16_t* y, float* yf, int f32_channels, hipStream_t s) {
    const int64_t tiles = int64_t(a.batch) * a.out[2] * a.out[1] * ((a.out[0] + 15) / 16);
    const int choice = g_conv_kernels.load(std::memory_order_relaxed);
    if (choice != SMOKE_NN_CONV_KERNELS_GATHER && (choice == SMOKE_NN_CONV_KERNELS_STAGED || tiles >= kSplitTiles) &&
        launch_conv_staged(a, x, w, bias, add, mask, relu, y, yf, f32_channels, s)) {
        hip_check(hipGetLastError(), "conv_staged_kernel");
        return;
    }
    if (tiles < kSplit16Tiles) launch_conv_split<16>(a, unsigned(tiles), x, w, bias, add, mask, relu, y, yf, f32_channels, s);
    else if (tiles < kSplit8Tiles) launch_conv_split<8>(a, unsigned(tiles), x, w, bias, add, mask, relu, y, yf, f32_channels, s);
    else if (tiles < kSplitTiles) launch_conv_split<4>(a, unsigned(tiles), x, w, bias, add, mask, relu, y, yf, f32_channels, s);
    else launch_conv_split<1>(a, unsigned((tiles + 3) / 4), x, w, bias, add, mask, relu, y, yf, f32_channels, s);
    hip_check(hipGetLastError(), "conv_kernel");
}

// The staged weight gradient's shapes: (mode, cin, cout) of every layer, with
// the region (RY rows x XB voxels) and taps per wave (TW) that keep the
// stage within about 80 KB of LDS and the accumulators within 28 tiles.
// Returns false for any other shape.
constexpr int64_t kWgradStagedVoxels = 65536;
constexpr int kWgradBlocks = 1024;

template <int MODE, int CIN, int COUT, int RY, int XB, int TW>
void launch_wgrad(const ConvArgs& a, const uint16_t* x, const uint16_t* dy, float* dw, hipStream_t s) {
    const int64_t regions = int64_t(a.batch) * a.out[2] * ((a.out[1] + RY - 1) / RY) * ((a.out[0] + XB - 1) / XB);
    if (regions > int64_t(0x7fffffff)) throw HipError("conv: volume too large");
    const int blocks = int(std::min<int64_t>(regions, kWgradBlocks));
    wgrad_staged_kernel<MODE, CIN, COUT, RY, XB, TW><<<unsigned(blocks), 256, 0, s>>>(a, x, dy, dw, int(regions));
}

bool launch_wgrad_staged(const ConvArgs& a, const uint16_t* x, const uint16_t* dy, float* dw, hipStream_t s) {
#define SMOKE_WGRAD(MODE, CIN, COUT, RY, XB, TW)                          \
    if (a.mode == MODE && a.cin == CIN && a.cout == COUT) {               \
        launch_wgrad<MODE, CIN, COUT, RY, XB, TW>(a, x, dy, dw, s);       \
        return true;                                                      \
    }
    SMOKE_WGRAD(SMOKE_NN_CONV_SAME, 8, 16, 4, 32, 7)
    SMOKE_WGRAD(SMOKE_NN_CONV_SAME, 16, 16, 4, 32, 7)
    SMOKE_WGRAD(SMOKE_NN_CONV_SAME, 32, 32, 2, 16, 7)
    SMOKE_WGRAD(SMOKE_NN_CONV_SAME, 64, 64, 2, 16, 1)
    SMOKE_WGRAD(SMOKE_NN_CONV_DOWN, 16, 32, 2, 32, 7)
    SMOKE_WGRAD(SMOKE_NN_CONV_DOWN, 32, 64, 2, 16, 3)
    SMOKE_WGRAD(SMOKE_NN_CONV_UP, 64, 32, 2, 16, 3)
    SMOKE_WGRAD(SMOKE_NN_CONV_UP, 32, 16, 4, 32, 7)
#undef SMOKE_WGRAD
    return false;
}

}  // namespace

// C++ entry points used by the preconditioner (same contracts as the C ABI).
uint64_t packed_bytes(const SmokeNnConvDesc& d) {
    return uint64_t(std::max<int64_t>(int64_t(d.cout) * padded_k(d.cin), int64_t(d.cin) * padded_k(d.cout))) * 2;
}

void pack(const SmokeNnConvDesc& d, const float* w, uint16_t* fwd, uint16_t* bwd, hipStream_t s) {
    const ConvArgs a = make(d);
    const int64_t n = std::max<int64_t>(int64_t(a.cout) * a.kpad, int64_t(a.cin) * padded_k(a.cout));
    pack_kernel<<<unsigned((n + 255) / 256), 256, 0, s>>>(a, w, fwd, bwd);
    hip_check(hipGetLastError(), "pack_kernel");
}

void forward(const SmokeNnConvDesc& d, const uint16_t* x, const uint16_t* w, const float* bias, const uint16_t* skip, int relu, uint16_t* y, float* yf,
             int f32_channels, hipStream_t s) {
    launch_conv(make(d), x, w, bias, skip, nullptr, relu, y, yf, f32_channels, s);
}

void backward_data(const SmokeNnConvDesc& d, const uint16_t* dy, const uint16_t* wb, const uint16_t* add, const uint16_t* mask, uint16_t* dx,
                   hipStream_t s) {
    const ConvArgs a = make(d);
    if (a.cin % 16 != 0) throw HipError("conv: the data gradient needs cin % 16 == 0");
    launch_conv(dual(a), dy, wb, nullptr, add, mask, 0, dx, nullptr, 0, s);
}

void backward_weights(const SmokeNnConvDesc& d, const uint16_t* x, const uint16_t* dy, float* dw, float* db, hipStream_t s) {
    const ConvArgs a = make(d);
    const int64_t total_rows = int64_t(a.batch) * a.out[2] * a.out[1];
    const int choice = g_conv_kernels.load(std::memory_order_relaxed);
    const bool staged = choice != SMOKE_NN_CONV_KERNELS_GATHER &&
                        (choice == SMOKE_NN_CONV_KERNELS_STAGED || (total_rows * a.out[0] >= kWgradStagedVoxels && a.out[0] >= 16)) &&
                        launch_wgrad_staged(a, x, dy, dw, s);
    if (staged) {
        hip_check(hipGetLastError(), "wgrad_staged_kernel");
    } else {
        const int rows = 8;
        const int64_t waves = (total_rows + rows - 1) / rows * 27 * ((a.cin + 15) / 16) * (a.cout / 16);
        wgrad_kernel<<<unsigned((waves + 3) / 4), 256, 0, s>>>(a, x, dy, dw, rows);
        hip_check(hipGetLastError(), "wgrad_kernel");
    }
    if (db) {
        const int64_t voxels = total_rows * a.out[0];
        const int64_t per_block = 256 / (a.cout / 8);
        const int64_t blocks = std::min<int64_t>(kBiasBlocks, (voxels + per_block - 1) / per_block);
        bias_grad_kernel<<<unsigned(blocks), 256, 0, s>>>(dy, voxels, a.cout, db);
        hip_check(hipGetLastError(), "bias_grad_kernel");
    }
}

}  // namespace nn
}  // namespace smoke

extern "C" {

uint64_t smoke_nn_packed_weight_bytes(const SmokeNnConvDesc* d) { return d ? smoke::nn::packed_bytes(*d) : 0; }

#define SMOKE_NN_GUARD(body)                                 \
    try {                                                    \
        body;                                                \
        return SMOKE_SIMULATION_RESULT_OK;                   \
    } catch (const std::bad_alloc&) {                        \
        return SMOKE_SIMULATION_RESULT_OUT_OF_MEMORY;        \
    } catch (...) {                                          \
        return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;      \
    }

SmokeSimulationResult smoke_nn_pack_weights(const SmokeNnConvDesc* d, const float* w, uint16_t* fwd, uint16_t* bwd, void* stream) {
    if (!d || !w) return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
    SMOKE_NN_GUARD(smoke::nn::pack(*d, w, fwd, bwd, static_cast<hipStream_t>(stream)))
}

SmokeSimulationResult smoke_nn_conv_forward(const SmokeNnConvDesc* d, const uint16_t* x, const uint16_t* w, const float* bias, const uint16_t* skip,
                                            int32_t relu, uint16_t* y, float* yf, int32_t f32_channels, void* stream) {
    if (!d || !x || !w || (!y && !yf)) return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
    SMOKE_NN_GUARD(smoke::nn::forward(*d, x, w, bias, skip, relu, y, yf, f32_channels, static_cast<hipStream_t>(stream)))
}

SmokeSimulationResult smoke_nn_conv_backward_data(const SmokeNnConvDesc* d, const uint16_t* dy, const uint16_t* wb, const uint16_t* add,
                                                  const uint16_t* mask, uint16_t* dx, void* stream) {
    if (!d || !dy || !wb || !dx) return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
    SMOKE_NN_GUARD(smoke::nn::backward_data(*d, dy, wb, add, mask, dx, static_cast<hipStream_t>(stream)))
}

SmokeSimulationResult smoke_nn_set_conv_kernels(int32_t choice) {
    if (choice < SMOKE_NN_CONV_KERNELS_AUTO || choice > SMOKE_NN_CONV_KERNELS_STAGED) return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
    smoke::nn::g_conv_kernels.store(choice, std::memory_order_relaxed);
    return SMOKE_SIMULATION_RESULT_OK;
}

SmokeSimulationResult smoke_nn_conv_backward_weights(const SmokeNnConvDesc* d, const uint16_t* x, const uint16_t* dy, float* dw, float* db,
                                                     void* stream) {
    if (!d || !x || !dy || !dw) return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
    SMOKE_NN_GUARD(smoke::nn::backward_weights(*d, x, dy, dw, db, static_cast<hipStream_t>(stream)))
}

}  // extern "C"
