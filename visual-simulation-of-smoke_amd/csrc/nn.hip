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
template <int NT, int SPLIT>
__global__ __launch_bounds__(SPLIT == 1 ? 256 : 64 * SPLIT) void conv_kernel(ConvArgs a, const uint16_t* __restrict__ x, const uint16_t* __restrict__ w,
                                                   const float* __restrict__ bias, const uint16_t* __restrict__ add,
                                                   const uint16_t* __restrict__ mask, int relu, uint16_t* __restrict__ y,
                                                   float* __restrict__ yf, int f32_channels) {
    const int lane = threadIdx.x & 63;
    const int wib = threadIdx.x >> 6;   // wave in block
    const int64_t wave = SPLIT == 1 ? int64_t(blockIdx.x) * 4 + wib : int64_t(blockIdx.x);
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
    if (first < ksteps) {
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
    // lane holds D[4 g + i][m]: voxel xtile * 16 + 4 g + i, channel nt * 16 + m
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
        const int n = nt * 16 + m;
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

// db[n] += sum over voxels of dy[.][n]
__global__ __launch_bounds__(256) void bias_grad_kernel(const uint16_t* __restrict__ dy, int64_t voxels, int cout, float* __restrict__ db) {
    const int n = threadIdx.x % cout;
    const int lanes_per_n = blockDim.x / cout;
    const int64_t chunk = 4096;
    const int64_t v0 = int64_t(blockIdx.x) * chunk;
    const int64_t v1 = min(voxels, v0 + chunk);
    float s = 0.0f;
    for (int64_t v = v0 + threadIdx.x / cout; v < v1; v += lanes_per_n) s += bf16_to_f32(dy[v * cout + n]);
    if (threadIdx.x < lanes_per_n * cout) atomicAdd(&db[n], s);
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
    switch (a.cout / 16) {
        case 1: conv_kernel<1, SPLIT><<<blocks, threads, 0, s>>>(a, x, w, bias, add, mask, relu, y, yf, f32_channels); break;
        case 2: conv_kernel<2, SPLIT><<<blocks, threads, 0, s>>>(a, x, w, bias, add, mask, relu, y, yf, f32_channels); break;
        case 3: conv_kernel<3, SPLIT><<<blocks, threads, 0, s>>>(a, x, w, bias, add, mask, relu, y, yf, f32_channels); break;
        case 4: conv_kernel<4, SPLIT><<<blocks, threads, 0, s>>>(a, x, w, bias, add, mask, relu, y, yf, f32_channels); break;
        default: throw HipError("conv: cout must be 16..64");
    }
}

void launch_conv(const ConvArgs& a, const uint16_t* x, const uint16_t* w, const float* bias, const uint16_t* add, const uint16_t* mask, int relu,
                 uint16_t* y, float* yf, int f32_channels, hipStream_t s) {
    const int64_t tiles = int64_t(a.batch) * a.out[2] * a.out[1] * ((a.out[0] + 15) / 16);
    if (tiles < kSplit16Tiles) launch_conv_split<16>(a, unsigned(tiles), x, w, bias, add, mask, relu, y, yf, f32_channels, s);
    else if (tiles < kSplit8Tiles) launch_conv_split<8>(a, unsigned(tiles), x, w, bias, add, mask, relu, y, yf, f32_channels, s);
    else if (tiles < kSplitTiles) launch_conv_split<4>(a, unsigned(tiles), x, w, bias, add, mask, relu, y, yf, f32_channels, s);
    else launch_conv_split<1>(a, unsigned((tiles + 3) / 4), x, w, bias, add, mask, relu, y, yf, f32_channels, s);
    hip_check(hipGetLastError(), "conv_kernel");
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
    const int rows = 8;
    const int64_t total_rows = int64_t(a.batch) * a.out[2] * a.out[1];
    const int64_t waves = (total_rows + rows - 1) / rows * 27 * ((a.cin + 15) / 16) * (a.cout / 16);
    wgrad_kernel<<<unsigned((waves + 3) / 4), 256, 0, s>>>(a, x, dy, dw, rows);
    hip_check(hipGetLastError(), "wgrad_kernel");
    if (db) {
        const int64_t voxels = total_rows * a.out[0];
        bias_grad_kernel<<<unsigned((voxels + 4095) / 4096), 256, 0, s>>>(dy, voxels, a.cout, db);
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

SmokeSimulationResult smoke_nn_conv_backward_weights(const SmokeNnConvDesc* d, const uint16_t* x, const uint16_t* dy, float* dw, float* db,
                                                     void* stream) {
    if (!d || !x || !dy || !dw) return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
    SMOKE_NN_GUARD(smoke::nn::backward_weights(*d, x, dy, dw, db, static_cast<hipStream_t>(stream)))
}

}  // extern "C"
