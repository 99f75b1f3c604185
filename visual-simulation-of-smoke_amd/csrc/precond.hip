// The learned pressure preconditioner: a 3-level convolutional
// encoder-decoder (architecture, normalisation and loss documented in
// visual-simulation-of-smoke-3d-nn.h) trained and applied with the bfloat16
// matrix-core convolutions of nn.hip.
#include "context.hpp"
#include "nn.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <random>
#include <vector>

using smoke::HipError;
using smoke::hip_check;

// Workspace of smoke_nn_preconditioner_apply_static: the activations, the
// normalised residual, the output and the sums of one inference live in this
// array instead of buffers the network allocates.  Each GPU holds its own copy
// until the process ends; calls on one GPU share it.
constexpr size_t kStaticWorkspaceBytes = size_t(16) << 20;
static __device__ __align__(256) unsigned char scratch[kStaticWorkspaceBytes];

struct SmokeNnPreconditioner {
    struct Layer {
        int cin, cout, mode, in_level, out_level, relu;
        int cin_used;          // input channels that carry data (the rest are zero padding)
        int cout_used;         // output channels that are used
        int64_t w_off, b_off;  // offsets into the parameter array
        uint16_t* fwd = nullptr;
        uint16_t* bwd = nullptr;
    };
    std::vector<Layer> layers;
    int64_t nparams = 0;
    float* params = nullptr;
    float* grads = nullptr;
    float* adam_m = nullptr;
    float* adam_v = nullptr;
    int adam_t = 0;
    bool packed_dirty = true;
    int loss = SMOKE_NN_LOSS_RESIDUAL;

    // workspace for one (batch, dims); grads only when trained
    int ws_batch = 0, ws_dims[3] = {0, 0, 0};
    bool ws_train = false;
    std::vector<void*> ws;
    uint16_t* T[12] = {};   // activations, see the enum below
    uint16_t* G[7] = {};    // activation gradients
    float* rn = nullptr;    // normalised residual (the loss target)
    float* z = nullptr;     // network output (channel 0, FP32)
    float* q = nullptr;     // loss residual A z - r
    double* sums = nullptr; // per volume: sum r^2, ||r||^2 over fluid rows, ||q||^2
    double* partials = nullptr;   // per volume and block of the sum-of-squares pass: its partial sum of r^2
};

namespace smoke {
namespace {

using Net = SmokeNnPreconditioner;

enum Tensor { kIn, kA0, kE0, kD1, kE1, kD2, kE2, kUp1, kC1u, kUp0, kC0u, kTensorCount };
enum Grad { kGOut, kG0a, kG0b, kG1a, kG1b, kG2a, kG2b };

// layer table: mode, cin, cout, input level, output level, relu, used in/out channels
struct LayerSpec {
    int mode, cin, cout, in_level, out_level, relu, cin_used, cout_used;
};
constexpr LayerSpec kLayers[] = {
    {SMOKE_NN_CONV_SAME, 8, 16, 0, 0, 1, 2, 16},    // c0a
    {SMOKE_NN_CONV_SAME, 16, 16, 0, 0, 1, 16, 16},  // c0b -> e0
    {SMOKE_NN_CONV_DOWN, 16, 32, 0, 1, 1, 16, 32},  // d1
    {SMOKE_NN_CONV_SAME, 32, 32, 1, 1, 1, 32, 32},  // c1 -> e1
    {SMOKE_NN_CONV_DOWN, 32, 64, 1, 2, 1, 32, 64},  // d2
    {SMOKE_NN_CONV_SAME, 64, 64, 2, 2, 1, 64, 64},  // c2 -> e2
    {SMOKE_NN_CONV_UP, 64, 32, 2, 1, 1, 64, 32},    // u1 (+ e1)
    {SMOKE_NN_CONV_SAME, 32, 32, 1, 1, 1, 32, 32},  // c1u
    {SMOKE_NN_CONV_UP, 32, 16, 1, 0, 1, 32, 16},    // u0 (+ e0)
    {SMOKE_NN_CONV_SAME, 16, 16, 0, 0, 1, 16, 16},  // c0u
    {SMOKE_NN_CONV_SAME, 16, 16, 0, 0, 0, 16, 1},   // out
};
constexpr uint8_t kIdentity = 0x40;
constexpr int kSumBlocks = 256;   // blocks of the sum-of-squares pass per volume (fixed: the sum's order does not depend on the size)

struct Dims {
    int n[3][3];   // level -> (W, H, D)
};

Dims level_dims(const int dims[3]) {
    Dims d{};
    for (int k = 0; k < 3; ++k) {
        d.n[0][k] = dims[k];
        d.n[1][k] = (dims[k] + 1) / 2;
        d.n[2][k] = (d.n[1][k] + 1) / 2;
    }
    return d;
}

int64_t voxels(const Dims& d, int level, int batch) { return int64_t(batch) * d.n[level][0] * d.n[level][1] * d.n[level][2]; }

SmokeNnConvDesc desc_of(const Net::Layer& L, const Dims& d, int batch) {
    SmokeNnConvDesc c{};
    c.batch = batch;
    for (int k = 0; k < 3; ++k) {
        c.in[k] = d.n[L.in_level][k];
        c.out[k] = d.n[L.out_level][k];
    }
    c.cin = L.cin;
    c.cout = L.cout;
    c.mode = L.mode;
    return c;
}

template <typename T>
T* dalloc(Net& net, size_t count) {
    void* p = nullptr;
    hip_check(hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)), "preconditioner buffer");
    net.ws.push_back(p);
    return static_cast<T*>(p);
}

void free_workspace(Net& net) {
    for (void* p : net.ws) (void)hipFree(p);
    net.ws.clear();
    std::memset(net.T, 0, sizeof(net.T));
    std::memset(net.G, 0, sizeof(net.G));
    net.rn = net.z = net.q = nullptr;
    net.sums = nullptr;
    net.partials = nullptr;
    net.ws_batch = 0;
}

void ensure_workspace(Net& net, int batch, const int dims[3], bool train) {
    if (net.ws_batch == batch && std::memcmp(net.ws_dims, dims, sizeof(net.ws_dims)) == 0 && (net.ws_train || !train)) return;
    free_workspace(net);
    const Dims d = level_dims(dims);
    const int level_of[kTensorCount] = {0, 0, 0, 1, 1, 2, 2, 1, 1, 0, 0};
    const int chan_of[kTensorCount] = {8, 16, 16, 32, 32, 64, 64, 32, 32, 16, 16};
    for (int t = 0; t < kTensorCount; ++t) net.T[t] = dalloc<uint16_t>(net, size_t(voxels(d, level_of[t], batch)) * chan_of[t]);
    if (train) {
        const int glevel[7] = {0, 0, 0, 1, 1, 2, 2};
        const int gchan[7] = {16, 16, 16, 32, 32, 64, 64};
        for (int g = 0; g < 7; ++g) net.G[g] = dalloc<uint16_t>(net, size_t(voxels(d, glevel[g], batch)) * gchan[g]);
        net.q = dalloc<float>(net, size_t(voxels(d, 0, batch)));
    }
    net.rn = dalloc<float>(net, size_t(voxels(d, 0, batch)));
    net.z = dalloc<float>(net, size_t(voxels(d, 0, batch)));
    net.sums = dalloc<double>(net, size_t(3 * batch));
    net.partials = dalloc<double>(net, size_t(kSumBlocks) * batch);
    net.ws_batch = batch;
    std::memcpy(net.ws_dims, dims, sizeof(net.ws_dims));
    net.ws_train = train;
}

__device__ __forceinline__ uint16_t to_bf16(float v) {
    uint32_t u = __float_as_uint(v);
    if ((u & 0x7fffffffu) > 0x7f800000u) return uint16_t((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return uint16_t(u >> 16);
}

__device__ __forceinline__ float from_bf16(uint16_t h) { return __uint_as_float(uint32_t(h) << 16); }

__global__ void zero_sums_kernel(double* __restrict__ p, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0.0;
}

// per volume (blockIdx.y): sum of r^2 in a fixed order -- each of the
// kSumBlocks blocks sums a strided share of the cells (wave shuffles, then the
// block's waves in order) into its partial, and finalize_sums_kernel adds the
// partials in block order -- so the normalisation s, and with it every
// bfloat16 input, is the same on every run
__global__ __launch_bounds__(256) void sum_squares_kernel(const float* __restrict__ r, int64_t cells, double* __restrict__ partials) {
    const int b = blockIdx.y;
    double s = 0.0;
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < cells; i += int64_t(gridDim.x) * 256) {
        const double v = r[int64_t(b) * cells + i];
        s += v * v;
    }
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    __shared__ double wave[4];
    if ((threadIdx.x & 63) == 0) wave[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) partials[int64_t(b) * gridDim.x + blockIdx.x] = ((wave[0] + wave[1]) + wave[2]) + wave[3];
}

// one block per volume: sums[3 b] = the volume's partials added in block order
__global__ __launch_bounds__(64) void finalize_sums_kernel(const double* __restrict__ partials, int blocks, double* __restrict__ sums) {
    const int b = blockIdx.x;
    double s = 0.0;
    for (int i = threadIdx.x; i < blocks; i += 64) s += partials[int64_t(b) * blocks + i];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if (threadIdx.x == 0) sums[3 * b] = s;
}

// input channels (r / s, identity mask, zeros), the normalised target rn and
// its squared norm over the fluid rows
__global__ __launch_bounds__(256) void input_kernel(const float* __restrict__ r, const uint8_t* __restrict__ flags, int64_t cells,
                                                    double* __restrict__ sums, uint16_t* __restrict__ in8, float* __restrict__ rn) {
    const int b = blockIdx.y;
    const double ss = sums[3 * b];
    const double s = ss > 0.0 ? sqrt(ss / double(cells)) : 1.0;
    double acc = 0.0;
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < cells; i += int64_t(gridDim.x) * 256) {
        const int64_t v = int64_t(b) * cells + i;
        const bool ident = (flags[v] & kIdentity) != 0;
        const uint16_t h = to_bf16(float(double(r[v]) / s));
        const float x = ident ? 0.0f : from_bf16(h);
        rn[v] = x;
        acc += double(x) * double(x);
        uint16_t* o = in8 + v * 8;
        o[0] = ident ? uint16_t(0) : h;
        o[1] = ident ? to_bf16(1.0f) : uint16_t(0);
#pragma unroll
        for (int c = 2; c < 8; ++c) o[c] = 0;
    }
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(&sums[3 * b + 1], acc);
}

// (A v)_i on one volume of a batch: identity rows give 0 here (the loss and
// the preconditioner only look at fluid rows), neighbours outside the volume
// read as zero.
__device__ __forceinline__ float apply_A(const float* __restrict__ v, const uint8_t* __restrict__ flags, int W, int H, int D, int64_t base, int x,
                                         int y, int z) {
    const int64_t i = base + (int64_t(z) * H + y) * W + x;
    const uint8_t f = flags[i];
    if (f & kIdentity) return 0.0f;
    const unsigned bits = f & 0x3fu;
    const float d = float(max(__popc(bits), 1));
    auto at = [&](int xx, int yy, int zz) -> float {
        if (xx < 0 || yy < 0 || zz < 0 || xx >= W || yy >= H || zz >= D) return 0.0f;
        const int64_t j = base + (int64_t(zz) * H + yy) * W + xx;
        return (flags[j] & kIdentity) ? 0.0f : v[j];
    };
    float s = d * v[i];
    if (bits & 1u) s -= at(x - 1, y, z);
    if (bits & 2u) s -= at(x + 1, y, z);
    if (bits & 4u) s -= at(x, y - 1, z);
    if (bits & 8u) s -= at(x, y + 1, z);
    if (bits & 16u) s -= at(x, y, z - 1);
    if (bits & 32u) s -= at(x, y, z + 1);
    return s;
}

// q = A z - rn on fluid rows (0 on identity rows); per volume ||q||^2
// (residual loss) or z.(q - rn) = z^T A z - 2 z^T rn (energy loss)
__global__ __launch_bounds__(256) void loss_residual_kernel(const float* __restrict__ z, const uint8_t* __restrict__ flags, int W, int H, int D,
                                                            const float* __restrict__ rn, float* __restrict__ q, double* __restrict__ sums,
                                                            int energy) {
    const int b = blockIdx.y;
    const int64_t cells = int64_t(W) * H * D, base = int64_t(b) * cells;
    double acc = 0.0;
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < cells; i += int64_t(gridDim.x) * 256) {
        const int x = int(i % W), y = int((i / W) % H), zz = int(i / (int64_t(W) * H));
        const bool ident = (flags[base + i] & kIdentity) != 0;
        const float qi = ident ? 0.0f : apply_A(z, flags, W, H, D, base, x, y, zz) - rn[base + i];
        q[base + i] = qi;
        if (energy) acc += ident ? 0.0 : double(z[base + i]) * (double(qi) - double(rn[base + i]));
        else acc += double(qi) * double(qi);
    }
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(&sums[3 * b + 2], acc);
}

// d loss / d z on fluid rows, written as the bfloat16 gradient of the output
// layer's 16 channels (channel 0 only): (2 / batch) A q / ||r||^2 for the
// residual loss, q / (batch ||r||^2) for the energy loss (A is symmetric)
__global__ __launch_bounds__(256) void loss_grad_kernel(const float* __restrict__ q, const uint8_t* __restrict__ flags, int W, int H, int D, int batch,
                                                        const double* __restrict__ sums, uint16_t* __restrict__ gout, int energy) {
    const int b = blockIdx.y;
    const int64_t cells = int64_t(W) * H * D, base = int64_t(b) * cells;
    const double rr = sums[3 * b + 1];
    const float scale = rr > 0.0 ? float((energy ? 1.0 : 2.0) / (double(batch) * rr)) : 0.0f;
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < cells; i += int64_t(gridDim.x) * 256) {
        const int x = int(i % W), y = int((i / W) % H), zz = int(i / (int64_t(W) * H));
        const float g = energy ? scale * q[base + i] : scale * apply_A(q, flags, W, H, D, base, x, y, zz);
        uint16_t* o = gout + (base + i) * 16;
        o[0] = to_bf16(g);
#pragma unroll
        for (int c = 1; c < 16; ++c) o[c] = 0;
    }
}

// z = s * out on fluid rows, 0 on identity rows (inference)
__global__ __launch_bounds__(256) void output_kernel(const float* __restrict__ out, const uint8_t* __restrict__ flags, int64_t cells,
                                                     const double* __restrict__ sums, float* __restrict__ z) {
    const double ss = sums[0];
    const float s = float(ss > 0.0 ? sqrt(ss / double(cells)) : 1.0);
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < cells; i += int64_t(gridDim.x) * 256)
        z[i] = (flags[i] & kIdentity) ? 0.0f : s * out[i];
}

// ---- the multilevel preconditioner's coarse-grid solve in few launches ----
//
// The coarsest level is small (16^3 at 256^3 with five levels), where every
// launch costs more than its work: the deflation of v = 1 (c = v.r / v.Av,
// rd = r - c A v), the network input (rd / s with s the RMS of rd) and, after
// the network, z = s out + (c - Av.z / v.Av) v run as one single-block
// kernel before the convolutions and one after.  Sums run in a fixed order
// (each thread's cells in order, wave shuffles, then the waves in order).
constexpr int kCoarseThreads = 1024;

__device__ double block_sum_fixed(double v, double* red) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    __syncthreads();   // the previous sum's reads of red are done
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0.0;
    for (int k = 0; k < kCoarseThreads / 64; ++k) t += red[k];
    return t;
}

// Each thread owns cells t, t + 1024, ... (at most kCoarsePerThread of them),
// loaded once into registers: one memory latency per kernel.
constexpr int kCoarsePerThread = 8;

// defl[0] = v.Av (set up with the hierarchy); writes defl[1] = c, rd, the
// network input in8 and sums[0] = ||rd||^2 (the output's scale)
__global__ __launch_bounds__(kCoarseThreads) void coarse_prep_kernel(const float* __restrict__ r, const float* __restrict__ av,
                                                                     const uint8_t* __restrict__ flags, int cells, double* __restrict__ defl,
                                                                     float* __restrict__ rd, double* __restrict__ sums,
                                                                     uint16_t* __restrict__ in8) {
    __shared__ double red[kCoarseThreads / 64];
    const int t = int(threadIdx.x);
    float rv[kCoarsePerThread], avv[kCoarsePerThread];
    bool fluid[kCoarsePerThread];
#pragma unroll
    for (int k = 0; k < kCoarsePerThread; ++k) {
        const int i = t + k * kCoarseThreads;
        const bool in = i < cells;
        rv[k] = in ? r[i] : 0.0f;
        avv[k] = in ? av[i] : 0.0f;
        fluid[k] = in && !(flags[i] & kIdentity);
    }
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < kCoarsePerThread; ++k) s += fluid[k] ? double(rv[k]) : 0.0;
    const double vr = block_sum_fixed(s, red);
    const double vav = defl[0];
    const double c = vav > 0.0 ? vr / vav : 0.0;
    if (t == 0) defl[1] = c;
    const float cf = float(c);
    float v[kCoarsePerThread];
    double ss = 0.0;
#pragma unroll
    for (int k = 0; k < kCoarsePerThread; ++k) {
        v[k] = fluid[k] ? fmaf(-cf, avv[k], rv[k]) : 0.0f;
        ss += double(v[k]) * double(v[k]);
    }
    const double tot = block_sum_fixed(ss, red);
    if (t == 0) sums[0] = tot;
    const double sc = tot > 0.0 ? sqrt(tot / double(cells)) : 1.0;
#pragma unroll
    for (int k = 0; k < kCoarsePerThread; ++k) {
        const int i = t + k * kCoarseThreads;
        if (i >= cells) break;
        rd[i] = v[k];
        // channels 0-7: r / s (0 on identity rows), the identity mask, zeros
        const uint32_t c0 = fluid[k] ? uint32_t(to_bf16(float(double(v[k]) / sc))) : 0u;
        const uint32_t c1 = fluid[k] ? 0u : uint32_t(to_bf16(1.0f));
        *reinterpret_cast<uint4*>(in8 + int64_t(i) * 8) = make_uint4(c0 | (c1 << 16), 0u, 0u, 0u);
    }
}

// z = s out on fluid rows (0 on identity rows); then `sweeps.count` damped
// Jacobi sweeps z <- z + omega D^-1 (rd - A z) on the fluid rows (the
// multilevel cycle's sweep: neighbours from the flag byte, periodic images
// at the domain edge, OUTFLOW faces on the diagonal, taps summed in the
// order x-, x+, y-, y+, z-, z+), z kept in LDS; then z += (c - Av.z / v.Av)
// on fluid rows
__global__ __launch_bounds__(kCoarseThreads) void coarse_finish_kernel(const float* __restrict__ out, const uint8_t* __restrict__ flags,
                                                                       const float* __restrict__ av, const float* __restrict__ rd, int n0, int n1,
                                                                       int n2, CoarseSweeps sweeps, const double* __restrict__ sums,
                                                                       const double* __restrict__ defl, float* __restrict__ z) {
    __shared__ double red[kCoarseThreads / 64];
    __shared__ float zs[kCoarseSolveMaxCells];
    const int t = int(threadIdx.x);
    const int cells = n0 * n1 * n2;
    const double tot = sums[0];
    const float sc = float(tot > 0.0 ? sqrt(tot / double(cells)) : 1.0);
    float zv[kCoarsePerThread];
    bool fluid[kCoarsePerThread];
#pragma unroll
    for (int k = 0; k < kCoarsePerThread; ++k) {
        const int i = t + k * kCoarseThreads;
        const bool in = i < cells;
        fluid[k] = in && !(flags[i] & kIdentity);
        zv[k] = fluid[k] ? sc * out[i] : 0.0f;
    }
    if (sweeps.count > 0) {
        unsigned fl[kCoarsePerThread];
        float rv[kCoarsePerThread], inv[kCoarsePerThread], dg[kCoarsePerThread];
#pragma unroll
        for (int k = 0; k < kCoarsePerThread; ++k) {
            const int i = t + k * kCoarseThreads;
            const bool in = i < cells;
            fl[k] = in ? unsigned(flags[i]) : 0u;
            rv[k] = in ? rd[i] : 0.0f;
            if (in) zs[i] = zv[k];
            const int x = i % n0, y = (i / n0) % n1, zz = i / (n0 * n1);
            const int of = (x == 0 && sweeps.outflow[0]) + (x == n0 - 1 && sweeps.outflow[1]) + (y == 0 && sweeps.outflow[2]) +
                           (y == n1 - 1 && sweeps.outflow[3]) + (zz == 0 && sweeps.outflow[4]) + (zz == n2 - 1 && sweeps.outflow[5]);
            const int d = __popc(fl[k] & 0x3fu) + of;
            dg[k] = float(d > 1 ? d : 1);
            inv[k] = __builtin_amdgcn_rcpf(dg[k]);
        }
        const int sy = n0, sz = n0 * n1;
        for (int sw = 0; sw < sweeps.count; ++sw) {
            __syncthreads();   // zs holds the previous iterate
#pragma unroll
            for (int k = 0; k < kCoarsePerThread; ++k) {
                const int i = t + k * kCoarseThreads;
                if (!fluid[k]) continue;
                const unsigned f = fl[k];
                const int x = i % n0, y = (i / n0) % n1, zz = i / (n0 * n1);
                // neighbour taps (set flag bits only: the neighbour exists, at the edge its periodic image)
                const float t0 = (f & 1u) ? zs[x > 0 ? i - 1 : i + n0 - 1] : 0.0f;
                const float t1 = (f & 2u) ? zs[x < n0 - 1 ? i + 1 : i - (n0 - 1)] : 0.0f;
                const float t2 = (f & 4u) ? zs[y > 0 ? i - sy : i + (n1 - 1) * sy] : 0.0f;
                const float t3 = (f & 8u) ? zs[y < n1 - 1 ? i + sy : i - (n1 - 1) * sy] : 0.0f;
                const float t4 = (f & 16u) ? zs[zz > 0 ? i - sz : i + (n2 - 1) * sz] : 0.0f;
                const float t5 = (f & 32u) ? zs[zz < n2 - 1 ? i + sz : i - (n2 - 1) * sz] : 0.0f;
                const float s = t0 + t1 + t2 + t3 + t4 + t5;
                const float xi = zv[k];
                const float res = rv[k] - (dg[k] * xi - s);
                zv[k] = xi + sweeps.omega * res * inv[k];
            }
            __syncthreads();   // every read of the previous iterate is done
#pragma unroll
            for (int k = 0; k < kCoarsePerThread; ++k) {
                const int i = t + k * kCoarseThreads;
                if (fluid[k]) zs[i] = zv[k];
            }
        }
    }
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < kCoarsePerThread; ++k) {
        const int i = t + k * kCoarseThreads;
        s += i < cells ? double(av[i]) * double(zv[k]) : 0.0;
    }
    const double az = block_sum_fixed(s, red);
    const double vav = defl[0];
    const float coef = vav > 0.0 ? float(defl[1] - az / vav) : 0.0f;
#pragma unroll
    for (int k = 0; k < kCoarsePerThread; ++k) {
        const int i = t + k * kCoarseThreads;
        if (i >= cells) break;
        z[i] = fluid[k] ? zv[k] + coef : 0.0f;
    }
}

__global__ void adam_kernel(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v, int64_t n, float lr,
                            float bc1, float bc2) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = 0.9f * m[i] + 0.1f * gi;
    const float vi = 0.999f * v[i] + 0.001f * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + 1e-8f);
}

unsigned grid_for(int64_t cells) { return unsigned(std::min<int64_t>((cells + 255) / 256, 4096)); }

void pack_all(Net& net, hipStream_t s) {
    if (!net.packed_dirty) return;
    const int dims1[3] = {16, 16, 16};
    const Dims d = level_dims(dims1);
    for (auto& L : net.layers) nn::pack(desc_of(L, d, 1), net.params + L.w_off, L.fwd, L.bwd, s);
    net.packed_dirty = false;
}

// forward pass; the output layer's channel 0 goes to net.z (FP32)
void forward(Net& net, int batch, const Dims& d, hipStream_t s) {
    auto conv = [&](int li, Tensor in, Tensor out, const uint16_t* skip) {
        const Net::Layer& L = net.layers[li];
        nn::forward(desc_of(L, d, batch), net.T[in], L.fwd, net.params + L.b_off, skip, L.relu, net.T[out], nullptr, 0, s);
    };
    conv(0, kIn, kA0, nullptr);
    conv(1, kA0, kE0, nullptr);
    conv(2, kE0, kD1, nullptr);
    conv(3, kD1, kE1, nullptr);
    conv(4, kE1, kD2, nullptr);
    conv(5, kD2, kE2, nullptr);
    conv(6, kE2, kUp1, net.T[kE1]);
    conv(7, kUp1, kC1u, nullptr);
    conv(8, kC1u, kUp0, net.T[kE0]);
    conv(9, kUp0, kC0u, nullptr);
    const Net::Layer& L = net.layers[10];
    nn::forward(desc_of(L, d, batch), net.T[kC0u], L.fwd, net.params + L.b_off, nullptr, 0, nullptr, net.z, 1, s);
}

// backward pass from the output gradient in G[kGOut]; accumulates into grads
void backward(Net& net, int batch, const Dims& d, hipStream_t s) {
    auto wgrad = [&](int li, const uint16_t* x, const uint16_t* dy) {
        const Net::Layer& L = net.layers[li];
        nn::backward_weights(desc_of(L, d, batch), x, dy, net.grads + L.w_off, net.grads + L.b_off, s);
    };
    auto dgrad = [&](int li, const uint16_t* dy, const uint16_t* add, const uint16_t* mask, uint16_t* dx) {
        const Net::Layer& L = net.layers[li];
        nn::backward_data(desc_of(L, d, batch), dy, L.bwd, add, mask, dx, s);
    };
    uint16_t** T = net.T;
    uint16_t** G = net.G;
    wgrad(10, T[kC0u], G[kGOut]);
    dgrad(10, G[kGOut], nullptr, T[kC0u], G[kG0a]);            // d c0u (pre-activation)
    wgrad(9, T[kUp0], G[kG0a]);
    dgrad(9, G[kG0a], nullptr, T[kUp0], G[kG0b]);              // d up0 (pre); also d e0 through the skip
    wgrad(8, T[kC1u], G[kG0b]);
    dgrad(8, G[kG0b], nullptr, T[kC1u], G[kG1a]);              // d c1u
    wgrad(7, T[kUp1], G[kG1a]);
    dgrad(7, G[kG1a], nullptr, T[kUp1], G[kG1b]);              // d up1 (pre); also d e1 through the skip
    wgrad(6, T[kE2], G[kG1b]);
    dgrad(6, G[kG1b], nullptr, T[kE2], G[kG2a]);               // d e2
    wgrad(5, T[kD2], G[kG2a]);
    dgrad(5, G[kG2a], nullptr, T[kD2], G[kG2b]);               // d d2
    wgrad(4, T[kE1], G[kG2b]);
    dgrad(4, G[kG2b], G[kG1b], T[kE1], G[kG1a]);               // d e1 = (down path + skip) * [e1 > 0]
    wgrad(3, T[kD1], G[kG1a]);
    dgrad(3, G[kG1a], nullptr, T[kD1], G[kG1b]);               // d d1
    wgrad(2, T[kE0], G[kG1b]);
    dgrad(2, G[kG1b], G[kG0b], T[kE0], G[kG0a]);               // d e0 = (down path + skip) * [e0 > 0]
    wgrad(1, T[kA0], G[kG0a]);
    dgrad(1, G[kG0a], nullptr, T[kA0], G[kG0b]);               // d a0
    wgrad(0, T[kIn], G[kG0b]);
}

void prepare_input(Net& net, const float* r, const uint8_t* flags, int batch, int64_t cells, hipStream_t s) {
    zero_sums_kernel<<<1, 64, 0, s>>>(net.sums, 3 * batch);
    sum_squares_kernel<<<dim3(kSumBlocks, batch), 256, 0, s>>>(r, cells, net.partials);
    finalize_sums_kernel<<<batch, 64, 0, s>>>(net.partials, kSumBlocks, net.sums);
    const dim3 grid(grid_for(cells), batch);
    input_kernel<<<grid, 256, 0, s>>>(r, flags, cells, net.sums, net.T[kIn], net.rn);
    hip_check(hipGetLastError(), "preconditioner input");
}

void init_parameters(Net& net, uint64_t seed) {
    std::vector<float> h(size_t(net.nparams), 0.0f);
    std::mt19937_64 rng(seed);
    std::normal_distribution<double> normal(0.0, 1.0);
    for (const auto& L : net.layers) {
        const double sd = std::sqrt(2.0 / (27.0 * L.cin_used));
        for (int t = 0; t < 27; ++t)
            for (int c = 0; c < L.cin; ++c)
                for (int n = 0; n < L.cout; ++n) {
                    const double v = normal(rng);
                    if (c < L.cin_used && n < L.cout_used) h[size_t(L.w_off + (int64_t(t) * L.cin + c) * L.cout + n)] = float(sd * v);
                }
    }
    hip_check(hipMemcpy(net.params, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice), "init parameters");
    hip_check(hipMemset(net.adam_m, 0, h.size() * sizeof(float)), "adam m");
    hip_check(hipMemset(net.adam_v, 0, h.size() * sizeof(float)), "adam v");
    net.adam_t = 0;
    net.packed_dirty = true;
}

void destroy(Net* net) {
    if (!net) return;
    free_workspace(*net);
    for (auto& L : net->layers) {
        if (L.fwd) (void)hipFree(L.fwd);
        if (L.bwd) (void)hipFree(L.bwd);
    }
    for (float* p : {net->params, net->grads, net->adam_m, net->adam_v})
        if (p) (void)hipFree(p);
    delete net;
}

Net* create(uint64_t seed) {
    Net* net = new Net();
    try {
        int64_t off = 0;
        const int dims1[3] = {16, 16, 16};
        const Dims d = level_dims(dims1);
        for (const LayerSpec& sp : kLayers) {
            Net::Layer L{};
            L.cin = sp.cin;
            L.cout = sp.cout;
            L.mode = sp.mode;
            L.in_level = sp.in_level;
            L.out_level = sp.out_level;
            L.relu = sp.relu;
            L.cin_used = sp.cin_used;
            L.cout_used = sp.cout_used;
            L.w_off = off;
            off += 27LL * sp.cin * sp.cout;
            L.b_off = off;
            off += sp.cout;
            net->layers.push_back(L);
        }
        net->nparams = off;
        for (float** p : {&net->params, &net->grads, &net->adam_m, &net->adam_v})
            hip_check(hipMalloc(reinterpret_cast<void**>(p), size_t(off) * sizeof(float)), "parameters");
        for (auto& L : net->layers) {
            const uint64_t bytes = nn::packed_bytes(desc_of(L, d, 1));
            hip_check(hipMalloc(reinterpret_cast<void**>(&L.fwd), bytes), "packed weights");
            hip_check(hipMalloc(reinterpret_cast<void**>(&L.bwd), bytes), "packed weights");
        }
        init_parameters(*net, seed);
        return net;
    } catch (...) {
        destroy(net);
        throw;
    }
}

}  // namespace

void preconditioner_apply(SmokeNnPreconditioner& net, const float* r, const uint8_t* flags, const int dims[3], float* z, hipStream_t s) {
    ensure_workspace(net, 1, dims, false);
    pack_all(net, s);
    const Dims d = level_dims(dims);
    const int64_t cells = voxels(d, 0, 1);
    prepare_input(net, r, flags, 1, cells, s);
    forward(net, 1, d, s);
    output_kernel<<<grid_for(cells), 256, 0, s>>>(net.z, flags, cells, net.sums, z);
    hip_check(hipGetLastError(), "preconditioner output");
}

void preconditioner_coarse_solve(SmokeNnPreconditioner& net, const float* r, const float* av, const uint8_t* flags, double* defl, const int dims[3],
                                 const CoarseSweeps& sweeps, float* rd, float* z, hipStream_t s) {
    const int64_t cells = int64_t(dims[0]) * dims[1] * dims[2];
    if (cells > kCoarseSolveMaxCells) throw HipError("coarse solve: volume too large for the single-block kernels");
    if (sweeps.count < 0 || sweeps.count > kCoarseSolveMaxSweeps) throw HipError("coarse solve: sweep count out of range");
    ensure_workspace(net, 1, dims, false);
    pack_all(net, s);
    const Dims d = level_dims(dims);
    coarse_prep_kernel<<<1, kCoarseThreads, 0, s>>>(r, av, flags, int(cells), defl, rd, net.sums, net.T[kIn]);
    hip_check(hipGetLastError(), "coarse_prep_kernel");
    forward(net, 1, d, s);
    coarse_finish_kernel<<<1, kCoarseThreads, 0, s>>>(net.z, flags, av, rd, dims[0], dims[1], dims[2], sweeps, net.sums, defl, z);
    hip_check(hipGetLastError(), "coarse_finish_kernel");
}

// Device address of `scratch` on the current device, looked up once per
// device (the host-side symbol is not a device address).
void* static_workspace() {
    constexpr int kMaxDevices = 64;
    static std::mutex m;
    static void* ptr[kMaxDevices] = {};
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) throw HipError("device index out of range");
    std::lock_guard<std::mutex> lock(m);
    if (!ptr[dev]) {
        void* p = nullptr;
        hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        ptr[dev] = p;
    }
    return ptr[dev];
}

// Offsets of one inference's buffers in the static workspace (256-byte
// aligned): the activations, rn, z, the sums and the sum partials.
struct StaticLayout {
    size_t tensor[kTensorCount];
    size_t rn, z, sums, partials, total;
};

StaticLayout static_layout(const int dims[3]) {
    const Dims d = level_dims(dims);
    const int level_of[kTensorCount] = {0, 0, 0, 1, 1, 2, 2, 1, 1, 0, 0};
    const int chan_of[kTensorCount] = {8, 16, 16, 32, 32, 64, 64, 32, 32, 16, 16};
    StaticLayout L{};
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t at = off;
        off += (bytes + 255) / 256 * 256;
        return at;
    };
    for (int t = 0; t < kTensorCount; ++t) L.tensor[t] = take(size_t(voxels(d, level_of[t], 1)) * chan_of[t] * sizeof(uint16_t));
    L.rn = take(size_t(voxels(d, 0, 1)) * sizeof(float));
    L.z = take(size_t(voxels(d, 0, 1)) * sizeof(float));
    L.sums = take(3 * sizeof(double));
    L.partials = take(size_t(kSumBlocks) * sizeof(double));
    L.total = off;
    return L;
}

// The same kernels as preconditioner_apply, on buffers in the static
// workspace: no device memory is allocated.
void preconditioner_apply_static(SmokeNnPreconditioner& net, const float* r, const uint8_t* flags, const int dims[3], float* z, hipStream_t s) {
    const StaticLayout L = static_layout(dims);
    if (L.total > kStaticWorkspaceBytes) throw HipError("volume too large for the static workspace");
    unsigned char* base = static_cast<unsigned char*>(static_workspace());
    // the network's own workspace stands aside for this call (kernel
    // arguments are taken at launch, so it is restored right after)
    uint16_t* saved_T[kTensorCount];
    std::memcpy(saved_T, net.T, sizeof(saved_T));
    float* saved_rn = net.rn;
    float* saved_z = net.z;
    double* saved_sums = net.sums;
    double* saved_partials = net.partials;
    auto restore = [&] {
        std::memcpy(net.T, saved_T, sizeof(saved_T));
        net.rn = saved_rn;
        net.z = saved_z;
        net.sums = saved_sums;
        net.partials = saved_partials;
    };
    for (int t = 0; t < kTensorCount; ++t) net.T[t] = reinterpret_cast<uint16_t*>(base + L.tensor[t]);
    net.rn = reinterpret_cast<float*>(base + L.rn);
    net.z = reinterpret_cast<float*>(base + L.z);
    net.sums = reinterpret_cast<double*>(base + L.sums);
    net.partials = reinterpret_cast<double*>(base + L.partials);
    try {
        pack_all(net, s);
        const Dims d = level_dims(dims);
        const int64_t cells = voxels(d, 0, 1);
        prepare_input(net, r, flags, 1, cells, s);
        forward(net, 1, d, s);
        output_kernel<<<grid_for(cells), 256, 0, s>>>(net.z, flags, cells, net.sums, z);
        hip_check(hipGetLastError(), "preconditioner output");
    } catch (...) {
        restore();
        throw;
    }
    restore();
}

void preconditioner_train_step(SmokeNnPreconditioner& net, const float* r, const uint8_t* flags, int batch, const int dims[3], float lr, double* loss,
                               float* host_grads, hipStream_t s) {
    ensure_workspace(net, batch, dims, true);
    pack_all(net, s);
    const Dims d = level_dims(dims);
    const int64_t cells = voxels(d, 0, 1);
    prepare_input(net, r, flags, batch, cells, s);
    forward(net, batch, d, s);
    const dim3 grid(grid_for(cells), batch);
    const int energy = net.loss == SMOKE_NN_LOSS_ENERGY ? 1 : 0;
    loss_residual_kernel<<<grid, 256, 0, s>>>(net.z, flags, dims[0], dims[1], dims[2], net.rn, net.q, net.sums, energy);
    loss_grad_kernel<<<grid, 256, 0, s>>>(net.q, flags, dims[0], dims[1], dims[2], batch, net.sums, net.G[kGOut], energy);
    hip_check(hipGetLastError(), "preconditioner loss");
    hip_check(hipMemsetAsync(net.grads, 0, size_t(net.nparams) * sizeof(float), s), "zero gradients");
    backward(net, batch, d, s);
    std::vector<double> sums(size_t(3 * batch));
    hip_check(hipMemcpyAsync(sums.data(), net.sums, sums.size() * sizeof(double), hipMemcpyDeviceToHost, s), "loss sums");
    if (host_grads) hip_check(hipMemcpyAsync(host_grads, net.grads, size_t(net.nparams) * sizeof(float), hipMemcpyDeviceToHost, s), "gradients");
    if (lr > 0.0f) {
        net.adam_t += 1;
        const float bc1 = 1.0f - std::pow(0.9f, float(net.adam_t)), bc2 = 1.0f - std::pow(0.999f, float(net.adam_t));
        adam_kernel<<<unsigned((net.nparams + 255) / 256), 256, 0, s>>>(net.params, net.grads, net.adam_m, net.adam_v, net.nparams, lr, bc1, bc2);
        hip_check(hipGetLastError(), "adam");
        net.packed_dirty = true;
    }
    hip_check(hipStreamSynchronize(s), "train step");
    double total = 0.0;
    // energy: sums[2] = z^T A z - 2 z^T rn, half of it is the loss
    const double f = energy ? 0.5 : 1.0;
    for (int b = 0; b < batch; ++b) total += sums[3 * b + 1] > 0.0 ? f * sums[3 * b + 2] / sums[3 * b + 1] : 0.0;
    if (loss) *loss = total / batch;
}

}  // namespace smoke

extern "C" {

#define SMOKE_NN_GUARD(body)                            \
    try {                                               \
        body;                                           \
        return SMOKE_SIMULATION_RESULT_OK;              \
    } catch (const std::bad_alloc&) {                   \
        return SMOKE_SIMULATION_RESULT_OUT_OF_MEMORY;   \
    } catch (...) {                                     \
        return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE; \
    }

SmokeSimulationResult smoke_nn_preconditioner_create(uint64_t seed, SmokeNnPreconditioner** out) {
    if (!out) return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
    SMOKE_NN_GUARD(*out = smoke::create(seed))
}

SmokeSimulationResult smoke_nn_preconditioner_destroy(SmokeNnPreconditioner* net) {
    SMOKE_NN_GUARD(smoke::destroy(net))
}

int64_t smoke_nn_preconditioner_parameter_count(const SmokeNnPreconditioner* net) { return net ? net->nparams : 0; }

SmokeSimulationResult smoke_nn_preconditioner_get_parameters(SmokeNnPreconditioner* net, float* out) {
    if (!net || !out) return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
    SMOKE_NN_GUARD(hip_check(hipMemcpy(out, net->params, size_t(net->nparams) * sizeof(float), hipMemcpyDeviceToHost), "get parameters"))
}

SmokeSimulationResult smoke_nn_preconditioner_set_parameters(SmokeNnPreconditioner* net, const float* values) {
    if (!net || !values) return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
    SMOKE_NN_GUARD(hip_check(hipMemcpy(net->params, values, size_t(net->nparams) * sizeof(float), hipMemcpyHostToDevice), "set parameters");
                   net->packed_dirty = true)
}

SmokeSimulationResult smoke_nn_preconditioner_save(SmokeNnPreconditioner* net, const char* path) {
    if (!net || !path) return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
    SMOKE_NN_GUARD({
        std::vector<float> h(size_t(net->nparams));
        hip_check(hipMemcpy(h.data(), net->params, h.size() * sizeof(float), hipMemcpyDeviceToHost), "save parameters");
        FILE* f = std::fopen(path, "wb");
        if (!f) throw HipError("cannot open weights file");
        const uint32_t version = 1;
        const uint64_t count = uint64_t(net->nparams);
        bool ok = std::fwrite("SMOKENN1", 1, 8, f) == 8 && std::fwrite(&version, 4, 1, f) == 1 && std::fwrite(&count, 8, 1, f) == 1 &&
                  std::fwrite(h.data(), sizeof(float), h.size(), f) == h.size();
        ok = std::fclose(f) == 0 && ok;
        if (!ok) throw HipError("weights write failed");
    })
}

SmokeSimulationResult smoke_nn_preconditioner_load(SmokeNnPreconditioner* net, const char* path) {
    if (!net || !path) return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
    SMOKE_NN_GUARD({
        FILE* f = std::fopen(path, "rb");
        if (!f) throw HipError("cannot open weights file");
        char magic[8];
        uint32_t version = 0;
        uint64_t count = 0;
        std::vector<float> h(size_t(net->nparams));
        bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, "SMOKENN1", 8) == 0 && std::fread(&version, 4, 1, f) == 1 && version == 1 &&
                  std::fread(&count, 8, 1, f) == 1 && count == uint64_t(net->nparams) &&
                  std::fread(h.data(), sizeof(float), h.size(), f) == h.size();
        std::fclose(f);
        if (!ok) throw HipError("weights read failed");
        hip_check(hipMemcpy(net->params, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice), "load parameters");
        net->packed_dirty = true;
    })
}

SmokeSimulationResult smoke_nn_preconditioner_train_step(SmokeNnPreconditioner* net, const float* r, const uint8_t* flags, int32_t batch,
                                                         const int32_t dims[3], float lr, double* loss, float* host_grads, void* stream) {
    if (!net || !r || !flags || batch <= 0 || !dims || dims[0] < 4 || dims[1] < 4 || dims[2] < 4) return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
    SMOKE_NN_GUARD(smoke::preconditioner_train_step(*net, r, flags, batch, dims, lr, loss, host_grads, static_cast<hipStream_t>(stream)))
}

SmokeSimulationResult smoke_nn_preconditioner_set_loss(SmokeNnPreconditioner* net, uint32_t loss) {
    if (!net || loss > 1u) return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
    net->loss = int(loss);
    return SMOKE_SIMULATION_RESULT_OK;
}

SmokeSimulationResult smoke_nn_preconditioner_apply(SmokeNnPreconditioner* net, const float* r, const uint8_t* flags, const int32_t dims[3], float* z,
                                                    void* stream) {
    if (!net || !r || !flags || !z || !dims || dims[0] < 4 || dims[1] < 4 || dims[2] < 4) return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
    SMOKE_NN_GUARD(smoke::preconditioner_apply(*net, r, flags, dims, z, static_cast<hipStream_t>(stream)))
}

SmokeSimulationResult smoke_nn_preconditioner_apply_static(SmokeNnPreconditioner* net, const float* r, const uint8_t* flags, const int32_t dims[3],
                                                           float* z, void* stream) {
    if (!net || !r || !flags || !z || !dims || dims[0] < 4 || dims[1] < 4 || dims[2] < 4) return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
    if (smoke::static_layout(dims).total > kStaticWorkspaceBytes) {
        std::fprintf(stderr, "smoke_nn_preconditioner_apply_static: a %dx%dx%d volume needs more than the %zu-byte static workspace\n", dims[0],
                     dims[1], dims[2], kStaticWorkspaceBytes);
        return SMOKE_SIMULATION_RESULT_BACKEND_FAILURE;
    }
    SMOKE_NN_GUARD(smoke::preconditioner_apply_static(*net, r, flags, dims, z, static_cast<hipStream_t>(stream)))
}

}  // extern "C"
