// Training-pair crops of the pressure solve (right-hand side, solution and
// operator flags), packed to bfloat16 on the device and appended to a file.
// The record layout is documented with smoke_simulation_ext_write_training_crops
// in visual-simulation-of-smoke-3d-ext.h.
#include "context.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

namespace smoke {
namespace {

constexpr char kMagic[8] = {'S', 'M', 'O', 'K', 'E', 'T', 'D', '1'};
constexpr uint32_t kVersion = 1;
constexpr int kChunk = 4096;   // crop cells per block of the RMS pass

struct FileHeader {
    char magic[8];
    uint32_t version;
    uint32_t crop[3];
    uint32_t grid[3];
    uint32_t record_bytes;
    uint32_t reserved[6];
};
static_assert(sizeof(FileHeader) == 64, "header layout");

struct RecordHeader {
    int32_t origin[3];
    uint32_t step;
    double scale;
    uint32_t reserved[2];
};
static_assert(sizeof(RecordHeader) == 32, "record layout");

struct CropArgs {
    int n[3];       // grid
    int c[3];       // crop extent
    int64_t cells;  // cells per crop
};

// Grid index of cell q (x fastest) of crop `crop`.
__device__ __forceinline__ int64_t crop_cell(const CropArgs& a, const int* __restrict__ origins, int crop, int64_t q) {
    const int x = int(q % a.c[0]);
    const int y = int((q / a.c[0]) % a.c[1]);
    const int z = int(q / (int64_t(a.c[0]) * a.c[1]));
    const int* o = origins + 3 * crop;
    return (int64_t(o[2] + z) * a.n[1] + (o[1] + y)) * a.n[0] + (o[0] + x);
}

// Per-block partial sums of b^2 over kChunk cells of one crop (blockIdx.y).
__global__ __launch_bounds__(256) void crop_rr_kernel(CropArgs a, const int* __restrict__ origins, const float* __restrict__ b,
                                                      double* __restrict__ partials) {
    const int crop = blockIdx.y;
    const int64_t q0 = int64_t(blockIdx.x) * kChunk;
    const int64_t q1 = min(a.cells, q0 + kChunk);
    double s = 0.0;
    for (int64_t q = q0 + threadIdx.x; q < q1; q += 256) {
        const double v = b[crop_cell(a, origins, crop, q)];
        s += v * v;
    }
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    __shared__ double red[4];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) partials[int64_t(crop) * gridDim.x + blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

// scale = RMS of b over the crop (1 when b is zero there), summed in block order.
__global__ void crop_scale_kernel(const double* __restrict__ partials, int blocks, int64_t cells, double* __restrict__ scales) {
    if (threadIdx.x != 0) return;
    const int crop = blockIdx.x;
    double s = 0.0;
    for (int i = 0; i < blocks; ++i) s += partials[int64_t(crop) * blocks + i];
    const double rms = sqrt(s / double(cells));
    scales[crop] = rms > 0.0 ? rms : 1.0;
}

// bfloat16 of v, rounded to nearest even; NaN stays a quiet NaN.
__device__ __forceinline__ uint16_t to_bf16(float v) {
    uint32_t u = __float_as_uint(v);
    if ((u & 0x7fffffffu) > 0x7f800000u) return uint16_t((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return uint16_t(u >> 16);
}

template <typename Z>
__global__ __launch_bounds__(256) void crop_pack_kernel(CropArgs a, const int* __restrict__ origins, const float* __restrict__ b,
                                                        const Z* __restrict__ x, const uint8_t* __restrict__ flags,
                                                        const double* __restrict__ scales, uint16_t* __restrict__ r_out,
                                                        uint16_t* __restrict__ z_out, uint8_t* __restrict__ f_out) {
    const int crop = blockIdx.y;
    const int64_t q = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (q >= a.cells) return;
    const int64_t i = crop_cell(a, origins, crop, q);
    const double scale = scales[crop];
    const int64_t o = int64_t(crop) * a.cells + q;
    r_out[o] = to_bf16(float(double(b[i]) / scale));
    z_out[o] = to_bf16(float(double(x[i]) / scale));
    f_out[o] = flags[i];
}

__global__ __launch_bounds__(256) void crop_copy_kernel(CropArgs a, const int* __restrict__ origins, const float* __restrict__ b,
                                                        const uint8_t* __restrict__ flags, const double* __restrict__ scales,
                                                        float* __restrict__ r_out, uint8_t* __restrict__ f_out) {
    const int crop = blockIdx.y;
    const int64_t q = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (q >= a.cells) return;
    const int64_t i = crop_cell(a, origins, crop, q);
    const int64_t o = int64_t(crop) * a.cells + q;
    r_out[o] = float(double(b[i]) / scales[crop]);
    f_out[o] = flags[i];
}

// Device buffer freed on scope exit (this call is offline logging, so it may
// allocate).
template <typename T>
struct DeviceBuffer {
    T* p = nullptr;
    explicit DeviceBuffer(size_t n) { hip_check(hipMalloc(&p, n * sizeof(T)), "training crops buffer"); }
    ~DeviceBuffer() {
        if (p) (void)hipFree(p);
    }
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
};

struct File {
    FILE* f = nullptr;
    File(const char* path, const char* mode) : f(std::fopen(path, mode)) {}
    ~File() {
        if (f) std::fclose(f);
    }
};

}  // namespace

namespace {

CropArgs crop_args(const Context& c, const int32_t* origins, uint32_t count, const uint32_t size[3]) {
    if (!origins || count == 0) throw HipError("training crops: missing origins");
    CropArgs a{};
    for (int k = 0; k < 3; ++k) {
        a.n[k] = c.grid.n[k];
        a.c[k] = int(size[k]);
        if (size[k] == 0 || size[k] > uint32_t(c.grid.n[k])) throw HipError("training crops: bad crop size");
    }
    for (uint32_t i = 0; i < count; ++i)
        for (int k = 0; k < 3; ++k) {
            const int o = origins[3 * i + k];
            if (o < 0 || int64_t(o) + a.c[k] > a.n[k]) throw HipError("training crops: crop outside the grid");
        }
    a.cells = int64_t(a.c[0]) * a.c[1] * a.c[2];
    return a;
}

}  // namespace

void copy_training_crops(Context& c, const SmokeSimulationDeviceCrops& d) {
    if (!d.r || !d.flags) throw HipError("training crops: missing output buffers");
    const CropArgs a = crop_args(c, d.origins, d.count, d.size);
    const int rr_blocks = int((a.cells + kChunk - 1) / kChunk);
    DeviceBuffer<int> origins(3 * size_t(d.count));
    DeviceBuffer<double> partials(size_t(rr_blocks) * d.count);
    DeviceBuffer<double> scales(d.count);
    hip_check(hipMemcpyAsync(origins.p, d.origins, 3 * size_t(d.count) * sizeof(int), hipMemcpyHostToDevice, c.stream), "crop origins");
    crop_rr_kernel<<<dim3(unsigned(rr_blocks), d.count), 256, 0, c.stream>>>(a, origins.p, c.pressure_rhs, partials.p);
    crop_scale_kernel<<<d.count, 64, 0, c.stream>>>(partials.p, rr_blocks, a.cells, scales.p);
    crop_copy_kernel<<<dim3(unsigned((a.cells + 255) / 256), d.count), 256, 0, c.stream>>>(a, origins.p, c.pressure_rhs, c.stencil_flags,
                                                                                          scales.p, d.r, d.flags);
    hip_check(hipGetLastError(), "training crop kernels");
    // the scratch buffers are freed on return: wait for the kernels that read them
    hip_check(hipStreamSynchronize(c.stream), "training crops sync");
}

void write_training_crops(Context& c, const SmokeSimulationTrainingCrops& d) {
    if (!d.path) throw HipError("training crops: missing path");
    const CropArgs a = crop_args(c, d.origins, d.count, d.size);
    const size_t n = size_t(a.cells) * d.count;
    const int rr_blocks = int((a.cells + kChunk - 1) / kChunk);

    DeviceBuffer<int> origins(3 * size_t(d.count));
    DeviceBuffer<double> partials(size_t(rr_blocks) * d.count);
    DeviceBuffer<double> scales(d.count);
    DeviceBuffer<uint16_t> r_out(n), z_out(n);
    DeviceBuffer<uint8_t> f_out(n);
    hip_check(hipMemcpyAsync(origins.p, d.origins, 3 * size_t(d.count) * sizeof(int), hipMemcpyHostToDevice, c.stream), "crop origins");
    crop_rr_kernel<<<dim3(unsigned(rr_blocks), d.count), 256, 0, c.stream>>>(a, origins.p, c.pressure_rhs, partials.p);
    crop_scale_kernel<<<d.count, 64, 0, c.stream>>>(partials.p, rr_blocks, a.cells, scales.p);
    const dim3 grid(unsigned((a.cells + 255) / 256), d.count);
    if (c.options.pressure_mode == SMOKE_SIMULATION_PRESSURE_TOLERANCE)
        crop_pack_kernel<double><<<grid, 256, 0, c.stream>>>(a, origins.p, c.pressure_rhs, c.x64[c.x64_current], c.stencil_flags, scales.p,
                                                             r_out.p, z_out.p, f_out.p);
    else
        crop_pack_kernel<float><<<grid, 256, 0, c.stream>>>(a, origins.p, c.pressure_rhs, c.pressure, c.stencil_flags, scales.p, r_out.p,
                                                            z_out.p, f_out.p);
    hip_check(hipGetLastError(), "training crop kernels");
    std::vector<uint16_t> r(n), z(n);
    std::vector<uint8_t> fl(n);
    std::vector<double> sc(d.count);
    hip_check(hipMemcpyAsync(r.data(), r_out.p, n * sizeof(uint16_t), hipMemcpyDeviceToHost, c.stream), "crop r");
    hip_check(hipMemcpyAsync(z.data(), z_out.p, n * sizeof(uint16_t), hipMemcpyDeviceToHost, c.stream), "crop z");
    hip_check(hipMemcpyAsync(fl.data(), f_out.p, n, hipMemcpyDeviceToHost, c.stream), "crop flags");
    hip_check(hipMemcpyAsync(sc.data(), scales.p, d.count * sizeof(double), hipMemcpyDeviceToHost, c.stream), "crop scales");
    hip_check(hipStreamSynchronize(c.stream), "training crops sync");

    const uint32_t record_bytes = uint32_t(sizeof(RecordHeader) + size_t(a.cells) * 5);
    FileHeader h{};
    std::memcpy(h.magic, kMagic, sizeof(kMagic));
    h.version = kVersion;
    for (int k = 0; k < 3; ++k) {
        h.crop[k] = d.size[k];
        h.grid[k] = uint32_t(c.grid.n[k]);
    }
    h.record_bytes = record_bytes;

    long offset = 0;
    {
        File existing(d.path, "rb");
        if (existing.f) {
            FileHeader e{};
            const size_t got = std::fread(&e, 1, sizeof(e), existing.f);
            if (got != 0) {
                if (got != sizeof(e) || std::memcmp(e.magic, kMagic, sizeof(kMagic)) != 0 || e.version != kVersion ||
                    std::memcmp(e.crop, h.crop, sizeof(h.crop)) != 0 || std::memcmp(e.grid, h.grid, sizeof(h.grid)) != 0)
                    throw HipError("training crops: existing file has another layout");
            }
        }
    }
    File out(d.path, "ab");
    if (!out.f) throw HipError("training crops: cannot open file");
    std::fseek(out.f, 0, SEEK_END);
    offset = std::ftell(out.f);
    if (offset == 0) {
        if (std::fwrite(&h, sizeof(h), 1, out.f) != 1) throw HipError("training crops: write failed");
        offset = long(sizeof(h));
    }
    const std::string manifest_path = std::string(d.path) + ".jsonl";
    File manifest(manifest_path.c_str(), "a");
    if (!manifest.f) throw HipError("training crops: cannot open manifest");
    for (uint32_t i = 0; i < d.count; ++i) {
        RecordHeader rh{};
        for (int k = 0; k < 3; ++k) rh.origin[k] = d.origins[3 * i + k];
        rh.step = d.step;
        rh.scale = sc[i];
        const size_t base = size_t(i) * size_t(a.cells);
        bool ok = std::fwrite(&rh, sizeof(rh), 1, out.f) == 1;
        ok = ok && std::fwrite(r.data() + base, sizeof(uint16_t), size_t(a.cells), out.f) == size_t(a.cells);
        ok = ok && std::fwrite(z.data() + base, sizeof(uint16_t), size_t(a.cells), out.f) == size_t(a.cells);
        ok = ok && std::fwrite(fl.data() + base, 1, size_t(a.cells), out.f) == size_t(a.cells);
        if (!ok) throw HipError("training crops: write failed");
        std::fprintf(manifest.f,
                     "{\"offset\": %ld, \"origin\": [%d, %d, %d], \"step\": %u, \"scale\": %.17g, \"crop\": [%u, %u, %u], "
                     "\"grid\": [%d, %d, %d], \"z_source\": \"%s\"}\n",
                     offset, rh.origin[0], rh.origin[1], rh.origin[2], rh.step, rh.scale, d.size[0], d.size[1], d.size[2], c.grid.n[0],
                     c.grid.n[1], c.grid.n[2], c.options.pressure_mode == SMOKE_SIMULATION_PRESSURE_TOLERANCE ? "fp64_solution" : "fp32_pressure");
        offset += long(record_bytes);
    }
    if (std::fflush(out.f) != 0 || std::fflush(manifest.f) != 0) throw HipError("training crops: flush failed");
}

}  // namespace smoke
