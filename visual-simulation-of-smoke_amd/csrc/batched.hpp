// One launch over many grids ("plumes" of a batch).  A kernel body takes its
// block index as an argument (VBlock) instead of reading blockIdx / gridDim;
// the ordinary launch passes the hardware's, and batched_kernel passes each
// grid's own numbering: the blocks of a batched launch are the concatenation
// of every grid's launch, a block finds its grid by binary search over the
// grids' first blocks, and runs the body with that grid's arguments and its
// block index within that grid's launch.  Every grid therefore computes
// exactly what its own launch would, bit for bit.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstring>
#include <utility>
#include <vector>

#include "device_common.hpp"

namespace smoke {

// A flat, trivially copyable argument pack (filled on the host, read by the
// device with the same layout).
template <size_t I, class T>
struct PackLeaf {
    T v;
};
template <class Seq, class... T>
struct PackBase;
template <size_t... I, class... T>
struct PackBase<std::index_sequence<I...>, T...> : PackLeaf<I, T>... {};
template <class... T>
struct Pack : PackBase<std::index_sequence_for<T...>, T...> {};

template <size_t I, class T>
__host__ __device__ __forceinline__ const T& leaf(const PackLeaf<I, T>& l) {
    return l.v;
}
template <size_t I, class T>
__host__ __device__ __forceinline__ T& leaf(PackLeaf<I, T>& l) {
    return l.v;
}

// One grid of a batched launch.
template <class... T>
struct Slot {
    Pack<T...> args;
    unsigned first;   // first block of this grid in the batched launch
    unsigned plume;   // index into the launch's gate array
    uint3 grid;       // blocks of the grid's own launch
};

template <class K, class... T, size_t... I>
__device__ __forceinline__ void run_slot(const Slot<T...>& s, const VBlock& vb, std::index_sequence<I...>) {
    K::run(vb, leaf<I>(s.args)...);
}

// K::run(VBlock, T...) over every grid of the table; grids whose gate entry
// is 0 are skipped (gate may be null: every grid runs).  The table starts
// with the grids' first blocks (`count` increasing uint32, padded to 256
// bytes), then the slots.  Each wave finds its block's grid on its own: lane
// l compares first[l] with the block index and the count of set ballot bits
// (the first blocks are increasing, so they form a prefix) is the grid, in
// one vector load per 64 grids; every wave reaches the same answer, so no
// barrier is needed.  Blocks of fewer than 64 threads bisect instead.
template <class K, int THREADS, class... T>
__global__ __launch_bounds__(THREADS) void batched_kernel(const unsigned* __restrict__ firsts, const Slot<T...>* __restrict__ t, int count,
                                                          const int* __restrict__ gate) {
    const unsigned b = blockIdx.x;
    int below = 0;
    if (blockDim.x * blockDim.y * blockDim.z >= 64) {
        const int lane = int(__lane_id());
        for (int base = 0; base < count; base += 64) {
            const int i = base + lane;
            const bool le = i < count && firsts[i] <= b;
            const int n = __popcll(__ballot(le));
            below += n;
            if (n < 64) break;
        }
    } else {
        // a block narrower than a wave (single-thread launches): bisection
        int lo = 0, hi = count - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (firsts[mid] <= b) lo = mid;
            else hi = mid - 1;
        }
        below = lo + 1;
    }
    const Slot<T...>& s = t[below - 1];
    if (gate && gate[s.plume] == 0) return;
    const unsigned local = b - s.first;
    VBlock vb;
    vb.dim = s.grid;
    vb.idx.x = local % s.grid.x;
    const unsigned rest = local / s.grid.x;
    vb.idx.y = rest % s.grid.y;
    vb.idx.z = rest / s.grid.y;
    run_slot<K>(s, vb, std::index_sequence_for<T...>{});
}

// Host side: the tables of a batch's launches live in one device buffer (an
// "arena"); a BatchedLaunch names its table by offset.
struct TableArena {
    std::vector<unsigned char> image;
    unsigned char* dev = nullptr;
    size_t capacity = 0;

    size_t add(const void* data, size_t bytes) {
        const size_t at = (image.size() + 255) / 256 * 256;
        image.resize(at + bytes);
        std::memcpy(image.data() + at, data, bytes);
        return at;
    }
    void clear() { image.clear(); }
};

struct BatchedLaunch {
    void (*fn)(const void* table, int count, unsigned blocks, dim3 block, const int* gate, hipStream_t s) = nullptr;
    size_t offset = 0;   // of the table in its arena
    int count = 0;       // grids
    unsigned blocks = 0; // total blocks
    dim3 block;
};

inline size_t firsts_bytes(int count) { return (size_t(count) * sizeof(unsigned) + 255) / 256 * 256; }

template <class K, int THREADS, class... T>
void launch_table(const void* table, int count, unsigned blocks, dim3 block, const int* gate, hipStream_t s) {
    const unsigned char* base = static_cast<const unsigned char*>(table);
    batched_kernel<K, THREADS, T...><<<blocks, block, 0, s>>>(reinterpret_cast<const unsigned*>(base), reinterpret_cast<const Slot<T...>*>(base + firsts_bytes(count)),
                                                            count, gate);
}

// Collects the grids of one batched launch of kernel K.
template <class K, int THREADS, class... T>
struct LaunchBuilder {
    std::vector<Slot<T...>> slots;
    unsigned blocks = 0;

    void add(unsigned plume, dim3 grid, const T&... args) {
        Slot<T...> s{};
        set(s.args, std::index_sequence_for<T...>{}, args...);
        s.first = blocks;
        s.plume = plume;
        s.grid = make_uint3(grid.x, grid.y, grid.z);
        const unsigned n = grid.x * grid.y * grid.z;
        if (n == 0) return;
        slots.push_back(s);
        blocks += n;
    }
    bool empty() const { return slots.empty(); }
    BatchedLaunch finish(TableArena& arena, dim3 block) const {
        BatchedLaunch L;
        L.fn = &launch_table<K, THREADS, T...>;
        const int n = int(slots.size());
        std::vector<unsigned char> image(firsts_bytes(n) + slots.size() * sizeof(Slot<T...>), 0);
        for (int i = 0; i < n; ++i) std::memcpy(image.data() + i * sizeof(unsigned), &slots[size_t(i)].first, sizeof(unsigned));
        std::memcpy(image.data() + firsts_bytes(n), slots.data(), slots.size() * sizeof(Slot<T...>));
        L.offset = arena.add(image.data(), image.size());
        L.count = int(slots.size());
        L.blocks = blocks;
        L.block = block;
        return L;
    }

   private:
    template <size_t... I>
    static void set(Pack<T...>& p, std::index_sequence<I...>, const T&... v) {
        ((leaf<I>(p) = v), ...);
    }
};

inline void run_launch(const BatchedLaunch& L, const unsigned char* arena_dev, const int* gate, hipStream_t s) {
    if (L.count > 0 && L.blocks > 0) L.fn(arena_dev + L.offset, L.count, L.blocks, L.block, gate, s);
}

}  // namespace smoke
