// prim.hip — fill, scan and radix sort for gfx950 (64-lane wavefronts).
#include <hip/hip_runtime.h>

#include "prim.hpp"

namespace dio {

namespace {

constexpr int kWave = 64;
constexpr int kScanBlock = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanBlock * kScanItems;  // 2048 elements per block

__global__ void __launch_bounds__(256) k_fill_u32(uint32_t* p, uint32_t value, uint32_t n) {
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) p[i] = value;
}

// inclusive wave scan of one value per lane
__device__ __forceinline__ uint32_t wave_inclusive(uint32_t x) {
    const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        uint32_t y = __shfl_up(x, off, kWave);
        if (lane >= off) x += y;
    }
    return x;
}

// Block-wide exclusive scan of one value per thread; returns the exclusive
// prefix and writes the block total to *total.
template <int NT>
__device__ __forceinline__ uint32_t block_exclusive(uint32_t x, uint32_t* lds_waves, uint32_t* total) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    constexpr int kWaves = NT / kWave;
    uint32_t inc = wave_inclusive(x);
    if (lane == kWave - 1) lds_waves[wave] = inc;
    __syncthreads();
    // wave 0 scans the wave totals (each lane reads its slot before any lane
    // of the wave writes)
    if (wave == 0) {
        const uint32_t t = lane < kWaves ? lds_waves[lane] : 0u;
        const uint32_t ti = wave_inclusive(t);
        if (lane < kWaves) lds_waves[lane] = ti - t;
        if (lane == kWaves - 1) lds_waves[kWaves] = ti;
    }
    __syncthreads();
    uint32_t r = lds_waves[wave] + inc - x;
    *total = lds_waves[kWaves];
    __syncthreads();
    return r;
}

// Eight consecutive words from base: two 16-byte loads when all are in range
// and the array is 16-byte aligned (base is a multiple of 8), else guarded
// scalar loads (zero past n).
__device__ __forceinline__ void load8(const uint32_t* in, uint32_t base, uint32_t n, bool aligned, uint32_t v[8]) {
    if (aligned && base + 8 <= n) {
        const uint4 a = *reinterpret_cast<const uint4*>(in + base);
        const uint4 b = *reinterpret_cast<const uint4*>(in + base + 4);
        v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
        v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
    } else {
#pragma unroll
        for (int k = 0; k < 8; ++k) v[k] = base + k < n ? in[base + k] : 0u;
    }
}
// out[base + k] = ex + v[0] + ... + v[k-1] for the k < 8 with base + k < n
__device__ __forceinline__ void store8_exclusive(uint32_t* out, uint32_t base, uint32_t n, bool aligned, uint32_t ex,
                                                 const uint32_t v[8]) {
    uint32_t e[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        e[k] = ex;
        ex += v[k];
    }
    if (aligned && base + 8 <= n) {
        *reinterpret_cast<uint4*>(out + base) = make_uint4(e[0], e[1], e[2], e[3]);
        *reinterpret_cast<uint4*>(out + base + 4) = make_uint4(e[4], e[5], e[6], e[7]);
    } else {
#pragma unroll
        for (int k = 0; k < 8; ++k)
            if (base + k < n) out[base + k] = e[k];
    }
}
__device__ __forceinline__ bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0u; }

__device__ __forceinline__ uint32_t read_n(uint32_t n_cap, const uint32_t* n_dev) {
    if (!n_dev) return n_cap;
    uint32_t n = *n_dev;
    return n < n_cap ? n : n_cap;
}

// pass 1: per-tile sums
__global__ void __launch_bounds__(kScanBlock) k_scan_partials(const uint32_t* in, uint32_t n_cap,
                                                              const uint32_t* n_dev, uint32_t* partials) {
    __shared__ uint32_t lds[kScanBlock / kWave + 1];
    const uint32_t n = read_n(n_cap, n_dev);
    const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    load8(in, base, n, aligned16(in), v);
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) s += v[k];
    uint32_t total;
    block_exclusive<kScanBlock>(s, lds, &total);
    if (threadIdx.x == 0) partials[blockIdx.x] = total;
}

// pass 2: scan each tile and add its offset, the sum of the tile sums before
// it, which each workgroup adds up itself (integer sums: any order gives the
// same result); workgroup 0 also writes the total.  Two launches in all.
__global__ void __launch_bounds__(kScanBlock) k_scan_apply(const uint32_t* in, uint32_t* out, uint32_t n_cap,
                                                           const uint32_t* n_dev, const uint32_t* partials,
                                                           uint32_t tiles) {
    __shared__ uint32_t lds[kScanBlock / kWave + 1];
    const uint32_t n = read_n(n_cap, n_dev);
    const uint32_t used = (n + kScanTile - 1) / kScanTile;  // tiles holding elements (<= tiles)
    const uint32_t upto = blockIdx.x == 0 ? used : (blockIdx.x < used ? blockIdx.x : used);
    uint32_t part = 0;
    for (uint32_t t = threadIdx.x; t < upto; t += kScanBlock) part += partials[t];
    uint32_t before;
    block_exclusive<kScanBlock>(part, lds, &before);  // the workgroup's sum of tile sums
    const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    const bool al = aligned16(in) && aligned16(out);
    uint32_t v[kScanItems];
    load8(in, base, n, al, v);
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) s += v[k];
    uint32_t total;
    const uint32_t ex = block_exclusive<kScanBlock>(s, lds, &total) + (blockIdx.x == 0 ? 0u : before);
    store8_exclusive(out, base, n, al, ex, v);
    if (blockIdx.x == 0 && threadIdx.x == 0) out[n] = before;  // workgroup 0 added every used tile
}

// One workgroup scans up to kSingleMax elements in tiles of kSingleTile with a
// running carry: one launch instead of three for the per-body arrays.
constexpr int kSingleThreads = 1024;
constexpr int kSingleItems = 8;
static_assert(kScanItems == 8 && kSingleItems == 8, "load8 / store8_exclusive move eight words per thread");
constexpr uint32_t kSingleTile = kSingleThreads * kSingleItems;
constexpr uint32_t kSingleMax = 4 * kSingleTile;

__global__ void __launch_bounds__(kSingleThreads) k_scan_single(const uint32_t* in, uint32_t* out, uint32_t n_cap,
                                                                const uint32_t* n_dev) {
    __shared__ uint32_t lds[kSingleThreads / kWave + 1];
    const uint32_t n = read_n(n_cap, n_dev);
    const bool al = aligned16(in) && aligned16(out);
    uint32_t carry = 0;
    for (uint32_t tile = 0; tile < n; tile += kSingleTile) {
        const uint32_t base = tile + threadIdx.x * kSingleItems;
        uint32_t v[kSingleItems];
        load8(in, base, n, al, v);
        uint32_t sum = 0;
#pragma unroll
        for (int k = 0; k < kSingleItems; ++k) sum += v[k];
        uint32_t total;
        const uint32_t ex = block_exclusive<kSingleThreads>(sum, lds, &total) + carry;
        store8_exclusive(out, base, n, al, ex, v);
        carry += total;
    }
    if (threadIdx.x == 0) out[n] = carry;
}

__global__ void __launch_bounds__(256) k_fill2_u32(uint32_t* a, uint32_t* b, uint32_t value, uint32_t n) {
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        a[i] = value;
        b[i] = value;
    }
}

// ------------------------------------------------------------ radix sort
constexpr int kSortBlock = 256;
constexpr int kSortItems = 8;
constexpr int kSortTile = kSortBlock * kSortItems;  // striped: element k*256 + t of a tile
constexpr int kRadix = 256;

__global__ void __launch_bounds__(kSortBlock) k_radix_hist(const uint32_t* keys, uint32_t n, int shift,
                                                           uint32_t* hist, uint32_t tiles) {
    __shared__ uint32_t h[kRadix];
    h[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * kSortTile;
#pragma unroll
    for (int k = 0; k < kSortItems; ++k) {
        const uint32_t i = base + k * kSortBlock + threadIdx.x;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & (kRadix - 1)], 1u);
    }
    __syncthreads();
    hist[threadIdx.x * tiles + blockIdx.x] = h[threadIdx.x];
}

// Stable scatter: ranks within a tile follow element order (round k, then
// lane), via 8 ballots per round that group the lanes of a wave by digit.
__global__ void __launch_bounds__(kSortBlock) k_radix_scatter(const uint32_t* keys_in, const uint32_t* vals_in,
                                                              uint32_t* keys_out, uint32_t* vals_out, uint32_t n,
                                                              int shift, const uint32_t* hist_scan,
                                                              uint32_t tiles) {
    constexpr int kWaves = kSortBlock / kWave;
    __shared__ uint32_t run[kRadix];
    __shared__ uint32_t wcnt[kWaves][kRadix];
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    const uint64_t lt_mask = (lane == 0) ? 0ull : ((~0ull) >> (kWave - lane));
    run[threadIdx.x] = hist_scan[threadIdx.x * tiles + blockIdx.x];
    const uint32_t base = blockIdx.x * kSortTile;
    for (int k = 0; k < kSortItems; ++k) {
#pragma unroll
        for (int w = 0; w < kWaves; ++w) wcnt[w][threadIdx.x] = 0;
        __syncthreads();
        const uint32_t i = base + k * kSortBlock + threadIdx.x;
        const bool valid = i < n;
        const uint32_t key = valid ? keys_in[i] : 0u;
        const uint32_t val = valid ? vals_in[i] : 0u;
        const uint32_t digit = (key >> shift) & (kRadix - 1);
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (digit >> b) & 1u;
            const uint64_t bal = __ballot(bit);
            peers &= bit ? bal : ~bal;
        }
        const uint32_t rank = __popcll(peers & lt_mask);
        if (valid && rank == 0) wcnt[wave][digit] = __popcll(peers);
        __syncthreads();
        if (valid) {
            uint32_t pos = run[digit] + rank;
            for (int w = 0; w < wave; ++w) pos += wcnt[w][digit];
            keys_out[pos] = key;
            vals_out[pos] = val;
        }
        __syncthreads();
        uint32_t add = 0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) add += wcnt[w][threadIdx.x];
        run[threadIdx.x] += add;
        __syncthreads();
    }
}

inline uint32_t ceil_div(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

}  // namespace

void launch_fill_u32(uint32_t* p, uint32_t value, uint32_t n, hipStream_t s) {
    if (n == 0) return;
    uint32_t blocks = ceil_div(n, 256);
    if (blocks > 2048) blocks = 2048;
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks), dim3(256), 0, s, p, value, n);
}

size_t scan_partials_size(uint32_t n_cap) { return ceil_div(n_cap > 0 ? n_cap : 1, kScanTile) + 1; }

void launch_fill2_u32(uint32_t* a, uint32_t* b, uint32_t value, uint32_t n, hipStream_t s) {
    if (n == 0) return;
    uint32_t blocks = ceil_div(n, 256);
    if (blocks > 2048) blocks = 2048;
    hipLaunchKernelGGL(k_fill2_u32, dim3(blocks), dim3(256), 0, s, a, b, value, n);
}

void launch_scan_exclusive(const uint32_t* in, uint32_t* out, uint32_t n_cap, const uint32_t* n_dev,
                           uint32_t* partials, hipStream_t s) {
    if (n_cap <= kSingleMax) {
        hipLaunchKernelGGL(k_scan_single, dim3(1), dim3(kSingleThreads), 0, s, in, out, n_cap, n_dev);
        return;
    }
    const uint32_t tiles = ceil_div(n_cap > 0 ? n_cap : 1, kScanTile);
    hipLaunchKernelGGL(k_scan_partials, dim3(tiles), dim3(kScanBlock), 0, s, in, n_cap, n_dev, partials);
    hipLaunchKernelGGL(k_scan_apply, dim3(tiles), dim3(kScanBlock), 0, s, in, out, n_cap, n_dev, partials, tiles);
}

size_t radix_scratch_size(uint32_t n) {
    const uint32_t tiles = ceil_div(n > 0 ? n : 1, kSortTile);
    const uint32_t hn = tiles * kRadix;
    return 2 * (size_t)hn + 1 + scan_partials_size(hn);
}

int radix_sort_pairs(uint32_t* k0, uint32_t* v0, uint32_t* k1, uint32_t* v1, uint32_t n, int bits,
                     uint32_t* scratch, hipStream_t s) {
    if (n <= 1) return 0;
    const uint32_t tiles = ceil_div(n, kSortTile);
    const uint32_t hn = tiles * kRadix;
    uint32_t* hist = scratch;
    uint32_t* hscan = scratch + hn;
    uint32_t* partials = scratch + 2 * (size_t)hn + 1;
    int cur = 0;
    for (int shift = 0; shift < bits; shift += 8) {
        uint32_t* ki = cur ? k1 : k0;
        uint32_t* vi = cur ? v1 : v0;
        uint32_t* ko = cur ? k0 : k1;
        uint32_t* vo = cur ? v0 : v1;
        hipLaunchKernelGGL(k_radix_hist, dim3(tiles), dim3(kSortBlock), 0, s, ki, n, shift, hist, tiles);
        launch_scan_exclusive(hist, hscan, hn, nullptr, partials, s);
        hipLaunchKernelGGL(k_radix_scatter, dim3(tiles), dim3(kSortBlock), 0, s, ki, vi, ko, vo, n, shift, hscan,
                           tiles);
        cur ^= 1;
    }
    return cur;
}

}  // namespace dio
