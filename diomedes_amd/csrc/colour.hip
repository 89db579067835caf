// colour.hip — Jones-Plassmann colouring of the body graph and per-colour
// body lists for the coloured Gauss-Seidel primal (no counterpart in the
// reference).  One thread per body per round; a round exits at once when the
// one before left nothing uncoloured, so a fixed batch of round launches can
// run ahead of the host's check.
#include <hip/hip_runtime.h>

#include "colour.hpp"
#include "dio_math.hpp"
#include "prim.hpp"

namespace dio {

namespace {

constexpr int kBlock = 256;
inline uint32_t blocks_for(uint32_t n) { return (n + kBlock - 1) / kBlock; }

// 32-bit integer hash (C. Wellons' "lowbias32" finaliser constants)
__device__ __forceinline__ uint32_t colour_priority(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}
__device__ __forceinline__ uint32_t local_index(const ColourGraph& g, uint32_t i) {
    return g.scene_of ? i - g.scene_off[g.scene_of[i]] : i;
}
// j and i are in one scene (the graph has no edge between scenes)
__device__ __forceinline__ bool colour_beats(const ColourGraph& g, uint32_t j, uint32_t i) {
    const uint32_t lj = local_index(g, j), li = local_index(g, i);
    const uint32_t pj = colour_priority(lj), pi = colour_priority(li);
    return pj > pi || (pj == pi && lj > li);
}

__global__ void __launch_bounds__(kBlock) k_colour_init(ColourGraph g, ColourBuffers b) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < g.n) {
        const bool dyn = g.inv_mm[i].w != 0.0f;
        b.round_of[i] = dyn ? kColourUncoloured : kColourNotInGraph;
        b.colour[i] = kColourNotInGraph;
    }
    for (uint32_t k = i; k < kColourRoundSlots; k += gridDim.x * kBlock) b.pending[k] = 0u;
}

// One neighbour o of body i in round r: its colour joins the mask when it was
// coloured in an earlier round; otherwise i is beaten when o's priority wins.
__device__ __forceinline__ bool colour_visit(const ColourGraph& g, const ColourBuffers& b, uint32_t i, uint32_t o,
                                             uint32_t r, uint64_t m[4]) {
    if (o == i || g.inv_mm[o].w == 0.0f) return true;
    const uint32_t ro = b.round_of[o];
    if (ro < r) {
        const uint32_t c = b.colour[o];
        const uint64_t bit = 1ull << (c & 63u);
        const uint32_t w = c >> 6;
        m[0] |= w == 0 ? bit : 0ull;
        m[1] |= w == 1 ? bit : 0ull;
        m[2] |= w == 2 ? bit : 0ull;
        m[3] |= w == 3 ? bit : 0ull;
        return true;
    }
    return !colour_beats(g, o, i);
}

// Body i's decision in round r.
__device__ __forceinline__ void colour_round_body(const ColourGraph& g, const ColourBuffers& b, uint32_t i,
                                                  uint32_t r) {
    if (b.round_of[i] != kColourUncoloured) return;
    uint64_t m[4] = {0ull, 0ull, 0ull, 1ull << 63};  // colour 255 is never free
    bool best = true;
    for (uint32_t e = g.off[i], e1 = g.off[i + 1]; e < e1 && best; ++e) {
        const uint32_t pc = g.peer[e];
        if (pc == kContainer) continue;
        best = colour_visit(g, b, i, pc & ~kPeerIsB, r, m);
    }
    if (g.j_off)
        for (uint32_t e = g.j_off[i], e1 = g.j_off[i + 1]; e < e1 && best; ++e) {
            const uint4 abk = g.j_abk[g.j_idx[e]];
            best = colour_visit(g, b, i, abk.x == i ? abk.y : abk.x, r, m);
        }
    if (!best) {
        b.pending[r] = 1u;
        return;
    }
    uint32_t c = kMaxColours - 1;
    bool found = false;
    for (int w = 0; w < 4 && !found; ++w)
        if (~m[w] != 0ull) {
            c = 64u * (uint32_t)w + (uint32_t)__builtin_ctzll(~m[w]);
            found = true;
        }
    if (!found || c >= kMaxColours) {
        b.pending[kColourRoundSlots - 1] = 1u;  // overflow: more colours than the lists hold
        c = kMaxColours - 1;
    }
    b.colour[i] = c;
    b.round_of[i] = r;
}

__global__ void __launch_bounds__(kBlock) k_colour_round(ColourGraph g, ColourBuffers b, uint32_t r) {
    if (r > 0 && b.pending[r - 1] == 0u) return;
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < g.n) colour_round_body(g, b, i, r);
}

// Rounds first, first + 1, ... for as long as the round before left a body
// uncoloured, in one workgroup with a barrier between rounds: the decisions
// of the separate round launches (each reads only the state before its
// round).  It finishes a colouring that a fixed batch of round launches (a
// captured step) did not; when the batch sufficed it exits at once.  Past the
// last round slot it sets the overflow flag.
constexpr uint32_t kFinishBlock = 1024;
__global__ void __launch_bounds__(kFinishBlock) k_colour_finish(ColourGraph g, ColourBuffers b, uint32_t first) {
    __shared__ uint32_t more;
    for (uint32_t r = first; r < kColourRoundSlots; ++r) {
        __syncthreads();  // the previous round's flags and colours are written
        if (threadIdx.x == 0) more = b.pending[r - 1];
        __syncthreads();
        if (more == 0u) return;  // uniform
        if (r == kColourRoundSlots - 1) {
            if (threadIdx.x == 0) b.pending[kColourRoundSlots - 1] = 1u;  // out of round slots
            return;
        }
        for (uint32_t i = threadIdx.x; i < g.n; i += kFinishBlock) colour_round_body(g, b, i, r);
    }
}

// Colour ranges of the sorted list: starts[c] .. starts[c + 1] holds colour c
// (c < kMaxColours), and starts[kMaxColours + 1] the number of colours (the
// largest colour in use + 1).  One workgroup, one count per thread.
static_assert(kMaxColours + 1 == kBlock, "one colour count per thread");
__global__ void __launch_bounds__(kBlock) k_colour_starts(ColourBuffers b, uint32_t* starts) {
    __shared__ uint32_t sums[kBlock];
    __shared__ uint32_t ncol;
    const uint32_t t = threadIdx.x;
    const uint32_t cnt = b.counts[t];
    sums[t] = cnt;
    if (t == 0) ncol = 0u;
    __syncthreads();
    for (uint32_t off = 1; off < (uint32_t)kBlock; off <<= 1) {
        const uint32_t x = t >= off ? sums[t - off] : 0u;
        __syncthreads();
        sums[t] += x;
        __syncthreads();
    }
    starts[t + 1] = sums[t];
    if (t == 0) starts[0] = 0u;
    if (t < kMaxColours && cnt != 0u) atomicMax(&ncol, t + 1);
    __syncthreads();
    if (t == 0) starts[kMaxColours + 1] = ncol;
}

// Sort keys (colour, or kMaxColours for a body not in the graph) and the
// count per key: a block histogram in LDS, then one global add per (block,
// key) -- a handful of colours would otherwise put every thread's atomic on
// the same few words.
static_assert(kMaxColours + 1 == kBlock, "one histogram bin per thread");
__global__ void __launch_bounds__(kBlock) k_colour_keys(ColourGraph g, ColourBuffers b) {
    __shared__ uint32_t hist[kMaxColours + 1];
    hist[threadIdx.x] = 0u;
    __syncthreads();
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < g.n) {
        // a body left uncoloured (only past the last round slot, with the
        // overflow flag set) still gets a key within the histogram
        const uint32_t c = b.colour[i];
        const uint32_t key = b.round_of[i] == kColourNotInGraph ? kMaxColours : (c < kMaxColours ? c : kMaxColours - 1);
        b.keys[0][i] = key;
        b.vals[0][i] = i;
        atomicAdd(&hist[key], 1u);
    }
    __syncthreads();
    const uint32_t h = hist[threadIdx.x];
    if (h) atomicAdd(&b.counts[threadIdx.x], h);
}

}  // namespace

size_t colour_buffer_words(uint32_t n_cap) {
    const size_t n = n_cap ? n_cap : 1;
    return 6 * n + kColourRoundSlots + (kMaxColours + 1) + radix_scratch_size((uint32_t)n) + 64;
}

void colour_buffers_bind(ColourBuffers& b, uint32_t* base, uint32_t n_cap) {
    const size_t n = n_cap ? n_cap : 1;
    uint32_t* p = base;
    b.colour = p; p += n;
    b.round_of = p; p += n;
    b.keys[0] = p; p += n;
    b.keys[1] = p; p += n;
    b.vals[0] = p; p += n;
    b.vals[1] = p; p += n;
    b.pending = p; p += kColourRoundSlots;
    b.counts = p; p += kMaxColours + 1;
    b.sort_scratch = p;
}

void launch_colour_init(const ColourGraph& g, const ColourBuffers& b, hipStream_t s) {
    const uint32_t blocks = blocks_for(g.n > kColourRoundSlots ? g.n : kColourRoundSlots);
    hipLaunchKernelGGL(k_colour_init, dim3(blocks), dim3(kBlock), 0, s, g, b);
}

void launch_colour_rounds(const ColourGraph& g, const ColourBuffers& b, uint32_t first, uint32_t count,
                          hipStream_t s) {
    if (g.n == 0) return;
    for (uint32_t r = first; r < first + count && r < kColourRoundSlots - 1; ++r)
        hipLaunchKernelGGL(k_colour_round, dim3(blocks_for(g.n)), dim3(kBlock), 0, s, g, b, r);
}

void launch_colour_finish(const ColourGraph& g, const ColourBuffers& b, uint32_t first, hipStream_t s) {
    if (g.n == 0 || first == 0) return;
    hipLaunchKernelGGL(k_colour_finish, dim3(1), dim3(kFinishBlock), 0, s, g, b, first);
}

void launch_colour_starts(const ColourBuffers& b, uint32_t* starts, hipStream_t s) {
    hipLaunchKernelGGL(k_colour_starts, dim3(1), dim3(kBlock), 0, s, b, starts);
}

int launch_colour_lists(const ColourGraph& g, const ColourBuffers& b, hipStream_t s) {
    launch_fill_u32(b.counts, 0u, kMaxColours + 1, s);
    if (g.n == 0) return 0;
    hipLaunchKernelGGL(k_colour_keys, dim3(blocks_for(g.n)), dim3(kBlock), 0, s, g, b);
    return radix_sort_pairs(b.keys[0], b.vals[0], b.keys[1], b.vals[1], g.n, 8, b.sort_scratch, s);
}

}  // namespace dio
