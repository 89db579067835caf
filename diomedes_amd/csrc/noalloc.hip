// noalloc.hip — the narrow phase over a pair list without any allocation
// (dio_narrow_pairs_noalloc).
//
// For callers that cannot allocate device memory (tight serving loops, code
// that must not call hipMalloc).  It computes what dio_narrow_submit_pairs
// followed by dio_narrow_read computes, with the same kernels: the bodies'
// records unpacked to structure-of-arrays, the per-pair dispatch (SAT, the
// sphere closed form, GJK), both EPA passes, the scan of the per-pair counts
// and the packing into the reference's 4-slots-per-pair records.  Every buffer
// beyond the caller's host inputs and outputs lives in `scratch`, a file-scope
// device array of fixed size: each GPU the library runs on holds its own copy
// from the library's load until the process ends.  Its device address is taken
// once per device with hipGetSymbolAddress, on the first call on that device.
//
// Calls on one device share that device's array.  A host mutex per device
// serialises them, and each call is synchronous (it returns after the results
// are copied back), so calls from several host threads do not overlap; work
// the caller enqueues on other streams may run beside it, since it touches
// nothing of the caller's on the device.
#include <hip/hip_runtime.h>

#include <stddef.h>
#include <stdint.h>

#include <mutex>

#include "../../include/diomedes_hip.h"
#include "abi_checks.hpp"
#include "broad.hpp"
#include "narrow.hpp"
#include "prim.hpp"

namespace {

constexpr uint32_t kMaxBodies = DIO_NOALLOC_MAX_BODIES;
constexpr uint32_t kMaxPairs = DIO_NOALLOC_MAX_PAIRS;
constexpr uint32_t kMaxHull = DIO_NOALLOC_MAX_HULL_VERTICES;
constexpr uint32_t kPartialWords = 1024;  // scan partials; scan_partials_size(kMaxPairs + 1) is 34
constexpr int kMaxDevices = 64;

constexpr size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Byte offsets of the parts of `scratch`, each 256-byte aligned.
struct Layout {
    size_t rec, pos, rot, tag, corners, hull_raw, hull, pairs, np, n, pa, pb, cnt, off, partials, gjk_list,
        gjk_count, retry_list, retry_count, simplex, slots, total;
};
constexpr Layout make_layout() {
    Layout L{};
    size_t at = 0;
    auto take = [&at](size_t bytes) {
        const size_t here = at;
        at = align256(at + bytes);
        return here;
    };
    L.rec = take((size_t)kMaxBodies * sizeof(dio_gpu_body));
    L.pos = take((size_t)kMaxBodies * 16);
    L.rot = take((size_t)kMaxBodies * 16);
    L.tag = take((size_t)kMaxBodies * 4);
    L.corners = take((size_t)kMaxBodies * 64);
    L.hull_raw = take((size_t)kMaxHull * 12);
    L.hull = take((size_t)kMaxHull * 16);
    L.pairs = take((size_t)kMaxPairs * 8);
    L.np = take(4);
    L.n = take((size_t)kMaxPairs * 64);
    L.pa = take((size_t)kMaxPairs * 64);
    L.pb = take((size_t)kMaxPairs * 64);
    L.cnt = take((size_t)kMaxPairs * 4);
    L.off = take(((size_t)kMaxPairs + 1) * 4);
    L.partials = take((size_t)kPartialWords * 4);
    L.gjk_list = take((size_t)kMaxPairs * 4);
    L.gjk_count = take(4);
    L.retry_list = take((size_t)kMaxPairs * 4);
    L.retry_count = take(4);
    L.simplex = take((size_t)kMaxPairs * 8 * 16);
    L.slots = take((size_t)kMaxPairs * 4 * sizeof(dio_gpu_contact_out));
    L.total = at;
    return L;
}
constexpr Layout kLayout = make_layout();
static_assert(kLayout.total == DIO_NOALLOC_SCRATCH_BYTES, "the header's scratch size must match the layout");

std::mutex g_lock[kMaxDevices];
unsigned char* g_scratch[kMaxDevices] = {};

__global__ void k_hull_to_f4(const float* raw, uint32_t count, float4* out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) out[i] = make_float4(raw[3 * i], raw[3 * i + 1], raw[3 * i + 2], 0.0f);
}

// Restores the calling thread's current device when the call returns.
struct DeviceGuard {
    int prev = -1;
    ~DeviceGuard() {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

}  // namespace

// The device array itself.  Host code never uses this name as an address: the
// address of each device's copy comes from hipGetSymbolAddress (g_scratch).
static __device__ __align__(256) unsigned char scratch[DIO_NOALLOC_SCRATCH_BYTES];

extern "C" dio_status dio_narrow_pairs_noalloc(int device, const dio_gpu_body* bodies, uint32_t n,
                                               const dio_gpu_pair* pairs, uint32_t pair_count,
                                               const float* hull_vertices, uint32_t hull_count,
                                               dio_gpu_contact_out* slots, uint32_t cap_slots) {
    using namespace dio;
    if (device < 0 || device >= kMaxDevices) return DIO_INVALID;
    if ((n > 0 && !bodies) || (pair_count > 0 && (!pairs || !slots)) || (hull_count > 0 && !hull_vertices))
        return DIO_INVALID;
    if (n > kMaxBodies || pair_count > kMaxPairs || hull_count > kMaxHull) return DIO_CAPACITY;
    if (cap_slots < 4ull * pair_count) return DIO_CAPACITY;
    for (uint32_t k = 0; k < pair_count; ++k)
        if (pairs[k].a >= n || pairs[k].b >= n) return DIO_INVALID;
    const int64_t need = hull_pool_need(bodies, n);
    if (need < 0 || need > (int64_t)hull_count) return DIO_INVALID;
    if (pair_count == 0) return DIO_OK;

    std::lock_guard<std::mutex> lock(g_lock[device]);
    DeviceGuard guard;
    if (hipGetDevice(&guard.prev) != hipSuccess) return DIO_HIP;
    if (hipSetDevice(device) != hipSuccess) {
        (void)hipGetLastError();
        return DIO_INVALID;
    }
    auto fail = []() {
        (void)hipGetLastError();
        return DIO_HIP;
    };
    if (!g_scratch[device]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return fail();
        g_scratch[device] = static_cast<unsigned char*>(p);
    }
    unsigned char* base = g_scratch[device];
    const Layout& L = kLayout;
    auto at = [base](size_t off) { return base + off; };
    hipStream_t s = nullptr;  // the null stream: the call is synchronous

    // inputs
    if (n > 0 && hipMemcpy(at(L.rec), bodies, (size_t)n * sizeof(dio_gpu_body), hipMemcpyHostToDevice) != hipSuccess)
        return fail();
    if (hipMemcpy(at(L.pairs), pairs, (size_t)pair_count * 8, hipMemcpyHostToDevice) != hipSuccess) return fail();
    if (hipMemcpy(at(L.np), &pair_count, 4, hipMemcpyHostToDevice) != hipSuccess) return fail();
    if (hull_count > 0) {
        if (hipMemcpy(at(L.hull_raw), hull_vertices, (size_t)hull_count * 12, hipMemcpyHostToDevice) != hipSuccess)
            return fail();
        hipLaunchKernelGGL(k_hull_to_f4, dim3((hull_count + 255) / 256), dim3(256), 0, s,
                           reinterpret_cast<const float*>(at(L.hull_raw)), hull_count,
                           reinterpret_cast<float4*>(at(L.hull)));
    }
    float4* pos = reinterpret_cast<float4*>(at(L.pos));
    float4* rot = reinterpret_cast<float4*>(at(L.rot));
    uint32_t* tag = reinterpret_cast<uint32_t*>(at(L.tag));
    float4* corners = reinterpret_cast<float4*>(at(L.corners));
    launch_unpack_gpu_bodies(reinterpret_cast<const dio_gpu_body*>(at(L.rec)), n, pos, rot, tag, corners, s);

    // the narrow phase and the packing, as the context runs them
    NarrowOut o{};
    o.n = reinterpret_cast<float4*>(at(L.n));
    o.pa = reinterpret_cast<float4*>(at(L.pa));
    o.pb = reinterpret_cast<float4*>(at(L.pb));
    o.cnt = reinterpret_cast<uint32_t*>(at(L.cnt));
    o.off = reinterpret_cast<uint32_t*>(at(L.off));
    o.scan_partials = reinterpret_cast<uint32_t*>(at(L.partials));
    o.gjk_list = reinterpret_cast<uint32_t*>(at(L.gjk_list));
    o.gjk_count = reinterpret_cast<uint32_t*>(at(L.gjk_count));
    o.epa_retry_list = reinterpret_cast<uint32_t*>(at(L.retry_list));
    o.epa_retry_count = reinterpret_cast<uint32_t*>(at(L.retry_count));
    o.pairs = reinterpret_cast<const uint2*>(at(L.pairs));
    o.p_dev = reinterpret_cast<const uint32_t*>(at(L.np));
    o.p_max = pair_count;
    if (scan_partials_size(pair_count + 1) > kPartialWords) return DIO_CAPACITY;
    const BodyView bv{pos, rot, tag, corners, reinterpret_cast<const float4*>(at(L.hull)), hull_count};
    // counters_cleared = false: the EPA counters left by the last call are
    // zeroed by a fill kernel before the dispatch
    run_narrow(bv, o, reinterpret_cast<float4*>(at(L.simplex)), s, false, true);
    dio_gpu_contact_out* packed = reinterpret_cast<dio_gpu_contact_out*>(at(L.slots));
    launch_pack_contact_out(o, pair_count, packed, s);
    if (hipGetLastError() != hipSuccess) return DIO_HIP;
    if (hipMemcpy(slots, packed, (size_t)pair_count * 4 * sizeof(dio_gpu_contact_out), hipMemcpyDeviceToHost) !=
        hipSuccess)
        return fail();
    return DIO_OK;
}
