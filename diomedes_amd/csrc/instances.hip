// instances.hip — instance transforms for an instanced renderer, written on
// the device from the world's poses.
//
// Counterpart of the reference's per-frame host work: the demo copies each
// body's pose into its scene instance (demos/avbd_box/src/main.rs:274-280),
// the renderer builds each instance's model matrix in its grouped order
// (engine/src/render/renderer.rs:222-255) and uploads (projection * view *
// model, model) pairs into a 256-entry uniform buffer
// (engine/src/render/context.rs:171-199).  Here one thread writes one
// instance slot's pair into device memory the caller names (an imported
// renderer buffer), with no instance limit.
//
// Matrix arithmetic follows glam 0.33.3 (the reference's math crate) in its
// operation order, compiled without contraction, so each result is the
// FP32 oracle's bit for bit: Mat4::from_scale_rotation_translation (the
// rotation's axes from the doubled components, each axis scaled, the
// translation column with w = 1) and Mat4 * Mat4 (column j of a * b is
// ((a0 b_j.x + a1 b_j.y) + a2 b_j.z) + a3 b_j.w per component).
//
// Bytes per instance slot: 4 (order) + 16 (position) + 16 (orientation) + 16
// (scale) read, 128 written; HBM-bound.
#include <hip/hip_runtime.h>

#include "instances.hpp"

namespace dio {

namespace {

constexpr int kInstBlock = 256;

__device__ __forceinline__ float4 scale4(float4 a, float s) { return make_float4(a.x * s, a.y * s, a.z * s, a.w * s); }

// column j of view_proj * m, for m's column c
__device__ __forceinline__ float4 mul_col(const Mat4Cols& a, float4 c) {
    return make_float4(((a.c[0].x * c.x + a.c[1].x * c.y) + a.c[2].x * c.z) + a.c[3].x * c.w,
                       ((a.c[0].y * c.x + a.c[1].y * c.y) + a.c[2].y * c.z) + a.c[3].y * c.w,
                       ((a.c[0].z * c.x + a.c[1].z * c.y) + a.c[2].z * c.z) + a.c[3].z * c.w,
                       ((a.c[0].w * c.x + a.c[1].w * c.y) + a.c[2].w * c.z) + a.c[3].w * c.w);
}

__global__ void __launch_bounds__(kInstBlock) k_instances(const float4* __restrict__ pos,
                                                          const float4* __restrict__ rot,
                                                          const float4* __restrict__ scale,
                                                          const uint32_t* __restrict__ order, uint32_t count,
                                                          Mat4Cols vp, float4* __restrict__ dst) {
    const uint32_t k = blockIdx.x * kInstBlock + threadIdx.x;
    if (k >= count) return;
    const uint32_t b = order ? order[k] : k;
    const float4 p = pos[b];
    const float4 q = rot[b];
    const float4 s = scale[b];
    // Mat4::from_scale_rotation_translation
    const float x2 = q.x + q.x, y2 = q.y + q.y, z2 = q.z + q.z;
    const float xx = q.x * x2, xy = q.x * y2, xz = q.x * z2;
    const float yy = q.y * y2, yz = q.y * z2, zz = q.z * z2;
    const float wx = q.w * x2, wy = q.w * y2, wz = q.w * z2;
    float4 m[4];
    m[0] = scale4(make_float4(1.0f - (yy + zz), xy + wz, xz - wy, 0.0f), s.x);
    m[1] = scale4(make_float4(xy - wz, 1.0f - (xx + zz), yz + wx, 0.0f), s.y);
    m[2] = scale4(make_float4(xz + wy, yz - wx, 1.0f - (xx + yy), 0.0f), s.z);
    m[3] = make_float4(p.x, p.y, p.z, 1.0f);
    float4* out = dst + 8 * (size_t)k;
#pragma unroll
    for (int j = 0; j < 4; ++j) out[j] = mul_col(vp, m[j]);
#pragma unroll
    for (int j = 0; j < 4; ++j) out[4 + j] = m[j];
}

}  // namespace

void launch_instances(const float4* pos, const float4* rot, const float4* scale, const uint32_t* order,
                      uint32_t count, const Mat4Cols& view_proj, float4* dst, hipStream_t s) {
    if (count == 0) return;
    hipLaunchKernelGGL(k_instances, dim3((count + kInstBlock - 1) / kInstBlock), dim3(kInstBlock), 0, s, pos, rot,
                       scale, order, count, view_proj, dst);
}

}  // namespace dio
