// rt_image.hip — image export: float RGB -> BGRA8 UNORM, the conversion an OpenCL write_imagef makes
// into the reference's render targets (rendering/_core.py:340: RGBA images are CL_BGRA /
// CL_UNORM_INT8; OpenCL 1.2 section 8.3.1.1: channel = convert_uchar_sat_rte(f * 255.0f)).
// Saturating round-to-nearest-even of f * 255; NaN -> 0; alpha = 255 (the shading output is opaque).
#include "rt_common.h"

namespace rt {
namespace {

RT_DEV uint32_t unorm8(float f) {
    const float x = f * 255.0f;
    if (!(x > 0.0f)) return 0u;  // negative, zero or NaN
    if (x >= 255.0f) return 255u;
    return (uint32_t)rintf(x);  // ties to even
}

// One pixel per thread: 12 bytes read, one 4-byte store (B, G, R, A in memory order).
__global__ void __launch_bounds__(256) k_rgb_to_bgra8(const float* __restrict__ rgb, uint32_t n,
                                                       uint32_t* __restrict__ bgra) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t k = 3 * (uint64_t)i;
    const uint32_t r = unorm8(rgb[k]), g = unorm8(rgb[k + 1]), b = unorm8(rgb[k + 2]);
    bgra[i] = b | (g << 8) | (r << 16) | (255u << 24);
}

// One texel per thread: float4 RGBA (16 B) -> 4 UNORM8 bytes in R, G, B, A order.
__global__ void __launch_bounds__(256) k_texels_to_rgba8(const float4* __restrict__ in, uint64_t n,
                                                          uint32_t* __restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 c = in[i];
    out[i] = unorm8(c.x) | (unorm8(c.y) << 8) | (unorm8(c.z) << 16) | (unorm8(c.w) << 24);
}

}  // namespace
}  // namespace rt

using namespace rt;

extern "C" int rt_texels_to_rgba8(const float* rgba, uint64_t n_texels, uint8_t* out, void* stream) {
    if (n_texels == 0) return RT_OK;
    if (!rgba || !out) return set_error(RT_EINVAL, "rt_texels_to_rgba8: input and output are required");
    if (((uintptr_t)rgba & 15) || ((uintptr_t)out & 3))
        return set_error(RT_EINVAL, "rt_texels_to_rgba8: input must be 16-byte, output 4-byte aligned");
    if (n_texels > (1ull << 36)) return set_error(RT_ECAPACITY, "rt_texels_to_rgba8: too many texels");
    hipLaunchKernelGGL(k_texels_to_rgba8, dim3((uint32_t)((n_texels + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, (const float4*)rgba, n_texels, (uint32_t*)out);
    return check_launch("rt_texels_to_rgba8");
}

extern "C" int rt_image_to_bgra8(const float* rgb, uint32_t n_pixels, uint8_t* bgra, void* stream) {
    if (n_pixels == 0) return RT_OK;
    if (!rgb || !bgra) return set_error(RT_EINVAL, "rt_image_to_bgra8: rgb and bgra are required");
    if (((uintptr_t)rgb & 3) || ((uintptr_t)bgra & 3))
        return set_error(RT_EINVAL, "rt_image_to_bgra8: buffers must be 4-byte aligned");
    hipLaunchKernelGGL(k_rgb_to_bgra8, dim3((n_pixels + 255) / 256), dim3(256), 0, (hipStream_t)stream,
                       rgb, n_pixels, (uint32_t*)bgra);
    return check_launch("rt_image_to_bgra8");
}
