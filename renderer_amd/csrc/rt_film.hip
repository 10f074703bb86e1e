/*
 * rt_film.hip — the tonemapper (src/tonemapper.cpp:39-56) on gfx950.
 *
 * Per channel value c of the float film: byte = clamp(255 * pow(c, 1/gamma))
 * with the reference's arithmetic: std::pow(float, double) is a double pow,
 * the product is rounded to float for Tonemapper::clamp(float), which
 * truncates to int; qRgb keeps the low 8 bits, so a NaN (x86 converts it to
 * INT_MIN) gives 0.
 *
 * Purely streaming: 12 B read and 3 B written per pixel.  One lane takes
 * four consecutive floats (one 16-B load) and writes their four bytes as one
 * 32-bit store, so a wave moves 1 KB in and 256 B out per iteration,
 * coalesced; a grid-stride loop over a fixed grid keeps the launch small.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rt_kernels.h"

namespace {

__device__ __forceinline__ uint32_t tone8(float c, double inv_gamma)
{
    float v = (float)(255.0 * pow((double)c, inv_gamma));
    if (v <= 0.f) return 0u;
    if (v >= 255.f) return 255u;
    if (v != v) return 0u;
    return (uint32_t)(int)v;
}

__global__ void __launch_bounds__(256) k_tonemap4(const float4* __restrict__ in, uint32_t* __restrict__ out,
                                                  int64_t n4, double inv_gamma)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride) {
        float4 c = in[i];
        out[i] = tone8(c.x, inv_gamma) | tone8(c.y, inv_gamma) << 8 | tone8(c.z, inv_gamma) << 16 |
                 tone8(c.w, inv_gamma) << 24;
    }
}

__global__ void __launch_bounds__(256) k_tonemap1(const float* __restrict__ in, uint8_t* __restrict__ out,
                                                  int64_t begin, int64_t n, double inv_gamma)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = begin + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        out[i] = (uint8_t)tone8(in[i], inv_gamma);
}

}  // namespace

hipError_t launch_tonemap(const float* film, uint8_t* out, int64_t n, float gamma, int grid, hipStream_t st)
{
    const double inv_gamma = 1. / (double)gamma;  // 1./gamma with gamma a float
    int64_t done = 0;
    if (((uintptr_t)film & 15) == 0 && ((uintptr_t)out & 3) == 0 && n >= 4) {
        const int64_t n4 = n / 4;
        int g = (int)std::min<int64_t>(grid, (n4 + 255) / 256);
        k_tonemap4<<<g, 256, 0, st>>>((const float4*)film, (uint32_t*)out, n4, inv_gamma);
        done = 4 * n4;
    }
    if (done < n) {
        int g = (int)std::min<int64_t>(grid, (n - done + 255) / 256);
        k_tonemap1<<<g, 256, 0, st>>>(film, out, done, n, inv_gamma);
    }
    return hipGetLastError();
}
