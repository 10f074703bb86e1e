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
 *
 * A double pow per value would make the kernel compute-bound (FP64), so the
 * value is first computed in float with the hardware log2 / exp2 (about 1 ulp
 * each): 255 * exp2(log2(c) / gamma) lands within 4e-4 of the exact value for
 * every c that maps into [1, 255] (the exponent log2(x / 255) is at most 8 in
 * magnitude, so its rounding costs at most 2^-19 relative).  The byte is the
 * truncation of that value, so the float result decides it unless it lies
 * within 1e-3 of an integer, or c <= 0 (powers of negative values and zero
 * follow the double pow).  Those values, about 0.2% of a film, are listed
 * and a second kernel evaluates the reference's double arithmetic for them,
 * so no wave diverges into the FP64 path.  Each block owns a segment of the
 * list, counted with an LDS atomic (one per wave that has such values): a
 * single global counter would serialise hundreds of thousands of
 * same-address atomics in L2.  Should a segment fill up, the remaining
 * values are evaluated in place.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rt_kernels.h"

namespace {

struct Gamma {
    float f;   /* 1/gamma rounded to float */
    double d;  /* 1./gamma as the reference computes it */
};

/* Tonemapper::clamp of the reference's float value */
__device__ __forceinline__ uint32_t clamp8(float v)
{
    if (v <= 0.f) return 0u;
    if (v >= 255.f) return 255u;
    if (v != v) return 0u;
    return (uint32_t)(int)v;
}

__device__ __forceinline__ uint32_t tone_exact(float c, Gamma g) { return clamp8((float)(255.0 * pow((double)c, g.d))); }

/* The float estimate and whether it decides the byte. */
__device__ __forceinline__ bool tone_fast(float c, Gamma g, uint32_t& byte)
{
    float v = 255.f * __builtin_amdgcn_exp2f(g.f * __builtin_amdgcn_logf(c));
    byte = clamp8(v);
    return c > 0.f && !(fabsf(v - rintf(v)) < 1e-3f);
}

struct Deferred {
    int64_t* list;        /* gridDim.x segments of seg entries */
    unsigned int* count;  /* entries used per segment */
    int seg;
};

/* Lanes whose value needs the double path append its index to the block's
 * segment; returns false when the segment was full for this lane, which then
 * evaluates the value itself. */
__device__ __forceinline__ bool defer(bool need, int64_t idx, const Deferred& d, unsigned int* cnt)
{
    const unsigned long long m = __ballot(need);
    if (m == 0) return true;
    const int lane = __lane_id();
    const int leader = __ffsll((long long)m) - 1;
    unsigned int base = 0;
    if (lane == leader) base = atomicAdd(cnt, (unsigned int)__popcll(m));
    base = __shfl(base, leader);
    if (!need) return true;
    const unsigned int slot = base + (unsigned int)__popcll(m & ((1ull << lane) - 1ull));
    if (slot >= (unsigned int)d.seg) return false;
    d.list[(size_t)blockIdx.x * d.seg + slot] = idx;
    return true;
}

/* Each block iteration covers TM_UNROLL x 256 consecutive float4: a lane
 * issues its TM_UNROLL loads (256 float4 apart, so every load instruction of
 * a wave is one contiguous 1 KB) before converting any, which keeps
 * TM_UNROLL x 16 B per lane in flight. */
#define TM_UNROLL 4
__global__ void __launch_bounds__(256) k_tonemap4(const float4* __restrict__ in, uint32_t* __restrict__ out,
                                                  int64_t n4, Gamma g, Deferred d)
{
    __shared__ unsigned int cnt;
    if (threadIdx.x == 0) cnt = 0;
    __syncthreads();
    const int64_t stride = (int64_t)gridDim.x * 256 * TM_UNROLL;
    for (int64_t i0 = (int64_t)blockIdx.x * 256 * TM_UNROLL + threadIdx.x; i0 < n4; i0 += stride) {
        float4 c[TM_UNROLL];
#pragma unroll
        for (int u = 0; u < TM_UNROLL; ++u) {
            const int64_t i = i0 + 256 * u;
            c[u] = i < n4 ? in[i] : make_float4(0.f, 0.f, 0.f, 0.f);
        }
#pragma unroll
        for (int u = 0; u < TM_UNROLL; ++u) {
            const int64_t i = i0 + 256 * u;
            const float cc[4] = {c[u].x, c[u].y, c[u].z, c[u].w};
            uint32_t word = 0;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                uint32_t b;
                const bool ok = tone_fast(cc[k], g, b) || i >= n4;
                if (!defer(!ok, 4 * i + k, d, &cnt)) b = tone_exact(cc[k], g);
                word |= b << (8 * k);
            }
            if (i < n4) out[i] = word;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) d.count[blockIdx.x] = min(cnt, (unsigned int)d.seg);
}

__global__ void __launch_bounds__(256) k_tonemap1(const float* __restrict__ in, uint8_t* __restrict__ out,
                                                  int64_t begin, int64_t n, Gamma g)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = begin + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        out[i] = (uint8_t)tone_exact(in[i], g);
}

/* block b: the values listed in segment b, in the reference's double arithmetic */
__global__ void __launch_bounds__(256) k_tonemap_deferred(const float* __restrict__ in, uint8_t* __restrict__ out,
                                                          Gamma g, Deferred d)
{
    const unsigned int n = d.count[blockIdx.x];
    const int64_t* seg = d.list + (size_t)blockIdx.x * d.seg;
    for (unsigned int k = threadIdx.x; k < n; k += blockDim.x) {
        const int64_t i = seg[k];
        out[i] = (uint8_t)tone_exact(in[i], g);
    }
}

}  // namespace

int tonemap_grid(int64_t n, int grid)
{
    return (int)std::max<int64_t>(1, std::min<int64_t>(grid, (n / 4 + 256 * TM_UNROLL - 1) / (256 * TM_UNROLL)));
}

hipError_t launch_tonemap(const float* film, uint8_t* out, int64_t n, float gamma, int64_t* list,
                          unsigned int* count, int64_t cap, int grid, hipStream_t st)
{
    Gamma g;
    g.d = 1. / (double)gamma;  // 1./gamma with gamma a float
    g.f = (float)g.d;
    int64_t done = 0;
    if (((uintptr_t)film & 15) == 0 && ((uintptr_t)out & 3) == 0 && n >= 4) {
        const int gr = tonemap_grid(n, grid);
        const Deferred d{list, count, (int)std::min<int64_t>(cap / gr, 1 << 30)};
        k_tonemap4<<<gr, 256, 0, st>>>((const float4*)film, (uint32_t*)out, n / 4, g, d);
        k_tonemap_deferred<<<gr, 256, 0, st>>>(film, out, g, d);
        done = 4 * (n / 4);
    }
    if (done < n) {
        int gr = (int)std::min<int64_t>(grid, (n - done + 255) / 256);
        k_tonemap1<<<gr, 256, 0, st>>>(film, out, done, n, g);
    }
    return hipGetLastError();
}
