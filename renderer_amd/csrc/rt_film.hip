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
 * and the block evaluates the reference's double arithmetic for them at the
 * end of its loop, so no wave diverges into the FP64 path while it streams.
 * The list is the block's own, in LDS (counted with an LDS atomic, one per
 * wave that has such values): a global counter would serialise hundreds of
 * thousands of same-address atomics in L2, and a second launch over a
 * global list cost 22 us of 240 on an 8192^2 film.  Should the list fill
 * up, the remaining values are evaluated in place.
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

/* Each block iteration covers TM_UNROLL x 256 consecutive float4: a lane
 * issues its TM_UNROLL loads (256 float4 apart, so every load instruction of
 * a wave is one contiguous 1 KB) before converting any, which keeps
 * TM_UNROLL x 16 B per lane in flight.  The film is read and the bytes
 * written once, as nontemporal accesses (no cache keeps lines that are not
 * used again): 0.228 -> 0.199 ms on an 8192^2 film; 8 loads in flight
 * instead of 4 gain nothing. */
typedef float vf4 __attribute__((ext_vector_type(4)));
#ifndef TM_UNROLL
#define TM_UNROLL 4
#endif
/* deferred values per block: about 0.2% of a film's, 200 per block on an
 * 8192^2 film with the default grid */
#define TM_LIST 1024
__global__ void __launch_bounds__(256) k_tonemap4(const float4* __restrict__ in, uint32_t* __restrict__ out,
                                                  int64_t n4, Gamma g)
{
    __shared__ unsigned int cnt;
    __shared__ int64_t list[TM_LIST];
    if (threadIdx.x == 0) cnt = 0;
    __syncthreads();
    const int lane = __lane_id();
    const int64_t stride = (int64_t)gridDim.x * 256 * TM_UNROLL;
    for (int64_t i0 = (int64_t)blockIdx.x * 256 * TM_UNROLL + threadIdx.x; i0 < n4; i0 += stride) {
        float4 c[TM_UNROLL];
#pragma unroll
        for (int u = 0; u < TM_UNROLL; ++u) {
            const int64_t i = i0 + 256 * u;
            const vf4 t = i < n4 ? __builtin_nontemporal_load(reinterpret_cast<const vf4*>(in) + i) : vf4{0.f, 0.f, 0.f, 0.f};
            c[u] = make_float4(t.x, t.y, t.z, t.w);
        }
#pragma unroll
        for (int u = 0; u < TM_UNROLL; ++u) {
            const int64_t i = i0 + 256 * u;
            const float cc[4] = {c[u].x, c[u].y, c[u].z, c[u].w};
            uint32_t word = 0;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                uint32_t b;
                const bool need = !(tone_fast(cc[k], g, b) || i >= n4);
                /* values the float estimate does not decide go to the list */
                const unsigned long long m = __ballot(need);
                if (m) {
                    const int leader = __ffsll((long long)m) - 1;
                    unsigned int base = 0;
                    if (lane == leader) base = atomicAdd(&cnt, (unsigned int)__popcll(m));
                    base = __shfl(base, leader);
                    if (need) {
                        const unsigned int slot = base + (unsigned int)__popcll(m & ((1ull << lane) - 1ull));
                        if (slot < TM_LIST) list[slot] = 4 * i + k;
                        else b = tone_exact(cc[k], g);
                    }
                }
                word |= b << (8 * k);
            }
            if (i < n4) __builtin_nontemporal_store(word, out + i);
        }
    }
    /* the listed values in the reference's double arithmetic; the barrier
     * orders the words written above before these byte stores */
    __syncthreads();
    const unsigned int n = min(cnt, (unsigned int)TM_LIST);
    const float* inf = reinterpret_cast<const float*>(in);
    uint8_t* outb = reinterpret_cast<uint8_t*>(out);
    for (unsigned int k = threadIdx.x; k < n; k += blockDim.x) {
        const int64_t i = list[k];
        outb[i] = (uint8_t)tone_exact(inf[i], g);
    }
}

__global__ void __launch_bounds__(256) k_tonemap1(const float* __restrict__ in, uint8_t* __restrict__ out,
                                                  int64_t begin, int64_t n, Gamma g)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = begin + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        out[i] = (uint8_t)tone_exact(in[i], g);
}

}  // namespace

int tonemap_grid(int64_t n, int grid)
{
    return (int)std::max<int64_t>(1, std::min<int64_t>(grid, (n / 4 + 256 * TM_UNROLL - 1) / (256 * TM_UNROLL)));
}

hipError_t launch_tonemap(const float* film, uint8_t* out, int64_t n, float gamma, int grid, hipStream_t st)
{
    Gamma g;
    g.d = 1. / (double)gamma;  // 1./gamma with gamma a float
    g.f = (float)g.d;
    int64_t done = 0;
    if (((uintptr_t)film & 15) == 0 && ((uintptr_t)out & 3) == 0 && n >= 4) {
        k_tonemap4<<<tonemap_grid(n, grid), 256, 0, st>>>((const float4*)film, (uint32_t*)out, n / 4, g);
        done = 4 * (n / 4);
    }
    if (done < n) {
        int gr = (int)std::min<int64_t>(grid, (n - done + 255) / 256);
        k_tonemap1<<<gr, 256, 0, st>>>(film, out, done, n, g);
    }
    return hipGetLastError();
}
