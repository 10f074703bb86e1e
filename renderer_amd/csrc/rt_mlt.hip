/*
 * rt_mlt.hip — chain kernels of the Markov-chain renderers on gfx950:
 * primary sample space Metropolis (MetropolisRenderer,
 * src/metropolisrenderer.cpp:56-200; MetropolisSample,
 * src/metropolissample.cpp) and energy redistribution path tracing
 * (EnergyRedistributionRenderer, src/energyredistributionrenderer.cpp:38-131).
 *
 * A chain is one path slot of the wavefront pipeline.  Its state is the
 * primary-sample vector (G float4 groups, stored group-major so that a wave
 * reads 64 consecutive float4 of one group) and the radiance of that state.
 * One mutation of every chain of a batch is: k_mlt_propose writes the
 * proposals, the path kernels trace them (PtParams::pss), k_mlt_accept
 * deposits and accepts.  Chains never interact, so a batch needs no
 * synchronisation beyond kernel order.
 *
 * Deposits land on random pixels from many chains at once.  They are added
 * as unsigned 64-bit fixed point (the host picks the unit from a bound on a
 * single deposit and the number of deposits), so the sum is exact and
 * independent of the order in which atomics arrive: a render is
 * reproducible bit for bit.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "rt_device.h"
#include "rt_kernels.h"
#include "rt_mlt.h"

#define MLT_BLOCK 256

namespace {

__device__ __forceinline__ float4 as_f4(U4 u) { return make_float4(u.u[0], u.u[1], u.u[2], u.u[3]); }

__device__ __forceinline__ U4 draw(uint32_t seed, uint32_t stream, uint32_t index, uint32_t step, uint32_t group)
{
    return rng4(seed, stream, index, step, group);
}

/* mutate(qreal&), src/metropolissample.cpp:62-78: a step of length
 * b * exp(-log(b / a) * u) between a = 1/1024 and b = 1/64, negative when
 * v < 0.5, wrapped into [0, 1]. */
__device__ __forceinline__ float perturb(float s, float u, float v)
{
    const float a = 1.f / 1024.f, b = 1.f / 64.f;
    const float log_ratio = -logf(b / a);
    float delta = (float)((double)b * exp((double)log_ratio * (double)u));
    if (v < 0.5f) delta = -delta;
    s += delta;
    if (s < 0.f) s += 1.f;
    if (s > 1.f) s -= 1.f;
    return s;
}

/* lum() of src/mathhelper.cpp:102-105 (double arithmetic, float result); the
 * _rn intrinsics keep the compiler from fusing into FMAs, as on the host. */
__device__ __forceinline__ float lum_ref(float r, float g, float b)
{
    return (float)__dadd_rn(__dadd_rn(__dmul_rn(.299, (double)r), __dmul_rn(.587, (double)g)), __dmul_rn(.114, (double)b));
}
/* cv::norm of a Vec3f */
__device__ __forceinline__ double norm_ref(float r, float g, float b)
{
    return sqrt(__dadd_rn(__dadd_rn(__dmul_rn(r, r), __dmul_rn(g, g)), __dmul_rn(b, b)));
}

/* MetropolisSample::getPos, src/metropolissample.cpp:129-136 */
__device__ __forceinline__ int pixel_of(float4 g0, int W, int H)
{
    int x = max(0, min((int)((double)g0.x * W), W - 1));
    int y = max(0, min((int)((double)g0.y * H), H - 1));
    return y * W + x;
}

/* ERPT initial sample i = pixel * spp + s: pixel plus the jitter of stratum
 * (s / spp_y, s % spp_y) over the image size
 * (src/energyredistributionrenderer.cpp:66-75, src/jitteredsampler.cpp:23-35) */
__device__ __forceinline__ float4 erpt_group(uint32_t seed, uint32_t i, int g, int W, int H, int sx, int sy)
{
    float4 v = as_f4(draw(seed, RT_MLT_STREAM_SEED, i, 0, (uint32_t)g));
    if (g == 0) {
        const int spp = sx * sy;
        const uint32_t pix = i / (uint32_t)spp;
        const int s = (int)(i % (uint32_t)spp), a = s / sy, b = s % sy;
        double jx = ((double)v.x + a) / sx, jy = ((double)v.y + b) / sy;
        v.x = (float)(((double)(pix % (uint32_t)W) + jx) / W);
        v.y = (float)(((double)(pix / (uint32_t)W) + jy) / H);
    }
    return v;
}

__device__ __forceinline__ void deposit(unsigned long long* acc, int pix, float r, float g, float b, double inv_unit)
{
    unsigned long long* a = acc + 3 * (size_t)pix;
    if (r > 0.f) atomicAdd(a, __double2ull_rn((double)r * inv_unit));
    if (g > 0.f) atomicAdd(a + 1, __double2ull_rn((double)g * inv_unit));
    if (b > 0.f) atomicAdd(a + 2, __double2ull_rn((double)b * inv_unit));
}

/* addSample's film deposit w * value and the weight sums of the running
 * variance, in fixed point */
__device__ __forceinline__ void flt_deposit(const MltDev& m, int film, int pix, float w, float x, float y, float z)
{
    const size_t pw = (size_t)film * m.W * m.H + pix;
    atomicAdd(m.acc_w + pw, __double2ull_rn((double)w * m.inv_unit_w));
    const float v[3] = {x, y, z};
    for (int c = 0; c < 3; ++c) {
        if (!(v[c] > 0.f)) continue;
        atomicAdd(m.acc + 3 * pw + c, __double2ull_rn((double)(w * v[c]) * m.inv_unit));
        atomicAdd(m.acc_q + 3 * pw + c, __double2ull_rn((double)w * v[c] * v[c] * m.inv_unit_q));
    }
}

__global__ __launch_bounds__(MLT_BLOCK) void k_flt_boot_weight(MltDev m, const float4* __restrict__ val,
                                                              float4* out, int n)
{
    for (int j = blockIdx.x * MLT_BLOCK + threadIdx.x; j < n; j += gridDim.x * MLT_BLOCK) {
        const float4 v = val[j];
        const int pos = pixel_of(m.prop[j], m.W, m.H);
        out[j] = make_float4(v.x, v.y, v.z, (float)(norm_ref(v.x, v.y, v.z) * (double)m.imp[pos]));
    }
}

__global__ __launch_bounds__(MLT_BLOCK) void k_mlt_fill(MltDev m, uint32_t stream, uint32_t index0, int n, int mode)
{
    for (int j = blockIdx.x * MLT_BLOCK + threadIdx.x; j < n; j += gridDim.x * MLT_BLOCK) {
        const uint32_t i = index0 + (uint32_t)j;
        for (int g = 0; g < m.G; ++g) {
            float4 v;
            if (mode == 2) {
                v = erpt_group(m.seed, i, g, m.W, m.H, m.spp_x, m.spp_y);
            } else {
                v = as_f4(draw(m.seed, stream, i, m.step, (uint32_t)g));
                if (mode == 1 && g == 0) {  /* the pixel in pixel units (camera.getRay(pixelCoord)) */
                    v.x = (float)(i % (uint32_t)m.W);
                    v.y = (float)(i / (uint32_t)m.W);
                }
            }
            m.prop[(size_t)g * m.cap + j] = v;
        }
    }
}

__global__ __launch_bounds__(MLT_BLOCK) void k_mlt_values(PathState ps, int n, float4* out, int metric)
{
    for (int j = blockIdx.x * MLT_BLOCK + threadIdx.x; j < n; j += gridDim.x * MLT_BLOCK) {
        float4 L = ps.radiance[j], c = ps.pend_c[j];
        float r = L.x + c.x, g = L.y + c.y, b = L.z + c.z;
        float w = metric == 0 ? lum_ref(r, g, b) : (float)norm_ref(r, g, b);
        out[j] = make_float4(r, g, b, w);
    }
}

/* The chain's start: the first bootstrap path whose running luminance sum
 * exceeds u * total (src/metropolisrenderer.cpp:97-106). */
__global__ __launch_bounds__(MLT_BLOCK) void k_mlt_start(MltDev m, const double* __restrict__ cdf, int64_t nb,
                                                        double total, const float4* __restrict__ bval, int n)
{
    for (int j = blockIdx.x * MLT_BLOCK + threadIdx.x; j < n; j += gridDim.x * MLT_BLOCK) {
        const uint32_t c = m.chain0 + (uint32_t)j;
        const double target = (double)draw(m.seed, RT_MLT_STREAM_START, c, 0, 0).u[0] * total;
        int64_t lo = 0, hi = nb;  /* first index with cdf > target */
        while (lo < hi) {
            int64_t mid = (lo + hi) >> 1;
            if (cdf[mid] > target) hi = mid;
            else lo = mid + 1;
        }
        const int64_t i = lo < nb ? lo : nb - 1;
        for (int g = 0; g < m.G; ++g)
            m.cur[(size_t)g * m.cap + j] = as_f4(draw(m.seed, RT_MLT_STREAM_BOOTSTRAP, (uint32_t)i, m.step, (uint32_t)g));
        float4 v = bval[i];
        if (m.kind == RT_MLT_FILTERED) v.w = (float)norm_ref(v.x, v.y, v.z);
        m.cur_val[j] = v;
    }
}

__global__ __launch_bounds__(MLT_BLOCK) void k_erpt_count(const float4* __restrict__ val, int n, uint32_t index0,
                                                         uint32_t seed, float ed, int M, int32_t* nch)
{
    for (int j = blockIdx.x * MLT_BLOCK + threadIdx.x; j < n; j += gridDim.x * MLT_BLOCK) {
        const float4 v = val[j];
        int k = 0;
        if (ed > 0.f) {
            double u = (double)draw(seed, RT_MLT_STREAM_COUNT, index0 + (uint32_t)j, 0, 0).u[0];
            k = (int)(u + norm_ref(v.x, v.y, v.z) / (double)((float)M * ed));
        }
        nch[j] = k;
    }
}

/* depVal = c * (e_d / norm(c) / chains) (src/energyredistributionrenderer.cpp:88-90) */
__global__ __launch_bounds__(MLT_BLOCK) void k_erpt_start(MltDev m, const int32_t* __restrict__ src,
                                                         const int32_t* __restrict__ nch,
                                                         const float4* __restrict__ val, float ed, int n)
{
    for (int j = blockIdx.x * MLT_BLOCK + threadIdx.x; j < n; j += gridDim.x * MLT_BLOCK) {
        const int32_t i = src[j];
        for (int g = 0; g < m.G; ++g)
            m.cur[(size_t)g * m.cap + j] = erpt_group(m.seed, (uint32_t)i, g, m.W, m.H, m.spp_x, m.spp_y);
        const float4 c = val[i];
        const double nr = norm_ref(c.x, c.y, c.z);
        const double k = (double)ed / nr / (double)nch[i];
        m.cur_val[j] = make_float4(c.x, c.y, c.z, (float)nr);
        m.energy[j] = norm_ref((float)(c.x * k), (float)(c.y * k), (float)(c.z * k));
    }
}

/* MetropolisSample::mutated (src/metropolissample.cpp:104-118): a fresh path
 * length, then a large step with probability large_prob, else a small step.
 * ERPT takes small steps only and keeps the full length. */
__global__ __launch_bounds__(MLT_BLOCK) void k_mlt_propose(MltDev m, uint32_t step, int n)
{
    for (int j = blockIdx.x * MLT_BLOCK + threadIdx.x; j < n; j += gridDim.x * MLT_BLOCK) {
        const uint32_t c = m.chain0 + (uint32_t)j;
        const U4 ctl = draw(m.seed, RT_MLT_STREAM_CONTROL, c, step, 0);
        const bool large = m.kind != RT_MLT_ERPT && ctl.u[0] < m.large_prob;
        for (int g = 0; g < m.G; ++g) {
            float4 v;
            if (large) {
                v = as_f4(draw(m.seed, RT_MLT_STREAM_LARGE, c, step, (uint32_t)g));
            } else {
                const U4 d = draw(m.seed, RT_MLT_STREAM_SMALL, c, step, (uint32_t)g);
                v = m.cur[(size_t)g * m.cap + j];
                v.x = perturb(v.x, d.u[0], d.u[1]);
                v.y = perturb(v.y, d.u[2], d.u[3]);
            }
            if (g == 0 && m.kind != RT_MLT_ERPT) v.z = ctl.u[1];
            m.prop[(size_t)g * m.cap + j] = v;
        }
    }
}

__global__ __launch_bounds__(MLT_BLOCK) void k_mlt_accept(MltDev m, PathState ps, uint32_t step, int n)
{
    __shared__ unsigned int block_acc;
    if (threadIdx.x == 0) block_acc = 0;
    __syncthreads();
    for (int j = blockIdx.x * MLT_BLOCK + threadIdx.x; j < n; j += gridDim.x * MLT_BLOCK) {
        const uint32_t c = m.chain0 + (uint32_t)j;
        const float4 L = ps.radiance[j], pc = ps.pend_c[j];
        const float r = L.x + pc.x, g = L.y + pc.y, b = L.z + pc.z;
        const float u = draw(m.seed, RT_MLT_STREAM_CONTROL, c, step, 0).u[2];
        const float4 cv = m.cur_val[j];
        const float4 g0c = m.cur[j], g0p = m.prop[j];
        bool move;
        if (m.kind == RT_MLT_METROPOLIS) {
            /* src/metropolisrenderer.cpp:156-182 (importance map all ones) */
            const float ln = lum_ref(r, g, b), lc = cv.w;
            const float a = lc > 0.f ? fminf(1.f, ln / lc) : 1.f;
            if (lc > 0.f) {
                const float s = (1.f / lc) * m.scale, w = 1.f - a;
                deposit(m.acc, pixel_of(g0c, m.W, m.H), (cv.x * s) * w, (cv.y * s) * w, (cv.z * s) * w, m.inv_unit);
            }
            if (ln > 0.f) {
                const float s = (1.f / ln) * m.scale;
                deposit(m.acc, pixel_of(g0p, m.W, m.H), (r * s) * a, (g * s) * a, (b * s) * a, m.inv_unit);
            }
            move = u < a;
            if (move) m.cur_val[j] = make_float4(r, g, b, ln);
        } else if (m.kind == RT_MLT_FILTERED) {
            /* src/metropolisfltrenderer.cpp:166-186: norm x importance */
            const double zn = norm_ref(r, g, b), cn = (double)cv.w;
            const int cp = pixel_of(g0c, m.W, m.H), np = pixel_of(g0p, m.W, m.H);
            const int film = (int)((c / m.film_chains) & 1u);
            const float a = (float)fmin(1.0, zn * (double)m.imp[np] / (cn * (double)m.imp[cp]));
            if (cn > 0.0 && a < 1.f) {
                const float w = (1.f - a) / m.newsum[cp];
                const double k = 1.0 / cn * (double)m.scale;
                flt_deposit(m, film, cp, w, (float)(cv.x * k), (float)(cv.y * k), (float)(cv.z * k));
            }
            if (zn > 0.0 && a > 0.f) {
                const float w = a / m.newsum[np];
                const double k = 1.0 / zn * (double)m.scale;
                flt_deposit(m, film, np, w, (float)(r * k), (float)(g * k), (float)(b * k));
            }
            move = u < a;
            if (move) m.cur_val[j] = make_float4(r, g, b, (float)zn);
        } else {
            /* src/energyredistributionrenderer.cpp:92-111 */
            const float zl = (float)norm_ref(r, g, b), yl = cv.w;
            const float q = fminf(1.f, zl / yl);
            move = q > u;
            if (move) m.cur_val[j] = make_float4(r, g, b, zl);
            const double E = m.energy[j], inv = 1.0 / (double)yl;
            float dr = (float)(E * cv.x), dg = (float)(E * cv.y), db = (float)(E * cv.z);
            dr = (float)(dr * inv);
            dg = (float)(dg * inv);
            db = (float)(db * inv);
            deposit(m.acc, pixel_of(move ? g0p : g0c, m.W, m.H), dr * m.scale, dg * m.scale, db * m.scale,
                    m.inv_unit);
        }
        if (move) {
            for (int k = 0; k < m.G; ++k) m.cur[(size_t)k * m.cap + j] = m.prop[(size_t)k * m.cap + j];
        }
        /* one LDS atomic per wave, one global atomic per block */
        const unsigned long long bal = __ballot(move);
        if ((threadIdx.x & 63) == 0 && bal) atomicAdd(&block_acc, (unsigned int)__popcll(bal));
    }
    __syncthreads();
    if (threadIdx.x == 0 && block_acc) atomicAdd(m.accepted, (unsigned long long)block_acc);
}

__global__ __launch_bounds__(MLT_BLOCK) void k_mlt_resolve(const unsigned long long* __restrict__ acc,
                                                          float* __restrict__ film, int64_t n, double unit)
{
    for (int64_t i = (int64_t)blockIdx.x * MLT_BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * MLT_BLOCK)
        film[i] += (float)((double)acc[i] * unit);
}

int blocks_for(int64_t n, int grid)
{
    int64_t b = (n + MLT_BLOCK - 1) / MLT_BLOCK;
    return (int)std::max<int64_t>(1, std::min<int64_t>(b, grid));
}

}  // namespace

hipError_t launch_mlt_fill(const MltDev& m, uint32_t stream, uint32_t index0, int n, int mode, int grid,
                           hipStream_t st)
{
    hipLaunchKernelGGL(k_mlt_fill, dim3(blocks_for(n, grid)), dim3(MLT_BLOCK), 0, st, m, stream, index0, n, mode);
    return hipGetLastError();
}

hipError_t launch_flt_boot_weight(const MltDev& m, const float4* val, float4* out, int n, int grid, hipStream_t st)
{
    hipLaunchKernelGGL(k_flt_boot_weight, dim3(blocks_for(n, grid)), dim3(MLT_BLOCK), 0, st, m, val, out, n);
    return hipGetLastError();
}

hipError_t launch_mlt_values(const PathState& ps, int n, float4* out, int metric, int grid, hipStream_t st)
{
    hipLaunchKernelGGL(k_mlt_values, dim3(blocks_for(n, grid)), dim3(MLT_BLOCK), 0, st, ps, n, out, metric);
    return hipGetLastError();
}

hipError_t launch_mlt_start(const MltDev& m, const double* cdf, int64_t nb, double total, const float4* bval,
                            int n, int grid, hipStream_t st)
{
    hipLaunchKernelGGL(k_mlt_start, dim3(blocks_for(n, grid)), dim3(MLT_BLOCK), 0, st, m, cdf, nb, total, bval, n);
    return hipGetLastError();
}

hipError_t launch_erpt_count(const float4* val, int n, uint32_t index0, uint32_t seed, float ed, int M,
                             int32_t* nch, int grid, hipStream_t st)
{
    hipLaunchKernelGGL(k_erpt_count, dim3(blocks_for(n, grid)), dim3(MLT_BLOCK), 0, st, val, n, index0, seed, ed, M,
                       nch);
    return hipGetLastError();
}

hipError_t launch_erpt_start(const MltDev& m, const int32_t* src, const int32_t* nch, const float4* val, float ed,
                             int n, int grid, hipStream_t st)
{
    hipLaunchKernelGGL(k_erpt_start, dim3(blocks_for(n, grid)), dim3(MLT_BLOCK), 0, st, m, src, nch, val, ed, n);
    return hipGetLastError();
}

hipError_t launch_mlt_propose(const MltDev& m, uint32_t step, int n, int grid, hipStream_t st)
{
    hipLaunchKernelGGL(k_mlt_propose, dim3(blocks_for(n, grid)), dim3(MLT_BLOCK), 0, st, m, step, n);
    return hipGetLastError();
}

hipError_t launch_mlt_accept(const MltDev& m, const PathState& ps, uint32_t step, int n, int grid, hipStream_t st)
{
    hipLaunchKernelGGL(k_mlt_accept, dim3(blocks_for(n, grid)), dim3(MLT_BLOCK), 0, st, m, ps, step, n);
    return hipGetLastError();
}

hipError_t launch_mlt_resolve(const unsigned long long* acc, float* film, int64_t n, double unit, int grid,
                              hipStream_t st)
{
    hipLaunchKernelGGL(k_mlt_resolve, dim3(blocks_for(n, grid)), dim3(MLT_BLOCK), 0, st, acc, film, n, unit);
    return hipGetLastError();
}
