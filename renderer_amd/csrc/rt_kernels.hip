/*
 * rt_kernels.hip — gfx950 kernels of the path-tracing hot path.
 *
 * Wavefront organisation (one pass = a set of path slots):
 *   generate  -> camera rays for every slot            (src/camera.cpp:50-58)
 *   repeat max_depth times:
 *     extend  -> closest hit of every live path        (src/bvh.cpp:135-149)
 *     shade   -> emission, BSDF sample, next-event ray (src/renderer.cpp:69-106,
 *                src/unidipathtracingintegrator.cpp:48-107)
 *     shadow  -> any-hit test of next-event rays       (src/arealight.cpp:35-51)
 *     advance -> swap the live-path queues
 *   accumulate -> film += radiance / spp in sample order (src/perpixelrenderer.cpp:92)
 * Every kernel is persistent (a fixed grid strides over a queue whose length
 * it reads from device memory), so the host never waits inside a pass and
 * the whole sequence can be replayed as one hipGraph.  Queues are compacted
 * with a ballot + mbcnt prefix per wave, an LDS prefix over the block's
 * waves and one atomic per block.
 */
#include "rt_device.h"
#include "rt_guide.h"
#include "rt_guide_dev.h"
#include "rt_kernels.h"

#include <cstring>
#include <mutex>

#define PT_BLOCK 256
#define RT_PI 3.14159265358979f
#define RT_INV_PI 0.318309886183791f
#define RT_INV_2PI 0.159154943091895f

static_assert(sizeof(HostDevScene) == sizeof(DevScene), "scene view layouts differ");

/* Waves per SIMD the traversal kernels are compiled for, i.e. a register
 * budget of 512 / waves VGPRs.  Their latency-bound gathers gain from
 * occupancy: 6 waves (80 VGPRs, no spills in k_shadow, 36 B in k_extend)
 * measured 4% faster than the unconstrained 89 VGPRs (5 waves); 7 and 8
 * waves spill more than they gain. */
#ifndef RT_TRAV_WAVES
#define RT_TRAV_WAVES 6
#endif
#if RT_TRAV_WAVES > 0
#define RT_TRAV_OCCUPANCY __attribute__((amdgpu_waves_per_eu(RT_TRAV_WAVES, 8)))
#else
#define RT_TRAV_OCCUPANCY
#endif
/* Register budget of k_shade (waves per SIMD; 0 = the compiler's choice).
 * Left alone the unbiased shader takes 145 VGPRs (3 waves); 4 waves (128
 * VGPRs, 8 bytes of spill) shade a 1024^2 x 64 spp render 10% faster, 5
 * waves spill 108 bytes and gain nothing more. */
#ifndef RT_SHADE_WAVES
#define RT_SHADE_WAVES 4
#endif
#ifndef RT_SHADE_GUIDED_WAVES
#define RT_SHADE_GUIDED_WAVES 0
#endif
#if RT_SHADE_GUIDED_WAVES > 0
#define RT_SHADE_GUIDED_OCCUPANCY __attribute__((amdgpu_waves_per_eu(RT_SHADE_GUIDED_WAVES, 8)))
#else
#define RT_SHADE_GUIDED_OCCUPANCY
#endif
/* Register budget of k_finish (waves per SIMD; 0 = the compiler's choice).
 * Left alone it takes 184 VGPRs (2 waves), too few lanes for the paths
 * alive at the fifth bounce of a 32M-path pass; 4 waves (128 VGPRs, 69
 * spilled) finish them in 4.0 instead of 4.2 ms per render, 5 and 6 waves
 * spill more and gain nothing. */
#ifndef RT_FINISH_WAVES
#define RT_FINISH_WAVES 4
#endif
#if RT_FINISH_WAVES > 0
#define RT_FINISH_OCCUPANCY __attribute__((amdgpu_waves_per_eu(RT_FINISH_WAVES, 8)))
#else
#define RT_FINISH_OCCUPANCY
#endif
#if RT_SHADE_WAVES > 0
#define RT_SHADE_OCCUPANCY __attribute__((amdgpu_waves_per_eu(RT_SHADE_WAVES, 8)))
#else
#define RT_SHADE_OCCUPANCY
#endif

int64_t trav_lanes(int grid) { return (int64_t)grid * PT_BLOCK; }

static inline DevScene as_dev(const HostDevScene& s)
{
    DevScene d;
    std::memcpy(&d, &s, sizeof(d));
    return d;
}

/* Append to NQ queues from a whole block (PT_BLOCK threads, uniform control
 * flow across the block): per queue one atomic per block on its counter,
 * wave offsets from an LDS prefix over the block's 4 waves, lane offsets
 * from the ballot mask.  Same-address atomics serialise in L2, so one per
 * block instead of one per wave matters: the shading kernels issue one per
 * 256 paths and queue.  pos[k] is this thread's slot where pred[k]. */
template <int NQ>
__device__ __forceinline__ void block_append(const bool* pred, int32_t* const* counter, int* pos)
{
    __shared__ int cnt[NQ][PT_BLOCK / 64];
    __shared__ int base[NQ];
    const int w = threadIdx.x >> 6, lane = __lane_id();
    unsigned long long m[NQ];
#pragma unroll
    for (int k = 0; k < NQ; ++k) {
        m[k] = __ballot(pred[k]);
        if (lane == 0) cnt[k][w] = __popcll(m[k]);
    }
    __syncthreads();
    if (threadIdx.x < NQ) {
        int k = threadIdx.x, t = 0;
#pragma unroll
        for (int j = 0; j < PT_BLOCK / 64; ++j) t += cnt[k][j];
        base[k] = t > 0 ? atomicAdd(counter[k], t) : 0;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NQ; ++k) {
        int off = base[k];
        for (int j = 0; j < w; ++j) off += cnt[k][j];
        uint32_t below =
            __builtin_amdgcn_mbcnt_hi((uint32_t)(m[k] >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m[k], 0u));
        pos[k] = off + (int)below;
    }
    __syncthreads();  /* cnt and base are reused by the next call */
}

/* Dynamic work distribution of the persistent traversal kernels: the queue
 * is cut into RT_XCDS segments of consecutive positions, and a wave claims
 * positions from the segment of its block's XCD, in chunks of 64 (one atomic
 * per chunk), handing them out in lane order to the lanes that have finished
 * their previous ray.  Consecutive queue positions hold rays of neighbouring
 * pixels, so each XCD's L2 caches the part of the scene its own screen
 * region needs.  A wave that finds its segment drained moves on to the next
 * one; once it has found all of them drained it hands out positions >= n.
 * Call in wave-uniform control flow.  The chunk and segment are uniform
 * across the wave; returns the position for lanes with `want`. */
struct WaveFeed {
    int base = 0, used = 0, lim = 0;  /* the chunk: positions base+used .. base+lim-1 are unclaimed */
    int seg, left = RT_XCDS;          /* the segment claimed from; segments not yet found drained */
    __device__ WaveFeed() : seg(blockIdx.x % RT_XCDS) {}
};
__device__ __forceinline__ int feed_seg_start(int n, int s) { return (int)(((long long)n * s) / RT_XCDS); }
__device__ __forceinline__ int feed_take(WaveFeed& f, bool want, int32_t* counters, int n)
{
    const int lane = __lane_id();
    const unsigned long long below = (1ull << lane) - 1ull;
    int idx = 0x7fffffff;
    unsigned long long m = __ballot(want);
    while (m) {
        if (f.used == f.lim) {
            if (f.left == 0) {
                f.base = n, f.used = 0, f.lim = 64;  /* all drained */
            } else {
                const int lo = feed_seg_start(n, f.seg), len = feed_seg_start(n, f.seg + 1) - lo;
                int b = 0;
                if (lane == 0) b = atomicAdd(counters + f.seg * RT_CLAIM_STRIDE, 64);
                b = __builtin_amdgcn_readfirstlane(b);
                if (b < len) {
                    f.base = lo + b, f.used = 0, f.lim = min(64, len - b);
                } else {
                    f.seg = f.seg + 1 == RT_XCDS ? 0 : f.seg + 1;
                    f.left -= 1;
                    continue;
                }
            }
        }
        int avail = f.lim - f.used;
        int rank = __popcll(m & below);
        bool mine = ((m >> lane) & 1ull) && rank < avail;
        if (mine) idx = f.base + f.used + rank;
        f.used += min(__popcll(m), avail);
        m &= ~__ballot(mine);
    }
    return idx;
}
__device__ __forceinline__ void clear_claims(DevCounters* c)
{
    for (int x = 0; x < RT_XCDS; ++x) c->ext_next[x * RT_CLAIM_STRIDE] = c->sh_next[x * RT_CLAIM_STRIDE] = 0;
}

/* Bounce `depth` (queue qsel) is the one from which k_finish traces the
 * pass's remaining paths; the counters it reads change only between the
 * launches of a bounce, so every kernel of the bounce sees the same. */
__device__ __forceinline__ bool finish_here(const DevCounters* c, int qsel, int depth, FinishRule fr)
{
    return c->finished < 0 && depth >= fr.min_depth && c->q_count[qsel] <= fr.limit;
}
/* ... or a later one: the per-bounce kernels have nothing to do */
__device__ __forceinline__ bool bounce_finished(const DevCounters* c, int qsel, int depth, FinishRule fr)
{
    return c->finished >= 0 || finish_here(c, qsel, depth, fr);
}

/* ===========================================================================
 * Batch traversal kernels
 * ========================================================================= */
/* The block's LDS copy of the tree's first nodes (RT_TOP_NODES), which
 * trav_step reads for the rays at the top of the tree; call before any
 * traversal, in uniform control flow. */
__device__ __forceinline__ const char* stage_top_nodes(const DevScene& S, float4* top)
{
    const int nt = min(S.n_nodes, RT_TOP_NODES) * (RT_NODE_FLOATS / 4);
    const float4* src = reinterpret_cast<const float4*>(S.nodes);
    for (int i = threadIdx.x; i < nt; i += blockDim.x) top[i] = src[i];
    __syncthreads();
    return reinterpret_cast<const char*>(top);
}

template <bool ANY, bool COUNT>
__global__ __launch_bounds__(PT_BLOCK) void k_trace(DevScene S, const float4* __restrict__ o,
                                                   const float4* __restrict__ d, int64_t n, int32_t* prim, float* tout,
                                                   uint8_t* occ, int32_t* nn, int32_t* np)
{
    __shared__ int32_t stack[RT_STACK_LDS * PT_BLOCK];
    __shared__ float4 top_nodes[RT_TOP_NODES * (RT_NODE_FLOATS / 4)];
    const char* top = stage_top_nodes(S, top_nodes);
    LaneStack my(stack + threadIdx.x, PT_BLOCK, S.stack_ovf);
    for (int64_t i = (int64_t)blockIdx.x * PT_BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * PT_BLOCK) {
        float4 a = o[i], b = d[i];
        int cn = 0, cp = 0;
        TraceOut h = traverse<ANY, COUNT>(S, xyz(a), xyz(b), a.w, b.w, my, &cn, &cp, top);
        /* counting launches write only the counts (the hit outputs are null) */
        if (COUNT) {
            nn[i] = cn;
            np[i] = cp;
        } else if (ANY) {
            occ[i] = h.g >= 0 ? 1 : 0;
        } else {
            prim[i] = h.g >= 0 ? S.prim_id[h.g] : -1;
            tout[i] = h.g >= 0 ? h.t : __builtin_nanf("");
        }
    }
}

hipError_t launch_trace_closest(const HostDevScene& s, const float4* o, const float4* d, int64_t n, int32_t* prim,
                                float* t, int block, int grid, hipStream_t st)
{
    (void)block;
    hipLaunchKernelGGL((k_trace<false, false>), dim3(grid), dim3(PT_BLOCK), 0, st, as_dev(s), o, d, n, prim, t,
                       (uint8_t*)nullptr, (int32_t*)nullptr, (int32_t*)nullptr);
    return hipGetLastError();
}
hipError_t launch_trace_any(const HostDevScene& s, const float4* o, const float4* d, int64_t n, uint8_t* occ,
                            int block, int grid, hipStream_t st)
{
    (void)block;
    hipLaunchKernelGGL((k_trace<true, false>), dim3(grid), dim3(PT_BLOCK), 0, st, as_dev(s), o, d, n,
                       (int32_t*)nullptr, (float*)nullptr, occ, (int32_t*)nullptr, (int32_t*)nullptr);
    return hipGetLastError();
}
hipError_t launch_trace_count(const HostDevScene& s, const float4* o, const float4* d, int64_t n, int32_t* nodes,
                              int32_t* prims, int block, int grid, hipStream_t st, bool any)
{
    (void)block;
    if (any)
        hipLaunchKernelGGL((k_trace<true, true>), dim3(grid), dim3(PT_BLOCK), 0, st, as_dev(s), o, d, n,
                           (int32_t*)nullptr, (float*)nullptr, (uint8_t*)nullptr, nodes, prims);
    else
        hipLaunchKernelGGL((k_trace<false, true>), dim3(grid), dim3(PT_BLOCK), 0, st, as_dev(s), o, d, n,
                           (int32_t*)nullptr, (float*)nullptr, (uint8_t*)nullptr, nodes, prims);
    return hipGetLastError();
}

/* ---------------------------------------------------------------------------
 * Allocation-free closest hit (rt_trace_closest_device_noalloc): the stack
 * overflow of a launch of at most RT_NOALLOC_BLOCKS blocks lives in this
 * file-scope array instead of memory the context allocates.  Each GPU holds
 * its own copy (28 MiB) from the first use of the library on it until the
 * process ends.  Host code never passes `scratch` itself to a kernel: that
 * name is a host-side stand-in, not a device address.  rt_create looks up the
 * array's address on its device once (noalloc_scratch_init) and every launch
 * takes that pointer.  Nothing in it needs clearing: a stack entry is written
 * before it is read.  Calls that overlap on one GPU share the array and race.
 * ------------------------------------------------------------------------- */
static __device__ __align__(256) int32_t scratch[(size_t)RT_NOALLOC_BLOCKS * PT_BLOCK * RT_STACK_OVF];

static std::mutex g_scratch_mu;
static int32_t* g_scratch_ptr[RT_NOALLOC_MAX_DEVICES];

hipError_t noalloc_scratch_init(int device)
{
    if (device < 0 || device >= RT_NOALLOC_MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return hipSuccess;
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[device] = (int32_t*)p;
    return hipSuccess;
}

int32_t* noalloc_scratch(int device)
{
    if (device < 0 || device >= RT_NOALLOC_MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return g_scratch_ptr[device];
}

int64_t noalloc_scratch_entries() { return (int64_t)RT_NOALLOC_BLOCKS * PT_BLOCK * RT_STACK_OVF; }

/* ===========================================================================
 * Sampling and material helpers
 * ========================================================================= */

/* Cosine-power lobe about w, src/sampler.cpp:48-76: z = u0^(1/(n+1)),
 * r = sqrt(1 - u0^(2/(n+1))), phi = 2 pi u1, pdf = z^n (n+1) / 2pi, frame
 * vector chosen by the smallest component of w. */
__device__ __forceinline__ vec3 basis_v(vec3 w)
{
    vec3 v;
    if (fabsf(w.x) < fabsf(w.y) && fabsf(w.x) < fabsf(w.z))
        v = mk(0.f, w.z, -w.y);
    else if (fabsf(w.y) < fabsf(w.z))
        v = mk(w.z, 0.f, -w.x);
    else
        v = mk(w.y, -w.x, 0.f);
    return norm(v);
}
__device__ __forceinline__ vec3 cos_pow_dir(float u0, float u1, vec3 w, float n, float& pdf)
{
    w = norm(w);
    float z, r;
    if (n == 1.f) {
        z = sqrtf(u0);
        r = sqrtf(1.f - u0);
        pdf = z * RT_INV_PI;
    } else {
        z = powf(u0, 1.f / (n + 1.f));
        r = sqrtf(fmaxf(0.f, 1.f - powf(u0, 2.f / (n + 1.f))));
        pdf = powf(z, n) * (n + 1.f) * RT_INV_2PI;
    }
    float sp, cp;
    sincosf(2.f * RT_PI * u1, &sp, &cp);
    vec3 v = basis_v(w);
    vec3 u = cross(v, w);
    return norm(u * (cp * r) + v * (sp * r) + w * z);
}

/* Environment: lat-long lookup (nearest texel) and importance sampling by a
 * row-marginal / column-conditional table weighted by luminance x sin(theta). */
__device__ __forceinline__ int env_texel(const DevScene& S, vec3 d)
{
    float dy = fminf(1.f, fmaxf(-1.f, d.y));
    float theta = acosf(dy);
    float phi = atan2f(d.x, -d.z);
    float u = phi * RT_INV_2PI + 0.5f, v = theta * RT_INV_PI;
    int x = min(max((int)(u * S.env_w), 0), S.env_w - 1);
    int y = min(max((int)(v * S.env_h), 0), S.env_h - 1);
    return y * S.env_w + x;
}
__device__ __forceinline__ vec3 env_value(const DevScene& S, vec3 d)
{
    d = norm(d);
    if (S.env_kind == 1) return xyz(S.probe_rgb[rt_probe_texel(d.x, d.y, d.z, S.probe_w, S.probe_h)]);
    return xyz(S.env_rgb[env_texel(S, d)]);
}
__device__ __forceinline__ int cdf_pick(const float* cdf, int n, float u, float& frac)
{
    int lo = 0, hi = n; /* find largest i in [0,n) with cdf[i] <= u */
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (cdf[mid] <= u) lo = mid;
        else hi = mid;
    }
    float w = cdf[lo + 1] - cdf[lo];
    frac = w > 0.f ? (u - cdf[lo]) / w : 0.5f;
    frac = fminf(fmaxf(frac, 0.f), 0.99999988f);
    return lo;
}
__device__ __forceinline__ vec3 env_sample(const DevScene& S, float u0, float u1, float& pdf)
{
    float fy, fx;
    int y = cdf_pick(S.env_row_cdf, S.env_h, u0, fy);
    int x = cdf_pick(S.env_col_cdf + (size_t)y * (S.env_w + 1), S.env_w, u1, fx);
    float theta = ((float)y + fy) / (float)S.env_h * RT_PI;
    float phi = (((float)x + fx) / (float)S.env_w - 0.5f) * 2.f * RT_PI;
    float st, ct, sp, cp;
    sincosf(theta, &st, &ct);
    sincosf(phi, &sp, &cp);
    float p = S.env_texel_p[(size_t)y * S.env_w + x];
    pdf = st > 0.f ? p * (float)S.env_w * (float)S.env_h / (2.f * RT_PI * RT_PI * st) : 0.f;
    return mk(st * sp, ct, -st * cp);
}
/* Solid-angle density with which env_sample returns the unit direction d. */
__device__ __forceinline__ float env_pdf(const DevScene& S, vec3 d)
{
    float st = sqrtf(fmaxf(0.f, 1.f - d.y * d.y));
    float p = S.env_texel_p[env_texel(S, d)];
    return st > 0.f ? p * (float)S.env_w * (float)S.env_h / (2.f * RT_PI * RT_PI * st) : 0.f;
}
/* Power heuristic (exponent 2): weight of a sample drawn with density a when
 * a second strategy could have drawn it with density b.  Requires a > 0. */
__device__ __forceinline__ float mis_power(float a, float b)
{
    float r = b / a;
    return 1.f / (1.f + r * r);
}

/* Area-light intensity towards `direction` (from the light to the point),
 * src/arealight.cpp:53-59: I * max(0, cos) * |n| / |direction|^2. */
__device__ __forceinline__ vec3 area_intensity(const RtLight& L, vec3 direction)
{
    vec3 n = ld3(L.n);
    float c = dot(norm(direction), norm(n)) / dot(direction, direction) * len(n);
    c = fmaxf(c, 0.f);
    return ld3(L.I) * c;
}

__device__ __forceinline__ int sgn(float x) { return (0.f < x) - (x < 0.f); }

/* BlinnMaterial, src/blinnmaterial.cpp:34-62: normalised Blinn-Phong
 * distribution D and the Cook-Torrance shadowing term G, about the normal n
 * (facing wo). */
__device__ __forceinline__ float blinn_D(float e, vec3 wh, vec3 n) { return (e + 2.f) * RT_INV_2PI * powf(dot(wh, n), e); }
__device__ __forceinline__ float blinn_G(vec3 wi, vec3 wo, vec3 n)
{
    vec3 wh = norm(wi + wo);
    return fminf(1.f, 2.f * dot(n, wh) * fminf(dot(n, wo), dot(n, wi)) / dot(wo, wh));
}
/* BlinnMaterial::brdf(hit, direction): wo towards the viewer, wi = -direction
 * (the direction handed in points from the light to the surface), zero when
 * they lie on opposite sides.  For a direction sampled by outDirection this
 * is zero as well, which ends the throughput of the path there. */
__device__ __forceinline__ vec3 blinn_ref(const RtMaterial& m, vec3 n, vec3 din, vec3 dir)
{
    vec3 wo = -din, wi = -norm(dir);
    if (sgn(dot(n, wo)) != sgn(dot(n, wi))) return mk(0.f, 0.f, 0.f);
    if (dot(n, wo) < 0.f) n = -n;
    vec3 wh = norm(wi + wo);
    float s = blinn_D(m.exponent, wh, n) * blinn_G(wi, wo, n) / (4.f * dot(wo, n) * dot(wi, n));
    return ld3(m.kd) * RT_INV_PI + ld3(m.ks) * s;
}

/* TransparentMaterial::outDirection, src/transparentmaterial.cpp:34-72:
 * reflection with the unpolarised Fresnel reflectance as probability (and
 * always under total internal reflection), refraction otherwise; ior is the
 * refraction coefficient of the inside.  Throughput is unchanged (a
 * specular vertex).  Types follow the reference (the sines in double). */
__device__ __forceinline__ vec3 refractive_dir(vec3 d_in, vec3 n_in, float ior, float u)
{
    vec3 direction = norm(d_in), normal = norm(n_in);
    float ratio;
    if (dot(-normal, direction) < 0.f) {
        normal = -normal;
        ratio = ior;
    } else {
        ratio = 1.f / ior;
    }
    float cosi = dot(-normal, direction);
    double s2 = fmax(0.0, 1.0 - (double)(cosi * cosi));
    float sino = (float)((double)ratio * sqrt(s2));
    float coso = (float)sqrt(fmax(0.0, 1.0 - (double)(ratio * ratio) * s2));
    float rperp = (cosi - ratio * coso) / (cosi + ratio * coso);
    float rpara = (ratio * cosi - coso) / (ratio * cosi + coso);
    float refl = (float)(0.5 * (double)(rperp * rperp + rpara * rpara));
    if (sino > 1.f || (double)u < (double)refl) return reflect(direction, normal);
    return direction * ratio + normal * (ratio * cosi - (float)sgn(cosi) * coso);
}

/* ConeLight::getIntensity, src/conelight.cpp:31-45: lit when the angle
 * between the light's axis and the direction from the light to the point is
 * at most the opening angle (v[0]). */
__device__ __forceinline__ bool cone_admits(const RtLight& L, vec3 dir)
{
    float a = acosf(dot(norm(dir), norm(ld3(L.u))));
    return 0.f <= a && a <= L.v[0];
}

/* Reference BRDF value, src/diffusematerial.cpp:25-28, src/phongmaterial.cpp:31-49.
 * `n` is the unflipped unit normal, `din` the unit incoming ray direction,
 * `dir` the direction from the light towards the point. */
__device__ __forceinline__ vec3 brdf_ref(const RtMaterial& m, vec3 n, vec3 din, vec3 dir)
{
    if (m.type == RT_MAT_DIFFUSE) return ld3(m.kd) * RT_INV_PI;
    if (m.type == RT_MAT_PHONG) {
        vec3 refl = reflect(dir, n);
        float dp = fmaxf(dot(norm(refl), -din), 0.f);
        float spec = (m.exponent + 2.f) * RT_INV_2PI * powf(dp, m.exponent);
        return ld3(m.kd) * RT_INV_PI + ld3(m.ks) * spec;
    }
    if (m.type == RT_MAT_BLINN) return blinn_ref(m, n, din, dir);
    return mk(0.f, 0.f, 0.f);
}
/* Phong specular term alone, src/phongmaterial.cpp:42-49. */
__device__ __forceinline__ vec3 phong_spec_ref(const RtMaterial& m, vec3 n, vec3 din, vec3 dir)
{
    vec3 refl = reflect(dir, n);
    float dp = fmaxf(dot(norm(refl), -din), 0.f);
    return ld3(m.ks) * ((m.exponent + 2.f) * RT_INV_2PI * powf(dp, m.exponent));
}

/* Unbiased-mode BSDF helpers: Phong as a luminance-weighted mixture of the
 * diffuse lobe and the cos^n lobe about the mirror direction. */
__device__ __forceinline__ vec3 lobe_dir(float u0, float u1, vec3 axis, float expo)
{
    float z = expo == 1.f ? sqrtf(u0) : powf(u0, 1.f / (expo + 1.f));
    float r = sqrtf(fmaxf(0.f, 1.f - z * z));
    float sp, cp;
    sincosf(2.f * RT_PI * u1, &sp, &cp);
    vec3 v = basis_v(axis);
    vec3 t = cross(v, axis);
    return norm(t * (r * cp) + v * (r * sp) + axis * z);
}
/* Blinn in the unbiased estimators: the same BRDF, with wo and wi both
 * leaving the surface; sampled as a luminance-weighted mixture of the
 * diffuse lobe and the reference's cos^n half-vector lobe. */
__device__ __forceinline__ vec3 bsdf_eval(const RtMaterial& m, vec3 wo, vec3 wi, vec3 nf)
{
    if (m.type == RT_MAT_DIFFUSE) return ld3(m.kd) * RT_INV_PI;
    if (m.type == RT_MAT_PHONG) {
        vec3 r = reflect(-wo, nf);
        float c = fmaxf(0.f, dot(r, wi));
        float s = (m.exponent + 2.f) * RT_INV_2PI * powf(c, m.exponent);
        return ld3(m.kd) * RT_INV_PI + ld3(m.ks) * s;
    }
    if (m.type == RT_MAT_BLINN) {
        float co = dot(wo, nf), ci = dot(wi, nf);
        if (!(co > 0.f) || !(ci > 0.f)) return ld3(m.kd) * RT_INV_PI;
        vec3 wh = norm(wo + wi);
        float s = blinn_D(m.exponent, wh, nf) * blinn_G(wi, wo, nf) / (4.f * co * ci);
        return ld3(m.kd) * RT_INV_PI + ld3(m.ks) * s;
    }
    return mk(0.f, 0.f, 0.f);
}
__device__ __forceinline__ float bsdf_pdf(const RtMaterial& m, vec3 wo, vec3 wi, vec3 nf)
{
    float cs = dot(wi, nf);
    if (cs <= 0.f) return 0.f;
    if (m.type == RT_MAT_DIFFUSE) return cs * RT_INV_PI;
    if (m.type == RT_MAT_PHONG) {
        float pd = m.rho_d / (m.rho_d + m.rho_s);
        vec3 r = reflect(-wo, nf);
        float c = fmaxf(0.f, dot(r, wi));
        return pd * cs * RT_INV_PI + (1.f - pd) * (m.exponent + 1.f) * RT_INV_2PI * powf(c, m.exponent);
    }
    if (m.type == RT_MAT_BLINN) {
        float pd = m.rho_d / (m.rho_d + m.rho_s);
        vec3 wh = norm(wo + wi);
        float ch = fmaxf(0.f, dot(wh, nf));
        float ph = (m.exponent + 1.f) * RT_INV_2PI * powf(ch, m.exponent) / (4.f * fmaxf(dot(wo, wh), 1e-4f));
        return pd * cs * RT_INV_PI + (1.f - pd) * ph;
    }
    return 0.f;
}
__device__ __forceinline__ bool bsdf_sample(const RtMaterial& m, vec3 wo, vec3 nf, float u0, float u1, float u2,
                                            vec3& wi)
{
    if (m.type == RT_MAT_DIFFUSE) {
        wi = lobe_dir(u0, u1, nf, 1.f);
        return true;
    }
    if (m.type == RT_MAT_PHONG) {
        if (!(m.rho_d + m.rho_s > 0.f)) return false;
        float pd = m.rho_d / (m.rho_d + m.rho_s);
        wi = u2 < pd ? lobe_dir(u0, u1, nf, 1.f) : lobe_dir(u0, u1, reflect(-wo, nf), m.exponent);
        return dot(wi, nf) > 0.f;
    }
    if (m.type == RT_MAT_BLINN) {
        if (!(m.rho_d + m.rho_s > 0.f)) return false;
        float pd = m.rho_d / (m.rho_d + m.rho_s);
        if (u2 < pd) {
            wi = lobe_dir(u0, u1, nf, 1.f);
        } else {
            vec3 wh = lobe_dir(u0, u1, nf, m.exponent);
            float c = dot(wo, wh);
            if (!(c > 0.f)) return false;
            wi = wh * (2.f * c) - wo;
        }
        return dot(wi, nf) > 0.f;
    }
    return false;
}

/* Packed path bookkeeping in info.w: stored vertices (8 bits), path length
 * (8 bits), specular-chain flag (bit 16). */
__device__ __forceinline__ int pk(int k, int L, int spec) { return k | (L << 8) | (spec << 16); }

/* Random numbers of group g of the path in slot p: its pixel sample's
 * Philox stream, or its primary-sample vector when the paths are driven by
 * Markov chains. */
__device__ __forceinline__ U4 path_u4(const PtParams& P, int4 info, int p, uint32_t g)
{
    if (P.pss) {
        const float4 v = P.pss[(size_t)g * P.pss_stride + p];
        U4 r;
        r.u[0] = v.x;
        r.u[1] = v.y;
        r.u[2] = v.z;
        r.u[3] = v.w;
        return r;
    }
    return rng4(P.seed, (uint32_t)info.z, (uint32_t)info.x, (uint32_t)info.y, g);
}

/* Next-event sample of the unbiased estimators: picks light floor(u2 N) and a
 * direction towards it.  Returns false when nothing can arrive.  `le` is the
 * light's contribution over the direction density (without the 1/N of the
 * light choice), `pl` the solid-angle density of the direction including the
 * 1/N (0 for a point light, which no other strategy can hit). */
__device__ __forceinline__ bool light_sample(const DevScene& S, vec3 x, const U4& gl, vec3& wi, vec3& le,
                                             float& dist, float& pl)
{
    int li = min((int)(gl.u[2] * (float)S.n_lights), S.n_lights - 1);
    const RtLight& Lt = S.lights[li];
    const float nl = (float)S.n_lights;
    dist = RT_INF;
    pl = 0.f;
    if (Lt.type == RT_LIGHT_AREA) {
        vec3 y = ld3(Lt.p) + ld3(Lt.u) * gl.u[0] + ld3(Lt.v) * gl.u[1];
        vec3 dv = y - x;
        float d2 = dot(dv, dv);
        dist = sqrtf(d2);
        wi = dv * (1.f / dist);
        vec3 ln = ld3(Lt.n);
        float an = sqrtf(dot(ln, ln));
        float cl = -dot(wi, ln) / an;
        if (!(cl > 0.f)) return false;
        le = ld3(Lt.I) * (cl * an / d2);
        pl = d2 / (cl * an) / nl;
        return true;
    }
    if (Lt.type == RT_LIGHT_POINT || Lt.type == RT_LIGHT_CONE) {
        vec3 dv = ld3(Lt.p) - x;
        if (Lt.type == RT_LIGHT_CONE && !cone_admits(Lt, -dv)) return false;
        float d2 = dot(dv, dv);
        dist = sqrtf(d2);
        wi = dv * (1.f / dist);
        le = ld3(Lt.I) * (1.f / d2);
        return true;
    }
    float pdf;
    wi = env_sample(S, gl.u[0], gl.u[1], pdf);
    if (!(pdf > 0.f)) return false;
    le = env_value(S, wi) * (1.f / pdf);
    pl = pdf / nl;
    return true;
}

/* Emission reached by an extension ray of the unbiased estimators: the
 * environment at a miss, or a front-facing area light.  Camera rays and rays
 * leaving a mirror (`spec`) count it in full.  A ray sampled at a
 * non-specular vertex carries its sampling density in alpha.w and counts the
 * emission with the power-heuristic weight against next-event estimation,
 * which could have picked the same direction.  `d` is the unit ray direction
 * and `t` the hit distance.  Returns true when the path ends here. */
__device__ __forceinline__ bool add_emission(const DevScene& S, const PathState& ps, int p, int g, bool spec,
                                             float t, vec3 d, float4 al)
{
    vec3 e;
    float pl = 0.f;
    if (g < 0) {
        if (S.env_light < 0) return true;
        e = env_value(S, d);
        if (!spec) pl = env_pdf(S, d) / (float)S.n_lights;
    } else {
        int m = S.mat[g];
        if (m >= 0) return false;
        const RtLight& Lt = S.lights[-m - 1];
        float dn = dot(d, ld3(Lt.n));
        if (!(dn < 0.f)) return true;
        e = ld3(Lt.I);
        /* area density t^2 / (A cos) with A cos = -dot(d, u x v) */
        if (!spec) pl = t * t / (-dn) / (float)S.n_lights;
    }
    float w = spec ? 1.f : mis_power(al.w, pl);
    e = mul3(xyz(al), e) * w;
    float4 r = ps.radiance[p];
    ps.radiance[p] = make_float4(r.x + e.x, r.y + e.y, r.z + e.z, 0.f);
    return true;
}

/* The q-th pixel of a tile's walk.  A tile of whole image rows is walked in
 * 8x8 pixel blocks (bands of 8 rows, blocks left to right; a narrower last
 * block and a shorter last band keep row order), so the camera rays of a
 * wave start from one compact block rather than a row and share more BVH
 * nodes.  Any other tile is walked in row order.
 * This is a bijection on [0, npix); the random stream is keyed by the pixel,
 * so the film does not depend on the order. */
__device__ __forceinline__ int tile_pixel(const DevTile& T, const DevView& V, int q)
{
    const int W = V.W;
    if (W < 8 || T.pix0 % W != 0 || T.npix % W != 0) return T.pix0 + q;
    const int band = 8 * W, g = q / band, r = q - g * band;
    const int rows = min(8, T.npix / W - 8 * g);
    int x, y;
    if (rows < 8) {
        x = r % W;
        y = r / W;
    } else {
        const int full = W - W % 8;  /* columns covered by whole 8x8 blocks */
        if (r < 8 * full) {
            x = (r >> 6) * 8 + (r & 7);
            y = (r >> 3) & 7;
        } else {
            const int rr = r - 8 * full, w = W - full;
            x = full + rr % w;
            y = rr / w;
        }
    }
    return T.pix0 + (8 * g + y) * W + x;
}

/* ===========================================================================
 * generate
 * ========================================================================= */
/* Camera::getRay (src/camera.cpp:52-57) with the reference's arithmetic: the
 * image-plane point in double (QPointF), rounded to float, then the
 * direction x u + y v - w with every product and sum rounded on its own. */
__device__ __forceinline__ vec3 camera_dir(const DevView& V, double px, double py)
{
    const float x = (float)__dadd_rn((double)V.l, __ddiv_rn(__dmul_rn((double)(V.r - V.l), px + 0.5), (double)V.W));
    const float y = (float)__dadd_rn((double)V.t, __ddiv_rn(__dmul_rn((double)(V.b - V.t), py + 0.5), (double)V.H));
    vec3 d;
    d.x = __fadd_rn(__fadd_rn(__fmul_rn(x, V.u[0]), __fmul_rn(y, V.v[0])), -V.w[0]);
    d.y = __fadd_rn(__fadd_rn(__fmul_rn(x, V.u[1]), __fmul_rn(y, V.v[1])), -V.w[1]);
    d.z = __fadd_rn(__fadd_rn(__fmul_rn(x, V.u[2]), __fmul_rn(y, V.v[2])), -V.w[2]);
    return d;
}

/* QVector3D::normalized(): the length in double, components rounded to float */
__device__ __forceinline__ vec3 norm_qt(vec3 a)
{
    const double l = __dadd_rn(__dadd_rn(__dmul_rn(a.x, a.x), __dmul_rn(a.y, a.y)), __dmul_rn(a.z, a.z));
    if (l == 0.0) return mk(0.f, 0.f, 0.f);
    const double r = sqrt(l);
    return mk((float)(a.x / r), (float)(a.y / r), (float)(a.z / r));
}
__global__ __launch_bounds__(PT_BLOCK) void k_generate(PathState ps, const DevView* __restrict__ views,
                                                      const DevTile* __restrict__ tiles,
                                                      const DevPass* __restrict__ passes, PtParams P)
{
    const DevPass pass = passes[ps.cnt->pass];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ps.cnt->q_count[0] = pass.slots;
        ps.cnt->q_count[1] = 0;
        ps.cnt->shadow_count = 0;
        clear_claims(ps.cnt);
        ps.cnt->finished = -1;
        ps.cnt->camera_rays += (unsigned long long)pass.slots;
    }
    for (int slot = blockIdx.x * PT_BLOCK + threadIdx.x; slot < pass.slots; slot += gridDim.x * PT_BLOCK) {
        int lo = pass.tile0, hi = pass.tile1;
        while (hi - lo > 1) {
            int mid = (lo + hi) >> 1;
            if (tiles[mid].slot0 <= slot) lo = mid;
            else hi = mid;
        }
        const DevTile T = tiles[lo];
        const DevView& V = views[T.view];
        int j = slot - T.slot0;
        int pixel = tile_pixel(T, V, j / T.ns);
        int s = T.s0 + j % T.ns;
        U4 g0 = rng4(P.seed, (uint32_t)T.view, (uint32_t)pixel, (uint32_t)s, 0u);
        int a = s / V.spp_y, b = s % V.spp_y;
        /* src/jitteredsampler.cpp:23-35 and src/perpixelrenderer.cpp:79-84 in
         * double (QPointF) */
        double sx = ((double)g0.u[0] + a) / V.spp_x, sy = ((double)g0.u[1] + b) / V.spp_y;
        vec3 d = camera_dir(V, (double)(pixel % V.W) + sx, (double)(pixel / V.W) + sy);
        if (P.mode != RT_MODE_COMPAT) d = norm_qt(d);
        int L = P.max_depth;
        if (P.mode == RT_MODE_COMPAT && P.rr_prob > 0.f) {
            int nb = 0;
            if (P.rr_prob < 1.f) {
                float q = floorf(logf(1.f - g0.u[2]) / logf(1.f - P.rr_prob));
                nb = q > 1000.f ? 1000 : (int)q;
            }
            L = min(P.rr_start + nb, P.max_depth);
        }
        ps.ray_o[slot] = make_float4(V.eye[0], V.eye[1], V.eye[2], RT_EPS);
        ps.ray_d[slot] = make_float4(d.x, d.y, d.z, RT_INF);
        ps.alpha[slot] = make_float4(1.f, 1.f, 1.f, 0.f);
        ps.radiance[slot] = make_float4(0.f, 0.f, 0.f, 0.f);
        if (P.mode == RT_MODE_COMPAT) ps.pend_c[slot] = make_float4(0.f, 0.f, 0.f, 0.f); /* only compat keeps one */
        ps.info[slot] = make_int4(pixel, s, T.view, pk(0, L, 1));
        ps.queue[0][slot] = slot;
    }
}

hipError_t launch_generate(const PathState& ps, const DevView* views, const DevTile* tiles, const DevPass* passes,
                           PtParams p, int grid, hipStream_t st)
{
    hipLaunchKernelGGL(k_generate, dim3(grid), dim3(PT_BLOCK), 0, st, ps, views, tiles, passes, p);
    return hipGetLastError();
}

/* Camera paths from primary samples (MetropolisSample::cameraPathFromSample,
 * src/metropolissample.cpp:120-127): the image point is group 0's (x, y)
 * times the resolution, the path length is drawn from its z as in
 * k_generate.  The pixel recorded in info is MetropolisSample::getPos. */
__global__ __launch_bounds__(PT_BLOCK) void k_generate_pss(PathState ps, const DevView* __restrict__ view, int n,
                                                          PtParams P, int pixel_units)
{
    const DevView V = *view;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ps.cnt->q_count[0] = n;
        ps.cnt->q_count[1] = 0;
        ps.cnt->shadow_count = 0;
        clear_claims(ps.cnt);
        ps.cnt->finished = -1;
        ps.cnt->camera_rays += (unsigned long long)n;
    }
    for (int slot = blockIdx.x * PT_BLOCK + threadIdx.x; slot < n; slot += gridDim.x * PT_BLOCK) {
        const float4 g0 = P.pss[slot];
        double px = pixel_units ? (double)g0.x : (double)g0.x * V.W;
        double py = pixel_units ? (double)g0.y : (double)g0.y * V.H;
        vec3 d = camera_dir(V, px, py);
        int L = P.max_depth;
        if (P.rr_prob > 0.f) {
            int nb = 0;
            if (P.rr_prob < 1.f) {
                float q = floorf(logf(1.f - g0.z) / logf(1.f - P.rr_prob));
                nb = q > 1000.f ? 1000 : (int)q;
            }
            L = min(P.rr_start + nb, P.max_depth);
        }
        int ix = pixel_units ? (int)g0.x : (int)((double)g0.x * V.W);
        int iy = pixel_units ? (int)g0.y : (int)((double)g0.y * V.H);
        ix = max(0, min(ix, V.W - 1));
        iy = max(0, min(iy, V.H - 1));
        ps.ray_o[slot] = make_float4(V.eye[0], V.eye[1], V.eye[2], RT_EPS);
        ps.ray_d[slot] = make_float4(d.x, d.y, d.z, RT_INF);
        ps.alpha[slot] = make_float4(1.f, 1.f, 1.f, 0.f);
        ps.radiance[slot] = make_float4(0.f, 0.f, 0.f, 0.f);
        ps.pend_c[slot] = make_float4(0.f, 0.f, 0.f, 0.f);
        ps.info[slot] = make_int4(iy * V.W + ix, 0, 0, pk(0, L, 1));
        ps.queue[0][slot] = slot;
    }
}

hipError_t launch_generate_pss(const PathState& ps, const DevView* view, int n, PtParams p, int pixel_units,
                               int grid, hipStream_t st)
{
    hipLaunchKernelGGL(k_generate_pss, dim3(grid), dim3(PT_BLOCK), 0, st, ps, view, n, p, pixel_units);
    return hipGetLastError();
}

/* Paths that start with caller rays (UniDiPathTracingIntegrator::integrate(
 * const Ray&, ...), src/unidipathtracingintegrator.cpp:34-46): ray j of the
 * batch fills slot j; its random numbers are the Philox stream of (seed,
 * view 0, index key0 + base + j, sample 0), group 0 giving the path-length
 * uniform as for a camera sample.  The reference estimator takes the ray as
 * given; the unbiased one normalises the direction (Qt arithmetic) and
 * scales the range by the same length, so the range covers the same points. */
__global__ __launch_bounds__(PT_BLOCK) void k_generate_rays(PathState ps, const float4* __restrict__ o,
                                                           const float4* __restrict__ d, int n, PtParams P,
                                                           uint32_t key0)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ps.cnt->q_count[0] = n;
        ps.cnt->q_count[1] = 0;
        ps.cnt->shadow_count = 0;
        clear_claims(ps.cnt);
        ps.cnt->finished = -1;
        ps.cnt->camera_rays += (unsigned long long)n;
    }
    for (int slot = blockIdx.x * PT_BLOCK + threadIdx.x; slot < n; slot += gridDim.x * PT_BLOCK) {
        float4 a = o[slot], b = d[slot];
        const uint32_t key = key0 + (uint32_t)slot;
        int L = P.max_depth;
        if (P.mode == RT_MODE_COMPAT && P.rr_prob > 0.f) {
            const U4 g0 = rng4(P.seed, 0u, key, 0u, 0u);
            int nb = 0;
            if (P.rr_prob < 1.f) {
                float q = floorf(logf(1.f - g0.u[2]) / logf(1.f - P.rr_prob));
                nb = q > 1000.f ? 1000 : (int)q;
            }
            L = min(P.rr_start + nb, P.max_depth);
        }
        if (P.mode != RT_MODE_COMPAT) {
            const double l = sqrt(__dadd_rn(__dadd_rn(__dmul_rn(b.x, b.x), __dmul_rn(b.y, b.y)), __dmul_rn(b.z, b.z)));
            if (l > 0.0) {
                b = make_float4((float)(b.x / l), (float)(b.y / l), (float)(b.z / l),
                                b.w == RT_INF ? RT_INF : (float)(b.w * l));
                a.w = (float)(a.w * l);
            }
        }
        ps.ray_o[slot] = a;
        ps.ray_d[slot] = b;
        ps.alpha[slot] = make_float4(1.f, 1.f, 1.f, 0.f);
        ps.radiance[slot] = make_float4(0.f, 0.f, 0.f, 0.f);
        ps.pend_c[slot] = make_float4(0.f, 0.f, 0.f, 0.f);
        ps.info[slot] = make_int4((int)key, 0, 0, pk(0, L, 1));
        ps.queue[0][slot] = slot;
    }
}

hipError_t launch_generate_rays(const PathState& ps, const float4* o, const float4* d, int n, PtParams p,
                                uint32_t key0, int grid, hipStream_t st)
{
    hipLaunchKernelGGL(k_generate_rays, dim3(grid), dim3(PT_BLOCK), 0, st, ps, o, d, n, p, key0);
    return hipGetLastError();
}

/* rgb[3 j ..] = the radiance of path slot j (with a pending next-event
 * estimate of the reference estimator committed) */
__global__ __launch_bounds__(PT_BLOCK) void k_ray_values(PathState ps, int n, float* __restrict__ rgb)
{
    for (int j = blockIdx.x * PT_BLOCK + threadIdx.x; j < n; j += gridDim.x * PT_BLOCK) {
        const float4 L = ps.radiance[j], c = ps.pend_c[j];
        rgb[3 * (size_t)j] = L.x + c.x;
        rgb[3 * (size_t)j + 1] = L.y + c.y;
        rgb[3 * (size_t)j + 2] = L.z + c.z;
    }
}

hipError_t launch_ray_values(const PathState& ps, int n, float* rgb, int grid, hipStream_t st)
{
    hipLaunchKernelGGL(k_ray_values, dim3(grid), dim3(PT_BLOCK), 0, st, ps, n, rgb);
    return hipGetLastError();
}

/* ===========================================================================
 * extend: closest hit for every live path
 * ========================================================================= */
/* Persistent traversal with per-lane ray replacement: a lane whose ray is
 * finished stores its result and takes the next unclaimed ray at once
 * (WaveFeed), instead of idling until the slowest ray of its wave finishes;
 * waves claim rays dynamically, so the launch ends evenly.
 * The traversal stack (LaneStack) keeps RT_STACK_LDS entries per lane in
 * LDS: 16 KB per block, and 8.6 KB for the top nodes: room for the 6
 * blocks per CU that the registers allow. */
template <bool COUNT, bool OBJ>
__global__ __launch_bounds__(PT_BLOCK) RT_TRAV_OCCUPANCY void k_extend(DevScene S, PathState ps, int qsel, int depth,
                                                                       FinishRule fr)
{
    if (bounce_finished(ps.cnt, qsel, depth, fr)) return;
    __shared__ int32_t stack[RT_STACK_LDS * PT_BLOCK];
    __shared__ float4 top_nodes[RT_TOP_NODES * (RT_NODE_FLOATS / 4)];
    const char* top = stage_top_nodes(S, top_nodes);
    LaneStack my(stack + threadIdx.x, PT_BLOCK, S.stack_ovf);
    const int n = ps.cnt->q_count[qsel];
    const int32_t* q = ps.queue[qsel];
    int cn = 0, cp = 0, ct = 0, p = 0;
    bool live = false, more = true;
    WaveFeed feed;
    Trav T;
    while (true) {
        bool want = !live && more;
        if (__ballot(want)) {
            int i = feed_take(feed, want, ps.cnt->ext_next, n);
            if (want) {
                if (i < n) {
                    p = q[i];
                    float4 a = ps.ray_o[p], b = ps.ray_d[p];
                    trav_init(S, T, xyz(a), xyz(b), a.w, b.w);
                    if (OBJ && a.w <= b.w) obj_closest(S, T, b.w);
                    live = true;
                } else {
                    more = false;
                }
            }
        }
        if (!__ballot(live)) break;  /* queue drained and every lane idle */
        if (live && (T.cur < 0 || trav_step<false, COUNT>(S, T, my, cn, cp, top, &ct))) {
            ps.hit[p] = make_float4(T.g >= 0 ? T.best : RT_INF, __int_as_float(T.g), T.u, T.v);
            live = false;
        }
    }
    if (COUNT && (cn || cp)) {
        atomicAdd(&ps.cnt->node_visits, (unsigned long long)cn);
        atomicAdd(&ps.cnt->prim_tests, (unsigned long long)cp);
        atomicAdd(&ps.cnt->node_visits_lds, (unsigned long long)ct);
    }
}

hipError_t launch_extend(const HostDevScene& s, const PathState& ps, int qsel, int depth, FinishRule fr, int count,
                         int grid, hipStream_t st)
{
    /* scenes without analytic objects run the variant without their test */
    const bool obj = s.n_obj > 0;
    if (count && obj)
        hipLaunchKernelGGL((k_extend<true, true>), dim3(grid), dim3(PT_BLOCK), 0, st, as_dev(s), ps, qsel, depth, fr);
    else if (count)
        hipLaunchKernelGGL((k_extend<true, false>), dim3(grid), dim3(PT_BLOCK), 0, st, as_dev(s), ps, qsel, depth, fr);
    else if (obj)
        hipLaunchKernelGGL((k_extend<false, true>), dim3(grid), dim3(PT_BLOCK), 0, st, as_dev(s), ps, qsel, depth, fr);
    else
        hipLaunchKernelGGL((k_extend<false, false>), dim3(grid), dim3(PT_BLOCK), 0, st, as_dev(s), ps, qsel, depth,
                           fr);
    return hipGetLastError();
}

/* ===========================================================================
 * shade, reference ("compat") estimator
 *
 * Restates Renderer::createPath (src/renderer.cpp:69-106) and
 * UniDiPathTracingIntegrator::integrate(Path) (src/unidipathtracingintegrator.cpp:48-107)
 * incrementally.  The reference first builds the whole path and then adds,
 * for every stored vertex, either its next-event estimate or - when the next
 * stored vertex is an emitter - that emitter's emission.  Here the
 * next-event estimate of the latest stored vertex is held "pending" in the
 * path state (with that vertex's position, normal, direction, material and
 * throughput) until the next stored vertex is known: it is committed when
 * that vertex is an ordinary surface or the path ends, and replaced by the
 * emission term when it is an emitter.
 * ========================================================================= */
__device__ void shade_compat(const DevScene& S, const PathState& ps, const PtParams& P, int depth, int p,
                             bool& cont, bool& shadow, float4& so, float4& sd, float4& sc)
{
    cont = false;
    shadow = false;
    int4 info = ps.info[p];
    int k = info.w & 0xff, L = (info.w >> 8) & 0xff;
    float4 h = ps.hit[p];
    int g = __float_as_int(h.y);
    float4 ro = ps.ray_o[p], rd = ps.ray_d[p];
    vec3 o = xyz(ro), d = xyz(rd);
    float4 pc = ps.pend_c[p];

    bool emitter = false;
    vec3 E, edir;  /* emission, and direction light -> previous stored vertex */
    if (g < 0) {
        if (S.env_light >= 0 && rd.w == RT_INF) {
            /* src/environmentmap.cpp:73-83: background hit at t = inf,
             * point = (unit direction, w = 0) */
            emitter = true;
            vec3 dn = norm(d);
            E = env_value(S, dn);
            edir = -dn;
        } else {
            return; /* miss: pending estimate stays and is committed at accumulate */
        }
    } else {
        int m = S.mat[g];
        if (m < 0) {
            const RtLight& Lt = S.lights[-m - 1];
            emitter = true;
            vec3 x = o + h.x * d;
            E = area_intensity(Lt, -d);
            edir = xyz(ps.pend_x[p]) - x;
        }
    }
    float4 al = ps.alpha[p];
    vec3 alpha = xyz(al);
    if (emitter) {
        if (k == 0) {
            /* first stored vertex emits: return its emission (:55-58) */
            if (!is_zero(E)) ps.radiance[p] = make_float4(E.x, E.y, E.z, 0.f);
            return;
        }
        /* replace the pending estimate of the previous stored vertex (:70-82) */
        float4 px = ps.pend_x[p];
        float4 pn4 = ps.pend_n[p], pd4 = ps.pend_d[p], pa4 = ps.pend_a[p];
        vec3 pn = xyz(pn4), pd = xyz(pd4), pa = xyz(pa4);
        RtMaterial pm = S.mats[__float_as_int(px.w)];
        if (pm.type == RT_MAT_TEXTURE) {  /* the albedo at that vertex, kept in the w components */
            pm.type = RT_MAT_DIFFUSE;
            pm.kd[0] = pn4.w, pm.kd[1] = pd4.w, pm.kd[2] = pa4.w;
        }
        vec3 f = brdf_ref(pm, pn, pd, edir);
        float in_cos = dot(-norm(edir), pn), out_cos = dot(-pd, pn);
        float4 rad = ps.radiance[p];
        if (sgn(in_cos) == sgn(out_cos)) {
            vec3 c = mul3(mul3(pa, f), E) * fabsf(in_cos);
            rad.x += c.x; rad.y += c.y; rad.z += c.z;
        }
        ps.radiance[p] = rad;
        ps.pend_c[p] = make_float4(0.f, 0.f, 0.f, 0.f);
        return;
    }

    vec3 x = o + h.x * d;
    const RtMaterial M = hit_material(S, g, x, h.z, h.w);
    vec3 n = hit_normal(S, g, x, h.z, h.w);
    U4 gs = path_u4(P, info, p, 1u + 2u * depth);

    if (M.type == RT_MAT_MIRROR || M.type == RT_MAT_REFRACTIVE) {
        /* specular vertex: not stored; extend the path length once (:87-91) */
        if (depth == L - 1 && L < P.max_depth) ++L;
        vec3 out = M.type == RT_MAT_MIRROR ? reflect(d, n) : refractive_dir(d, n, M.exponent, gs.u[0]);
        if (depth + 1 < L) {
            ps.ray_o[p] = make_float4(x.x, x.y, x.z, RT_EPS);
            ps.ray_d[p] = make_float4(out.x, out.y, out.z, RT_INF);
            ps.info[p] = make_int4(info.x, info.y, info.z, pk(k, L, 0));
            cont = true;
        }
        return;
    }

    /* ordinary stored vertex: the previous pending estimate is committed */
    float4 rad = ps.radiance[p];
    rad.x += pc.x; rad.y += pc.y; rad.z += pc.z;
    ps.radiance[p] = rad;
    ps.pend_c[p] = make_float4(0.f, 0.f, 0.f, 0.f);

    vec3 dn = norm(d);
    /* BSDF sample, Material::outDirection */
    float pdf = 0.f;
    vec3 brdf = mk(0.f, 0.f, 0.f), out = mk(0.f, 0.f, 0.f);
    if (M.type == RT_MAT_DIFFUSE) {
        vec3 nf = dot(n, dn) > 0.f ? -n : n;
        out = cos_pow_dir(gs.u[0], gs.u[1], nf, 1.f, pdf);
        brdf = ld3(M.kd) * RT_INV_PI;
    } else if (M.type == RT_MAT_PHONG) {
        vec3 nrm = dot(d, n) > 0.f ? -n : n;
        vec3 spec_dir = reflect(d, nrm);
        float u0 = gs.u[0];
        if (u0 < M.rho_d) {
            out = cos_pow_dir(u0 / M.rho_d, gs.u[1], nrm, 1.f, pdf);
            brdf = ld3(M.kd) * RT_INV_PI;
        } else if (u0 < M.rho_d + M.rho_s) {
            out = cos_pow_dir((u0 - M.rho_d) / M.rho_s, gs.u[1], spec_dir, M.exponent, pdf);
            if (dot(out, nrm) < 1.f) pdf = 0.f;
            brdf = phong_spec_ref(M, n, dn, out);
        }
    } else if (M.type == RT_MAT_BLINN) {
        /* BlinnMaterial::outDirection, src/blinnmaterial.cpp:64-84 */
        vec3 wo = -dn;
        vec3 nrm = dot(n, wo) < 0.f ? -n : n;
        vec3 wh = cos_pow_dir(gs.u[0], gs.u[1], nrm, M.exponent, pdf);
        if (dot(wo, wh) < 0.f) {
            pdf = 0.f;
        } else {
            out = -reflect(wo, wh);
            pdf /= 4.f * fmaxf(dot(wo, wh), 1e-4f);
            brdf = blinn_ref(M, n, dn, out);
        }
    }

    /* next-event estimate for this stored vertex k (Light::getIntensity) */
    if (S.n_lights > 0) {
        U4 gl = path_u4(P, info, p, 2u + 2u * k);
        int li = min((int)(gl.u[2] * (float)S.n_lights), S.n_lights - 1);
        const RtLight& Lt = S.lights[li];
        vec3 dir, Li = mk(0.f, 0.f, 0.f);
        bool has_ray = false;
        if (Lt.type == RT_LIGHT_AREA) {
            vec3 y = ld3(Lt.p) + ld3(Lt.u) * gl.u[0] + ld3(Lt.v) * gl.u[1];
            dir = x - y;
            float dl = len(dir);
            so = make_float4(x.x, x.y, x.z, RT_EPS);
            vec3 sdir = -norm(dir);
            sd = make_float4(sdir.x, sdir.y, sdir.z, dl - RT_EPS);
            Li = area_intensity(Lt, dir);
            has_ray = true;
        } else if (Lt.type == RT_LIGHT_POINT || Lt.type == RT_LIGHT_CONE) {
            dir = x - ld3(Lt.p);
            float dl = len(dir);
            so = make_float4(x.x, x.y, x.z, RT_EPS);
            vec3 sdir = -norm(dir);
            sd = make_float4(sdir.x, sdir.y, sdir.z, dl - RT_EPS);
            Li = ld3(Lt.I) * (1.f / dot(dir, dir));
            has_ray = Lt.type == RT_LIGHT_POINT || cone_admits(Lt, dir);
        } else {
            /* src/environmentmap.cpp:40-55: cosine sample about the unflipped
             * normal, shadow ray over negative t, value / pdf looked up at
             * the negated sample.  The backwards ray is traced forwards. */
            float epdf;
            vec3 s = cos_pow_dir(gl.u[0], gl.u[1], n, 1.f, epdf);
            dir = -s;
            so = make_float4(x.x, x.y, x.z, RT_EPS);
            sd = make_float4(s.x, s.y, s.z, RT_INF);
            Li = env_value(S, dir) * (1.f / epdf);
            has_ray = true;
        }
        vec3 f = brdf_ref(M, n, dn, dir);
        float in_cos = dot(-norm(dir), n), out_cos = dot(-dn, n);
        if (has_ray && sgn(in_cos) == sgn(out_cos)) {
            vec3 c = mul3(mul3(alpha, f), Li) * fabsf(in_cos);
            if (!is_zero(c)) {
                shadow = true;
                sc = make_float4(c.x, c.y, c.z, __int_as_float(p));
            }
        }
    }
    ps.pend_x[p] = make_float4(x.x, x.y, x.z, __int_as_float(S.mat[g]));
    ps.pend_n[p] = make_float4(n.x, n.y, n.z, M.kd[0]);
    ps.pend_d[p] = make_float4(dn.x, dn.y, dn.z, M.kd[1]);
    ps.pend_a[p] = make_float4(al.x, al.y, al.z, M.kd[2]);
    ++k;
    if (pdf == 0.f || is_zero(alpha)) {
        ps.info[p] = make_int4(info.x, info.y, info.z, pk(k, L, 0));
        return;
    }
    float c = fabsf(dot(norm(out), n));
    alpha = mul3(alpha, brdf) * (c / pdf);
    /* src/renderer.cpp:53-61,101: a termination probability of 0 means
     * "no roulette" and a compensation factor of 1 */
    if (depth > P.rr_start) alpha = alpha * (1.f / (P.rr_prob > 0.f ? P.rr_prob : 1.f));
    ps.alpha[p] = make_float4(alpha.x, alpha.y, alpha.z, 0.f);
    ps.info[p] = make_int4(info.x, info.y, info.z, pk(k, L, 0));
    if (depth + 1 < L) {
        ps.ray_o[p] = make_float4(x.x, x.y, x.z, RT_EPS);
        ps.ray_d[p] = make_float4(out.x, out.y, out.z, RT_INF);
        cont = true;
    }
}

/* ===========================================================================
 * shade, unbiased estimator: next-event estimation and BSDF sampling combined
 * by multiple importance sampling (power heuristic), compensated roulette.
 * The next-event weight is 1 at the last vertex, whose BSDF ray is not
 * traced, and for point lights.
 * ========================================================================= */
__device__ void shade_unbiased(const DevScene& S, const PathState& ps, const PtParams& P, int depth, int p,
                               bool& cont, bool& shadow, float4& so, float4& sd, float4& sc)
{
    cont = false;
    shadow = false;
    int4 info = ps.info[p];
    int spec = (info.w >> 16) & 1;
    float4 h = ps.hit[p];
    int g = __float_as_int(h.y);
    float4 ro = ps.ray_o[p], rd = ps.ray_d[p];
    vec3 o = xyz(ro), d = xyz(rd);
    float4 al = ps.alpha[p];
    vec3 alpha = xyz(al);
    if (add_emission(S, ps, p, g, spec, h.x, d, al)) return;
    vec3 x = o + h.x * d;
    const RtMaterial M = hit_material(S, g, x, h.z, h.w);
    vec3 n = hit_normal(S, g, x, h.z, h.w);
    vec3 nf = dot(n, d) > 0.f ? -n : n;
    if (M.type == RT_MAT_MIRROR || M.type == RT_MAT_REFRACTIVE) {
        if (depth + 1 < P.max_depth) {
            vec3 out = reflect(d, n);
            if (M.type == RT_MAT_REFRACTIVE) {
                U4 gr = rng4(P.seed, (uint32_t)info.z, (uint32_t)info.x, (uint32_t)info.y, 1u + 2u * depth);
                out = refractive_dir(d, n, M.exponent, gr.u[0]);
            }
            ps.ray_o[p] = make_float4(x.x, x.y, x.z, RT_EPS);
            ps.ray_d[p] = make_float4(out.x, out.y, out.z, RT_INF);
            ps.info[p] = make_int4(info.x, info.y, info.z, pk(0, 0, 1));
            cont = true;
        }
        return;
    }
    if (M.type != RT_MAT_DIFFUSE && M.type != RT_MAT_PHONG && M.type != RT_MAT_BLINN) return;
    vec3 wo = -d;
    const bool last = depth + 1 >= P.max_depth;
    if (S.n_lights > 0) {
        U4 gl = rng4(P.seed, (uint32_t)info.z, (uint32_t)info.x, (uint32_t)info.y, 2u + 2u * depth);
        vec3 wi, le;
        float dist, pl;
        if (light_sample(S, x, gl, wi, le, dist, pl)) {
            float cx = dot(wi, nf);
            if (cx > 0.f) {
                float w = (pl > 0.f && !last) ? mis_power(pl, bsdf_pdf(M, wo, wi, nf)) : 1.f;
                vec3 f = bsdf_eval(M, wo, wi, nf);
                vec3 c = mul3(mul3(alpha, f), le) * (cx * (float)S.n_lights * w);
                if (!is_zero(c)) {
                    shadow = true;
                    so = make_float4(x.x, x.y, x.z, RT_EPS);
                    sd = make_float4(wi.x, wi.y, wi.z, dist == RT_INF ? RT_INF : dist - RT_EPS);
                    sc = make_float4(c.x, c.y, c.z, __int_as_float(p));
                }
            }
        }
    }
    if (last) return;
    /* drawn after next-event estimation: fewer values live across it */
    U4 gb = rng4(P.seed, (uint32_t)info.z, (uint32_t)info.x, (uint32_t)info.y, 1u + 2u * depth);
    vec3 wi;
    if (!bsdf_sample(M, wo, nf, gb.u[0], gb.u[1], gb.u[2], wi)) return;
    float pdf = bsdf_pdf(M, wo, wi, nf);
    if (!(pdf > 0.f)) return;
    vec3 f = bsdf_eval(M, wo, wi, nf);
    alpha = mul3(alpha, f) * (dot(wi, nf) / pdf);
    if (depth + 1 >= P.rr_start) {
        if (gb.u[3] < P.rr_prob) return;
        alpha = alpha * (1.f / (1.f - P.rr_prob));
    }
    ps.alpha[p] = make_float4(alpha.x, alpha.y, alpha.z, pdf);
    ps.ray_o[p] = make_float4(x.x, x.y, x.z, RT_EPS);
    ps.ray_d[p] = make_float4(wi.x, wi.y, wi.z, RT_INF);
    ps.info[p] = make_int4(info.x, info.y, info.z, pk(0, 0, 0));
    cont = true;
}

/* ===========================================================================
 * shade, guided estimator: the unbiased estimator with the BSDF sample
 * replaced by a one-sample mixture of the learned vMF proposal (probability
 * alpha) and the BSDF (1 - alpha), weighted by the mixture pdf
 * alpha q(w) + (1 - alpha) p_bsdf(w), so the estimator stays unbiased for
 * any network output.  That mixture density is also the one next-event
 * estimation is weighted against.  Optionally records (position, normal, wo, w,
 * p_bsdf(w), throughput) for online training; the record is "closed" at the
 * next shade of the path, when the radiance gathered so far (including this
 * vertex's next-event estimate) is known.
 * ========================================================================= */
struct GuideRecord {
    vec3 x, nf, wo, wi;
    float pb, la;
};

__device__ void shade_guided(const DevScene& S, const PathState& ps, const GuideState& G, const PtParams& P,
                             int depth, int p, int qi, bool& cont, bool& shadow, float4& so, float4& sd, float4& sc,
                             bool& want_rec, GuideRecord& rec)
{
    cont = false;
    shadow = false;
    want_rec = false;
    if (P.guide_train) {
        int open = G.open_rec[p];
        if (open >= 0) {
            float4 r = ps.radiance[p];
            float4 a = G.rec_aux[open];
            G.rec_aux[open] = make_float4(r.x, r.y, r.z, a.w);
            G.open_rec[p] = -1;
        }
    }
    int4 info = ps.info[p];
    int spec = (info.w >> 16) & 1;
    float4 h = ps.hit[p];
    int g = __float_as_int(h.y);
    float4 ro = ps.ray_o[p], rd = ps.ray_d[p];
    vec3 o = xyz(ro), d = xyz(rd);
    float4 al = ps.alpha[p];
    vec3 alpha = xyz(al);
    if (add_emission(S, ps, p, g, spec, h.x, d, al)) return;
    vec3 x = o + h.x * d;
    const RtMaterial M = hit_material(S, g, x, h.z, h.w);
    vec3 n = hit_normal(S, g, x, h.z, h.w);
    vec3 nf = dot(n, d) > 0.f ? -n : n;
    if (M.type == RT_MAT_MIRROR || M.type == RT_MAT_REFRACTIVE) {
        if (depth + 1 < P.max_depth) {
            vec3 out = reflect(d, n);
            if (M.type == RT_MAT_REFRACTIVE) {
                U4 gr = rng4(P.seed, (uint32_t)info.z, (uint32_t)info.x, (uint32_t)info.y, 1u + 2u * depth);
                out = refractive_dir(d, n, M.exponent, gr.u[0]);
            }
            ps.ray_o[p] = make_float4(x.x, x.y, x.z, RT_EPS);
            ps.ray_d[p] = make_float4(out.x, out.y, out.z, RT_INF);
            ps.info[p] = make_int4(info.x, info.y, info.z, pk(0, 0, 1));
            cont = true;
        }
        return;
    }
    if (M.type != RT_MAT_DIFFUSE && M.type != RT_MAT_PHONG && M.type != RT_MAT_BLINN) return;
    vec3 wo = -d;
    const bool last = depth + 1 >= P.max_depth;
    /* the network is not evaluated for the last bounce; its outputs are
     * stored by queue position */
    const _Float16* go = G.out + (size_t)qi * GD_PARAMS;
    float ag = last ? 0.f : (float)go[GD_ALPHA_OFF];
    if (S.n_lights > 0) {
        U4 gl = rng4(P.seed, (uint32_t)info.z, (uint32_t)info.x, (uint32_t)info.y, 2u + 2u * depth);
        vec3 wi, le;
        float dist, pl;
        if (light_sample(S, x, gl, wi, le, dist, pl)) {
            float cx = dot(wi, nf);
            if (cx > 0.f) {
                /* the competing strategy is the one-sample mixture */
                float w = 1.f;
                if (pl > 0.f && !last)
                    w = mis_power(pl, ag * guide_pdf(go, wi) + (1.f - ag) * bsdf_pdf(M, wo, wi, nf));
                vec3 f = bsdf_eval(M, wo, wi, nf);
                vec3 c = mul3(mul3(alpha, f), le) * (cx * (float)S.n_lights * w);
                if (!is_zero(c)) {
                    shadow = true;
                    so = make_float4(x.x, x.y, x.z, RT_EPS);
                    sd = make_float4(wi.x, wi.y, wi.z, dist == RT_INF ? RT_INF : dist - RT_EPS);
                    sc = make_float4(c.x, c.y, c.z, __int_as_float(p));
                }
            }
        }
    }
    if (last) return;
    /* drawn after next-event estimation: fewer values live across it */
    U4 gb = rng4(P.seed, (uint32_t)info.z, (uint32_t)info.x, (uint32_t)info.y, 1u + 2u * depth);
    U4 gx = rng4(P.seed, (uint32_t)info.z, (uint32_t)info.x, (uint32_t)info.y, 1u + 2u * depth, 1u);
    vec3 wi;
    if (gx.u[0] < ag) {
        wi = guide_sample(go, gx.u[1], gb.u[0], gb.u[1]);
    } else if (!bsdf_sample(M, wo, nf, gb.u[0], gb.u[1], gb.u[2], wi)) {
        return;
    }
    float cs = dot(wi, nf);
    if (!(cs > 0.f)) return;
    float pb = bsdf_pdf(M, wo, wi, nf);
    float pdf = ag * guide_pdf(go, wi) + (1.f - ag) * pb;
    if (!(pdf > 0.f)) return;
    vec3 f = bsdf_eval(M, wo, wi, nf);
    float la = lum(alpha);
    alpha = mul3(alpha, f) * (cs / pdf);
    if (depth + 1 >= P.rr_start) {
        if (gb.u[3] < P.rr_prob) return;
        alpha = alpha * (1.f / (1.f - P.rr_prob));
    }
    if (!(lum(alpha) < 3.0e38f)) return;
    ps.alpha[p] = make_float4(alpha.x, alpha.y, alpha.z, pdf);
    ps.ray_o[p] = make_float4(x.x, x.y, x.z, RT_EPS);
    ps.ray_d[p] = make_float4(wi.x, wi.y, wi.z, RT_INF);
    ps.info[p] = make_int4(info.x, info.y, info.z, pk(0, 0, 0));
    cont = true;
    if (P.guide_train && gx.u[2] < *G.keep && la > 0.f) {
        want_rec = true;
        rec.x = x;
        rec.nf = nf;
        rec.wo = wo;
        rec.wi = wi;
        rec.pb = pb;
        rec.la = la;
    }
}

__global__ __launch_bounds__(PT_BLOCK) RT_SHADE_GUIDED_OCCUPANCY void k_shade_guided(DevScene S, PathState ps, GuideState G, PtParams P,
                                                          int depth, int qsel)
{
    const int n = ps.cnt->q_count[qsel];
    const int32_t* q = ps.queue[qsel];
    int32_t* qn = ps.queue[qsel ^ 1];
    for (int base = blockIdx.x * PT_BLOCK; base < n; base += gridDim.x * PT_BLOCK) {
        int i = base + threadIdx.x;
        bool cont = false, shadow = false, want = false;
        float4 so, sd, sc;
        GuideRecord rec;
        int p = -1;
        if (i < n) {
            p = q[i];
            shade_guided(S, ps, G, P, depth, p, i, cont, shadow, so, sd, sc, want, rec);
        }
        const bool pred[3] = {cont, shadow, want && P.guide_train};
        int32_t* const ctr[3] = {&ps.cnt->q_count[qsel ^ 1], &ps.cnt->shadow_count, G.rec_count};
        int pos[3];
        block_append<3>(pred, ctr, pos);
        if (cont) qn[pos[0]] = p;
        if (shadow) {
            ps.sh_o[pos[1]] = so;
            ps.sh_d[pos[1]] = sd;
            ps.sh_c[pos[1]] = sc;
        }
        if (pred[2]) {
            int ar = pos[2];
            if (ar < G.rec_cap) {
                float* r = G.rec + (size_t)ar * GD_REC;
                r[0] = rec.x.x; r[1] = rec.x.y; r[2] = rec.x.z;
                r[3] = rec.nf.x; r[4] = rec.nf.y; r[5] = rec.nf.z;
                r[6] = rec.wo.x; r[7] = rec.wo.y; r[8] = rec.wo.z;
                r[9] = rec.wi.x; r[10] = rec.wi.y; r[11] = rec.wi.z;
                r[12] = rec.pb; r[13] = 0.f; r[14] = rec.la; r[15] = 0.f;
                float4 rad = ps.radiance[p];
                G.rec_aux[ar] = make_float4(rad.x, rad.y, rad.z, __int_as_float(p));
                G.open_rec[p] = ar;
            }
        }
    }
}

hipError_t launch_shade_guided(const HostDevScene& s, const PathState& ps, const GuideState& g, PtParams p,
                               int depth, int qsel, int grid, hipStream_t st)
{
    hipLaunchKernelGGL(k_shade_guided, dim3(grid), dim3(PT_BLOCK), 0, st, as_dev(s), ps, g, p, depth, qsel);
    return hipGetLastError();
}

template <int MODE>
__global__ __launch_bounds__(PT_BLOCK) RT_SHADE_OCCUPANCY void k_shade(DevScene S, PathState ps, PtParams P, int depth, int qsel,
                                                                       FinishRule fr)
{
    if (bounce_finished(ps.cnt, qsel, depth, fr)) return;
    const int n = ps.cnt->q_count[qsel];
    const int32_t* q = ps.queue[qsel];
    int32_t* qn = ps.queue[qsel ^ 1];
    for (int base = blockIdx.x * PT_BLOCK; base < n; base += gridDim.x * PT_BLOCK) {
        int i = base + threadIdx.x;
        bool cont = false, shadow = false;
        float4 so, sd, sc;
        int p = -1;
        if (i < n) {
            p = q[i];
            if (MODE == RT_MODE_COMPAT)
                shade_compat(S, ps, P, depth, p, cont, shadow, so, sd, sc);
            else
                shade_unbiased(S, ps, P, depth, p, cont, shadow, so, sd, sc);
        }
        const bool pred[2] = {cont, shadow};
        int32_t* const ctr[2] = {&ps.cnt->q_count[qsel ^ 1], &ps.cnt->shadow_count};
        int pos[2];
        block_append<2>(pred, ctr, pos);
        if (cont) qn[pos[0]] = p;
        if (shadow) {
            ps.sh_o[pos[1]] = so;
            ps.sh_d[pos[1]] = sd;
            ps.sh_c[pos[1]] = sc;
        }
    }
}

hipError_t launch_shade(const HostDevScene& s, const PathState& ps, PtParams p, int depth, int qsel, FinishRule fr,
                        int grid, hipStream_t st)
{
    if (p.mode == RT_MODE_COMPAT)
        hipLaunchKernelGGL(k_shade<RT_MODE_COMPAT>, dim3(grid), dim3(PT_BLOCK), 0, st, as_dev(s), ps, p, depth, qsel,
                           fr);
    else
        hipLaunchKernelGGL(k_shade<RT_MODE_UNBIASED>, dim3(grid), dim3(PT_BLOCK), 0, st, as_dev(s), ps, p, depth,
                           qsel, fr);
    return hipGetLastError();
}

/* ===========================================================================
 * shadow: any-hit test of next-event rays
 * ========================================================================= */
template <bool OBJ, bool COUNT>
__global__ __launch_bounds__(PT_BLOCK) RT_TRAV_OCCUPANCY void k_shadow(DevScene S, PathState ps, int compat, int qsel,
                                                                       int depth, FinishRule fr)
{
    if (bounce_finished(ps.cnt, qsel, depth, fr)) return;
    __shared__ int32_t stack[RT_STACK_LDS * PT_BLOCK];
    __shared__ float4 top_nodes[RT_TOP_NODES * (RT_NODE_FLOATS / 4)];
    const char* top = stage_top_nodes(S, top_nodes);
    LaneStack my(stack + threadIdx.x, PT_BLOCK, S.stack_ovf);
    const int n = ps.cnt->shadow_count;
    int cn = 0, cp = 0, ct = 0, item = 0;
    bool live = false, more = true;
    WaveFeed feed;
    Trav T;
    while (true) {
        bool want = !live && more;
        if (__ballot(want)) {
            int i = feed_take(feed, want, ps.cnt->sh_next, n);
            if (want) {
                if (i < n) {
                    float4 a = ps.sh_o[i], b = ps.sh_d[i];
                    item = i;
                    trav_init(S, T, xyz(a), xyz(b), a.w, b.w);
                    if (OBJ && a.w <= b.w && obj_any(S, T.o, T.d, a.w, b.w)) {
                        T.g = 0;  /* occluded by an object: no traversal */
                        T.cur = -1;
                    }
                    live = true;
                } else {
                    more = false;
                }
            }
        }
        if (!__ballot(live)) break;  /* queue drained and every lane idle */
        if (live && (T.cur < 0 || trav_step<true, COUNT>(S, T, my, cn, cp, top, &ct))) {
            if (T.g < 0) {
                /* unoccluded: add the contribution to its path */
                float4 c = ps.sh_c[item];
                int p = __float_as_int(c.w);
                if (compat) {
                    ps.pend_c[p] = make_float4(c.x, c.y, c.z, 0.f);
                } else {
                    float4 r = ps.radiance[p];
                    ps.radiance[p] = make_float4(r.x + c.x, r.y + c.y, r.z + c.z, 0.f);
                }
            }
            live = false;
        }
    }
    if (COUNT && (cn || cp)) {
        atomicAdd(&ps.cnt->sh_node_visits, (unsigned long long)cn);
        atomicAdd(&ps.cnt->sh_prim_tests, (unsigned long long)cp);
        atomicAdd(&ps.cnt->sh_node_visits_lds, (unsigned long long)ct);
    }
}

hipError_t launch_shadow(const HostDevScene& s, const PathState& ps, PtParams p, int qsel, int depth, FinishRule fr,
                         int count, int grid, hipStream_t st)
{
    int compat = p.mode == RT_MODE_COMPAT ? 1 : 0;
    const bool obj = s.n_obj > 0;
    if (count && obj)
        hipLaunchKernelGGL((k_shadow<true, true>), dim3(grid), dim3(PT_BLOCK), 0, st, as_dev(s), ps, compat, qsel,
                           depth, fr);
    else if (count)
        hipLaunchKernelGGL((k_shadow<false, true>), dim3(grid), dim3(PT_BLOCK), 0, st, as_dev(s), ps, compat, qsel,
                           depth, fr);
    else if (obj)
        hipLaunchKernelGGL((k_shadow<true, false>), dim3(grid), dim3(PT_BLOCK), 0, st, as_dev(s), ps, compat, qsel,
                           depth, fr);
    else
        hipLaunchKernelGGL((k_shadow<false, false>), dim3(grid), dim3(PT_BLOCK), 0, st, as_dev(s), ps, compat, qsel,
                           depth, fr);
    return hipGetLastError();
}

/* ===========================================================================
 * finish: the last bounces of the compat and unbiased estimators in one launch
 * ========================================================================= */
/* Late bounces hold few paths (1.5% of the camera rays from the fifth
 * bounce on in a 1024^2, 64 spp render of a 1M-triangle mesh), and a launch
 * of the separate kernels costs about the latency chain of its slowest ray
 * whatever its size, 300-450 us each there.  Here a lane takes a path from
 * the queue and runs it to its end: closest hit, shade (the same per-path
 * functions as k_shade, so the path's state changes in the same order), its
 * shadow ray, the next bounce; a lane whose path has ended takes the next
 * one (WaveFeed).  A lane whose closest hit is found waits until its wave's
 * waiting lanes are at least as many as those still traversing (or none
 * traverses), so the wave runs the shading code once for many lanes rather
 * than once per lane: 4.0 -> 3.1 ms per 1024^2 x 64 spp render of the
 * 1M-triangle mesh, Metropolis 815 -> 930 Mrays/s.  The film is the one the
 * separate kernels give.  Rays are counted as extension and shadow rays and,
 * separately, as the ones traced here. */
__device__ __forceinline__ unsigned wave_sum(unsigned v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

template <int MODE, bool OBJ>
__global__ __launch_bounds__(PT_BLOCK) RT_FINISH_OCCUPANCY void k_finish(DevScene S, PathState ps, PtParams P,
                                                                         int depth0, int qsel, FinishRule fr)
{
    if (!finish_here(ps.cnt, qsel, depth0, fr)) return;
    __shared__ int32_t stack[RT_STACK_LDS * PT_BLOCK];
    __shared__ float4 top_nodes[RT_TOP_NODES * (RT_NODE_FLOATS / 4)];
    const char* top = stage_top_nodes(S, top_nodes);
    LaneStack my(stack + threadIdx.x, PT_BLOCK, S.stack_ovf);
    const int n = ps.cnt->q_count[qsel];
    const int32_t* q = ps.queue[qsel];
    int cn = 0, cp = 0, p = 0, depth = 0;
    bool live = false, more = true, in_shadow = false, cont = false, to_shade = false;
    float4 sc = make_float4(0.f, 0.f, 0.f, 0.f);
    unsigned n_ext = 0, n_sh = 0;
    WaveFeed feed;
    Trav T;
    auto start_extend = [&]() {
        const float4 a = ps.ray_o[p], b = ps.ray_d[p];
        trav_init(S, T, xyz(a), xyz(b), a.w, b.w);
        if (OBJ && a.w <= b.w) obj_closest(S, T, b.w);
        if (depth > 0) ++n_ext;  /* bounce 0 traces the camera rays, counted by k_generate */
    };
    while (true) {
        bool want = !live && more;
        if (__ballot(want)) {
            int i = feed_take(feed, want, ps.cnt->ext_next, n);
            if (want) {
                if (i < n) {
                    p = q[i];
                    depth = depth0;
                    in_shadow = false;
                    start_extend();
                    live = true;
                } else {
                    more = false;
                }
            }
        }
        if (!__ballot(live)) break;  /* queue drained and every lane idle */
        const bool done = live && !to_shade &&
                          (T.cur < 0 || (in_shadow ? trav_step<true, false>(S, T, my, cn, cp, top)
                                                   : trav_step<false, false>(S, T, my, cn, cp, top)));
        if (done && !in_shadow) {
            ps.hit[p] = make_float4(T.g >= 0 ? T.best : RT_INF, __int_as_float(T.g), T.u, T.v);
            to_shade = true;
        }
        /* a wave runs the shading code once for all the lanes that wait for
         * it: when they are at least as many as the lanes still traversing,
         * or no lane traverses */
        const unsigned long long waiting = __ballot(to_shade), tracing = __ballot(live && !to_shade);
        const bool shade_now = to_shade && (tracing == 0 || __popcll(waiting) >= __popcll(tracing));
        if (done && in_shadow) {
            if (T.g < 0) {  /* unoccluded: as k_shadow */
                if (MODE == RT_MODE_COMPAT) {
                    ps.pend_c[p] = make_float4(sc.x, sc.y, sc.z, 0.f);
                } else {
                    const float4 r = ps.radiance[p];
                    ps.radiance[p] = make_float4(r.x + sc.x, r.y + sc.y, r.z + sc.z, 0.f);
                }
            }
            in_shadow = false;
            if (cont) {
                ++depth;
                start_extend();
            } else {
                live = false;
            }
        } else if (shade_now) {
            to_shade = false;
            bool shadow = false;
            float4 so, sd;
            if (MODE == RT_MODE_COMPAT)
                shade_compat(S, ps, P, depth, p, cont, shadow, so, sd, sc);
            else
                shade_unbiased(S, ps, P, depth, p, cont, shadow, so, sd, sc);
            if (shadow) {
                in_shadow = true;
                ++n_sh;
                trav_init(S, T, xyz(so), xyz(sd), so.w, sd.w);
                if (OBJ && so.w <= sd.w && obj_any(S, T.o, T.d, so.w, sd.w)) {
                    T.g = 0;  /* occluded by an object: no traversal */
                    T.cur = -1;
                }
            } else if (cont) {
                ++depth;
                start_extend();
            } else {
                live = false;
            }
        }
    }
    n_ext = wave_sum(n_ext);
    n_sh = wave_sum(n_sh);
    if (__lane_id() == 0 && (n_ext | n_sh)) {
        atomicAdd(&ps.cnt->extension_rays, (unsigned long long)n_ext);
        atomicAdd(&ps.cnt->finish_rays, (unsigned long long)n_ext);
        atomicAdd(&ps.cnt->shadow_rays, (unsigned long long)n_sh);
        atomicAdd(&ps.cnt->finish_shadow_rays, (unsigned long long)n_sh);
    }
}

hipError_t launch_finish(const HostDevScene& s, const PathState& ps, PtParams p, int depth, int qsel, FinishRule fr,
                         int grid, hipStream_t st)
{
    const bool obj = s.n_obj > 0;
    const bool compat = p.mode == RT_MODE_COMPAT;
    if (compat && obj)
        hipLaunchKernelGGL((k_finish<RT_MODE_COMPAT, true>), dim3(grid), dim3(PT_BLOCK), 0, st, as_dev(s), ps, p,
                           depth, qsel, fr);
    else if (compat)
        hipLaunchKernelGGL((k_finish<RT_MODE_COMPAT, false>), dim3(grid), dim3(PT_BLOCK), 0, st, as_dev(s), ps, p,
                           depth, qsel, fr);
    else if (obj)
        hipLaunchKernelGGL((k_finish<RT_MODE_UNBIASED, true>), dim3(grid), dim3(PT_BLOCK), 0, st, as_dev(s), ps, p,
                           depth, qsel, fr);
    else
        hipLaunchKernelGGL((k_finish<RT_MODE_UNBIASED, false>), dim3(grid), dim3(PT_BLOCK), 0, st, as_dev(s), ps, p,
                           depth, qsel, fr);
    return hipGetLastError();
}

/* End of a bounce: count the rays the bounce traced (its queue length and
 * its shadow-ray count; the queue of bounce 0 holds the camera rays, counted
 * by k_generate), then clear the consumed queue and the work counters.  At
 * the finishing bounce, record it (k_finish counted its own rays). */
__global__ void k_advance(PathState ps, int qsel, int depth, FinishRule fr)
{
    DevCounters* c = ps.cnt;
    if (finish_here(c, qsel, depth, fr)) {
        c->finished = depth;
    } else if (c->finished < 0) {
        if (depth > 0) c->extension_rays += (unsigned long long)c->q_count[qsel];
        c->shadow_rays += (unsigned long long)c->shadow_count;
        c->busy_bounces += 1;
    }
    c->q_count[qsel] = 0;
    c->shadow_count = 0;
    clear_claims(c);
}
hipError_t launch_advance(const PathState& ps, int qsel, int depth, FinishRule fr, hipStream_t st)
{
    hipLaunchKernelGGL(k_advance, dim3(1), dim3(1), 0, st, ps, qsel, depth, fr);
    return hipGetLastError();
}

/* ===========================================================================
 * accumulate: film += (radiance + pending) / spp, samples of a pixel added in
 * increasing sample order (src/perpixelrenderer.cpp:92) - deterministic and
 * independent of how samples were split into passes.
 * ========================================================================= */
/* film += the pass's samples of each pixel, in sample order (the
 * reference's film[i][j] += s/spp, src/perpixelrenderer.cpp:92).  A
 * pixel's samples occupy consecutive slots, so a thread reading its own
 * would touch a different cache line than its neighbours at every load; the
 * block instead reads its pixels' samples eight at a time, eight lanes to
 * each 128-byte group of slots (one L2 line of each array), into LDS, and
 * each thread then adds its pixel's terms in order. */
#define ACC_CHUNK 8
__global__ __launch_bounds__(PT_BLOCK) void k_accumulate(PathState ps, const DevView* __restrict__ views,
                                                        const DevTile* __restrict__ tiles,
                                                        const DevPass* __restrict__ passes, int compat)
{
    __shared__ int s_base[PT_BLOCK], s_ns[PT_BLOCK];
    __shared__ float s_inv[PT_BLOCK];
    __shared__ float s_term[PT_BLOCK][ACC_CHUNK][3];
    __shared__ int s_max;
    const DevPass pass = passes[ps.cnt->pass];
    const int t = threadIdx.x;
    for (int j0 = blockIdx.x * PT_BLOCK; j0 < pass.pixels; j0 += gridDim.x * PT_BLOCK) {
        const int j = j0 + t;
        float* f = nullptr;
        float r = 0.f, g = 0.f, b = 0.f;
        int ns = 0;
        __syncthreads(); /* the previous round's readers of the block's tables are done */
        if (t == 0) s_max = 0;
        if (j < pass.pixels) {
            int lo = pass.tile0, hi = pass.tile1;
            while (hi - lo > 1) {
                int mid = (lo + hi) >> 1;
                if (tiles[mid].acc0 <= j) lo = mid;
                else hi = mid;
            }
            const DevTile T = tiles[lo];
            const DevView& V = views[T.view];
            const int local = j - T.acc0; /* pixel of the tile's walk */
            f = V.film + 3 * (size_t)tile_pixel(T, V, local);
            r = f[0], g = f[1], b = f[2];
            ns = T.ns;
            s_base[t] = T.slot0 + local * T.ns;
            s_inv[t] = V.inv_spp;
        }
        s_ns[t] = ns;
        __syncthreads();
        if (ns > 0) atomicMax(&s_max, ns);
        __syncthreads();
        const int maxns = s_max;
        for (int c0 = 0; c0 < maxns; c0 += ACC_CHUNK) {
#pragma unroll
            for (int k = 0; k < ACC_CHUNK; ++k) {
                const int e = k * PT_BLOCK + t, pe = e / ACC_CHUNK, se = c0 + e % ACC_CHUNK;
                if (se < s_ns[pe]) {
                    const int slot = s_base[pe] + se;
                    const float4 L = ps.radiance[slot];
                    const float4 cc = compat ? ps.pend_c[slot] : make_float4(0.f, 0.f, 0.f, 0.f);
                    const float inv = s_inv[pe];
                    s_term[pe][e % ACC_CHUNK][0] = (L.x + cc.x) * inv;
                    s_term[pe][e % ACC_CHUNK][1] = (L.y + cc.y) * inv;
                    s_term[pe][e % ACC_CHUNK][2] = (L.z + cc.z) * inv;
                }
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < ACC_CHUNK; ++k)
                if (c0 + k < ns) {
                    r += s_term[t][k][0];
                    g += s_term[t][k][1];
                    b += s_term[t][k][2];
                }
            __syncthreads();
        }
        if (f) {
            f[0] = r;
            f[1] = g;
            f[2] = b;
        }
    }
}
hipError_t launch_accumulate(const PathState& ps, const DevView* views, const DevTile* tiles,
                             const DevPass* passes, int compat, int grid, hipStream_t st)
{
    hipLaunchKernelGGL(k_accumulate, dim3(grid), dim3(PT_BLOCK), 0, st, ps, views, tiles, passes, compat);
    return hipGetLastError();
}

__global__ void k_pass_end(PathState ps) { ps.cnt->pass += 1; }
hipError_t launch_pass_end(const PathState& ps, hipStream_t st)
{
    hipLaunchKernelGGL(k_pass_end, dim3(1), dim3(1), 0, st, ps);
    return hipGetLastError();
}

__global__ void k_reset(PathState ps)
{
    DevCounters* c = ps.cnt;
    c->q_count[0] = c->q_count[1] = 0;
    c->shadow_count = 0;
    c->pass = 0;
    c->finished = -1;
    c->busy_bounces = 0;
    clear_claims(c);
    c->camera_rays = c->extension_rays = c->shadow_rays = 0ull;
    c->node_visits = c->prim_tests = 0ull;
    c->sh_node_visits = c->sh_prim_tests = 0ull;
    c->node_visits_lds = c->sh_node_visits_lds = 0ull;
    c->finish_rays = c->finish_shadow_rays = 0ull;
}
hipError_t launch_reset_counters(const PathState& ps, hipStream_t st)
{
    hipLaunchKernelGGL(k_reset, dim3(1), dim3(1), 0, st, ps);
    return hipGetLastError();
}
