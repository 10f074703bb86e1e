/*
 * rt_guide.hip — gfx950 kernels of the path-guiding network.
 *
 * Inference (k_guide_net) uses v_mfma_f32_16x16x32_f16 in the transposed
 * form Z[out x samples] = W[out x in] * X[in x samples].  A wave carries 64
 * samples as four tiles of 16, so every weight tile read from LDS feeds four
 * MFMAs.  Lane l holds sample l % 16 of a tile; the D tile of a layer gives
 * it outputs 4 (l / 16) .. +3, and two D tiles are exactly the lane's B
 * operand of the next layer once the next layer's weight columns are stored
 * in that order (rt_guide.h), so activations never leave registers.  The
 * last layer's rows are ordered so that each lane ends with two whole
 * lobes, and the head is evaluated in registers too.
 *
 * Training (k_guide_train) uses the same instruction and forward pass on 64
 * samples per wave, back-propagates with transposed weight tiles laid out
 * so that a layer's output gradients are the next B operand, and forms the
 * weight gradients (a reduction over samples) with samples along K.  The
 * transposition that this needs - samples from lanes into the registers of
 * one lane - is itself an MFMA: a B operand (8 features of one sample per
 * lane) used as the A operand of a product with an identity selection gives
 * a D tile holding 4 samples of one feature per lane.  Everything stays in
 * registers; the gradients are summed over the block's waves with LDS
 * atomics and written once per block.
 */
#include "rt_device.h"
#include "rt_guide.h"
#include "rt_guide_dev.h"
#include "rt_kernels.h"

#include <cstring>

typedef _Float16 h2 __attribute__((ext_vector_type(2)));
typedef _Float16 h4 __attribute__((ext_vector_type(4)));
typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef float f2 __attribute__((ext_vector_type(2)));
typedef float f4 __attribute__((ext_vector_type(4)));

#define GB 256  /* threads per block: 4 waves */

static inline DevScene as_dev_g(const HostDevScene& s)
{
    DevScene d;
    std::memcpy(&d, &s, sizeof(d));
    return d;
}

__device__ __forceinline__ f4 mfma32(h8 a, h8 b, f4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

/* conversion of a gradient term to FP16 that saturates instead of overflowing */
__device__ __forceinline__ _Float16 to_h_sat(float v)
{
    return (_Float16)fminf(fmaxf(v, -65000.f), 65000.f);
}

/* Head of one lobe from its five raw outputs. */
__device__ __forceinline__ void lobe_head(const float* r, vec3& mu, float& mlen, float& kappa, bool& kin)
{
    vec3 m = mk(r[0], r[1], r[2]);
    mlen = fmaxf(sqrtf(dot(m, m)), 1e-6f);
    mu = m * (1.f / mlen);
    kin = r[3] > GD_KLO && r[3] < GD_KHI;
    kappa = expf(fminf(fmaxf(r[3], GD_KLO), GD_KHI));
}

/* The same head with the hardware's reciprocal square root and exponent:
 * its outputs are stored as FP16, far coarser than these approximations. */
__device__ __forceinline__ void lobe_head_fast(const float* r, vec3& mu, float& kappa)
{
    const vec3 m = mk(r[0], r[1], r[2]);
    mu = m * __builtin_amdgcn_rsqf(fmaxf(dot(m, m), 1e-12f));
    kappa = __expf(fminf(fmaxf(r[3], GD_KLO), GD_KHI));
}

/* Reductions over the four lanes l, l ^ 16, l ^ 32, l ^ 48 (the four lane
 * groups that hold one sample) with gfx950's lane swaps, which are VALU
 * instructions (no LDS round trip).  permlane32_swap(v, v) returns
 * {v of lane l % 32, v of lane 32 + l % 32}; permlane16_swap(v, v) returns
 * {v of row 2 (r / 2), v of row 2 (r / 2) + 1} for a lane in row r of 16
 * lanes.  Either pair sums (or maxes) to the reduction over its two lanes. */
__device__ __forceinline__ f2 swap32(float v)
{
    const auto p = __builtin_amdgcn_permlane32_swap(__float_as_uint(v), __float_as_uint(v), false, false);
    return f2{__uint_as_float(p[0]), __uint_as_float(p[1])};
}
__device__ __forceinline__ f2 swap16(float v)
{
    const auto p = __builtin_amdgcn_permlane16_swap(__float_as_uint(v), __float_as_uint(v), false, false);
    return f2{__uint_as_float(p[0]), __uint_as_float(p[1])};
}
__device__ __forceinline__ float quad_max(float v)
{
    const f2 a = swap32(v);
    const f2 b = swap16(fmaxf(a[0], a[1]));
    return fmaxf(b[0], b[1]);
}
__device__ __forceinline__ float quad_sum(float v)
{
    const f2 a = swap32(v);
    const f2 b = swap16(a[0] + a[1]);
    return b[0] + b[1];
}
/* v of lane l % 16 (row 0) in every lane */
__device__ __forceinline__ float row0_bcast(float v)
{
    return swap16(swap32(v)[0])[0];
}

/* ===========================================================================
 * Inference: the network on 64 samples per wave
 * ========================================================================= */
#define GD_INS 9  /* inputs staged per sample: position in [0,1]^3, normal, outgoing direction */
/* sample tiles of 16 a wave carries per iteration, and threads per block */
#ifndef GD_NET_ST
#define GD_NET_ST 4
#endif
#ifndef GD_NET_THREADS
#define GD_NET_THREADS 256
#endif
#define GD_NET_CH (16 * GD_NET_ST) /* samples per wave and iteration */
/* waves per SIMD the network kernel is compiled for (0: the compiler's choice) */
#ifndef RT_GUIDE_NET_WAVES
#define RT_GUIDE_NET_WAVES 3
#endif
#if RT_GUIDE_NET_WAVES > 0
#define RT_GUIDE_NET_OCCUPANCY __attribute__((amdgpu_waves_per_eu(RT_GUIDE_NET_WAVES, 8)))
#else
#define RT_GUIDE_NET_OCCUPANCY
#endif

/* FP16 inference tiles, FP32 biases in D-tile order (b4 in the row order of
 * the inference W4), and per wave the staged inputs of its 64 samples (an
 * odd stride keeps the 16 samples a tile reads on distinct banks). */
struct alignas(16) InferLds {
    _Float16 w[GD_NW16I];
    float b[4 * GD_H];
    float in[GD_NET_THREADS / 64][GD_NET_CH][GD_INS];
};

/* Bias i of the D-tile order: b1, b2, b3, then b4 in the row order of the
 * forward W4 tiles (zero for rows without an output). */
__device__ __forceinline__ float bias_value(const GuideState& g, int i)
{
    if (i < 3 * GD_H) return g.theta[(i < 64 ? GD_OFF_B1 : i < 128 ? GD_OFF_B2 - 64 : GD_OFF_B3 - 128) + i];
    if (i < 3 * GD_H + GD_OUT) {
        const int r = i - 3 * GD_H, ot = r >> 4, m = r & 15;
        const int o = gd_head_output(m >> 2, 4 * ot + (m & 3));
        if (o >= 0) return g.theta[GD_OFF_B4 + o];
    }
    return 0.f;
}

__device__ __forceinline__ void load_infer_weights(InferLds& L, const GuideState& g)
{
    const uint4* src = reinterpret_cast<const uint4*>(g.w16i);
    uint4* dst = reinterpret_cast<uint4*>(L.w);
    for (int i = threadIdx.x; i < GD_NW16I / 8; i += blockDim.x) dst[i] = src[i];
    for (int i = threadIdx.x; i < 4 * GD_H; i += blockDim.x) L.b[i] = bias_value(g, i);
    __syncthreads();
}

/* LDS written by some lanes of a wave and read by others: a wave's LDS
 * operations execute in order, so only the compiler must not move them. */
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

/* sin and cos of 2 pi x for x in [0, 1) (the hardware's sine and cosine
 * take revolutions) */
__device__ __forceinline__ void sincos_rev(float x, float& s, float& c)
{
    s = __builtin_amdgcn_sinf(x);
    c = __builtin_amdgcn_cosf(x);
}

/* Eight features of one coordinate p in [0, 1]: (sin, cos) of 2^k pi p for
 * k = k0 .. k0 + 3.  The base angle is reduced exactly (p 2^(k0-1) has at
 * most 7 integer bits, so its fractional part is exact in FP32), evaluated
 * in revolutions, then doubled three times (sin 2a = 2 sin a cos a,
 * cos 2a = (cos a - sin a)(cos a + sin a)). */
__device__ __forceinline__ h8 freq_group(float p, int k0)
{
    const float x = ldexpf(p, k0 - 1);
    float sn, cs;
    sincos_rev(x - floorf(x), sn, cs);
    h8 r;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        r[2 * j] = (_Float16)sn;
        r[2 * j + 1] = (_Float16)cs;
        const float s2 = 2.f * sn * cs, c2 = (cs - sn) * (cs + sn);
        sn = s2;
        cs = c2;
    }
    return r;
}

/* Features 32 kb + 8 q .. +7 of one sample, i.e. group g = 4 kb + q of the
 * encoding: g < 9 is coordinate g / 3 at frequencies 4 (g % 3) .. +3, g = 9
 * the normal and outgoing direction, g = 10, 11 the zero padding.  v: the
 * sample's staged inputs. */
__device__ __forceinline__ h8 feature_group(int kb, int q, const float* v)
{
    if (kb == 0) return freq_group(q < 3 ? v[0] : v[1], 4 * (q < 3 ? q : 0));
    if (kb == 1) return freq_group(q < 2 ? v[1] : v[2], 4 * ((q + 1) % 3));
    /* kb 2 without branches on the lane group (a divergent branch would run
     * every case for the whole wave anyway): the frequency group of lane
     * group 0, the six direction values of group 1, zeros elsewhere */
    const h8 f = freq_group(v[2], 8);
    typedef uint32_t u4v __attribute__((ext_vector_type(4)));
    h8 d = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int j = 0; j < 6; ++j) d[j] = (_Float16)v[3 + j];
    const u4v fu = __builtin_bit_cast(u4v, f), du = __builtin_bit_cast(u4v, d);
    u4v r;
#pragma unroll
    for (int j = 0; j < 4; ++j) r[j] = q == 0 ? fu[j] : q == 1 ? du[j] : 0u;
    return __builtin_bit_cast(h8, r);
}

/* bias of D tile ot in every sample tile (ST tiles of 16 samples) */
template <int ST, int OT>
__device__ __forceinline__ void bias_init(const float* b, f4 (&out)[ST][OT])
{
    const int q = __lane_id() >> 4;
#pragma unroll
    for (int ot = 0; ot < OT; ++ot) {
        const f4 bias = {b[16 * ot + 4 * q], b[16 * ot + 4 * q + 1], b[16 * ot + 4 * q + 2], b[16 * ot + 4 * q + 3]};
#pragma unroll
        for (int st = 0; st < ST; ++st) out[st][ot] = bias;
    }
}

/* out[st][ot] += W tile (ot, kb) x in[st] for every output tile: one weight
 * read from LDS per tile, ST MFMAs on it */
template <int ST, int KB, int OT>
__device__ __forceinline__ void dense32_k(const _Float16* W, int kb, const h8 (&in)[ST], f4 (&out)[ST][OT])
{
    const int lane = __lane_id();
#pragma unroll
    for (int ot = 0; ot < OT; ++ot) {
        const h8 a = *reinterpret_cast<const h8*>(W + ((((ot * KB + kb) << 6) + lane) << 3));
#pragma unroll
        for (int st = 0; st < ST; ++st) out[st][ot] = mfma32(a, in[st], out[st][ot]);
    }
}

/* out = W x in + b for a hidden or output layer whose input is two B operands */
template <int ST, int OT>
__device__ __forceinline__ void dense32(const _Float16* W, const float* b, const h8 (&in)[ST][2], f4 (&out)[ST][OT])
{
    bias_init<ST, OT>(b, out);
#pragma unroll
    for (int kb = 0; kb < 2; ++kb) {
        h8 x[ST];
#pragma unroll
        for (int st = 0; st < ST; ++st) x[st] = in[st][kb];
        dense32_k<ST, 2, OT>(W, kb, x, out);
    }
}

/* ReLU of four D tiles into the two B operands of the next layer: pairs
 * converted to FP16 first, then clamped at zero as pairs (rounding keeps the
 * sign, so this is the ReLU of the FP32 value, rounded) */
template <int ST>
__device__ __forceinline__ void relu_pack(const f4 (&z)[ST][4], h8 (&h)[ST][2])
{
    const h2 zero = {(_Float16)0.f, (_Float16)0.f};
#pragma unroll
    for (int st = 0; st < ST; ++st)
#pragma unroll
        for (int kb = 0; kb < 2; ++kb)
#pragma unroll
            for (int j = 0; j < 8; j += 2) {
                const f4& src = z[st][2 * kb + (j >> 2)];
                const f2 pair = {src[j & 3], src[(j & 3) + 1]};
                h2 v = __builtin_convertvector(pair, h2);
                v = __builtin_elementwise_max(v, zero);
                h[st][kb][j] = v[0];
                h[st][kb][j + 1] = v[1];
            }
}

/* The network's inputs of every path vertex of the bounce, by queue
 * position: position in the scene box (FP32, the features need its full
 * precision), viewer-side normal and outgoing direction (FP16, as the
 * network reads them), 24 bytes.  A separate kernel with few registers and
 * many waves hides the latency of the dependent loads (queue -> slot ->
 * hit -> shading normal), and the network kernel then reads its inputs
 * coalesced. */
__global__ __launch_bounds__(GB) void k_guide_gather(DevScene S, PathState ps, GuideState g, int qsel)
{
    const int cnt = ps.cnt->q_count[qsel];
    for (int i = blockIdx.x * GB + threadIdx.x; i < cnt; i += gridDim.x * GB) {
        const int p = ps.queue[qsel][i];
        const float4 h = ps.hit[p];
        const int gg = __float_as_int(h.y);
        float4 a = make_float4(0.f, 0.f, 0.f, 0.f);
        float2 b = make_float2(0.f, 0.f);
        if (gg >= 0) {
            const vec3 o = xyz(ps.ray_o[p]), d = xyz(ps.ray_d[p]);
            const vec3 x = o + h.x * d;
            const vec3 nn = hit_normal(S, gg, x, h.z, h.w);
            const vec3 nf = dot(nn, d) > 0.f ? -nn : nn;
            const h2 n01 = {(_Float16)nf.x, (_Float16)nf.y}, n2w0 = {(_Float16)nf.z, (_Float16)-d.x},
                     w12 = {(_Float16)-d.y, (_Float16)-d.z};
            a = make_float4((x.x - g.lo[0]) * g.inv_ext[0], (x.y - g.lo[1]) * g.inv_ext[1],
                            (x.z - g.lo[2]) * g.inv_ext[2], __builtin_bit_cast(float, n01));
            b = make_float2(__builtin_bit_cast(float, n2w0), __builtin_bit_cast(float, w12));
        }
        g.in_a[i] = a;
        g.in_b[i] = b;
    }
}

__device__ __forceinline__ void unpack_inputs(float4 a, float2 b, float* v)
{
    const h2 n01 = __builtin_bit_cast(h2, a.w), n2w0 = __builtin_bit_cast(h2, b.x), w12 = __builtin_bit_cast(h2, b.y);
    v[0] = a.x, v[1] = a.y, v[2] = a.z;
    v[3] = (float)n01[0], v[4] = (float)n01[1], v[5] = (float)n2w0[0];
    v[6] = (float)n2w0[1], v[7] = (float)w12[0], v[8] = (float)w12[1];
}

/* one lane group's 24 bytes of head output (4-byte aligned: a 16-byte and
 * an 8-byte store) */
struct __attribute__((packed, aligned(4))) HeadGroup {
    uint32_t v[6];
};

/* Head outputs of the path vertex (or record) at queue position i: 48
 * halves, lobe j at gd_lobe_off(j) .. +4 = (mu.xyz, kappa, weight),
 * [GD_ALPHA_OFF] the probability of sampling the guide (rt_guide.h).  RAW:
 * the 48 raw network outputs as floats instead (records only, the parity
 * surface).  Path vertices take their inputs from k_guide_gather's compact
 * records, the next chunk's read while this one is computed. */
template <bool FROM_REC, bool RAW>
__global__ __launch_bounds__(GD_NET_THREADS) RT_GUIDE_NET_OCCUPANCY void k_guide_net(PathState ps, GuideState g, int qsel,
                                                                        const float* rec, int nrec, _Float16* hout,
                                                                        float* yout)
{
    __shared__ InferLds L;
    load_infer_weights(L, g);
    const int lane = __lane_id(), n = lane & 15, q = lane >> 4, w = threadIdx.x >> 6;
    const int cnt = FROM_REC ? nrec : ps.cnt->q_count[qsel];
    if (!FROM_REC && blockIdx.x == 0 && threadIdx.x == 0 && cnt > 0) atomicAdd(g.inferred, (unsigned long long)cnt);
    float (*IN)[GD_INS] = L.in[w];
    const int stride = gridDim.x * (GD_NET_THREADS / 64) * GD_NET_CH;
    int base = (blockIdx.x * (GD_NET_THREADS / 64) + w) * GD_NET_CH;
    float4 na = make_float4(0.f, 0.f, 0.f, 0.f);
    float2 nb = make_float2(0.f, 0.f);
    if (!FROM_REC && lane < GD_NET_CH && base + lane < cnt) {
        na = g.in_a[base + lane];
        nb = g.in_b[base + lane];
    }
    for (; base < cnt; base += stride) {
        /* stage: lane l holds sample base + l */
        {
            const int i = lane < GD_NET_CH ? base + lane : cnt;
            float v[GD_INS];
#pragma unroll
            for (int c = 0; c < GD_INS; ++c) v[c] = 0.f;
            if (FROM_REC) {
                if (i < cnt) {
                    const float* r = rec + (size_t)i * GD_REC;
#pragma unroll
                    for (int k = 0; k < 3; ++k) v[k] = (r[k] - g.lo[k]) * g.inv_ext[k];
#pragma unroll
                    for (int k = 3; k < 9; ++k) v[k] = r[k];
                }
            } else {
                unpack_inputs(na, nb, v);
                /* the next chunk's inputs, in flight during this chunk's work */
                const int j = i + stride;
                na = make_float4(0.f, 0.f, 0.f, 0.f);
                nb = make_float2(0.f, 0.f);
                if (lane < GD_NET_CH && j < cnt) {
                    na = g.in_a[j];
                    nb = g.in_b[j];
                }
            }
            if (lane < GD_NET_CH)
#pragma unroll
                for (int c = 0; c < GD_INS; ++c) IN[lane][c] = v[c];
        }
        wave_lds_sync();
        /* layer 1, one K block of features at a time */
        f4 z[GD_NET_ST][4];
        bias_init<GD_NET_ST, 4>(L.b, z);
#pragma unroll
        for (int kb = 0; kb < 3; ++kb) {
            h8 x[GD_NET_ST];
#pragma unroll
            for (int st = 0; st < GD_NET_ST; ++st) x[st] = feature_group(kb, q, IN[16 * st + n]);
            dense32_k<GD_NET_ST, 3, 4>(L.w + GD_W16I_W1, kb, x, z);
        }
        wave_lds_sync();
        h8 h[GD_NET_ST][2];
        relu_pack<GD_NET_ST>(z, h);
        dense32<GD_NET_ST, 4>(L.w + GD_W16I_W2, L.b + GD_H, h, z);
        relu_pack<GD_NET_ST>(z, h);
        dense32<GD_NET_ST, 4>(L.w + GD_W16I_W3, L.b + 2 * GD_H, h, z);
        relu_pack<GD_NET_ST>(z, h);
        f4 y[GD_NET_ST][3];
        dense32<GD_NET_ST, 3>(L.w + GD_W16I_W4, L.b + 3 * GD_H, h, y);
#pragma unroll
        for (int st = 0; st < GD_NET_ST; ++st) {
            const int i = base + 16 * st + n;
            const bool valid = i < cnt;
            float r[12];
#pragma unroll
            for (int ot = 0; ot < 3; ++ot)
#pragma unroll
                for (int e = 0; e < 4; ++e) r[4 * ot + e] = y[st][ot][e];
            if (RAW) {
#pragma unroll
                for (int sl = 0; sl < 12; ++sl) {
                    const int o = gd_head_output(q, sl);
                    if (valid && o >= 0) yout[(size_t)i * GD_OUT + o] = r[sl];
                }
                continue;
            }
            vec3 mu0, mu1;
            float k0, k1;
            lobe_head_fast(r, mu0, k0);
            lobe_head_fast(r + 5, mu1, k1);
            const float mx = quad_max(fmaxf(r[4], r[9]));
            const float e0 = __expf(r[4] - mx), e1 = __expf(r[9] - mx);
            const float inv = __builtin_amdgcn_rcpf(quad_sum(e0 + e1));
            /* the probability of sampling the guide (lane group 0), zero
             * in the other groups' spare half */
            const float ga = q == 0 ? GD_ALPHA_LO + GD_ALPHA_SPAN * __builtin_amdgcn_rcpf(1.f + __expf(-r[10])) : 0.f;
            if (valid) {
                const h2 v[6] = {h2{(_Float16)mu0.x, (_Float16)mu0.y}, h2{(_Float16)mu0.z, (_Float16)k0},
                                 h2{(_Float16)(e0 * inv), (_Float16)mu1.x}, h2{(_Float16)mu1.y, (_Float16)mu1.z},
                                 h2{(_Float16)k1, (_Float16)(e1 * inv)}, h2{(_Float16)ga, (_Float16)0.f}};
                HeadGroup hg;
#pragma unroll
                for (int k = 0; k < 6; ++k) hg.v[k] = __builtin_bit_cast(uint32_t, v[k]);
                *reinterpret_cast<HeadGroup*>(hout + (size_t)i * GD_PARAMS + 12 * q) = hg;
            }
        }
    }
}

hipError_t launch_guide_infer(const HostDevScene& s, const PathState& ps, const GuideState& g, int qsel, int grid,
                              hipStream_t st)
{
    hipLaunchKernelGGL(k_guide_gather, dim3(grid * 4), dim3(GB), 0, st, as_dev_g(s), ps, g, qsel);
    hipLaunchKernelGGL((k_guide_net<false, false>), dim3(grid), dim3(GD_NET_THREADS), 0, st, ps, g, qsel, (const float*)nullptr, 0,
                       g.out, (float*)nullptr);
    return hipGetLastError();
}

/* the network on records (parity surface): raw outputs y [n x 48] */
hipError_t launch_guide_forward(const GuideState& g, const float* rec, int n, float* y, int grid, hipStream_t st)
{
    hipLaunchKernelGGL((k_guide_net<true, true>), dim3(grid), dim3(GD_NET_THREADS), 0, st, PathState{}, g, 0, rec, n,
                       (_Float16*)nullptr, y);
    return hipGetLastError();
}

/* head outputs of records as the shading kernel receives them [n x 48 halves] */
hipError_t launch_guide_head(const GuideState& g, const float* rec, int n, _Float16* out, int grid, hipStream_t st)
{
    hipLaunchKernelGGL((k_guide_net<true, false>), dim3(grid), dim3(GD_NET_THREADS), 0, st, PathState{}, g, 0, rec, n, out,
                       (float*)nullptr);
    return hipGetLastError();
}

/* ===========================================================================
 * record weights: w = lum(radiance_final - radiance_at_close) / lum(alpha),
 * and the weight sum of the records a training step will use
 * ========================================================================= */
/* records of one training chunk: [off, off + max_n) clipped to the stored ones */
__device__ __forceinline__ int chunk_len(const GuideState& g, const int32_t* count, int off, int max_n)
{
    return max(0, min(min(*count, g.rec_cap) - off, max_n));
}

__global__ __launch_bounds__(GB) void k_guide_weights(GuideState g, PathState ps, int from_paths, int off, int max_n)
{
    const int N = off + chunk_len(g, g.rec_count, off, max_n);
    float local = 0.f;
    for (int i = off + blockIdx.x * GB + threadIdx.x; i < N; i += gridDim.x * GB) {
        float* r = g.rec + (size_t)i * GD_REC;
        float w;
        if (from_paths) {
            float4 aux = g.rec_aux[i];
            int p = __float_as_int(aux.w);
            float4 L = ps.radiance[p];
            vec3 inc = mk(L.x - aux.x, L.y - aux.y, L.z - aux.z);
            w = fmaxf(lum(inc), 0.f) / fmaxf(r[14], 1e-8f);
            if (!(w < 3.0e38f)) w = 0.f;
            r[13] = w;
        } else {
            w = r[13];
        }
        local += w;
    }
    /* one atomic per block: same-address atomics serialise */
    __shared__ float part[GB / 64];
    for (int o = 32; o > 0; o >>= 1) local += __shfl_xor(local, o);
    if (__lane_id() == 0) part[threadIdx.x >> 6] = local;
    __syncthreads();
    if (threadIdx.x == 0) {
        float s = 0.f;
        for (int k = 0; k < GB / 64; ++k) s += part[k];
        if (s != 0.f) atomicAdd(&g.acc[GD_WSUM_OFF + blockIdx.x % GD_WSUM_N], s);
    }
}

hipError_t launch_guide_weights(const GuideState& g, const PathState& ps, int off, int max_n, int grid,
                                hipStream_t st)
{
    /* from_paths = 1 when the records come from a render (radiance known) */
    hipLaunchKernelGGL(k_guide_weights, dim3(grid), dim3(GB), 0, st, g, ps, ps.radiance != nullptr ? 1 : 0, off,
                       max_n);
    return hipGetLastError();
}

__global__ void k_guide_zero_acc(GuideState g)
{
    if (threadIdx.x < 3) g.acc[threadIdx.x] = 0.f;
    if (threadIdx.x < GD_WSUM_N) g.acc[GD_WSUM_OFF + threadIdx.x] = 0.f;
}

hipError_t launch_guide_zero_acc(const GuideState& g, hipStream_t st)
{
    hipLaunchKernelGGL(k_guide_zero_acc, dim3(1), dim3(64), 0, st, g);
    return hipGetLastError();
}

/* ===========================================================================
 * training: loss -sum s_i log p_mix_i with s_i = w_i N / sum w (capped),
 * gradients summed over the records into g.grad (Adam divides by N)
 * ========================================================================= */
#define GD_TW (GD_TRAIN_THREADS / 64) /* waves per training block */
#define GD_TST 2                       /* sample tiles per iteration: 32 records per wave */
#define GD_TCH (16 * GD_TST)           /* records per wave and iteration */
#define GD_TIN 15                      /* staged floats per sample (odd stride: distinct banks) */
#define GD_NTILE 64                    /* weight-gradient tiles: W1 20, W2 16, W3 16, W4 12 */
#define GD_NBT 15                      /* bias-gradient tiles: b1..b3 4 each, b4 3 */
#define GD_NSUM (GD_NTILE + 1)         /* register sums per wave: the weight tiles, then all bias tiles */
static_assert(GD_TST == 2, "the training kernel's operand lists name two sample tiles");

/* FP16 forward and backward tiles, FP32 biases in D-tile order and per wave
 * the staged inputs of its 32 samples (network inputs (9), sampled direction
 * (3), BSDF pdf, normalised weight); after the last iteration the same space
 * holds the block's gradient sums [tile][lane][e].  A wave's own sums stay
 * in its registers until then. */
struct alignas(16) TrainLds {
    union {
        struct {
            _Float16 w[GD_NW16];
            float b[4 * GD_H];
            float in[GD_TW][GD_TCH][GD_TIN];
        } s;
        float acc[GD_NSUM * 256];
    };
};

/* Identity selection of one 16-feature tile of a K block, as the B operand
 * of a transposing product (lane c + 16 q).  Hidden layers hold their
 * features in the permuted order of the forward tiles (feature 16 t + c of
 * the block at slot c - 4 q + 4 t of lane group c / 4); the network inputs in
 * natural order (slot 16 t + c - 8 q of lane group (16 t + c) / 8). */
__device__ __forceinline__ h8 sel_op(int t, bool natural)
{
    const int lane = __lane_id(), c = lane & 15, q = lane >> 4;
    const int d = c - 4 * q;
    const int j = natural ? 16 * t + c - 8 * q : (d >= 0 && d < 4 ? d + 4 * t : -1);
    h8 r;
#pragma unroll
    for (int k = 0; k < 8; ++k) r[k] = (_Float16)(k == j ? 1.f : 0.f);
    return r;
}

__device__ __forceinline__ h8 pack8(f4 a, f4 b)
{
    const h4 lo = __builtin_convertvector(a, h4), hi = __builtin_convertvector(b, h4);
    return __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7);
}

/* One 16-feature tile of a matrix held as B operands M[st] (the K block
 * that contains it; sel picks it) with the wave's 32 samples along K: lane
 * c + 16 q gets feature c of the tile for samples 4 q .. 4 q + 3 of both
 * sample tiles.  Every operand made this way orders the samples alike, so
 * two of them are the A and B operands of a reduction over the samples.
 * FP16 values pass the product exactly. */
template <int T, bool NATURAL>
__device__ __forceinline__ h8 to_sil(const h8 (&M)[GD_TST])
{
    const h8 sel = sel_op(T, NATURAL); /* rematerialised at each use: fewer live registers */
    const f4 zero = {0.f, 0.f, 0.f, 0.f};
    return pack8(mfma32(M[0], sel, zero), mfma32(M[1], sel, zero));
}

/* Bias-gradient tile `tile`: its output gradients summed over the samples,
 * as the product with a B operand whose column `tile` is ones and whose
 * other columns are zero, so that all fifteen bias tiles add into one
 * accumulator (lane c + 16 q: rows 4 q .. 4 q + 3 of tile c). */
__device__ __forceinline__ void bias_tile(f4& bacc, int tile, h8 zS)
{
    const _Float16 one = (__lane_id() & 15) == tile ? (_Float16)1.f : (_Float16)0.f;
    const h8 sel = {one, one, one, one, one, one, one, one};
    bacc = mfma32(zS, sel, bacc);
}

/* Saturating FP16 pair of two gradient terms, zeroed where the ReLU input
 * pair m (FP16, >= 0) is zero: min(m 2^24, 1) is 1 for every positive FP16
 * value (subnormals included) and 0 for zero, so the mask is two packed
 * products and a packed minimum */
__device__ __forceinline__ h2 relu_grad_pair(float a, float b, h2 m)
{
    const f2 v = {fminf(fmaxf(a, -65000.f), 65000.f), fminf(fmaxf(b, -65000.f), 65000.f)};
    const h2 big = {(_Float16)4096.f, (_Float16)4096.f}, one = {(_Float16)1.f, (_Float16)1.f};
    const h2 mask = __builtin_elementwise_min(m * big * big, one);
    return __builtin_convertvector(v, h2) * mask;
}

/* Output gradients of the layer below (ReLU activations h, B operands) from
 * this layer's output gradients `in`: W^T (backward tiles (it, kb)) x in,
 * masked by h > 0 and packed as B operands (slot j of K block kb is input
 * 16 (2 kb + j / 4) + 4 q + j % 4), one K block of the result at a time. */
__device__ __forceinline__ void back_relu(const _Float16* WT, const h8 (&in)[GD_TST][2], const h8 (&h)[GD_TST][2],
                                          h8 (&out)[GD_TST][2])
{
    const int lane = __lane_id();
#pragma unroll
    for (int ko = 0; ko < 2; ++ko) {
        f4 d[GD_TST][2];
#pragma unroll
        for (int st = 0; st < GD_TST; ++st) d[st][0] = d[st][1] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kb = 0; kb < 2; ++kb)
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                const int it = 2 * ko + t;
                const h8 a = *reinterpret_cast<const h8*>(WT + ((((it * 2 + kb) << 6) + lane) << 3));
#pragma unroll
                for (int st = 0; st < GD_TST; ++st) d[st][t] = mfma32(a, in[st][kb], d[st][t]);
            }
#pragma unroll
        for (int st = 0; st < GD_TST; ++st)
#pragma unroll
            for (int j = 0; j < 8; j += 2) {
                const h2 m = {h[st][ko][j], h[st][ko][j + 1]};
                const h2 v = relu_grad_pair(d[st][j >> 2][j & 3], d[st][j >> 2][(j & 3) + 1], m);
                out[st][ko][j] = v[0];
                out[st][ko][j + 1] = v[1];
            }
    }
}

/* One hidden layer of the backward pass: the gradients of its weights
 * (tiles TILE0 + 4 ot + it) and biases (tiles BTILE0 + ot) from its output
 * gradients dz and inputs a, then dz <- the output gradients of the layer
 * below (through W^T and that layer's ReLU, whose activations a are). */
template <int TILE0, int BTILE0>
__device__ __forceinline__ void hidden_back(f4 (&acc)[GD_NSUM], const _Float16* WT, const h8 (&a)[GD_TST][2],
                                            h8 (&dz)[GD_TST][2])
{
    h8 zS[4];
#pragma unroll
    for (int ot = 0; ot < 4; ++ot) {
        const h8 m[GD_TST] = {dz[0][ot >> 1], dz[1][ot >> 1]};
        zS[ot] = (ot & 1) ? to_sil<1, false>(m) : to_sil<0, false>(m);
        bias_tile(acc[GD_NTILE], BTILE0 + ot, zS[ot]);
    }
#pragma unroll
    for (int it = 0; it < 4; ++it) {
        const h8 ma[GD_TST] = {a[0][it >> 1], a[1][it >> 1]};
        const h8 aS = (it & 1) ? to_sil<1, false>(ma) : to_sil<0, false>(ma);
#pragma unroll
        for (int ot = 0; ot < 4; ++ot) acc[TILE0 + 4 * ot + it] = mfma32(zS[ot], aS, acc[TILE0 + 4 * ot + it]);
    }
    h8 below[GD_TST][2];
    back_relu(WT, dz, a, below);
#pragma unroll
    for (int st = 0; st < GD_TST; ++st) dz[st][0] = below[st][0], dz[st][1] = below[st][1];
}

/* Master parameter of gradient tile `tile`, element e, lane (-1: none). */
__device__ __forceinline__ int grad_param(int tile, int e, int lane)
{
    const int n = lane & 15, q = lane >> 4;
    if (tile < 20) return GD_OFF_W1 + (16 * (tile / 5) + 4 * q + e) * GD_IN + 16 * (tile % 5) + n;
    if (tile < 52) {
        const int t = tile < 36 ? tile - 20 : tile - 36;
        return (tile < 36 ? GD_OFF_W2 : GD_OFF_W3) + (16 * (t >> 2) + 4 * q + e) * GD_H + 16 * (t & 3) + n;
    }
    const int t = tile - 52, o = gd_head_output(q, 4 * (t >> 2) + e);
    return o < 0 ? -1 : GD_OFF_W4 + o * GD_H + 16 * (t & 3) + n;
}
__device__ __forceinline__ int bias_param(int tile, int e, int q)
{
    if (tile < 12) return (tile < 4 ? GD_OFF_B1 : tile < 8 ? GD_OFF_B2 : GD_OFF_B3) + 16 * (tile & 3) + 4 * q + e;
    const int o = gd_head_output(q, 4 * (tile - 12) + e);
    return o < 0 ? -1 : GD_OFF_B4 + o;
}

__global__ __launch_bounds__(GD_TRAIN_THREADS) __attribute__((amdgpu_waves_per_eu(1, 1))) void k_guide_train(
    GuideState g, const float* rec0, const int32_t* count, int off, int max_n)
{
    __shared__ TrainLds L;
    {
        const uint4* src = reinterpret_cast<const uint4*>(g.w16i);
        uint4* dst = reinterpret_cast<uint4*>(L.s.w);
        for (int i = threadIdx.x; i < GD_NW16 / 8; i += blockDim.x) dst[i] = src[i];
        for (int i = threadIdx.x; i < 4 * GD_H; i += blockDim.x) L.s.b[i] = bias_value(g, i);
    }
    __syncthreads();
    const _Float16* WF = L.s.w;
    const _Float16* WB = L.s.w + GD_NW16I;
    const int lane = __lane_id(), n = lane & 15, q = lane >> 4, w = threadIdx.x >> 6;
    const int N = chunk_len(g, count, off, max_n);
    const float* rec = rec0 + (size_t)off * GD_REC;
    float wsum = 0.f;
#pragma unroll
    for (int j = 0; j < GD_WSUM_N; ++j) wsum += g.acc[GD_WSUM_OFF + j];
    const float wscale = wsum > 0.f ? (float)N / wsum : 0.f;
    float (*IN)[GD_TIN] = L.s.in[w];
    float lacc = 0.f;
    f4 acc[GD_NSUM];
#pragma unroll
    for (int t = 0; t < GD_NSUM; ++t) acc[t] = f4{0.f, 0.f, 0.f, 0.f};
    /* record loads: lane l reads floats 8 (l / 32) .. + 7 of record l % 32,
     * the next iteration's while this one is computed */
    const int half = lane >> 5, rl = lane & 31;
    const int stride = gridDim.x * GD_TW * GD_TCH;
    int base = (blockIdx.x * GD_TW + w) * GD_TCH;
    float4 pa = make_float4(0.f, 0.f, 0.f, 0.f), pb4 = pa;
    if (base + rl < N) {
        const float4* r = reinterpret_cast<const float4*>(rec + (size_t)(base + rl) * GD_REC) + 2 * half;
        pa = r[0];
        pb4 = r[1];
    }
    for (; base < N; base += stride) {
        /* stage: record base + rl, floats 8 half .. (float 15 is not staged) */
        {
            /* the position mapped into the scene box (floats 0..2), the
             * weight normalised by the step's weight sum and capped (float
             * 13), the rest as stored; the kernel arguments indexed by
             * compile-time constants only */
            const float v[8] = {pa.x, pa.y, pa.z, pa.w, pb4.x, pb4.y, pb4.z, pb4.w};
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                float val = v[k];
                if (k < 3) val = half == 0 ? (val - g.lo[k]) * g.inv_ext[k] : val;
                if (k == 13 - 8) val = half == 1 ? fminf(val * wscale, GD_WCLAMP) : val;
                if (k < GD_TIN - 8 || half == 0) IN[rl][8 * half + k] = val;
            }
            const int j = base + stride + rl;
            pa = make_float4(0.f, 0.f, 0.f, 0.f);
            pb4 = pa;
            if (j < N) {
                const float4* r = reinterpret_cast<const float4*>(rec + (size_t)j * GD_REC) + 2 * half;
                pa = r[0];
                pb4 = r[1];
            }
        }
        wave_lds_sync();

        /* forward (as inference), keeping the activations as B operands */
        h8 h1[GD_TST][2], h2[GD_TST][2], h3[GD_TST][2];
        f4 y[GD_TST][3];
        {
            f4 z[GD_TST][4];
            bias_init<GD_TST, 4>(L.s.b, z);
#pragma unroll
            for (int kb = 0; kb < 3; ++kb) {
                h8 x[GD_TST];
#pragma unroll
                for (int st = 0; st < GD_TST; ++st) x[st] = feature_group(kb, q, IN[16 * st + n]);
                dense32_k<GD_TST, 3, 4>(WF + GD_W16I_W1, kb, x, z);
            }
            relu_pack<GD_TST>(z, h1);
            dense32<GD_TST, 4>(WF + GD_W16I_W2, L.s.b + GD_H, h1, z);
            relu_pack<GD_TST>(z, h2);
            dense32<GD_TST, 4>(WF + GD_W16I_W3, L.s.b + 2 * GD_H, h2, z);
            relu_pack<GD_TST>(z, h3);
            dense32<GD_TST, 3>(WF + GD_W16I_W4, L.s.b + 3 * GD_H, h3, y);
        }

        /* head: loss and its gradient in the lanes that hold the outputs
         * (lane group q: lobes 2q, 2q+1; group 0 also the guide logit) */
        h8 dy[GD_TST][2];
#pragma unroll
        for (int st = 0; st < GD_TST; ++st) {
            const int i = base + 16 * st + n;
            const bool valid = i < N;
            const float* v = IN[16 * st + n];
            const vec3 om = mk(v[9], v[10], v[11]);
            const float pb = v[12], sw = v[13];
            float r[12];
#pragma unroll
            for (int ot = 0; ot < 3; ++ot)
#pragma unroll
                for (int e = 0; e < 4; ++e) r[4 * ot + e] = y[st][ot][e];
            /* the hardware exponential, logarithm and reciprocal (about 1
             * ulp): the gradients are rounded to FP16 right after */
            vec3 mu[2];
            float rlen[2], kap[2], vv[2], om1[2];
            bool kin[2];
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                const float* rs = r + 5 * s;
                const vec3 m = mk(rs[0], rs[1], rs[2]);
                rlen[s] = __builtin_amdgcn_rcpf(fmaxf(sqrtf(dot(m, m)), 1e-6f));
                mu[s] = m * rlen[s];
                kin[s] = rs[3] > GD_KLO && rs[3] < GD_KHI;
                kap[s] = __expf(fminf(fmaxf(rs[3], GD_KLO), GD_KHI));
                om1[s] = 1.f - __expf(-2.f * kap[s]); /* >= 0.094 for kappa >= e^-3 */
                vv[s] = kap[s] * __builtin_amdgcn_rcpf(6.28318530718f * om1[s]) * __expf(kap[s] * (dot(mu[s], om) - 1.f));
            }
            const float mx = quad_max(fmaxf(r[4], r[9]));
            const float e0 = __expf(r[4] - mx), e1 = __expf(r[9] - mx);
            const float inv = __builtin_amdgcn_rcpf(quad_sum(e0 + e1));
            const float pi[2] = {e0 * inv, e1 * inv};
            const float qd = quad_sum(pi[0] * vv[0] + pi[1] * vv[1]);
            const float sg = __builtin_amdgcn_rcpf(1.f + __expf(-row0_bcast(r[10])));
            const float alpha = GD_ALPHA_LO + GD_ALPHA_SPAN * sg;
            const float pm = alpha * qd + (1.f - alpha) * pb;
            const float c = valid ? -sw * __builtin_amdgcn_rcpf(pm) : 0.f;
            const float dq = c * alpha;
            if (valid && q == 0) lacc += -sw * __logf(pm);
            float dr[12];
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                const float dv = dq * pi[s];
                const float cw = dot(mu[s], om);
                const float dlogc = 1.f - 2.f * kap[s] * (1.f - om1[s]) * __builtin_amdgcn_rcpf(om1[s]);
                const float dk = kin[s] ? dv * vv[s] * (dlogc + kap[s] * (cw - 1.f)) : 0.f;
                const float gm = dv * vv[s] * kap[s] * rlen[s];
                const vec3 dm = (om - mu[s] * cw) * gm;
                dr[5 * s] = dm.x;
                dr[5 * s + 1] = dm.y;
                dr[5 * s + 2] = dm.z;
                dr[5 * s + 3] = dk;
                dr[5 * s + 4] = dq * pi[s] * (vv[s] - qd);
            }
            dr[10] = q == 0 ? c * (qd - pb) * GD_ALPHA_SPAN * sg * (1.f - sg) : 0.f;
            dr[11] = 0.f;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                dy[st][0][j] = to_h_sat(dr[j]);
                dy[st][1][j] = j < 4 ? to_h_sat(dr[8 + j]) : (_Float16)0.f;
            }
        }

        /* W4 and b4: gradients (rows: the head's slots as the forward W4 tile rows) */
        {
            h8 dyS[3];
            {
                const h8 m0[GD_TST] = {dy[0][0], dy[1][0]};
                const h8 m1[GD_TST] = {dy[0][1], dy[1][1]};
                dyS[0] = to_sil<0, false>(m0);
                dyS[1] = to_sil<1, false>(m0);
                dyS[2] = to_sil<0, false>(m1);
            }
#pragma unroll
            for (int ot = 0; ot < 3; ++ot) bias_tile(acc[GD_NTILE], 12 + ot, dyS[ot]);
#pragma unroll
            for (int it = 0; it < 4; ++it) {
                const h8 m[GD_TST] = {h3[0][it >> 1], h3[1][it >> 1]};
                const h8 aS = (it & 1) ? to_sil<1, false>(m) : to_sil<0, false>(m);
#pragma unroll
                for (int ot = 0; ot < 3; ++ot) acc[52 + 4 * ot + it] = mfma32(dyS[ot], aS, acc[52 + 4 * ot + it]);
            }
        }
        /* back through W4, W3, W2; each layer's gradients as soon as its
         * output gradients exist */
        h8 dz[GD_TST][2];
        back_relu(WB + GD_W16B_W4, dy, h3, dz);
        hidden_back<36, 8>(acc, WB + GD_W16B_W3, h2, dz);
        hidden_back<20, 4>(acc, WB + GD_W16B_W2, h1, dz);
        /* W1 and b1: output gradients dz (layer 1), inputs recomputed */
        {
            h8 zS[4];
#pragma unroll
            for (int ot = 0; ot < 4; ++ot) {
                const h8 m[GD_TST] = {dz[0][ot >> 1], dz[1][ot >> 1]};
                zS[ot] = (ot & 1) ? to_sil<1, false>(m) : to_sil<0, false>(m);
                bias_tile(acc[GD_NTILE], ot, zS[ot]);
            }
#pragma unroll
            for (int kb = 0; kb < 3; ++kb) {
                h8 x[GD_TST];
#pragma unroll
                for (int st = 0; st < GD_TST; ++st) x[st] = feature_group(kb, q, IN[16 * st + n]);
#pragma unroll
                for (int t = 0; t < 2; ++t) {
                    const int it = 2 * kb + t;
                    if (it >= 5) continue; /* inputs 80..95 are padding */
                    const h8 xS = t ? to_sil<1, true>(x) : to_sil<0, true>(x);
#pragma unroll
                    for (int ot = 0; ot < 4; ++ot) acc[5 * ot + it] = mfma32(zS[ot], xS, acc[5 * ot + it]);
                }
            }
        }
        wave_lds_sync();
    }
    /* the block's sums: the waves add their register sums in turn (into the
     * weights' space, free once every wave is past its last iteration) */
    __syncthreads();
    for (int turn = 0; turn < GD_TW; ++turn) {
        if (w == turn)
#pragma unroll
            for (int t = 0; t < GD_NSUM; ++t) {
                f4* dst = reinterpret_cast<f4*>(&L.acc[(t * 64 + lane) * 4]);
                *dst = turn == 0 ? acc[t] : *dst + acc[t];
            }
        __syncthreads();
    }
    float* part = g.part + (size_t)blockIdx.x * GD_NPARAM;
    static_assert(GD_TRAIN_THREADS == 256, "one thread per (lane, element) of a tile");
    {
        /* tile-major, so the tile index is uniform and its arithmetic scalar */
        const int e = threadIdx.x & 3, l = threadIdx.x >> 2;
        for (int t = 0; t < GD_NSUM; ++t) {
            const int pi = t < GD_NTILE ? grad_param(t, e, l) : ((l & 15) < GD_NBT ? bias_param(l & 15, e, l >> 4) : -1);
            if (pi >= 0) part[pi] = L.acc[t * 256 + threadIdx.x];
        }
    }
    for (int o = 32; o > 0; o >>= 1) lacc += __shfl_xor(lacc, o);
    if (lane == 0 && lacc != 0.f) atomicAdd(&g.acc[1], lacc);
}

/* grad += the sum of the training blocks' partial sums; blockIdx.y takes a
 * slice of the blocks */
__global__ __launch_bounds__(GB) void k_guide_reduce(GuideState g, int nb)
{
    const int i = blockIdx.x * GB + threadIdx.x;
    if (i >= GD_NPARAM || !gd_trained(i)) return;
    const int per = (nb + gridDim.y - 1) / gridDim.y, b0 = blockIdx.y * per, b1 = min(nb, b0 + per);
    float s = 0.f;
    for (int b = b0; b < b1; ++b) s += g.part[(size_t)b * GD_NPARAM + i];
    if (s != 0.f) atomicAdd(&g.grad[i], s);
}

hipError_t launch_guide_train(const GuideState& g, const float* rec, const int32_t* count, int off, int max_n,
                              int grid, hipStream_t st)
{
    if (grid > g.part_blocks) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_guide_train, dim3(grid), dim3(GD_TRAIN_THREADS), 0, st, g, rec, count, off, max_n);
    hipLaunchKernelGGL(k_guide_reduce, dim3((GD_NPARAM + GB - 1) / GB, 8), dim3(GB), 0, st, g, grid);
    return hipGetLastError();
}

/* ===========================================================================
 * Adam (Kingma & Ba, 2015) on the FP32 master copy, refresh FP16 weights,
 * clear gradients and accumulators
 * ========================================================================= */
__global__ __launch_bounds__(GB) void k_guide_adam(GuideState g, float lr, int step, const int32_t* count, int off,
                                                  int max_n)
{
    const int N = chunk_len(g, count, off, max_n);
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
    if (blockIdx.x == 0 && threadIdx.x == 0 && N > 0) atomicAdd(g.trained, (unsigned long long)N);
    for (int i = blockIdx.x * GB + threadIdx.x; i < GD_NPARAM; i += gridDim.x * GB) {
        float gr = g.grad[i];
        g.grad[i] = 0.f;
        if (N <= 0) continue;
        gr /= (float)N;
        float m = b1 * g.adam_m[i] + (1.f - b1) * gr;
        float v = b2 * g.adam_v[i] + (1.f - b2) * gr * gr;
        g.adam_m[i] = m;
        g.adam_v[i] = v;
        float th = g.theta[i] - lr * (m / c1) / (sqrtf(v / c2) + eps);
        g.theta[i] = th;
        const int iw = gd_w16i_slot(i), ib = gd_w16b_slot(i);
        if (iw >= 0) g.w16i[iw] = (_Float16)th;
        if (ib >= 0) g.w16b[ib] = (_Float16)th;
    }
}

hipError_t launch_guide_adam(const GuideState& g, float lr, int step, const int32_t* count, int off, int max_n,
                             hipStream_t st)
{
    hipLaunchKernelGGL(k_guide_adam, dim3(32), dim3(GB), 0, st, g, lr, step, count, off, max_n);
    return hipGetLastError();
}

__global__ void k_guide_keep(GuideState g, float scale)
{
    if (threadIdx.x != 0) return;
    const int n = *g.rec_count;
    const float k = *g.keep;
    if (scale > 0.f) {
        *g.keep = fminf(1.f, fmaxf(k * scale, 1e-7f));
    } else if (n > 0 && k > 0.f) {
        const float offered = (float)n / k;
        *g.keep = fminf(1.f, fmaxf(0.9f * (float)g.rec_cap / offered, 1e-7f));
    }
}

hipError_t launch_guide_keep(const GuideState& g, float scale, hipStream_t st)
{
    hipLaunchKernelGGL(k_guide_keep, dim3(1), dim3(64), 0, st, g, scale);
    return hipGetLastError();
}

__global__ void k_guide_reset(GuideState g, int slots)
{
    int i = blockIdx.x * GB + threadIdx.x;
    if (i == 0) {
        *g.rec_count = 0;
        g.acc[0] = g.acc[1] = g.acc[2] = 0.f;
        for (int j = 0; j < GD_WSUM_N; ++j) g.acc[GD_WSUM_OFF + j] = 0.f;
    }
    for (; i < slots; i += gridDim.x * GB) g.open_rec[i] = -1;
}

hipError_t launch_guide_reset_records(const GuideState& g, int slots, hipStream_t st)
{
    int grid = std::max(1, std::min(1024, (slots + GB - 1) / GB));
    hipLaunchKernelGGL(k_guide_reset, dim3(grid), dim3(GB), 0, st, g, slots);
    return hipGetLastError();
}
