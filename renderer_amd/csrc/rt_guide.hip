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
 * Training (k_guide_train) uses v_mfma_f32_16x16x16_f16 on 16 samples per
 * wave.  Weight gradients (a reduction over samples) read activations and
 * deltas back from LDS in [neuron][sample] order, which gives both MFMA
 * operands as contiguous 8-byte reads.
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

__device__ __forceinline__ f4 mfma16(h4 a, h4 b, f4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x16f16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f4 mfma32(h8 a, h8 b, f4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

/* FP16 weights (the first NW values of the tiled copy) and FP32 biases in LDS */
template <int NW>
struct alignas(16) WLds {
    _Float16 w[NW];
    float b[3 * GD_H + GD_OUT];
};

template <int NW>
__device__ __forceinline__ void load_weights(WLds<NW>& L, const GuideState& g)
{
    const uint4* src = reinterpret_cast<const uint4*>(g.w16);
    uint4* dst = reinterpret_cast<uint4*>(L.w);
    for (int i = threadIdx.x; i < NW / 8; i += blockDim.x) dst[i] = src[i];
    for (int i = threadIdx.x; i < 3 * GD_H + GD_OUT; i += blockDim.x) {
        int off = i < 64 ? GD_OFF_B1 + i : i < 128 ? GD_OFF_B2 + i - 64 : i < 192 ? GD_OFF_B3 + i - 128 : GD_OFF_B4 + i - 192;
        L.b[i] = g.theta[off];
    }
    __syncthreads();
}

/* conversion of a gradient term to FP16 that saturates instead of overflowing */
__device__ __forceinline__ _Float16 to_h_sat(float v)
{
    return (_Float16)fminf(fmaxf(v, -65000.f), 65000.f);
}

/* Input features of a wave's 16 samples, staged through LDS rows of
 * GD_FROW halves (the padding keeps the 8-byte reads of 16 rows on distinct
 * banks).  Lane (n, q) evaluates sin/cos(2^k pi p_d) for k = 3q .. 3q+2 and
 * the three coordinates of sample n: one sincospi at k = 3q, then two angle
 * doublings (sin 2a = 2 sin a cos a, cos 2a = (cos a - sin a)(cos a + sin a));
 * lane q = 3 adds the normal, the outgoing direction and the zero padding.
 * After a barrier, x[kb] (features 16 kb + 4 q .. +3 of sample n, the MFMA
 * B-operand quad) is one 8-byte read. */
#define GD_FROW (GD_IN + 4)
__device__ __forceinline__ void features_to_lds(_Float16 (*F)[GD_FROW], int n, int q, const float* p01, vec3 nf,
                                                vec3 wo)
{
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        float s, c;
        sincospif(ldexpf(p01[d], 3 * q), &s, &c);
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            int f = 24 * d + 2 * (3 * q + j);
            F[n][f] = (_Float16)s;
            F[n][f + 1] = (_Float16)c;
            float s2 = 2.f * s * c, c2 = (c - s) * (c + s);
            s = s2;
            c = c2;
        }
    }
    if (q == 3) {
        F[n][72] = (_Float16)nf.x;
        F[n][73] = (_Float16)nf.y;
        F[n][74] = (_Float16)nf.z;
        F[n][75] = (_Float16)wo.x;
        F[n][76] = (_Float16)wo.y;
        F[n][77] = (_Float16)wo.z;
        F[n][78] = (_Float16)0.f;
        F[n][79] = (_Float16)0.f;
    }
}
__device__ __forceinline__ void features_from_lds(const _Float16 (*F)[GD_FROW], int n, int q, h4* x)
{
#pragma unroll
    for (int kb = 0; kb < 5; ++kb) x[kb] = *reinterpret_cast<const h4*>(&F[n][16 * kb + 4 * q]);
}

/* out[ot] = W[16 ot .. +15][:] * in (+ bias), W as forward tiles (ot, kb). */
template <int KB, int OT>
__device__ __forceinline__ void dense(const _Float16* W, const float* b, const h4* in, f4* out)
{
    const int lane = __lane_id(), q = lane >> 4;
#pragma unroll
    for (int ot = 0; ot < OT; ++ot) {
        f4 acc = {b[16 * ot + 4 * q], b[16 * ot + 4 * q + 1], b[16 * ot + 4 * q + 2], b[16 * ot + 4 * q + 3]};
#pragma unroll
        for (int kb = 0; kb < KB; ++kb) {
            h4 a = *reinterpret_cast<const h4*>(W + (((ot * KB + kb) << 6) + lane) * 4);
            acc = mfma16(a, in[kb], acc);
        }
        out[ot] = acc;
    }
}

/* out[it] = W^T[16 it .. +15][:] * in, W^T as transposed tiles (it, kb)
 * (back-propagation through a layer with KB * 16 outputs). */
template <int KB, int IT>
__device__ __forceinline__ void dense_t(const _Float16* WT, const h4* in, f4* out)
{
    const int lane = __lane_id();
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kb = 0; kb < KB; ++kb) {
            h4 a = *reinterpret_cast<const h4*>(WT + (((it * KB + kb) << 6) + lane) * 4);
            acc = mfma16(a, in[kb], acc);
        }
        out[it] = acc;
    }
}

template <int T>
__device__ __forceinline__ void relu_h(const f4* z, h4* h)
{
#pragma unroll
    for (int t = 0; t < T; ++t)
#pragma unroll
        for (int i = 0; i < 4; ++i) h[t][i] = (_Float16)fmaxf(z[t][i], 0.f);
}

/* Forward pass of 16 samples held by one wave.  Returns y tiles (3 x f4). */
__device__ __forceinline__ void forward16(const _Float16* w, const float* b, const h4* x, h4* h1, h4* h2, h4* h3,
                                          f4* y)
{
    f4 z[4];
    dense<5, 4>(w + GD_W16_W1, b, x, z);
    relu_h<4>(z, h1);
    dense<4, 4>(w + GD_W16_W2, b + 64, h1, z);
    relu_h<4>(z, h2);
    dense<4, 4>(w + GD_W16_W3, b + 128, h2, z);
    relu_h<4>(z, h3);
    dense<4, 3>(w + GD_W16_W4, b + 192, h3, y);
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

__device__ __forceinline__ float quad_max(float v)
{
    v = fmaxf(v, __shfl_xor(v, 16));
    return fmaxf(v, __shfl_xor(v, 32));
}
__device__ __forceinline__ float quad_sum(float v)
{
    v += __shfl_xor(v, 16);
    return v + __shfl_xor(v, 32);
}

/* ===========================================================================
 * Inference: the network on 64 samples per wave
 * ========================================================================= */
#define GD_INS 9  /* inputs staged per sample: position in [0,1]^3, normal, outgoing direction */
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
    float in[GB / 64][64][GD_INS];
};

__device__ __forceinline__ void load_infer_weights(InferLds& L, const GuideState& g)
{
    const uint4* src = reinterpret_cast<const uint4*>(g.w16i);
    uint4* dst = reinterpret_cast<uint4*>(L.w);
    for (int i = threadIdx.x; i < GD_NW16I / 8; i += blockDim.x) dst[i] = src[i];
    for (int i = threadIdx.x; i < 4 * GD_H; i += blockDim.x) {
        float v = 0.f;
        if (i < 3 * GD_H) {
            v = g.theta[(i < 64 ? GD_OFF_B1 : i < 128 ? GD_OFF_B2 - 64 : GD_OFF_B3 - 128) + i];
        } else if (i < 3 * GD_H + GD_OUT) {
            const int r = i - 3 * GD_H, ot = r >> 4, m = r & 15;
            const int o = gd_head_output(m >> 2, 4 * ot + (m & 3));
            if (o >= 0) v = g.theta[GD_OFF_B4 + o];
        }
        L.b[i] = v;
    }
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
    if (q == 0) return freq_group(v[2], 8);
    h8 r = {0, 0, 0, 0, 0, 0, 0, 0};
    if (q == 1)
#pragma unroll
        for (int j = 0; j < 6; ++j) r[j] = (_Float16)v[3 + j];
    return r;
}

/* bias of D tile ot in every sample tile */
template <int OT>
__device__ __forceinline__ void bias_init(const float* b, f4 (&out)[4][OT])
{
    const int q = __lane_id() >> 4;
#pragma unroll
    for (int ot = 0; ot < OT; ++ot) {
        const f4 bias = {b[16 * ot + 4 * q], b[16 * ot + 4 * q + 1], b[16 * ot + 4 * q + 2], b[16 * ot + 4 * q + 3]};
#pragma unroll
        for (int st = 0; st < 4; ++st) out[st][ot] = bias;
    }
}

/* out[st][ot] += W tile (ot, kb) x in[st] for every output tile: one weight
 * read from LDS per tile, four MFMAs on it */
template <int KB, int OT>
__device__ __forceinline__ void dense32_k(const _Float16* W, int kb, const h8 (&in)[4], f4 (&out)[4][OT])
{
    const int lane = __lane_id();
#pragma unroll
    for (int ot = 0; ot < OT; ++ot) {
        const h8 a = *reinterpret_cast<const h8*>(W + ((((ot * KB + kb) << 6) + lane) << 3));
#pragma unroll
        for (int st = 0; st < 4; ++st) out[st][ot] = mfma32(a, in[st], out[st][ot]);
    }
}

/* out = W x in + b for a hidden or output layer whose input is two B operands */
template <int OT>
__device__ __forceinline__ void dense32(const _Float16* W, const float* b, const h8 (&in)[4][2], f4 (&out)[4][OT])
{
    bias_init<OT>(b, out);
#pragma unroll
    for (int kb = 0; kb < 2; ++kb) {
        const h8 x[4] = {in[0][kb], in[1][kb], in[2][kb], in[3][kb]};
        dense32_k<2, OT>(W, kb, x, out);
    }
}

/* ReLU of four D tiles into the two B operands of the next layer: pairs
 * converted to FP16 first, then clamped at zero as pairs (rounding keeps the
 * sign, so this is the ReLU of the FP32 value, rounded) */
__device__ __forceinline__ void relu_pack(const f4 (&z)[4][4], h8 (&h)[4][2])
{
    const h2 zero = {(_Float16)0.f, (_Float16)0.f};
#pragma unroll
    for (int st = 0; st < 4; ++st)
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

/* Head outputs of the path vertex (or record) at queue position i: 48
 * halves, lobe j at [5j .. 5j+4] = (mu.xyz, kappa, weight), [40] the
 * probability of sampling the guide.  RAW: the 48 raw network outputs as
 * floats instead (records only, the parity surface).  Path vertices take
 * their inputs from k_guide_gather's compact records, the next chunk's read
 * while this one is computed. */
template <bool FROM_REC, bool RAW>
__global__ __launch_bounds__(GB) RT_GUIDE_NET_OCCUPANCY void k_guide_net(PathState ps, GuideState g, int qsel,
                                                                        const float* rec, int nrec, _Float16* hout,
                                                                        float* yout)
{
    __shared__ InferLds L;
    load_infer_weights(L, g);
    const int lane = __lane_id(), n = lane & 15, q = lane >> 4, w = threadIdx.x >> 6;
    const int cnt = FROM_REC ? nrec : ps.cnt->q_count[qsel];
    if (!FROM_REC && blockIdx.x == 0 && threadIdx.x == 0 && cnt > 0) atomicAdd(g.inferred, (unsigned long long)cnt);
    float (*IN)[GD_INS] = L.in[w];
    const int stride = gridDim.x * (GB / 64) * 64;
    int base = (blockIdx.x * (GB / 64) + w) * 64;
    float4 na = make_float4(0.f, 0.f, 0.f, 0.f);
    float2 nb = make_float2(0.f, 0.f);
    if (!FROM_REC && base + lane < cnt) {
        na = g.in_a[base + lane];
        nb = g.in_b[base + lane];
    }
    for (; base < cnt; base += stride) {
        /* stage: lane l holds sample base + l */
        {
            const int i = base + lane;
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
                if (j < cnt) {
                    na = g.in_a[j];
                    nb = g.in_b[j];
                }
            }
#pragma unroll
            for (int c = 0; c < GD_INS; ++c) IN[lane][c] = v[c];
        }
        wave_lds_sync();
        /* layer 1, one K block of features at a time */
        f4 z[4][4];
        bias_init<4>(L.b, z);
#pragma unroll
        for (int kb = 0; kb < 3; ++kb) {
            h8 x[4];
#pragma unroll
            for (int st = 0; st < 4; ++st) x[st] = feature_group(kb, q, IN[16 * st + n]);
            dense32_k<3, 4>(L.w + GD_W16I_W1, kb, x, z);
        }
        wave_lds_sync();
        h8 h[4][2];
        relu_pack(z, h);
        dense32<4>(L.w + GD_W16I_W2, L.b + GD_H, h, z);
        relu_pack(z, h);
        dense32<4>(L.w + GD_W16I_W3, L.b + 2 * GD_H, h, z);
        relu_pack(z, h);
        f4 y[4][3];
        dense32<3>(L.w + GD_W16I_W4, L.b + 3 * GD_H, h, y);
#pragma unroll
        for (int st = 0; st < 4; ++st) {
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
            if (valid) {
                h2* o = reinterpret_cast<h2*>(hout + (size_t)i * GD_PARAMS + 10 * q);
                o[0] = h2{(_Float16)mu0.x, (_Float16)mu0.y};
                o[1] = h2{(_Float16)mu0.z, (_Float16)k0};
                o[2] = h2{(_Float16)(e0 * inv), (_Float16)mu1.x};
                o[3] = h2{(_Float16)mu1.y, (_Float16)mu1.z};
                o[4] = h2{(_Float16)k1, (_Float16)(e1 * inv)};
                if (q == 0) hout[(size_t)i * GD_PARAMS + 40] = (_Float16)(0.05f + 0.9f * __builtin_amdgcn_rcpf(1.f + __expf(-r[10])));
            }
        }
    }
}

hipError_t launch_guide_infer(const HostDevScene& s, const PathState& ps, const GuideState& g, int qsel, int grid,
                              hipStream_t st)
{
    hipLaunchKernelGGL(k_guide_gather, dim3(grid * 4), dim3(GB), 0, st, as_dev_g(s), ps, g, qsel);
    hipLaunchKernelGGL((k_guide_net<false, false>), dim3(grid), dim3(GB), 0, st, ps, g, qsel, (const float*)nullptr, 0,
                       g.out, (float*)nullptr);
    return hipGetLastError();
}

/* the network on records (parity surface): raw outputs y [n x 48] */
hipError_t launch_guide_forward(const GuideState& g, const float* rec, int n, float* y, int grid, hipStream_t st)
{
    hipLaunchKernelGGL((k_guide_net<true, true>), dim3(grid), dim3(GB), 0, st, PathState{}, g, 0, rec, n,
                       (_Float16*)nullptr, y);
    return hipGetLastError();
}

/* head outputs of records as the shading kernel receives them [n x 48 halves] */
hipError_t launch_guide_head(const GuideState& g, const float* rec, int n, _Float16* out, int grid, hipStream_t st)
{
    hipLaunchKernelGGL((k_guide_net<true, false>), dim3(grid), dim3(GB), 0, st, PathState{}, g, 0, rec, n, out,
                       (float*)nullptr);
    return hipGetLastError();
}

__device__ __forceinline__ void record_inputs(const GuideState& g, const float* r, float* p01, vec3& nf, vec3& wo)
{
    p01[0] = (r[0] - g.lo[0]) * g.inv_ext[0];
    p01[1] = (r[1] - g.lo[1]) * g.inv_ext[1];
    p01[2] = (r[2] - g.lo[2]) * g.inv_ext[2];
    nf = mk(r[3], r[4], r[5]);
    wo = mk(r[6], r[7], r[8]);
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
        if (s != 0.f) atomicAdd(&g.acc[0], s);
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
/* Rows of 64 samples padded by 4 halves: a 16-lane group reading 8 bytes
 * from 16 consecutive rows then touches 32 distinct 4-byte banks. */
#define GD_ROW (64 + 4)
struct TrainLds {
    WLds<GD_NW16> wl;
    float Y[4][GD_OUT][16];
    _Float16 act[GD_IN + 3 * GD_H][GD_ROW];   /* x, h1, h2, h3 of the block's 64 samples */
    _Float16 dz[3 * GD_H + GD_OUT][GD_ROW];   /* dz1, dz2, dz3, dy */
};

/* tile t of the block's 64 weight-gradient tiles -> (layer, out tile, in tile) */
__device__ __forceinline__ void grad_tile(int t, int& layer, int& ot, int& it)
{
    if (t < 20) { layer = 0; ot = t / 5; it = t % 5; }
    else if (t < 36) { layer = 1; ot = (t - 20) / 4; it = (t - 20) % 4; }
    else if (t < 52) { layer = 2; ot = (t - 36) / 4; it = (t - 36) % 4; }
    else { layer = 3; ot = (t - 52) / 4; it = (t - 52) % 4; }
}

__global__ __launch_bounds__(GB) void k_guide_train(GuideState g, const float* rec0, const int32_t* count, int off,
                                                   int max_n)
{
    __shared__ TrainLds S;
    load_weights(S.wl, g);
    const int lane = __lane_id(), n = lane & 15, q = lane >> 4, w = threadIdx.x >> 6;
    const int N = chunk_len(g, count, off, max_n);
    const float* rec = rec0 + (size_t)off * GD_REC;
    const float wsum = g.acc[0];
    const float wscale = wsum > 0.f ? (float)N / wsum : 0.f;
    f4 gacc[16];
#pragma unroll
    for (int t = 0; t < 16; ++t) gacc[t] = f4{0.f, 0.f, 0.f, 0.f};
    float bacc = 0.f, lacc = 0.f;
    const int col = 16 * w + n;
    for (int base = blockIdx.x * 64; base < N; base += gridDim.x * 64) {
        int i = base + col;
        bool valid = i < N;
        const float* r = rec + (size_t)(valid ? i : 0) * GD_REC;
        float p01[3] = {0.f, 0.f, 0.f};
        vec3 nf = mk(0.f, 0.f, 0.f), wo = nf, om = mk(0.f, 0.f, 1.f);
        float pb = 1.f, sw = 0.f;
        if (valid) {
            record_inputs(g, r, p01, nf, wo);
            om = mk(r[9], r[10], r[11]);
            pb = r[12];
            sw = fminf(r[13] * wscale, GD_WCLAMP);
        }
        h4 x[5], h1[4], h2[4], h3[4];
        f4 y[3];
        _Float16 (*F)[GD_FROW] = reinterpret_cast<_Float16 (*)[GD_FROW]>(&S.Y[w][0][0]);
        features_to_lds(F, n, q, p01, nf, wo);
        __syncthreads();
        features_from_lds(F, n, q, x);
        forward16(S.wl.w, S.wl.b, x, h1, h2, h3, y);
        /* activations to LDS, [neuron][sample] */
#pragma unroll
        for (int kb = 0; kb < 5; ++kb)
#pragma unroll
            for (int e = 0; e < 4; ++e) S.act[16 * kb + 4 * q + e][col] = x[kb][e];
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                S.act[GD_IN + 16 * t + 4 * q + e][col] = h1[t][e];
                S.act[GD_IN + 64 + 16 * t + 4 * q + e][col] = h2[t][e];
                S.act[GD_IN + 128 + 16 * t + 4 * q + e][col] = h3[t][e];
            }
#pragma unroll
        for (int t = 0; t < 3; ++t)
#pragma unroll
            for (int e = 0; e < 4; ++e) S.Y[w][16 * t + 4 * q + e][n] = y[t][e];
        __syncthreads();

        /* head: loss and dL/dy of lobes 2q, 2q+1 (and alpha on q == 0) */
        float rr[2][5];
        vec3 mu[2];
        float mlen[2], kap[2], v[2];
        bool kin[2];
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            int j = 2 * q + s;
#pragma unroll
            for (int c = 0; c < 5; ++c) rr[s][c] = S.Y[w][5 * j + c][n];
            lobe_head(rr[s], mu[s], mlen[s], kap[s], kin[s]);
            v[s] = vmf_pdf(mu[s], kap[s], om);
        }
        float mx = quad_max(fmaxf(rr[0][4], rr[1][4]));
        float e0 = expf(rr[0][4] - mx), e1 = expf(rr[1][4] - mx);
        float inv = 1.f / quad_sum(e0 + e1);
        float pi0 = e0 * inv, pi1 = e1 * inv;
        float qd = quad_sum(pi0 * v[0] + pi1 * v[1]);
        float sg = 1.f / (1.f + expf(-S.Y[w][40][n]));
        float alpha = 0.05f + 0.9f * sg;
        float pm = alpha * qd + (1.f - alpha) * pb;
        float c = valid ? -sw / pm : 0.f;
        float dq = c * alpha;
        if (valid && q == 0) lacc += -sw * logf(pm);
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            int j = 2 * q + s;
            float pij = s ? pi1 : pi0;
            float dv = dq * pij;
            float cw = dot(mu[s], om);
            float em = expf(-2.f * kap[s]);
            float dlogc = 1.f - 2.f * kap[s] * em / (-expm1f(-2.f * kap[s]));
            float dk = kin[s] ? dv * v[s] * (dlogc + kap[s] * (cw - 1.f)) : 0.f;
            float gm = dv * v[s] * kap[s] / mlen[s];
            vec3 dm = (om - mu[s] * cw) * gm;
            float da = dq * pij * (v[s] - qd);
            S.dz[3 * GD_H + 5 * j][col] = to_h_sat(dm.x);
            S.dz[3 * GD_H + 5 * j + 1][col] = to_h_sat(dm.y);
            S.dz[3 * GD_H + 5 * j + 2][col] = to_h_sat(dm.z);
            S.dz[3 * GD_H + 5 * j + 3][col] = to_h_sat(dk);
            S.dz[3 * GD_H + 5 * j + 4][col] = to_h_sat(da);
        }
        if (q == 0) S.dz[3 * GD_H + 40][col] = to_h_sat(c * (qd - pb) * 0.9f * sg * (1.f - sg));
        if (q == 1)
            for (int o = 41; o < 48; ++o) S.dz[3 * GD_H + o][col] = (_Float16)0.f;
        __syncthreads();

        /* back-propagate: dy -> dz3 -> dz2 -> dz1 (B-operand layout in registers) */
        h4 dy[3], d3[4], d2[4], d1[4];
        f4 z[4];
#pragma unroll
        for (int t = 0; t < 3; ++t)
#pragma unroll
            for (int e = 0; e < 4; ++e) dy[t][e] = S.dz[3 * GD_H + 16 * t + 4 * q + e][col];
        dense_t<3, 4>(S.wl.w + GD_W16T_W4, dy, z);
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                d3[t][e] = h3[t][e] > (_Float16)0.f ? to_h_sat(z[t][e]) : (_Float16)0.f;
                S.dz[2 * GD_H + 16 * t + 4 * q + e][col] = d3[t][e];
            }
        dense_t<4, 4>(S.wl.w + GD_W16T_W3, d3, z);
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                d2[t][e] = h2[t][e] > (_Float16)0.f ? to_h_sat(z[t][e]) : (_Float16)0.f;
                S.dz[GD_H + 16 * t + 4 * q + e][col] = d2[t][e];
            }
        dense_t<4, 4>(S.wl.w + GD_W16T_W2, d2, z);
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                d1[t][e] = h1[t][e] > (_Float16)0.f ? to_h_sat(z[t][e]) : (_Float16)0.f;
                S.dz[16 * t + 4 * q + e][col] = d1[t][e];
            }
        __syncthreads();

        /* weight gradients: wave w owns tiles 16 w .. 16 w + 15, K = 64 samples */
#pragma unroll
        for (int tt = 0; tt < 16; ++tt) {
            int layer, ot, it;
            grad_tile(16 * w + tt, layer, ot, it);
            int zrow = layer == 3 ? 3 * GD_H : layer * GD_H;   /* delta rows */
            int arow = layer == 0 ? 0 : GD_IN + (layer - 1) * GD_H; /* input rows */
#pragma unroll
            for (int ks = 0; ks < 4; ++ks) {
                h4 a = *reinterpret_cast<const h4*>(&S.dz[zrow + 16 * ot + n][16 * ks + 4 * q]);
                h4 b = *reinterpret_cast<const h4*>(&S.act[arow + 16 * it + n][16 * ks + 4 * q]);
                gacc[tt] = mfma16(a, b, gacc[tt]);
            }
        }
        /* bias gradients: thread t < 240 sums one delta row */
        if (threadIdx.x < 3 * GD_H + GD_OUT) {
            int row = threadIdx.x;  /* dz1 0..63, dz2 64..127, dz3 128..191, dy 192..239 */
            float s = 0.f;
#pragma unroll
            for (int k = 0; k < 64; k += 4) {
                h4 v = *reinterpret_cast<const h4*>(&S.dz[row][k]);
                s += ((float)v[0] + (float)v[1]) + ((float)v[2] + (float)v[3]);
            }
            bacc += s;
        }
        __syncthreads();
    }
    /* flush */
#pragma unroll
    for (int tt = 0; tt < 16; ++tt) {
        int layer, ot, it;
        grad_tile(16 * w + tt, layer, ot, it);
        int in_dim = layer == 0 ? GD_IN : GD_H;
        int off = layer == 0 ? GD_OFF_W1 : layer == 1 ? GD_OFF_W2 : layer == 2 ? GD_OFF_W3 : GD_OFF_W4;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            int o = 16 * ot + 4 * q + e, ii = 16 * it + n;
            float val = gacc[tt][e];
            if (val != 0.f) atomicAdd(&g.grad[off + o * in_dim + ii], val);
        }
    }
    if (threadIdx.x < 3 * GD_H + GD_OUT && bacc != 0.f) {
        int row = threadIdx.x;
        int off = row < 64 ? GD_OFF_B1 + row : row < 128 ? GD_OFF_B2 + row - 64 : row < 192 ? GD_OFF_B3 + row - 128 : GD_OFF_B4 + row - 192;
        atomicAdd(&g.grad[off], bacc);
    }
    for (int o = 32; o > 0; o >>= 1) lacc += __shfl_xor(lacc, o);
    if (lane == 0 && lacc != 0.f) atomicAdd(&g.acc[1], lacc);
}

hipError_t launch_guide_train(const GuideState& g, const float* rec, const int32_t* count, int off, int max_n,
                              int grid, hipStream_t st)
{
    hipLaunchKernelGGL(k_guide_train, dim3(grid), dim3(GB), 0, st, g, rec, count, off, max_n);
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
        int fw, bw;
        gd_w16_slots(i, &fw, &bw);
        if (fw >= 0) g.w16[fw] = (_Float16)th;
        if (bw >= 0) g.w16[bw] = (_Float16)th;
        const int iw = gd_w16i_slot(i);
        if (iw >= 0) g.w16i[iw] = (_Float16)th;
    }
}

hipError_t launch_guide_adam(const GuideState& g, float lr, int step, const int32_t* count, int off, int max_n,
                             hipStream_t st)
{
    hipLaunchKernelGGL(k_guide_adam, dim3(32), dim3(GB), 0, st, g, lr, step, count, off, max_n);
    return hipGetLastError();
}

__global__ void k_guide_reset(GuideState g, int slots)
{
    int i = blockIdx.x * GB + threadIdx.x;
    if (i == 0) {
        *g.rec_count = 0;
        g.acc[0] = g.acc[1] = g.acc[2] = 0.f;
    }
    for (; i < slots; i += gridDim.x * GB) g.open_rec[i] = -1;
}

hipError_t launch_guide_reset_records(const GuideState& g, int slots, hipStream_t st)
{
    int grid = std::max(1, std::min(1024, (slots + GB - 1) / GB));
    hipLaunchKernelGGL(k_guide_reset, dim3(grid), dim3(GB), 0, st, g, slots);
    return hipGetLastError();
}
