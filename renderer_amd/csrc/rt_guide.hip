/*
 * rt_guide.hip — gfx950 kernels of the path-guiding network.
 *
 * Matrix products use v_mfma_f32_16x16x16_f16 in the transposed form
 * Z[out x samples] = W[out x in] * X[in x samples]: a wave carries 16
 * samples; lane l holds sample l % 16 and the feature quad 4 (l / 16)..+3.
 * With that orientation the D tile of one layer (4 consecutive outputs of
 * one sample per lane) is exactly the B operand of the next layer's K block,
 * so activations never leave registers between layers.  Weight gradients
 * (a reduction over samples) read activations and deltas back from LDS in
 * [neuron][sample] order, which gives both MFMA operands as contiguous
 * 8-byte reads.
 */
#include "rt_device.h"
#include "rt_guide.h"
#include "rt_guide_dev.h"
#include "rt_kernels.h"

#include <cstring>

typedef _Float16 h4 __attribute__((ext_vector_type(4)));
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
 * inference: head outputs for every live path at a surface hit
 * ========================================================================= */
__global__ __launch_bounds__(GB) void k_guide_infer(DevScene S, PathState ps, GuideState g, int qsel)
{
    __shared__ WLds<GD_NW16F> L;
    __shared__ float Y[4][GD_OUT][16];
    load_weights(L, g);
    const int lane = __lane_id(), n = lane & 15, q = lane >> 4, w = threadIdx.x >> 6;
    const int cnt = ps.cnt->q_count[qsel];
    if (blockIdx.x == 0 && threadIdx.x == 0 && cnt > 0) atomicAdd(g.inferred, (unsigned long long)cnt);
    for (int base = blockIdx.x * 64; base < cnt; base += gridDim.x * 64) {
        int i = base + 16 * w + n;
        bool valid = i < cnt;
        int p = valid ? ps.queue[qsel][i] : 0;
        float p01[3] = {0.f, 0.f, 0.f};
        vec3 nf = mk(0.f, 0.f, 0.f), wo = mk(0.f, 0.f, 0.f);
        if (valid) {
            float4 h = ps.hit[p];
            int gg = __float_as_int(h.y);
            if (gg >= 0) {
                vec3 o = xyz(ps.ray_o[p]), d = xyz(ps.ray_d[p]);
                vec3 x = o + h.x * d;
                vec3 nn = hit_normal(S, gg, x, h.z, h.w);
                nf = dot(nn, d) > 0.f ? -nn : nn;
                wo = -d;
                p01[0] = (x.x - g.lo[0]) * g.inv_ext[0];
                p01[1] = (x.y - g.lo[1]) * g.inv_ext[1];
                p01[2] = (x.z - g.lo[2]) * g.inv_ext[2];
            } else {
                valid = false;
            }
        }
        h4 x[5], h1[4], h2[4], h3[4];
        f4 y[3];
        /* the wave's feature rows share its output staging area */
        _Float16 (*F)[GD_FROW] = reinterpret_cast<_Float16 (*)[GD_FROW]>(&Y[w][0][0]);
        features_to_lds(F, n, q, p01, nf, wo);
        __syncthreads();
        features_from_lds(F, n, q, x);
        forward16(L.w, L.b, x, h1, h2, h3, y);
#pragma unroll
        for (int t = 0; t < 3; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r) Y[w][16 * t + 4 * q + r][n] = y[t][r];
        __syncthreads();
        float a[2];
        vec3 mu[2];
        float kap[2];
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            int j = 2 * q + s;
            float r[5];
#pragma unroll
            for (int c = 0; c < 5; ++c) r[c] = Y[w][5 * j + c][n];
            float mlen;
            bool kin;
            lobe_head(r, mu[s], mlen, kap[s], kin);
            a[s] = r[4];
        }
        float mx = quad_max(fmaxf(a[0], a[1]));
        float e0 = expf(a[0] - mx), e1 = expf(a[1] - mx);
        float inv = 1.f / quad_sum(e0 + e1);
        if (valid) {
            float* o = g.out + (size_t)p * GD_PARAMS;
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                int j = 2 * q + s;
                o[5 * j] = mu[s].x;
                o[5 * j + 1] = mu[s].y;
                o[5 * j + 2] = mu[s].z;
                o[5 * j + 3] = kap[s];
                o[5 * j + 4] = (s ? e1 : e0) * inv;
            }
            if (q == 0) o[40] = 0.05f + 0.9f / (1.f + expf(-Y[w][40][n]));
        }
        __syncthreads();
    }
}

hipError_t launch_guide_infer(const HostDevScene& s, const PathState& ps, const GuideState& g, int qsel, int grid,
                              hipStream_t st)
{
    hipLaunchKernelGGL(k_guide_infer, dim3(grid), dim3(GB), 0, st, as_dev_g(s), ps, g, qsel);
    return hipGetLastError();
}

/* ===========================================================================
 * forward only on records (parity surface): raw outputs y [n x 48]
 * ========================================================================= */
__device__ __forceinline__ void record_inputs(const GuideState& g, const float* r, float* p01, vec3& nf, vec3& wo)
{
    p01[0] = (r[0] - g.lo[0]) * g.inv_ext[0];
    p01[1] = (r[1] - g.lo[1]) * g.inv_ext[1];
    p01[2] = (r[2] - g.lo[2]) * g.inv_ext[2];
    nf = mk(r[3], r[4], r[5]);
    wo = mk(r[6], r[7], r[8]);
}

__global__ __launch_bounds__(GB) void k_guide_forward(GuideState g, const float* rec, int N, float* yout)
{
    __shared__ WLds<GD_NW16F> L;
    __shared__ _Float16 F[4][16][GD_FROW];
    load_weights(L, g);
    const int lane = __lane_id(), n = lane & 15, q = lane >> 4, w = threadIdx.x >> 6;
    for (int base = blockIdx.x * 64; base < N; base += gridDim.x * 64) {
        int i = base + 16 * w + n;
        bool valid = i < N;
        float p01[3] = {0.f, 0.f, 0.f};
        vec3 nf = mk(0.f, 0.f, 0.f), wo = nf;
        if (valid) record_inputs(g, rec + (size_t)i * GD_REC, p01, nf, wo);
        h4 x[5], h1[4], h2[4], h3[4];
        f4 y[3];
        features_to_lds(F[w], n, q, p01, nf, wo);
        __syncthreads();
        features_from_lds(F[w], n, q, x);
        forward16(L.w, L.b, x, h1, h2, h3, y);
        if (valid)
#pragma unroll
            for (int t = 0; t < 3; ++t)
#pragma unroll
                for (int r = 0; r < 4; ++r) yout[(size_t)i * GD_OUT + 16 * t + 4 * q + r] = y[t][r];
    }
}

hipError_t launch_guide_forward(const GuideState& g, const float* rec, int n, float* y, int grid, hipStream_t st)
{
    hipLaunchKernelGGL(k_guide_forward, dim3(grid), dim3(GB), 0, st, g, rec, n, y);
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
