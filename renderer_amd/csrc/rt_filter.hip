/*
 * rt_filter.hip — the image passes of MetropolisFltRenderer
 * (src/metropolisfltrenderer.cpp:207-320) and its symmetric filter
 * (src/symmetricfilter.cpp:40-137) on gfx950.
 *
 * Images are H x W x 3 floats (RGB interleaved, like films) or H x W floats.
 * The symmetric filter visits the offsets d of half its window (the pair
 * +d / -d at once) one after the other; per offset it computes three patch
 * distances per pixel (the guide at p - d, at p + d and their mean against
 * the guide at p, less the variances), box-blurs them (separable, two
 * passes), turns them into weights, blurs those, and adds the weighted
 * neighbours into per-pixel sums.  Every pass is an element-wise or
 * row / column sweep over the image: each lane owns a pixel, so loads of a
 * wave are consecutive.  The arithmetic keeps the order of the oracle's
 * restatement (no contraction), so the two agree up to expf's last bit.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "rt_kernels.h"

#define FB 256

namespace {

#pragma clang fp contract(off)

__device__ __forceinline__ int reflect(int i, int n)  /* BORDER_REFLECT: fedcba|abcdef|fedcba */
{
    while (i < 0 || i >= n) i = i < 0 ? -i - 1 : 2 * n - i - 1;
    return i;
}
__device__ __forceinline__ int reflect101(int i, int n)  /* BORDER_REFLECT_101: edcb|abcde|dcba */
{
    int r = i < 0 ? -i : i >= n ? 2 * n - i - 2 : i;
    return min(max(r, 0), n - 1);
}
__device__ __forceinline__ float chan_mean3(float a, float b, float c) { return (a / 3.f + b / 3.f) + c / 3.f; }

/* mathhelper var() of two images */
__global__ void k_var2(const float* __restrict__ a, const float* __restrict__ b, float* mean, float* var, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * FB + threadIdx.x; i < n; i += (int64_t)gridDim.x * FB) {
        float sum = a[i] + b[i], sq = a[i] * a[i] + b[i] * b[i];
        float m = sum * 0.5f;
        mean[i] = m;
        var[i] = fmaxf((sq - sum * m) / 2.f, 0.f);
    }
}

/* offset (dx, dy): patch distances of the three candidate sources and their
 * denominators (computeWeights, src/symmetricfilter.cpp:46-58) */
__global__ void k_sf_dist(const float* __restrict__ guide, const float* __restrict__ var, int W, int H, int dx, int dy,
                          float* d2, float* den)
{
    const int64_t n = (int64_t)W * H;
    const float h2 = 0.4f * 0.4f;
    const int x = blockIdx.x * FB + threadIdx.x;
    for (int y = blockIdx.y; y < H && x < W; y += gridDim.y) {
        const int64_t i = (int64_t)y * W + x;
        const bool in1 = y - dy >= 0 && y - dy < H && x - dx >= 0 && x - dx < W;
        const bool in2 = y + dy >= 0 && y + dy < H && x + dx >= 0 && x + dx < W;
        const int64_t j1 = in1 ? (int64_t)(y - dy) * W + (x - dx) : 0, j2 = in2 ? (int64_t)(y + dy) * W + (x + dx) : 0;
        float dd[3][3], vv[3][3];
        for (int c = 0; c < 3; ++c) {
            float s1 = in1 ? guide[3 * j1 + c] : 0.f, s2 = in2 ? guide[3 * j2 + c] : 0.f;
            float v1 = in1 ? var[3 * j1 + c] : 0.f, v2 = in2 ? var[3 * j2 + c] : 0.f;
            float sS = (s1 + s2) / 2.f, vS = (v1 + v2) / 4.f;
            float t = guide[3 * i + c], v = var[3 * i + c];
            const float src[3] = {s1, s2, sS}, sv[3] = {v1, v2, vS};
            for (int k = 0; k < 3; ++k) {
                float df = src[k] - t;
                dd[k][c] = df * df - (v + fminf(sv[k], v));
                vv[k][c] = v + sv[k];
            }
        }
        for (int k = 0; k < 3; ++k) {
            d2[k * n + i] = chan_mean3(dd[k][0], dd[k][1], dd[k][2]);
            den[k * n + i] = 1e-8f + h2 * chan_mean3(vv[k][0], vv[k][1], vv[k][2]);
        }
    }
}

/* normalised p x p box filter of three planes, BORDER_REFLECT: row sums,
 * then column sums times 1 / p^2 */
__global__ void k_box_rows(const float* __restrict__ in, float* out, int W, int H, int p)
{
    const int64_t n = (int64_t)W * H;
    const int r = p / 2;
    const int x = blockIdx.x * FB + threadIdx.x;
    for (int y = blockIdx.y; y < H && x < W; y += gridDim.y)
    for (int k = 0; k < 3; ++k) {
        const float* row = in + k * n + (int64_t)y * W;
        float s = 0.f;
        for (int q = -r; q <= r; ++q) s += row[reflect(x + q, W)];
        out[k * n + (int64_t)y * W + x] = s;
    }
}
__global__ void k_box_cols(const float* __restrict__ in, float* out, int W, int H, int p)
{
    const int64_t n = (int64_t)W * H;
    const int r = p / 2;
    const float inv = 1.f / (float)(p * p);
    const int x = blockIdx.x * FB + threadIdx.x;
    for (int y = blockIdx.y; y < H && x < W; y += gridDim.y)
    for (int k = 0; k < 3; ++k) {
        const float* pl = in + k * n;
        float s = 0.f;
        for (int q = -r; q <= r; ++q) s += pl[(int64_t)reflect(y + q, H) * W + x];
        out[k * n + (int64_t)y * W + x] = s * inv;
    }
}

/* raw weights exp(-max(d2, 0) / den) of the three planes */
__global__ void k_sf_exp(const float* __restrict__ d2, const float* __restrict__ den, float* out, int64_t n3)
{
    for (int64_t i = (int64_t)blockIdx.x * FB + threadIdx.x; i < n3; i += (int64_t)gridDim.x * FB)
        out[i] = expf(-fmaxf(d2[i], 0.f) / den[i]);
}

/* weights (src/symmetricfilter.cpp:62-69, 114-125) and the weighted
 * neighbours p - d, p + d added to the sums */
__global__ void k_sf_accum(const float* __restrict__ w, const float* __restrict__ image, int W, int H, int dx, int dy,
                           float* area, float* acc)
{
    const int64_t n = (int64_t)W * H;
    const int x = blockIdx.x * FB + threadIdx.x;
    for (int y = blockIdx.y; y < H && x < W; y += gridDim.y) {
        const int64_t i = (int64_t)y * W + x;
        float ww[3];
        for (int k = 0; k < 3; ++k) {
            float x = fmaxf(w[k * n + i], 0.f);
            ww[k] = x < 0.05f ? 0.f : x;
        }
        float w1 = ww[0], w2 = ww[1];
        if (ww[2] > fmaxf(w1, w2)) w1 = w2 = ww[2];
        const bool in1 = y - dy >= 0 && y - dy < H && x - dx >= 0 && x - dx < W;
        const bool in2 = y + dy >= 0 && y + dy < H && x + dx >= 0 && x + dx < W;
        const int64_t j1 = (int64_t)(y - dy) * W + (x - dx), j2 = (int64_t)(y + dy) * W + (x + dx);
        area[i] += w1 + w2;
        for (int c = 0; c < 3; ++c)
            acc[3 * i + c] += w1 * (in1 ? image[3 * j1 + c] : 0.f) + w2 * (in2 ? image[3 * j2 + c] : 0.f);
    }
}

__global__ void k_sf_init(float* area, float* acc, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * FB + threadIdx.x; i < n; i += (int64_t)gridDim.x * FB) {
        area[i] = 0.f;
        for (int c = 0; c < 3; ++c) acc[3 * i + c] = 0.f;
    }
}

/* the offset d = 0, reached last (src/symmetricfilter.cpp:104-110) */
__global__ void k_sf_centre(const float* __restrict__ image, float* area, float* acc, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * FB + threadIdx.x; i < n; i += (int64_t)gridDim.x * FB) {
        area[i] += 1.f;
        for (int c = 0; c < 3; ++c) acc[3 * i + c] += image[3 * i + c];
    }
}

__global__ void k_sf_final(const float* __restrict__ area, const float* __restrict__ acc, float* out, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * FB + threadIdx.x; i < n; i += (int64_t)gridDim.x * FB)
        for (int c = 0; c < 3; ++c) out[3 * i + c] = acc[3 * i + c] / area[i];
}

/* sepFilter2D(noisy, .., [.1 .8 .1], [.1 .8 .1]) with BORDER_REFLECT_101 */
__global__ void k_sep3(const float* __restrict__ in, float* out, int W, int H, int vertical)
{
    const int64_t n = (int64_t)W * H;
    const float k3[3] = {.1f, .8f, .1f};
    for (int64_t i = (int64_t)blockIdx.x * FB + threadIdx.x; i < n; i += (int64_t)gridDim.x * FB) {
        const int y = (int)(i / W), x = (int)(i % W);
        for (int c = 0; c < 3; ++c) {
            float s = 0.f;
            for (int j = -1; j <= 1; ++j) {
                int64_t q = vertical ? (int64_t)reflect101(y + j, H) * W + x : (int64_t)y * W + reflect101(x + j, W);
                s += k3[j + 1] * in[3 * q + c];
            }
            out[3 * i + c] = s;
        }
    }
}

/* importance from the relative variance (src/metropolisfltrenderer.cpp:231-236),
 * before normalisation */
__global__ void k_flt_importance(const float* __restrict__ vof, const float* __restrict__ fmean, float* imp, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * FB + threadIdx.x; i < n; i += (int64_t)gridDim.x * FB) {
        float vm = chan_mean3(vof[3 * i], vof[3 * i + 1], vof[3 * i + 2]);
        float mm = chan_mean3(fmean[3 * i], fmean[3 * i + 1], fmean[3 * i + 2]);
        float rel = vm / (mm * mm + 1e-2f) + 1e-8f;
        imp[i] = -20.f * (expf(-rel) - 1.f) + 1.f;
    }
}

__global__ void k_scale(float* x, float k, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * FB + threadIdx.x; i < n; i += (int64_t)gridDim.x * FB) x[i] *= k;
}

/* per pixel: channelNorm(film) * importance (b of a later pass) */
__global__ void k_flt_bnorm(const float* __restrict__ film, const float* __restrict__ imp, float* out, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * FB + threadIdx.x; i < n; i += (int64_t)gridDim.x * FB) {
        float r = film[3 * i], g = film[3 * i + 1], b = film[3 * i + 2];
        out[i] = sqrtf(r * r + g * g + b * b) * imp[i];
    }
}

/* newsum = sumimp + imp; films and m2 rescaled by sumimp / newsum (:126-134) */
__global__ void k_flt_rescale(const float* __restrict__ sumimp, const float* __restrict__ imp, float* newsum,
                              float* f0, float* f1, float* m0, float* m1, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * FB + threadIdx.x; i < n; i += (int64_t)gridDim.x * FB) {
        float ns = sumimp[i] + imp[i];
        newsum[i] = ns;
        float r = sumimp[i] / ns;
        for (int c = 0; c < 3; ++c) {
            f0[3 * i + c] = r * f0[3 * i + c];
            f1[3 * i + c] = r * f1[3 * i + c];
            m0[3 * i + c] = r * (r * m0[3 * i + c]);
            m1[3 * i + c] = r * (r * m1[3 * i + c]);
        }
    }
}

/* a pass's deposits (sum w, sum w x, sum w x^2 in fixed point) merged into
 * the film and the weighted running mean / variance of addSample
 * (src/metropolisfltrenderer.cpp:54-75): the merge of two weighted sets,
 * which equals continuing the running update with the pass's samples */
__global__ void k_flt_merge(const unsigned long long* __restrict__ aw, const unsigned long long* __restrict__ ax,
                            const unsigned long long* __restrict__ aq, double uw, double ux, double uq, float* film,
                            float* mean, float* m2, float* sw, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * FB + threadIdx.x; i < n; i += (int64_t)gridDim.x * FB) {
        const double Wb = (double)aw[i] * uw;
        if (!(Wb > 0.0)) continue;
        const double Wa = sw[i], W = Wa + Wb;
        for (int c = 0; c < 3; ++c) {
            const double S = (double)ax[3 * i + c] * ux, Q = (double)aq[3 * i + c] * uq;
            const double mb = S / Wb, m2b = fmax(Q - Wb * mb * mb, 0.0);
            const double ma = mean[3 * i + c], delta = mb - ma;
            mean[3 * i + c] = (float)(ma + delta * (Wb / W));
            m2[3 * i + c] = (float)((double)m2[3 * i + c] + m2b + delta * delta * (Wa * Wb / W));
            film[3 * i + c] = (float)((double)film[3 * i + c] + S);
        }
        sw[i] = (float)W;
    }
}

__global__ void k_add(float* dst, const float* __restrict__ src, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * FB + threadIdx.x; i < n; i += (int64_t)gridDim.x * FB) dst[i] += src[i];
}

/* biased variance m2 / (sumweight + 1e-8) (src/metropolisfltrenderer.cpp:199-203) */
__global__ void k_flt_bvar(const float* __restrict__ m2, const float* __restrict__ sw, float* out, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * FB + threadIdx.x; i < n; i += (int64_t)gridDim.x * FB)
        for (int c = 0; c < 3; ++c) out[3 * i + c] = m2[3 * i + c] / (sw[i] + 1e-8f);
}

#pragma clang fp contract(fast)

int blocks(int64_t n, int grid)
{
    return (int)std::max<int64_t>(1, std::min<int64_t>((n + FB - 1) / FB, grid));
}

}  // namespace

hipError_t launch_var2(const float* a, const float* b, float* mean, float* var, int64_t n, int grid, hipStream_t st)
{
    hipLaunchKernelGGL(k_var2, dim3(blocks(n, grid)), dim3(FB), 0, st, a, b, mean, var, n);
    return hipGetLastError();
}

/* SymmetricFilter(patch, radius).filter(image, guide, pixvar) into out.
 * tmp: RT_SYM_FILTER_TMP * W * H floats of scratch (three distance planes,
 * three denominators, two blur buffers of three planes, the weight sums of
 * one channel and the accumulated values of three). */
hipError_t launch_sym_filter(const float* image, const float* guide, const float* pixvar, int W, int H, int patch,
                             int radius, float* tmp, float* out, int grid, hipStream_t st)
{
    const int64_t n = (int64_t)W * H;
    float *d2 = tmp, *den = tmp + 3 * n, *b0 = tmp + 6 * n, *b1 = tmp + 9 * n, *area = tmp + 12 * n,
          *acc = tmp + 13 * n;  /* acc ends at 16 n = RT_SYM_FILTER_TMP * n */
    static_assert(RT_SYM_FILTER_TMP == 16, "scratch layout");
    const int g1 = blocks(n, grid), g3 = blocks(3 * n, grid);
    const dim3 rows((W + FB - 1) / FB, std::min(H, 65535)); /* block rows walk image rows: no index division */
    hipLaunchKernelGGL(k_sf_init, dim3(g1), dim3(FB), 0, st, area, acc, n);
    for (int dx = -radius; dx <= 0; ++dx)
        for (int dy = -radius; dy <= radius; ++dy) {
            if (dx == 0 && dy == 0) {  /* the centre; dx = 0, dy > 0 are the pairs of dy < 0 */
                hipLaunchKernelGGL(k_sf_centre, dim3(g1), dim3(FB), 0, st, image, area, acc, n);
                break;
            }
            hipLaunchKernelGGL(k_sf_dist, rows, dim3(FB), 0, st, guide, pixvar, W, H, dx, dy, d2, den);
            const float* db = d2;
            if (patch > 1) {
                hipLaunchKernelGGL(k_box_rows, rows, dim3(FB), 0, st, d2, b0, W, H, patch);
                hipLaunchKernelGGL(k_box_cols, rows, dim3(FB), 0, st, b0, b1, W, H, patch);
                db = b1;
            }
            hipLaunchKernelGGL(k_sf_exp, dim3(g3), dim3(FB), 0, st, db, den, b0, 3 * n);
            const float* wb = b0;
            if (patch > 1) {
                hipLaunchKernelGGL(k_box_rows, rows, dim3(FB), 0, st, b0, d2, W, H, patch);
                hipLaunchKernelGGL(k_box_cols, rows, dim3(FB), 0, st, d2, b1, W, H, patch);
                wb = b1;
            }
            hipLaunchKernelGGL(k_sf_accum, rows, dim3(FB), 0, st, wb, image, W, H, dx, dy, area, acc);
        }
    hipLaunchKernelGGL(k_sf_final, dim3(g1), dim3(FB), 0, st, area, acc, out, n);
    return hipGetLastError();
}

hipError_t launch_sep3(const float* in, float* tmp, float* out, int W, int H, int grid, hipStream_t st)
{
    const int64_t n = (int64_t)W * H;
    hipLaunchKernelGGL(k_sep3, dim3(blocks(n, grid)), dim3(FB), 0, st, in, tmp, W, H, 0);
    hipLaunchKernelGGL(k_sep3, dim3(blocks(n, grid)), dim3(FB), 0, st, tmp, out, W, H, 1);
    return hipGetLastError();
}

hipError_t launch_flt_importance(const float* vof, const float* fmean, float* imp, int64_t n, int grid, hipStream_t st)
{
    hipLaunchKernelGGL(k_flt_importance, dim3(blocks(n, grid)), dim3(FB), 0, st, vof, fmean, imp, n);
    return hipGetLastError();
}

hipError_t launch_scale(float* x, float k, int64_t n, int grid, hipStream_t st)
{
    hipLaunchKernelGGL(k_scale, dim3(blocks(n, grid)), dim3(FB), 0, st, x, k, n);
    return hipGetLastError();
}

hipError_t launch_flt_bnorm(const float* film, const float* imp, float* out, int64_t n, int grid, hipStream_t st)
{
    hipLaunchKernelGGL(k_flt_bnorm, dim3(blocks(n, grid)), dim3(FB), 0, st, film, imp, out, n);
    return hipGetLastError();
}

hipError_t launch_flt_rescale(const float* sumimp, const float* imp, float* newsum, float* f0, float* f1, float* m0,
                              float* m1, int64_t n, int grid, hipStream_t st)
{
    hipLaunchKernelGGL(k_flt_rescale, dim3(blocks(n, grid)), dim3(FB), 0, st, sumimp, imp, newsum, f0, f1, m0, m1, n);
    return hipGetLastError();
}

hipError_t launch_flt_merge(const unsigned long long* aw, const unsigned long long* ax, const unsigned long long* aq,
                            double uw, double ux, double uq, float* film, float* mean, float* m2, float* sw, int64_t n,
                            int grid, hipStream_t st)
{
    hipLaunchKernelGGL(k_flt_merge, dim3(blocks(n, grid)), dim3(FB), 0, st, aw, ax, aq, uw, ux, uq, film, mean, m2, sw,
                       n);
    return hipGetLastError();
}

hipError_t launch_flt_bvar(const float* m2, const float* sw, float* out, int64_t n, int grid, hipStream_t st)
{
    hipLaunchKernelGGL(k_flt_bvar, dim3(blocks(n, grid)), dim3(FB), 0, st, m2, sw, out, n);
    return hipGetLastError();
}

hipError_t launch_add(float* dst, const float* src, int64_t n, int grid, hipStream_t st)
{
    hipLaunchKernelGGL(k_add, dim3(blocks(n, grid)), dim3(FB), 0, st, dst, src, n);
    return hipGetLastError();
}
