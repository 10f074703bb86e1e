/*
 * rt_guide_dev.h — device helpers of the guiding distribution: von
 * Mises-Fisher lobes (density and sampling) and the one-sample mixture
 * with the BSDF.
 */
#ifndef RT_GUIDE_DEV_H
#define RT_GUIDE_DEV_H

#include "rt_device.h"
#include "rt_guide.h"

/* vMF density in its numerically stable form
 * kappa / (2 pi (1 - exp(-2 kappa))) * exp(kappa (mu.w - 1)). */
__device__ __forceinline__ float vmf_pdf(vec3 mu, float kappa, vec3 w)
{
    float c = kappa / (6.28318530718f * (-expm1f(-2.f * kappa)));
    return c * expf(kappa * (dot(mu, w) - 1.f));
}

/* Sample a vMF lobe (W. Jakob, "Numerically stable sampling of the von
 * Mises-Fisher distribution on S^2", 2012): cos = 1 + log(u + (1-u) e^{-2k}) / k. */
__device__ __forceinline__ vec3 vmf_sample(vec3 mu, float kappa, float u0, float u1)
{
    float w = 1.f + logf(u0 + (1.f - u0) * expf(-2.f * kappa)) / kappa;
    w = fminf(1.f, fmaxf(-1.f, w));
    float r = sqrtf(fmaxf(0.f, 1.f - w * w));
    float sp, cp;
    sincosf(6.28318530718f * u1, &sp, &cp);
    vec3 v;
    if (fabsf(mu.x) < fabsf(mu.y) && fabsf(mu.x) < fabsf(mu.z))
        v = mk(0.f, mu.z, -mu.y);
    else if (fabsf(mu.y) < fabsf(mu.z))
        v = mk(mu.z, 0.f, -mu.x);
    else
        v = mk(mu.y, -mu.x, 0.f);
    v = norm(v);
    vec3 t = cross(v, mu);
    return norm(t * (r * cp) + v * (r * sp) + mu * w);
}

/* Head output of one path vertex: 48 halves, lobe j at gd_lobe_off(j) ..
 * +4 = (mu.xyz, kappa, weight), [GD_ALPHA_OFF] the probability of sampling
 * the guide.  The
 * lobes are read as the distribution the rounded halves define: mu
 * renormalised and the weights divided by their sum, the same values in
 * guide_sample and guide_pdf, so that the density is exactly that of the
 * samples. */
__device__ __forceinline__ void guide_lobe(const _Float16* o, int j, vec3& mu, float& kappa, float& w)
{
    const _Float16* l = o + gd_lobe_off(j);
    mu = norm(mk((float)l[0], (float)l[1], (float)l[2]));
    kappa = (float)l[3];
    w = (float)l[4];
}

__device__ __forceinline__ float guide_pdf(const _Float16* o, vec3 w)
{
    float q = 0.f, ws = 0.f;
#pragma unroll
    for (int j = 0; j < GD_LOBES; ++j) {
        vec3 mu;
        float k, wt;
        guide_lobe(o, j, mu, k, wt);
        q += wt * vmf_pdf(mu, k, w);
        ws += wt;
    }
    return q / ws;
}

__device__ __forceinline__ vec3 guide_sample(const _Float16* o, float ul, float u0, float u1)
{
    float ws = 0.f;
#pragma unroll
    for (int j = 0; j < GD_LOBES; ++j) ws += (float)o[gd_lobe_off(j) + 4];
    const float t = ul * ws;
    int j = 0;
    float acc = 0.f;
#pragma unroll
    for (int k = 0; k < GD_LOBES - 1; ++k) {
        acc += (float)o[gd_lobe_off(k) + 4];
        if (t >= acc) j = k + 1;
    }
    vec3 mu;
    float kappa, wt;
    guide_lobe(o, j, mu, kappa, wt);
    return vmf_sample(mu, kappa, u0, u1);
}

#endif
