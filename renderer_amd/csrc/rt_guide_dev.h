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

/* Head output record of one path: lobe j at [5j .. 5j+4] = (mu.xyz, kappa,
 * weight), [40] = probability of sampling the guide. */
__device__ __forceinline__ float guide_pdf(const float* o, vec3 w)
{
    float q = 0.f;
#pragma unroll
    for (int j = 0; j < GD_LOBES; ++j)
        q += o[5 * j + 4] * vmf_pdf(mk(o[5 * j], o[5 * j + 1], o[5 * j + 2]), o[5 * j + 3], w);
    return q;
}

__device__ __forceinline__ vec3 guide_sample(const float* o, float ul, float u0, float u1)
{
    int j = 0;
    float acc = 0.f;
#pragma unroll
    for (int k = 0; k < GD_LOBES - 1; ++k) {
        acc += o[5 * k + 4];
        if (ul >= acc) j = k + 1;
    }
    return vmf_sample(mk(o[5 * j], o[5 * j + 1], o[5 * j + 2]), o[5 * j + 3], u0, u1);
}

#endif
