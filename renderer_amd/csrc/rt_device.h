/*
 * rt_device.h — device-side math, random stream, scene view and BVH
 * traversal for gfx950.
 */
#ifndef RT_DEVICE_H
#define RT_DEVICE_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/rt_scene.h"
#include "rt_layout.h"

#define RT_INF __builtin_inff()
#define RT_EPS 0.001f  /* src/ray.h:34 */

struct vec3 {
    float x, y, z;
};
__device__ __forceinline__ vec3 mk(float x, float y, float z) { return vec3{x, y, z}; }
__device__ __forceinline__ vec3 operator+(vec3 a, vec3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ vec3 operator-(vec3 a, vec3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ vec3 operator-(vec3 a) { return mk(-a.x, -a.y, -a.z); }
__device__ __forceinline__ vec3 operator*(float s, vec3 a) { return mk(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ vec3 operator*(vec3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ vec3 mul3(vec3 a, vec3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ float dot(vec3 a, vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ vec3 cross(vec3 a, vec3 b)
{
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float len(vec3 a) { return sqrtf(dot(a, a)); }
__device__ __forceinline__ vec3 norm(vec3 a)
{
    float l2 = dot(a, a);
    return l2 > 0.f ? a * (1.0f / sqrtf(l2)) : mk(0.f, 0.f, 0.f);
}
/* src/vechelper.cpp:3-6 */
__device__ __forceinline__ vec3 reflect(vec3 v, vec3 n) { return v - (2.f * dot(v, n)) * n; }
__device__ __forceinline__ bool is_zero(vec3 a) { return a.x == 0.f && a.y == 0.f && a.z == 0.f; }
__device__ __forceinline__ vec3 ld3(const float* p) { return mk(p[0], p[1], p[2]); }
__device__ __forceinline__ vec3 xyz(float4 a) { return mk(a.x, a.y, a.z); }
/* src/mathhelper.cpp:102-105 with RGB channel order */
__device__ __forceinline__ float lum(vec3 c) { return 0.299f * c.x + 0.587f * c.y + 0.114f * c.z; }

/* ---------------------------------------------------------------------------
 * Philox-4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), counter-based:
 * counter = (pixel, sample, dimension group, stream), key = (seed, view).
 * The dimension groups are fixed per use (see the path kernels), so every
 * draw of a path is addressable without carried RNG state.
 * ------------------------------------------------------------------------- */
struct U4 {
    float u[4];
};
__device__ __forceinline__ uint4 philox4x32_10(uint4 c, uint2 k)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k.x += 0x9E3779B9u;
            k.y += 0xBB67AE85u;
        }
        uint32_t lo0 = 0xD2511F53u * c.x, hi0 = __umulhi(0xD2511F53u, c.x);
        uint32_t lo1 = 0xCD9E8D57u * c.z, hi1 = __umulhi(0xCD9E8D57u, c.z);
        c = make_uint4(hi1 ^ c.y ^ k.x, lo1, hi0 ^ c.w ^ k.y, lo0);
    }
    return c;
}
__device__ __forceinline__ U4 rng4(uint32_t seed, uint32_t view, uint32_t pixel, uint32_t sample, uint32_t group,
                                   uint32_t stream = 0)
{
    uint4 o = philox4x32_10(make_uint4(pixel, sample, group, stream), make_uint2(seed, view));
    const float s = 1.0f / 16777216.0f;
    U4 r;
    r.u[0] = (float)(o.x >> 8) * s;
    r.u[1] = (float)(o.y >> 8) * s;
    r.u[2] = (float)(o.z >> 8) * s;
    r.u[3] = (float)(o.w >> 8) * s;
    return r;
}

/* ---------------------------------------------------------------------------
 * Scene as seen by the kernels (all pointers are device pointers).
 * ------------------------------------------------------------------------- */
struct DevScene {
    const float4* nodes;    /* 8 float4 (128 B) per 4-wide interior node */
    const float4* prims;    /* 3 float4 per primitive record */
    const int32_t* prim_id; /* shading index g -> reference primitive ID */
    const int32_t* mat;     /* g -> material index, or -(light+1) for an area light */
    const float* vnormals;  /* 9 floats per triangle g, or null (flat) */
    const float4* geo;      /* g -> flat normal (tri, para) or centre+radius (sphere) */
    int32_t n_nodes;
    int32_t n_tri, n_sphere, n_para; /* g ranges: tri [0,n_tri), sphere, parallelogram */
    const RtMaterial* mats;
    const RtLight* lights;
    int32_t n_lights;
    int32_t env_light;      /* index of the env light, -1 if none */
    int32_t env_w, env_h;
    const float4* env_rgb;  /* texels (rgb * coefficient) */
    const float* env_row_cdf;  /* env_h + 1 */
    const float* env_col_cdf;  /* env_h * (env_w + 1) */
    const float* env_texel_p;  /* env_w * env_h */
    int32_t* stack_ovf;        /* traversal stack overflow: RT_STACK_OVF entries per lane of the launch */
    int32_t env_kind;          /* 0: lat-long texels; 1: angular probe (values from probe_rgb) */
    int32_t probe_w, probe_h;
    const float4* probe_rgb;   /* probe texels (rgb * coefficient) */
    /* analytic objects outside the BVH: postfix programs of 2-float4 nodes
     * (p0..p3), (p4, p5, op, leaf g); object i = nodes [obj_first[i],
     * obj_first[i + 1]); the leaves have shading indices from g_obj on */
    const float4* obj_nodes;
    const int32_t* obj_first;
    int32_t n_obj, g_obj;
    /* objects under an instance transform: per object three rows (A row,
     * b) of the world-to-object map (null when no object has one), and per
     * object leaf (shading index - g_obj) its object, or -1 for a leaf of an
     * object in world space */
    const float4* obj_xf;
    const int32_t* leaf_obj;
    /* textures: texels (rgb, 0) of all images; per image (first texel,
     * width, height, 0); texture coordinates of the triangles (6 floats per
     * triangle) or null for the default ones */
    const float4* tex_rgb;
    const int4* tex_info;
    const float* tri_tc;
};

/* ---------------------------------------------------------------------------
 * Primitive tests.  The arithmetic follows the reference formulation
 * (Cramer's rule on the un-normalised plane normal, src/triangle.cpp:30-56;
 * quadratic with both roots, src/sphere.cpp:29-56 + src/csgobject.cpp:25-44),
 * not Moller-Trumbore, so hit/miss decisions and t agree with it to rounding.
 * Accept: tmin <= t <= tmax and t < best.
 * ------------------------------------------------------------------------- */
/* Primitive arithmetic is compiled without FMA contraction so that every
 * product and sum is rounded as in the reference's written expression
 * order; t then agrees bit for bit with a CPU evaluation of the same
 * formulas.  (The pragma marks the instructions of these functions, and the
 * marks survive inlining.) */
#pragma clang fp contract(off)
__device__ __forceinline__ float dot_x(vec3 a, vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ vec3 cross_x(vec3 a, vec3 b)
{
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ bool hit_tri(vec3 o, vec3 d, float4 r0, float4 r1, float4 r2, float tmin, float tmax,
                                        float best, float& t, float& u, float& v, bool para)
{
    vec3 p1 = xyz(r0), e1 = xyz(r1), e2 = xyz(r2);
    vec3 ad = o - p1;
    vec3 n = cross_x(e1, e2);
    float w = -dot_x(n, d);
    float tn = dot_x(n, ad);
    vec3 adxde = cross_x(ad, d);
    float un = dot_x(adxde, e2);
    float vn = -dot_x(adxde, e1);
    if (w == 0.f) return false;
    /* Prefilter with the hardware reciprocal (within 1 ulp): x * rcp(w)
     * differs from x / w by at most 2 ulps and has the same sign, so these
     * widened tests accept every ray the exact tests accept; only the
     * candidates pay for the three correctly rounded divisions. */
    const float m = 1e-6f;
    float rw = __builtin_amdgcn_rcpf(w);
    float ta = tn * rw, ua = un * rw, va = vn * rw;
    float bound = fminf(tmax, best);
    bool cand = (ta >= tmin - m * fabsf(tmin)) && (ta <= bound + m * fabsf(bound)) && (ua >= 0.f) &&
                (ua <= 1.f + m) && (va >= 0.f) && (para ? (va <= 1.f + m) : (ua + va <= 1.f + 2.f * m));
    if (!cand) return false;
    t = tn / w;
    u = un / w;
    v = vn / w;
    bool ok = (t >= tmin) && (t <= tmax) && (t < best) && (u >= 0.f) && (u <= 1.f) && (v >= 0.f);
    return ok && (para ? (v <= 1.f) : (u + v <= 1.f));
}

__device__ __forceinline__ bool hit_sphere(vec3 o, vec3 d, float4 r0, float4 r1, float tmin, float tmax,
                                           float best, float& t)
{
    vec3 c = xyz(r0);
    float rad = r1.x;
    vec3 diff = o - c;
    float a = dot_x(d, d);
    float b = 2.f * dot_x(d, diff);
    float cc = dot_x(diff, diff) - rad * rad;
    float disc = b * b - 4.f * a * cc;
    if (!(disc >= 0.f)) return false;
    float root = sqrtf(disc);
    float t0 = (-b - root) / (2.f * a), t1 = (-b + root) / (2.f * a);
    if (tmin > t1) return false;
    t = t0 >= tmin ? t0 : t1;
    return t <= tmax && t < best;
}
/* Plane::intersect, src/plane.cpp:27-40: t = (P . (o, 1)) / -(d . n),
 * accepted when from <= t <= to (t = +inf for a parallel ray counts, as in
 * the reference). */
__device__ __forceinline__ bool hit_plane(vec3 o, vec3 d, float4 P, float tmin, float tmax, float& t)
{
    float dd = -(d.x * P.x + d.y * P.y + d.z * P.z);
    float u = o.x * P.x + o.y * P.y + o.z * P.z + 1.f * P.w;
    t = u / dd;
    return tmin <= t && t <= tmax;
}

/* AxisAlignedBox::getCSGIntersection, src/axisalignedbox.cpp:30-71: entry
 * and exit parameters; both carry the normal of the entry slab (code
 * axis * 2 + (normal > 0)). */
__device__ __forceinline__ bool box_params(vec3 o, vec3 d, float4 lo, float4 hi, float& t0, float& t1, int& code)
{
    const float mn[3] = {lo.x, lo.y, lo.z}, mx[3] = {lo.w, hi.x, hi.y};
    const float oo[3] = {o.x, o.y, o.z}, dv[3] = {d.x, d.y, d.z};
    float tn = -RT_INF, tf = RT_INF;
    code = -1;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        float a = 1.f / dv[i];
        float x0, x1;
        int c;
        if (a >= 0.f) {
            x0 = mn[i] - oo[i];
            x1 = mx[i] - oo[i];
            c = 2 * i;
        } else {
            x0 = mx[i] - oo[i];
            x1 = mn[i] - oo[i];
            c = 2 * i + 1;
        }
        x0 *= a;
        x1 *= a;
        if (x0 > tn) {
            tn = x0;
            code = c;
        }
        if (x1 < tf) tf = x1;
        if (tn > tf) return false;
    }
    t0 = tn;
    t1 = tf;
    return true;
}

/* Sphere::getCSGIntersection, src/sphere.cpp:29-56: both roots */
__device__ __forceinline__ bool sphere_params(vec3 o, vec3 d, float4 c, float& t0, float& t1)
{
    vec3 diff = o - xyz(c);
    float a = dot_x(d, d);
    float b = 2.f * dot_x(d, diff);
    float cc = dot_x(diff, diff) - c.w * c.w;
    float disc = b * b - 4.f * a * cc;
    if (!(disc >= 0.f)) return false;
    float root = sqrtf(disc);
    t0 = (-b - root) / (2.f * a);
    t1 = (-b + root) / (2.f * a);
    return true;
}
/* The world-to-object map of an instance, rows (A row, b): points A x + b,
 * directions A d, and normals back to world space as A^T n (the normal
 * matrix of the instance transform, src/hitrecord.h:79-84).  Written out so
 * that the oracle's evaluation in the same order gives the same bits. */
__device__ __forceinline__ vec3 xf_point(const float4* r, vec3 x)
{
    return mk(((r[0].x * x.x + r[0].y * x.y) + r[0].z * x.z) + r[0].w,
              ((r[1].x * x.x + r[1].y * x.y) + r[1].z * x.z) + r[1].w,
              ((r[2].x * x.x + r[2].y * x.y) + r[2].z * x.z) + r[2].w);
}
__device__ __forceinline__ vec3 xf_dir(const float4* r, vec3 x)
{
    return mk((r[0].x * x.x + r[0].y * x.y) + r[0].z * x.z, (r[1].x * x.x + r[1].y * x.y) + r[1].z * x.z,
              (r[2].x * x.x + r[2].y * x.y) + r[2].z * x.z);
}
__device__ __forceinline__ vec3 xf_normal(const float4* r, vec3 n)
{
    return mk((r[0].x * n.x + r[1].x * n.y) + r[2].x * n.z, (r[0].y * n.x + r[1].y * n.y) + r[2].y * n.z,
              (r[0].z * n.x + r[1].z * n.y) + r[2].z * n.z);
}
#pragma clang fp contract(fast)

/* Value of a CSG program for the leaves whose bits are set in `inside`
 * (CSGIntersection / CSGUnion / CSGDifference / CSGXor::isInside,
 * src/csgoperation.cpp:73-91). */
__device__ __forceinline__ bool csg_value(const float4* __restrict__ nodes, int first, int count, uint32_t inside)
{
    uint32_t st = 0;
    int leaf = 0;
    for (int i = 0; i < count; ++i) {
        int op = __float_as_int(nodes[2 * (first + i) + 1].z);
        if (op <= RT_CSG_PLANE) {
            st = (st << 1) | ((inside >> leaf) & 1u);
            ++leaf;
        } else {
            uint32_t r = st & 1u, l = (st >> 1) & 1u, v;
            if (op == RT_CSG_ISECT) v = l & r;
            else if (op == RT_CSG_UNION) v = l | r;
            else if (op == RT_CSG_DIFF) v = l & (r ^ 1u);
            else v = l ^ r;
            st = ((st >> 2) << 1) | v;
        }
    }
    return st & 1u;
}

/* First hit of object `obj` in [tmin, tmax] (CSGObject::intersect,
 * src/csgobject.cpp:25-44, on the interval list of CSGOperation::
 * getCSGIntersection, src/csgoperation.cpp:39-63).  The nested pairwise
 * merges of the reference (left first on equal t) give the same list as one
 * sweep over all leaf events ordered by t and then by leaf order, keeping
 * the events at which the program's value changes: every operation maps
 * (outside, outside) to outside, so a parent's view of a child is the
 * child's value.  Returns the leaf's shading index and the box normal code. */
__device__ __forceinline__ bool obj_hit(const float4* __restrict__ nd, const int32_t* __restrict__ obj_first, int obj,
                                     const float4* xf, vec3 o, vec3 d, float tmin, float tmax, float& t, int& g,
                                     float& code)
{
    const int first = obj_first[obj], count = obj_first[obj + 1] - first;
    /* an instance (src/intersectableinstance.cpp:24-30): the program in
     * object space, the ray mapped there; t is the same on both sides */
    if (xf) {
        o = xf_point(xf, o);
        d = xf_dir(xf, d);
    }
    if (count == 1 && __float_as_int(nd[2 * first + 1].z) == RT_CSG_PLANE) {
        if (!hit_plane(o, d, nd[2 * first], tmin, tmax, t)) return false;
        g = __float_as_int(nd[2 * first + 1].w);
        code = -1.f;
        return true;
    }
    float et[2 * RT_CSG_MAX_LEAVES];
    int8_t el[2 * RT_CSG_MAX_LEAVES], ec[2 * RT_CSG_MAX_LEAVES];
    int ne = 0, leaf = 0;
    for (int i = 0; i < count; ++i) {
        float4 a = nd[2 * (first + i)], b = nd[2 * (first + i) + 1];
        int op = __float_as_int(b.z);
        if (op > RT_CSG_PLANE) continue;
        float t0, t1;
        int c = -1;
        bool any = op == RT_CSG_SPHERE ? sphere_params(o, d, a, t0, t1) : box_params(o, d, a, b, t0, t1, c);
        if (any) {
            /* insertion after the events of equal t: stable in leaf order */
            for (int k = 0; k < 2; ++k) {
                float tv = k ? t1 : t0;
                int j = ne++;
                while (j > 0 && et[j - 1] > tv) {
                    et[j] = et[j - 1];
                    el[j] = el[j - 1];
                    ec[j] = ec[j - 1];
                    --j;
                }
                et[j] = tv;
                el[j] = (int8_t)leaf;
                ec[j] = (int8_t)c;
            }
        }
        ++leaf;
    }
    uint32_t inside = 0;
    bool val = false;
    for (int j = 0; j < ne; ++j) {
        inside ^= 1u << el[j];
        bool nv = csg_value(nd, first, count, inside);
        if (nv == val) continue;
        val = nv;
        if (!(et[j] >= tmin)) continue;
        if (et[j] > tmax) return false;
        t = et[j];
        /* the leaf's shading index: leaf k of the program */
        int k = 0;
        for (int i = 0; i < count; ++i) {
            float4 b = nd[2 * (first + i) + 1];
            if (__float_as_int(b.z) > RT_CSG_PLANE) continue;
            if (k++ == el[j]) {
                g = __float_as_int(b.w);
                break;
            }
        }
        code = (float)ec[j];
        return true;
    }
    return false;
}


/* Leaf word decoding (rt_layout.h). */
__device__ __forceinline__ void leaf_decode(int32_t w, uint32_t& first, uint32_t& count, uint32_t& kind)
{
    uint32_t x = ~(uint32_t)w;
    count = x & 7u;
    kind = (x >> 3) & 3u;
    first = x >> 5;
}

struct TraceOut {
    float t, u, v;
    int32_t g;
};

/*
 * Resumable traversal state of one ray: nearest child first, pruned by the
 * closest hit so far, the stack of deferred subtrees in LDS laid out
 * [depth][thread] (the 64 lanes of a wave hit 64 different banks).
 *
 * Box tests use t = (lo - o) * inv (no per-ray o * inv kept: three fewer
 * VGPRs let 8 waves per SIMD fit) with a zero
 * direction component replaced by a tiny signed value, and accept a box
 * when  max(near planes, tmin) <= min(far planes, closest) * (1 + 2^-20):
 * the relative slack absorbs the rounding of the plane distances, so a box
 * is never rejected where the reference's exact-order slab test
 * (src/axisalignedbox.cpp:30-73) would accept it.  Box tests only prune;
 * the reported hit comes from the primitive tests above.
 */
/* `best` starts one ulp above tmax, so that "t < best" alone enforces the
 * inclusive t <= tmax of the reference's range test (no float lies between
 * tmax and its successor) and the range end needs no register of its own. */
struct Trav {
    vec3 o, d, inv;
    float tmin, best, u, v;
    int32_t g, cur, sp;
    /* per axis, the byte offset (0 or 16) of the near planes within a node's
     * lo/hi pair: the hi planes when the direction component is negative */
    uint32_t nx, ny, nz;
};

__device__ __forceinline__ float safe_inv(float x)
{
    const float tiny = 1e-30f;
    return 1.0f / (fabsf(x) < tiny ? copysignf(tiny, x) : x);
}

__device__ __forceinline__ void trav_init(const DevScene& S, Trav& T, vec3 o, vec3 d, float tmin, float tmax)
{
    T.o = o;
    T.d = d;
    T.inv = mk(safe_inv(d.x), safe_inv(d.y), safe_inv(d.z));
    T.tmin = tmin;
    T.best = nextafterf(tmax, RT_INF);
    T.u = T.v = 0.f;
    T.g = -1;
    T.sp = 0;
    T.cur = S.n_nodes > 0 && tmin <= tmax ? 0 : -1;
    T.nx = T.inv.x >= 0.f ? 0u : 16u;
    T.ny = T.inv.y >= 0.f ? 0u : 16u;
    T.nz = T.inv.z >= 0.f ? 0u : 16u;
}

/* The analytic objects against a freshly initialised ray (tmax: its range
 * end): the closest object hit becomes the current hit, which the BVH
 * traversal then only replaces with a strictly closer one. */
__device__ __forceinline__ void obj_closest(const DevScene& S, Trav& T, float tmax)
{
    for (int i = 0; i < S.n_obj; ++i) {
        float t, code;
        int g;
        const float4* xf = S.obj_xf ? S.obj_xf + 3 * i : nullptr;
        if (obj_hit(S.obj_nodes, S.obj_first, i, xf, T.o, T.d, T.tmin, T.g >= 0 ? T.best : tmax, t, g, code) &&
            (T.g < 0 || t < T.best)) {
            T.best = t;
            T.g = g;
            T.u = code;
            T.v = 0.f;
        }
    }
}

/* Any object hit in [tmin, tmax] (shadow rays). */
__device__ __forceinline__ bool obj_any(const DevScene& S, vec3 o, vec3 d, float tmin, float tmax)
{
    for (int i = 0; i < S.n_obj; ++i) {
        float t, code;
        int g;
        const float4* xf = S.obj_xf ? S.obj_xf + 3 * i : nullptr;
        if (obj_hit(S.obj_nodes, S.obj_first, i, xf, o, d, tmin, tmax, t, g, code)) return true;
    }
    return false;
}

/* Monotone map of a float to uint32 (negative values reversed below the
 * positive ones), so keys sort with integer min / max. */
__device__ __forceinline__ uint32_t sortable(float f)
{
    uint32_t b = __float_as_uint(f);
    return b ^ ((uint32_t)((int32_t)b >> 31) | 0x80000000u);
}

__device__ __forceinline__ void box2(const Trav& T, float lo, float hi, float inv, float o, float& n, float& f)
{
    float a = (lo - o) * inv, b = (hi - o) * inv;
    bool pos = inv >= 0.f;
    n = pos ? a : b;
    f = pos ? b : a;
}

/* Traversal stack of one lane, RT_STACK_DEPTH entries: the first
 * RT_STACK_LDS in LDS (one column per lane, `stride` words apart, so a wave's
 * accesses hit 64 different banks), the next RT_STACK_REG in registers, the
 * rest in a global overflow array laid out [entry][lane] over the launch
 * grid.  The LDS part bounds how many waves fit on a CU (16 entries: 16 KB
 * per 256-lane block), so it is kept short; only rays that defer more than
 * RT_STACK_LDS + RT_STACK_REG subtrees at once reach the overflow. */
struct LaneStack {
    int32_t* lds;
    int stride;
    int32_t* ovf;    /* this lane's first overflow entry */
    size_t ostride;  /* lanes of the launch grid */
    int32_t r0 = 0, r1 = 0, r2 = 0, r3 = 0;  /* register entries (scalars, so they stay in VGPRs) */
    static constexpr int capacity = RT_STACK_DEPTH;
    __device__ __forceinline__ LaneStack(int32_t* column, int s, int32_t* overflow)
        : lds(column), stride(s), ovf(overflow + (size_t)blockIdx.x * blockDim.x + threadIdx.x),
          ostride((size_t)gridDim.x * blockDim.x)
    {
    }
    __device__ __forceinline__ void put(int sp, int32_t v)
    {
        if (sp < RT_STACK_LDS) {
            lds[sp * stride] = v;
            return;
        }
        int k = sp - RT_STACK_LDS;
        if (RT_STACK_REG > 0 && k < RT_STACK_REG) {
            r0 = k == 0 ? v : r0;
            r1 = k == 1 ? v : r1;
            r2 = k == 2 ? v : r2;
            r3 = k == 3 ? v : r3;
            return;
        }
        ovf[(size_t)(k - RT_STACK_REG) * ostride] = v;
    }
    __device__ __forceinline__ int32_t get(int sp) const
    {
        if (sp < RT_STACK_LDS) return lds[sp * stride];
        int k = sp - RT_STACK_LDS;
        if (RT_STACK_REG > 0 && k < RT_STACK_REG) return k == 0 ? r0 : k == 1 ? r1 : k == 2 ? r2 : r3;
        return ovf[(size_t)(k - RT_STACK_REG) * ostride];
    }
};

/* Visit 4-wide node T.cur (T.cur >= 0): test the 2-4 child boxes, test
 * the primitives of hit leaves (nearest box first), continue with the
 * nearest hit interior child and defer the others, nearest on top; pop when
 * none is hit.  Children whose box entry lies beyond the closest hit found
 * meanwhile are dropped.  Any-hit rays skip the ordering.  Returns true
 * when the traversal of this ray is complete.  The stack needs room for the
 * builder's stack_need. */
template <bool ANY, bool COUNT, class Stack>
__device__ __forceinline__ bool trav_step(const DevScene& S, Trav& T, Stack& stack, int& cn, int& cp,
                                          const char* top = nullptr, int* ct = nullptr)
{
    /* The near and far planes of each axis are fetched directly (the pair
     * lo/hi sits 16 B apart, so near = base + T.n?, far = that ^ 16): no
     * per-child selects.  Offsets are 32-bit from the node base, which
     * bounds the tree at 2^25 nodes (the upload checks it).  `top`: the
     * block's LDS copy of nodes 0 .. RT_TOP_NODES - 1 (a generic pointer, so
     * that one load instruction serves the lanes at the top of the tree from
     * LDS and the others from memory), or none. */
    const char* nb = top && T.cur < RT_TOP_NODES ? top : reinterpret_cast<const char*>(S.nodes);
    const uint32_t c = (uint32_t)T.cur << 7;
    const uint32_t ax = c + T.nx, ay = c + 32u + T.ny, az = c + 64u + T.nz;
    const float4 nx4 = *reinterpret_cast<const float4*>(nb + ax);
    const float4 fx4 = *reinterpret_cast<const float4*>(nb + (ax ^ 16u));
    const float4 ny4 = *reinterpret_cast<const float4*>(nb + ay);
    const float4 fy4 = *reinterpret_cast<const float4*>(nb + (ay ^ 16u));
    const float4 nz4 = *reinterpret_cast<const float4*>(nb + az);
    const float4 fz4 = *reinterpret_cast<const float4*>(nb + (az ^ 16u));
    const float4 cw = *reinterpret_cast<const float4*>(nb + c + 96u);
    if (COUNT) {
        ++cn;
        if (ct && nb == top) ++*ct; /* a visit read from the LDS copy */
    }
    const float slack = 9.5367431640625e-07f; /* 2^-20 */
    float bound = T.best;
    bound = bound + fabsf(bound) * slack;
    const float n_x[4] = {nx4.x, nx4.y, nx4.z, nx4.w}, f_x[4] = {fx4.x, fx4.y, fx4.z, fx4.w};
    const float n_y[4] = {ny4.x, ny4.y, ny4.z, ny4.w}, f_y[4] = {fy4.x, fy4.y, fy4.z, fy4.w};
    const float n_z[4] = {nz4.x, nz4.y, nz4.z, nz4.w}, f_z[4] = {fz4.x, fz4.y, fz4.z, fz4.w};
    float key[4];
    int32_t word[4] = {__float_as_int(cw.x), __float_as_int(cw.y), __float_as_int(cw.z), __float_as_int(cw.w)};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        float tl = fmaxf(fmaxf((n_x[k] - T.o.x) * T.inv.x, (n_y[k] - T.o.y) * T.inv.y),
                         fmaxf((n_z[k] - T.o.z) * T.inv.z, T.tmin));
        float th = fminf(fminf((f_x[k] - T.o.x) * T.inv.x, (f_y[k] - T.o.y) * T.inv.y),
                         fminf((f_z[k] - T.o.z) * T.inv.z, bound));
        key[k] = tl <= th + fabsf(th) * slack ? tl : RT_INF;
    }
    uint32_t uk[4];
    if (!ANY) {
        /* ascending box entry, misses (key = inf) last: the keys as sortable
         * integers with the slot in the two low bits, through a 5-comparator
         * network of integer min / max (no branches); then the words follow
         * their slots.  Clearing the low bits lowers a key by < 4 ulp, so
         * comparisons against a bound stay conservative. */
        uint32_t u[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) u[k] = (sortable(key[k]) & ~3u) | (uint32_t)k;
#define RT_MINMAX(i, j)                 \
    {                                   \
        const uint32_t lo_ = min(u[i], u[j]); \
        u[j] = max(u[i], u[j]);         \
        u[i] = lo_;                     \
    }
        RT_MINMAX(0, 1) RT_MINMAX(2, 3) RT_MINMAX(0, 2) RT_MINMAX(1, 3) RT_MINMAX(1, 2)
#undef RT_MINMAX
        const int32_t w0 = word[0], w1 = word[1], w2 = word[2], w3 = word[3];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const bool b0 = (u[j] & 1u) != 0, b1 = (u[j] & 2u) != 0;
            const int32_t wl = b0 ? w1 : w0, wh = b0 ? w3 : w2;
            word[j] = b1 ? wh : wl;
            uk[j] = u[j] & ~3u;
        }
    } else {
#pragma unroll
        for (int k = 0; k < 4; ++k) uk[k] = sortable(key[k]);
    }
    /* key comparisons below are on the sortable integers; a missed box has
     * key = inf, every hit box a finite entry */
    const uint32_t miss = sortable(RT_INF) & ~3u;
    /* leaves, nearest first: records two at a time, both loads in flight
     * before the first test (a single last record is loaded twice, the
     * second from cache); tests stay in record order within a leaf */
    auto test = [&](float4 r0, float4 r1, float4 r2, uint32_t kind) {
        if (COUNT) ++cp;
        float t, u = 0.f, v = 0.f;
        bool hit;
        if (kind == RT_KIND_SPHERE)
            hit = hit_sphere(T.o, T.d, r0, r1, T.tmin, RT_INF, T.best, t);
        else
            hit = hit_tri(T.o, T.d, r0, r1, r2, T.tmin, RT_INF, T.best, t, u, v, kind == RT_KIND_PARA);
        if (hit) {
            T.best = t;
            T.u = u;
            T.v = v;
            T.g = __float_as_int(r0.w);
        }
        return hit;
    };
    /* one loop body over the hit leaves (a bit mask in slot order), so the
     * record registers are allocated once rather than per slot;
     * key = inf marks a missed box (a hit box is entered at a finite t) */
    uint32_t leaves = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (word[k] < 0 && uk[k] < miss) leaves |= 1u << k;
    while (leaves) {
        const int k = __builtin_ctz(leaves);
        leaves &= leaves - 1u;
        const bool b0 = (k & 1) != 0, b1 = (k & 2) != 0;
        const uint32_t kk = b1 ? (b0 ? uk[3] : uk[2]) : (b0 ? uk[1] : uk[0]);
        const int32_t wk = b1 ? (b0 ? word[3] : word[2]) : (b0 ? word[1] : word[0]);
        if (!(kk <= sortable(T.best + fabsf(T.best) * slack))) continue;
        uint32_t first, count, kind;
        leaf_decode(wk, first, count, kind);
        for (uint32_t j = 0; j < count; j += 2) {
            const float4* pa = S.prims + 3 * (size_t)(first + j);
            const bool two = j + 1 < count;
            const float4* pb = two ? pa + 3 : pa;
            float4 a0 = pa[0], a1 = pa[1], a2 = pa[2];
            float4 b0 = pb[0], b1 = pb[1], b2 = pb[2];
            if (test(a0, a1, a2, kind) && ANY) return true;
            if (two && test(b0, b1, b2, kind) && ANY) return true;
        }
    }
    /* interior children, farthest first: each earlier choice is deferred,
     * so the nearest continues and the next nearest is on top */
    int32_t next = -1;
    const uint32_t cut = sortable(T.best + fabsf(T.best) * slack);
#pragma unroll
    for (int k = 3; k >= 0; --k) {
        if (word[k] < 0 || !(uk[k] < miss) || !(uk[k] <= cut)) continue;
        if (next >= 0) {
            stack.put(T.sp, next);
            ++T.sp;
        }
        next = word[k];
    }
    if (next >= 0) {
        T.cur = next;
        return false;
    }
    if (T.sp == 0) return true;
    --T.sp;
    T.cur = stack.get(T.sp);
    return false;
}

/* Whole traversal of one ray (batch kernels and counting). */
template <bool ANY, bool COUNT, class Stack>
__device__ __forceinline__ TraceOut traverse(const DevScene& S, vec3 o, vec3 d, float tmin, float tmax,
                                             Stack& stack, int* n_nodes = nullptr, int* n_prims = nullptr,
                                             const char* top = nullptr)
{
    Trav T;
    trav_init(S, T, o, d, tmin, tmax);
    int cn = 0, cp = 0;
    if (S.n_obj > 0 && tmin <= tmax) {
        if (ANY) {
            if (obj_any(S, o, d, tmin, tmax)) {
                T.g = 0;
                T.cur = -1;
            }
        } else {
            obj_closest(S, T, tmax);
        }
    }
    if (T.cur >= 0)
        while (!trav_step<ANY, COUNT>(S, T, stack, cn, cp, top)) {
        }
    if (COUNT) {
        *n_nodes = cn;
        *n_prims = cp;
    }
    TraceOut out;
    out.t = T.g >= 0 ? T.best : RT_INF;
    out.u = T.u;
    out.v = T.v;
    out.g = T.g;
    return out;
}

/* Shading normal of a hit (normalised), src/hitrecord.h:66-70: triangles
 * interpolate vertex normals or use the flat normal (p1-p2)x(p1-p3)
 * (src/triangle.h:35-38, precomputed in geo[g]); spheres use p - c
 * (src/sphere.cpp:46); parallelograms u x v (src/arealight.cpp:31-33). */
__device__ __forceinline__ vec3 hit_normal(const DevScene& S, int32_t g, vec3 p, float u, float v)
{
    float4 gg = S.geo[g];
    if (g < S.n_tri) {
        if (S.vnormals) {
            const float* n = S.vnormals + 9 * (size_t)g;
            vec3 nn = (1.f - u - v) * ld3(n) + u * ld3(n + 3) + v * ld3(n + 6);
            return norm(nn);
        }
        return xyz(gg);
    }
    if (g < S.n_tri + S.n_sphere) return norm(p - xyz(gg));
    if (g < S.g_obj) return xyz(gg);
    /* object leaves: sphere (centre, radius), plane (unit normal, -2), box
     * (-1: the normal of the entry slab, coded in u); a leaf of an instance
     * has them in object space */
    if (gg.w == -2.f) return xyz(gg);
    const int lo = S.leaf_obj ? S.leaf_obj[g - S.g_obj] : -1;
    const float4* xf = lo >= 0 ? S.obj_xf + 3 * lo : nullptr;
    vec3 n;
    if (gg.w > 0.f) {
        n = (xf ? xf_point(xf, p) : p) - xyz(gg);
    } else {
        const int code = (int)u, axis = code >> 1;
        const float sgn = (code & 1) ? 1.f : -1.f;
        n = mk(axis == 0 ? sgn : 0.f, axis == 1 ? sgn : 0.f, axis == 2 ? sgn : 0.f);
        if (!xf) return n;
    }
    return norm(xf ? xf_normal(xf, n) : n);
}

/* Texture coordinates of a hit: triangles interpolate their vertices'
 * (src/triangle.cpp:54), spheres map the unnormalised normal p - c
 * (src/sphere.cpp:47), everything else reports (0, 0). */
__device__ __forceinline__ void hit_texcoord(const DevScene& S, int32_t g, vec3 p, float u, float v, float& s,
                                             float& t)
{
    s = t = 0.f;
    if (g < S.n_tri) {
        float t1x = 0.f, t1y = 0.f, t2x = 0.f, t2y = 1.f, t3x = 1.f, t3y = 0.f;
        if (S.tri_tc) {
            const float* q = S.tri_tc + 6 * (size_t)g;
            t1x = q[0], t1y = q[1], t2x = q[2], t2y = q[3], t3x = q[4], t3y = q[5];
        }
        const float w = 1.f - u - v;
        s = w * t1x + u * t2x + v * t3x;
        t = w * t1y + u * t2y + v * t3y;
        return;
    }
    const float4 gg = S.geo[g];
    const bool sphere = g < S.n_tri + S.n_sphere || (g >= S.g_obj && gg.w > 0.f);
    if (!sphere) return;
    /* an instance reports the object-space coordinates of its hit */
    const int lo = g >= S.g_obj && S.leaf_obj ? S.leaf_obj[g - S.g_obj] : -1;
    const vec3 n = (lo >= 0 ? xf_point(S.obj_xf + 3 * lo, p) : p) - xyz(gg);
    s = (float)((double)acosf(n.z / gg.w) / M_PI);
    t = (float)((double)atan2f(n.x, n.y) / M_PI / 2.0 + 0.5);
}

/* TextureMaterial::get (src/texturematerial.cpp:62-72) and the gamma of
 * its brdf: the coordinate wrapped into [0, 1) (clamp), the point
 * (s (W - 1), (1 - t) (H - 1)) sampled bilinearly with the border
 * replicated (cv::getRectSubPix of a 1x1 patch), times the coefficient,
 * raised to gamma.  The diffuse albedo of the hit (the brdf is it / pi). */
__device__ __forceinline__ vec3 texture_albedo(const DevScene& S, const RtMaterial& M, int32_t g, vec3 p, float u,
                                               float v)
{
    float s, t;
    hit_texcoord(S, g, p, u, v, s, t);
    auto wrap = [](float x) {
        float f = fmodf(x, 1.f);
        return f < 0.f ? f + 1.f : f;
    };
    const int4 info = S.tex_info[M.texture];
    const int W = info.y, H = info.z;
    const float cx = wrap(s) * (float)(W - 1), cy = wrap(1.f - t) * (float)(H - 1);
    const float fx = floorf(cx), fy = floorf(cy), a = cx - fx, b = cy - fy;
    const int x0 = min(max((int)fx, 0), W - 1), y0 = min(max((int)fy, 0), H - 1);
    const int x1 = min(x0 + 1, W - 1), y1 = min(y0 + 1, H - 1);
    const float4* img = S.tex_rgb + info.x;
    const float4 p00 = img[(size_t)y0 * W + x0], p01 = img[(size_t)y0 * W + x1];
    const float4 p10 = img[(size_t)y1 * W + x0], p11 = img[(size_t)y1 * W + x1];
    const float w00 = (1.f - a) * (1.f - b), w01 = a * (1.f - b), w10 = (1.f - a) * b, w11 = a * b;
    vec3 c = mk(p00.x * w00 + p01.x * w01 + p10.x * w10 + p11.x * w11,
                p00.y * w00 + p01.y * w01 + p10.y * w10 + p11.y * w11,
                p00.z * w00 + p01.z * w01 + p10.z * w10 + p11.z * w11);
    const float gamma = M.exponent;
    return mk(powf(c.x * M.kd[0], gamma), powf(c.y * M.kd[1], gamma), powf(c.z * M.kd[2], gamma));
}

/* The material of a hit as the shading code sees it: a texture becomes a
 * diffuse material with the texture's albedo at the hit. */
__device__ __forceinline__ RtMaterial hit_material(const DevScene& S, int32_t g, vec3 p, float u, float v)
{
    RtMaterial M = S.mats[S.mat[g]];
    if (M.type == RT_MAT_TEXTURE) {
        const vec3 a = texture_albedo(S, M, g, p, u, v);
        M.type = RT_MAT_DIFFUSE;
        M.kd[0] = a.x, M.kd[1] = a.y, M.kd[2] = a.z;
    }
    return M;
}

#endif
