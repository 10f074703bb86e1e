/*
 * rt_layout.h — HBM data layout shared by the host builder and the gfx950
 * kernels.  Plain structs only; included by C++ host code and HIP code.
 *
 * BVH: 4-wide tree (the median-split binary tree with every node absorbing
 * its interior children), flattened depth-first, one 128-byte node (one
 * cache line) per interior node holding the boxes of its 2-4 children as
 * structure of arrays, so one fetch answers all child tests, then the 4
 * child words.  Leaves are not nodes: a child word < 0 names a run of at
 * most 4 primitive records of one kind.
 *
 *   node[i].f0 = lo.x of children 0..3     node[i].f1 = hi.x of children 0..3
 *   node[i].f2 = lo.y                      node[i].f3 = hi.y
 *   node[i].f4 = lo.z                      node[i].f5 = hi.z
 *   node[i].f6 = child words 0..3 (ints)   node[i].f7 = 0
 *     word >= 0 : interior node index
 *     word <  0 : leaf, ~word = first record << 5 | kind << 3 | count
 *   An unused child slot has the empty box (lo = +inf, hi = -inf) and a
 *   leaf word of count 0.
 *
 * Primitive records: 48 bytes = 3 x float4, in leaf order.
 *   triangle      (p1.xyz, g) (p2-p1, 0) (p3-p1, 0)   src/triangle.cpp:32-44
 *   sphere        (c.xyz,  g) (r, 0, 0, 0) (0, 0, 0, 0) src/sphere.cpp:29-56
 *   parallelogram (o.xyz,  g) (u, 0) (v, 0)            src/arealight.cpp:73-90
 * g (stored as int bits) indexes the shading arrays: the reference primitive
 * ID, material, and vertex normals of triangles.
 */
#ifndef RT_LAYOUT_H
#define RT_LAYOUT_H

#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#define RT_HD __host__ __device__
#else
#define RT_HD
#endif

/* Angular light-probe lookup of EnvironmentMap::get
 * (src/environmentmap.cpp:57-66) for a unit direction d: the image point is
 * (d.y, -d.z) scaled to length acos(d.x) / pi, mapped to pixel
 * ((c + 1) / 2) * size - 1 and truncated; row from the second coordinate,
 * column from the first.  Returns row * W + col.  Indices the reference would
 * read outside the image (-1 on the rim) are clamped; the two directions
 * along the probe axis (where the scale is undefined) map to the centre. */
static inline RT_HD int rt_probe_texel(float dx, float dy, float dz, int W, int H)
{
    float icx = dy, icy = -dz;
    float len = (float)sqrt((double)icx * icx + (double)icy * icy);
    double ax = acos((double)fminf(1.f, fmaxf(-1.f, dx)));
    float scale = len > 0.f ? (float)(ax / ((double)len * M_PI)) : 0.f;
    icx *= scale;
    icy *= scale;
    int x = (int)(float)(((icx + 1.0) / 2.0) * W - 1);
    int y = (int)(float)(((icy + 1.0) / 2.0) * H - 1);
    int row = y < W - 1 ? y : W - 1, col = x < H - 1 ? x : H - 1;
    row = row < 0 ? 0 : row > H - 1 ? H - 1 : row;
    col = col < 0 ? 0 : col > W - 1 ? W - 1 : col;
    return row * W + col;
}
/* lat-long proxy of a probe, used only for the importance-sampling tables */
#define RT_PROBE_PROXY_W 512
#define RT_PROBE_PROXY_H 256

#define RT_KIND_TRI 0
#define RT_KIND_SPHERE 1
#define RT_KIND_PARA 2

#define RT_LEAF_MAX 4
#define RT_STACK_DEPTH 48
/* traversal stack entries per lane: in LDS, in registers, in global overflow */
#ifndef RT_STACK_LDS
#define RT_STACK_LDS 16
#endif
#ifndef RT_STACK_REG
#define RT_STACK_REG 4
#endif
#define RT_STACK_OVF (RT_STACK_DEPTH - RT_STACK_LDS - RT_STACK_REG)
#define RT_NODE_FLOATS 32  /* 128 B per 4-wide node */
/* Nodes 0 .. RT_TOP_NODES - 1 are the tree's first levels in breadth-first
 * order (the root, its children, theirs, and the first 48 nodes of the
 * fourth level); the traversal kernels keep them in LDS (8.6 KB per block),
 * since every ray starts there.  With the 16 KB stack a block takes 24.6 KB
 * of LDS, so the 6 blocks per CU that the registers allow still fit. */
#ifndef RT_TOP_NODES
#define RT_TOP_NODES 69
#endif

/* Material record (64 B).  type: rt_material_type. */
struct RtMaterial {
    int32_t type;
    float kd[3];
    float ks[3];
    float exponent;
    float rho_d, rho_s;  /* luminance of kd / ks (Phong lobe pick) */
    int32_t texture;     /* RT_MAT_TEXTURE: texture index (kd = coefficient, exponent = gamma) */
    float pad[5];
};

/* Light record (64 B).  type: rt_light_type.  n = u x v (not normalised,
 * |n| = area), as src/arealight.cpp:31-33. */
struct RtLight {
    int32_t type;
    float p[3];
    float u[3];
    float v[3];
    float n[3];
    float I[3];
    float area;
    int32_t g;  /* shading index of the parallelogram record, -1 if none */
};

#endif
