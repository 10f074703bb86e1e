/*
 * rt_layout.h — HBM data layout shared by the host builder and the gfx950
 * kernels.  Plain structs only; included by C++ host code and HIP code.
 *
 * BVH: binary tree, flattened depth-first, one 64-byte node per interior
 * node holding the boxes of BOTH children (one 64-B fetch answers both
 * child tests), then 4 child words.  Leaves are not nodes: a child word < 0
 * names a run of at most 4 primitive records of one kind.
 *
 *   node[i].f0 = (c0.lo.x, c0.hi.x, c0.lo.y, c0.hi.y)
 *   node[i].f1 = (c0.lo.z, c0.hi.z, c1.lo.x, c1.hi.x)
 *   node[i].f2 = (c1.lo.y, c1.hi.y, c1.lo.z, c1.hi.z)
 *   node[i].f3 = (child0, child1, meta0, meta1)  (ints)
 *     child >= 0 : interior node index
 *     child <  0 : leaf, first record = -child - 1, meta = count | kind << 8
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

#include <stdint.h>

#define RT_KIND_TRI 0
#define RT_KIND_SPHERE 1
#define RT_KIND_PARA 2

#define RT_LEAF_MAX 4
#define RT_STACK_DEPTH 32

/* Material record (64 B).  type: rt_material_type. */
struct RtMaterial {
    int32_t type;
    float kd[3];
    float ks[3];
    float exponent;
    float rho_d, rho_s;  /* luminance of kd / ks (Phong lobe pick) */
    float pad[6];
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
