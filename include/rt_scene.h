/*
 * rt_scene.h — flat, pointer-free-at-rest scene description shared by the
 * MI355X path tracer (librt_hip.so) and its CPU checker.
 *
 * The reference renderer builds its scene as a graph of heap objects
 * (src/sceneast.cpp:673-695: root list = [lights..., "intersectable"],
 * then BVHNode::create, src/bvh.cpp:67-94).  Crossing a C ABI needs plain
 * arrays instead, so a scene is handed over as this struct of typed arrays.
 * The caller owns every array; rt_scene_upload copies what it needs.
 *
 * Primitive IDs: the reference has no primitive ID.  Here it is the index
 * of the primitive in the flattened root list of the reference
 * (src/sceneast.cpp:676-678 then src/intersectablelist.cpp:67-78): lights
 * first, in file order, then the components of "intersectable".  A quad is
 * one primitive (two triangles that carry the same ID, src/quad.h:32).
 *
 * Colours are RGB.  The reference keeps OpenCV BGR order in cv::Vec3f
 * (src/sceneast.cpp:410-414); channel-wise arithmetic is identical, only
 * the luminance weights (src/mathhelper.cpp:102-105) are applied to the
 * matching channels.
 */
#ifndef RT_SCENE_H
#define RT_SCENE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Material kinds: src/diffusematerial.cpp, src/mirrormaterial.cpp,
 * src/phongmaterial.cpp, src/darkmatter.cpp, src/blinnmaterial.cpp,
 * src/transparentmaterial.cpp, src/texturematerial.cpp. */
enum rt_material_type {
    RT_MAT_DIFFUSE = 0,
    RT_MAT_MIRROR = 1,
    RT_MAT_PHONG = 2,
    RT_MAT_DARK = 3,
    RT_MAT_BLINN = 4,
    RT_MAT_REFRACTIVE = 5,
    RT_MAT_TEXTURE = 6   /* src/texturematerial.cpp: kd = coefficient, exponent = gamma, texture = image index */
};

/* Light kinds: src/arealight.cpp, src/pointlight.cpp, src/environmentmap.cpp,
 * src/conelight.cpp */
enum rt_light_type {
    RT_LIGHT_AREA = 0,
    RT_LIGHT_POINT = 1,
    RT_LIGHT_ENV = 2,
    RT_LIGHT_CONE = 3
};

/* Analytic objects (src/plane.cpp, src/axisalignedbox.cpp, src/sphere.cpp,
 * src/csgoperation.cpp): planes, boxes and CSG trees over spheres and
 * boxes, each given as a postfix program of nodes.  A leaf pushes its
 * interval list (the reference's getCSGIntersection), an operation pops two
 * (left below right) and pushes their combination; the program of a plane is
 * the plane alone. */
enum rt_csg_op {
    RT_CSG_SPHERE = 0,  /* p: centre, radius */
    RT_CSG_BOX = 1,     /* p: min corner, max corner */
    RT_CSG_PLANE = 2,   /* p: (a, b, c, d), points x with a x + b y + c z + d = 0; normal (a, b, c) */
    RT_CSG_ISECT = 3,   /* isect(l, r) */
    RT_CSG_UNION = 4,   /* union(l, r) */
    RT_CSG_DIFF = 5,    /* diff(l, r) */
    RT_CSG_XOR = 6      /* xor(l, r) */
};

typedef struct rt_csg_node {
    int32_t op;         /* rt_csg_op */
    int32_t material;   /* leaves: material index */
    float p[6];
} rt_csg_node;

typedef struct rt_material {
    int32_t type;        /* rt_material_type */
    float kd[3];         /* diffuse colour (RGB): Diffuse, Phong, Blinn */
    float ks[3];         /* specular colour (RGB): Phong, Blinn */
    float exponent;      /* specular exponent: Phong, Blinn; refraction coefficient: refractive; gamma: texture */
    int32_t texture;     /* texture: index into rt_scene_desc.textures */
} rt_material;

/* A texture image: height * width * 3 RGB floats (byte / 255 as the
 * reference's convertTo, src/texturematerial.cpp:46-60), row 0 at the top. */
typedef struct rt_texture {
    int32_t width, height;
    const float* rgb;
} rt_texture;

typedef struct rt_light {
    int32_t type;        /* rt_light_type */
    float p[3];          /* area: parallelogram origin; point, cone: position */
    float u[3];          /* area: first edge  (normal = u x v, |u x v| = area); cone: axis direction */
    float v[3];          /* area: second edge; cone: v[0] = opening angle in radians */
    float intensity[3];  /* RGB; env: coefficient multiplying the map */
} rt_light;

typedef struct rt_scene_desc {
    /* indexed triangle meshes (quads are given as two triangles (p1,p2,p4),
     * (p4,p2,p3) in that order, src/quad.h:32) */
    int32_t n_vertices;
    const float* positions;      /* 3 * n_vertices */
    const float* normals;        /* 3 * n_vertices, or NULL = flat normals */
    int32_t n_triangles;
    const int32_t* tri_vertices; /* 3 * n_triangles */
    const int32_t* tri_material; /* n_triangles */
    const int32_t* tri_prim_id;  /* n_triangles */

    /* spheres (cx, cy, cz, r) */
    int32_t n_spheres;
    const float* spheres;         /* 4 * n_spheres */
    const int32_t* sphere_material;
    const int32_t* sphere_prim_id;

    int32_t n_materials;
    const rt_material* materials;

    /* lights, in reference file order.  The primitive ID of light i is i. */
    int32_t n_lights;
    const rt_light* lights;

    /* environment map, row-major, RGB float, used by an RT_LIGHT_ENV light;
     * NULL when there is none.  env_kind 0: latitude-longitude, row 0 = +y
     * pole; 1: angular light probe looked up as the reference's
     * EnvironmentMap::get (src/environmentmap.cpp:57-66). */
    int32_t env_width, env_height;
    const float* env_rgb;
    int32_t env_kind;

    /* analytic objects: object i is the program csg_nodes[object_first[i] ..
     * object_first[i + 1]) (object_first has n_objects + 1 entries) with
     * primitive ID object_prim_id[i]; a CSG program has at most
     * RT_CSG_MAX_LEAVES leaves.  n_objects = 0 when there are none. */
    int32_t n_objects;
    const rt_csg_node* csg_nodes;
    const int32_t* object_first;
    const int32_t* object_prim_id;

    /* texture images of RT_MAT_TEXTURE materials, and texture coordinates of
     * the triangles (u, v at p1, p2, p3: 6 floats per triangle; NULL = the
     * reference's default (0, 0), (0, 1), (1, 0), src/triangle.h:35-36) */
    int32_t n_textures;
    const rt_texture* textures;
    const float* tri_texcoords;

    /* analytic objects under an instance transform (IntersectableInstance,
     * src/intersectableinstance.cpp:24-30): 12 floats per object, the
     * world-to-object affine map [A | b] row-major (object point =
     * A x + b), or NULL when every object is given in world space.  A ray is
     * mapped into object space before the object's program is evaluated
     * (its parameter t is unchanged by an affine map) and the object-space
     * normal n comes back as A^T n (the normal matrix of the instance).  An
     * identity row set (A = I, b = 0) leaves the object in world space.
     * Planes are always given in world space. */
    const float* object_xform;

    /* Split rule of the BVH that rt_scene_upload builds: RT_BVH_MEDIAN, the
     * reference's (BVHNode::create, src/bvh.cpp:67-94: median of the box
     * centres along the longest axis; leaves of up to 4 records here), or
     * RT_BVH_SAH, the surface-area heuristic.  Hits do not depend on it
     * (up to ties in t); the traversal work per ray does. */
    int32_t bvh_build;
} rt_scene_desc;

enum { RT_BVH_MEDIAN = 0, RT_BVH_SAH = 1 };

#define RT_CSG_MAX_LEAVES 16

/* Pinhole camera, src/camera.h:34-101: eye, look-at, up, vertical field of
 * view in radians, resolution. */
typedef struct rt_camera {
    float eye[3];
    float look_at[3];
    float up[3];
    float fov;
    int32_t width, height;
} rt_camera;

/* One view of a (possibly batched) render: camera, stratified samples per
 * pixel spp_x * spp_y (src/perpixelrenderer.cpp:78, pt-x/y-samples). */
typedef struct rt_view {
    rt_camera camera;
    int32_t spp_x, spp_y;
} rt_view;

/* Estimator modes.
 *  COMPAT:   the reference estimator, quirks included
 *            (src/renderer.cpp:47-106, src/unidipathtracingintegrator.cpp:34-107)
 *            with max_depth as a runtime parameter.
 *  UNBIASED: next-event estimation with light-selection weight, emission only
 *            on camera / specular chains, compensated Russian roulette.
 *  GUIDED:   UNBIASED with BSDF sampling replaced by a one-sample MIS mixture
 *            of the BSDF and a learned von Mises-Fisher mixture. */
enum rt_mode { RT_MODE_COMPAT = 0, RT_MODE_UNBIASED = 1, RT_MODE_GUIDED = 2 };

typedef struct rt_params {
    int32_t mode;          /* rt_mode */
    int32_t max_depth;     /* reference MAX_DEPTH (src/renderer.h:38) */
    int32_t rr_start;      /* Russian roulette start (src/renderer.cpp:52) */
    float rr_prob;         /* termination probability (--pt-termination-prob) */
    uint32_t seed;         /* --seed */
    int32_t guide_train;   /* GUIDED: 1 = train the network online after every pass */
    float guide_lr;        /* GUIDED: Adam learning rate (0 = 1e-3) */
    int32_t guide_batch;   /* GUIDED: records per training step (0 = 262144) */
} rt_params;

#ifdef __cplusplus
}
#endif

#endif /* RT_SCENE_H */
