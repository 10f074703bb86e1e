/*
 * rt_abi.h — C ABI of the MI355X path-tracing library (librt_hip.so).
 *
 * The reference renderer has no device boundary; its hot path is the
 * virtual call Integrator::integrate(const Ray&, const Intersectable&,
 * std::vector<const Light*>, gsl_rng*) (src/integrator.h:32-37) made per
 * camera sample by PerPixelRenderer::render (src/perpixelrenderer.cpp:43-101),
 * which reaches BVHNode::intersect (src/bvh.cpp:135-149) for every ray.
 * This ABI replaces that path at batch granularity:
 *   - rt_scene_upload  replaces scene_builder::getScene + BVHNode::create
 *                      (src/sceneast.cpp:673-695, src/bvh.cpp:67-94)
 *   - rt_trace_closest replaces Intersectable::intersect on the scene root
 *                      (src/intersectable.h:40, src/bvh.cpp:135-149)
 *   - rt_trace_any     replaces the shadow queries inside Light::getIntensity
 *                      (src/arealight.cpp:35-51, src/pointlight.cpp:60-65)
 *   - rt_render        replaces PerPixelRenderer::render's loop over pixels and
 *                      samples calling UniDiPathTracingIntegrator::integrate
 *                      (src/perpixelrenderer.cpp:62-98,
 *                       src/unidipathtracingintegrator.cpp:34-107)
 *
 * Conventions: every function returns 0 on success and a negative rt_status
 * on failure; no C++ exception crosses the ABI.  All calls are synchronous
 * (they return after the GPU work finished).  A context is not re-entrant:
 * one thread uses it at a time.  The caller owns every array it passes.
 * "_device" variants take device pointers (HBM-resident inputs / outputs);
 * the others take host pointers and copy.
 */
#ifndef RT_ABI_H
#define RT_ABI_H

#include <stdint.h>
#include "rt_scene.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rt_ctx rt_ctx;

enum rt_status {
    RT_OK = 0,
    RT_ERR_INVALID = -1,     /* bad argument */
    RT_ERR_NO_SCENE = -2,    /* no scene uploaded */
    RT_ERR_HIP = -3,         /* a HIP runtime call failed */
    RT_ERR_BUILD = -4,       /* BVH construction failed (e.g. depth limit) */
    RT_ERR_CANCELLED = -5,   /* rt_request_stop was called during a render */
    RT_ERR_FAULT = -6,       /* injected by rt_set_fault (test hook) */
    RT_ERR_NO_GUIDE = -7,    /* guided mode without a guiding network */
    RT_ERR_PARSE = -8,       /* scene text or OBJ file rejected (rt_scene_file.h) */
    RT_ERR_UNSUPPORTED = -9, /* scene construct the GPU path does not render */
    RT_ERR_IO = -10          /* file could not be read or written */
};

/* Counters of the last call (and totals since rt_create). */
typedef struct rt_stats {
    int64_t camera_rays;      /* primary rays traced */
    int64_t extension_rays;   /* closest-hit rays after the first bounce */
    int64_t shadow_rays;      /* any-hit rays */
    double render_ms;         /* wall time of the last rt_render, GPU events */
    double extend_ms;         /* summed duration of closest-hit kernel launches */
    double shadow_ms;         /* summed duration of any-hit kernel launches */
    int64_t extend_launches;  /* number of per-bounce closest-hit launches that traced rays (timed renders) */
    int64_t bvh_nodes;        /* interior nodes of the uploaded BVH */
    int64_t bvh_prims;        /* primitive records (triangles, spheres, parallelograms) */
    int32_t bvh_depth;        /* depth of the uploaded BVH */
    int32_t paths_per_pass;   /* path slots used by the last render */
    int32_t passes;           /* passes of the last render */
    int32_t bvh_stack_need;   /* most traversal-stack entries a ray of the uploaded BVH can hold */
    int64_t node_visits;      /* interior nodes fetched by closest-hit traversal */
    int64_t prim_tests;       /* primitive records tested by closest-hit traversal */
    /* RT_MODE_GUIDED only (0 otherwise): network time and work of the last render */
    double guide_infer_ms;        /* summed duration of the per-bounce inference launches */
    double guide_train_ms;        /* summed duration of the training steps (weights, gradients, Adam) */
    int64_t guide_infer_samples;  /* path vertices the network was evaluated for */
    int64_t guide_train_samples;  /* records summed over all training steps */
    int64_t guide_train_steps;    /* Adam steps */
    double tonemap_ms;            /* duration of the last rt_film_tonemap kernel (GPU events) */
    int64_t shadow_node_visits;   /* interior nodes fetched by any-hit traversal (count_traversal) */
    int64_t shadow_prim_tests;    /* primitive records tested by any-hit traversal (count_traversal) */
    /* of node_visits and shadow_node_visits, the visits to the tree's first
     * nodes, which the traversal kernels read from their LDS copy (no
     * memory traffic; count_traversal) */
    int64_t node_visits_lds;
    int64_t shadow_node_visits_lds;
    /* the last bounces of a pass, run in one launch (rt_launch_config
     * finish_depth, finish_paths): its summed duration and, of
     * extension_rays and shadow_rays, the rays it traced (the rest were
     * traced by the per-bounce launches that extend_ms and shadow_ms time) */
    double finish_ms;         /* (extend_ms and shadow_ms include the per-bounce launches from
                                 there on, which return at once) */
    int64_t finish_rays;
    int64_t finish_shadow_rays;
    int64_t finish_bounce;    /* the bounce the last pass moved to it from, -1 if it did not */
} rt_stats;

/* Library-wide. */
const char* rt_error_string(int status);
int rt_version(void);

/* Context. */
int rt_create(int device, rt_ctx** out);
int rt_destroy(rt_ctx* ctx);
const char* rt_last_error(rt_ctx* ctx);

/* Scene: builds the flattened BVH on the host (median split on the longest
 * centroid axis, leaves of at most 4 primitives of one kind) and uploads
 * nodes, primitive records, shading data, materials, lights and the
 * environment map with its sampling tables. */
int rt_scene_upload(rt_ctx* ctx, const rt_scene_desc* desc);

/* Ray batches.  Ray i: origin o_tmin[4i..4i+2], range start o_tmin[4i+3],
 * direction d_tmax[4i..4i+2], range end d_tmax[4i+3].  Results: prim[i] =
 * primitive ID of the closest hit in [tmin, tmax] or -1, t[i] = its ray
 * parameter or NaN.  The environment is not a primitive here. */
int rt_trace_closest(rt_ctx* ctx, const float* o_tmin, const float* d_tmax, int64_t n,
                     int32_t* prim, float* t);
int rt_trace_closest_device(rt_ctx* ctx, const float* o_tmin, const float* d_tmax, int64_t n,
                            int32_t* prim, float* t);
int rt_trace_any(rt_ctx* ctx, const float* o_tmin, const float* d_tmax, int64_t n, uint8_t* occluded);
int rt_trace_any_device(rt_ctx* ctx, const float* o_tmin, const float* d_tmax, int64_t n,
                        uint8_t* occluded);

/* rt_trace_closest_device for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  It calls no hipMalloc or
 * hipFree: the traversal-stack overflow is a static array of 28 MiB per GPU
 * (held from the library's first use on that GPU until the process ends),
 * whose address rt_create looks up.  Same arguments and results as
 * rt_trace_closest_device; n above RT_NOALLOC_MAX_RAYS is RT_ERR_INVALID.
 * Calls that overlap on one GPU (from several contexts or threads) share the
 * array and race. */
#define RT_NOALLOC_MAX_RAYS ((int64_t)1 << 31)
int rt_trace_closest_device_noalloc(rt_ctx* ctx, const float* o_tmin, const float* d_tmax, int64_t n,
                                    int32_t* prim, float* t);

/* The 4-wide BVH that rt_scene_upload builds for a description, built on
 * the host (no context, no device), over its triangles, spheres and area
 * lights (analytic objects are tested outside the BVH).  Node i (root 0):
 * 32 floats = lo.x[4] hi.x[4] lo.y[4] hi.y[4] lo.z[4] hi.z[4] of its child
 * boxes, 4 child words (int bits), 4 zeros; a word >= 0 is an interior node,
 * a word < 0 a leaf, ~word = first record << 5 | kind << 3 | count (kind 0
 * triangle, 1 sphere, 2 parallelogram; an unused slot has an empty box and
 * count 0).  Record (12 floats, leaf order): (p, g bits) (e1, 0) (e2, 0) for
 * a triangle (vertex 0, edges to vertices 1 and 2), (c, g bits) (r, 0, 0, 0)
 * (0) for a sphere, (corner, g bits) (u, 0) (v, 0) for a parallelogram; g is
 * the primitive's index: triangles, then spheres, then area lights in light
 * order.  On entry *n_nodes and *n_records are the capacities of
 * the buffers in nodes and records, on return the counts; null buffers ask
 * for the counts only, and a buffer below its count is RT_ERR_INVALID.
 * stack_need (may be null): most stack entries a traversal holds at once.
 * No counterpart in the reference (its BVH is private to the CPU renderer,
 * src/bvh.cpp): it lets a CPU traversal of the same layout count what the
 * device traversal visits. */
int rt_bvh_export(const rt_scene_desc* desc, float* nodes, int64_t* n_nodes, float* records, int64_t* n_records,
                  int32_t* stack_need);

/* Traversal statistics of a ray batch (device pointers): interior nodes
 * visited and primitives tested per ray by the closest-hit traversal. */
int rt_trace_count_device(rt_ctx* ctx, const float* o_tmin, const float* d_tmax, int64_t n,
                          int32_t* nodes, int32_t* prims);
/* The same for the any-hit traversal of shadow rays (stops at the first hit,
 * children unordered). */
int rt_trace_any_count_device(rt_ctx* ctx, const float* o_tmin, const float* d_tmax, int64_t n,
                              int32_t* nodes, int32_t* prims);

/* Render.  films[v] receives view v's image, height*width*3 RGB float,
 * accumulated as film += sample / spp in sample order (the caller zeroes it
 * or passes an image to add to).  Several views of different resolution and
 * sample count are traced together in one pipeline of path slots. */
int rt_render(rt_ctx* ctx, const rt_view* views, int32_t n_views, const rt_params* params,
              float* const* films);
int rt_render_device(rt_ctx* ctx, const rt_view* views, int32_t n_views, const rt_params* params,
                     float* const* films);

/* UniDiPathTracingIntegrator::integrate(const Ray&, const Intersectable&,
 * std::vector<const Light*>, gsl_rng*) (src/integrator.h:35,
 * src/unidipathtracingintegrator.cpp:34-46) for n caller rays: rgb[3i..3i+2]
 * = the estimate of params->mode (RT_MODE_COMPAT or RT_MODE_UNBIASED; the
 * view fields and the guide fields of params are not used) along ray i,
 * given as for rt_trace_closest (origin, range start; direction, range end).
 * The reference estimator takes the ray as given (its direction length
 * enters the emission term as in src/arealight.cpp:102-106); the unbiased
 * one normalises the direction and scales the range by its length.  Ray i
 * draws the random numbers of the Philox stream (params->seed, view 0,
 * index key0 + i, sample 0), the groups a camera sample of rt_render draws
 * (group 0: only its path-length uniform), so a ray and key give the same
 * value on every call. */
int rt_integrate_rays(rt_ctx* ctx, const float* o_tmin, const float* d_tmax, int64_t n, const rt_params* params,
                      uint32_t key0, float* rgb);
int rt_integrate_rays_device(rt_ctx* ctx, const float* o_tmin, const float* d_tmax, int64_t n,
                             const rt_params* params, uint32_t key0, float* rgb);

/* Launch configuration used by rt_render; 0 leaves path_slots, block,
 * blocks_per_cu and guide_epochs unchanged, while the two flags are set as
 * given.  path_slots: paths in flight per pass; block: threads per block;
 * blocks_per_cu: persistent blocks per CU of the pipeline kernels (the
 * grid is CUs x blocks_per_cu). */
typedef struct rt_launch_config {
    int32_t path_slots;
    int32_t block;
    int32_t blocks_per_cu;
    int32_t use_graph;     /* 1 = capture one pass's kernel sequence once and replay it as a
                              hipGraph (ignored while the guide trains); the per-kernel
                              times in rt_stats are then 0 */
    int32_t count_traversal; /* 1 = closest-hit and any-hit kernels also count nodes and primitives */
    int32_t guide_epochs;  /* GUIDED + training: passes over each render pass's records (0 = 1) */
    int32_t finish_depth;  /* COMPAT and UNBIASED: a pass's remaining paths are traced, shaded
                              and shadow-tested to their ends in one launch, instead of one
                              launch per kernel and bounce, from the first bounce >=
                              finish_depth whose queue holds at most finish_paths paths (late
                              bounces hold few paths, and each launch costs about its slowest
                              ray's latency); the image is the same either way.  Default 1;
                              0 leaves it unchanged, negative = every bounce in separate
                              launches.  Not used with count_traversal or in GUIDED mode. */
    int32_t finish_paths;  /* that queue bound; default 4096 per CU, 0 leaves it unchanged */
} rt_launch_config;
int rt_set_launch_config(rt_ctx* ctx, const rt_launch_config* cfg);
int rt_get_launch_config(rt_ctx* ctx, rt_launch_config* cfg);

int rt_get_stats(rt_ctx* ctx, rt_stats* out);

/* Cooperative cancel, polled between passes (the reference's doStop,
 * src/renderer.cpp:125-129). */
int rt_request_stop(rt_ctx* ctx, int32_t stop);

/* Test hook: when set, the next ABI call that would launch GPU work
 * returns RT_ERR_FAULT instead. */
int rt_set_fault(rt_ctx* ctx, int32_t fault);

/* ---------------------------------------------------------------------------
 * Learned path guiding (no counterpart in the reference, which samples
 * directions from the BSDF only, src/renderer.cpp:80): a 3 x 64 MLP with
 * FP16 weights emitting 8 von Mises-Fisher lobes and the probability of
 * sampling them instead of the BSDF.  RT_MODE_GUIDED renders sample the
 * one-sample mixture and, with rt_params.guide_train, train the network
 * online (Adam) on the radiance their own paths carry.
 *
 * Parameters: a flat FP32 vector of rt_guide_param_count() values,
 * W1[64][80] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[48][64] b4[48].
 * Training records: 16 floats each - position (3), viewer-side normal (3),
 * outgoing direction (3), sampled direction (3), BSDF pdf of it, weight,
 * luminance of the throughput, unused.  Positions are encoded relative to
 * the scene box of the last rt_scene_upload unless rt_guide_set_bounds set
 * another box after it (an upload resets the box to the new scene's).
 * ------------------------------------------------------------------------- */
int rt_guide_param_count(void);
int rt_guide_init(rt_ctx* ctx, uint32_t seed);
int rt_guide_set_params(rt_ctx* ctx, const float* theta, int32_t n);
int rt_guide_get_params(rt_ctx* ctx, float* theta, int32_t n);
int rt_guide_set_bounds(rt_ctx* ctx, const float* lo, const float* hi);
/* raw network outputs y[n][48] for n records (host arrays); outputs 41..47,
 * which the head does not use, are returned as 0 */
int rt_guide_forward(rt_ctx* ctx, const float* records, int32_t n, float* y);
/* head outputs out[n][48] for n records, as the guided shading kernel
 * receives them (stored as FP16): lobe j at [5j .. 5j+4] = (mean direction,
 * concentration, mixture weight), [40] the probability of sampling the
 * guide; the directions are read renormalised and the weights divided by
 * their sum */
int rt_guide_head(rt_ctx* ctx, const float* records, int32_t n, float* out);
/* gradient of L = -sum_i w_i log p_mix_i / sum_i w_i over n records, and L */
int rt_guide_gradient(rt_ctx* ctx, const float* records, int32_t n, float* grad, float* loss);
/* `steps` Adam steps on n records with learning rate lr */
int rt_guide_train(rt_ctx* ctx, const float* records, int32_t n, int32_t steps, float lr);

#ifdef __cplusplus
}
#endif

#endif /* RT_ABI_H */
