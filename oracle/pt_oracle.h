/*
 * pt_oracle.h — CPU oracle for the path-tracing hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: tests, the smoke
 * check and the CPU-baseline leg of the benchmark call it; the product
 * (librt_hip.so and renderer_amd/) never links, loads or calls it.
 *
 * It restates the reference renderer's algorithms on the CPU, following
 * (all paths relative to the reference tree):
 *   src/ray.h:34-111, src/hitrecord.h:62-84         rays and hit records
 *   src/triangle.cpp:30-56, src/quad.cpp:23-34      triangle / quad tests
 *   src/sphere.cpp:29-56, src/csgobject.cpp:25-44   sphere test
 *   src/axisalignedbox.cpp:30-104, src/bvh.cpp:34-149, src/intersectablelist.cpp:30-78
 *   src/abstractdiffusematerial.cpp:8-15, src/diffusematerial.cpp:25-28,
 *   src/mirrormaterial.cpp:31-41, src/phongmaterial.cpp:31-81
 *   src/arealight.cpp:31-119, src/pointlight.cpp:29-60, src/environmentmap.cpp:35-108
 *   src/sampler.cpp:33-91, src/jitteredsampler.cpp:23-35, src/camera.cpp:34-58
 *   src/renderer.cpp:47-106, src/unidipathtracingintegrator.cpp:34-107,
 *   src/perpixelrenderer.cpp:43-101
 * with two deliberate changes: the GSL taus2 stream is replaced by the
 * Philox-4x32-10 stream keyed by (seed, view, pixel, sample, dimension) that
 * the GPU path uses, and MAX_DEPTH / RR start are runtime parameters.
 *
 * The environment light is a latitude-longitude map, a format of this
 * library's own (the reference reads an angular light probe through OpenCV,
 * which is restated too); its lookup and importance sampling follow the
 * textbook method and are pinned by a white furnace and by the consistency
 * of the sampling density with the sampled directions, there being no
 * reference counterpart to compare with.
 */
#ifndef PT_ORACLE_H
#define PT_ORACLE_H

#include <stdint.h>
#include "../include/rt_scene.h"
#include "../include/rt_mlt.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Philox-4x32-10 block function (Salmon, Moraes, Dror, Shaw, SC'11). */
void oc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* accel: 0 = linear IntersectableList (reference --no-bvh),
 *        1 = reference BVH (leaf = 1 primitive, both children visited). */
void* oc_scene_create(const rt_scene_desc* desc, int accel);
void oc_scene_destroy(void* scene);

/* Closest hit for n rays; ray i = (o[4i..4i+2], d[4i..4i+2]) over
 * [o[4i+3], d[4i+3]].  Miss: prim = -1, t = NaN.  Area lights are hit like
 * any primitive (their ID is their light index). */
void oc_trace_closest(void* scene, int64_t n, const float* o_tmin, const float* d_tmax,
                      int32_t* prim, float* t, int threads);

/* Any hit in range: occluded[i] = 1 if some primitive is hit. */
void oc_trace_any(void* scene, int64_t n, const float* o_tmin, const float* d_tmax,
                  uint8_t* occluded, int threads);

/* Traversal of a flattened 4-wide BVH as rt_bvh_export writes it (nodes,
 * records), in the order the device traversal visits it: closest hit
 * (any = 0: g[i] = record's primitive index or -1, t[i] = its parameter or
 * NaN) or any hit (any = 1: g[i] = 1 occluded, 0 not; t untouched).
 * visits[i] / tests[i]: interior nodes fetched and primitive records tested
 * by ray i.  Rays as for oc_trace_closest. */
void oc_flat_trace(const float* nodes, int64_t n_nodes, const float* records, int64_t n, const float* o_tmin,
                   const float* d_tmax, int any, int32_t* g, float* t, int32_t* visits, int32_t* tests, int threads);

/* Texture coordinates of each ray's closest hit (NaN for a miss), and the
 * albedo pow(image(s, t) * coefficient, gamma) of texture material `mat`
 * (returns -1 when it is not a texture). */
void oc_hit_texcoords(void* scene, int64_t n, const float* o_tmin, const float* d_tmax, float* tc);
int oc_texture_albedo(void* scene, int mat, float s, float t, float rgb[3]);

/* Surface normal (normalised, as HitRecord keeps it) of each ray's closest
 * hit, 0 for a miss. */
void oc_hit_normals(void* scene, int64_t n, const float* o_tmin, const float* d_tmax, float* normals);

/* Single-primitive test with the reference arithmetic, used to judge hit
 * mismatches at shared edges: returns t or NaN. */
float oc_intersect_prim(void* scene, int32_t prim, const float* o_tmin, const float* d_tmax);

/* Render rows [row0, row1) of one view into film (width*height*3 RGB, the
 * caller zeroes it).  Returns the number of rays traced (camera, extension
 * and shadow rays).  view_id keys the random stream like the GPU path. */
int64_t oc_render(void* scene, const rt_view* view, const rt_params* params, uint32_t view_id,
                  float* film, int row0, int row1, int threads);

/* Radiance of one camera sample (pixel (px,py), sample index s) — the
 * per-sample Integrator::integrate entry of the reference. */
/* UniDiPathTracingIntegrator::integrate(const Ray&, ...) for n rays (the
 * layout of oc_trace_closest); ray i draws the stream (seed, view 0,
 * key0 + i, sample 0); rgb[3i..3i+2] */
void oc_integrate_rays(void* scene, int64_t n, const float* o, const float* d, const rt_params* params,
                       uint32_t key0, float* rgb, int threads);
void oc_render_sample(void* scene, const rt_view* view, const rt_params* params, uint32_t view_id,
                      int px, int py, int s, float out_rgb[3]);

/* Environment importance sampling: direction and solid-angle pdf for
 * uniforms (u0, u1); and pdf of a direction. */
void oc_env_sample(void* scene, float u0, float u1, float dir[3], float* pdf);
float oc_env_pdf(void* scene, const float dir[3]);
/* Environment radiance towards a direction (EnvironmentMap::get). */
void oc_env_value(void* scene, const float dir[3], float rgb[3]);

/* Tonemapper::map per channel value (src/tonemapper.cpp:31-56): n floats to
 * n bytes. */
void oc_tonemap(const float* film, int64_t n, float gamma, uint8_t* out);

/* TransparentMaterial::outDirection for ray direction d, surface normal n,
 * refraction coefficient ior and sample coordinate u. */
void oc_transparent_dir(const float d[3], const float n[3], float ior, double u, float out[3]);

/* The unbiased estimators' BSDF of scene material `mat`: value and density
 * of wo -> wi about the viewer-side normal nf, and a sample from uniforms. */
void oc_bsdf(void* scene, int mat, const float wo[3], const float wi[3], const float nf[3], float f[3], float* pdf);
int oc_bsdf_sample(void* scene, int mat, const float wo[3], const float nf[3], float u0, float u1, float u2,
                   float wi[3]);

/* Markov-chain renderers (src/metropolisrenderer.cpp:56-200,
 * src/energyredistributionrenderer.cpp:38-131) with the chain layout and
 * Philox streams of include/rt_mlt.h.  film (height * width * 3) is added
 * to.  Returns 0, or -1 for bad parameters. */
int oc_render_mlt(void* scene, const rt_camera* camera, const rt_mlt_params* params, float* film,
                  rt_mlt_result* result, int threads);
/* The sample-driven integrator (src/unidipathtracingintegrator.cpp:48-107 on
 * the path of src/renderer.cpp:68-106): samples [n][2*max_depth+1][4]. */
void oc_integrate_samples(void* scene, const rt_camera* camera, int max_depth, int rr_start, float termination_prob,
                          const float* samples, int64_t n, float* rgb, int threads);
/* SymmetricFilter(patch, radius).filter(image, guide, pixvar)
 * (src/symmetricfilter.cpp:72-137) on h * w * 3 float images. */
void oc_sym_filter(const float* image, const float* guide, const float* pixvar, int w, int h, int patch, int radius,
                   float* out);
/* The primary-sample perturbation (src/metropolissample.cpp:62-78) of n coordinates. */
void oc_mlt_perturb(const float* s, const float* u, const float* v, int64_t n, float* out);
/* Pieces of the filtered Metropolis renderer (src/metropolisfltrenderer.cpp), on
 * h * w (* 3) float images: cv::blur with BORDER_REFLECT (odd p), the [.1 .8 .1]
 * sepFilter2D with BORDER_REFLECT_101, mathhelper var() of two images, the
 * importance map of a later pass (:231-239), and one film's running sums
 * rescaled to a new pass (:241-250, sumimp/newsum may be null) plus deposits
 * added by addSample (:54-75). */
void oc_box_blur(const float* in, int w, int h, int p, float* out);
void oc_sep3_reflect101(const float* in, int w, int h, float* out);
void oc_var2(const float* a, const float* b, int64_t n, float* mean, float* var);
void oc_flt_importance(const float* var_filtered, const float* mean_filtered, int w, int h, float* imp);
void oc_flt_accumulate(int w, int h, float* film, float* mean, float* m2, float* sw, const float* sumimp,
                       const float* newsum, int64_t ndep, const int32_t* pos, const float* wt, const float* rgb);

#ifdef __cplusplus
}
#endif

#endif
