/*
 * rt_mlt.h — the reference's Markov-chain renderers on the GPU, over the
 * sample-driven form of its integrator.
 *
 *   rt_integrate_samples   UniDiPathTracingIntegrator::integrate(Path,
 *                          lightSamples, lightIndex) on the path that
 *                          Renderer::createPath builds from given samples
 *                          (src/unidipathtracingintegrator.cpp:48-107,
 *                          src/renderer.cpp:68-106,
 *                          MetropolisSample::cameraPathFromSample,
 *                          src/metropolissample.cpp:120-127): one path per
 *                          primary-sample vector
 *   rt_render_mlt          RT_MLT_METROPOLIS: MetropolisRenderer::render
 *                          (src/metropolisrenderer.cpp:56-200), primary
 *                          sample space Metropolis (Kelemen et al. 2002);
 *                          RT_MLT_ERPT: EnergyRedistributionRenderer::render
 *                          (src/energyredistributionrenderer.cpp:38-131),
 *                          energy redistribution path tracing (Cline,
 *                          Talbot, Egbert 2005);
 *                          RT_MLT_FILTERED: MetropolisFltRenderer::render
 *                          (src/metropolisfltrenderer.cpp:60-320): passes of
 *                          two Metropolis films driven by an importance map
 *                          from the relative variance of the previous pass,
 *                          each pass cross-filtered by the symmetric filter
 *                          (src/symmetricfilter.cpp:40-137)
 *
 * Both use the reference estimator (RT_MODE_COMPAT semantics, quirks
 * included).  The scene is the one uploaded with rt_scene_upload.
 *
 * Primary samples.  A path's random numbers are G = 2 * max_depth + 1
 * groups of four floats in [0, 1]:
 *   group 0        camera sample (x, y: the image point, x * width and
 *                  y * height in pixels), path-length uniform (z), unused (w)
 *   group 1 + 2i   direction sample of path vertex i (x, y:
 *                  MetropolisSample::cameraPathSamples[i]); z, w unused
 *   group 2 + 2k   light sample of stored vertex k (x, y: lightSample1[k]) and
 *                  the light-index uniform (z: lightIndex[k] =
 *                  min(floor(z * lights), lights - 1)); w unused
 * The path length is min(rr_start + NB(termination_prob), max_depth) drawn
 * from the group-0 z by inversion (MetropolisSample::mutatePathLength), or
 * max_depth when termination_prob is 0.
 *
 * Chains.  The reference runs one chain per OpenMP thread; here thousands of
 * chains advance in lockstep, one path slot each, through the same wavefront
 * kernels as rt_render.  Every random number is a Philox draw addressed by
 * (seed, chain, mutation, group), so a render is reproducible bit for bit,
 * and deposits are summed in 64-bit fixed point, so their order does not
 * matter.
 */
#ifndef RT_MLT_H
#define RT_MLT_H

#include <stdint.h>
#include "rt_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

enum rt_mlt_kind { RT_MLT_METROPOLIS = 0, RT_MLT_ERPT = 1, RT_MLT_FILTERED = 2 };

/* RT_MLT_FILTERED options (MetropolisFltRenderer, --metflt-*) */
enum rt_mlt_filter_flags {
    RT_FLT_NO_FILTER = 1,              /* --metflt-no-filter */
    RT_FLT_NO_CROSSFILTERING = 2,      /* --metflt-no-crossfiltering */
    RT_FLT_NO_VARIANCE_FILTER = 4,     /* --metflt-no-variance-filter */
    RT_FLT_SIMPLE_VARIANCE_FILTER = 8  /* --metflt-simple-variance-filter */
};

/* Philox-4x32-10 streams of the chains: key (seed, stream), counter
 * (index, step, group, 0), four uniforms (x >> 8) * 2^-24 per draw - the
 * generator of rt_render with the stream in place of the view index.
 *   BOOTSTRAP  bootstrap path `index`, all groups (a large step)
 *   START      chain `index`: x picks its bootstrap path
 *   CONTROL    chain `index`, mutation `step`: x large-step test, y the new
 *              path-length uniform, z the acceptance test
 *   LARGE      chain, mutation, group: the values of a large step
 *   SMALL      chain, mutation, group: (x, y) perturb the group's x and
 *              (z, w) its y
 *   ED         ERPT e_d path of pixel `index`, all groups
 *   SEED       ERPT initial sample `index` = pixel * spp + s, all groups
 *              (group 0 x, y: the jitter inside its stratum)
 *   COUNT      ERPT initial sample `index`: x rounds its chain count
 * ERPT chains are numbered in initial-sample order. */
enum rt_mlt_stream {
    RT_MLT_STREAM_BOOTSTRAP = 0x4d4c5401,
    RT_MLT_STREAM_START = 0x4d4c5402,
    RT_MLT_STREAM_CONTROL = 0x4d4c5403,
    RT_MLT_STREAM_LARGE = 0x4d4c5404,
    RT_MLT_STREAM_SMALL = 0x4d4c5405,
    RT_MLT_STREAM_ED = 0x4d4c5406,
    RT_MLT_STREAM_SEED = 0x4d4c5407,
    RT_MLT_STREAM_COUNT = 0x4d4c5408
};

typedef struct rt_mlt_params {
    int32_t kind;              /* rt_mlt_kind */
    int32_t max_depth;         /* MAX_DEPTH (src/renderer.h:38; 5 in the reference) */
    int32_t rr_start;          /* MetropolisSample::terminateRoussianRouletteIndex (3) */
    float termination_prob;    /* METROPOLIS: --met-termination-prob (0.5) */
    float large_step_prob;     /* METROPOLIS: --met-large-step-prob (0.1) */
    int32_t bootstrap;         /* METROPOLIS: --met-bootstrap (10000); at least one per chain is used */
    int32_t mutations;         /* METROPOLIS: --met-mutations, per pixel (16);
                                  ERPT: --erpt-mutations, per chain (100) */
    int32_t chains;            /* METROPOLIS: chains run side by side (0: mutations * pixels / 16,
                                  at most 2^22); ERPT: ignored (the chain count follows from the image) */
    int32_t erpt_spp_x;        /* ERPT: --erpt-x-samples (2) */
    int32_t erpt_spp_y;        /* ERPT: --erpt-y-samples (2) */
    float erpt_ed_termination; /* ERPT: termination probability of the paths that estimate e_d
                                  (UniDiPathTracingIntegrator(0.1), src/energyredistributionrenderer.cpp:118) */
    uint32_t seed;             /* --seed */
    int32_t flt_passes;        /* FILTERED: --metflt-num-passes (3) */
    int32_t flt_flags;         /* FILTERED: rt_mlt_filter_flags */
} rt_mlt_params;

typedef struct rt_mlt_result {
    double b;                  /* METROPOLIS: mean luminance of the bootstrap paths */
    double ed;                 /* ERPT: deposit energy e_d */
    int64_t chains;            /* chains run */
    int64_t mutations;         /* proposals evaluated */
    int64_t accepted;          /* proposals accepted */
    int64_t seed_paths;        /* METROPOLIS: bootstrap paths; ERPT: e_d paths + initial samples */
    int64_t rays;              /* camera + extension + shadow rays traced */
    double render_ms;          /* GPU time of the whole render */
    int64_t chain_seeds;       /* ERPT: initial samples that started at least one chain */
} rt_mlt_result;

/* Radiance of n paths given by their primary samples: samples holds
 * n * (2 * max_depth + 1) * 4 floats ([path][group][4], host memory), rgb
 * receives n * 3 floats.  Camera rays follow rt_camera. */
int rt_integrate_samples(rt_ctx* ctx, const rt_camera* camera, int32_t max_depth, int32_t rr_start,
                         float termination_prob, const float* samples, int64_t n, float* rgb);

/* Render into film (height * width * 3 RGB floats, added to what it holds:
 * the reference clears its film first, src/renderer.cpp:128).  result may
 * be NULL.  RT_ERR_INVALID: bad parameters; RT_ERR_NO_SCENE. */
int rt_render_mlt(rt_ctx* ctx, const rt_camera* camera, const rt_mlt_params* params, float* film,
                  rt_mlt_result* result);
int rt_render_mlt_device(rt_ctx* ctx, const rt_camera* camera, const rt_mlt_params* params, float* film,
                         rt_mlt_result* result);

/* SymmetricFilter::filter (src/symmetricfilter.cpp:72-137), the filter of
 * RT_MLT_FILTERED: image filtered with weights from patch distances of
 * guide less the per-pixel variance pixvar, over a (2 radius + 1)^2 window
 * (patch odd).  All arrays height * width * 3 floats in host memory. */
int rt_symmetric_filter(rt_ctx* ctx, const float* image, const float* guide, const float* pixvar, int32_t width,
                        int32_t height, int32_t patch, int32_t radius, float* out);

#ifdef __cplusplus
}
#endif

#endif /* RT_MLT_H */
