/*
 * rt_kernels.h — host-side launch interface of the gfx950 kernels
 * (rt_kernels.hip) used by the C ABI layer (rt_abi.cpp).
 */
#ifndef RT_KERNELS_H
#define RT_KERNELS_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rt_layout.h"

/* Device scene view; mirrors DevScene in rt_device.h field by field. */
struct HostDevScene {
    const void* nodes;
    const void* prims;
    const int32_t* prim_id;
    const int32_t* mat;
    const float* vnormals;
    const void* geo;
    int32_t n_nodes;
    int32_t n_tri, n_sphere, n_para;
    const RtMaterial* mats;
    const RtLight* lights;
    int32_t n_lights;
    int32_t env_light;
    int32_t env_w, env_h;
    const void* env_rgb;
    const float* env_row_cdf;
    const float* env_col_cdf;
    const float* env_texel_p;
    int32_t* stack_ovf;
    int32_t env_kind;
    int32_t probe_w, probe_h;
    const void* probe_rgb;
    const void* obj_nodes;
    const int32_t* obj_first;
    int32_t n_obj, g_obj;
    const void* obj_xf;
    const int32_t* leaf_obj;
    const void* tex_rgb;
    const void* tex_info;
    const float* tri_tc;
};

/* One view of a render: camera basis computed on the host with the
 * reference's arithmetic (src/camera.cpp:34-46). */
struct DevView {
    float eye[3], u[3], v[3], w[3];
    float l, r, t, b;
    int32_t W, H, spp_x, spp_y;
    float inv_spp;
    int32_t pad;
    float* film;
};

/* A tile: samples [s0, s0+ns) of pixels [pix0, pix0+npix) of one view,
 * occupying path slots [slot0, slot0 + npix*ns) of its pass.  Slot
 * slot0 + j is sample s0 + j % ns of the (j / ns)-th pixel of the tile's
 * walk (tile_pixel): a pixel's samples are consecutive, so the camera rays
 * of a wave leave from one or a few pixels. */
struct DevTile {
    int32_t view, pix0, npix, s0, ns, slot0;
    int32_t acc0;  /* first pixel of this tile among the pass's pixels */
    int32_t pad;
};

/* A pass: tiles [tile0, tile1), `slots` path slots in total. */
struct DevPass {
    int32_t tile0, tile1, slots, pixels;  /* pixels = sum of npix over its tiles */
};

/* Device counters of the wavefront pipeline. */
/* Work counters of the persistent traversal launches: the queue is cut into
 * RT_XCDS equal segments, one per XCD (workgroups go to the XCDs round-robin,
 * so block b runs on XCD b % RT_XCDS), each with its own claim counter on a
 * 128-byte line of its own. */
#define RT_XCDS 8
#define RT_CLAIM_STRIDE 32

/* When a pass's remaining paths go to the finishing launch (k_finish): at
 * the first bounce >= min_depth whose queue holds at most `limit` paths
 * (min_depth = INT_MAX: never).  The host enqueues k_finish and the
 * per-bounce kernels for every bounce; each evaluates this rule on the same
 * device counters, so they agree on the bounce without a round trip to the
 * host, and the kernels of the bounces from there on return at once. */
struct FinishRule {
    int32_t min_depth, limit;
};
/* Defaults (rt_launch_config.finish_depth, finish_paths): from bounce 1, at
 * most 4096 paths per CU.  The finishing launch runs a path per lane at 4
 * waves per SIMD (4096 lanes per CU), and a launch of the per-bounce kernels
 * costs about its slowest ray's latency chain however few rays it holds.  A
 * 1024^2, 64 spp render of a 1M-triangle mesh (max depth 8, roulette from
 * bounce 3) then finishes from bounce 4 (0.34M paths of a 32M-path pass):
 * 83.9 ms per render, where every bounce in separate launches took 86.2 ms
 * and finishing from bounce 3 (1.45M paths) 85.5 ms; Metropolis batches of
 * 1M paths finish from bounce 1 (612 to 825 Mrays/s). */
#ifndef RT_FINISH_DEPTH
#define RT_FINISH_DEPTH 1
#endif
#ifndef RT_FINISH_PATHS_PER_CU
#define RT_FINISH_PATHS_PER_CU 4096
#endif

struct DevCounters {
    int32_t q_count[2];      /* path queue lengths (ping-pong) */
    int32_t shadow_count;
    int32_t pass;            /* current pass index */
    int32_t finished;        /* the bounce from which this pass's paths went to k_finish, -1 before */
    int32_t busy_bounces;    /* bounces traced by per-bounce launches (not skipped), summed over passes */
    int32_t ext_next[RT_XCDS * RT_CLAIM_STRIDE]; /* closest-hit and any-hit launches: the next unclaimed */
    int32_t sh_next[RT_XCDS * RT_CLAIM_STRIDE];  /* position in each segment (zero before each launch) */
    unsigned long long camera_rays, extension_rays, shadow_rays;
    unsigned long long node_visits, prim_tests;
    unsigned long long sh_node_visits, sh_prim_tests;  /* any-hit traversal (count_traversal) */
    unsigned long long node_visits_lds, sh_node_visits_lds; /* of those, visits read from the LDS copy */
    unsigned long long finish_rays, finish_shadow_rays; /* of the extension and shadow rays, those k_finish traced */
};

/* Path state in HBM (structure of arrays, capacity = path slots). */
struct PathState {
    float4* ray_o;   /* o.xyz, tmin */
    float4* ray_d;   /* d.xyz, tmax */
    float4* hit;     /* t, g (int bits), u, v */
    float4* alpha;   /* throughput rgb */
    float4* radiance;/* accumulated radiance rgb */
    int4* info;      /* pixel, sample, view, packed depth/stored/length/flags */
    float4* pend_x;  /* compat: previous stored vertex: position, material */
    float4* pend_n;  /* compat: its normal */
    float4* pend_d;  /* compat: its incoming direction (unit) */
    float4* pend_a;  /* compat: its throughput */
    float4* pend_c;  /* compat: its pending next-event contribution */
    float4* sh_o;    /* shadow rays: o, tmin */
    float4* sh_d;    /* d, tmax */
    float4* sh_c;    /* contribution rgb, path slot (int bits) */
    int32_t* queue[2];
    DevCounters* cnt;
};

struct PtParams {
    int32_t mode, max_depth, rr_start;
    float rr_prob;
    uint32_t seed;
    int32_t guide_train;  /* record training samples (kept with probability *GuideState::keep) */
    /* Compat paths driven by primary samples instead of Philox streams
     * (Markov-chain renderers): group g of the path in slot p is
     * pss[g * pss_stride + p].  Null for the path-tracing renderer. */
    const float4* pss;
    int32_t pss_stride;
};

/* Chains of the Markov-chain renderers (rt_mlt.hip), one path slot each. */
struct MltDev {
    float4* cur;        /* G x cap: primary samples of the current states (group-major) */
    float4* prop;       /* G x cap: proposals (PtParams::pss of the traced paths) */
    float4* cur_val;    /* cap: radiance of the current state, w = its luminance (Metropolis) or norm (ERPT) */
    double* energy;     /* cap (ERPT): norm of the chain's deposit value */
    unsigned long long* acc;       /* W * H * 3 deposits in fixed point */
    unsigned long long* accepted;  /* accepted proposals */
    int32_t cap, G, W, H;
    uint32_t seed;
    int32_t kind;       /* rt_mlt_kind */
    float large_prob;
    float scale;        /* Metropolis: b * pixels / mutations; ERPT: 1 / samples per pixel */
    double inv_unit;    /* fixed-point steps per unit of radiance */
    uint32_t chain0;    /* chain number of slot 0 */
    int32_t spp_x, spp_y;  /* ERPT initial samples per pixel */
    /* filtered Metropolis: importance map and summed importance per pixel;
     * sums of deposit weights and of weight * value^2 in fixed point */
    const float* imp;
    const float* newsum;
    unsigned long long* acc_w;
    unsigned long long* acc_q;
    double inv_unit_w, inv_unit_q;
    uint32_t step;      /* bootstrap draws: the pass number (0 outside the filtered renderer) */
    uint32_t film_chains;  /* filtered: chains per film; chain id / film_chains odd = the second film,
                              whose sums follow the first film's (W * H weights, W * H * 3 values) */
};

/* Fill the proposal vectors of n paths with fresh draws of `stream` for
 * items index0 .. index0 + n - 1; mode 0: as drawn, 1: group 0 (x, y) = the
 * pixel's corner in pixels (ERPT e_d), 2: ERPT initial sample. */
hipError_t launch_mlt_fill(const MltDev& m, uint32_t stream, uint32_t index0, int n, int mode, int grid,
                           hipStream_t st);
/* out[j] = (radiance of path j, its luminance (metric 0) or norm (1)). */
hipError_t launch_mlt_values(const PathState& ps, int n, float4* out, int metric, int grid, hipStream_t st);
/* Metropolis chain starts: bootstrap path drawn by luminance from cdf. */
hipError_t launch_mlt_start(const MltDev& m, const double* cdf, int64_t nb, double total, const float4* bval,
                            int n, int grid, hipStream_t st);
/* ERPT chain counts of initial samples index0 ..: floor(u + |c| / (M e_d)). */
hipError_t launch_erpt_count(const float4* val, int n, uint32_t index0, uint32_t seed, float ed, int M,
                             int32_t* nch, int grid, hipStream_t st);
/* ERPT chain starts: chain j begins at initial sample src[j]. */
hipError_t launch_erpt_start(const MltDev& m, const int32_t* src, const int32_t* nch, const float4* val, float ed,
                             int n, int grid, hipStream_t st);
hipError_t launch_mlt_propose(const MltDev& m, uint32_t step, int n, int grid, hipStream_t st);
hipError_t launch_mlt_accept(const MltDev& m, const PathState& ps, uint32_t step, int n, int grid, hipStream_t st);
/* filtered Metropolis: bootstrap weights norm x importance at the path's
 * pixel into out[j].w (groups 0 of the proposals give the pixel) */
hipError_t launch_flt_boot_weight(const MltDev& m, const float4* val, float4* out, int n, int grid, hipStream_t st);
/* image passes (rt_filter.hip) */
hipError_t launch_var2(const float* a, const float* b, float* mean, float* var, int64_t n, int grid, hipStream_t st);
/* scratch of launch_sym_filter in floats per pixel: distance and denominator
 * planes (3 + 3), two blur buffers (3 + 3), weight sums (1), accumulated
 * colours (3) */
constexpr int64_t RT_SYM_FILTER_TMP = 16;
hipError_t launch_sym_filter(const float* image, const float* guide, const float* pixvar, int W, int H, int patch,
                             int radius, float* tmp, float* out, int grid, hipStream_t st);
hipError_t launch_sep3(const float* in, float* tmp, float* out, int W, int H, int grid, hipStream_t st);
hipError_t launch_flt_importance(const float* vof, const float* fmean, float* imp, int64_t n, int grid, hipStream_t st);
hipError_t launch_scale(float* x, float k, int64_t n, int grid, hipStream_t st);
hipError_t launch_flt_bnorm(const float* film, const float* imp, float* out, int64_t n, int grid, hipStream_t st);
hipError_t launch_flt_rescale(const float* sumimp, const float* imp, float* newsum, float* f0, float* f1, float* m0,
                              float* m1, int64_t n, int grid, hipStream_t st);
hipError_t launch_flt_merge(const unsigned long long* aw, const unsigned long long* ax, const unsigned long long* aq,
                            double uw, double ux, double uq, float* film, float* mean, float* m2, float* sw, int64_t n,
                            int grid, hipStream_t st);
hipError_t launch_flt_bvar(const float* m2, const float* sw, float* out, int64_t n, int grid, hipStream_t st);
/* dst[i] += src[i] */
hipError_t launch_add(float* dst, const float* src, int64_t n, int grid, hipStream_t st);

/* film[i] += acc[i] * unit */
hipError_t launch_mlt_resolve(const unsigned long long* acc, float* film, int64_t n, double unit, int grid,
                              hipStream_t st);

struct GuideState;
hipError_t launch_shade_guided(const HostDevScene& s, const PathState& ps, const GuideState& g, PtParams p,
                               int depth, int qsel, int grid, hipStream_t st);

/* Batch traversal kernels (parity surface and benchmarks). */
hipError_t launch_trace_closest(const HostDevScene& s, const float4* o, const float4* d, int64_t n, int32_t* prim,
                                float* t, int block, int grid, hipStream_t st);
hipError_t launch_trace_any(const HostDevScene& s, const float4* o, const float4* d, int64_t n, uint8_t* occ,
                            int block, int grid, hipStream_t st);
/* Nodes fetched and records tested per ray, by the closest-hit (any = false)
 * or the any-hit traversal. */
hipError_t launch_trace_count(const HostDevScene& s, const float4* o, const float4* d, int64_t n, int32_t* nodes,
                              int32_t* prims, int block, int grid, hipStream_t st, bool any);

/* Allocation-free closest hit.  The traversal-stack overflow of launches of
 * at most RT_NOALLOC_BLOCKS blocks in a static device array (RT_STACK_OVF
 * entries per lane): noalloc_scratch_init(device), with that device current,
 * looks up its address once per device; noalloc_scratch(device) returns it
 * (nullptr before the lookup). */
#define RT_NOALLOC_BLOCKS 1024
#define RT_NOALLOC_MAX_DEVICES 64
hipError_t noalloc_scratch_init(int device);
int32_t* noalloc_scratch(int device);
int64_t noalloc_scratch_entries();

/* Wavefront pipeline kernels. */
hipError_t launch_generate(const PathState& ps, const DevView* views, const DevTile* tiles, const DevPass* passes,
                           PtParams p, int grid, hipStream_t st);
/* Camera rays of n paths from their primary samples (group 0: image point,
 * scaled by the resolution unless pixel_units, and the length uniform);
 * path slot j = sample j.  View: the camera of `view` (film unused). */
hipError_t launch_generate_pss(const PathState& ps, const DevView* view, int n, PtParams p, int pixel_units,
                               int grid, hipStream_t st);
/* Paths from caller rays: slot j = ray j, random numbers keyed by key0 + j. */
hipError_t launch_generate_rays(const PathState& ps, const float4* o, const float4* d, int n, PtParams p,
                                uint32_t key0, int grid, hipStream_t st);
/* rgb[3 j .. 3 j + 2] = radiance of path slot j, j < n */
hipError_t launch_ray_values(const PathState& ps, int n, float* rgb, int grid, hipStream_t st);
/* The per-bounce kernels of bounce `depth` (queue qsel); under FinishRule
 * `fr` they return at once from the finishing bounce on. */
hipError_t launch_extend(const HostDevScene& s, const PathState& ps, int qsel, int depth, FinishRule fr, int count,
                         int grid, hipStream_t st);
hipError_t launch_shade(const HostDevScene& s, const PathState& ps, PtParams p, int depth, int qsel, FinishRule fr,
                        int grid, hipStream_t st);
hipError_t launch_shadow(const HostDevScene& s, const PathState& ps, PtParams p, int qsel, int depth, FinishRule fr,
                         int count, int grid, hipStream_t st);
/* Compat and unbiased estimators: if bounce `depth` is the finishing bounce
 * under `fr`, every path of queue qsel traced, shaded and shadow-tested to
 * its end in one launch, a lane per path (else it returns at once); counts
 * its rays after bounce 0 as extension and shadow rays. */
hipError_t launch_finish(const HostDevScene& s, const PathState& ps, PtParams p, int depth, int qsel, FinishRule fr,
                         int grid, hipStream_t st);
/* lanes of a traversal launch of `grid` blocks (the stack overflow array
 * holds RT_STACK_OVF entries for each) */
int64_t trav_lanes(int grid);
hipError_t launch_advance(const PathState& ps, int qsel, int depth, FinishRule fr, hipStream_t st);
hipError_t launch_accumulate(const PathState& ps, const DevView* views, const DevTile* tiles,
                             const DevPass* passes, int compat, int grid, hipStream_t st);
hipError_t launch_pass_end(const PathState& ps, hipStream_t st);
hipError_t launch_reset_counters(const PathState& ps, hipStream_t st);

/* rt_film.hip: n floats of a film -> n bytes (src/tonemapper.cpp:50-56) */
int tonemap_grid(int64_t n, int grid);
hipError_t launch_tonemap(const float* film, uint8_t* out, int64_t n, float gamma, int grid, hipStream_t st);

#endif
