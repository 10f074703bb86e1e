/*
 * rt_abi.cpp — the C ABI of librt_hip.so (declared in include/rt_abi.h):
 * context and device memory, scene upload (host BVH build + flattening),
 * ray-batch entry points, and the wavefront render driver.
 */
#include <climits>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/rt_abi.h"
#include "../../include/rt_film.h"
#include "../../include/rt_mlt.h"
#include "bvh_build.h"
#include "rt_guide.h"
#include "rt_kernels.h"

#include <random>

namespace {

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

}  // namespace

struct rt_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    int fault = 0;
    volatile int stop = 0;
    int n_cu = 256;

    // scene
    bool has_scene = false;
    HostDevScene scene{};
    DevBuf nodes, prims, prim_id, mat, vnormals, geo, mats, lights, env_rgb, env_row, env_col, env_p, probe_rgb;
    DevBuf obj_nodes, obj_first, obj_xf, leaf_obj, tex_rgb, tex_info, tri_tc;
    int bvh_depth = 0;
    int bvh_stack = 0;  /* most stack entries a traversal can need */

    // ray batches (host entry points)
    DevBuf rb_o, rb_d, rb_out0, rb_out1;
    DevBuf stack_ovf;  /* traversal stack overflow of the largest launch so far */

    // render
    PathState ps{};
    size_t ps_cap = 0;
    std::vector<DevBuf> ps_bufs;
    DevBuf counters, views, tiles, passes, films;
    // Markov-chain renderers: chain states, proposals, seed-path values, deposits
    DevBuf mlt_cur, mlt_prop, mlt_val, mlt_energy, mlt_acc, mlt_cnt, mlt_seed, mlt_cdf, mlt_nch, mlt_src, mlt_view;
    DevBuf flt_img, flt_acc;  /* filtered Metropolis: per-pixel images, fixed-point pass sums */
    /* defaults from a sweep on MI355X (1M-triangle mesh, 1024^2 x 64 spp):
     * 32M path slots per pass (see render_impl), 8 persistent blocks per CU
     * (all resident: the traversal kernels claim work dynamically, so extra
     * blocks only add launch overhead) */
    rt_launch_config cfg{0, 256, 8, 0, 0, 1, RT_FINISH_DEPTH, 256 * RT_FINISH_PATHS_PER_CU};

    // guiding network
    GuideState guide{};
    bool guide_ready = false;
    bool guide_bounds_set = false;
    int guide_step = 0;
    size_t guide_slots = 0;
    int64_t guide_keep_slots = 0; /* pass size (path slots) the device keep probability was set for */
    DevBuf g_theta, g_w16, g_m, g_v, g_grad, g_part, g_acc, g_count, g_rec, g_aux, g_open, g_out, g_in, g_tmp_rec, g_tmp_y;
    float scene_lo[3] = {0, 0, 0}, scene_hi[3] = {1, 1, 1};
    int timing = 1;
    hipEvent_t ev_a = nullptr, ev_b = nullptr;
    std::vector<hipEvent_t> ev_pool;
    std::vector<hipEvent_t> ev_flight;  /* ends of the units of work in flight (InFlight) */
    rt_stats stats{};
};

namespace {

int set_err(rt_ctx* c, int code, const std::string& msg)
{
    if (c) c->err = msg;
    return code;
}

#define HIP_TRY(ctx, call)                                                                         \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return set_err(ctx, RT_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));    \
    } while (0)

int ensure(rt_ctx* c, DevBuf& b, size_t bytes)
{
    if (b.bytes >= bytes && b.p) return 0;
    if (b.p) {
        hipFree(b.p);
        b.p = nullptr;
        b.bytes = 0;
    }
    if (bytes == 0) return 0;
    hipError_t e = hipMalloc(&b.p, bytes);
    if (e != hipSuccess) return set_err(c, RT_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
    b.bytes = bytes;
    return 0;
}

void release(DevBuf& b)
{
    if (b.p) hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
}

int upload(rt_ctx* c, DevBuf& b, const void* src, size_t bytes)
{
    if (bytes == 0) {
        release(b);
        return 0;
    }
    int r = ensure(c, b, bytes);
    if (r) return r;
    HIP_TRY(c, hipMemcpy(b.p, src, bytes, hipMemcpyHostToDevice));
    return 0;
}

int check_fault(rt_ctx* c)
{
    if (c->fault) {
        c->fault = 0;
        return set_err(c, RT_ERR_FAULT, "injected fault");
    }
    return 0;
}

// Qt 5 QVector3D::normalized(): length in double, components rounded to float.
void qt_normalize(const float in[3], float out[3])
{
    double l = (double)in[0] * in[0] + (double)in[1] * in[1] + (double)in[2] * in[2];
    if (l == 0.0) {
        out[0] = out[1] = out[2] = 0.f;
        return;
    }
    double s = std::sqrt(l);
    for (int k = 0; k < 3; ++k) out[k] = (float)(in[k] / s);
}
void cross3(const float a[3], const float b[3], float o[3])
{
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}
float lum_rgb(const float c[3]) { return (float)(.299 * c[0] + .587 * c[1] + .114 * c[2]); }

int grid_for(rt_ctx* c) { return c->n_cu * std::max(1, c->cfg.blocks_per_cu); }

/* When the compat and unbiased estimators move a pass's remaining paths to
 * the finishing launch (cfg.finish_depth <= 0: never). */
FinishRule finish_rule(const rt_ctx* c)
{
    if (c->cfg.finish_depth <= 0) return FinishRule{INT_MAX, 0};
    return FinishRule{c->cfg.finish_depth, c->cfg.finish_paths};
}

/* Bounds the units of work (render passes, Markov-chain mutations) queued on
 * the context's stream ahead of the GPU: after each unit an event is
 * recorded, and before recording the unit `depth` places later the host
 * waits for it.  The GPU always has the next units queued, and a stop
 * request (rt_request_stop from another thread) is seen within about
 * `depth` units instead of after everything was queued. */
struct InFlight {
    rt_ctx* c;
    int depth;
    size_t n = 0;
    InFlight(rt_ctx* ctx, int d) : c(ctx), depth(d) {}
    hipError_t unit_done()
    {
        while ((int)c->ev_flight.size() < depth) {
            hipEvent_t e;
            hipError_t r = hipEventCreateWithFlags(&e, hipEventDisableTiming);
            if (r != hipSuccess) return r;
            c->ev_flight.push_back(e);
        }
        hipEvent_t slot = c->ev_flight[n % (size_t)depth];
        if (n >= (size_t)depth) {
            hipError_t r = hipEventSynchronize(slot);  /* the unit `depth` before this one */
            if (r != hipSuccess) return r;
        }
        ++n;
        return hipEventRecord(slot, c->stream);
    }
};

/* A stop request seen between units: wait for the queued work, so that no
 * kernel writes into the caller's memory after the call returns. */
int cancel(rt_ctx* c)
{
    hipStreamSynchronize(c->stream);
    return set_err(c, RT_ERR_CANCELLED, "render cancelled");
}

/* The overflow part of the traversal stacks for a launch of `grid` blocks:
 * RT_STACK_OVF int32 entries per lane.  Grows, never shrinks. */
int ensure_stack_overflow(rt_ctx* c, int grid)
{
    size_t bytes = (size_t)trav_lanes(grid) * RT_STACK_OVF * sizeof(int32_t);
    if (int r = ensure(c, c->stack_ovf, std::max<size_t>(bytes, 4))) return r;
    c->scene.stack_ovf = (int32_t*)c->stack_ovf.p;
    return 0;
}

}  // namespace

extern "C" {

const char* rt_error_string(int s)
{
    switch (s) {
    case RT_OK: return "ok";
    case RT_ERR_INVALID: return "invalid argument";
    case RT_ERR_NO_SCENE: return "no scene uploaded";
    case RT_ERR_HIP: return "HIP runtime error";
    case RT_ERR_BUILD: return "BVH build failed";
    case RT_ERR_CANCELLED: return "cancelled";
    case RT_ERR_FAULT: return "injected fault";
    case RT_ERR_NO_GUIDE: return "no guiding network";
    case RT_ERR_PARSE: return "scene parse error";
    case RT_ERR_UNSUPPORTED: return "unsupported scene construct";
    case RT_ERR_IO: return "file error";
    default: return "unknown error";
    }
}

int rt_version(void) { return 1; }

int rt_create(int device, rt_ctx** out)
{
    if (!out) return RT_ERR_INVALID;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return RT_ERR_HIP;
    rt_ctx* c = new (std::nothrow) rt_ctx();
    if (!c) return RT_ERR_INVALID;
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return RT_ERR_HIP;
    }
    int cu = 0;
    if (hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cu > 0) c->n_cu = cu;
    c->cfg.finish_paths = c->n_cu * RT_FINISH_PATHS_PER_CU;
    /* the allocation-free trace's static scratch: its address on this device */
    if (noalloc_scratch_init(device) != hipSuccess) {
        hipStreamDestroy(c->stream);
        delete c;
        return RT_ERR_HIP;
    }
    hipEventCreate(&c->ev_a);
    hipEventCreate(&c->ev_b);
    *out = c;
    return RT_OK;
}

int rt_destroy(rt_ctx* c)
{
    if (!c) return RT_ERR_INVALID;
    hipSetDevice(c->device);
    hipStreamSynchronize(c->stream);
    DevBuf* all[] = {&c->nodes, &c->prims, &c->prim_id, &c->mat, &c->vnormals, &c->geo, &c->mats, &c->lights,
                     &c->env_rgb, &c->env_row, &c->env_col, &c->env_p, &c->probe_rgb, &c->rb_o, &c->rb_d, &c->rb_out0,
                     &c->rb_out1, &c->stack_ovf, &c->counters, &c->views, &c->tiles, &c->passes, &c->films, &c->g_theta,
                     &c->g_w16, &c->g_m, &c->g_v, &c->g_grad, &c->g_part, &c->g_acc, &c->g_count, &c->g_rec, &c->g_aux,
                     &c->g_open, &c->g_out, &c->g_in, &c->g_tmp_rec, &c->g_tmp_y, &c->obj_nodes, &c->obj_first, &c->obj_xf, &c->leaf_obj,
                     &c->tex_rgb, &c->tex_info, &c->tri_tc,
                     &c->mlt_cur, &c->mlt_prop, &c->mlt_val, &c->mlt_energy, &c->mlt_acc, &c->mlt_cnt,
                     &c->mlt_seed, &c->mlt_cdf, &c->mlt_nch, &c->mlt_src, &c->mlt_view, &c->flt_img, &c->flt_acc};
    for (DevBuf* b : all) release(*b);
    for (auto& b : c->ps_bufs) release(b);
    for (auto e : c->ev_pool) hipEventDestroy(e);
    for (auto e : c->ev_flight) hipEventDestroy(e);
    if (c->ev_a) hipEventDestroy(c->ev_a);
    if (c->ev_b) hipEventDestroy(c->ev_b);
    hipStreamDestroy(c->stream);
    delete c;
    return RT_OK;
}

const char* rt_last_error(rt_ctx* c) { return c ? c->err.c_str() : "null context"; }

int rt_set_fault(rt_ctx* c, int32_t f)
{
    if (!c) return RT_ERR_INVALID;
    c->fault = f;
    return RT_OK;
}

int rt_request_stop(rt_ctx* c, int32_t s)
{
    if (!c) return RT_ERR_INVALID;
    c->stop = s;
    return RT_OK;
}

int rt_set_launch_config(rt_ctx* c, const rt_launch_config* cfg)
{
    if (!c || !cfg) return RT_ERR_INVALID;
    if (cfg->path_slots < 0 || cfg->block < 0 || cfg->blocks_per_cu < 0) return set_err(c, RT_ERR_INVALID, "negative launch config");
    if (cfg->path_slots) c->cfg.path_slots = cfg->path_slots;
    if (cfg->block) c->cfg.block = cfg->block;
    if (cfg->blocks_per_cu) c->cfg.blocks_per_cu = std::min(cfg->blocks_per_cu, 64);
    c->cfg.use_graph = cfg->use_graph;
    c->cfg.count_traversal = cfg->count_traversal;
    if (cfg->guide_epochs > 0) c->cfg.guide_epochs = std::min(cfg->guide_epochs, 64);
    if (cfg->finish_depth) c->cfg.finish_depth = cfg->finish_depth;
    if (cfg->finish_paths > 0) c->cfg.finish_paths = cfg->finish_paths;
    return RT_OK;
}

int rt_get_launch_config(rt_ctx* c, rt_launch_config* cfg)
{
    if (!c || !cfg) return RT_ERR_INVALID;
    *cfg = c->cfg;
    return RT_OK;
}

int rt_get_stats(rt_ctx* c, rt_stats* out)
{
    if (!c || !out) return RT_ERR_INVALID;
    *out = c->stats;
    return RT_OK;
}

/* Checks the primitive, material and light arrays of a scene description;
 * *env_out receives the index of the environment light or -1. */
static int check_prims(rt_ctx* c, const rt_scene_desc* d, int* env_out)
{
    if (d->n_vertices < 0 || d->n_triangles < 0 || d->n_spheres < 0 || d->n_materials < 0 || d->n_lights < 0)
        return set_err(c, RT_ERR_INVALID, "negative count");
    if (d->bvh_build != RT_BVH_MEDIAN && d->bvh_build != RT_BVH_SAH) return set_err(c, RT_ERR_INVALID, "unknown bvh_build");
    if ((d->n_triangles && (!d->tri_vertices || !d->tri_material || !d->tri_prim_id || !d->positions)) ||
        (d->n_spheres && (!d->spheres || !d->sphere_material || !d->sphere_prim_id)) ||
        (d->n_materials && !d->materials) || (d->n_lights && !d->lights))
        return set_err(c, RT_ERR_INVALID, "missing array");
    const int nT = d->n_triangles, nS = d->n_spheres;
    int env = -1;
    for (int i = 0; i < d->n_lights; ++i) {
        if (d->lights[i].type == RT_LIGHT_ENV) {
            if (env >= 0 || !d->env_rgb || d->env_width <= 0 || d->env_height <= 0)
                return set_err(c, RT_ERR_INVALID, "environment light needs exactly one map");
            env = i;
        } else if (d->lights[i].type != RT_LIGHT_AREA && d->lights[i].type != RT_LIGHT_POINT &&
                   d->lights[i].type != RT_LIGHT_CONE)
            return set_err(c, RT_ERR_INVALID, "unknown light type");
    }
    for (int i = 0; i < nT; ++i) {
        for (int k = 0; k < 3; ++k)
            if (d->tri_vertices[3 * i + k] < 0 || d->tri_vertices[3 * i + k] >= d->n_vertices)
                return set_err(c, RT_ERR_INVALID, "triangle vertex index out of range");
        if (d->tri_material[i] < 0 || d->tri_material[i] >= d->n_materials)
            return set_err(c, RT_ERR_INVALID, "triangle material out of range");
    }
    for (int i = 0; i < nS; ++i)
        if (d->sphere_material[i] < 0 || d->sphere_material[i] >= d->n_materials)
            return set_err(c, RT_ERR_INVALID, "sphere material out of range");
    *env_out = env;
    return RT_OK;
}

/* The 4-wide BVH over a checked description's triangles, spheres and area
 * lights (primitive index g: triangles, then spheres, then area lights in
 * light order), and its primitive records in leaf order: 12 floats each,
 * (p, g bits) (e1 or radius, 0) (e2, 0) for triangles (p = vertex 0,
 * e = edges from it), spheres (centre, radius) and parallelograms (corner,
 * u, v).  lo/hi: the union of the primitive boxes (+inf/-inf when empty). */
static int build_bvh(rt_ctx* c, const rt_scene_desc* d, rtb::BuildResult& br, std::vector<float>& rec, float lo[3],
                     float hi[3])
{
    const int nT = d->n_triangles, nS = d->n_spheres;
    std::vector<rtb::PrimRef> refs;
    std::vector<int> light_of;  /* light index of area light j */
    refs.reserve(nT + nS);
    auto P = [d](int v, int k) { return d->positions[3 * v + k]; };
    for (int i = 0; i < nT; ++i) {
        const int32_t* tv = d->tri_vertices + 3 * i;
        rtb::PrimRef r;
        for (int k = 0; k < 3; ++k) {
            r.lo[k] = std::min(P(tv[0], k), std::min(P(tv[1], k), P(tv[2], k)));
            r.hi[k] = std::max(P(tv[0], k), std::max(P(tv[1], k), P(tv[2], k)));
        }
        r.kind = RT_KIND_TRI;
        r.g = i;
        refs.push_back(r);
    }
    for (int i = 0; i < nS; ++i) {
        const float* s = d->spheres + 4 * i;
        rtb::PrimRef r;
        for (int k = 0; k < 3; ++k) {
            r.lo[k] = s[k] - s[3];
            r.hi[k] = s[k] + s[3];
        }
        r.kind = RT_KIND_SPHERE;
        r.g = nT + i;
        refs.push_back(r);
    }
    for (int i = 0; i < d->n_lights; ++i) {
        const rt_light& l = d->lights[i];
        if (l.type != RT_LIGHT_AREA) continue;
        rtb::PrimRef r;
        for (int k = 0; k < 3; ++k) {
            float c0 = l.p[k], c1 = l.p[k] + l.u[k], c2 = l.p[k] + l.v[k], c3 = l.p[k] + l.u[k] + l.v[k];
            r.lo[k] = std::min(std::min(c0, c1), std::min(c2, c3));
            r.hi[k] = std::max(std::max(c0, c1), std::max(c2, c3));
        }
        r.kind = RT_KIND_PARA;
        r.g = nT + nS + (int)light_of.size();
        light_of.push_back(i);
        refs.push_back(r);
    }
    const int G = (int)refs.size();
    for (int k = 0; k < 3; ++k) {
        lo[k] = INFINITY;
        hi[k] = -INFINITY;
    }
    for (const auto& r : refs)
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], r.lo[k]);
            hi[k] = std::max(hi[k], r.hi[k]);
        }
    int br_status = d->bvh_build == RT_BVH_SAH ? rtb::build_best(refs, br) : rtb::build(refs, br);
    if (br_status != 0)
        return set_err(c, RT_ERR_BUILD, br_status == -2 ? "BVH deeper than the traversal stack" : "too many primitives");
    /* node fetches use 32-bit byte offsets from the node base */
    if (br.nodes.size() * sizeof(float) > 0xffffffffull)
        return set_err(c, RT_ERR_BUILD, "BVH larger than 4 GiB of nodes");
    rec.assign((size_t)12 * G, 0.f);
    for (size_t i = 0; i < refs.size(); ++i) {
        int g = refs[i].g;
        float* o = &rec[12 * i];
        std::memcpy(&o[3], &g, 4);
        if (g < nT) {
            const int32_t* tv = d->tri_vertices + 3 * g;
            for (int k = 0; k < 3; ++k) {
                o[k] = P(tv[0], k);
                o[4 + k] = P(tv[1], k) - P(tv[0], k);
                o[8 + k] = P(tv[2], k) - P(tv[0], k);
            }
        } else if (g < nT + nS) {
            const float* s = d->spheres + 4 * (g - nT);
            for (int k = 0; k < 3; ++k) o[k] = s[k];
            o[4] = s[3];
        } else {
            const rt_light& L = d->lights[light_of[g - nT - nS]];
            for (int k = 0; k < 3; ++k) {
                o[k] = L.p[k];
                o[4 + k] = L.u[k];
                o[8 + k] = L.v[k];
            }
        }
    }
    return RT_OK;
}

int rt_bvh_export(const rt_scene_desc* d, float* nodes, int64_t* n_nodes, float* records, int64_t* n_records,
                  int32_t* stack_need)
{
    if (!d || !n_nodes || !n_records) return RT_ERR_INVALID;
    int env = -1;
    if (int r = check_prims(nullptr, d, &env)) return r;
    rtb::BuildResult br;
    std::vector<float> rec;
    float lo[3], hi[3];
    if (int r = build_bvh(nullptr, d, br, rec, lo, hi)) return r;
    const int64_t nn = (int64_t)(br.nodes.size() / RT_NODE_FLOATS), nr = (int64_t)(rec.size() / 12);
    const int64_t cap_n = *n_nodes, cap_r = *n_records;
    *n_nodes = nn;
    *n_records = nr;
    if (stack_need) *stack_need = br.stack_need;
    if ((nodes && cap_n < nn) || (records && cap_r < nr)) return RT_ERR_INVALID;
    if (nodes) std::memcpy(nodes, br.nodes.data(), br.nodes.size() * sizeof(float));
    if (records) std::memcpy(records, rec.data(), rec.size() * sizeof(float));
    return RT_OK;
}

int rt_scene_upload(rt_ctx* c, const rt_scene_desc* d)
{
    if (!c || !d) return RT_ERR_INVALID;
    if (int r = check_fault(c)) return r;
    int env0 = -1;
    if (int r = check_prims(c, d, &env0)) return r;
    hipSetDevice(c->device);
    const int nT = d->n_triangles, nS = d->n_spheres;
    int nP = 0;
    const int env = env0;
    for (int i = 0; i < d->n_lights; ++i) nP += d->lights[i].type == RT_LIGHT_AREA;
    const int G = nT + nS + nP;

    std::vector<float> geo((size_t)4 * G, 0.f);
    std::vector<int32_t> pid(G), mat(G);
    auto P = [d](int v, int k) { return d->positions[3 * v + k]; };
    for (int i = 0; i < nT; ++i) {
        const int32_t* tv = d->tri_vertices + 3 * i;
        float a[3], b[3], n[3], nn[3];
        for (int k = 0; k < 3; ++k) {
            a[k] = P(tv[0], k) - P(tv[1], k);
            b[k] = P(tv[0], k) - P(tv[2], k);
        }
        cross3(a, b, n);
        qt_normalize(n, nn);
        geo[4 * i] = nn[0];
        geo[4 * i + 1] = nn[1];
        geo[4 * i + 2] = nn[2];
        pid[i] = d->tri_prim_id[i];
        mat[i] = d->tri_material[i];
    }
    for (int i = 0; i < nS; ++i) {
        const float* s = d->spheres + 4 * i;
        for (int k = 0; k < 4; ++k) geo[4 * (nT + i) + k] = s[k];
        pid[nT + i] = d->sphere_prim_id[i];
        mat[nT + i] = d->sphere_material[i];
    }
    std::vector<RtLight> lights(d->n_lights);
    {
        int j = 0;
        for (int i = 0; i < d->n_lights; ++i) {
            const rt_light& l = d->lights[i];
            RtLight& L = lights[i];
            std::memset(&L, 0, sizeof(L));
            L.type = l.type;
            for (int k = 0; k < 3; ++k) {
                L.p[k] = l.p[k];
                L.u[k] = l.u[k];
                L.v[k] = l.v[k];
                L.I[k] = l.intensity[k];
            }
            cross3(L.u, L.v, L.n);
            L.area = std::sqrt(L.n[0] * L.n[0] + L.n[1] * L.n[1] + L.n[2] * L.n[2]);
            L.g = -1;
            if (l.type != RT_LIGHT_AREA) continue;
            int g = nT + nS + j++;
            L.g = g;
            float nn[3];
            qt_normalize(L.n, nn);
            for (int k = 0; k < 3; ++k) geo[4 * g + k] = nn[k];
            pid[g] = i;
            mat[g] = -(i + 1);
        }
    }
    /* analytic objects: validated postfix programs; every leaf gets a shading
     * index after the BVH primitives */
    std::vector<float> onodes, oxf;
    std::vector<int32_t> ofirst, lobj;
    int n_leaves = 0;
    if (d->n_objects < 0) return set_err(c, RT_ERR_INVALID, "negative object count");
    if (d->n_objects > 0) {
        if (!d->csg_nodes || !d->object_first || !d->object_prim_id)
            return set_err(c, RT_ERR_INVALID, "missing object array");
        if (d->object_first[0] != 0) return set_err(c, RT_ERR_INVALID, "object_first[0] must be 0");
        for (int i = 0; i < d->n_objects; ++i) {
            const int f = d->object_first[i], e = d->object_first[i + 1];
            if (e <= f) return set_err(c, RT_ERR_INVALID, "empty object program");
            int depth = 0, leaves = 0;
            for (int k = f; k < e; ++k) {
                const rt_csg_node& nd = d->csg_nodes[k];
                if (nd.op == RT_CSG_SPHERE || nd.op == RT_CSG_BOX || nd.op == RT_CSG_PLANE) {
                    if (nd.material < 0 || nd.material >= d->n_materials)
                        return set_err(c, RT_ERR_INVALID, "object material out of range");
                    if (nd.op == RT_CSG_PLANE && e - f != 1)
                        return set_err(c, RT_ERR_INVALID, "a plane is an object of its own, not a CSG operand");
                    if (nd.op == RT_CSG_SPHERE && !(nd.p[3] > 0.f))
                        return set_err(c, RT_ERR_INVALID, "sphere radius must be positive");
                    ++depth;
                    ++leaves;
                } else if (nd.op >= RT_CSG_ISECT && nd.op <= RT_CSG_XOR) {
                    if (depth < 2) return set_err(c, RT_ERR_INVALID, "CSG operation without two operands");
                    --depth;
                } else {
                    return set_err(c, RT_ERR_INVALID, "unknown CSG node");
                }
            }
            if (depth != 1) return set_err(c, RT_ERR_INVALID, "CSG program does not reduce to one object");
            if (leaves > RT_CSG_MAX_LEAVES) return set_err(c, RT_ERR_UNSUPPORTED, "CSG object with too many leaves");
            n_leaves += leaves;
        }
        const int n_nodes = d->object_first[d->n_objects];
        onodes.assign((size_t)8 * n_nodes, 0.f);
        ofirst.assign(d->object_first, d->object_first + d->n_objects + 1);
        geo.resize((size_t)4 * (G + n_leaves), 0.f);
        pid.resize(G + n_leaves);
        mat.resize(G + n_leaves);
        int g = G;
        for (int i = 0; i < d->n_objects; ++i)
            for (int k = d->object_first[i]; k < d->object_first[i + 1]; ++k) {
                const rt_csg_node& nd = d->csg_nodes[k];
                float* o = &onodes[8 * (size_t)k];
                for (int j = 0; j < 6; ++j) o[j] = nd.p[j];
                std::memcpy(&o[6], &nd.op, 4);
                int32_t leaf_g = -1;
                if (nd.op <= RT_CSG_PLANE) {
                    leaf_g = g++;
                    pid[leaf_g] = d->object_prim_id[i];
                    mat[leaf_g] = nd.material;
                    float* q = &geo[4 * (size_t)leaf_g];
                    if (nd.op == RT_CSG_SPHERE) {
                        for (int j = 0; j < 4; ++j) q[j] = nd.p[j];
                    } else if (nd.op == RT_CSG_PLANE) {
                        /* HitRecord normalises the normal (src/hitrecord.h:74) */
                        qt_normalize(nd.p, q);
                        q[3] = -2.f;
                    } else {
                        q[3] = -1.f;
                    }
                }
                std::memcpy(&o[7], &leaf_g, 4);
            }
        /* instance transforms of objects (world-to-object rows), and the
         * object of every leaf of a transformed object */
        if (d->object_xform) {
            static const float identity[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
            bool any = false;
            for (int i = 0; i < d->n_objects && !any; ++i) {
                const float* x = d->object_xform + 12 * (size_t)i;
                if (std::memcmp(x, identity, sizeof identity) == 0) continue;
                if (d->csg_nodes[d->object_first[i]].op == RT_CSG_PLANE)
                    return set_err(c, RT_ERR_INVALID, "a plane is given in world space, without a transform");
                any = true;
            }
            if (any) {
                oxf.assign(d->object_xform, d->object_xform + 12 * (size_t)d->n_objects);
                lobj.assign(n_leaves, -1);
                int leaf = 0;
                for (int i = 0; i < d->n_objects; ++i) {
                    const bool mapped = std::memcmp(d->object_xform + 12 * (size_t)i, identity, sizeof identity) != 0;
                    for (int k = d->object_first[i]; k < d->object_first[i + 1]; ++k)
                        if (d->csg_nodes[k].op <= RT_CSG_PLANE) {
                            if (mapped) lobj[leaf] = i;
                            ++leaf;
                        }
                }
            }
        }
    }

    rtb::BuildResult br;
    std::vector<float> rec;
    float box_lo[3], box_hi[3];
    if (int r = build_bvh(c, d, br, rec, box_lo, box_hi)) return r;
    std::vector<float> vn;
    if (d->normals && nT) {
        vn.resize((size_t)9 * nT);
        for (int i = 0; i < nT; ++i)
            for (int j = 0; j < 3; ++j) qt_normalize(d->normals + 3 * d->tri_vertices[3 * i + j], &vn[9 * i + 3 * j]);
    }
    std::vector<RtMaterial> mats(d->n_materials);
    for (int i = 0; i < d->n_materials; ++i)
        if (d->materials[i].type < RT_MAT_DIFFUSE || d->materials[i].type > RT_MAT_TEXTURE)
            return set_err(c, RT_ERR_INVALID, "unknown material type");
        else if (d->materials[i].type == RT_MAT_TEXTURE &&
                 (d->materials[i].texture < 0 || d->materials[i].texture >= d->n_textures))
            return set_err(c, RT_ERR_INVALID, "texture index out of range");
    /* textures: all texels in one (rgb, 0) array, a table of (first, w, h) */
    std::vector<float> tex;
    std::vector<int32_t> tex_info;
    if (d->n_textures < 0 || (d->n_textures > 0 && !d->textures)) return set_err(c, RT_ERR_INVALID, "bad textures");
    for (int i = 0; i < d->n_textures; ++i) {
        const rt_texture& t = d->textures[i];
        if (t.width <= 0 || t.height <= 0 || !t.rgb) return set_err(c, RT_ERR_INVALID, "bad texture");
        const size_t first = tex.size() / 4;
        tex_info.insert(tex_info.end(), {(int32_t)first, t.width, t.height, 0});
        for (size_t k = 0; k < (size_t)t.width * t.height; ++k)
            tex.insert(tex.end(), {t.rgb[3 * k], t.rgb[3 * k + 1], t.rgb[3 * k + 2], 0.f});
    }
    std::vector<float> tri_tc;
    if (d->tri_texcoords && nT) tri_tc.assign(d->tri_texcoords, d->tri_texcoords + (size_t)6 * nT);
    for (int i = 0; i < d->n_materials; ++i) {
        const rt_material& m = d->materials[i];
        RtMaterial& M = mats[i];
        std::memset(&M, 0, sizeof(M));
        M.type = m.type;
        for (int k = 0; k < 3; ++k) {
            M.kd[k] = m.kd[k];
            M.ks[k] = m.ks[k];
        }
        M.exponent = m.exponent;
        M.texture = m.type == RT_MAT_TEXTURE ? m.texture : -1;
        M.rho_d = lum_rgb(m.kd);
        M.rho_s = lum_rgb(m.ks);
    }
    // environment: texels x coefficient, sampling tables in double precision.
    // A light probe keeps its own texels for lookups; its sampling tables come
    // from a lat-long proxy (the probe looked up at each proxy texel centre),
    // so the density stays exact whatever the proxy's resolution.
    std::vector<float> erow, ecol, ep, ergb, prgb;
    std::vector<float> proxy;
    const float* esrc = d->env_rgb;
    int W = 0, H = 0, PW = 0, PH = 0;
    const bool probe = env >= 0 && d->env_kind == 1;
    if (env >= 0 && d->env_kind != 0 && d->env_kind != 1) return set_err(c, RT_ERR_INVALID, "unknown env_kind");
    if (probe) {
        PW = d->env_width;
        PH = d->env_height;
        const float* coef = d->lights[env].intensity;
        prgb.resize((size_t)4 * PW * PH);
        for (size_t i = 0; i < (size_t)PW * PH; ++i)
            for (int k = 0; k < 3; ++k) prgb[4 * i + k] = d->env_rgb[3 * i + k] * coef[k];
        proxy.resize((size_t)3 * RT_PROBE_PROXY_W * RT_PROBE_PROXY_H);
        for (int y = 0; y < RT_PROBE_PROXY_H; ++y)
            for (int x = 0; x < RT_PROBE_PROXY_W; ++x) {
                double th = (y + 0.5) / RT_PROBE_PROXY_H * M_PI, ph = ((x + 0.5) / RT_PROBE_PROXY_W - 0.5) * 2 * M_PI;
                float dx = (float)(std::sin(th) * std::sin(ph)), dy = (float)std::cos(th);
                float dz = (float)(-std::sin(th) * std::cos(ph));
                int t = rt_probe_texel(dx, dy, dz, PW, PH);
                for (int k = 0; k < 3; ++k)
                    proxy[3 * ((size_t)y * RT_PROBE_PROXY_W + x) + k] = d->env_rgb[3 * (size_t)t + k];
            }
        esrc = proxy.data();
    }
    if (env >= 0) {
        W = probe ? RT_PROBE_PROXY_W : d->env_width;
        H = probe ? RT_PROBE_PROXY_H : d->env_height;
        const float* coef = d->lights[env].intensity;
        ergb.resize((size_t)4 * W * H);
        std::vector<double> tp((size_t)W * H), roww(H, 0.0);
        std::vector<double> colc((size_t)H * (W + 1), 0.0);
        double total = 0;
        for (int y = 0; y < H; ++y) {
            double st = std::sin((y + 0.5) / H * M_PI);
            double acc = 0;
            for (int x = 0; x < W; ++x) {
                float rgb[3];
                for (int k = 0; k < 3; ++k) {
                    rgb[k] = esrc[3 * ((size_t)y * W + x) + k] * coef[k];
                    ergb[4 * ((size_t)y * W + x) + k] = rgb[k];
                }
                double w = (double)lum_rgb(rgb) * st;
                tp[(size_t)y * W + x] = w;
                acc += w;
                colc[(size_t)y * (W + 1) + x + 1] = acc;
            }
            for (int x = 0; x <= W; ++x) {
                double& v = colc[(size_t)y * (W + 1) + x];
                v = acc > 0 ? v / acc : (double)x / W;
            }
            roww[y] = acc;
            total += acc;
        }
        erow.resize(H + 1);
        erow[0] = 0.f;
        double acc = 0;
        for (int y = 0; y < H; ++y) {
            acc += roww[y];
            erow[y + 1] = (float)(total > 0 ? acc / total : (double)(y + 1) / H);
        }
        ecol.resize(colc.size());
        for (size_t i = 0; i < colc.size(); ++i) ecol[i] = (float)colc[i];
        ep.resize(tp.size());
        for (size_t i = 0; i < tp.size(); ++i) ep[i] = (float)(total > 0 ? tp[i] / total : 1.0 / ((double)W * H));
    }

    int r = 0;
    if ((r = upload(c, c->nodes, br.nodes.data(), br.nodes.size() * 4)) ||
        (r = upload(c, c->prims, rec.data(), rec.size() * 4)) ||
        (r = upload(c, c->prim_id, pid.data(), pid.size() * 4)) || (r = upload(c, c->mat, mat.data(), mat.size() * 4)) ||
        (r = upload(c, c->geo, geo.data(), geo.size() * 4)) ||
        (r = upload(c, c->vnormals, vn.data(), vn.size() * 4)) ||
        (r = upload(c, c->obj_nodes, onodes.data(), onodes.size() * 4)) ||
        (r = upload(c, c->tex_rgb, tex.data(), tex.size() * 4)) ||
        (r = upload(c, c->tex_info, tex_info.data(), tex_info.size() * 4)) ||
        (r = upload(c, c->tri_tc, tri_tc.data(), tri_tc.size() * 4)) ||
        (r = upload(c, c->obj_first, ofirst.data(), ofirst.size() * 4)) ||
        (r = upload(c, c->obj_xf, oxf.data(), oxf.size() * 4)) ||
        (r = upload(c, c->leaf_obj, lobj.data(), lobj.size() * 4)) ||
        (r = upload(c, c->mats, mats.data(), mats.size() * sizeof(RtMaterial))) ||
        (r = upload(c, c->lights, lights.data(), lights.size() * sizeof(RtLight))) ||
        (r = upload(c, c->env_rgb, ergb.data(), ergb.size() * 4)) ||
        (r = upload(c, c->env_row, erow.data(), erow.size() * 4)) ||
        (r = upload(c, c->env_col, ecol.data(), ecol.size() * 4)) || (r = upload(c, c->env_p, ep.data(), ep.size() * 4)) ||
        (r = upload(c, c->probe_rgb, prgb.data(), prgb.size() * 4)))
        return r;
    HostDevScene& s = c->scene;
    s.nodes = c->nodes.p;
    s.prims = c->prims.p;
    s.prim_id = (const int32_t*)c->prim_id.p;
    s.mat = (const int32_t*)c->mat.p;
    s.vnormals = vn.empty() ? nullptr : (const float*)c->vnormals.p;
    s.geo = c->geo.p;
    s.n_nodes = (int32_t)(br.nodes.size() / RT_NODE_FLOATS);
    s.n_tri = nT;
    s.n_sphere = nS;
    s.n_para = nP;
    s.mats = (const RtMaterial*)c->mats.p;
    s.lights = (const RtLight*)c->lights.p;
    s.n_lights = d->n_lights;
    s.env_light = env;
    s.env_w = W;
    s.env_h = H;
    s.env_rgb = c->env_rgb.p;
    s.env_row_cdf = (const float*)c->env_row.p;
    s.env_col_cdf = (const float*)c->env_col.p;
    s.env_texel_p = (const float*)c->env_p.p;
    s.env_kind = probe ? 1 : 0;
    s.probe_w = PW;
    s.probe_h = PH;
    s.probe_rgb = c->probe_rgb.p;
    s.obj_nodes = onodes.empty() ? nullptr : c->obj_nodes.p;
    s.obj_first = ofirst.empty() ? nullptr : (const int32_t*)c->obj_first.p;
    s.n_obj = d->n_objects;
    s.g_obj = G;
    s.obj_xf = oxf.empty() ? nullptr : c->obj_xf.p;
    s.leaf_obj = lobj.empty() ? nullptr : (const int32_t*)c->leaf_obj.p;
    s.tex_rgb = tex.empty() ? nullptr : c->tex_rgb.p;
    s.tex_info = tex_info.empty() ? nullptr : c->tex_info.p;
    s.tri_tc = tri_tc.empty() ? nullptr : (const float*)c->tri_tc.p;
    c->has_scene = true;
    c->bvh_depth = br.depth;
    c->bvh_stack = br.stack_need;
    c->stats.bvh_nodes = s.n_nodes;
    c->stats.bvh_prims = G;
    c->stats.bvh_depth = br.depth;
    c->stats.bvh_stack_need = br.stack_need;
    // scene box (finite primitives) for the guiding network's position encoding
    {
        float lo[3] = {box_lo[0], box_lo[1], box_lo[2]}, hi[3] = {box_hi[0], box_hi[1], box_hi[2]};
        for (int k = 0; k < 3; ++k) {
            if (!(lo[k] <= hi[k])) { lo[k] = 0.f; hi[k] = 1.f; }
            float pad = 1e-3f * std::max(1e-3f, hi[k] - lo[k]);
            c->scene_lo[k] = lo[k] - pad;
            c->scene_hi[k] = hi[k] + pad;
        }
        /* a new scene replaces any box given by rt_guide_set_bounds */
        c->guide_bounds_set = false;
        for (int k = 0; k < 3; ++k) {
            c->guide.lo[k] = c->scene_lo[k];
            c->guide.inv_ext[k] = 1.f / (c->scene_hi[k] - c->scene_lo[k]);
        }
    }
    return RT_OK;
}

static int trace_device(rt_ctx* c, int kind, const float* o, const float* d, int64_t n, void* out0, void* out1)
{
    if (!c->has_scene) return set_err(c, RT_ERR_NO_SCENE, "no scene uploaded");
    if (n == 0) return RT_OK;
    hipSetDevice(c->device);
    int grid = (int)std::min<int64_t>((n + 255) / 256, (int64_t)grid_for(c) * 4);
    if (int r = ensure_stack_overflow(c, grid)) return r;
    hipError_t e;
    HIP_TRY(c, hipEventRecord(c->ev_a, c->stream));
    if (kind == 0)
        e = launch_trace_closest(c->scene, (const float4*)o, (const float4*)d, n, (int32_t*)out0, (float*)out1, 256,
                                 grid, c->stream);
    else if (kind == 1)
        e = launch_trace_any(c->scene, (const float4*)o, (const float4*)d, n, (uint8_t*)out0, 256, grid, c->stream);
    else
        e = launch_trace_count(c->scene, (const float4*)o, (const float4*)d, n, (int32_t*)out0, (int32_t*)out1, 256,
                               grid, c->stream, kind == 3);
    if (e != hipSuccess) return set_err(c, RT_ERR_HIP, std::string("trace launch: ") + hipGetErrorString(e));
    HIP_TRY(c, hipEventRecord(c->ev_b, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    float ms = 0;
    hipEventElapsedTime(&ms, c->ev_a, c->ev_b);
    c->stats.extend_ms = ms;
    c->stats.extend_launches = 1;
    return RT_OK;
}

int rt_trace_closest_device(rt_ctx* c, const float* o, const float* d, int64_t n, int32_t* prim, float* t)
{
    if (!c || n < 0 || (n && (!o || !d || !prim || !t))) return RT_ERR_INVALID;
    if (int r = check_fault(c)) return r;
    return trace_device(c, 0, o, d, n, prim, t);
}

int rt_trace_any_device(rt_ctx* c, const float* o, const float* d, int64_t n, uint8_t* occ)
{
    if (!c || n < 0 || (n && (!o || !d || !occ))) return RT_ERR_INVALID;
    if (int r = check_fault(c)) return r;
    return trace_device(c, 1, o, d, n, occ, nullptr);
}

int rt_trace_count_device(rt_ctx* c, const float* o, const float* d, int64_t n, int32_t* nodes, int32_t* prims)
{
    if (!c || n < 0 || (n && (!o || !d || !nodes || !prims))) return RT_ERR_INVALID;
    if (int r = check_fault(c)) return r;
    return trace_device(c, 2, o, d, n, nodes, prims);
}

int rt_trace_any_count_device(rt_ctx* c, const float* o, const float* d, int64_t n, int32_t* nodes, int32_t* prims)
{
    if (!c || n < 0 || (n && (!o || !d || !nodes || !prims))) return RT_ERR_INVALID;
    if (int r = check_fault(c)) return r;
    return trace_device(c, 3, o, d, n, nodes, prims);
}

/* rt_trace_closest_device without allocation: the launch is capped at
 * RT_NOALLOC_BLOCKS blocks (any ray count is covered by the kernel's
 * grid-stride loop) and its stack overflow is the static scratch array whose
 * address rt_create looked up on this device. */
int rt_trace_closest_device_noalloc(rt_ctx* c, const float* o, const float* d, int64_t n, int32_t* prim, float* t)
{
    if (!c || n < 0 || (n && (!o || !d || !prim || !t))) return RT_ERR_INVALID;
    if (n > RT_NOALLOC_MAX_RAYS)
        return set_err(c, RT_ERR_INVALID, "rt_trace_closest_device_noalloc: more than RT_NOALLOC_MAX_RAYS rays");
    if (int r = check_fault(c)) return r;
    if (!c->has_scene) return set_err(c, RT_ERR_NO_SCENE, "no scene uploaded");
    if (n == 0) return RT_OK;
    hipSetDevice(c->device);
    int32_t* ovf = noalloc_scratch(c->device);
    if (!ovf) return set_err(c, RT_ERR_HIP, "allocation-free trace: no scratch address on this device");
    const int grid = (int)std::min<int64_t>((n + 255) / 256, RT_NOALLOC_BLOCKS);
    if (trav_lanes(grid) * RT_STACK_OVF > noalloc_scratch_entries())
        return set_err(c, RT_ERR_INVALID, "allocation-free trace: launch exceeds the scratch array");
    HostDevScene s = c->scene;
    s.stack_ovf = ovf;
    HIP_TRY(c, hipEventRecord(c->ev_a, c->stream));
    hipError_t e = launch_trace_closest(s, (const float4*)o, (const float4*)d, n, prim, t, 256, grid, c->stream);
    if (e != hipSuccess) return set_err(c, RT_ERR_HIP, std::string("trace launch: ") + hipGetErrorString(e));
    HIP_TRY(c, hipEventRecord(c->ev_b, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    float ms = 0;
    hipEventElapsedTime(&ms, c->ev_a, c->ev_b);
    c->stats.extend_ms = ms;
    c->stats.extend_launches = 1;
    return RT_OK;
}

static int trace_host(rt_ctx* c, int kind, const float* o, const float* d, int64_t n, void* out0, void* out1)
{
    if (!c->has_scene) return set_err(c, RT_ERR_NO_SCENE, "no scene uploaded");
    if (n == 0) return RT_OK;
    hipSetDevice(c->device);
    size_t rb = (size_t)n * 16;
    int r;
    if ((r = ensure(c, c->rb_o, rb)) || (r = ensure(c, c->rb_d, rb)) || (r = ensure(c, c->rb_out0, (size_t)n * 4)) ||
        (r = ensure(c, c->rb_out1, (size_t)n * 4)))
        return r;
    HIP_TRY(c, hipMemcpyAsync(c->rb_o.p, o, rb, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->rb_d.p, d, rb, hipMemcpyHostToDevice, c->stream));
    if ((r = trace_device(c, kind, (const float*)c->rb_o.p, (const float*)c->rb_d.p, n, c->rb_out0.p, c->rb_out1.p)))
        return r;
    size_t b0 = kind == 1 ? (size_t)n : (size_t)n * 4;
    HIP_TRY(c, hipMemcpy(out0, c->rb_out0.p, b0, hipMemcpyDeviceToHost));
    if (out1) HIP_TRY(c, hipMemcpy(out1, c->rb_out1.p, (size_t)n * 4, hipMemcpyDeviceToHost));
    return RT_OK;
}

int rt_trace_closest(rt_ctx* c, const float* o, const float* d, int64_t n, int32_t* prim, float* t)
{
    if (!c || n < 0 || (n && (!o || !d || !prim || !t))) return RT_ERR_INVALID;
    if (int r = check_fault(c)) return r;
    return trace_host(c, 0, o, d, n, prim, t);
}

int rt_trace_any(rt_ctx* c, const float* o, const float* d, int64_t n, uint8_t* occ)
{
    if (!c || n < 0 || (n && (!o || !d || !occ))) return RT_ERR_INVALID;
    if (int r = check_fault(c)) return r;
    return trace_host(c, 1, o, d, n, occ, nullptr);
}

/* ---------------------------------------------------------------------------
 * Tonemapper (include/rt_film.h)
 * ------------------------------------------------------------------------- */
static int tonemap_impl(rt_ctx* c, const float* film, int64_t n, float gamma, uint8_t* out)
{
    const int grid = c->n_cu * 8;
    HIP_TRY(c, hipEventRecord(c->ev_a, c->stream));
    hipError_t e = launch_tonemap(film, out, n, gamma, grid, c->stream);
    if (e != hipSuccess) return set_err(c, RT_ERR_HIP, std::string("tonemap launch: ") + hipGetErrorString(e));
    HIP_TRY(c, hipEventRecord(c->ev_b, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    float ms = 0;
    hipEventElapsedTime(&ms, c->ev_a, c->ev_b);
    c->stats.tonemap_ms = ms;
    return RT_OK;
}

int rt_film_tonemap_device(rt_ctx* c, const float* film, int32_t w, int32_t h, float gamma, uint8_t* rgb8)
{
    if (!c || w < 0 || h < 0 || ((int64_t)w * h && (!film || !rgb8))) return RT_ERR_INVALID;
    if (int r = check_fault(c)) return r;
    const int64_t n = (int64_t)w * h * 3;
    if (n == 0) return RT_OK;
    hipSetDevice(c->device);
    return tonemap_impl(c, film, n, gamma, rgb8);
}

int rt_film_tonemap(rt_ctx* c, const float* film, int32_t w, int32_t h, float gamma, uint8_t* rgb8)
{
    if (!c || w < 0 || h < 0 || ((int64_t)w * h && (!film || !rgb8))) return RT_ERR_INVALID;
    if (int r = check_fault(c)) return r;
    const int64_t n = (int64_t)w * h * 3;
    if (n == 0) return RT_OK;
    hipSetDevice(c->device);
    int r;
    if ((r = ensure(c, c->rb_o, (size_t)n * 4)) || (r = ensure(c, c->rb_out0, (size_t)n))) return r;
    HIP_TRY(c, hipMemcpyAsync(c->rb_o.p, film, (size_t)n * 4, hipMemcpyHostToDevice, c->stream));
    if ((r = tonemap_impl(c, (const float*)c->rb_o.p, n, gamma, (uint8_t*)c->rb_out0.p))) return r;
    HIP_TRY(c, hipMemcpy(rgb8, c->rb_out0.p, (size_t)n, hipMemcpyDeviceToHost));
    return RT_OK;
}

/* ---------------------------------------------------------------------------
 * Render driver
 * ------------------------------------------------------------------------- */
namespace {

// src/camera.cpp:34-46 with the Qt arithmetic of the reference.
DevView make_view(const rt_view& v, float* film)
{
    DevView V;
    std::memset(&V, 0, sizeof(V));
    const rt_camera& c = v.camera;
    float wv[3] = {c.eye[0] - c.look_at[0], c.eye[1] - c.look_at[1], c.eye[2] - c.look_at[2]};
    qt_normalize(wv, V.w);
    float uc[3];
    cross3(c.up, V.w, uc);
    qt_normalize(uc, V.u);
    cross3(V.w, V.u, V.v);
    for (int k = 0; k < 3; ++k) V.eye[k] = c.eye[k];
    float aspect = (float)c.width / c.height;
    V.t = (float)std::tan((double)(c.fov / 2));
    V.b = -V.t;
    V.r = aspect * V.t;
    V.l = -V.r;
    V.W = c.width;
    V.H = c.height;
    V.spp_x = v.spp_x;
    V.spp_y = v.spp_y;
    V.inv_spp = 1.f / (float)(v.spp_x * v.spp_y);
    V.film = film;
    return V;
}

int alloc_path_state(rt_ctx* c, size_t cap)
{
    if (c->ps_cap >= cap) return 0;
    for (auto& b : c->ps_bufs) release(b);
    c->ps_bufs.assign(16, DevBuf());
    float4** f4[] = {&c->ps.ray_o, &c->ps.ray_d, &c->ps.hit, &c->ps.alpha, &c->ps.radiance, &c->ps.pend_x,
                     &c->ps.pend_n, &c->ps.pend_d, &c->ps.pend_a, &c->ps.pend_c, &c->ps.sh_o, &c->ps.sh_d,
                     &c->ps.sh_c};
    int i = 0;
    for (float4** p : f4) {
        if (int r = ensure(c, c->ps_bufs[i], cap * sizeof(float4))) return r;
        *p = (float4*)c->ps_bufs[i++].p;
    }
    if (int r = ensure(c, c->ps_bufs[i], cap * sizeof(int4))) return r;
    c->ps.info = (int4*)c->ps_bufs[i++].p;
    for (int q = 0; q < 2; ++q) {
        if (int r = ensure(c, c->ps_bufs[i], cap * sizeof(int32_t))) return r;
        c->ps.queue[q] = (int32_t*)c->ps_bufs[i++].p;
    }
    if (int r = ensure(c, c->counters, sizeof(DevCounters))) return r;
    c->ps.cnt = (DevCounters*)c->counters.p;
    c->ps_cap = cap;
    return 0;
}

}  // namespace

/* ---------------------------------------------------------------------------
 * Guiding network state
 * ------------------------------------------------------------------------- */
static int guide_alloc_base(rt_ctx* c)
{
    int r;
    if ((r = ensure(c, c->g_theta, GD_NPARAM * 4)) || (r = ensure(c, c->g_w16, GD_NW16 * 2)) ||
        (r = ensure(c, c->g_m, GD_NPARAM * 4)) || (r = ensure(c, c->g_v, GD_NPARAM * 4)) ||
        (r = ensure(c, c->g_grad, GD_NPARAM * 4)) || (r = ensure(c, c->g_acc, GD_ACC_FLOATS * 4)) || (r = ensure(c, c->g_count, 32)) ||
        (r = ensure(c, c->g_part, (size_t)c->n_cu * GD_TRAIN_BLOCKS_PER_CU * GD_NPARAM * 4)))
        return r;
    GuideState& g = c->guide;
    g.theta = (float*)c->g_theta.p;
    g.w16i = (_Float16*)c->g_w16.p;
    g.w16b = g.w16i + GD_NW16I;
    g.part = (float*)c->g_part.p;
    g.part_blocks = c->n_cu * GD_TRAIN_BLOCKS_PER_CU;
    g.adam_m = (float*)c->g_m.p;
    g.adam_v = (float*)c->g_v.p;
    g.grad = (float*)c->g_grad.p;
    g.acc = (float*)c->g_acc.p;
    g.keep = g.acc + 4;
    g.rec_count = (int32_t*)c->g_count.p;
    g.trained = (unsigned long long*)((char*)c->g_count.p + 8);
    g.inferred = (unsigned long long*)((char*)c->g_count.p + 16);
    return 0;
}

static int guide_alloc_slots(rt_ctx* c, size_t slots, int batch)
{
    int r;
    size_t cap = (size_t)GD_TRAIN_CHUNKS * std::max(batch, 1 << 14);
    if ((r = ensure(c, c->g_rec, cap * GD_REC * 4)) || (r = ensure(c, c->g_aux, cap * 16)) ||
        (r = ensure(c, c->g_open, slots * 4)) || (r = ensure(c, c->g_out, slots * GD_PARAMS * sizeof(_Float16))) ||
        (r = ensure(c, c->g_in, slots * (sizeof(float4) + sizeof(float2)))))
        return r;
    GuideState& g = c->guide;
    g.rec = (float*)c->g_rec.p;
    g.rec_aux = (float4*)c->g_aux.p;
    g.open_rec = (int32_t*)c->g_open.p;
    g.out = (_Float16*)c->g_out.p;
    g.in_a = (float4*)c->g_in.p;
    g.in_b = (float2*)((char*)c->g_in.p + slots * sizeof(float4));
    g.rec_cap = (int32_t)std::min<size_t>(cap, c->g_rec.bytes / (GD_REC * 4));
    if (c->guide_slots < slots) {
        HIP_TRY(c, launch_guide_reset_records(g, (int)slots, c->stream));
        c->guide_slots = slots;
    } else {
        HIP_TRY(c, launch_guide_reset_records(g, (int)c->guide_slots, c->stream));
    }
    return 0;
}

static int guide_upload_params(rt_ctx* c, const float* theta)
{
    /* forward tiles, then backward tiles (their padding stays zero) */
    std::vector<_Float16> w16(GD_NW16, (_Float16)0.f);
    for (int i = 0; i < GD_NPARAM; ++i) {
        const int iw = gd_w16i_slot(i), ib = gd_w16b_slot(i);
        if (iw >= 0) w16[iw] = (_Float16)theta[i];
        if (ib >= 0) w16[GD_NW16I + ib] = (_Float16)theta[i];
    }
    HIP_TRY(c, hipMemcpy(c->guide.theta, theta, GD_NPARAM * 4, hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemcpy(c->guide.w16i, w16.data(), w16.size() * 2, hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemset(c->guide.adam_m, 0, GD_NPARAM * 4));
    HIP_TRY(c, hipMemset(c->guide.adam_v, 0, GD_NPARAM * 4));
    HIP_TRY(c, hipMemset(c->guide.grad, 0, GD_NPARAM * 4));
    HIP_TRY(c, hipMemset(c->guide.acc, 0, 16));
    HIP_TRY(c, hipMemset(c->guide.acc + GD_WSUM_OFF, 0, GD_WSUM_N * 4));
    HIP_TRY(c, hipMemset(c->guide.rec_count, 0, 32));
    c->guide_step = 0;
    c->guide_ready = true;
    return 0;
}

int rt_guide_param_count(void) { return GD_NPARAM; }

int rt_guide_init(rt_ctx* c, uint32_t seed)
{
    if (!c) return RT_ERR_INVALID;
    if (int r = check_fault(c)) return r;
    hipSetDevice(c->device);
    if (int r = guide_alloc_base(c)) return r;
    // Xavier-uniform weights, zero biases
    std::mt19937 gen(seed);
    std::vector<float> th(GD_NPARAM, 0.f);
    const int dims[4][3] = {{GD_OFF_W1, 64, 80}, {GD_OFF_W2, 64, 64}, {GD_OFF_W3, 64, 64}, {GD_OFF_W4, 48, 64}};
    for (auto& L : dims) {
        float lim = std::sqrt(6.f / (float)(L[1] + L[2]));
        std::uniform_real_distribution<float> U(-lim, lim);
        for (int i = 0; i < L[1] * L[2]; ++i) th[L[0] + i] = U(gen);
    }
    /* guide fraction starts at 0.05 + 0.9 sigmoid(-1.609) = 0.2 */
    th[GD_OFF_B4 + 40] = -1.609f;
    if (!c->guide_bounds_set)
        for (int k = 0; k < 3; ++k) {
            c->guide.lo[k] = c->scene_lo[k];
            c->guide.inv_ext[k] = 1.f / (c->scene_hi[k] - c->scene_lo[k]);
        }
    return guide_upload_params(c, th.data());
}

int rt_guide_set_params(rt_ctx* c, const float* theta, int32_t n)
{
    if (!c || !theta || n != GD_NPARAM) return RT_ERR_INVALID;
    hipSetDevice(c->device);
    if (int r = guide_alloc_base(c)) return r;
    if (!c->guide_ready && !c->guide_bounds_set)
        for (int k = 0; k < 3; ++k) {
            c->guide.lo[k] = c->scene_lo[k];
            c->guide.inv_ext[k] = 1.f / (c->scene_hi[k] - c->scene_lo[k]);
        }
    return guide_upload_params(c, theta);
}

int rt_guide_get_params(rt_ctx* c, float* theta, int32_t n)
{
    if (!c || !theta || n != GD_NPARAM) return RT_ERR_INVALID;
    if (!c->guide_ready) return set_err(c, RT_ERR_NO_GUIDE, "no guiding network");
    hipSetDevice(c->device);
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    HIP_TRY(c, hipMemcpy(theta, c->guide.theta, GD_NPARAM * 4, hipMemcpyDeviceToHost));
    return RT_OK;
}

int rt_guide_set_bounds(rt_ctx* c, const float* lo, const float* hi)
{
    if (!c || !lo || !hi) return RT_ERR_INVALID;
    for (int k = 0; k < 3; ++k) {
        if (!(hi[k] > lo[k])) return set_err(c, RT_ERR_INVALID, "empty bounds");
        c->guide.lo[k] = lo[k];
        c->guide.inv_ext[k] = 1.f / (hi[k] - lo[k]);
    }
    c->guide_bounds_set = true;
    return RT_OK;
}

static int guide_records_to_device(rt_ctx* c, const float* rec, int32_t n)
{
    if (int r = ensure(c, c->g_tmp_rec, (size_t)std::max(n, 1) * GD_REC * 4)) return r;
    HIP_TRY(c, hipMemcpy(c->g_tmp_rec.p, rec, (size_t)n * GD_REC * 4, hipMemcpyHostToDevice));
    return 0;
}

int rt_guide_forward(rt_ctx* c, const float* rec, int32_t n, float* y)
{
    if (!c || n < 0 || (n && (!rec || !y))) return RT_ERR_INVALID;
    if (int r = check_fault(c)) return r;
    if (!c->guide_ready) return set_err(c, RT_ERR_NO_GUIDE, "no guiding network");
    if (n == 0) return RT_OK;
    hipSetDevice(c->device);
    int r;
    if ((r = guide_records_to_device(c, rec, n)) || (r = ensure(c, c->g_tmp_y, (size_t)n * GD_OUT * 4))) return r;
    int grid = std::max(1, std::min(c->n_cu * GD_INFER_BLOCKS_PER_CU, (n + 255) / 256));
    HIP_TRY(c, hipMemsetAsync(c->g_tmp_y.p, 0, (size_t)n * GD_OUT * 4, c->stream));
    HIP_TRY(c, launch_guide_forward(c->guide, (const float*)c->g_tmp_rec.p, n, (float*)c->g_tmp_y.p, grid, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    HIP_TRY(c, hipMemcpy(y, c->g_tmp_y.p, (size_t)n * GD_OUT * 4, hipMemcpyDeviceToHost));
    return RT_OK;
}

int rt_guide_head(rt_ctx* c, const float* rec, int32_t n, float* out)
{
    if (!c || n < 0 || (n && (!rec || !out))) return RT_ERR_INVALID;
    if (int r = check_fault(c)) return r;
    if (!c->guide_ready) return set_err(c, RT_ERR_NO_GUIDE, "no guiding network");
    if (n == 0) return RT_OK;
    hipSetDevice(c->device);
    int r;
    if ((r = guide_records_to_device(c, rec, n)) || (r = ensure(c, c->g_tmp_y, (size_t)n * GD_PARAMS * 2))) return r;
    int grid = std::max(1, std::min(c->n_cu * GD_INFER_BLOCKS_PER_CU, (n + 255) / 256));
    HIP_TRY(c, hipMemsetAsync(c->g_tmp_y.p, 0, (size_t)n * GD_PARAMS * 2, c->stream));
    HIP_TRY(c, launch_guide_head(c->guide, (const float*)c->g_tmp_rec.p, n, (_Float16*)c->g_tmp_y.p, grid, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    std::vector<_Float16> h((size_t)n * GD_PARAMS);
    HIP_TRY(c, hipMemcpy(h.data(), c->g_tmp_y.p, h.size() * 2, hipMemcpyDeviceToHost));
    /* from the kernel's lane-group layout to lobe j at [5j .. 5j+4], the
     * guide probability at [40], zeros after it */
    for (int32_t i = 0; i < n; ++i) {
        const _Float16* src = h.data() + (size_t)i * GD_PARAMS;
        float* dst = out + (size_t)i * GD_PARAMS;
        for (int k = 0; k < GD_PARAMS; ++k) dst[k] = 0.f;
        for (int j = 0; j < GD_LOBES; ++j)
            for (int e = 0; e < 5; ++e) dst[5 * j + e] = (float)src[gd_lobe_off(j) + e];
        dst[40] = (float)src[GD_ALPHA_OFF];
    }
    return RT_OK;
}

/* Records given by the caller: weights are taken as given (record[13]). */
static int guide_grad_on(rt_ctx* c, int32_t n, GuideState* used)
{
    GuideState g = c->guide;
    g.rec = (float*)c->g_tmp_rec.p;
    g.rec_cap = n;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    HIP_TRY(c, hipMemcpy(g.rec_count, &n, 4, hipMemcpyHostToDevice));
    HIP_TRY(c, launch_guide_zero_acc(g, c->stream));
    PathState none{};
    int grid = std::max(1, std::min(grid_for(c), (n + 255) / 256));
    HIP_TRY(c, launch_guide_weights(g, none, 0, n, grid, c->stream));
    int tgrid = std::max(1, std::min(c->n_cu * GD_TRAIN_BLOCKS_PER_CU, (n + 63) / 64));
    HIP_TRY(c, launch_guide_train(g, g.rec, g.rec_count, 0, n, tgrid, c->stream));
    if (used) *used = g;
    return 0;
}

int rt_guide_gradient(rt_ctx* c, const float* rec, int32_t n, float* grad, float* loss)
{
    if (!c || n <= 0 || !rec || !grad) return RT_ERR_INVALID;
    if (int r = check_fault(c)) return r;
    if (!c->guide_ready) return set_err(c, RT_ERR_NO_GUIDE, "no guiding network");
    hipSetDevice(c->device);
    if (int r = guide_records_to_device(c, rec, n)) return r;
    HIP_TRY(c, hipMemsetAsync(c->guide.grad, 0, GD_NPARAM * 4, c->stream));
    if (int r = guide_grad_on(c, n, nullptr)) return r;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    HIP_TRY(c, hipMemcpy(grad, c->guide.grad, GD_NPARAM * 4, hipMemcpyDeviceToHost));
    for (int i = 0; i < GD_NPARAM; ++i) grad[i] /= (float)n;
    float acc[4];
    HIP_TRY(c, hipMemcpy(acc, c->guide.acc, 16, hipMemcpyDeviceToHost));
    if (loss) *loss = acc[1] / (float)n;
    HIP_TRY(c, hipMemset(c->guide.grad, 0, GD_NPARAM * 4));
    return RT_OK;
}

int rt_guide_train(rt_ctx* c, const float* rec, int32_t n, int32_t steps, float lr)
{
    if (!c || n <= 0 || !rec || steps < 0) return RT_ERR_INVALID;
    if (int r = check_fault(c)) return r;
    if (!c->guide_ready) return set_err(c, RT_ERR_NO_GUIDE, "no guiding network");
    hipSetDevice(c->device);
    if (int r = guide_records_to_device(c, rec, n)) return r;
    for (int s = 0; s < steps; ++s) {
        HIP_TRY(c, hipMemsetAsync(c->guide.grad, 0, GD_NPARAM * 4, c->stream));
        GuideState g;
        if (int r = guide_grad_on(c, n, &g)) return r;
        ++c->guide_step;
        HIP_TRY(c, launch_guide_adam(g, lr > 0.f ? lr : 1e-3f, c->guide_step, g.rec_count, 0, n, c->stream));
    }
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return RT_OK;
}

static int render_impl(rt_ctx* c, const rt_view* views, int32_t nv, const rt_params* prm, float* const* films,
                       bool device_films)
{
    if (!c || !views || nv <= 0 || !prm || !films) return RT_ERR_INVALID;
    if (int r = check_fault(c)) return r;
    if (!c->has_scene) return set_err(c, RT_ERR_NO_SCENE, "no scene uploaded");
    if (prm->mode != RT_MODE_COMPAT && prm->mode != RT_MODE_UNBIASED && prm->mode != RT_MODE_GUIDED)
        return set_err(c, RT_ERR_INVALID, "unsupported mode");
    if (prm->mode == RT_MODE_GUIDED && !c->guide_ready)
        return set_err(c, RT_ERR_NO_GUIDE, "guided mode needs rt_guide_init first");
    if (prm->max_depth < 1 || prm->max_depth > 64 || prm->rr_prob < 0.f || prm->rr_prob >= 1.f || prm->rr_start < 0)
        return set_err(c, RT_ERR_INVALID, "bad render parameters");
    int64_t total_samples = 0;
    int64_t max_pix = 0;
    for (int i = 0; i < nv; ++i) {
        const rt_view& v = views[i];
        if (v.camera.width <= 0 || v.camera.height <= 0 || v.spp_x <= 0 || v.spp_y <= 0 || !films[i])
            return set_err(c, RT_ERR_INVALID, "bad view");
        int64_t npix = (int64_t)v.camera.width * v.camera.height;
        if (npix > (1 << 26)) return set_err(c, RT_ERR_INVALID, "view too large");
        total_samples += npix * v.spp_x * v.spp_y;
        max_pix = std::max(max_pix, npix);
    }
    hipSetDevice(c->device);
    /* default 64M path slots (~15.6 GB of path state, 28 GB with the guide's):
     * every pass pays its bounces' launch tails and its finishing launch's
     * latency chain once, so fewer passes are faster; the 1024^2 x 64 spp
     * film takes one pass (83.2 -> 80.2 ms against 32M slots and two) */
    int64_t P = c->cfg.path_slots > 0 ? c->cfg.path_slots : (int64_t)1 << 26;
    P = std::min<int64_t>(P, std::max<int64_t>(total_samples, 1));
    P = std::max<int64_t>(P, 256);

    // film storage on the device for the host entry point
    std::vector<float*> dfilm(nv);
    if (!device_films) {
        size_t tot = 0;
        for (int i = 0; i < nv; ++i) tot += (size_t)views[i].camera.width * views[i].camera.height * 3;
        if (int r = ensure(c, c->films, tot * 4)) return r;
        size_t off = 0;
        for (int i = 0; i < nv; ++i) {
            size_t n = (size_t)views[i].camera.width * views[i].camera.height * 3;
            dfilm[i] = (float*)c->films.p + off;
            HIP_TRY(c, hipMemcpyAsync(dfilm[i], films[i], n * 4, hipMemcpyHostToDevice, c->stream));
            off += n;
        }
    } else {
        for (int i = 0; i < nv; ++i) dfilm[i] = films[i];
    }

    // tiles and passes: each pass holds at most P slots and at most one tile
    // per (view, pixel); tiles of a view run in increasing sample order.
    std::vector<DevView> dv(nv);
    for (int i = 0; i < nv; ++i) dv[i] = make_view(views[i], dfilm[i]);
    std::vector<DevTile> tiles;
    std::vector<DevPass> passes;
    DevPass cur{0, 0, 0, 0};
    std::vector<int> in_pass;  // views present in the current pass
    auto close_pass = [&]() {
        if (cur.tile1 > cur.tile0) passes.push_back(cur);
        cur = DevPass{(int32_t)tiles.size(), (int32_t)tiles.size(), 0, 0};
        in_pass.clear();
    };
    for (int v = 0; v < nv; ++v) {
        int64_t npix = (int64_t)views[v].camera.width * views[v].camera.height;
        int spp = views[v].spp_x * views[v].spp_y;
        int64_t chunk = std::min<int64_t>(npix, P);
        for (int s0 = 0; s0 < spp;) {
            int ns = (int)std::max<int64_t>(1, std::min<int64_t>(spp - s0, P / chunk));
            for (int64_t p0 = 0; p0 < npix; p0 += chunk) {
                int64_t np_ = std::min<int64_t>(chunk, npix - p0);
                int64_t slots = np_ * ns;
                bool dup = std::find(in_pass.begin(), in_pass.end(), v) != in_pass.end() && chunk == npix;
                if (cur.slots + slots > P || dup) close_pass();
                DevTile t{v, (int32_t)p0, (int32_t)np_, s0, ns, cur.slots, cur.pixels, 0};
                tiles.push_back(t);
                cur.tile1 = (int32_t)tiles.size();
                cur.slots += (int32_t)slots;
                cur.pixels += (int32_t)np_;
                in_pass.push_back(v);
            }
            s0 += ns;
            if (chunk < npix) close_pass();  // pixel-chunked views: one sample block per pass
        }
    }
    close_pass();

    int r;
    if ((r = alloc_path_state(c, (size_t)P))) return r;
    const bool guided = prm->mode == RT_MODE_GUIDED;
    const bool train = guided && prm->guide_train != 0;
    const int batch = prm->guide_batch > 0 ? prm->guide_batch : (1 << 18);
    const float lr = prm->guide_lr > 0.f ? prm->guide_lr : 1e-3f;
    if (guided) {
        if ((r = guide_alloc_slots(c, (size_t)P, batch))) return r;
    }
    if ((r = upload(c, c->views, dv.data(), dv.size() * sizeof(DevView))) ||
        (r = upload(c, c->tiles, tiles.data(), tiles.size() * sizeof(DevTile))) ||
        (r = upload(c, c->passes, passes.data(), passes.size() * sizeof(DevPass))))
        return r;
    const DevView* d_views = (const DevView*)c->views.p;
    const DevTile* d_tiles = (const DevTile*)c->tiles.p;
    const DevPass* d_passes = (const DevPass*)c->passes.p;

    PtParams pp{prm->mode, prm->max_depth, prm->rr_start, prm->rr_prob, prm->seed, train ? 1 : 0};
    if (train) {
        /* Records are kept with a probability that spreads the record buffer
         * over all path depths (a full buffer would keep only the first
         * vertices appended, i.e. the shallow ones).  The first training pass
         * of a context assumes 6 vertices per path slot; every training pass
         * then sets it from the vertices it offered (k_guide_keep), so that
         * the next fills about 90% of the buffer: GD_TRAIN_CHUNKS steps of
         * `batch` records each.  The value carries over to later renders,
         * scaled by the ratio of their pass sizes. */
        if (c->guide_keep_slots == 0) {
            const float keep0 = (float)std::min(1.0, (double)c->guide.rec_cap / std::max(1.0, (double)P * 6.0));
            HIP_TRY(c, hipMemcpyAsync(c->guide.keep, &keep0, sizeof(float), hipMemcpyHostToDevice, c->stream));
        } else if (c->guide_keep_slots != P) {
            HIP_TRY(c, launch_guide_keep(c->guide, (float)((double)c->guide_keep_slots / (double)P), c->stream));
        }
        c->guide_keep_slots = P;
    }
    const int grid = grid_for(c);
    const int depth = prm->max_depth;
    if (int r = ensure_stack_overflow(c, grid)) return r;
    /* Every pass enqueues the same kernels with the same arguments (the pass
     * index lives in the device counters), so one captured pass can be
     * replayed as a hipGraph - except with training, whose Adam step number
     * is a per-step argument.  A graph carries no per-kernel timing events. */
    const bool graph = c->cfg.use_graph != 0 && !train;
    const bool timing = c->timing != 0 && !graph;
    /* per bounce: before and after the finishing launch (from
     * fr.min_depth on), before extend, after extend, after inference, after
     * shade, after shadow; per pass with training: before and after the
     * steps */
    size_t need_ev = timing ? (size_t)passes.size() * (depth * 7 + (train ? 2 : 0)) : 0;
    while (c->ev_pool.size() < need_ev) {
        hipEvent_t e;
        HIP_TRY(c, hipEventCreate(&e));
        c->ev_pool.push_back(e);
    }
    HIP_TRY(c, hipGetLastError());
    if ((r = (launch_reset_counters(c->ps, c->stream) == hipSuccess ? 0 : RT_ERR_HIP))) return set_err(c, r, "reset");
    unsigned long long gcount0[2] = {0, 0};  /* records trained, vertices inferred */
    if (guided) HIP_TRY(c, hipMemcpy(gcount0, c->guide.trained, sizeof(gcount0), hipMemcpyDeviceToHost));
    const int steps0 = c->guide_step;
    HIP_TRY(c, hipEventRecord(c->ev_a, c->stream));
    size_t ev = 0;
    std::vector<size_t> bounce_ev, train_ev, finish_ev;  /* index of the first event of each group */
    const FinishRule fr = !guided && !c->cfg.count_traversal ? finish_rule(c) : FinishRule{INT_MAX, 0};
    bool cancelled = false;
    /* one pass: generate, max_depth x (extend, [inference], shade, shadow,
     * advance), accumulate, [training steps], pass end */
    auto enqueue_pass = [&]() -> hipError_t {
        hipError_t e;
#define QCHK(x)                              \
    do {                                     \
        if ((e = (x)) != hipSuccess) return e; \
    } while (0)
        auto mark = [&]() {
            if (!timing) return hipSuccess;
            if (ev >= c->ev_pool.size()) return hipErrorInvalidValue; /* the pool is sized for the marks above */
            return hipEventRecord(c->ev_pool[ev++], c->stream);
        };
        QCHK(launch_generate(c->ps, d_views, d_tiles, d_passes, pp, grid, c->stream));
        for (int dd = 0; dd < depth; ++dd) {
            int q = dd & 1;
            if (dd >= fr.min_depth) {
                /* the remaining bounces in one launch once the queue is
                 * short (decided on the device; a no-op otherwise) */
                if (timing) finish_ev.push_back(ev);
                QCHK(mark());
                QCHK(launch_finish(c->scene, c->ps, pp, dd, q, fr, grid, c->stream));
                QCHK(mark());
            }
            if (timing) bounce_ev.push_back(ev);
            QCHK(mark());
            QCHK(launch_extend(c->scene, c->ps, q, dd, fr, c->cfg.count_traversal, grid, c->stream));
            QCHK(mark());
            if (guided) {
                /* the last bounce neither samples the guide nor weights against it */
                if (dd + 1 < depth)
                    QCHK(launch_guide_infer(c->scene, c->ps, c->guide, q, c->n_cu * GD_INFER_BLOCKS_PER_CU,
                                            c->stream));
                QCHK(mark());
                QCHK(launch_shade_guided(c->scene, c->ps, c->guide, pp, dd, q, grid, c->stream));
            } else {
                QCHK(mark());
                QCHK(launch_shade(c->scene, c->ps, pp, dd, q, fr, grid, c->stream));
            }
            QCHK(mark());
            QCHK(launch_shadow(c->scene, c->ps, pp, q, dd, fr, c->cfg.count_traversal, grid, c->stream));
            QCHK(mark());
            QCHK(launch_advance(c->ps, q, dd, fr, c->stream));
        }
        QCHK(launch_accumulate(c->ps, d_views, d_tiles, d_passes, pp.mode == RT_MODE_COMPAT, grid, c->stream));
        if (train) {
            if (timing) train_ev.push_back(ev);
            QCHK(mark());
            /* the pass's records in GD_TRAIN_CHUNKS chunks of `batch`: one Adam step each */
            for (int k = 0; k < GD_TRAIN_CHUNKS * std::max(1, c->cfg.guide_epochs); ++k) {
                const int off = (k % GD_TRAIN_CHUNKS) * batch;
                ++c->guide_step;
                QCHK(launch_guide_zero_acc(c->guide, c->stream));
                QCHK(launch_guide_weights(c->guide, c->ps, off, batch, c->n_cu * 4, c->stream));
                QCHK(launch_guide_train(c->guide, c->guide.rec, c->guide.rec_count, off, batch,
                                        c->n_cu * GD_TRAIN_BLOCKS_PER_CU, c->stream));
                QCHK(launch_guide_adam(c->guide, lr, c->guide_step, c->guide.rec_count, off, batch, c->stream));
            }
            QCHK(mark());
            QCHK(launch_guide_keep(c->guide, 0.f, c->stream));
            QCHK(launch_guide_reset_records(c->guide, (int)P, c->stream));
        }
        QCHK(launch_pass_end(c->ps, c->stream));
#undef QCHK
        return hipSuccess;
    };
    hipGraphExec_t gexec = nullptr;
    if (graph) {
        hipGraph_t gr = nullptr;
        hipError_t e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) {
            hipError_t e1 = enqueue_pass();
            hipError_t e2 = hipStreamEndCapture(c->stream, &gr);
            e = e1 != hipSuccess ? e1 : e2;
        }
        if (e == hipSuccess) e = hipGraphInstantiate(&gexec, gr, nullptr, nullptr, 0);
        if (gr) hipGraphDestroy(gr);
        if (e != hipSuccess) return set_err(c, RT_ERR_HIP, std::string("pass graph: ") + hipGetErrorString(e));
    }
    InFlight flight(c, 2);
    for (size_t pi = 0; pi < passes.size(); ++pi) {
        if (c->stop) {
            cancelled = true;
            break;
        }
        hipError_t e = graph ? hipGraphLaunch(gexec, c->stream) : enqueue_pass();
        if (e == hipSuccess) e = flight.unit_done();
        if (e != hipSuccess) {
            if (gexec) hipGraphExecDestroy(gexec);
            return set_err(c, RT_ERR_HIP, std::string("render pass: ") + hipGetErrorString(e));
        }
    }
    HIP_TRY(c, hipEventRecord(c->ev_b, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (gexec) HIP_TRY(c, hipGraphExecDestroy(gexec));
    float ms = 0;
    hipEventElapsedTime(&ms, c->ev_a, c->ev_b);
    c->stats.render_ms = ms;
    auto span = [&](size_t i, size_t j) {
        float t = 0;
        hipEventElapsedTime(&t, c->ev_pool[i], c->ev_pool[j]);
        return (double)t;
    };
    double ext = 0, sh = 0, inf = 0, trn = 0;
    for (size_t i : bounce_ev) {
        ext += span(i, i + 1);
        inf += span(i + 1, i + 2);
        sh += span(i + 3, i + 4);
    }
    for (size_t i : train_ev) trn += span(i, i + 1);
    double fin = 0;
    for (size_t i : finish_ev) fin += span(i, i + 1);
    c->stats.finish_ms = fin;
    c->stats.extend_ms = ext;
    c->stats.shadow_ms = sh;
    c->stats.extend_launches = 0;  /* set from the device counters below */
    c->stats.guide_infer_ms = guided ? inf : 0.0;
    c->stats.guide_train_ms = trn;
    c->stats.guide_train_steps = c->guide_step - steps0;
    c->stats.guide_train_samples = 0;
    c->stats.guide_infer_samples = 0;
    if (guided) {
        unsigned long long gcount1[2] = {0, 0};
        HIP_TRY(c, hipMemcpy(gcount1, c->guide.trained, sizeof(gcount1), hipMemcpyDeviceToHost));
        c->stats.guide_train_samples = (int64_t)(gcount1[0] - gcount0[0]);
        c->stats.guide_infer_samples = (int64_t)(gcount1[1] - gcount0[1]);
    }
    DevCounters hc;
    HIP_TRY(c, hipMemcpy(&hc, c->ps.cnt, sizeof(hc), hipMemcpyDeviceToHost));
    c->stats.camera_rays = (int64_t)hc.camera_rays;
    c->stats.extension_rays = (int64_t)hc.extension_rays;
    c->stats.shadow_rays = (int64_t)hc.shadow_rays;
    c->stats.node_visits = (int64_t)hc.node_visits;
    c->stats.prim_tests = (int64_t)hc.prim_tests;
    c->stats.shadow_node_visits = (int64_t)hc.sh_node_visits;
    c->stats.shadow_prim_tests = (int64_t)hc.sh_prim_tests;
    c->stats.node_visits_lds = (int64_t)hc.node_visits_lds;
    c->stats.shadow_node_visits_lds = (int64_t)hc.sh_node_visits_lds;
    c->stats.finish_rays = (int64_t)hc.finish_rays;
    c->stats.finish_shadow_rays = (int64_t)hc.finish_shadow_rays;
    c->stats.finish_bounce = hc.finished;
    c->stats.extend_launches = timing ? (int64_t)hc.busy_bounces : 0;
    c->stats.paths_per_pass = (int32_t)P;
    c->stats.passes = (int32_t)passes.size();
    if (!device_films) {
        for (int i = 0; i < nv; ++i) {
            size_t n = (size_t)views[i].camera.width * views[i].camera.height * 3;
            HIP_TRY(c, hipMemcpy(films[i], dfilm[i], n * 4, hipMemcpyDeviceToHost));
        }
    }
    if (cancelled) return set_err(c, RT_ERR_CANCELLED, "render cancelled");
    return RT_OK;
}

/* ---------------------------------------------------------------------------
 * Markov-chain renderers (include/rt_mlt.h)
 * ------------------------------------------------------------------------- */
namespace {

/* The bounces of the paths in queue 0 (compat or unbiased estimator):
 * per bounce, the finishing launch (a no-op until the queue is short) and
 * the extend, shade, shadow and advance launches (no-ops from then on). */
hipError_t trace_bounces(rt_ctx* c, const PtParams& pp, int grid)
{
    const FinishRule fr = finish_rule(c);
    hipError_t e;
    for (int dd = 0; dd < pp.max_depth; ++dd) {
        const int q = dd & 1;
        if (dd >= fr.min_depth && (e = launch_finish(c->scene, c->ps, pp, dd, q, fr, grid, c->stream))) return e;
        if ((e = launch_extend(c->scene, c->ps, q, dd, fr, 0, grid, c->stream))) return e;
        if ((e = launch_shade(c->scene, c->ps, pp, dd, q, fr, grid, c->stream))) return e;
        if ((e = launch_shadow(c->scene, c->ps, pp, q, dd, fr, 0, grid, c->stream))) return e;
        if ((e = launch_advance(c->ps, q, dd, fr, c->stream))) return e;
    }
    return hipSuccess;
}

/* Trace the n paths whose primary samples are in c->mlt_prop (stride cap):
 * the compat path kernels, depth bounces, with the given length rule. */
hipError_t mlt_trace(rt_ctx* c, int n, int cap, int max_depth, int rr_start, float rr_prob, uint32_t seed,
                     int pixel_units)
{
    PtParams pp{RT_MODE_COMPAT, max_depth, rr_start, rr_prob, seed, 0, (const float4*)c->mlt_prop.p, cap};
    const int grid = grid_for(c);
    hipError_t e;
    if ((e = launch_generate_pss(c->ps, (const DevView*)c->mlt_view.p, n, pp, pixel_units, grid, c->stream)))
        return e;
    return trace_bounces(c, pp, grid);
}

/* Path slots, proposal and state buffers for cap chains of G groups. */
int mlt_alloc(rt_ctx* c, int64_t cap, int G, int64_t npix)
{
    int r;
    if ((r = alloc_path_state(c, (size_t)cap))) return r;
    if ((r = ensure_stack_overflow(c, grid_for(c)))) return r;
    if ((r = ensure(c, c->mlt_cur, (size_t)G * cap * sizeof(float4))) ||
        (r = ensure(c, c->mlt_prop, (size_t)G * cap * sizeof(float4))) ||
        (r = ensure(c, c->mlt_val, (size_t)cap * sizeof(float4))) ||
        (r = ensure(c, c->mlt_energy, (size_t)cap * sizeof(double))) ||
        (r = ensure(c, c->mlt_src, (size_t)cap * sizeof(int32_t))) ||
        (r = ensure(c, c->mlt_acc, (size_t)npix * 3 * sizeof(unsigned long long))) ||
        (r = ensure(c, c->mlt_cnt, sizeof(unsigned long long))) || (r = ensure(c, c->mlt_view, sizeof(DevView))))
        return r;
    return 0;
}

MltDev mlt_dev(rt_ctx* c, int cap, int G, int W, int H, const rt_mlt_params& p)
{
    MltDev m{};
    m.cur = (float4*)c->mlt_cur.p;
    m.prop = (float4*)c->mlt_prop.p;
    m.cur_val = (float4*)c->mlt_val.p;
    m.energy = (double*)c->mlt_energy.p;
    m.acc = (unsigned long long*)c->mlt_acc.p;
    m.accepted = (unsigned long long*)c->mlt_cnt.p;
    m.cap = cap;
    m.G = G;
    m.W = W;
    m.H = H;
    m.seed = p.seed;
    m.kind = p.kind;
    m.large_prob = p.large_step_prob;
    m.spp_x = p.erpt_spp_x;
    m.spp_y = p.erpt_spp_y;
    return m;
}

/* Fixed-point unit for deposits of at most `bound` each, n_dep of them in
 * total: 2^-30 of the bound, coarser only if the sum of n_dep largest
 * deposits could overflow 63 bits. */
double mlt_unit(double bound, double n_dep)
{
    int frac = 30;
    while (frac > 0 && std::ldexp(1.0, frac) * (n_dep + 1.0) >= std::ldexp(1.0, 62)) --frac;
    return bound > 0 ? bound * std::ldexp(1.0, -frac) : 1.0;
}

#define MLT_TRY(call)                                                                              \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess) return set_err(c, RT_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

/* MetropolisRenderer::render, src/metropolisrenderer.cpp:56-200, with
 * chains in lockstep: bootstrap, chain starts drawn by luminance, `iters`
 * mutations per chain. */
int run_metropolis(rt_ctx* c, const rt_mlt_params& p, int W, int H, int64_t cap_slots, rt_mlt_result& res,
                   double& unit)
{
    const int G = 2 * p.max_depth + 1;
    const int64_t area = (int64_t)W * H;
    const int64_t total = (int64_t)p.mutations * area;
    const int64_t C = p.chains > 0 ? p.chains : std::max<int64_t>(1, std::min<int64_t>(total / 16, (int64_t)1 << 22));
    const int64_t iters = std::max<int64_t>(1, (total + C / 2) / C);
    const int64_t nb = std::max<int64_t>(p.bootstrap, C);
    if (nb > INT32_MAX || C > INT32_MAX || iters > INT32_MAX) return set_err(c, RT_ERR_INVALID, "too many chains or mutations");
    const int cap = (int)std::min<int64_t>(cap_slots, std::max(C, nb));
    if (int r = mlt_alloc(c, cap, G, area)) return r;
    if (int r = ensure(c, c->mlt_seed, (size_t)nb * sizeof(float4))) return r;
    MltDev m = mlt_dev(c, cap, G, W, H, p);
    const int grid = grid_for(c);
    InFlight flight(c, 4);
    float4* bval = (float4*)c->mlt_seed.p;
    /* bootstrap: full-length paths (MetropolisSample's default pathLength =
     * MAX_DEPTH, src/metropolissample.h:35), no roulette compensation */
    for (int64_t base = 0; base < nb; base += cap) {
        const int n = (int)std::min<int64_t>(cap, nb - base);
        MLT_TRY(launch_mlt_fill(m, RT_MLT_STREAM_BOOTSTRAP, (uint32_t)base, n, 0, grid, c->stream));
        MLT_TRY(mlt_trace(c, n, cap, p.max_depth, p.rr_start, 0.f, p.seed, 0));
        MLT_TRY(launch_mlt_values(c->ps, n, bval + base, 0, grid, c->stream));
    }
    std::vector<float4> hv((size_t)nb);
    MLT_TRY(hipMemcpyAsync(hv.data(), bval, (size_t)nb * sizeof(float4), hipMemcpyDeviceToHost, c->stream));
    MLT_TRY(hipStreamSynchronize(c->stream));
    std::vector<double> cdf((size_t)nb);
    double acc = 0;
    for (int64_t i = 0; i < nb; ++i) cdf[i] = acc += (double)hv[i].w;
    const float b = (float)(acc / (double)nb);
    res.b = b;
    res.chains = C;
    res.seed_paths = nb;
    res.mutations = 0;
    if (!(acc > 0)) return RT_OK;  /* nothing reaches the camera: no deposit */
    if (int r = upload(c, c->mlt_cdf, cdf.data(), cdf.size() * sizeof(double))) return r;
    m.scale = (float)((double)b * (double)area / ((double)C * (double)iters));
    /* a deposit's channel is at most scale / 0.114 (luminance weights) */
    unit = mlt_unit((double)m.scale * 9.0, 2.0 * (double)C * (double)iters);
    m.inv_unit = 1.0 / unit;
    for (int64_t c0 = 0; c0 < C; c0 += cap) {
        const int n = (int)std::min<int64_t>(cap, C - c0);
        m.chain0 = (uint32_t)c0;
        MLT_TRY(launch_mlt_start(m, (const double*)c->mlt_cdf.p, nb, acc, bval, n, grid, c->stream));
        for (int64_t it = 0; it < iters; ++it) {
            if (c->stop) return cancel(c);
            MLT_TRY(launch_mlt_propose(m, (uint32_t)it, n, grid, c->stream));
            MLT_TRY(mlt_trace(c, n, cap, p.max_depth, p.rr_start, p.termination_prob, p.seed, 0));
            MLT_TRY(launch_mlt_accept(m, c->ps, (uint32_t)it, n, grid, c->stream));
            MLT_TRY(flight.unit_done());
        }
    }
    res.mutations = C * iters;
    return RT_OK;
}

/* EnergyRedistributionRenderer::render, src/energyredistributionrenderer.cpp:38-131 */
int run_erpt(rt_ctx* c, const rt_mlt_params& p, int W, int H, int64_t cap_slots, rt_mlt_result& res, double& unit)
{
    const int G = 2 * p.max_depth + 1;
    const int64_t area = (int64_t)W * H;
    const int spp = p.erpt_spp_x * p.erpt_spp_y, M = p.mutations;
    const int64_t n0 = area * spp;
    if (n0 > INT32_MAX) return set_err(c, RT_ERR_INVALID, "too many initial samples");
    const int seed_cap = (int)std::min<int64_t>(cap_slots, std::max(n0, area));
    if (int r = mlt_alloc(c, seed_cap, G, area)) return r;
    if (int r = ensure(c, c->mlt_seed, (size_t)std::max(n0, area) * sizeof(float4))) return r;
    if (int r = ensure(c, c->mlt_nch, (size_t)n0 * sizeof(int32_t))) return r;
    MltDev m = mlt_dev(c, seed_cap, G, W, H, p);
    const int grid = grid_for(c);
    InFlight flight(c, 4);
    float4* vals = (float4*)c->mlt_seed.p;
    /* e_d: one path through each pixel corner (computeEd) */
    for (int64_t base = 0; base < area; base += seed_cap) {
        const int n = (int)std::min<int64_t>(seed_cap, area - base);
        MLT_TRY(launch_mlt_fill(m, RT_MLT_STREAM_ED, (uint32_t)base, n, 1, grid, c->stream));
        MLT_TRY(mlt_trace(c, n, seed_cap, p.max_depth, p.rr_start, p.erpt_ed_termination, p.seed, 1));
        MLT_TRY(launch_mlt_values(c->ps, n, vals + base, 1, grid, c->stream));
    }
    std::vector<float4> hv((size_t)area);
    MLT_TRY(hipMemcpyAsync(hv.data(), vals, (size_t)area * sizeof(float4), hipMemcpyDeviceToHost, c->stream));
    MLT_TRY(hipStreamSynchronize(c->stream));
    double er = 0, eg = 0, eb = 0;
    for (const float4& v : hv) er += v.x, eg += v.y, eb += v.z;
    const float ed = (float)(std::sqrt(er * er + eg * eg + eb * eb) / ((double)area * M));
    res.ed = ed;
    res.seed_paths = area + n0;
    /* initial samples and their chain counts */
    int32_t* nch = (int32_t*)c->mlt_nch.p;
    for (int64_t base = 0; base < n0; base += seed_cap) {
        const int n = (int)std::min<int64_t>(seed_cap, n0 - base);
        MLT_TRY(launch_mlt_fill(m, RT_MLT_STREAM_SEED, (uint32_t)base, n, 2, grid, c->stream));
        MLT_TRY(mlt_trace(c, n, seed_cap, p.max_depth, p.rr_start, 0.f, p.seed, 0));
        MLT_TRY(launch_mlt_values(c->ps, n, vals + base, 1, grid, c->stream));
        MLT_TRY(launch_erpt_count(vals + base, n, (uint32_t)base, p.seed, ed, M, nch + base, grid, c->stream));
    }
    std::vector<int32_t> hn((size_t)n0);
    MLT_TRY(hipMemcpyAsync(hn.data(), nch, (size_t)n0 * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    MLT_TRY(hipStreamSynchronize(c->stream));
    int64_t C = 0;
    for (int32_t k : hn) {
        C += k;
        res.chain_seeds += k > 0;
    }
    res.chains = C;
    res.mutations = 0;
    if (C == 0) return RT_OK;
    if (C > INT32_MAX) return set_err(c, RT_ERR_INVALID, "too many chains");
    const int cap = (int)std::min<int64_t>(cap_slots, C);
    if (cap > seed_cap) {
        if (int r = mlt_alloc(c, cap, G, area)) return r;
    }
    const int stride = std::max(cap, seed_cap);
    m = mlt_dev(c, stride, G, W, H, p);
    m.scale = 1.f / (float)spp;
    /* a deposit's channel is at most e_d / spp */
    unit = mlt_unit((double)ed * m.scale * 1.01, (double)C * M);
    m.inv_unit = 1.0 / unit;
    /* chain -> initial sample, chains numbered in initial-sample order; all
     * of it uploaded before the batches are queued */
    std::vector<int32_t> src((size_t)C);
    for (int64_t i = 0, k = 0; i < n0; ++i)
        for (int32_t r = 0; r < hn[i]; ++r) src[k++] = (int32_t)i;
    if (int r = upload(c, c->mlt_src, src.data(), src.size() * sizeof(int32_t))) return r;
    for (int64_t c0 = 0; c0 < C; c0 += cap) {
        const int n = (int)std::min<int64_t>(cap, C - c0);
        m.chain0 = (uint32_t)c0;
        MLT_TRY(launch_erpt_start(m, (const int32_t*)c->mlt_src.p + c0, nch, vals, ed, n, grid, c->stream));
        for (int it = 0; it < M; ++it) {
            if (c->stop) return cancel(c);
            MLT_TRY(launch_mlt_propose(m, (uint32_t)it, n, grid, c->stream));
            MLT_TRY(mlt_trace(c, n, stride, p.max_depth, p.rr_start, 0.f, p.seed, 0));
            MLT_TRY(launch_mlt_accept(m, c->ps, (uint32_t)it, n, grid, c->stream));
            MLT_TRY(flight.unit_done());
        }
    }
    res.mutations = C * M;
    return RT_OK;
}

/* MetropolisFltRenderer::render, src/metropolisfltrenderer.cpp:60-320:
 * flt_passes passes; each rescales the two films by the summed importance,
 * runs `chains` chains per film with norm x importance acceptance, merges the
 * deposits into the films and the running variance, then filters.  The
 * result (the filtered mean of the films) is added to film_out (W * H * 3,
 * device memory). */
int run_filtered(rt_ctx* c, const rt_mlt_params& p, int W, int H, int64_t cap_slots, rt_mlt_result& res, float* film_out)
{
    const int G = 2 * p.max_depth + 1, P = p.flt_passes, F = p.flt_flags;
    const int64_t area = (int64_t)W * H;
    const int64_t per_film = std::max<int64_t>(1, (int64_t)p.mutations * area / P / 2);
    const int64_t C = p.chains > 0 ? p.chains : std::max<int64_t>(1, std::min<int64_t>(per_film / 16, (int64_t)1 << 22));
    const int64_t iters = std::max<int64_t>(1, (per_film + C / 2) / C);
    const int64_t nb = std::max<int64_t>(p.bootstrap, C);
    if (nb > INT32_MAX || (int64_t)P * 2 * C > INT32_MAX || iters > INT32_MAX)
        return set_err(c, RT_ERR_INVALID, "too many chains or mutations");
    /* both films' chains in one batch: chain ids (pass * 2 + film) * C + j */
    const int cap = (int)std::min<int64_t>(cap_slots, std::max(2 * C, nb));
    if (int r = mlt_alloc(c, cap, G, area)) return r;
    if (int r = ensure(c, c->mlt_seed, (size_t)nb * sizeof(float4) * 2)) return r;
    if (int r = ensure(c, c->mlt_acc, (size_t)area * 6 * sizeof(unsigned long long))) return r;
    /* images: films, means, m2 (3 each per film), sw (1 per film), sumimp,
     * imp, newsum, per-pixel scratch (1), newOut, noisy, bvar x2, meanvar,
     * varvar, filteredVar, filteredMean, varOfFiltered, filtered films x2,
     * filter scratch (RT_SYM_FILTER_TMP per pixel, last) */
    const size_t n3 = (size_t)area * 3;
    float* im = nullptr;
    const size_t total = n3 * 6 + (size_t)area * 6 + n3 * 11 + (size_t)area * RT_SYM_FILTER_TMP;
    if (int r = ensure(c, c->flt_img, total * 4)) return r;
    if (int r = ensure(c, c->flt_acc, (size_t)area * 8 * sizeof(unsigned long long))) return r;
    im = (float*)c->flt_img.p;
    float *film[2] = {im, im + n3}, *mean[2] = {im + 2 * n3, im + 3 * n3}, *m2[2] = {im + 4 * n3, im + 5 * n3};
    float* q = im + 6 * n3;
    float *sw[2] = {q, q + area}, *sumimp = q + 2 * area, *imp = q + 3 * area, *newsum = q + 4 * area,
          *pix = q + 5 * area;
    q += 6 * area;
    float *newOut = q, *noisy = q + n3, *bv[2] = {q + 2 * n3, q + 3 * n3}, *meanvar = q + 4 * n3, *varvar = q + 5 * n3,
          *fvar = q + 6 * n3, *fmean = q + 7 * n3, *vof = q + 8 * n3, *ff[2] = {q + 9 * n3, q + 10 * n3};
    float* scratch = q + 11 * n3;
    unsigned long long* aw = (unsigned long long*)c->flt_acc.p;  /* [film][pixel] */
    unsigned long long* aq = aw + 2 * area;                        /* [film][pixel][3] */
    unsigned long long* ax = (unsigned long long*)c->mlt_acc.p;
    const int grid = grid_for(c);
    InFlight flight(c, 4);
    MLT_TRY(hipMemsetAsync(im, 0, (size_t)(n3 * 6 + area * 3) * 4, c->stream));  /* films, stats, sumimp */
    float4* bval = (float4*)c->mlt_seed.p;
    float4* bw = bval + nb;
    MltDev m = mlt_dev(c, cap, G, W, H, p);
    m.imp = imp;
    m.newsum = newsum;
    m.acc_w = aw;
    m.acc_q = aq;
    m.film_chains = (uint32_t)C;
    std::vector<float> hbuf((size_t)area);
    std::vector<float4> hv((size_t)nb);
    int64_t mutations = 0;
    for (int pass = 0; pass < P; ++pass) {
        if (c->stop) return cancel(c);
        /* importance map */
        if (pass == 0) {
            std::fill(hbuf.begin(), hbuf.end(), 1.f);
            MLT_TRY(hipMemcpyAsync(imp, hbuf.data(), area * 4, hipMemcpyHostToDevice, c->stream));
        } else {
            MLT_TRY(launch_flt_importance(vof, fmean, imp, area, grid, c->stream));
            MLT_TRY(hipMemcpyAsync(hbuf.data(), imp, area * 4, hipMemcpyDeviceToHost, c->stream));
            MLT_TRY(hipStreamSynchronize(c->stream));
            double sum = 0;
            for (float v : hbuf) sum += v;
            MLT_TRY(launch_scale(imp, (float)(area / sum), area, grid, c->stream));
        }
        /* bootstrap of this pass (full-length large steps), weights norm x importance */
        m.step = (uint32_t)pass;
        for (int64_t base = 0; base < nb; base += cap) {
            const int n = (int)std::min<int64_t>(cap, nb - base);
            MLT_TRY(launch_mlt_fill(m, RT_MLT_STREAM_BOOTSTRAP, (uint32_t)base, n, 0, grid, c->stream));
            MLT_TRY(mlt_trace(c, n, cap, p.max_depth, p.rr_start, 0.f, p.seed, 0));
            MLT_TRY(launch_mlt_values(c->ps, n, bval + base, 1, grid, c->stream));
            MLT_TRY(launch_flt_boot_weight(m, bval + base, bw + base, n, grid, c->stream));
        }
        MLT_TRY(hipMemcpyAsync(hv.data(), bw, (size_t)nb * sizeof(float4), hipMemcpyDeviceToHost, c->stream));
        /* b: the bootstrap's mean norm, or the first film's norm x importance (:109-124) */
        float b;
        if (pass == 0) {
            std::vector<float4> hb((size_t)nb);
            MLT_TRY(hipMemcpyAsync(hb.data(), bval, (size_t)nb * sizeof(float4), hipMemcpyDeviceToHost, c->stream));
            MLT_TRY(hipStreamSynchronize(c->stream));
            double acc = 0;
            for (const float4& v : hb) acc += std::sqrt((double)v.x * v.x + (double)v.y * v.y + (double)v.z * v.z);
            b = (float)(acc / (double)nb);
            res.b = b;
        } else {
            MLT_TRY(launch_flt_bnorm(film[0], imp, pix, area, grid, c->stream));
            MLT_TRY(hipMemcpyAsync(hbuf.data(), pix, area * 4, hipMemcpyDeviceToHost, c->stream));
            MLT_TRY(hipStreamSynchronize(c->stream));
            double acc = 0;
            for (float v : hbuf) acc += (double)v;
            b = (float)(acc / (double)area);
        }
        std::vector<double> cdf((size_t)nb);
        double tot = 0;
        for (int64_t i = 0; i < nb; ++i) cdf[i] = tot += (double)hv[i].w;
        MLT_TRY(launch_flt_rescale(sumimp, imp, newsum, film[0], film[1], m2[0], m2[1], area, grid, c->stream));
        if (tot > 0) {
            if (int r = upload(c, c->mlt_cdf, cdf.data(), cdf.size() * sizeof(double))) return r;
            /* fixed point: a deposit is at most (1 / min newsum) * scale per channel */
            MLT_TRY(hipMemcpyAsync(hbuf.data(), newsum, area * 4, hipMemcpyDeviceToHost, c->stream));
            MLT_TRY(hipStreamSynchronize(c->stream));
            float mn = INFINITY;
            for (float v : hbuf) mn = std::min(mn, v);
            /* the reference scales each deposit by b * passes / mutations-per-pixel
             * over (mutations-per-pixel * area / passes / 2) mutations per film
             * (src/metropolisfltrenderer.cpp:139,158-165): b * area / 2 per film
             * and pass.  The chains here run C * iters mutations per film, so
             * the scale that keeps that total is b * area / (2 C iters). */
            m.scale = (float)((double)b * (double)area / (2.0 * (double)C * (double)iters));
            const double wmax = 1.0 / mn, deps = 2.0 * (double)C * (double)iters;
            const double ux = mlt_unit(wmax * m.scale * 1.01, deps), uw = mlt_unit(wmax * 1.01, deps),
                         uq = mlt_unit(wmax * (double)m.scale * m.scale * 1.01, deps);
            m.inv_unit = 1.0 / ux;
            m.inv_unit_w = 1.0 / uw;
            m.inv_unit_q = 1.0 / uq;
            MLT_TRY(hipMemsetAsync(ax, 0, n3 * 2 * 8, c->stream));
            MLT_TRY(hipMemsetAsync(aw, 0, (size_t)area * 8 * 8, c->stream));
            for (int64_t c0 = 0; c0 < 2 * C; c0 += cap) {
                const int nn = (int)std::min<int64_t>(cap, 2 * C - c0);
                m.chain0 = (uint32_t)((int64_t)pass * 2 * C + c0);
                MLT_TRY(launch_mlt_start(m, (const double*)c->mlt_cdf.p, nb, tot, bval, nn, grid, c->stream));
                for (int64_t it = 0; it < iters; ++it) {
                    if (c->stop) return cancel(c);
                    MLT_TRY(launch_mlt_propose(m, (uint32_t)it, nn, grid, c->stream));
                    MLT_TRY(mlt_trace(c, nn, cap, p.max_depth, p.rr_start, p.termination_prob, p.seed, 0));
                    MLT_TRY(launch_mlt_accept(m, c->ps, (uint32_t)it, nn, grid, c->stream));
                    MLT_TRY(flight.unit_done());
                }
            }
            for (int n = 0; n < 2; ++n)
                MLT_TRY(launch_flt_merge(aw + n * area, ax + n * n3, aq + n * n3, uw, ux, uq, film[n], mean[n], m2[n],
                                         sw[n], area, grid, c->stream));
            mutations += 2 * C * iters;
        }
        /* filtering (:252-300) */
        for (int n = 0; n < 2; ++n) MLT_TRY(launch_flt_bvar(m2[n], sw[n], bv[n], area, grid, c->stream));
        MLT_TRY(launch_var2(film[0], film[1], newOut, noisy, (int64_t)n3, grid, c->stream));
        MLT_TRY(launch_var2(bv[0], bv[1], meanvar, varvar, (int64_t)n3, grid, c->stream));
        if ((F & RT_FLT_NO_FILTER) || (F & RT_FLT_NO_VARIANCE_FILTER))
            MLT_TRY(hipMemcpyAsync(fvar, noisy, n3 * 4, hipMemcpyDeviceToDevice, c->stream));
        else if (F & RT_FLT_SIMPLE_VARIANCE_FILTER)
            MLT_TRY(launch_sep3(noisy, scratch, fvar, W, H, grid, c->stream));
        else
            MLT_TRY(launch_sym_filter(noisy, meanvar, varvar, W, H, 1, 2, scratch, fvar, grid, c->stream));
        if (F & RT_FLT_NO_FILTER) {
            MLT_TRY(hipMemcpyAsync(vof, fvar, n3 * 4, hipMemcpyDeviceToDevice, c->stream));
            MLT_TRY(hipMemcpyAsync(fmean, newOut, n3 * 4, hipMemcpyDeviceToDevice, c->stream));
        } else if (F & RT_FLT_NO_CROSSFILTERING) {
            MLT_TRY(hipMemcpyAsync(vof, fvar, n3 * 4, hipMemcpyDeviceToDevice, c->stream));
            MLT_TRY(launch_sym_filter(newOut, newOut, fvar, W, H, 7, 10, scratch, fmean, grid, c->stream));
        } else {
            MLT_TRY(launch_sym_filter(film[0], film[1], fvar, W, H, 7, 10, scratch, ff[0], grid, c->stream));
            MLT_TRY(launch_sym_filter(film[1], film[0], fvar, W, H, 7, 10, scratch, ff[1], grid, c->stream));
            MLT_TRY(launch_var2(ff[0], ff[1], fmean, vof, (int64_t)n3, grid, c->stream));
        }
        MLT_TRY(hipMemcpyAsync(sumimp, newsum, area * 4, hipMemcpyDeviceToDevice, c->stream));
    }
    MLT_TRY(launch_add(film_out, fmean, (int64_t)n3, grid, c->stream));
    res.chains = 2 * C * P;
    res.seed_paths = nb * P;
    res.mutations = mutations;
    return RT_OK;
}

int render_mlt_impl(rt_ctx* c, const rt_camera* cam, const rt_mlt_params* p, float* film, bool device_film,
                    rt_mlt_result* out)
{
    if (!c || !cam || !p || !film) return RT_ERR_INVALID;
    if (int r = check_fault(c)) return r;
    if (!c->has_scene) return set_err(c, RT_ERR_NO_SCENE, "no scene uploaded");
    const int W = cam->width, H = cam->height;
    if (W <= 0 || H <= 0 || (int64_t)W * H > (1 << 26)) return set_err(c, RT_ERR_INVALID, "bad camera");
    if (p->kind != RT_MLT_METROPOLIS && p->kind != RT_MLT_ERPT && p->kind != RT_MLT_FILTERED)
        return set_err(c, RT_ERR_INVALID, "unknown kind");
    if (p->kind == RT_MLT_FILTERED && (p->flt_passes < 1 || p->flt_passes > 64 || p->flt_flags < 0 || p->flt_flags > 15))
        return set_err(c, RT_ERR_INVALID, "bad filtered Metropolis parameters");
    if (p->max_depth < 1 || p->max_depth > 64 || p->rr_start < 0 || p->mutations < 1 ||
        !(p->termination_prob >= 0.f && p->termination_prob < 1.f) ||
        !(p->large_step_prob >= 0.f && p->large_step_prob <= 1.f) || p->bootstrap < 0 || p->chains < 0 ||
        (p->kind == RT_MLT_ERPT && (p->erpt_spp_x < 1 || p->erpt_spp_y < 1 ||
                                    !(p->erpt_ed_termination >= 0.f && p->erpt_ed_termination < 1.f))))
        return set_err(c, RT_ERR_INVALID, "bad Metropolis parameters");
    hipSetDevice(c->device);
    /* chains per batch: 4M path slots unless configured */
    const int64_t cap = c->cfg.path_slots > 0 ? c->cfg.path_slots : (int64_t)1 << 22;
    const int64_t area = (int64_t)W * H;
    rt_view v{*cam, 1, 1};
    DevView dv = make_view(v, nullptr);
    rt_mlt_result res{};
    HIP_TRY(c, hipEventRecord(c->ev_a, c->stream));
    int r;
    if ((r = alloc_path_state(c, 256)) || (r = ensure(c, c->mlt_view, sizeof(DevView))) ||
        (r = ensure(c, c->mlt_cnt, sizeof(unsigned long long))) ||
        (r = ensure(c, c->mlt_acc, (size_t)area * 3 * sizeof(unsigned long long))))
        return r;
    HIP_TRY(c, hipMemcpyAsync(c->mlt_view.p, &dv, sizeof(dv), hipMemcpyHostToDevice, c->stream));
    if ((r = (launch_reset_counters(c->ps, c->stream) == hipSuccess ? 0 : RT_ERR_HIP))) return set_err(c, r, "reset");
    MLT_TRY(hipMemsetAsync(c->mlt_acc.p, 0, (size_t)area * 3 * sizeof(unsigned long long), c->stream));
    MLT_TRY(hipMemsetAsync(c->mlt_cnt.p, 0, sizeof(unsigned long long), c->stream));
    double unit = 0;
    float* dfilm = film;
    if (!device_film) {
        if ((r = ensure(c, c->films, (size_t)area * 3 * 4))) return r;
        dfilm = (float*)c->films.p;
        HIP_TRY(c, hipMemcpyAsync(dfilm, film, (size_t)area * 3 * 4, hipMemcpyHostToDevice, c->stream));
    }
    if (p->kind == RT_MLT_FILTERED) {
        if ((r = run_filtered(c, *p, W, H, cap, res, dfilm))) return r;
    } else {
        r = p->kind == RT_MLT_METROPOLIS ? run_metropolis(c, *p, W, H, cap, res, unit)
                                         : run_erpt(c, *p, W, H, cap, res, unit);
        if (r) return r;
    }
    if (unit > 0)
        MLT_TRY(launch_mlt_resolve((const unsigned long long*)c->mlt_acc.p, dfilm, area * 3, unit, grid_for(c),
                                   c->stream));
    HIP_TRY(c, hipEventRecord(c->ev_b, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (!device_film) HIP_TRY(c, hipMemcpy(film, dfilm, (size_t)area * 3 * 4, hipMemcpyDeviceToHost));
    unsigned long long acc_n = 0;
    HIP_TRY(c, hipMemcpy(&acc_n, c->mlt_cnt.p, sizeof(acc_n), hipMemcpyDeviceToHost));
    DevCounters hc;
    HIP_TRY(c, hipMemcpy(&hc, c->ps.cnt, sizeof(hc), hipMemcpyDeviceToHost));
    float ms = 0;
    hipEventElapsedTime(&ms, c->ev_a, c->ev_b);
    res.accepted = (int64_t)acc_n;
    res.rays = (int64_t)(hc.camera_rays + hc.extension_rays + hc.shadow_rays);
    res.render_ms = ms;
    c->stats.render_ms = ms;
    c->stats.camera_rays = (int64_t)hc.camera_rays;
    c->stats.extension_rays = (int64_t)hc.extension_rays;
    c->stats.shadow_rays = (int64_t)hc.shadow_rays;
    c->stats.extend_ms = 0;
    c->stats.shadow_ms = 0;
    c->stats.extend_launches = 0;
    c->stats.finish_ms = 0;
    c->stats.finish_rays = (int64_t)hc.finish_rays;
    c->stats.finish_shadow_rays = (int64_t)hc.finish_shadow_rays;
    c->stats.finish_bounce = hc.finished;
    if (out) *out = res;
    return RT_OK;
}

/* Radiance of host-given primary-sample vectors ([n][G][4]). */
int integrate_samples_impl(rt_ctx* c, const rt_camera* cam, int max_depth, int rr_start, float tp,
                           const float* samples, int64_t n, float* rgb)
{
    if (!c || !cam || (!samples && n > 0) || (!rgb && n > 0) || n < 0) return RT_ERR_INVALID;
    if (int r = check_fault(c)) return r;
    if (!c->has_scene) return set_err(c, RT_ERR_NO_SCENE, "no scene uploaded");
    if (cam->width <= 0 || cam->height <= 0 || max_depth < 1 || max_depth > 64 || rr_start < 0 ||
        !(tp >= 0.f && tp < 1.f))
        return set_err(c, RT_ERR_INVALID, "bad parameters");
    if (n == 0) return RT_OK;
    hipSetDevice(c->device);
    const int G = 2 * max_depth + 1;
    const int cap = (int)std::min<int64_t>(n, c->cfg.path_slots > 0 ? c->cfg.path_slots : (int64_t)1 << 22);
    int r;
    if ((r = mlt_alloc(c, cap, G, 1))) return r;
    rt_view v{*cam, 1, 1};
    DevView dv = make_view(v, nullptr);
    HIP_TRY(c, hipMemcpyAsync(c->mlt_view.p, &dv, sizeof(dv), hipMemcpyHostToDevice, c->stream));
    if (launch_reset_counters(c->ps, c->stream) != hipSuccess) return set_err(c, RT_ERR_HIP, "reset");
    std::vector<float> soa((size_t)G * cap * 4), out((size_t)cap * 4);
    for (int64_t base = 0; base < n; base += cap) {
        const int m = (int)std::min<int64_t>(cap, n - base);
        for (int j = 0; j < m; ++j)
            for (int g = 0; g < G; ++g)
                std::memcpy(&soa[((size_t)g * cap + j) * 4], samples + ((size_t)(base + j) * G + g) * 4, 16);
        HIP_TRY(c, hipMemcpyAsync(c->mlt_prop.p, soa.data(), soa.size() * 4, hipMemcpyHostToDevice, c->stream));
        MLT_TRY(mlt_trace(c, m, cap, max_depth, rr_start, tp, 0, 0));
        MLT_TRY(launch_mlt_values(c->ps, m, (float4*)c->mlt_val.p, 0, grid_for(c), c->stream));
        HIP_TRY(c, hipMemcpyAsync(out.data(), c->mlt_val.p, (size_t)m * 16, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        for (int j = 0; j < m; ++j)
            for (int k = 0; k < 3; ++k) rgb[(base + j) * 3 + k] = out[(size_t)j * 4 + k];
    }
    return RT_OK;
}

}  // namespace

/* UniDiPathTracingIntegrator::integrate(const Ray&, ...) for a batch of rays
 * in device memory: slots of at most the configured path-slot count, each
 * chunk traced through the bounce pipeline of rt_render. */
static int integrate_rays_impl(rt_ctx* c, const float* o, const float* d, int64_t n, const rt_params* p,
                               uint32_t key0, float* rgb)
{
    if (!c->has_scene) return set_err(c, RT_ERR_NO_SCENE, "no scene uploaded");
    if (p->mode != RT_MODE_COMPAT && p->mode != RT_MODE_UNBIASED)
        return set_err(c, RT_ERR_INVALID, "rt_integrate_rays: mode must be RT_MODE_COMPAT or RT_MODE_UNBIASED");
    if (p->max_depth < 1 || p->max_depth > 64 || p->rr_prob < 0.f || p->rr_prob >= 1.f || p->rr_start < 0)
        return set_err(c, RT_ERR_INVALID, "bad render parameters");
    if (n > (int64_t)UINT32_MAX) return set_err(c, RT_ERR_INVALID, "too many rays");
    if (n == 0) return RT_OK;
    hipSetDevice(c->device);
    const int64_t cap = std::min<int64_t>(n, c->cfg.path_slots > 0 ? c->cfg.path_slots : (int64_t)1 << 25);
    if (int r = alloc_path_state(c, (size_t)cap)) return r;
    const int grid = grid_for(c);
    if (int r = ensure_stack_overflow(c, grid)) return r;
    PtParams pp{p->mode, p->max_depth, p->rr_start, p->rr_prob, p->seed, 0};
    HIP_TRY(c, launch_reset_counters(c->ps, c->stream));
    HIP_TRY(c, hipEventRecord(c->ev_a, c->stream));
    InFlight flight(c, 2);
    for (int64_t base = 0; base < n; base += cap) {
        if (c->stop) return cancel(c);
        const int m = (int)std::min<int64_t>(cap, n - base);
        HIP_TRY(c, launch_generate_rays(c->ps, (const float4*)o + base, (const float4*)d + base, m, pp,
                                        key0 + (uint32_t)base, grid, c->stream));
        HIP_TRY(c, trace_bounces(c, pp, grid));
        HIP_TRY(c, launch_ray_values(c->ps, m, rgb + 3 * base, grid, c->stream));
        HIP_TRY(c, flight.unit_done());
    }
    HIP_TRY(c, hipEventRecord(c->ev_b, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    float ms = 0;
    hipEventElapsedTime(&ms, c->ev_a, c->ev_b);
    c->stats.render_ms = ms;
    DevCounters hc;
    HIP_TRY(c, hipMemcpy(&hc, c->ps.cnt, sizeof(hc), hipMemcpyDeviceToHost));
    c->stats.camera_rays = (int64_t)hc.camera_rays;
    c->stats.extension_rays = (int64_t)hc.extension_rays;
    c->stats.shadow_rays = (int64_t)hc.shadow_rays;
    c->stats.finish_rays = (int64_t)hc.finish_rays;
    c->stats.finish_shadow_rays = (int64_t)hc.finish_shadow_rays;
    c->stats.finish_bounce = hc.finished;
    return RT_OK;
}

int rt_integrate_rays_device(rt_ctx* c, const float* o, const float* d, int64_t n, const rt_params* p, uint32_t key0,
                             float* rgb)
{
    if (!c || !p || n < 0 || (n && (!o || !d || !rgb))) return RT_ERR_INVALID;
    if (int r = check_fault(c)) return r;
    return integrate_rays_impl(c, o, d, n, p, key0, rgb);
}

int rt_integrate_rays(rt_ctx* c, const float* o, const float* d, int64_t n, const rt_params* p, uint32_t key0,
                      float* rgb)
{
    if (!c || !p || n < 0 || (n && (!o || !d || !rgb))) return RT_ERR_INVALID;
    if (int r = check_fault(c)) return r;
    if (!c->has_scene) return set_err(c, RT_ERR_NO_SCENE, "no scene uploaded");
    if (n == 0) return RT_OK;
    hipSetDevice(c->device);
    int r;
    if ((r = ensure(c, c->rb_o, (size_t)n * 16)) || (r = ensure(c, c->rb_d, (size_t)n * 16)) ||
        (r = ensure(c, c->rb_out0, (size_t)n * 12)))
        return r;
    HIP_TRY(c, hipMemcpyAsync(c->rb_o.p, o, (size_t)n * 16, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->rb_d.p, d, (size_t)n * 16, hipMemcpyHostToDevice, c->stream));
    if ((r = integrate_rays_impl(c, (const float*)c->rb_o.p, (const float*)c->rb_d.p, n, p, key0,
                                 (float*)c->rb_out0.p)))
        return r;
    HIP_TRY(c, hipMemcpy(rgb, c->rb_out0.p, (size_t)n * 12, hipMemcpyDeviceToHost));
    return RT_OK;
}

int rt_render(rt_ctx* c, const rt_view* views, int32_t nv, const rt_params* p, float* const* films)
{
    return render_impl(c, views, nv, p, films, false);
}

int rt_render_device(rt_ctx* c, const rt_view* views, int32_t nv, const rt_params* p, float* const* films)
{
    return render_impl(c, views, nv, p, films, true);
}

int rt_integrate_samples(rt_ctx* c, const rt_camera* camera, int32_t max_depth, int32_t rr_start,
                         float termination_prob, const float* samples, int64_t n, float* rgb)
{
    return integrate_samples_impl(c, camera, max_depth, rr_start, termination_prob, samples, n, rgb);
}

int rt_render_mlt(rt_ctx* c, const rt_camera* camera, const rt_mlt_params* p, float* film, rt_mlt_result* res)
{
    return render_mlt_impl(c, camera, p, film, false, res);
}

int rt_render_mlt_device(rt_ctx* c, const rt_camera* camera, const rt_mlt_params* p, float* film,
                         rt_mlt_result* res)
{
    return render_mlt_impl(c, camera, p, film, true, res);
}

int rt_symmetric_filter(rt_ctx* c, const float* image, const float* guide, const float* pixvar, int32_t width,
                        int32_t height, int32_t patch, int32_t radius, float* out)
{
    if (!c || !image || !guide || !pixvar || !out) return RT_ERR_INVALID;
    if (int r = check_fault(c)) return r;
    if (width <= 0 || height <= 0 || (int64_t)width * height > (1 << 26) || patch < 1 || patch % 2 == 0 ||
        patch > 63 || radius < 0 || radius > 63)
        return set_err(c, RT_ERR_INVALID, "bad filter size");
    hipSetDevice(c->device);
    const size_t n3 = (size_t)width * height * 3, n = (size_t)width * height;
    if (int r = ensure(c, c->flt_img, (n3 * 4 + n * RT_SYM_FILTER_TMP) * 4)) return r;
    float* d = (float*)c->flt_img.p;
    HIP_TRY(c, hipMemcpyAsync(d, image, n3 * 4, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(d + n3, guide, n3 * 4, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(d + 2 * n3, pixvar, n3 * 4, hipMemcpyHostToDevice, c->stream));
    MLT_TRY(launch_sym_filter(d, d + n3, d + 2 * n3, width, height, patch, radius, d + 4 * n3, d + 3 * n3,
                              grid_for(c), c->stream));
    HIP_TRY(c, hipMemcpyAsync(out, d + 3 * n3, n3 * 4, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return RT_OK;
}

}  // extern "C"
