/*
 * gpu_integrator.h — C++ host interface over librt_hip.so, shaped like the
 * reference renderer's own classes:
 *
 *   SceneBuilder                  the objects of a scene file (materials,
 *                                 quads, triangles, meshes, spheres, lights)
 *                                 in the reference's root-list order
 *                                 (src/sceneast.cpp:673-695)
 *   Camera                        src/camera.h:34-101 (eye, look-at, up,
 *                                 vertical fov in radians, resolution)
 *   Integrator                    the integrator interface, batch form
 *                                 (src/integrator.h:32-37)
 *   GpuUniDiPathTracingIntegrator UniDiPathTracingIntegrator(terminationProb)
 *                                 (src/unidipathtracingintegrator.h:30) on the GPU:
 *                                 films, and integrate(ray) per ray
 *                                 (src/unidipathtracingintegrator.cpp:34-46)
 *   PerPixelRenderer              PerPixelRenderer::render (src/perpixelrenderer.cpp:43-101)
 *                                 with pt-x-samples / pt-y-samples / seed
 *   SceneFile                     a scene file read by the library's front end
 *                                 (SceneGrammar::parse + resolveVars + buildScene,
 *                                 src/main.cpp:150-182; --dump-bvh)
 *   MetropolisRenderer            MetropolisRenderer::render
 *                                 (src/metropolisrenderer.cpp:56-212) with the
 *                                 met-* options, chains side by side on the GPU
 *   EnergyRedistributionRenderer  EnergyRedistributionRenderer::render
 *                                 (src/energyredistributionrenderer.cpp:38-147)
 *                                 with the erpt-* options
 *
 * Header-only; link with -lrt_hip.  Errors throw RtError (C++ callers) while
 * the library itself only returns status codes.
 */
#ifndef RENDERER_AMD_GPU_INTEGRATOR_H
#define RENDERER_AMD_GPU_INTEGRATOR_H

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/rt_abi.h"
#include "../../include/rt_mlt.h"
#include "../../include/rt_scene_file.h"

namespace rtgpu {

struct RtError : std::runtime_error {
    int status;
    RtError(int s, const std::string& what) : std::runtime_error(what), status(s) {}
};

inline void check(rt_ctx* ctx, int status, const char* call)
{
    if (status != RT_OK)
        throw RtError(status, std::string(call) + ": " + rt_error_string(status) + (ctx ? std::string(" (") + rt_last_error(ctx) + ")" : ""));
}

struct Vec3 {
    float x, y, z;
};

/* src/ray.h:36-111: origin, direction, parameter range [from, to] */
struct Ray {
    Vec3 o, d;
    float from = 0.001f, to = __builtin_inff();
};

/* Scene objects in file order; primitive IDs follow the reference's root
 * list: lights first, then intersectables in insertion order. */
class SceneBuilder {
public:
    int diffuse(Vec3 rgb) { return material(RT_MAT_DIFFUSE, rgb, Vec3{0, 0, 0}, 0.f); }
    int mirror() { return material(RT_MAT_MIRROR, Vec3{0, 0, 0}, Vec3{0, 0, 0}, 0.f); }
    int phong(Vec3 kd, Vec3 ks, float n) { return material(RT_MAT_PHONG, kd, ks, n); }
    int blinn(Vec3 kd, Vec3 ks, float n) { return material(RT_MAT_BLINN, kd, ks, n); }
    int refractive(float ior) { return material(RT_MAT_REFRACTIVE, Vec3{0, 0, 0}, Vec3{0, 0, 0}, ior); }

    void areaLight(Vec3 o, Vec3 u, Vec3 v, Vec3 intensity) { light(RT_LIGHT_AREA, o, u, v, intensity); }
    void pointLight(Vec3 p, Vec3 intensity) { light(RT_LIGHT_POINT, p, Vec3{0, 0, 0}, Vec3{0, 0, 0}, intensity); }
    /* src/conelight.h: opening angle in radians */
    void coneLight(Vec3 p, Vec3 direction, float angle, Vec3 intensity)
    {
        light(RT_LIGHT_CONE, p, direction, Vec3{angle, 0, 0}, intensity);
    }
    /* latitude-longitude RGB image, row 0 at +y */
    void environmentMap(int w, int h, const std::vector<float>& rgb, Vec3 coefficient)
    {
        env_w_ = w;
        env_h_ = h;
        env_ = rgb;
        light(RT_LIGHT_ENV, Vec3{0, 0, 0}, Vec3{0, 0, 0}, Vec3{0, 0, 0}, coefficient);
    }

    /* src/quad.h:32: triangles (p1,p2,p4), (p4,p2,p3), one primitive ID */
    void quad(Vec3 p1, Vec3 p2, Vec3 p3, Vec3 p4, int mat)
    {
        int id = nextId();
        int b = vertex(p1), c = vertex(p2), d = vertex(p3), e = vertex(p4);
        tri(b, c, e, mat, id);
        tri(e, c, d, mat, id);
    }
    void triangle(Vec3 p1, Vec3 p2, Vec3 p3, int mat) { tri(vertex(p1), vertex(p2), vertex(p3), mat, nextId()); }
    void sphere(Vec3 c, float r, int mat)
    {
        spheres_.insert(spheres_.end(), {c.x, c.y, c.z, r});
        sphere_mat_.push_back(mat);
        sphere_id_.push_back(nextId());
    }
    /* indexed mesh: every triangle is its own primitive (as an OBJ import) */
    void mesh(const std::vector<float>& positions, const std::vector<int32_t>& faces, int mat)
    {
        int base = (int)(pos_.size() / 3);
        pos_.insert(pos_.end(), positions.begin(), positions.end());
        for (size_t i = 0; i + 2 < faces.size(); i += 3) tri(base + faces[i], base + faces[i + 1], base + faces[i + 2], mat, nextId());
    }

    /* Pointer-free view of the arrays (valid while the builder lives and is
     * not modified). */
    rt_scene_desc desc() const
    {
        rt_scene_desc d{};
        d.n_vertices = (int32_t)(pos_.size() / 3);
        d.positions = pos_.data();
        d.normals = nullptr;
        d.n_triangles = (int32_t)tri_mat_.size();
        d.tri_vertices = tri_v_.data();
        d.tri_material = tri_mat_.data();
        d.tri_prim_id = tri_id_.data();
        d.n_spheres = (int32_t)sphere_mat_.size();
        d.spheres = spheres_.data();
        d.sphere_material = sphere_mat_.data();
        d.sphere_prim_id = sphere_id_.data();
        d.n_materials = (int32_t)mats_.size();
        d.materials = mats_.data();
        d.n_lights = (int32_t)lights_.size();
        d.lights = lights_.data();
        d.env_width = env_w_;
        d.env_height = env_h_;
        d.env_rgb = env_.empty() ? nullptr : env_.data();
        return d;
    }

private:
    int material(int type, Vec3 kd, Vec3 ks, float n)
    {
        rt_material m{};
        m.type = type;
        m.kd[0] = kd.x; m.kd[1] = kd.y; m.kd[2] = kd.z;
        m.ks[0] = ks.x; m.ks[1] = ks.y; m.ks[2] = ks.z;
        m.exponent = n;
        mats_.push_back(m);
        return (int)mats_.size() - 1;
    }
    void light(int type, Vec3 p, Vec3 u, Vec3 v, Vec3 i)
    {
        if (n_objects_ > 0) throw std::logic_error("lights precede intersectables in the root list");
        rt_light l{};
        l.type = type;
        l.p[0] = p.x; l.p[1] = p.y; l.p[2] = p.z;
        l.u[0] = u.x; l.u[1] = u.y; l.u[2] = u.z;
        l.v[0] = v.x; l.v[1] = v.y; l.v[2] = v.z;
        l.intensity[0] = i.x; l.intensity[1] = i.y; l.intensity[2] = i.z;
        lights_.push_back(l);
    }
    int vertex(Vec3 p)
    {
        pos_.insert(pos_.end(), {p.x, p.y, p.z});
        return (int)(pos_.size() / 3) - 1;
    }
    void tri(int a, int b, int c, int mat, int id)
    {
        tri_v_.insert(tri_v_.end(), {a, b, c});
        tri_mat_.push_back(mat);
        tri_id_.push_back(id);
    }
    int nextId() { return (int)lights_.size() + n_objects_++; }

    std::vector<float> pos_, spheres_, env_;
    std::vector<int32_t> tri_v_, tri_mat_, tri_id_, sphere_mat_, sphere_id_;
    std::vector<rt_material> mats_;
    std::vector<rt_light> lights_;
    int n_objects_ = 0, env_w_ = 0, env_h_ = 0;
};

struct Camera {
    Vec3 cop, lookAt, up;
    float fov;
    int width, height;
    rt_camera c() const
    {
        rt_camera r{};
        r.eye[0] = cop.x; r.eye[1] = cop.y; r.eye[2] = cop.z;
        r.look_at[0] = lookAt.x; r.look_at[1] = lookAt.y; r.look_at[2] = lookAt.z;
        r.up[0] = up.x; r.up[1] = up.y; r.up[2] = up.z;
        r.fov = fov;
        r.width = width;
        r.height = height;
        return r;
    }
};

/* A scene file in the reference's scene language.  scene() evaluates it; the
 * description stays valid while this object lives. */
class SceneFile {
public:
    explicit SceneFile(const std::string& path)
    {
        int st = rt_scene_file_load(path.c_str(), &f_);
        if (st != RT_OK) throw RtError(st, std::string("rt_scene_file_load: ") + rt_scene_file_last_error());
    }
    ~SceneFile() { rt_scene_file_free(f_); }
    SceneFile(const SceneFile&) = delete;
    SceneFile& operator=(const SceneFile&) = delete;

    const rt_scene_desc& scene()
    {
        int st = rt_scene_file_scene(f_, &desc_, &cam_);
        if (st != RT_OK) throw RtError(st, std::string("rt_scene_file_scene: ") + rt_scene_file_last_error());
        return desc_;
    }
    /* the "camera" variable (call scene() first) */
    Camera camera() const
    {
        return Camera{{cam_.eye[0], cam_.eye[1], cam_.eye[2]},
                      {cam_.look_at[0], cam_.look_at[1], cam_.look_at[2]},
                      {cam_.up[0], cam_.up[1], cam_.up[2]},
                      cam_.fov, cam_.width, cam_.height};
    }
    void dumpBvh(const std::string& path) const
    {
        int st = rt_scene_file_dump_bvh(f_, path.c_str());
        if (st != RT_OK) throw RtError(st, std::string("rt_scene_file_dump_bvh: ") + rt_scene_file_last_error());
    }
    std::string messages() const { return rt_scene_file_messages(f_); }

private:
    rt_scene_file* f_ = nullptr;
    rt_scene_desc desc_{};
    rt_camera cam_{};
};

/* Batch form of the reference's Integrator: one call renders a film. */
class Integrator {
public:
    virtual ~Integrator() {}
    virtual void render(const Camera& cam, int sppX, int sppY, unsigned long seed, float* film) = 0;
};

class GpuUniDiPathTracingIntegrator : public Integrator {
public:
    explicit GpuUniDiPathTracingIntegrator(float terminationProb = 0.5f, int mode = RT_MODE_COMPAT, int maxDepth = 5,
                                           int device = 0)
        : p_(terminationProb), mode_(mode), depth_(maxDepth)
    {
        check(nullptr, rt_create(device, &ctx_), "rt_create");
    }
    ~GpuUniDiPathTracingIntegrator() override { rt_destroy(ctx_); }
    GpuUniDiPathTracingIntegrator(const GpuUniDiPathTracingIntegrator&) = delete;
    GpuUniDiPathTracingIntegrator& operator=(const GpuUniDiPathTracingIntegrator&) = delete;

    void setScene(const SceneBuilder& scene)
    {
        rt_scene_desc d = scene.desc();
        check(ctx_, rt_scene_upload(ctx_, &d), "rt_scene_upload");
    }
    void setScene(const rt_scene_desc& d) { check(ctx_, rt_scene_upload(ctx_, &d), "rt_scene_upload"); }
    /* film: height * width * 3 RGB floats, accumulated as film += sample / spp */
    void render(const Camera& cam, int sppX, int sppY, unsigned long seed, float* film) override
    {
        rt_view v{};
        v.camera = cam.c();
        v.spp_x = sppX;
        v.spp_y = sppY;
        rt_params p{};
        p.mode = mode_;
        p.max_depth = depth_;
        p.rr_start = 3;
        p.rr_prob = p_;
        p.seed = (uint32_t)seed;
        float* films[1] = {film};
        check(ctx_, rt_render(ctx_, &v, 1, &p, films), "rt_render");
    }
    /* Integrator::integrate(const Ray&, scene, lights, rng): the radiance
     * along each ray (RGB, 3 floats per ray).  Ray i draws the random numbers
     * keyed by key + i, so a caller holding the reference's gsl_rng passes
     * gsl_rng_get(rng) as the key. */
    std::vector<float> integrate(const std::vector<Ray>& rays, uint32_t key, unsigned long seed)
    {
        std::vector<float> o(rays.size() * 4), d(rays.size() * 4), rgb(rays.size() * 3);
        for (size_t i = 0; i < rays.size(); ++i) {
            const Ray& r = rays[i];
            o[4 * i] = r.o.x, o[4 * i + 1] = r.o.y, o[4 * i + 2] = r.o.z, o[4 * i + 3] = r.from;
            d[4 * i] = r.d.x, d[4 * i + 1] = r.d.y, d[4 * i + 2] = r.d.z, d[4 * i + 3] = r.to;
        }
        rt_params p{};
        p.mode = mode_;
        p.max_depth = depth_;
        p.rr_start = 3;
        p.rr_prob = p_;
        p.seed = (uint32_t)seed;
        check(ctx_, rt_integrate_rays(ctx_, o.data(), d.data(), (int64_t)rays.size(), &p, key, rgb.data()),
              "rt_integrate_rays");
        return rgb;
    }
    Vec3 integrate(const Ray& ray, uint32_t key, unsigned long seed)
    {
        std::vector<float> c = integrate(std::vector<Ray>{ray}, key, seed);
        return Vec3{c[0], c[1], c[2]};
    }
    void stop() { rt_request_stop(ctx_, 1); }
    rt_ctx* context() { return ctx_; }

private:
    rt_ctx* ctx_ = nullptr;
    float p_;
    int mode_, depth_;
};

/* PerPixelRenderer with the reference's options: pt-x-samples, pt-y-samples, seed. */
class PerPixelRenderer {
public:
    PerPixelRenderer(Integrator& integrator, int xSamples = 4, int ySamples = 4, unsigned long seed = 0)
        : integ_(integrator), sx_(xSamples), sy_(ySamples), seed_(seed)
    {
    }
    std::vector<float> render(const Camera& cam)
    {
        std::vector<float> film((size_t)cam.width * cam.height * 3, 0.f);
        integ_.render(cam, sx_, sy_, seed_, film.data());
        return film;
    }

private:
    Integrator& integ_;
    int sx_, sy_;
    unsigned long seed_;
};

/* The Markov-chain renderers share a context and the uploaded scene. */
class MarkovChainRenderer {
public:
    explicit MarkovChainRenderer(int device = 0) { check(nullptr, rt_create(device, &ctx_), "rt_create"); }
    virtual ~MarkovChainRenderer() { rt_destroy(ctx_); }
    MarkovChainRenderer(const MarkovChainRenderer&) = delete;
    MarkovChainRenderer& operator=(const MarkovChainRenderer&) = delete;
    void setScene(const SceneBuilder& scene)
    {
        rt_scene_desc d = scene.desc();
        check(ctx_, rt_scene_upload(ctx_, &d), "rt_scene_upload");
    }
    void setScene(const rt_scene_desc& d) { check(ctx_, rt_scene_upload(ctx_, &d), "rt_scene_upload"); }
    /* a new film (height * width * 3 RGB floats) */
    std::vector<float> render(const Camera& cam)
    {
        std::vector<float> film((size_t)cam.width * cam.height * 3, 0.f);
        rt_camera c = cam.c();
        check(ctx_, rt_render_mlt(ctx_, &c, &p_, film.data(), &last_), "rt_render_mlt");
        return film;
    }
    const rt_mlt_result& lastResult() const { return last_; }
    void stop() { rt_request_stop(ctx_, 1); }
    rt_ctx* context() { return ctx_; }

protected:
    rt_ctx* ctx_ = nullptr;
    rt_mlt_params p_{};
    rt_mlt_result last_{};
};

/* met-large-step-prob, met-termination-prob, met-bootstrap, met-mutations, seed */
class MetropolisRenderer : public MarkovChainRenderer {
public:
    MetropolisRenderer(float largeStepProb = 0.1f, float terminationProb = 0.5f, int bootstrap = 10000,
                       int mutations = 16, unsigned long seed = 0, int chains = 0, int device = 0)
        : MarkovChainRenderer(device)
    {
        p_.kind = RT_MLT_METROPOLIS;
        p_.max_depth = 5;
        p_.rr_start = 3;
        p_.large_step_prob = largeStepProb;
        p_.termination_prob = terminationProb;
        p_.bootstrap = bootstrap;
        p_.mutations = mutations;
        p_.chains = chains;
        p_.seed = (uint32_t)seed;
    }
};

/* metflt-num-passes, the --metflt-* switches (rt_mlt_filter_flags), and the
 * met-* options of the chains */
class MetropolisFltRenderer : public MetropolisRenderer {
public:
    MetropolisFltRenderer(int passes = 3, int flags = 0, float largeStepProb = 0.1f, float terminationProb = 0.5f,
                          int bootstrap = 10000, int mutations = 16, unsigned long seed = 0, int chains = 0,
                          int device = 0)
        : MetropolisRenderer(largeStepProb, terminationProb, bootstrap, mutations, seed, chains, device)
    {
        p_.kind = RT_MLT_FILTERED;
        p_.flt_passes = passes;
        p_.flt_flags = flags;
    }
};

/* erpt-mutations, erpt-x-samples, erpt-y-samples, seed */
class EnergyRedistributionRenderer : public MarkovChainRenderer {
public:
    EnergyRedistributionRenderer(int mutations = 100, int xSamples = 2, int ySamples = 2, unsigned long seed = 0,
                                 int device = 0)
        : MarkovChainRenderer(device)
    {
        p_.kind = RT_MLT_ERPT;
        p_.max_depth = 5;
        p_.rr_start = 3;
        p_.mutations = mutations;
        p_.erpt_spp_x = xSamples;
        p_.erpt_spp_y = ySamples;
        p_.erpt_ed_termination = 0.1f;
        p_.seed = (uint32_t)seed;
    }
};

}  // namespace rtgpu

#endif
