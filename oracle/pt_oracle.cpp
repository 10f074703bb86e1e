/*
 * pt_oracle.cpp — CPU oracle for the path-tracing hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path and the CPU
 * baseline of the benchmark.  Nothing in the product links or calls it.
 *
 * This is a restatement of the reference renderer's own algorithms in its
 * own object model (virtual Intersectable / Material / Light classes, a
 * leaf-1 pointer BVH whose traversal visits both children, std::list paths,
 * an OpenMP loop over rows with a critical section per sample), written
 * without Qt, OpenCV, GSL or Boost.  Each function cites the reference
 * file:line it follows.  Vector arithmetic follows the Qt 5 semantics the
 * reference relies on: float component arithmetic, length() and
 * normalized() evaluated in double and rounded to float.
 *
 * Random numbers: the reference draws from one GSL taus2 stream per image
 * row (src/perpixelrenderer.cpp:67-68).  Here every draw comes from the
 * counter-based Philox-4x32-10 stream that the GPU path uses; the layout of
 * the stream is fixed by philox_uniforms() below and is the contract that
 * lets the GPU image be compared pixel by pixel with this one.
 *
 * The filtered Metropolis renderer's pieces that no reference vectors cover
 * are pinned one by one (tests/test_oracle_flt_pieces.py): the OpenCV blur
 * and sepFilter2D border arithmetic against scipy's reflect / mirror
 * filters, the importance map of later passes at hand-computed values, the
 * weighted running mean and variance against the direct weighted formulas;
 * the symmetric filter and one pass without filtering by test_oracle_mlt.py.
 * Parity unpinned: how the reference composes the passes (the order of
 * rescale, deposits and filtering) beyond GPU-oracle agreement.
 */
#include "pt_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <list>
#include <memory>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// ---------------------------------------------------------------------------
// Philox-4x32-10, Salmon et al., "Parallel random numbers: as easy as 1, 2, 3"
// (SC'11).  Constants from the paper: multipliers 0xD2511F53 / 0xCD9E8D57,
// Weyl key increments 0x9E3779B9 / 0xBB67AE85.
// ---------------------------------------------------------------------------
void philox(const uint32_t in[4], const uint32_t k[2], uint32_t out[4])
{
    uint32_t c0 = in[0], c1 = in[1], c2 = in[2], c3 = in[3];
    uint32_t k0 = k[0], k1 = k[1];
    for (int r = 0; r < 10; ++r) {
        if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint64_t a = (uint64_t)0xD2511F53u * c0;
        uint64_t b = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(b >> 32) ^ c1 ^ k0;
        uint32_t n2 = (uint32_t)(a >> 32) ^ c3 ^ k1;
        c0 = n0; c1 = (uint32_t)b; c2 = n2; c3 = (uint32_t)a;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// The random stream of one camera sample: counter = (pixel, sample index,
// dimension group, 0), key = (seed, view).  Group 0: pixel jitter (u0,u1)
// and the path-length draw (u2).  Group 1+2i: BSDF sample of bounce i (u0,u1),
// lobe pick (u2), roulette (u3).  Group 2+2k: light sample of stored vertex /
// bounce k (u0,u1), light index (u2).  Uniforms are the top 24 bits / 2^24.
struct U4 { float u[4]; };
U4 philox_uniforms(uint32_t seed, uint32_t view, uint32_t pixel, uint32_t sample, uint32_t group)
{
    uint32_t c[4] = {pixel, sample, group, 0u}, k[2] = {seed, view}, o[4];
    philox(c, k, o);
    U4 r;
    for (int i = 0; i < 4; ++i) r.u[i] = (float)(o[i] >> 8) * (1.0f / 16777216.0f);
    return r;
}

// ---------------------------------------------------------------------------
// Vector type with the Qt 5 QVector3D semantics used by the reference.
// ---------------------------------------------------------------------------
struct V3 {
    float x = 0, y = 0, z = 0;
    V3() = default;
    V3(float a, float b, float c) : x(a), y(b), z(c) {}
    float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    float& at(int i) { return i == 0 ? x : (i == 1 ? y : z); }
};
inline V3 operator+(V3 a, V3 b) { return V3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V3 operator-(V3 a, V3 b) { return V3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline V3 operator-(V3 a) { return V3(-a.x, -a.y, -a.z); }
inline V3 operator*(float s, V3 a) { return V3(a.x * s, a.y * s, a.z * s); }
inline V3 operator*(V3 a, float s) { return V3(a.x * s, a.y * s, a.z * s); }
inline bool is_zero(V3 a) { return a.x == 0 && a.y == 0 && a.z == 0; }
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b)
{
    return V3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
inline float length_sq(V3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
inline float length(V3 a)
{
    double l = (double)a.x * a.x + (double)a.y * a.y + (double)a.z * a.z;
    return (float)std::sqrt(l);
}
inline V3 normalized(V3 a)
{
    double l = (double)a.x * a.x + (double)a.y * a.y + (double)a.z * a.z;
    if (l == 0.0) return V3();
    double s = std::sqrt(l);
    return V3((float)(a.x / s), (float)(a.y / s), (float)(a.z / s));
}
// src/vechelper.cpp:3-6
inline V3 reflect(V3 v, V3 n) { return v - (2 * dot(v, n)) * n; }

// RGB colour (cv::Vec3f in the reference, stored there in BGR order).
struct Col {
    float r = 0, g = 0, b = 0;
    Col() = default;
    Col(float a, float b_, float c) : r(a), g(b_), b(c) {}
    bool zero() const { return r == 0 && g == 0 && b == 0; }
};
inline Col operator+(Col a, Col b) { return Col(a.r + b.r, a.g + b.g, a.b + b.b); }
inline Col operator*(Col a, float s) { return Col(a.r * s, a.g * s, a.b * s); }
inline Col mul(Col a, Col b) { return Col(a.r * b.r, a.g * b.g, a.b * b.b); }
// src/mathhelper.cpp:102-105 (.299 R + .587 G + .114 B, evaluated in double)
inline float lum(Col c) { return (float)(.299 * c.r + .587 * c.g + .114 * c.b); }

const float kEps = 0.001f;   // src/ray.h:34 (EPSILON 0.001, passed as float)
const float kInf = std::numeric_limits<float>::infinity();
const float kNaN = std::numeric_limits<float>::quiet_NaN();

// src/ray.h:36-111
struct Ray {
    V3 o, d;
    float from = kEps, to = kInf;
    Ray() = default;
    Ray(V3 o_, V3 d_, float f = kEps, float t = kInf) : o(o_), d(d_), from(f), to(t) {}
    bool in_range(float u) const { return from <= u && u <= to; }
};

struct Material;

// src/hitrecord.h:35-121: t = NaN is a miss; the point has w = 0 at t = inf.
struct Hit {
    float t = kNaN;
    V3 p;
    float pw = 1;
    V3 n;
    Ray ray;
    const Material* mat = nullptr;
    int prim = -1;
    float tc[2] = {0, 0};  // texture coordinates (cv::Point2f, default (0, 0))
    bool hits() const { return !std::isnan(t); }
};
Hit make_hit(float t, const Ray& ray, const Material* m, V3 normal, int prim, float ts = 0, float tt = 0)
{
    Hit h;
    h.tc[0] = ts;
    h.tc[1] = tt;
    h.t = t;
    h.ray = ray;
    h.mat = m;
    h.prim = prim;
    h.n = normalized(normal);
    if (std::isinf(t)) {  // src/ray.h:91-101
        h.p = (t > 0 ? 1.f : -1.f) * normalized(ray.d);
        h.pw = 0;
    } else {
        h.p = ray.o + t * ray.d;
        h.pw = 1;
    }
    return h;
}

// ---------------------------------------------------------------------------
// Sampling warps, src/sampler.cpp:33-86.  The sample is a QPointF (double).
// ---------------------------------------------------------------------------
struct Sample { double x, y; };

V3 cosine_power_dir(Sample s, V3 w, float& pdf, float n)
{
    const double rf = 1.0;  // sin(openingAngle = pi/2)
    w = normalized(w);
    float e2 = 2 / (n + 1), e1 = 1 / (n + 1);
    double rad = std::sqrt(1 - std::pow(s.x, (double)e2) * rf * rf);
    V3 nd((float)(std::cos(2 * M_PI * s.y) * rad), (float)(std::sin(2 * M_PI * s.y) * rad),
          (float)(std::pow(s.x, (double)e1) * rf));
    pdf = (float)(std::pow((double)nd.z, (double)n) * (n + 1) / (2 * M_PI));
    V3 v;
    if (std::fabs(w.x) < std::fabs(w.y) && std::fabs(w.x) < std::fabs(w.z))
        v = V3(0, w.z, -w.y);
    else if (std::fabs(w.y) < std::fabs(w.z))
        v = V3(w.z, 0, -w.x);
    else
        v = V3(w.y, -w.x, 0);
    v = normalized(v);
    V3 u = cross(v, w);
    V3 r(nd.x * u.x + nd.y * v.x + nd.z * w.x, nd.x * u.y + nd.y * v.y + nd.z * w.y,
         nd.x * u.z + nd.y * v.z + nd.z * w.z);
    return normalized(r);
}
inline V3 cosine_dir(Sample s, V3 w, float& pdf) { return cosine_power_dir(s, w, pdf, 1); }

// ---------------------------------------------------------------------------
// Intersectables
// ---------------------------------------------------------------------------
struct Box { V3 mn{kInf, kInf, kInf}, mx{-kInf, -kInf, -kInf}; };
void include_point(Box& b, V3 p)  // src/axisalignedbox.cpp:90-99
{
    for (int i = 0; i < 3; ++i) {
        if (p[i] < b.mn[i]) b.mn.at(i) = p[i];
        if (p[i] > b.mx[i]) b.mx.at(i) = p[i];
    }
}
void include_box(Box& b, const Box& o) { include_point(b, o.mn); include_point(b, o.mx); }

// Slab test, src/axisalignedbox.cpp:30-73.  Returns false on tmin > tmax.
bool slab(const Box& b, const Ray& ray, float& tmin, float& tmax)
{
    tmin = -kInf;
    tmax = kInf;
    for (int i = 0; i < 3; ++i) {
        float a = 1 / ray.d[i];
        float t0, t1;
        if (a >= 0) { t0 = b.mn[i] - ray.o[i]; t1 = b.mx[i] - ray.o[i]; }
        else { t0 = b.mx[i] - ray.o[i]; t1 = b.mn[i] - ray.o[i]; }
        t0 *= a;
        t1 *= a;
        if (t0 > tmin) tmin = t0;
        if (t1 < tmax) tmax = t1;
        if (tmin > tmax) return false;
    }
    return true;
}

struct Intersectable {
    Box box;
    virtual ~Intersectable() {}
    virtual Hit intersect(const Ray& ray) const = 0;
};

// src/triangle.cpp:30-56, ctor src/triangle.h:35-38
struct Triangle : Intersectable {
    V3 p1, p2, p3, n1, n2, n3;
    float t1[2] = {0, 0}, t2[2] = {0, 1}, t3[2] = {1, 0};  // default texcoords, src/triangle.h:35-36
    const Material* m;
    int prim;
    Triangle(V3 a, V3 b, V3 c, const Material* mat, int id) : p1(a), p2(b), p3(c), m(mat), prim(id)
    {
        n1 = n2 = n3 = normalized(cross(p1 - p2, p1 - p3));
        include_point(box, p1); include_point(box, p2); include_point(box, p3);
    }
    void set_normals(V3 a, V3 b, V3 c) { n1 = normalized(a); n2 = normalized(b); n3 = normalized(c); }
    Hit intersect(const Ray& ray) const override
    {
        V3 ad = ray.o - p1;
        V3 abxac = cross(p2 - p1, p3 - p1);
        float w = -dot(abxac, ray.d);
        if (w == 0) return Hit();
        float t = dot(abxac, ad) / w;
        if (!ray.in_range(t)) return Hit();
        V3 adxde = cross(ad, ray.d);
        float u = dot(adxde, p3 - p1) / w;
        if (u < 0 || u > 1) return Hit();
        float v = -dot(adxde, p2 - p1) / w;
        if (v < 0 || u + v > 1) return Hit();
        V3 normal = (1 - u - v) * n1 + u * n2 + v * n3;
        if (is_zero(normal)) return Hit();
        const float bw = 1 - u - v;  // (1 - u - v) * t1 + u * t2 + v * t3, src/triangle.cpp:54
        return make_hit(t, ray, m, normal, prim, bw * t1[0] + u * t2[0] + v * t3[0], bw * t1[1] + u * t2[1] + v * t3[1]);
    }
};

// src/quad.h:32, src/quad.cpp:23-41: t1 = (p1,p2,p4) first, then t2 = (p4,p2,p3)
struct Quad : Intersectable {
    Triangle t1, t2;
    Quad(Triangle a, Triangle b) : t1(a), t2(b) { box = t1.box; include_box(box, t2.box); }
    Hit intersect(const Ray& ray) const override
    {
        Hit h = t1.intersect(ray);
        return h.hits() ? h : t2.intersect(ray);
    }
};

// src/sphere.cpp:47: texture coordinates of the unnormalised normal n
inline float sphere_tc_s(V3 n, float r) { return (float)(std::acos(n.z / r) / M_PI); }
inline float sphere_tc_t(V3 n) { return (float)(std::atan2(n.x, n.y) / M_PI / 2 + 0.5f); }

// src/sphere.cpp:29-56 + src/csgobject.cpp:25-44
struct Sphere : Intersectable {
    V3 c;
    float r;
    const Material* m;
    int prim;
    Sphere(V3 c_, float r_, const Material* mat, int id) : c(c_), r(r_), m(mat), prim(id)
    {
        V3 rv(r, r, r);
        box.mn = c - rv;
        box.mx = c + rv;
    }
    Hit intersect(const Ray& ray) const override
    {
        V3 diff = ray.o - c;
        float r2 = r * r;
        float a = length_sq(ray.d);
        float b = 2 * dot(ray.d, diff);
        float cc = length_sq(diff) - r2;
        float disc = b * b - 4 * a * cc;
        if (!(disc >= 0)) return Hit();
        float root = std::sqrt(disc);
        float ts[2] = {(-b - root) / (2 * a), (-b + root) / (2 * a)};
        if (ray.from > ts[1]) return Hit();
        for (float t : ts) {
            if (t >= ray.from) {
                if (t > ray.to) return Hit();
                V3 n = (ray.o + t * ray.d) - c;
                return make_hit(t, ray, m, n, prim, sphere_tc_s(n, r), sphere_tc_t(n));
            }
        }
        return Hit();
    }
};

// src/plane.cpp:27-40: t = (P . (o, 1)) / -(d . n), accepted when
// from <= t <= to; the bounding box is infinite (kept out of the BVH here).
struct Plane : Intersectable {
    float P[4];
    const Material* m;
    int prim;
    Plane(const float* p, const Material* mat, int id) : m(mat), prim(id)
    {
        for (int k = 0; k < 4; ++k) P[k] = p[k];
    }
    Hit intersect(const Ray& ray) const override
    {
        float d = -(ray.d.x * P[0] + ray.d.y * P[1] + ray.d.z * P[2]);
        float u = ray.o.x * P[0] + ray.o.y * P[1] + ray.o.z * P[2] + 1.f * P[3];
        float t = u / d;
        if (!ray.in_range(t)) return Hit();
        return make_hit(t, ray, m, V3(P[0], P[1], P[2]), prim);
    }
};

// CSG objects: interval lists (getCSGIntersection) of spheres, boxes and
// boolean operations, src/sphere.cpp:29-56, src/axisalignedbox.cpp:30-71,
// src/csgoperation.cpp:26-91, resolved by src/csgobject.cpp:25-44.
struct CsgParam {
    float t;
    V3 n;
    const Material* m;
    float ts = 0, tt = 0;
};
struct CsgNode {
    Box box;
    virtual ~CsgNode() {}
    virtual std::vector<CsgParam> params(const Ray& ray) const = 0;
};
struct CsgSphere : CsgNode {
    V3 c;
    float r;
    const Material* m;
    CsgSphere(V3 c_, float r_, const Material* mat) : c(c_), r(r_), m(mat)
    {
        V3 rv(r, r, r);
        box.mn = c - rv;
        box.mx = c + rv;
    }
    std::vector<CsgParam> params(const Ray& ray) const override
    {
        std::vector<CsgParam> out;
        V3 diff = ray.o - c;
        float a = length_sq(ray.d);
        float b = 2 * dot(ray.d, diff);
        float cc = length_sq(diff) - r * r;
        float disc = b * b - 4 * a * cc;
        if (disc >= 0) {
            float root = std::sqrt(disc);
            for (float t : {(-b - root) / (2 * a), (-b + root) / (2 * a)}) {
                V3 n = (ray.o + t * ray.d) - c;
                out.push_back({t, n, m, sphere_tc_s(n, r), sphere_tc_t(n)});
            }
        }
        return out;
    }
};
struct CsgBox : CsgNode {
    const Material* m;
    CsgBox(V3 mn, V3 mx, const Material* mat) : m(mat)
    {
        box.mn = mn;
        box.mx = mx;
    }
    std::vector<CsgParam> params(const Ray& ray) const override
    {
        std::vector<CsgParam> out;
        float tmin = -kInf, tmax = kInf;
        V3 normal;
        const float mn[3] = {box.mn.x, box.mn.y, box.mn.z}, mx[3] = {box.mx.x, box.mx.y, box.mx.z};
        const float o[3] = {ray.o.x, ray.o.y, ray.o.z}, dv[3] = {ray.d.x, ray.d.y, ray.d.z};
        for (int i = 0; i < 3; ++i) {
            float a = 1 / dv[i];
            float n[3] = {0, 0, 0}, t0, t1;
            if (a >= 0) {
                t0 = mn[i] - o[i];
                t1 = mx[i] - o[i];
                n[i] = -1;
            } else {
                t0 = mx[i] - o[i];
                t1 = mn[i] - o[i];
                n[i] = 1;
            }
            t0 *= a;
            t1 *= a;
            if (t0 > tmin) {
                tmin = t0;
                normal = V3(n[0], n[1], n[2]);
            }
            if (t1 < tmax) tmax = t1;
            if (tmin > tmax) return out;
        }
        out.push_back({tmin, normal, m});
        out.push_back({tmax, normal, m});
        return out;
    }
};
struct CsgOp : CsgNode {
    int op;
    std::unique_ptr<CsgNode> l, r;
    CsgOp(int o, std::unique_ptr<CsgNode> a, std::unique_ptr<CsgNode> b) : op(o), l(std::move(a)), r(std::move(b))
    {
        box.mn = V3(std::min(l->box.mn.x, r->box.mn.x), std::min(l->box.mn.y, r->box.mn.y),
                    std::min(l->box.mn.z, r->box.mn.z));
        box.mx = V3(std::max(l->box.mx.x, r->box.mx.x), std::max(l->box.mx.y, r->box.mx.y),
                    std::max(l->box.mx.z, r->box.mx.z));
    }
    bool inside(bool a, bool b) const
    {
        switch (op) {
        case RT_CSG_ISECT: return a && b;
        case RT_CSG_UNION: return a || b;
        case RT_CSG_DIFF: return a && !b;
        default: return a != b;
        }
    }
    std::vector<CsgParam> params(const Ray& ray) const override
    {
        std::vector<CsgParam> L = l->params(ray), R = r->params(ray), out;
        bool li = false, ri = false;
        size_t a = 0, b = 0;
        while (a < L.size() || b < R.size()) {
            bool left = a < L.size() && (b == R.size() || L[a].t <= R[b].t);
            bool was = inside(li, ri);
            if (left) li = !li;
            else ri = !ri;
            if (was != inside(li, ri)) out.push_back(left ? L[a] : R[b]);
            if (left) ++a;
            else ++b;
        }
        return out;
    }
};
struct CsgObject : Intersectable {
    std::unique_ptr<CsgNode> root;
    int prim;
    CsgObject(std::unique_ptr<CsgNode> r, int id) : root(std::move(r)), prim(id) { box = root->box; }
    Hit intersect(const Ray& ray) const override
    {
        std::vector<CsgParam> ps = root->params(ray);
        if (ps.empty() || ray.from > ps.back().t) return Hit();
        for (const CsgParam& p : ps) {
            if (p.t >= ray.from) {
                if (p.t > ray.to) return Hit();
                return make_hit(p.t, ray, p.m, p.n, prim, p.ts, p.tt);
            }
        }
        return Hit();
    }
};

// IntersectableInstance::intersect, src/intersectableinstance.cpp:24-30: the
// ray mapped into object space by the inverse transform (points A x + b,
// directions A d; the range keeps its parameters), the object intersected
// there, and the hit reported in world space: the normal through the
// normal matrix A^T (src/hitrecord.h:79-84), the point o + t d of the world
// ray, the object-space texture coordinates.  The box is the transformed
// object box's 8 corners (createBoundingBox, :33-47).
struct Instance : Intersectable {
    Intersectable* inner;
    float r[12];  // rows (A row, b) of the world-to-object map
    Instance(Intersectable* obj, const float* w2o) : inner(obj)
    {
        std::copy(w2o, w2o + 12, r);
        // object-to-world: the inverse of [A | b], in double
        double a[3][4];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 4; ++j) a[i][j] = r[4 * i + j];
        const double det = a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) -
                           a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
                           a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
        double m[3][3];
        m[0][0] = (a[1][1] * a[2][2] - a[1][2] * a[2][1]) / det;
        m[0][1] = (a[0][2] * a[2][1] - a[0][1] * a[2][2]) / det;
        m[0][2] = (a[0][1] * a[1][2] - a[0][2] * a[1][1]) / det;
        m[1][0] = (a[1][2] * a[2][0] - a[1][0] * a[2][2]) / det;
        m[1][1] = (a[0][0] * a[2][2] - a[0][2] * a[2][0]) / det;
        m[1][2] = (a[0][2] * a[1][0] - a[0][0] * a[1][2]) / det;
        m[2][0] = (a[1][0] * a[2][1] - a[1][1] * a[2][0]) / det;
        m[2][1] = (a[0][1] * a[2][0] - a[0][0] * a[2][1]) / det;
        m[2][2] = (a[0][0] * a[1][1] - a[0][1] * a[1][0]) / det;
        const Box& ib = inner->box;
        box = Box();
        for (int c = 0; c < 8; ++c) {
            const double x[3] = {(c & 1) ? ib.mx.x : ib.mn.x, (c & 2) ? ib.mx.y : ib.mn.y, (c & 4) ? ib.mx.z : ib.mn.z};
            double w[3];
            for (int i = 0; i < 3; ++i)
                w[i] = m[i][0] * (x[0] - a[0][3]) + m[i][1] * (x[1] - a[1][3]) + m[i][2] * (x[2] - a[2][3]);
            include_point(box, V3((float)w[0], (float)w[1], (float)w[2]));
        }
    }
    V3 point(V3 x) const
    {
        return V3(((r[0] * x.x + r[1] * x.y) + r[2] * x.z) + r[3], ((r[4] * x.x + r[5] * x.y) + r[6] * x.z) + r[7],
                  ((r[8] * x.x + r[9] * x.y) + r[10] * x.z) + r[11]);
    }
    V3 dir(V3 x) const
    {
        return V3((r[0] * x.x + r[1] * x.y) + r[2] * x.z, (r[4] * x.x + r[5] * x.y) + r[6] * x.z,
                  (r[8] * x.x + r[9] * x.y) + r[10] * x.z);
    }
    V3 normal(V3 n) const
    {
        return V3((r[0] * n.x + r[4] * n.y) + r[8] * n.z, (r[1] * n.x + r[5] * n.y) + r[9] * n.z,
                  (r[2] * n.x + r[6] * n.y) + r[10] * n.z);
    }
    Hit intersect(const Ray& ray) const override
    {
        Hit h = inner->intersect(Ray(point(ray.o), dir(ray.d), ray.from, ray.to));
        if (!h.hits()) return h;
        return make_hit(h.t, ray, h.mat, normal(h.n), h.prim, h.tc[0], h.tc[1]);
    }
};

// src/intersectablelist.cpp:30-48: linear scan with ray shortening; ties go
// to the later component.
struct List : Intersectable {
    std::vector<Intersectable*> items;
    Hit intersect(const Ray& r0) const override
    {
        Ray ray = r0;
        Hit hit;
        for (auto* it : items) {
            Hit h = it->intersect(ray);
            if (h.hits()) {
                ray = Ray(ray.o, ray.d, ray.from, h.t);
                hit = h;
            }
        }
        return hit;
    }
};

// src/bvh.cpp:135-149: slab test of the node's own box with the unshortened
// ray, then both children; the smaller t wins, ties go to the right child.
struct BVHNode : Intersectable {
    Intersectable *l, *r;
    Hit intersect(const Ray& ray) const override
    {
        float t0, t1;
        if (!slab(box, ray, t0, t1) || t0 > ray.to || t1 < ray.from) return Hit();
        Hit lh = l->intersect(ray);
        Hit rh = r->intersect(ray);
        if (!rh.hits() || lh.t < rh.t) return lh;
        return rh;
    }
};

// src/bvh.cpp:67-94: longest axis (strict >, ties to the later axis), sort by
// box centre, split at n/2, one primitive per leaf.
Intersectable* build_bvh(std::vector<Intersectable*> v, std::vector<std::unique_ptr<Intersectable>>& pool)
{
    if (v.size() == 1) return v[0];
    Box bb;
    for (auto* it : v) include_box(bb, it->box);
    V3 diff = bb.mx - bb.mn;
    int axis = diff.x > diff.y && diff.x > diff.z ? 0 : (diff.y > diff.z ? 1 : 2);
    std::sort(v.begin(), v.end(), [axis](Intersectable* a, Intersectable* b) {
        return (a->box.mx + a->box.mn)[axis] < (b->box.mx + b->box.mn)[axis];
    });
    size_t half = v.size() / 2;
    std::vector<Intersectable*> left(v.begin(), v.begin() + half), right(v.begin() + half, v.end());
    auto* node = new BVHNode();
    pool.emplace_back(node);
    node->box = bb;
    node->l = build_bvh(left, pool);
    node->r = build_bvh(right, pool);
    return node;
}

// ---------------------------------------------------------------------------
// Materials, src/material.h:30-50
// ---------------------------------------------------------------------------
struct Material {
    virtual ~Material() {}
    virtual Col brdf(const Hit&, V3) const { return Col(); }
    virtual bool specular() const { return false; }
    virtual Col emission(const Hit&) const { return Col(); }
    virtual bool emits() const { return false; }
    virtual V3 out_direction(const Hit& hit, Sample s, float& pdf, Col& brdf) const = 0;
    virtual int kind() const = 0;
};

// src/darkmatter.cpp:32-39
struct DarkMatter : Material {
    V3 out_direction(const Hit&, Sample, float& pdf, Col& brdf) const override
    {
        brdf = Col();
        pdf = 0;
        return V3();
    }
    int kind() const override { return RT_MAT_DARK; }
};

// src/abstractdiffusematerial.cpp:8-15, src/diffusematerial.cpp:25-28
struct Diffuse : Material {
    Col kd;
    explicit Diffuse(Col c) : kd(c) {}
    Col brdf(const Hit&, V3) const override { return kd * (float)(1 / M_PI); }
    V3 out_direction(const Hit& hit, Sample s, float& pdf, Col& b) const override
    {
        V3 n = hit.n;
        if (dot(normalized(n), normalized(hit.ray.d)) > 0) n = n * -1.f;
        V3 out = cosine_dir(s, n, pdf);
        b = brdf(hit, out);
        return out;
    }
    int kind() const override { return RT_MAT_DIFFUSE; }
};

// src/texturematerial.cpp:33-72 (an AbstractDiffuseMaterial): brdf =
// pow(get(texcoord), gamma) / pi, get = the image sampled bilinearly at
// (clamp(s) (W - 1), clamp(1 - t) (H - 1)) with the border replicated
// (cv::getRectSubPix of a 1x1 patch) times the coefficient; clamp wraps into
// [0, 1) with fmod.
struct TextureMat : Material {
    int W = 0, H = 0;
    std::vector<float> rgb;
    Col coef;
    float gamma = 1;
    static float wrap(float x)
    {
        float f = std::fmod(x, 1.f);
        return f < 0 ? f + 1 : f;
    }
    Col albedo(const Hit& h) const
    {
        const float cx = wrap(h.tc[0]) * (W - 1), cy = wrap(1 - h.tc[1]) * (H - 1);
        const float fx = std::floor(cx), fy = std::floor(cy), a = cx - fx, b = cy - fy;
        const int x0 = std::min(std::max((int)fx, 0), W - 1), y0 = std::min(std::max((int)fy, 0), H - 1);
        const int x1 = std::min(x0 + 1, W - 1), y1 = std::min(y0 + 1, H - 1);
        auto px = [&](int x, int y, int k) { return rgb[3 * ((size_t)y * W + x) + k]; };
        const float w00 = (1 - a) * (1 - b), w01 = a * (1 - b), w10 = (1 - a) * b, w11 = a * b;
        float c[3];
        const float cf[3] = {coef.r, coef.g, coef.b};
        for (int k = 0; k < 3; ++k) {
            float v = px(x0, y0, k) * w00 + px(x1, y0, k) * w01 + px(x0, y1, k) * w10 + px(x1, y1, k) * w11;
            c[k] = std::pow(v * cf[k], gamma);
        }
        return Col(c[0], c[1], c[2]);
    }
    Col brdf(const Hit& h, V3) const override { return albedo(h) * (float)(1 / M_PI); }
    V3 out_direction(const Hit& hit, Sample s, float& pdf, Col& b) const override
    {
        V3 n = hit.n;
        if (dot(normalized(n), normalized(hit.ray.d)) > 0) n = n * -1.f;
        V3 out = cosine_dir(s, n, pdf);
        b = brdf(hit, out);
        return out;
    }
    int kind() const override { return RT_MAT_TEXTURE; }
};

// src/mirrormaterial.cpp:31-41
struct Mirror : Material {
    V3 out_direction(const Hit& hit, Sample, float& pdf, Col& b) const override
    {
        pdf = 1;
        b = Col();
        return reflect(hit.ray.d, normalized(hit.n));
    }
    bool specular() const override { return true; }
    int kind() const override { return RT_MAT_MIRROR; }
};

// src/phongmaterial.cpp:31-81
struct Phong : Material {
    Col kd, ks;
    float n;
    Phong(Col d, Col s, float e) : kd(d), ks(s), n(e) {}
    Col diffuse_brdf() const { return kd * (float)(1 / M_PI); }
    Col specular_brdf(const Hit& hit, V3 dir) const
    {
        V3 normal = normalized(hit.n);
        V3 refl = reflect(dir, normal);
        float dp = std::max(dot(normalized(refl), -normalized(hit.ray.d)), 0.f);
        float spec = (float)((n + 2) / (2 * M_PI) * std::pow((double)dp, (double)n));
        return ks * spec;
    }
    Col brdf(const Hit& hit, V3 dir) const override
    {
        return diffuse_brdf() + specular_brdf(hit, dir);
    }
    V3 out_direction(const Hit& hit, Sample s, float& pdf, Col& b) const override
    {
        V3 nrm = hit.n;
        if (dot(hit.ray.d, nrm) > 0) nrm = nrm * -1.f;
        nrm = normalized(nrm);
        V3 spec_dir = reflect(hit.ray.d, nrm);
        float rd = lum(kd), rs = lum(ks);
        if (s.x < rd) {
            s.x /= rd;
            V3 out = cosine_dir(s, nrm, pdf);
            b = diffuse_brdf();
            return out;
        } else if (s.x < rd + rs) {
            s.x = (s.x - rd) / rs;
            V3 out = cosine_power_dir(s, spec_dir, pdf, n);
            if (dot(out, nrm) < 1) pdf = 0;
            b = specular_brdf(hit, out);
            return out;
        }
        pdf = 0;
        b = Col();
        return V3();
    }
    int kind() const override { return RT_MAT_PHONG; }
};

inline int sgn_f(float x) { return (0.f < x) - (x < 0.f); }

// src/blinnmaterial.cpp:30-84
struct Blinn : Material {
    Col kd, ks;
    float e;
    Blinn(Col d, Col s, float ex) : kd(d), ks(s), e(ex) {}
    float D(V3 wh, V3 n) const { return (float)((e + 2) / (2 * M_PI) * std::pow((double)dot(wh, n), (double)e)); }
    static float G(V3 wi, V3 wo, V3 n)
    {
        V3 wh = normalized(wi + wo);
        return std::min(1.f, 2 * dot(n, wh) * std::min(dot(n, wo), dot(n, wi)) / dot(wo, wh));
    }
    Col brdf(const Hit& hit, V3 direction) const override
    {
        V3 n = normalized(hit.n);
        V3 wo = -normalized(hit.ray.d);
        V3 wi = -normalized(direction);
        if (sgn_f(dot(n, wo)) != sgn_f(dot(n, wi))) return Col();
        if (dot(n, wo) < 0) n = n * -1.f;
        V3 wh = normalized(wi + wo);
        return kd * (float)(1 / M_PI) + ks * (D(wh, n) * G(wi, wo, n) / (4 * dot(wo, n) * dot(wi, n)));
    }
    V3 out_direction(const Hit& hit, Sample s, float& pdf, Col& b) const override
    {
        V3 n = normalized(hit.n);
        V3 wo = -normalized(hit.ray.d);
        if (dot(n, wo) < 0) n = n * -1.f;
        V3 wh = cosine_power_dir(s, n, pdf, e);
        if (dot(wo, wh) < 0) {
            pdf = 0;
            b = Col();
            return V3();
        }
        V3 out = -reflect(wo, wh);
        pdf /= 4 * std::max(dot(wo, wh), .0001f);
        b = brdf(hit, out);
        return out;
    }
    int kind() const override { return RT_MAT_BLINN; }
};

// src/transparentmaterial.cpp:34-72: Fresnel-weighted choice between the
// mirror and the refracted direction; `u` is the sample's first coordinate.
V3 transparent_dir(V3 d_in, V3 n_in, float ior, double u)
{
    V3 direction = normalized(d_in), normal = normalized(n_in);
    float ratio;
    if (dot(-normal, direction) < 0.) {
        normal = normal * -1.f;
        ratio = ior;
    } else {
        ratio = 1 / ior;
    }
    float cosi = dot(-normal, direction);
    float sino = ratio * std::sqrt(std::max(0., 1. - cosi * cosi));
    float coso = std::sqrt(std::max(0., 1 - ratio * ratio * std::max(0., 1. - cosi * cosi)));
    float rperp = (cosi - ratio * coso) / (cosi + ratio * coso);
    float rpara = (ratio * cosi - coso) / (ratio * cosi + coso);
    float refl = 1 / 2. * (rperp * rperp + rpara * rpara);
    if (sino > 1 || u < refl) return reflect(direction, normal);
    return ratio * direction + (ratio * cosi - sgn_f(cosi) * coso) * normal;
}

struct Transparent : Material {
    float ior;
    explicit Transparent(float c) : ior(c) {}
    V3 out_direction(const Hit& hit, Sample s, float& pdf, Col& b) const override
    {
        pdf = 1;
        b = Col();
        return transparent_dir(hit.ray.d, hit.n, ior, s.x);
    }
    bool specular() const override { return true; }
    int kind() const override { return RT_MAT_REFRACTIVE; }
};

// ---------------------------------------------------------------------------
// Lights, src/light.h:31-37
// ---------------------------------------------------------------------------
struct Scene;

struct Light {
    virtual ~Light() {}
    // NEE: returns the light's contribution and sets `direction` to the
    // vector from the light towards the shaded point.
    virtual Col intensity(const Hit& hit, V3& direction, const Scene& sc, Sample s, int64_t& rays) const = 0;
};

struct EnvMap;

struct Scene {
    std::vector<std::unique_ptr<Intersectable>> owned;
    std::vector<std::unique_ptr<Material>> mats;
    std::vector<Light*> lights;
    std::vector<Intersectable*> prims;   // by primitive ID (null for non-intersectable lights)
    Intersectable* root = nullptr;
    const EnvMap* env = nullptr;
    Hit intersect(const Ray& ray) const;
};

// src/arealight.cpp:31-119
struct AreaLight : Intersectable, Material, Light {
    V3 o, u, v, nrm;
    Col I;
    int prim;
    AreaLight(V3 o_, V3 u_, V3 v_, Col i, int id) : o(o_), u(u_), v(v_), I(i), prim(id)
    {
        nrm = cross(u, v);
        include_point(box, o); include_point(box, o + u); include_point(box, o + v);
        include_point(box, o + u + v);
    }
    V3 location(Sample s) const { return o + (float)s.x * u + (float)s.y * v; }
    Col intensity_dir(V3 direction) const
    {
        float c = dot(normalized(direction), normalized(nrm)) / length_sq(direction) * length(nrm);
        if (c < 0) c = 0;
        return I * c;
    }
    Col intensity(const Hit& hit, V3& direction, const Scene& sc, Sample s, int64_t& rays) const override
    {
        V3 at = hit.p;
        V3 y = location(s);
        direction = at - y;
        Ray sr(at, -normalized(direction), kEps, (float)((double)length(direction) - 0.001));
        ++rays;
        Hit sh = sc.intersect(sr);
        if (sh.hits() && sh.mat != static_cast<const Material*>(this)) return Col();
        return intensity_dir(direction);
    }
    Hit intersect(const Ray& ray) const override
    {
        V3 ad = ray.o - o;
        V3 abxac = cross(u, v);
        float w = -dot(abxac, ray.d);
        if (w == 0) return Hit();
        float t = dot(abxac, ad) / w;
        if (!ray.in_range(t)) return Hit();
        V3 adxde = cross(ad, ray.d);
        float uu = dot(adxde, v) / w;
        if (uu < 0 || uu > 1) return Hit();
        float vv = -dot(adxde, u) / w;
        if (vv < 0 || vv > 1) return Hit();
        return make_hit(t, ray, this, nrm, prim);
    }
    Col emission(const Hit& hit) const override { return intensity_dir(-hit.ray.d); }
    bool emits() const override { return true; }
    V3 out_direction(const Hit&, Sample, float& pdf, Col& b) const override
    {
        b = Col();
        pdf = 0;
        return V3();
    }
    int kind() const override { return -1; }
};

// src/pointlight.cpp:29-70
struct PointLight : Intersectable, Light {
    V3 p;
    Col I;
    PointLight(V3 p_, Col i) : p(p_), I(i) { box.mn = V3(); box.mx = V3(); }
    Col intensity(const Hit& hit, V3& direction, const Scene& sc, Sample, int64_t& rays) const override
    {
        V3 at = hit.p;
        direction = at - p;
        Ray sr(at, -normalized(direction), kEps, (float)((double)length(direction) - 0.001));
        ++rays;
        if (sc.intersect(sr).hits()) return Col();
        return I * (1 / length_sq(direction));
    }
    Hit intersect(const Ray&) const override { return Hit(); }
};

// src/conelight.cpp:31-45: a point light lit only inside the cone of the
// given opening angle about its axis
struct ConeLight : Intersectable, Light {
    V3 p, axis;
    float angle;
    Col I;
    ConeLight(V3 p_, V3 a, float ang, Col i) : p(p_), axis(a), angle(ang), I(i) { box.mn = V3(); box.mx = V3(); }
    bool admits(V3 direction) const
    {
        float a = std::acos(dot(normalized(direction), normalized(axis)));
        return 0 <= a && a <= angle;
    }
    Col intensity(const Hit& hit, V3& direction, const Scene& sc, Sample, int64_t& rays) const override
    {
        V3 at = hit.p;
        direction = at - p;
        Ray sr(at, -normalized(direction), kEps, (float)((double)length(direction) - 0.001));
        ++rays;
        Hit sh = sc.intersect(sr);
        if (admits(direction) && !sh.hits()) return I * (1 / length_sq(direction));
        return Col();
    }
    Hit intersect(const Ray&) const override { return Hit(); }
};

// Latitude-longitude environment light.  Lookup direction -> texel:
// theta = acos(y), phi = atan2(x, -z), column = (phi/2pi + 1/2) W,
// row = theta/pi H (nearest texel).  NEE follows src/environmentmap.cpp:40-55
// (cosine sample about the unflipped normal, backwards shadow ray, value over
// pdf) in the reference-compatible mode.
struct EnvMap : Material, Light {
    int W = 0, H = 0;
    std::vector<float> rgb;
    Col coef;
    int prim;
    std::vector<double> row_cdf;              // H + 1
    std::vector<std::vector<double>> col_cdf; // H x (W + 1)
    std::vector<double> texel_p;              // W * H, probability of each texel
    Col texel(int x, int y) const
    {
        const float* c = &rgb[3 * ((size_t)y * W + x)];
        return Col(c[0], c[1], c[2]);
    }
    void texel_of(V3 d, int& x, int& y) const
    {
        double dy = std::max(-1.0, std::min(1.0, (double)d.y));
        double theta = std::acos(dy);
        double phi = std::atan2((double)d.x, -(double)d.z);
        double u = phi / (2 * M_PI) + 0.5, v = theta / M_PI;
        x = std::min(std::max((int)(u * W), 0), W - 1);
        y = std::min(std::max((int)(v * H), 0), H - 1);
    }
    // angular light probe (env_kind 1): its own texels for lookups; the
    // lat-long texels above then hold a proxy for the sampling tables only
    bool probe = false;
    int PW = 0, PH = 0;
    std::vector<float> prgb;
    // EnvironmentMap::get, src/environmentmap.cpp:35-66: (d.y, -d.z) scaled to
    // length acos(d.x) / pi; pixelFromNormalizedCoord = ((c + 1) / 2) size - 1,
    // truncated; row from the second coordinate, column from the first (the
    // rim index -1 and the undefined axis directions clamped / centred)
    int probe_index(V3 d) const
    {
        float icx = d.y, icy = -d.z;
        float len = (float)std::sqrt((double)icx * icx + (double)icy * icy);
        double ang = std::acos((double)std::max(-1.f, std::min(1.f, d.x)));
        float s = len > 0 ? (float)(ang / ((double)len * M_PI)) : 0.f;
        icx *= s;
        icy *= s;
        int x = (int)(float)(((icx + 1.0) / 2.0) * PW - 1);
        int y = (int)(float)(((icy + 1.0) / 2.0) * PH - 1);
        int row = std::min(y, PW - 1), col = std::min(x, PH - 1);
        row = std::min(std::max(row, 0), PH - 1);
        col = std::min(std::max(col, 0), PW - 1);
        return row * PW + col;
    }
    void set_probe(const float* img, int w, int h)
    {
        probe = true;
        PW = w;
        PH = h;
        prgb.assign(img, img + (size_t)3 * w * h);
        W = 512;
        H = 256;
        rgb.assign((size_t)3 * W * H, 0.f);
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x) {
                double th = (y + 0.5) / H * M_PI, ph = ((x + 0.5) / W - 0.5) * 2 * M_PI;
                V3 d((float)(std::sin(th) * std::sin(ph)), (float)std::cos(th), (float)(-std::sin(th) * std::cos(ph)));
                int t = probe_index(d);
                for (int k = 0; k < 3; ++k) rgb[3 * ((size_t)y * W + x) + k] = prgb[3 * (size_t)t + k];
            }
    }
    Col get(V3 d) const
    {
        d = normalized(d);
        if (probe) {
            const float* c = &prgb[3 * (size_t)probe_index(d)];
            return mul(Col(c[0], c[1], c[2]), coef);
        }
        int x, y;
        texel_of(d, x, y);
        return mul(texel(x, y), coef);
    }
    void build_sampling()
    {
        row_cdf.assign(H + 1, 0.0);
        col_cdf.assign(H, std::vector<double>(W + 1, 0.0));
        texel_p.assign((size_t)W * H, 0.0);
        std::vector<double> roww(H, 0.0);
        double total = 0;
        for (int y = 0; y < H; ++y) {
            double st = std::sin((y + 0.5) / H * M_PI);
            double acc = 0;
            for (int x = 0; x < W; ++x) {
                double w = (double)lum(mul(texel(x, y), coef)) * st;
                texel_p[(size_t)y * W + x] = w;
                acc += w;
                col_cdf[y][x + 1] = acc;
            }
            for (int x = 0; x <= W; ++x) col_cdf[y][x] = acc > 0 ? col_cdf[y][x] / acc : (double)x / W;
            roww[y] = acc;
            total += acc;
        }
        double acc = 0;
        for (int y = 0; y < H; ++y) { acc += roww[y]; row_cdf[y + 1] = total > 0 ? acc / total : (double)(y + 1) / H; }
        for (auto& p : texel_p) p = total > 0 ? p / total : 1.0 / ((double)W * H);
    }
    static int pick(const std::vector<double>& cdf, double u, double& frac)
    {
        int n = (int)cdf.size() - 1;
        int i = (int)(std::upper_bound(cdf.begin(), cdf.end(), u) - cdf.begin()) - 1;
        i = std::min(std::max(i, 0), n - 1);
        double w = cdf[i + 1] - cdf[i];
        frac = w > 0 ? (u - cdf[i]) / w : 0.5;
        frac = std::min(std::max(frac, 0.0), 1.0 - 1e-7);
        return i;
    }
    void sample(float u0, float u1, V3& d, float& pdf) const
    {
        double fy, fx;
        int y = pick(row_cdf, u0, fy);
        int x = pick(col_cdf[y], u1, fx);
        double theta = (y + fy) / H * M_PI, phi = ((x + fx) / W - 0.5) * 2 * M_PI;
        double st = std::sin(theta);
        d = V3((float)(st * std::sin(phi)), (float)std::cos(theta), (float)(-st * std::cos(phi)));
        pdf = st > 0 ? (float)(texel_p[(size_t)y * W + x] * W * H / (2 * M_PI * M_PI * st)) : 0.f;
    }
    float pdf(V3 d) const
    {
        d = normalized(d);
        int x, y;
        texel_of(d, x, y);
        double st = std::sqrt(std::max(0.0, 1.0 - (double)d.y * d.y));
        return st > 0 ? (float)(texel_p[(size_t)y * W + x] * W * H / (2 * M_PI * M_PI * st)) : 0.f;
    }
    Col intensity(const Hit& hit, V3& direction, const Scene& sc, Sample s, int64_t& rays) const override
    {
        float pdfv;
        direction = -cosine_dir(s, hit.n, pdfv);
        Ray sr(hit.p, direction, -kInf, -kEps);
        ++rays;
        if (sc.intersect(sr).hits()) return Col();
        return get(direction) * (1 / pdfv);
    }
    Col emission(const Hit& hit) const override { return get(normalized(hit.ray.d)); }
    bool emits() const override { return true; }
    V3 out_direction(const Hit&, Sample, float& pdf, Col& b) const override
    {
        b = Col();
        pdf = 0;
        return V3();
    }
    int kind() const override { return -2; }
};

// Closest hit against the scene; the environment answers rays that leave the
// scene with an unbounded range (src/environmentmap.cpp:73-83).
Hit Scene::intersect(const Ray& ray) const
{
    Hit h = root ? root->intersect(ray) : Hit();
    if (!h.hits() && env && std::isinf(ray.to) && ray.to > 0)
        h = make_hit(ray.to, ray, env, -normalized(ray.d), env->prim);
    return h;
}

// Shadow query of the unbiased estimator: geometry only, no environment.
Hit root_any(const Scene& sc, const Ray& r) { return sc.root ? sc.root->intersect(r) : Hit(); }

// ---------------------------------------------------------------------------
// Camera, src/camera.cpp:34-58
// ---------------------------------------------------------------------------
struct Camera {
    V3 eye, u, v, w;
    float t, b, l, r;
    int W, H;
    explicit Camera(const rt_camera& c)
    {
        eye = V3(c.eye[0], c.eye[1], c.eye[2]);
        V3 la(c.look_at[0], c.look_at[1], c.look_at[2]), up(c.up[0], c.up[1], c.up[2]);
        W = c.width;
        H = c.height;
        w = normalized(eye - la);
        u = normalized(cross(up, w));
        v = cross(w, u);
        float aspect = (float)W / H;
        t = (float)std::tan((double)(c.fov / 2));
        b = -t;
        r = aspect * t;
        l = -r;
    }
    Ray ray(double px, double py) const
    {
        float x = (float)(l + (r - l) * (px + 0.5) / W);
        float y = (float)(t + (b - t) * (py + 0.5) / H);
        V3 d(x * u.x + y * v.x + -1.f * w.x, x * u.y + y * v.y + -1.f * w.y, x * u.z + y * v.z + -1.f * w.z);
        return Ray(eye, d);
    }
};

// ---------------------------------------------------------------------------
// Reference path sampler and integrator ("compat" mode)
// ---------------------------------------------------------------------------
struct Path {  // src/path.h:31-36
    std::list<Col> alpha;
    std::list<Hit> hits;
};

struct SampleKey { uint32_t seed, view, pixel, sample; };

// Where a path's random numbers come from: the Philox streams of one pixel
// sample (the path-tracing renderer), or a primary-sample vector of
// 2 * max_depth + 1 groups (MetropolisSample, src/metropolissample.h:47-52:
// group 0 camera sample + length uniform, 1 + 2i cameraPathSamples[i],
// 2 + 2k lightSample1[k] + lightIndex uniform).
struct SampleSource {
    SampleKey k{};
    const U4* pss = nullptr;
    SampleSource(const SampleKey& key) : k(key) {}
    explicit SampleSource(const U4* v) : pss(v) {}
    U4 get(int g) const { return pss ? pss[g] : philox_uniforms(k.seed, k.view, k.pixel, k.sample, (uint32_t)g); }
};

// Renderer::createPath, src/renderer.cpp:47-106, with MAX_DEPTH = max_depth
// and the negative-binomial length draw made from one uniform by inversion.
Path create_path(const Scene& sc, const Ray& primary, const SampleSource& k, int max_depth, int rr_start,
                 float p, int64_t& rays)
{
    U4 g0 = k.get(0);
    int L;
    if (p > 0) {
        int nb = 0;
        if (p < 1) {
            double u = 1.0 - (double)g0.u[2];
            double q = std::floor(std::log(u) / std::log(1.0 - (double)p));
            nb = q > 1e6 ? 1000000 : (int)q;
        }
        L = std::min(rr_start + nb, max_depth);
    } else {
        p = 1;
        L = max_depth;
    }
    Path res;
    Col alpha(1, 1, 1);
    ++rays;
    Hit hit = sc.intersect(primary);
    for (int i = 0; i < L; ++i) {
        if (!hit.hits()) return res;
        U4 gs = k.get(1 + 2 * i);
        Sample s{gs.u[0], gs.u[1]};
        float pdf = 0;
        Col brdf;
        V3 out = hit.mat->out_direction(hit, s, pdf, brdf);
        if (hit.mat->emits()) {
            res.alpha.push_back(alpha);
            res.hits.push_back(hit);
            return res;
        } else if (hit.mat->specular()) {
            if (i == L - 1 && L < max_depth) ++L;
        } else {
            res.alpha.push_back(alpha);
            res.hits.push_back(hit);
            if (pdf == 0 || alpha.zero()) return res;
            float c = std::fabs(dot(normalized(out), normalized(hit.n)));
            alpha = mul(alpha, brdf) * (c / pdf);
            if (i > rr_start) alpha = alpha * (1 / p);
        }
        ++rays;
        hit = sc.intersect(Ray(hit.p, out));
    }
    return res;
}

inline int signum(float x) { return (0.f < x) - (x < 0.f); }

// UniDiPathTracingIntegrator::integrate(Ray) + integrate(Path),
// src/unidipathtracingintegrator.cpp:34-107
Col integrate_compat(const Scene& sc, const Ray& ray, const SampleSource& k, int max_depth, int rr_start, float p,
                     int64_t& rays)
{
    Path path = create_path(sc, ray, k, max_depth, rr_start, p, rays);
    Col color;
    auto ait = path.alpha.begin();
    auto hit = path.hits.begin();
    if (hit != path.hits.end() && !hit->mat->emission(*hit).zero()) return hit->mat->emission(*hit);
    int i = 0;
    const size_t nl = sc.lights.size();
    while (ait != path.alpha.end() && hit != path.hits.end()) {
        V3 direction;
        Col li;
        auto next = hit;
        ++next;
        if (next != path.hits.end() && next->mat->emits()) {
            li = next->mat->emission(*next);
            if (next->pw == 0)
                direction = -next->p;
            else
                direction = hit->p - next->p;
        } else if (nl > 0) {
            U4 gl = k.get(2 + 2 * i);
            size_t idx = std::min((size_t)(gl.u[2] * (float)nl), nl - 1);
            li = sc.lights[idx]->intensity(*hit, direction, sc, Sample{gl.u[0], gl.u[1]}, rays);
        }
        Col brdf = hit->mat->brdf(*hit, direction);
        float in_cos = dot(-normalized(direction), normalized(hit->n));
        float out_cos = dot(-normalized(hit->ray.d), normalized(hit->n));
        if (signum(in_cos) == signum(out_cos)) color = color + mul(mul(*ait, brdf), li) * std::fabs(in_cos);
        ++ait;
        ++hit;
        ++i;
    }
    return color;
}

// ---------------------------------------------------------------------------
// "Unbiased" estimator: next-event estimation with the light-selection
// weight, emission counted on camera and specular chains only, Russian
// roulette compensated by 1/(1-q), Phong sampled as a proper lobe mixture.
// This is not in the reference; it is the textbook estimator (Pharr, Jakob,
// Humphreys, PBRT 3rd ed., ch. 14) restated with the reference materials.
// ---------------------------------------------------------------------------
struct Frame { V3 t, b, n; };
Frame frame_about(V3 w)  // the same basis rule as src/sampler.cpp:55-68
{
    V3 v;
    if (std::fabs(w.x) < std::fabs(w.y) && std::fabs(w.x) < std::fabs(w.z))
        v = V3(0, w.z, -w.y);
    else if (std::fabs(w.y) < std::fabs(w.z))
        v = V3(w.z, 0, -w.x);
    else
        v = V3(w.y, -w.x, 0);
    v = normalized(v);
    return Frame{cross(v, w), v, w};
}
V3 lobe_dir(float u0, float u1, V3 axis, float expo)  // cos^expo lobe about axis (unit)
{
    float z = std::pow(u0, 1.f / (expo + 1.f));
    float r = std::sqrt(std::max(0.f, 1.f - z * z));
    float ph = 6.28318530718f * u1;
    Frame f = frame_about(axis);
    V3 d = (r * std::cos(ph)) * f.t + (r * std::sin(ph)) * f.b + z * f.n;
    return normalized(d);
}

struct UnbiasedCtx { const Scene* sc; int max_depth, rr_start; float rr_prob; };

// BSDF value for wo -> wi at a non-specular vertex with viewer-side normal nf
Col bsdf_eval(const Material* m, V3 wo, V3 wi, V3 nf)
{
    const float inv_pi = 0.318309886f;
    if (m->kind() == RT_MAT_DIFFUSE) return static_cast<const Diffuse*>(m)->kd * inv_pi;
    if (m->kind() == RT_MAT_PHONG) {
        auto* ph = static_cast<const Phong*>(m);
        V3 r = reflect(-wo, nf);
        float c = std::max(0.f, dot(r, wi));
        float s = (ph->n + 2) * (0.5f * inv_pi) * std::pow(c, ph->n);
        return ph->kd * inv_pi + ph->ks * s;
    }
    if (m->kind() == RT_MAT_BLINN) {
        auto* bl = static_cast<const Blinn*>(m);
        float co = dot(wo, nf), ci = dot(wi, nf);
        if (!(co > 0) || !(ci > 0)) return bl->kd * inv_pi;
        V3 wh = normalized(wo + wi);
        float s = bl->D(wh, nf) * Blinn::G(wi, wo, nf) / (4 * co * ci);
        return bl->kd * inv_pi + bl->ks * s;
    }
    return Col();
}
float bsdf_pdf(const Material* m, V3 wo, V3 wi, V3 nf)
{
    const float inv_pi = 0.318309886f;
    float cs = dot(wi, nf);
    if (cs <= 0) return 0;
    if (m->kind() == RT_MAT_DIFFUSE) return cs * inv_pi;
    if (m->kind() == RT_MAT_PHONG) {
        auto* ph = static_cast<const Phong*>(m);
        float rd = lum(ph->kd), rs = lum(ph->ks);
        float pd = rd / (rd + rs);
        V3 r = reflect(-wo, nf);
        float c = std::max(0.f, dot(r, wi));
        return pd * cs * inv_pi + (1 - pd) * (ph->n + 1) * (0.5f * inv_pi) * std::pow(c, ph->n);
    }
    if (m->kind() == RT_MAT_BLINN) {
        auto* bl = static_cast<const Blinn*>(m);
        float rd = lum(bl->kd), rs = lum(bl->ks);
        float pd = rd / (rd + rs);
        V3 wh = normalized(wo + wi);
        float ch = std::max(0.f, dot(wh, nf));
        float ph = (bl->e + 1) * (0.5f * inv_pi) * std::pow(ch, bl->e) / (4 * std::max(dot(wo, wh), 1e-4f));
        return pd * cs * inv_pi + (1 - pd) * ph;
    }
    return 0;
}
// Samples wi; returns false when the path must stop.
bool bsdf_sample(const Material* m, V3 wo, V3 nf, float u0, float u1, float u2, V3& wi)
{
    if (m->kind() == RT_MAT_DIFFUSE) {
        wi = lobe_dir(u0, u1, nf, 1.f);
        return true;
    }
    if (m->kind() == RT_MAT_PHONG) {
        auto* ph = static_cast<const Phong*>(m);
        float rd = lum(ph->kd), rs = lum(ph->ks);
        if (!(rd + rs > 0)) return false;
        float pd = rd / (rd + rs);
        if (u2 < pd) wi = lobe_dir(u0, u1, nf, 1.f);
        else wi = lobe_dir(u0, u1, reflect(-wo, nf), ph->n);
        return dot(wi, nf) > 0;
    }
    if (m->kind() == RT_MAT_BLINN) {
        auto* bl = static_cast<const Blinn*>(m);
        float rd = lum(bl->kd), rs = lum(bl->ks);
        if (!(rd + rs > 0)) return false;
        float pd = rd / (rd + rs);
        if (u2 < pd) {
            wi = lobe_dir(u0, u1, nf, 1.f);
        } else {
            V3 wh = lobe_dir(u0, u1, nf, bl->e);
            float c = dot(wo, wh);
            if (!(c > 0)) return false;
            wi = wh * (2 * c) - wo;
        }
        return dot(wi, nf) > 0;
    }
    return false;
}

// Power heuristic with exponent 2 (Veach, "Robust Monte Carlo Methods for
// Light Transport Simulation", 1997, sec. 9.2.1): weight of a sample drawn
// with density a when a second strategy has density b there.  Requires a > 0.
float mis_power(float a, float b)
{
    float r = b / a;
    return 1.f / (1.f + r * r);
}

// Next-event estimation and BSDF sampling are combined with the power
// heuristic.  The next-event weight is 1 at the last vertex (its BSDF ray is
// not traced) and for point lights; emission reached by a camera ray or
// through mirrors only has weight 1.
Col integrate_unbiased(const UnbiasedCtx& c, Ray ray, const SampleKey& k, int64_t& rays)
{
    const Scene& sc = *c.sc;
    Col L, alpha(1, 1, 1);
    bool spec = true;
    float pdf_prev = 0;  // solid-angle density of the current ray's direction (non-specular vertex)
    const int nl = (int)sc.lights.size();
    for (int depth = 0; depth < c.max_depth; ++depth) {
        ++rays;
        Hit hit = sc.intersect(ray);
        if (!hit.hits()) break;
        if (hit.mat->emits()) {
            Col e;
            float pl = 0;
            if (hit.pw == 0) {
                auto* em = static_cast<const EnvMap*>(hit.mat);
                e = em->get(ray.d);
                if (!spec) pl = em->pdf(ray.d) / (float)nl;
            } else {
                auto* al = static_cast<const AreaLight*>(hit.mat);
                float dn = dot(ray.d, al->nrm);
                if (dn < 0) {
                    e = al->I;
                    // area density t^2 / (A cos) with A cos = -dot(d, u x v)
                    if (!spec) pl = hit.t * hit.t / (-dn) / (float)nl;
                }
            }
            float w = spec ? 1.f : mis_power(pdf_prev, pl);
            L = L + mul(alpha, e) * w;
            break;
        }
        const Material* m = hit.mat;
        Diffuse tex_here(Col(0, 0, 0));  // a texture is a diffuse surface with the albedo at the hit
        if (m->kind() == RT_MAT_TEXTURE) {
            tex_here.kd = static_cast<const TextureMat*>(m)->albedo(hit);
            m = &tex_here;
        }
        V3 n = hit.n;
        V3 d = normalized(ray.d);
        V3 nf = dot(n, d) > 0 ? -n : n;
        V3 x = hit.p;
        if (m->kind() == RT_MAT_MIRROR) {
            ray = Ray(x, reflect(d, n));
            spec = true;
            continue;
        }
        if (m->kind() == RT_MAT_REFRACTIVE) {
            U4 gr = philox_uniforms(k.seed, k.view, k.pixel, k.sample, 1 + 2 * depth);
            ray = Ray(x, transparent_dir(ray.d, n, static_cast<const Transparent*>(m)->ior, gr.u[0]));
            spec = true;
            continue;
        }
        if (m->kind() == RT_MAT_DARK) break;
        V3 wo = -d;
        U4 gb = philox_uniforms(k.seed, k.view, k.pixel, k.sample, 1 + 2 * depth);
        U4 gl = philox_uniforms(k.seed, k.view, k.pixel, k.sample, 2 + 2 * depth);
        const bool last = depth + 1 >= c.max_depth;
        if (nl > 0) {
            int j = std::min((int)(gl.u[2] * (float)nl), nl - 1);
            const Light* lt = sc.lights[j];
            V3 wi;
            float dist = kInf;
            float pl = 0;  // solid-angle density incl. the 1/nl light choice; 0 for a point light
            Col le;
            bool ok = false;
            if (auto* al = dynamic_cast<const AreaLight*>(lt)) {
                V3 y = al->o + gl.u[0] * al->u + gl.u[1] * al->v;
                V3 dv = y - x;
                float d2 = dot(dv, dv);
                dist = std::sqrt(d2);
                wi = dv * (1 / dist);
                float an = std::sqrt(dot(al->nrm, al->nrm));
                float cl = -dot(wi, al->nrm) / an;
                if (cl > 0) {
                    le = al->I * (cl * an / d2);
                    pl = d2 / (cl * an) / (float)nl;
                    ok = true;
                }
            } else if (auto* pt = dynamic_cast<const PointLight*>(lt)) {
                V3 dv = pt->p - x;
                float d2 = dot(dv, dv);
                dist = std::sqrt(d2);
                wi = dv * (1 / dist);
                le = pt->I * (1 / d2);
                ok = true;
            } else if (auto* cl = dynamic_cast<const ConeLight*>(lt)) {
                V3 dv = cl->p - x;
                float d2 = dot(dv, dv);
                dist = std::sqrt(d2);
                wi = dv * (1 / dist);
                le = cl->I * (1 / d2);
                ok = cl->admits(-dv);
            } else if (auto* em = dynamic_cast<const EnvMap*>(lt)) {
                float pdf;
                em->sample(gl.u[0], gl.u[1], wi, pdf);
                if (pdf > 0) {
                    le = em->get(wi) * (1 / pdf);
                    pl = pdf / (float)nl;
                    ok = true;
                }
            }
            float cx = dot(wi, nf);
            if (ok && cx > 0) {
                float w = (pl > 0 && !last) ? mis_power(pl, bsdf_pdf(m, wo, wi, nf)) : 1.f;
                Col f = bsdf_eval(m, wo, wi, nf);
                Col contrib = mul(mul(alpha, f), le) * (cx * (float)nl * w);
                if (!contrib.zero()) {
                    float tmax = std::isinf(dist) ? kInf : dist - kEps;
                    ++rays;
                    Hit sh = root_any(sc, Ray(x, wi, kEps, tmax));
                    if (!sh.hits()) L = L + contrib;
                }
            }
        }
        V3 wi;
        if (!bsdf_sample(m, wo, nf, gb.u[0], gb.u[1], gb.u[2], wi)) break;
        float pdf = bsdf_pdf(m, wo, wi, nf);
        if (!(pdf > 0)) break;
        Col f = bsdf_eval(m, wo, wi, nf);
        alpha = mul(alpha, f) * (dot(wi, nf) / pdf);
        spec = false;
        pdf_prev = pdf;
        if (depth + 1 >= c.rr_start) {
            if (gb.u[3] < c.rr_prob) break;
            alpha = alpha * (1 / (1 - c.rr_prob));
        }
        ray = Ray(x, wi);
    }
    return L;
}

}  // namespace

// ---------------------------------------------------------------------------
// Markov-chain renderers over the sample-driven integrator
// (src/metropolissample.cpp, src/metropolisrenderer.cpp:56-200,
// src/energyredistributionrenderer.cpp:38-131).  Every random number is a
// Philox draw addressed by (seed, stream, index, step, group), the streams
// of include/rt_mlt.h, so chains here and on the GPU take the same steps.
// ---------------------------------------------------------------------------
struct Pss {
    std::vector<U4> g;
    explicit Pss(int groups = 0) : g(groups) {}
};

inline U4 draw(uint32_t seed, uint32_t stream, uint32_t index, uint32_t step, uint32_t group)
{
    return philox_uniforms(seed, stream, index, step, group);
}

// mutate(qreal&), src/metropolissample.cpp:62-78: a step of length
// b * exp(-log(b / a) * u) between a = 1/1024 and b = 1/64, negative when
// v < 0.5, wrapped into [0, 1].  The coordinate is a float here.
float mlt_perturb(float s, float u, float v)
{
    const float a = 1.f / 1024.f, b = 1.f / 64.f;
    const float log_ratio = -std::log(b / a);
    float delta = (float)(b * std::exp((double)log_ratio * u));
    if (v < 0.5f) delta = -delta;
    s += delta;
    if (s < 0) s += 1;
    if (s > 1) s -= 1;
    return s;
}

// Path through a primary-sample vector: the camera sample scaled to pixels
// (MetropolisSample::cameraPathFromSample, src/metropolissample.cpp:120-127),
// or given in pixels (the e_d estimate, src/energyredistributionrenderer.cpp:121-125).
Col integrate_pss(const Scene& sc, const Camera& cam, const Pss& x, int max_depth, int rr_start, float p,
                  bool pixel_units, int64_t& rays)
{
    const U4& g0 = x.g[0];
    double px = pixel_units ? (double)g0.u[0] : (double)g0.u[0] * cam.W;
    double py = pixel_units ? (double)g0.u[1] : (double)g0.u[1] * cam.H;
    return integrate_compat(sc, cam.ray(px, py), SampleSource(x.g.data()), max_depth, rr_start, p, rays);
}

// MetropolisSample::getPos, src/metropolissample.cpp:129-136
inline int mlt_pos(const Pss& x, int W, int H)
{
    int px = std::max(0, std::min((int)((double)x.g[0].u[0] * W), W - 1));
    int py = std::max(0, std::min((int)((double)x.g[0].u[1] * H), H - 1));
    return py * W + px;
}

// cv::norm of a Vec3f: the L2 norm in double
inline double norm2(Col c) { return std::sqrt((double)c.r * c.r + (double)c.g * c.g + (double)c.b * c.b); }

// largeStep (src/metropolissample.cpp:45-59) as fresh draws; smallStep
// (:86-97) perturbs x and y of every group and keeps z and w (the light
// index is only redrawn by a large step).
void mlt_propose(const Pss& cur, Pss& out, uint32_t seed, uint32_t chain, uint32_t step, bool large)
{
    for (size_t g = 0; g < cur.g.size(); ++g) {
        if (large) {
            out.g[g] = draw(seed, RT_MLT_STREAM_LARGE, chain, step, (uint32_t)g);
        } else {
            U4 m = draw(seed, RT_MLT_STREAM_SMALL, chain, step, (uint32_t)g);
            out.g[g] = cur.g[g];
            out.g[g].u[0] = mlt_perturb(cur.g[g].u[0], m.u[0], m.u[1]);
            out.g[g].u[1] = mlt_perturb(cur.g[g].u[1], m.u[2], m.u[3]);
        }
    }
}

// ERPT initial sample i = pixel * spp + s: the camera sample is the pixel
// plus the jitter of stratum (s / spp_y, s % spp_y) (JitteredSampler,
// src/jitteredsampler.cpp:23-35) over the image size
// (src/energyredistributionrenderer.cpp:66-75), the rest a large step.
void erpt_seed_sample(Pss& x, uint32_t seed, int64_t i, int W, int H, int sx, int sy)
{
    const int G = (int)x.g.size(), spp = sx * sy;
    for (int g = 0; g < G; ++g) x.g[g] = draw(seed, RT_MLT_STREAM_SEED, (uint32_t)i, 0, (uint32_t)g);
    const int64_t pix = i / spp;
    const int s = (int)(i % spp), a = s / sy, b = s % sy;
    double jx = ((double)x.g[0].u[0] + a) / sx, jy = ((double)x.g[0].u[1] + b) / sy;
    x.g[0].u[0] = (float)(((double)(pix % W) + jx) / W);
    x.g[0].u[1] = (float)(((double)(pix / W) + jy) / H);
}

struct MltRun {
    const Scene* sc;
    Camera cam;
    rt_mlt_params p;
    int G;
    std::vector<double> film;  // W * H * 3
    rt_mlt_result res{};
    int threads;
    MltRun(const Scene* s, const rt_camera& c, const rt_mlt_params& prm, int thr)
        : sc(s), cam(c), p(prm), G(2 * prm.max_depth + 1), film((size_t)c.width * c.height * 3, 0.0), threads(thr)
    {
    }
};

int64_t mlt_default_chains(int64_t total)
{
    return std::max<int64_t>(1, std::min<int64_t>(total / 16, (int64_t)1 << 22));
}

// MetropolisRenderer::render with the chains of the GPU: a bootstrap of
// max(bootstrap, chains) large-step paths at full length (the reference's
// bootstrap samples keep MetropolisSample's default pathLength = MAX_DEPTH,
// src/metropolissample.h:35), b = their mean luminance, each chain started
// from a bootstrap path drawn in proportion to luminance, then `iters`
// mutations per chain with both deposits of src/metropolisrenderer.cpp:160-175.
void run_metropolis(MltRun& R)
{
    const Scene& sc = *R.sc;
    const int W = R.cam.W, H = R.cam.H, G = R.G;
    const int64_t area = (int64_t)W * H;
    const int64_t total = (int64_t)R.p.mutations * area;
    const int64_t C = R.p.chains > 0 ? R.p.chains : mlt_default_chains(total);
    const int64_t iters = std::max<int64_t>(1, (total + C / 2) / C);
    const int64_t nb = std::max<int64_t>(R.p.bootstrap, C);
    const uint32_t seed = R.p.seed;
    std::vector<Col> bval(nb);
    std::vector<double> blum(nb);
    int64_t rays = 0;
#pragma omp parallel for schedule(dynamic, 64) num_threads(R.threads) reduction(+ : rays)
    for (int64_t i = 0; i < nb; ++i) {
        Pss x(G);
        for (int g = 0; g < G; ++g) x.g[g] = draw(seed, RT_MLT_STREAM_BOOTSTRAP, (uint32_t)i, 0, (uint32_t)g);
        bval[i] = integrate_pss(sc, R.cam, x, R.p.max_depth, R.p.rr_start, 0.f, false, rays);
        blum[i] = lum(bval[i]);
    }
    std::vector<double> cdf(nb);
    double acc = 0;
    for (int64_t i = 0; i < nb; ++i) cdf[i] = acc += blum[i];
    const float b = (float)(acc / (double)nb);
    R.res.b = b;
    R.res.chains = C;
    R.res.seed_paths = nb;
    const float scale = (float)((double)b * (double)area / ((double)C * (double)iters));
    const float lp = R.p.large_step_prob;
    int64_t accepted = 0;
    std::vector<double>& film = R.film;
#pragma omp parallel for schedule(dynamic, 16) num_threads(R.threads) reduction(+ : rays, accepted)
    for (int64_t c = 0; c < C; ++c) {
        double target = (double)draw(seed, RT_MLT_STREAM_START, (uint32_t)c, 0, 0).u[0] * acc;
        int64_t idx = std::upper_bound(cdf.begin(), cdf.end(), target) - cdf.begin();
        idx = std::min(idx, nb - 1);
        Pss cur(G), prop(G);
        for (int g = 0; g < G; ++g) cur.g[g] = draw(seed, RT_MLT_STREAM_BOOTSTRAP, (uint32_t)idx, 0, (uint32_t)g);
        Col cv = bval[idx];
        float lc = lum(cv);
        std::vector<std::pair<int, Col>> dep;
        dep.reserve(2 * iters);
        for (int64_t m = 0; m < iters; ++m) {
            U4 ctl = draw(seed, RT_MLT_STREAM_CONTROL, (uint32_t)c, (uint32_t)m, 0);
            mlt_propose(cur, prop, seed, (uint32_t)c, (uint32_t)m, ctl.u[0] < lp);
            prop.g[0].u[2] = ctl.u[1];  // mutatePathLength: a fresh length every mutation
            Col nv = integrate_pss(sc, R.cam, prop, R.p.max_depth, R.p.rr_start, R.p.termination_prob, false, rays);
            float ln = lum(nv);
            float a = lc > 0 ? std::min(1.f, ln / lc) : 1.f;
            if (lc > 0) dep.emplace_back(mlt_pos(cur, W, H), (cv * ((1.f / lc) * scale)) * (1 - a));
            if (ln > 0) dep.emplace_back(mlt_pos(prop, W, H), (nv * ((1.f / ln) * scale)) * a);
            if (ctl.u[2] < a) {
                std::swap(cur, prop);
                cv = nv;
                lc = ln;
                ++accepted;
            }
        }
#pragma omp critical
        for (auto& d : dep) {
            film[3 * (size_t)d.first] += d.second.r;
            film[3 * (size_t)d.first + 1] += d.second.g;
            film[3 * (size_t)d.first + 2] += d.second.b;
        }
    }
    R.res.mutations = C * iters;
    R.res.accepted = accepted;
    R.res.rays = rays;
}

// EnergyRedistributionRenderer::render.  e_d from one path per pixel centre
// with termination probability erpt_ed_termination (computeEd,
// src/energyredistributionrenderer.cpp:115-131); erpt_spp_x * erpt_spp_y
// jittered initial samples per pixel at full path length; floor(u +
// |c| / (M e_d)) chains each, which deposit e_d / chains per mutation
// (src/energyredistributionrenderer.cpp:84-112, with the reference's order:
// the deposit goes to the pixel of the state after the acceptance test and
// carries the colour of the state before it).
void run_erpt(MltRun& R)
{
    const Scene& sc = *R.sc;
    const int W = R.cam.W, H = R.cam.H, G = R.G;
    const int64_t area = (int64_t)W * H;
    const int M = R.p.mutations, sx = R.p.erpt_spp_x, sy = R.p.erpt_spp_y, spp = sx * sy;
    const uint32_t seed = R.p.seed;
    int64_t rays = 0;
    std::vector<Col> edv(area);
#pragma omp parallel for schedule(dynamic, 64) num_threads(R.threads) reduction(+ : rays)
    for (int64_t i = 0; i < area; ++i) {
        Pss x(G);
        for (int g = 0; g < G; ++g) x.g[g] = draw(seed, RT_MLT_STREAM_ED, (uint32_t)i, 0, (uint32_t)g);
        x.g[0].u[0] = (float)(i % W);
        x.g[0].u[1] = (float)(i / W);
        edv[i] = integrate_pss(sc, R.cam, x, R.p.max_depth, R.p.rr_start, R.p.erpt_ed_termination, true, rays);
    }
    double er = 0, eg = 0, eb = 0;
    for (int64_t i = 0; i < area; ++i) er += edv[i].r, eg += edv[i].g, eb += edv[i].b;
    const float ed = (float)(std::sqrt(er * er + eg * eg + eb * eb) / ((double)area * M));
    R.res.ed = ed;
    const int64_t n0 = area * spp;
    std::vector<Col> cval(n0);
    std::vector<int32_t> nch(n0);
#pragma omp parallel for schedule(dynamic, 64) num_threads(R.threads) reduction(+ : rays)
    for (int64_t i = 0; i < n0; ++i) {
        Pss x(G);
        erpt_seed_sample(x, seed, i, W, H, sx, sy);
        cval[i] = integrate_pss(sc, R.cam, x, R.p.max_depth, R.p.rr_start, 0.f, false, rays);
        double nr = norm2(cval[i]);
        nch[i] = ed > 0 ? (int32_t)((double)draw(seed, RT_MLT_STREAM_COUNT, (uint32_t)i, 0, 0).u[0] +
                                    nr / (double)(M * ed))
                        : 0;
    }
    std::vector<int64_t> first(n0 + 1, 0);
    for (int64_t i = 0; i < n0; ++i) first[i + 1] = first[i] + nch[i];
    const int64_t C = first[n0];
    R.res.chains = C;
    for (int64_t i = 0; i < n0; ++i) R.res.chain_seeds += nch[i] > 0;
    R.res.seed_paths = area + n0;
    const float inv_spp = 1.f / (float)spp;
    int64_t accepted = 0;
    std::vector<double>& film = R.film;
#pragma omp parallel for schedule(dynamic, 16) num_threads(R.threads) reduction(+ : rays, accepted)
    for (int64_t i = 0; i < n0; ++i) {
        if (nch[i] == 0) continue;
        Pss x0(G);
        erpt_seed_sample(x0, seed, i, W, H, sx, sy);
        const Col c0 = cval[i];
        const double n0r = norm2(c0);
        const double k = (double)ed / n0r / nch[i];
        const Col dv((float)(c0.r * k), (float)(c0.g * k), (float)(c0.b * k));
        const double E = norm2(dv);
        std::vector<std::pair<int, Col>> dep;
        dep.reserve((size_t)M * nch[i]);
        for (int64_t ch = first[i]; ch < first[i + 1]; ++ch) {
            Pss y = x0, z(G);
            Col yv = c0;
            float yl = (float)n0r;
            for (int m = 0; m < M; ++m) {
                mlt_propose(y, z, seed, (uint32_t)ch, (uint32_t)m, false);
                Col zv = integrate_pss(sc, R.cam, z, R.p.max_depth, R.p.rr_start, 0.f, false, rays);
                float zl = (float)norm2(zv);
                float q = std::min(1.f, zl / yl);
                const Col before = yv;
                const double inv_before = 1.0 / (double)yl;
                if (q > draw(seed, RT_MLT_STREAM_CONTROL, (uint32_t)ch, (uint32_t)m, 0).u[2]) {
                    std::swap(y, z);
                    yv = zv;
                    yl = zl;
                    ++accepted;
                }
                // norm(depVal) * yVal * (1 / norm(yVal)) * (1.f / spp), each product rounded to float
                Col d((float)(E * before.r), (float)(E * before.g), (float)(E * before.b));
                d = Col((float)(d.r * inv_before), (float)(d.g * inv_before), (float)(d.b * inv_before));
                d = d * inv_spp;
                dep.emplace_back(mlt_pos(y, W, H), d);
            }
        }
#pragma omp critical
        for (auto& d : dep) {
            film[3 * (size_t)d.first] += d.second.r;
            film[3 * (size_t)d.first + 1] += d.second.g;
            film[3 * (size_t)d.first + 2] += d.second.b;
        }
    }
    R.res.mutations = C * M;
    R.res.accepted = accepted;
    R.res.rays = rays;
}

// ---------------------------------------------------------------------------
// MetropolisFltRenderer (src/metropolisfltrenderer.cpp:60-320) and the
// symmetric filter (src/symmetricfilter.cpp:40-137), on images of H x W x C
// floats.  OpenCV's blur (a normalised box filter, BORDER_REFLECT) and
// element-wise Mat arithmetic are restated as loops; their float summation
// order is this restatement's own.
// ---------------------------------------------------------------------------
struct Img {
    int W = 0, H = 0, C = 1;
    std::vector<float> v;
    Img() = default;
    Img(int w, int h, int c, float x = 0) : W(w), H(h), C(c), v((size_t)w * h * c, x) {}
    float& at(int y, int x, int c = 0) { return v[((size_t)y * W + x) * C + c]; }
    float at(int y, int x, int c = 0) const { return v[((size_t)y * W + x) * C + c]; }
};

// mathhelper var() of two images: mean = (a + b) / 2, var = max((a^2 + b^2 - sum mean) / 2, 0)
void var2(const Img& a, const Img& b, Img& mean, Img& var)
{
    mean = Img(a.W, a.H, a.C);
    var = Img(a.W, a.H, a.C);
    for (size_t i = 0; i < a.v.size(); ++i) {
        float sum = a.v[i] + b.v[i], sq = a.v[i] * a.v[i] + b.v[i] * b.v[i];
        float m = sum * 0.5f;
        mean.v[i] = m;
        var.v[i] = std::max((sq - sum * m) / 2.f, 0.f);
    }
}

inline float chan_mean3(float a, float b, float c) { return (a / 3.f + b / 3.f) + c / 3.f; }

// cv::blur(.., Size(p, p), Point(-1, -1), BORDER_REFLECT) of a one-channel image
Img box_blur(const Img& in, int p)
{
    if (p <= 1) return in;
    const int r = p / 2, W = in.W, H = in.H;
    auto refl = [](int i, int n) {
        while (i < 0 || i >= n) i = i < 0 ? -i - 1 : 2 * n - i - 1;
        return i;
    };
    Img t(W, H, 1), out(W, H, 1);
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            float s = 0;
            for (int k = -r; k <= r; ++k) s += in.at(y, refl(x + k, W));
            t.at(y, x) = s;
        }
    const float inv = 1.f / (float)(p * p);
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            float s = 0;
            for (int k = -r; k <= r; ++k) s += t.at(refl(y + k, H), x);
            out.at(y, x) = s * inv;
        }
    return out;
}

// SymmetricFilter::filter (src/symmetricfilter.cpp:72-137) with computeWeights
// (:46-70): pairs of offsets +-d (half the window; d = 0 once), weights from
// the patch distance of the guide less the variance, the symmetric average
// when it beats both one-sided weights; outside the image everything is 0.
Img sym_filter(const Img& image, const Img& guide, const Img& pixvar, int patch, int R)
{
    const int W = image.W, H = image.H;
    const float h2 = 0.4f * 0.4f;
    Img area(W, H, 1), acc(W, H, 3);
    auto get = [&](const Img& m, int y, int x, int c) { return y < 0 || y >= H || x < 0 || x >= W ? 0.f : m.at(y, x, c); };
    for (int dx = -R; dx <= 0; ++dx)
        for (int dy = -R; dy <= R; ++dy) {
            if (dx == 0 && dy == 0) {
                for (int y = 0; y < H; ++y)
                    for (int x = 0; x < W; ++x) {
                        area.at(y, x) += 1.f;
                        for (int c = 0; c < 3; ++c) acc.at(y, x, c) += image.at(y, x, c);
                    }
                break;
            }
            Img d2[3] = {Img(W, H, 1), Img(W, H, 1), Img(W, H, 1)}, den[3] = {Img(W, H, 1), Img(W, H, 1), Img(W, H, 1)};
            for (int y = 0; y < H; ++y)
                for (int x = 0; x < W; ++x) {
                    float dd[3][3], vv[3][3];
                    for (int c = 0; c < 3; ++c) {
                        float s1 = get(guide, y - dy, x - dx, c), s2 = get(guide, y + dy, x + dx, c);
                        float v1 = get(pixvar, y - dy, x - dx, c), v2 = get(pixvar, y + dy, x + dx, c);
                        float sS = (s1 + s2) / 2.f, vS = (v1 + v2) / 4.f;
                        float t = guide.at(y, x, c), v = pixvar.at(y, x, c);
                        const float src[3] = {s1, s2, sS}, sv[3] = {v1, v2, vS};
                        for (int k = 0; k < 3; ++k) {
                            float df = src[k] - t;
                            dd[k][c] = df * df - (v + std::min(sv[k], v));
                            vv[k][c] = v + sv[k];
                        }
                    }
                    for (int k = 0; k < 3; ++k) {
                        d2[k].at(y, x) = chan_mean3(dd[k][0], dd[k][1], dd[k][2]);
                        den[k].at(y, x) = 1e-8f + h2 * chan_mean3(vv[k][0], vv[k][1], vv[k][2]);
                    }
                }
            Img w[3];
            for (int k = 0; k < 3; ++k) {
                Img b = box_blur(d2[k], patch), raw(W, H, 1);
                for (size_t i = 0; i < raw.v.size(); ++i) raw.v[i] = std::exp(-std::max(b.v[i], 0.f) / den[k].v[i]);
                w[k] = box_blur(raw, patch);
                for (float& x : w[k].v) {
                    x = std::max(x, 0.f);
                    if (x < 0.05f) x = 0.f;
                }
            }
            for (int y = 0; y < H; ++y)
                for (int x = 0; x < W; ++x) {
                    float w1 = w[0].at(y, x), w2 = w[1].at(y, x), wS = w[2].at(y, x);
                    if (wS > std::max(w1, w2)) w1 = w2 = wS;
                    area.at(y, x) += w1 + w2;
                    for (int c = 0; c < 3; ++c)
                        acc.at(y, x, c) += w1 * get(image, y - dy, x - dx, c) + w2 * get(image, y + dy, x + dx, c);
                }
        }
    Img out(W, H, 3);
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x)
            for (int c = 0; c < 3; ++c) out.at(y, x, c) = acc.at(y, x, c) / area.at(y, x);
    return out;
}

// The chains of one pass of MetropolisFltRenderer::renderStep
// (src/metropolisfltrenderer.cpp:86-205) with the GPU's chain layout: per
// film, `chains` chains started from bootstrap draws in proportion to
// norm x importance, norm-based acceptance weighted by the importance map,
// deposits weighted by 1 / (summed importance) into the film and the
// weighted running mean / variance of addSample (:54-75).
struct FltState {
    Img film[2], mean[2], m2[2], sw[2];
};

// addSample (src/metropolisfltrenderer.cpp:54-75) of film n at pixel pos:
// the weighted running mean and sum of weighted squared deviations (West's
// update), the weighted sum of the values and the sum of the weights
void flt_add_sample(FltState& st, int n, int pos, float w, const float vc[3])
{
    const int W = st.film[n].W;
    const float ow = st.sw[n].v[pos], nw = w + ow;
    for (int c = 0; c < 3; ++c) {
        float delta = vc[c] - st.mean[n].at(pos / W, pos % W, c);
        float Rr = delta * (w / nw);
        st.mean[n].at(pos / W, pos % W, c) += Rr;
        st.m2[n].at(pos / W, pos % W, c) += ow * (delta * Rr);
        st.film[n].at(pos / W, pos % W, c) += w * vc[c];
    }
    st.sw[n].v[pos] = nw;
}

void flt_pass(MltRun& R, int pass, const Img& imp, const Img& newsum, float b, FltState& st, int64_t& rays,
              int64_t& accepted, int64_t& mutations)
{
    const Scene& sc = *R.sc;
    const int W = R.cam.W, H = R.cam.H, G = R.G, P = R.p.flt_passes;
    const int64_t area = (int64_t)W * H;
    const int64_t per_film = std::max<int64_t>(1, (int64_t)R.p.mutations * area / P / 2);
    const int64_t C = R.p.chains > 0 ? R.p.chains : mlt_default_chains(per_film);
    const int64_t iters = std::max<int64_t>(1, (per_film + C / 2) / C);
    const int64_t nb = std::max<int64_t>(R.p.bootstrap, C);
    const uint32_t seed = R.p.seed;
    // bootstrap of this pass: full-length large steps, weighted by norm x importance
    std::vector<Col> bval(nb);
    std::vector<double> cdf(nb);
#pragma omp parallel for schedule(dynamic, 64) num_threads(R.threads) reduction(+ : rays)
    for (int64_t i = 0; i < nb; ++i) {
        Pss x(G);
        for (int g = 0; g < G; ++g) x.g[g] = draw(seed, RT_MLT_STREAM_BOOTSTRAP, (uint32_t)i, (uint32_t)pass, (uint32_t)g);
        bval[i] = integrate_pss(sc, R.cam, x, R.p.max_depth, R.p.rr_start, 0.f, false, rays);
        const int pos = mlt_pos(x, W, H);
        cdf[i] = (float)(norm2(bval[i]) * imp.v[pos]);
    }
    double acc = 0;
    for (int64_t i = 0; i < nb; ++i) cdf[i] = acc += cdf[i];
    R.res.chains += 2 * C;
    R.res.seed_paths += nb;
    if (!(acc > 0)) return;
    // b * area / 2 per film and pass, over the C * iters mutations that run
    // (src/metropolisfltrenderer.cpp:139,158-165 aim at that total with
    // b * passes / mutations-per-pixel per deposit)
    const float scale = (float)((double)b * (double)area / (2.0 * (double)C * (double)iters));
    for (int n = 0; n < 2; ++n) {
        std::vector<std::vector<std::pair<int, std::pair<float, Col>>>> deps(C);
#pragma omp parallel for schedule(dynamic, 16) num_threads(R.threads) reduction(+ : rays, accepted)
        for (int64_t c = 0; c < C; ++c) {
            const uint32_t id = (uint32_t)(((int64_t)pass * 2 + n) * C + c);
            double target = (double)draw(seed, RT_MLT_STREAM_START, id, 0, 0).u[0] * acc;
            int64_t idx = std::min<int64_t>(std::upper_bound(cdf.begin(), cdf.end(), target) - cdf.begin(), nb - 1);
            Pss cur(G), prop(G);
            for (int g = 0; g < G; ++g)
                cur.g[g] = draw(seed, RT_MLT_STREAM_BOOTSTRAP, (uint32_t)idx, (uint32_t)pass, (uint32_t)g);
            Col cv = bval[idx];
            float cn = (float)norm2(cv);
            auto& dep = deps[c];
            for (int64_t m = 0; m < iters; ++m) {
                U4 ctl = draw(seed, RT_MLT_STREAM_CONTROL, id, (uint32_t)m, 0);
                mlt_propose(cur, prop, seed, id, (uint32_t)m, ctl.u[0] < R.p.large_step_prob);
                prop.g[0].u[2] = ctl.u[1];
                Col nv = integrate_pss(sc, R.cam, prop, R.p.max_depth, R.p.rr_start, R.p.termination_prob, false, rays);
                double zn = norm2(nv);
                const int cp = mlt_pos(cur, W, H), np = mlt_pos(prop, W, H);
                float a = (float)std::min(1., zn * imp.v[np] / ((double)cn * imp.v[cp]));
                if (cn > 0 && a < 1) {
                    const double k = 1 / (double)cn * scale;
                    dep.push_back({cp, {(1 - a) / newsum.v[cp], Col((float)(cv.r * k), (float)(cv.g * k), (float)(cv.b * k))}});
                }
                if (zn > 0 && a > 0) {
                    const double k = 1 / zn * scale;
                    dep.push_back({np, {a / newsum.v[np], Col((float)(nv.r * k), (float)(nv.g * k), (float)(nv.b * k))}});
                }
                if (ctl.u[2] < a) {
                    std::swap(cur, prop);
                    cv = nv;
                    cn = (float)zn;
                    ++accepted;
                }
            }
        }
        // addSample, in chain order
        for (auto& dep : deps)
            for (auto& d : dep) {
                const Col v = d.second.second;
                const float vc[3] = {v.r, v.g, v.b};
                flt_add_sample(st, n, d.first, d.second.first, vc);
            }
        mutations += C * iters;
    }
}

// MetropolisFltRenderer::render (src/metropolisfltrenderer.cpp:207-320)
// relative variance of the filtered estimate -> importance (:231-239),
// normalised to mean 1
Img flt_importance(const Img& varOfFiltered, const Img& filteredMean)
{
    const int W = varOfFiltered.W, H = varOfFiltered.H;
    const int64_t area = (int64_t)W * H;
    Img imp(W, H, 1, 1.f);
    double sum = 0;
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            float vm = chan_mean3(varOfFiltered.at(y, x, 0), varOfFiltered.at(y, x, 1), varOfFiltered.at(y, x, 2));
            float mm = chan_mean3(filteredMean.at(y, x, 0), filteredMean.at(y, x, 1), filteredMean.at(y, x, 2));
            float rel = vm / (mm * mm + 1e-2f) + 1e-8f;
            float e = std::exp(-rel);
            imp.at(y, x) = -20 * (e - 1) + 1;
            sum += imp.at(y, x);
        }
    const float k = (float)(area / sum);
    for (float& x : imp.v) x *= k;
    return imp;
}

// the running sums of later passes: each sample so far weighted by r, the
// share of the importance it was deposited under (:241-250)
void flt_rescale(FltState& st, const Img& sumimp, const Img& newsum)
{
    const int64_t area = (int64_t)sumimp.W * sumimp.H;
    for (int n = 0; n < 2; ++n)
        for (int64_t k = 0; k < area; ++k) {
            const float r = sumimp.v[k] / newsum.v[k];
            for (int c = 0; c < 3; ++c) {
                st.film[n].v[3 * k + c] = r * st.film[n].v[3 * k + c];
                st.m2[n].v[3 * k + c] = r * (r * st.m2[n].v[3 * k + c]);
            }
        }
}

// sepFilter2D(.., [.1 .8 .1], [.1 .8 .1]) with BORDER_REFLECT_101 of a
// three-channel image (the FLT_SIMPLE_VARIANCE_FILTER path, :270-272)
Img sep3_reflect101(const Img& in)
{
    const int W = in.W, H = in.H;
    auto r101 = [](int i, int n) { return i < 0 ? -i : i >= n ? 2 * n - i - 2 : i; };
    Img t(W, H, 3), out(W, H, 3);
    const float k3[3] = {.1f, .8f, .1f};
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x)
            for (int c = 0; c < 3; ++c) {
                float s = 0;
                for (int j = -1; j <= 1; ++j) s += k3[j + 1] * in.at(y, std::min(std::max(r101(x + j, W), 0), W - 1), c);
                t.at(y, x, c) = s;
            }
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x)
            for (int c = 0; c < 3; ++c) {
                float s = 0;
                for (int j = -1; j <= 1; ++j) s += k3[j + 1] * t.at(std::min(std::max(r101(y + j, H), 0), H - 1), x, c);
                out.at(y, x, c) = s;
            }
    return out;
}

void run_filtered(MltRun& R)
{
    const int W = R.cam.W, H = R.cam.H, P = R.p.flt_passes, F = R.p.flt_flags;
    const int64_t area = (int64_t)W * H;
    FltState st;
    for (int n = 0; n < 2; ++n) {
        st.film[n] = Img(W, H, 3);
        st.mean[n] = Img(W, H, 3);
        st.m2[n] = Img(W, H, 3);
        st.sw[n] = Img(W, H, 1);
    }
    Img sumimp(W, H, 1), varOfFiltered, filteredMean;
    int64_t rays = 0, accepted = 0, mutations = 0;
    for (int i = 0; i < P; ++i) {
        const Img imp = i > 0 ? flt_importance(varOfFiltered, filteredMean) : Img(W, H, 1, 1.f);
        // b: the bootstrap's mean norm (first pass) or the first film's (:118-124)
        float b;
        if (i == 0) {
            const int64_t per_film = std::max<int64_t>(1, (int64_t)R.p.mutations * area / P / 2);
            const int64_t C = R.p.chains > 0 ? R.p.chains : mlt_default_chains(per_film);
            const int64_t nb = std::max<int64_t>(R.p.bootstrap, C);
            // the same paths as the pass's bootstrap, whose rays flt_pass counts
            double acc = 0;
            int64_t unused_rays = 0;
#pragma omp parallel for schedule(dynamic, 64) num_threads(R.threads) reduction(+ : unused_rays, acc)
            for (int64_t j = 0; j < nb; ++j) {
                Pss x(R.G);
                for (int g = 0; g < R.G; ++g) x.g[g] = draw(R.p.seed, RT_MLT_STREAM_BOOTSTRAP, (uint32_t)j, 0, (uint32_t)g);
                acc += norm2(integrate_pss(*R.sc, R.cam, x, R.p.max_depth, R.p.rr_start, 0.f, false, unused_rays));
            }
            b = (float)(acc / (double)nb);
        } else {
            double acc = 0;
            for (int y = 0; y < H; ++y)
                for (int x = 0; x < W; ++x) {
                    float r = st.film[0].at(y, x, 0), g = st.film[0].at(y, x, 1), bb = st.film[0].at(y, x, 2);
                    acc += (double)(std::sqrt(r * r + g * g + bb * bb) * imp.at(y, x));
                }
            b = (float)(acc / (double)area);
        }
        if (i == 0) R.res.b = b;
        Img newsum(W, H, 1);
        for (size_t k = 0; k < newsum.v.size(); ++k) newsum.v[k] = sumimp.v[k] + imp.v[k];
        flt_rescale(st, sumimp, newsum);
        flt_pass(R, i, imp, newsum, b, st, rays, accepted, mutations);
        // filtering (:252-300)
        Img bv[2];
        for (int n = 0; n < 2; ++n) {
            bv[n] = Img(W, H, 3);
            for (int64_t k = 0; k < area; ++k)
                for (int c = 0; c < 3; ++c) bv[n].v[3 * k + c] = st.m2[n].v[3 * k + c] / (st.sw[n].v[k] + 1e-8f);
        }
        Img newOut, noisy, meanvar, varvar, filteredVar;
        var2(st.film[0], st.film[1], newOut, noisy);
        var2(bv[0], bv[1], meanvar, varvar);
        if ((F & RT_FLT_NO_FILTER) || (F & RT_FLT_NO_VARIANCE_FILTER)) {
            filteredVar = noisy;
        } else if (F & RT_FLT_SIMPLE_VARIANCE_FILTER) {
            filteredVar = sep3_reflect101(noisy);
        } else {
            filteredVar = sym_filter(noisy, meanvar, varvar, 1, 2);
        }
        if (F & RT_FLT_NO_FILTER) {
            varOfFiltered = filteredVar;
            filteredMean = newOut;
        } else if (F & RT_FLT_NO_CROSSFILTERING) {
            varOfFiltered = filteredVar;
            filteredMean = sym_filter(newOut, newOut, filteredVar, 7, 10);
        } else {
            Img f0 = sym_filter(st.film[0], st.film[1], filteredVar, 7, 10);
            Img f1 = sym_filter(st.film[1], st.film[0], filteredVar, 7, 10);
            var2(f0, f1, filteredMean, varOfFiltered);
        }
        sumimp = newsum;
    }
    for (size_t k = 0; k < R.film.size(); ++k) R.film[k] = filteredMean.v[k];
    R.res.mutations = mutations;
    R.res.accepted = accepted;
    R.res.rays = rays;
}

// ---------------------------------------------------------------------------
// C API
// ---------------------------------------------------------------------------
extern "C" {

void oc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    philox(ctr, key, out);
}

void* oc_scene_create(const rt_scene_desc* d, int accel)
{
    auto* sc = new Scene();
    std::vector<Material*> mats;
    for (int i = 0; i < d->n_materials; ++i) {
        const rt_material& m = d->materials[i];
        Material* mm;
        switch (m.type) {
        case RT_MAT_DIFFUSE: mm = new Diffuse(Col(m.kd[0], m.kd[1], m.kd[2])); break;
        case RT_MAT_MIRROR: mm = new Mirror(); break;
        case RT_MAT_PHONG:
            mm = new Phong(Col(m.kd[0], m.kd[1], m.kd[2]), Col(m.ks[0], m.ks[1], m.ks[2]), m.exponent);
            break;
        case RT_MAT_BLINN:
            mm = new Blinn(Col(m.kd[0], m.kd[1], m.kd[2]), Col(m.ks[0], m.ks[1], m.ks[2]), m.exponent);
            break;
        case RT_MAT_REFRACTIVE: mm = new Transparent(m.exponent); break;
        case RT_MAT_TEXTURE: {
            auto* tm = new TextureMat();
            const rt_texture& t = d->textures[m.texture];
            tm->W = t.width;
            tm->H = t.height;
            tm->rgb.assign(t.rgb, t.rgb + (size_t)3 * t.width * t.height);
            tm->coef = Col(m.kd[0], m.kd[1], m.kd[2]);
            tm->gamma = m.exponent;
            mm = tm;
            break;
        }
        default: mm = new DarkMatter(); break;
        }
        sc->mats.emplace_back(mm);
        mats.push_back(mm);
    }
    int max_prim = d->n_lights - 1;
    for (int i = 0; i < d->n_triangles; ++i) max_prim = std::max(max_prim, d->tri_prim_id[i]);
    for (int i = 0; i < d->n_spheres; ++i) max_prim = std::max(max_prim, d->sphere_prim_id[i]);
    for (int i = 0; i < d->n_objects; ++i) max_prim = std::max(max_prim, d->object_prim_id[i]);
    sc->prims.assign(max_prim + 1, nullptr);

    std::vector<Intersectable*> root_items;
    for (int i = 0; i < d->n_lights; ++i) {
        const rt_light& l = d->lights[i];
        V3 p(l.p[0], l.p[1], l.p[2]), u(l.u[0], l.u[1], l.u[2]), v(l.v[0], l.v[1], l.v[2]);
        Col I(l.intensity[0], l.intensity[1], l.intensity[2]);
        if (l.type == RT_LIGHT_AREA) {
            auto* al = new AreaLight(p, u, v, I, i);
            sc->owned.emplace_back(al);
            sc->lights.push_back(al);
            sc->prims[i] = al;
            root_items.push_back(al);
        } else if (l.type == RT_LIGHT_POINT) {
            auto* pl = new PointLight(p, I);
            sc->owned.emplace_back(pl);
            sc->lights.push_back(pl);
            root_items.push_back(pl);
        } else if (l.type == RT_LIGHT_CONE) {
            auto* cl = new ConeLight(p, u, l.v[0], I);
            sc->owned.emplace_back(cl);
            sc->lights.push_back(cl);
            root_items.push_back(cl);
        } else {
            auto* em = new EnvMap();
            if (d->env_kind == 1) {
                em->set_probe(d->env_rgb, d->env_width, d->env_height);
            } else {
                em->W = d->env_width;
                em->H = d->env_height;
                em->rgb.assign(d->env_rgb, d->env_rgb + (size_t)3 * em->W * em->H);
            }
            em->coef = I;
            em->prim = i;
            em->build_sampling();
            sc->mats.emplace_back(em);
            sc->lights.push_back(em);
            sc->env = em;
        }
    }
    auto vert = [d](int idx) {
        return V3(d->positions[3 * idx], d->positions[3 * idx + 1], d->positions[3 * idx + 2]);
    };
    auto nrm = [d](int idx) {
        return V3(d->normals[3 * idx], d->normals[3 * idx + 1], d->normals[3 * idx + 2]);
    };
    auto make_tri = [&](int i) {
        const int32_t* tv = d->tri_vertices + 3 * i;
        Triangle t(vert(tv[0]), vert(tv[1]), vert(tv[2]), mats[d->tri_material[i]], d->tri_prim_id[i]);
        if (d->normals) t.set_normals(nrm(tv[0]), nrm(tv[1]), nrm(tv[2]));
        if (d->tri_texcoords) {
            const float* q = d->tri_texcoords + 6 * (size_t)i;
            t.t1[0] = q[0], t.t1[1] = q[1], t.t2[0] = q[2], t.t2[1] = q[3], t.t3[0] = q[4], t.t3[1] = q[5];
        }
        return t;
    };
    // Components in primitive-ID order; two consecutive triangles with one ID form a quad.
    struct Item { int id; Intersectable* obj; };
    std::vector<Item> items;
    for (int i = 0; i < d->n_triangles;) {
        int id = d->tri_prim_id[i];
        if (i + 1 < d->n_triangles && d->tri_prim_id[i + 1] == id) {
            auto* q = new Quad(make_tri(i), make_tri(i + 1));
            sc->owned.emplace_back(q);
            items.push_back({id, q});
            i += 2;
        } else {
            auto* t = new Triangle(make_tri(i));
            sc->owned.emplace_back(t);
            items.push_back({id, t});
            i += 1;
        }
    }
    for (int i = 0; i < d->n_spheres; ++i) {
        const float* s = d->spheres + 4 * i;
        auto* sp = new Sphere(V3(s[0], s[1], s[2]), s[3], mats[d->sphere_material[i]], d->sphere_prim_id[i]);
        sc->owned.emplace_back(sp);
        items.push_back({d->sphere_prim_id[i], sp});
    }
    // analytic objects from their postfix programs
    std::vector<Intersectable*> planes;
    for (int i = 0; i < d->n_objects; ++i) {
        const int f = d->object_first[i], e = d->object_first[i + 1];
        const int id = d->object_prim_id[i];
        if (e - f == 1 && d->csg_nodes[f].op == RT_CSG_PLANE) {
            auto* pl = new Plane(d->csg_nodes[f].p, mats[d->csg_nodes[f].material], id);
            sc->owned.emplace_back(pl);
            sc->prims[id] = pl;
            planes.push_back(pl);
            continue;
        }
        std::vector<std::unique_ptr<CsgNode>> st;
        for (int k = f; k < e; ++k) {
            const rt_csg_node& n = d->csg_nodes[k];
            if (n.op == RT_CSG_SPHERE) {
                st.emplace_back(new CsgSphere(V3(n.p[0], n.p[1], n.p[2]), n.p[3], mats[n.material]));
            } else if (n.op == RT_CSG_BOX) {
                st.emplace_back(new CsgBox(V3(n.p[0], n.p[1], n.p[2]), V3(n.p[3], n.p[4], n.p[5]), mats[n.material]));
            } else {
                auto r = std::move(st.back());
                st.pop_back();
                auto l = std::move(st.back());
                st.pop_back();
                st.emplace_back(new CsgOp(n.op, std::move(l), std::move(r)));
            }
        }
        auto* obj = new CsgObject(std::move(st.back()), id);
        sc->owned.emplace_back(obj);
        static const float identity[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
        if (d->object_xform && !std::equal(identity, identity + 12, d->object_xform + 12 * (size_t)i)) {
            auto* inst = new Instance(obj, d->object_xform + 12 * (size_t)i);
            sc->owned.emplace_back(inst);
            items.push_back({id, inst});
        } else {
            items.push_back({id, obj});
        }
    }
    std::stable_sort(items.begin(), items.end(), [](const Item& a, const Item& b) { return a.id < b.id; });
    for (auto& it : items) {
        sc->prims[it.id] = it.obj;
        root_items.push_back(it.obj);
    }
    if (root_items.empty()) {
        sc->root = nullptr;
    } else if (accel == 0 || root_items.size() == 1) {
        auto* l = new List();
        l->items = root_items;
        sc->owned.emplace_back(l);
        sc->root = l;
    } else {
        sc->root = build_bvh(root_items, sc->owned);
    }
    // planes have infinite bounds: tested after the rest, in a list (ties
    // with the rest go to the plane)
    if (!planes.empty()) {
        auto* l = new List();
        if (sc->root) l->items.push_back(sc->root);
        for (auto* p : planes) l->items.push_back(p);
        sc->owned.emplace_back(l);
        sc->root = l;
    }
    return sc;
}

void oc_scene_destroy(void* s) { delete static_cast<Scene*>(s); }

static inline Ray ray_from(const float* o, const float* d, int64_t i)
{
    return Ray(V3(o[4 * i], o[4 * i + 1], o[4 * i + 2]), V3(d[4 * i], d[4 * i + 1], d[4 * i + 2]), o[4 * i + 3],
               d[4 * i + 3]);
}

void oc_trace_closest(void* s, int64_t n, const float* o, const float* d, int32_t* prim, float* t, int threads)
{
    const Scene& sc = *static_cast<Scene*>(s);
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads > 0 ? threads : 1)
    for (int64_t i = 0; i < n; ++i) {
        Hit h = sc.root ? sc.root->intersect(ray_from(o, d, i)) : Hit();
        prim[i] = h.hits() ? h.prim : -1;
        t[i] = h.t;
    }
}

void oc_trace_any(void* s, int64_t n, const float* o, const float* d, uint8_t* occ, int threads)
{
    const Scene& sc = *static_cast<Scene*>(s);
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads > 0 ? threads : 1)
    for (int64_t i = 0; i < n; ++i) {
        Hit h = sc.root ? sc.root->intersect(ray_from(o, d, i)) : Hit();
        occ[i] = h.hits() ? 1 : 0;
    }
}

void oc_hit_texcoords(void* s, int64_t n, const float* o, const float* d, float* tc)
{
    const Scene& sc = *static_cast<Scene*>(s);
    for (int64_t i = 0; i < n; ++i) {
        Hit h = sc.root ? sc.root->intersect(ray_from(o, d, i)) : Hit();
        tc[2 * i] = h.hits() ? h.tc[0] : kNaN;
        tc[2 * i + 1] = h.hits() ? h.tc[1] : kNaN;
    }
}

int oc_texture_albedo(void* s, int mat, float ts, float tt, float rgb[3])
{
    const Scene& sc = *static_cast<Scene*>(s);
    if (mat < 0 || mat >= (int)sc.mats.size() || sc.mats[mat]->kind() != RT_MAT_TEXTURE) return -1;
    Hit h;
    h.tc[0] = ts;
    h.tc[1] = tt;
    Col c = static_cast<const TextureMat*>(sc.mats[mat].get())->albedo(h);
    rgb[0] = c.r, rgb[1] = c.g, rgb[2] = c.b;
    return 0;
}

void oc_hit_normals(void* s, int64_t n, const float* o, const float* d, float* normals)
{
    const Scene& sc = *static_cast<Scene*>(s);
    for (int64_t i = 0; i < n; ++i) {
        Hit h = sc.root ? sc.root->intersect(ray_from(o, d, i)) : Hit();
        V3 v = h.hits() ? h.n : V3();
        normals[3 * i] = v.x, normals[3 * i + 1] = v.y, normals[3 * i + 2] = v.z;
    }
}

float oc_intersect_prim(void* s, int32_t prim, const float* o, const float* d)
{
    const Scene& sc = *static_cast<Scene*>(s);
    if (prim < 0 || prim >= (int)sc.prims.size() || !sc.prims[prim]) return kNaN;
    return sc.prims[prim]->intersect(ray_from(o, d, 0)).t;
}

// UniDiPathTracingIntegrator::integrate(const Ray&, ...),
// src/unidipathtracingintegrator.cpp:34-46, for caller rays: ray i draws the
// Philox stream (seed, view 0, index key0 + i, sample 0).  The unbiased
// estimator normalises the direction (Qt: length in double, components
// rounded to float) and scales the range by that length.
void oc_integrate_rays(void* s, int64_t n, const float* o, const float* d, const rt_params* p, uint32_t key0,
                       float* rgb, int threads)
{
    const Scene& sc = *static_cast<Scene*>(s);
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads > 0 ? threads : 1)
    for (int64_t i = 0; i < n; ++i) {
        Ray ray = ray_from(o, d, i);
        SampleKey k{p->seed, 0u, key0 + (uint32_t)i, 0u};
        int64_t rays = 0;
        Col c;
        if (p->mode == RT_MODE_COMPAT) {
            c = integrate_compat(sc, ray, k, p->max_depth, p->rr_start, p->rr_prob, rays);
        } else {
            const double l = std::sqrt((double)ray.d.x * ray.d.x + (double)ray.d.y * ray.d.y + (double)ray.d.z * ray.d.z);
            if (l > 0) {
                ray.d = V3((float)(ray.d.x / l), (float)(ray.d.y / l), (float)(ray.d.z / l));
                ray.from = (float)(ray.from * l);
                if (!std::isinf(ray.to)) ray.to = (float)(ray.to * l);
            }
            UnbiasedCtx ctx{&sc, p->max_depth, p->rr_start, p->rr_prob};
            c = integrate_unbiased(ctx, ray, k, rays);
        }
        rgb[3 * i] = c.r, rgb[3 * i + 1] = c.g, rgb[3 * i + 2] = c.b;
    }
}

static Col render_one(const Scene& sc, const Camera& cam, const rt_view* view, const rt_params* p,
                      uint32_t view_id, int px, int py, int s, int64_t& rays)
{
    const int ny = view->spp_y;
    SampleKey k{p->seed, view_id, (uint32_t)(py * cam.W + px), (uint32_t)s};
    U4 g0 = philox_uniforms(k.seed, k.view, k.pixel, k.sample, 0);
    // src/jitteredsampler.cpp:23-35: stratum (a, b) with a = s / ny, b = s % ny
    int a = s / ny, b = s % ny;
    double sx = ((double)g0.u[0] + a) / view->spp_x, sy = ((double)g0.u[1] + b) / ny;
    // src/perpixelrenderer.cpp:79-84: point = pixel + jitter, camera adds 0.5
    Ray ray = cam.ray(px + sx, py + sy);
    if (p->mode == RT_MODE_COMPAT)
        return integrate_compat(sc, ray, k, p->max_depth, p->rr_start, p->rr_prob, rays);
    ray.d = normalized(ray.d);
    UnbiasedCtx c{&sc, p->max_depth, p->rr_start, p->rr_prob};
    return integrate_unbiased(c, ray, k, rays);
}

int64_t oc_render(void* s, const rt_view* view, const rt_params* p, uint32_t view_id, float* film, int row0,
                  int row1, int threads)
{
    const Scene& sc = *static_cast<Scene*>(s);
    Camera cam(view->camera);
    const int spp = view->spp_x * view->spp_y;
    int64_t total_rays = 0;
    Col sum;
    // src/perpixelrenderer.cpp:62-98: rows in parallel, film[i][j] += s / spp,
    // a critical section per sample for the luminance sum.
#pragma omp parallel for schedule(dynamic) num_threads(threads > 0 ? threads : 1) reduction(+ : total_rays)
    for (int i = row0; i < row1; ++i) {
        for (int j = 0; j < cam.W; ++j) {
            float* px = film + 3 * ((size_t)i * cam.W + j);
            for (int si = 0; si < spp; ++si) {
                int64_t rays = 0;
                Col c = render_one(sc, cam, view, p, view_id, j, i, si, rays);
                total_rays += rays;
                Col w = c * (1.f / spp);
                px[0] += w.r;
                px[1] += w.g;
                px[2] += w.b;
#pragma omp critical
                sum = sum + w;
            }
        }
    }
    return total_rays;
}

void oc_render_sample(void* s, const rt_view* view, const rt_params* p, uint32_t view_id, int px, int py, int si,
                      float out[3])
{
    const Scene& sc = *static_cast<Scene*>(s);
    Camera cam(view->camera);
    int64_t rays = 0;
    Col c = render_one(sc, cam, view, p, view_id, px, py, si, rays);
    out[0] = c.r;
    out[1] = c.g;
    out[2] = c.b;
}

void oc_env_sample(void* s, float u0, float u1, float dir[3], float* pdf)
{
    const Scene& sc = *static_cast<Scene*>(s);
    V3 d;
    *pdf = 0;
    dir[0] = dir[1] = dir[2] = 0;
    if (!sc.env) return;
    sc.env->sample(u0, u1, d, *pdf);
    dir[0] = d.x;
    dir[1] = d.y;
    dir[2] = d.z;
}

void oc_env_value(void* s, const float dir[3], float rgb[3])
{
    const Scene& sc = *static_cast<Scene*>(s);
    Col c = sc.env ? sc.env->get(V3(dir[0], dir[1], dir[2])) : Col();
    rgb[0] = c.r, rgb[1] = c.g, rgb[2] = c.b;
}

float oc_env_pdf(void* s, const float dir[3])
{
    const Scene& sc = *static_cast<Scene*>(s);
    if (!sc.env) return 0;
    return sc.env->pdf(V3(dir[0], dir[1], dir[2]));
}

void oc_transparent_dir(const float d[3], const float n[3], float ior, double u, float out[3])
{
    V3 o = transparent_dir(V3(d[0], d[1], d[2]), V3(n[0], n[1], n[2]), ior, u);
    out[0] = o.x, out[1] = o.y, out[2] = o.z;
}

void oc_bsdf(void* s, int mat, const float wo[3], const float wi[3], const float nf[3], float f[3], float* pdf)
{
    const Scene& sc = *static_cast<Scene*>(s);
    const Material* m = sc.mats[mat].get();
    V3 a(wo[0], wo[1], wo[2]), b(wi[0], wi[1], wi[2]), n(nf[0], nf[1], nf[2]);
    Col c = bsdf_eval(m, a, b, n);
    f[0] = c.r, f[1] = c.g, f[2] = c.b;
    *pdf = bsdf_pdf(m, a, b, n);
}

int oc_bsdf_sample(void* s, int mat, const float wo[3], const float nf[3], float u0, float u1, float u2, float wi[3])
{
    const Scene& sc = *static_cast<Scene*>(s);
    V3 w;
    bool ok = bsdf_sample(sc.mats[mat].get(), V3(wo[0], wo[1], wo[2]), V3(nf[0], nf[1], nf[2]), u0, u1, u2, w);
    wi[0] = w.x, wi[1] = w.y, wi[2] = w.z;
    return ok ? 1 : 0;
}

// Tonemapper::clamp (src/tonemapper.cpp:31-36): the float argument is
// converted to the int result by truncation.  x86 converts NaN and values
// outside int to INT_MIN; the explicit test keeps that defined here.
static int tonemap_clamp(float x)
{
    if (x <= 0) return 0;
    if (x >= 255) return 255;
    if (x != x) return INT32_MIN;
    return (int)x;
}

// Tonemapper::map (src/tonemapper.cpp:50-56): std::pow(float, double) is the
// double pow; qRgb keeps the low 8 bits of each channel.
void oc_tonemap(const float* film, int64_t n, float gamma, uint8_t* out)
{
    for (int64_t i = 0; i < n; ++i) out[i] = (uint8_t)(tonemap_clamp(255 * std::pow(film[i], 1. / gamma)) & 0xff);
}

/* The filtered Metropolis renderer's pieces, for the tests that pin them. */
void oc_box_blur(const float* in, int w, int h, int p, float* out)
{
    Img a(w, h, 1);
    std::memcpy(a.v.data(), in, a.v.size() * 4);
    const Img b = box_blur(a, p);
    std::memcpy(out, b.v.data(), b.v.size() * 4);
}
void oc_sep3_reflect101(const float* in, int w, int h, float* out)
{
    Img a(w, h, 3);
    std::memcpy(a.v.data(), in, a.v.size() * 4);
    const Img b = sep3_reflect101(a);
    std::memcpy(out, b.v.data(), b.v.size() * 4);
}
void oc_var2(const float* a, const float* b, int64_t n, float* mean, float* var)
{
    Img x(1, 1, 1), y(1, 1, 1), m, v;
    x.W = y.W = (int)n;
    x.v.assign(a, a + n);
    y.v.assign(b, b + n);
    var2(x, y, m, v);
    std::memcpy(mean, m.v.data(), n * 4);
    std::memcpy(var, v.v.data(), n * 4);
}
void oc_flt_importance(const float* var_filtered, const float* mean_filtered, int w, int h, float* imp)
{
    Img v(w, h, 3), m(w, h, 3);
    std::memcpy(v.v.data(), var_filtered, v.v.size() * 4);
    std::memcpy(m.v.data(), mean_filtered, m.v.size() * 4);
    const Img out = flt_importance(v, m);
    std::memcpy(imp, out.v.data(), out.v.size() * 4);
}
/* One film's running sums (film, mean, m2: w*h*3; sw: w*h): rescaled to a
 * new pass's importance (sumimp, newsum: w*h, or null for none), then the
 * deposits (pixel, weight, rgb) added in order. */
void oc_flt_accumulate(int w, int h, float* film, float* mean, float* m2, float* sw, const float* sumimp,
                       const float* newsum, int64_t ndep, const int32_t* pos, const float* wt, const float* rgb)
{
    FltState st;
    const size_t n3 = (size_t)w * h * 3, n1 = (size_t)w * h;
    for (int k = 0; k < 2; ++k) {
        st.film[k] = Img(w, h, 3);
        st.mean[k] = Img(w, h, 3);
        st.m2[k] = Img(w, h, 3);
        st.sw[k] = Img(w, h, 1);
    }
    std::memcpy(st.film[0].v.data(), film, n3 * 4);
    std::memcpy(st.mean[0].v.data(), mean, n3 * 4);
    std::memcpy(st.m2[0].v.data(), m2, n3 * 4);
    std::memcpy(st.sw[0].v.data(), sw, n1 * 4);
    if (sumimp && newsum) {
        Img a(w, h, 1), b(w, h, 1);
        std::memcpy(a.v.data(), sumimp, n1 * 4);
        std::memcpy(b.v.data(), newsum, n1 * 4);
        flt_rescale(st, a, b);
    }
    for (int64_t i = 0; i < ndep; ++i) flt_add_sample(st, 0, pos[i], wt[i], rgb + 3 * i);
    std::memcpy(film, st.film[0].v.data(), n3 * 4);
    std::memcpy(mean, st.mean[0].v.data(), n3 * 4);
    std::memcpy(m2, st.m2[0].v.data(), n3 * 4);
    std::memcpy(sw, st.sw[0].v.data(), n1 * 4);
}

void oc_sym_filter(const float* image, const float* guide, const float* pixvar, int w, int h, int patch, int radius,
                   float* out)
{
    Img a(w, h, 3), g(w, h, 3), v(w, h, 3);
    std::memcpy(a.v.data(), image, a.v.size() * 4);
    std::memcpy(g.v.data(), guide, g.v.size() * 4);
    std::memcpy(v.v.data(), pixvar, v.v.size() * 4);
    Img o = sym_filter(a, g, v, patch, radius);
    std::memcpy(out, o.v.data(), o.v.size() * 4);
}

void oc_mlt_perturb(const float* s, const float* u, const float* v, int64_t n, float* out)
{
    for (int64_t i = 0; i < n; ++i) out[i] = mlt_perturb(s[i], u[i], v[i]);
}

void oc_integrate_samples(void* s, const rt_camera* camera, int max_depth, int rr_start, float p,
                          const float* samples, int64_t n, float* rgb, int threads)
{
    const Scene& sc = *static_cast<Scene*>(s);
    Camera cam(*camera);
    const int G = 2 * max_depth + 1;
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads > 0 ? threads : 1)
    for (int64_t i = 0; i < n; ++i) {
        Pss x(G);
        for (int g = 0; g < G; ++g)
            for (int k = 0; k < 4; ++k) x.g[g].u[k] = samples[((size_t)i * G + g) * 4 + k];
        int64_t rays = 0;
        Col c = integrate_pss(sc, cam, x, max_depth, rr_start, p, false, rays);
        rgb[3 * i] = c.r, rgb[3 * i + 1] = c.g, rgb[3 * i + 2] = c.b;
    }
}

int oc_render_mlt(void* s, const rt_camera* camera, const rt_mlt_params* p, float* film, rt_mlt_result* res,
                  int threads)
{
    if (p->max_depth < 1 || p->mutations < 1) return -1;
    MltRun R(static_cast<Scene*>(s), *camera, *p, threads > 0 ? threads : 1);
    if (p->kind == RT_MLT_METROPOLIS)
        run_metropolis(R);
    else if (p->kind == RT_MLT_ERPT)
        run_erpt(R);
    else if (p->kind == RT_MLT_FILTERED && p->flt_passes >= 1)
        run_filtered(R);
    else
        return -1;
    for (size_t i = 0; i < R.film.size(); ++i) film[i] += (float)R.film[i];
    if (res) *res = R.res;
    return 0;
}

}  // extern "C"

/* ---------------------------------------------------------------------------
 * Traversal of the flattened 4-wide BVH (the layout rt_bvh_export writes),
 * restating the device traversal's visiting order (trav_step and traverse in
 * renderer_amd/csrc/rt_device.h) so that the nodes it fetches and the
 * records it tests can be counted on the CPU.  The reference has no such
 * tree (its BVH is binary, leaf = 1 primitive, both children visited,
 * src/bvh.cpp); the primitive tests are the reference's formulas
 * (src/triangle.cpp:30-56 on the un-normalised normal, the parallelogram of
 * src/arealight.cpp:73-90 as a triangle whose v runs to 1, the sphere of
 * src/sphere.cpp:29-56), evaluated in the written order, as on the device.
 * ------------------------------------------------------------------------- */
namespace flat {

enum { KIND_TRI = 0, KIND_SPHERE = 1, KIND_PARA = 2 }; /* leaf word kinds */

struct F3 {
    float x, y, z;
};
inline F3 sub(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline float dotf(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline F3 crossf(F3 a, F3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }

/* Accepts tmin <= t <= tmax and t < best (the device's exact test after its
 * widened reciprocal prefilter, which never drops an accepted hit). */
inline bool tri(F3 o, F3 d, const float* r, float tmin, float tmax, float best, bool para, float& t)
{
    F3 p1{r[0], r[1], r[2]}, e1{r[4], r[5], r[6]}, e2{r[8], r[9], r[10]};
    F3 ad = sub(o, p1);
    F3 n = crossf(e1, e2);
    float w = -dotf(n, d);
    float tn = dotf(n, ad);
    F3 adxde = crossf(ad, d);
    float un = dotf(adxde, e2);
    float vn = -dotf(adxde, e1);
    if (w == 0.f) return false;
    t = tn / w;
    float u = un / w, v = vn / w;
    bool ok = (t >= tmin) && (t <= tmax) && (t < best) && (u >= 0.f) && (u <= 1.f) && (v >= 0.f);
    return ok && (para ? (v <= 1.f) : (u + v <= 1.f));
}

inline bool sphere(F3 o, F3 d, const float* r, float tmin, float tmax, float best, float& t)
{
    F3 c{r[0], r[1], r[2]};
    float rad = r[4];
    F3 diff = sub(o, c);
    float a = dotf(d, d);
    float b = 2.f * dotf(d, diff);
    float cc = dotf(diff, diff) - rad * rad;
    float disc = b * b - 4.f * a * cc;
    if (!(disc >= 0.f)) return false;
    float root = std::sqrt(disc);
    float t0 = (-b - root) / (2.f * a), t1 = (-b + root) / (2.f * a);
    if (tmin > t1) return false;
    t = t0 >= tmin ? t0 : t1;
    return t <= tmax && t < best;
}

/* Order-preserving map of a float onto uint32 (negatives reversed below the
 * positives). */
inline uint32_t sortable(float f)
{
    uint32_t b;
    std::memcpy(&b, &f, 4);
    return b ^ ((uint32_t)((int32_t)b >> 31) | 0x80000000u);
}

inline float safe_inv(float x)
{
    const float tiny = 1e-30f;
    return 1.0f / (std::fabs(x) < tiny ? std::copysign(tiny, x) : x);
}

struct Result {
    float t;
    int32_t g;
    int visits, tests;
};

Result trace(const float* nodes, int64_t n_nodes, const float* recs, F3 o, F3 d, float tmin, float tmax, bool any)
{
    const float inf = std::numeric_limits<float>::infinity();
    const float slack = 9.5367431640625e-07f; /* 2^-20 */
    Result R{inf, -1, 0, 0};
    const F3 inv{safe_inv(d.x), safe_inv(d.y), safe_inv(d.z)};
    /* best starts one ulp above tmax: "t < best" then also enforces t <= tmax */
    float best = std::nextafter(tmax, inf);
    int64_t cur = n_nodes > 0 && tmin <= tmax ? 0 : -1;
    std::vector<int32_t> stack;
    const int near_x = inv.x >= 0.f ? 0 : 4, near_y = inv.y >= 0.f ? 0 : 4, near_z = inv.z >= 0.f ? 0 : 4;
    while (cur >= 0) {
        const float* nd = nodes + 32 * cur;
        ++R.visits;
        const float bound = best + std::fabs(best) * slack;
        float key[4];
        int32_t word[4];
        for (int k = 0; k < 4; ++k) {
            const float nx = nd[near_x + k], fx = nd[(4 - near_x) + k];
            const float ny = nd[8 + near_y + k], fy = nd[8 + (4 - near_y) + k];
            const float nz = nd[16 + near_z + k], fz = nd[16 + (4 - near_z) + k];
            float tl = std::fmax(std::fmax((nx - o.x) * inv.x, (ny - o.y) * inv.y), std::fmax((nz - o.z) * inv.z, tmin));
            float th = std::fmin(std::fmin((fx - o.x) * inv.x, (fy - o.y) * inv.y), std::fmin((fz - o.z) * inv.z, bound));
            key[k] = tl <= th + std::fabs(th) * slack ? tl : inf;
            std::memcpy(&word[k], &nd[24 + k], 4);
        }
        uint32_t uk[4];
        if (!any) {
            /* ascending entry distance, slot in the two low bits as the tie
             * break; the child words follow their slots */
            uint32_t u[4];
            for (int k = 0; k < 4; ++k) u[k] = (sortable(key[k]) & ~3u) | (uint32_t)k;
            std::sort(u, u + 4);
            int32_t w[4];
            for (int j = 0; j < 4; ++j) {
                w[j] = word[u[j] & 3u];
                uk[j] = u[j] & ~3u;
            }
            std::memcpy(word, w, sizeof w);
        } else {
            for (int k = 0; k < 4; ++k) uk[k] = sortable(key[k]);
        }
        const uint32_t miss = sortable(inf) & ~3u;
        bool done = false;
        /* leaves, nearest first; records in leaf order */
        for (int k = 0; k < 4 && !done; ++k) {
            if (!(word[k] < 0 && uk[k] < miss)) continue;
            if (!(uk[k] <= sortable(best + std::fabs(best) * slack))) continue;
            const uint32_t x = ~(uint32_t)word[k];
            const uint32_t count = x & 7u, kind = (x >> 3) & 3u, first = x >> 5;
            for (uint32_t j = 0; j < count; ++j) {
                const float* r = recs + 12 * (size_t)(first + j);
                ++R.tests;
                float t = 0.f;
                bool hit = kind == KIND_SPHERE ? sphere(o, d, r, tmin, inf, best, t)
                                               : tri(o, d, r, tmin, inf, best, kind == KIND_PARA, t);
                if (hit) {
                    best = t;
                    std::memcpy(&R.g, &r[3], 4);
                    if (any) {
                        done = true;
                        break;
                    }
                }
            }
        }
        if (done) break;
        /* interior children: the nearest continues, the others are deferred
         * with the next nearest on top */
        int32_t next = -1;
        const uint32_t cut = sortable(best + std::fabs(best) * slack);
        for (int k = 3; k >= 0; --k) {
            if (word[k] < 0 || !(uk[k] < miss) || !(uk[k] <= cut)) continue;
            if (next >= 0) stack.push_back(next);
            next = word[k];
        }
        if (next >= 0) {
            cur = next;
        } else if (stack.empty()) {
            cur = -1;
        } else {
            cur = stack.back();
            stack.pop_back();
        }
    }
    if (R.g >= 0) R.t = best;
    return R;
}

}  // namespace flat

extern "C" void oc_flat_trace(const float* nodes, int64_t n_nodes, const float* records, int64_t n,
                              const float* o_tmin, const float* d_tmax, int any, int32_t* g, float* t,
                              int32_t* visits, int32_t* tests, int threads)
{
#pragma omp parallel for schedule(dynamic, 256) num_threads(threads > 0 ? threads : 1)
    for (int64_t i = 0; i < n; ++i) {
        const float* a = o_tmin + 4 * i;
        const float* b = d_tmax + 4 * i;
        flat::Result R = flat::trace(nodes, n_nodes, records, {a[0], a[1], a[2]}, {b[0], b[1], b[2]}, a[3], b[3], any != 0);
        if (any) {
            g[i] = R.g >= 0 ? 1 : 0;
        } else {
            g[i] = R.g;
            t[i] = R.g >= 0 ? R.t : std::numeric_limits<float>::quiet_NaN();
        }
        if (visits) visits[i] = R.visits;
        if (tests) tests[i] = R.tests;
    }
}
