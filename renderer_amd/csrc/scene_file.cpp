/*
 * scene_file.cpp — scene-description front end (include/rt_scene_file.h).
 *
 * Host C++ with no dependency beyond the standard library.  It follows the
 * reference's front end step by step:
 *   grammar            src/sceneparser.cpp:37-125 (Boost.Spirit there; a
 *                      backtracking recursive-descent parser here, with the
 *                      same alternatives in the same order and the same
 *                      skipper: whitespace, block comments, line comments)
 *   obj(...) paths     src/sceneast.cpp:798-826 (relative to the scene file)
 *   OBJ / MTL import   src/objreader.cpp:142-366, std::istream extraction
 *                      semantics restated by LineStream below
 *   resolveVars        src/sceneast.cpp:731-747
 *   buildScene         src/sceneast.cpp:589-695 (assignments in order, the
 *                      last definition of a name wins, materials resolved
 *                      when the intersectable is assigned)
 *   createBVH + dump   src/sceneast.cpp:749-766, src/bvh.cpp:97-133,
 *                      src/scenedumper.cpp:20-220
 *
 * Numbers are converted with strtof (correctly rounded).  Spirit's float_
 * accumulates digits in float and may differ by one ulp on literals with
 * more than about eight significant digits.
 */
#include <algorithm>
#include <cctype>
#include <charconv>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <set>
#include <string>
#include <vector>

#include "rt_abi.h"
#include "rt_scene_file.h"

/* film_io.cpp: Radiance RGBE reader */
bool rt_read_rgbe(const std::string& path, int& w, int& h, std::vector<float>& rgb, std::string& err);
bool rt_read_image8(const std::string& path, int& w, int& h, std::vector<float>& rgb, std::string& err);

namespace {

thread_local std::string g_last_error;

// ---------------------------------------------------------------------------
// AST (src/sceneast.h:38-540)
// ---------------------------------------------------------------------------
struct V3 {
    float x = 0, y = 0, z = 0;
};

enum class MatKind { Diffuse, Phong, Blinn, Mirror, Texture, ScaledTexture, Refractive, Ident };
struct MatExpr {
    MatKind kind = MatKind::Diffuse;
    V3 a, b;             // diffuse colour | Phong/Blinn kd, ks | scaled texture coefficient (a)
    float f = 0, g = 1;  // exponent | mirror coefficient | ior ; scaled texture gamma (g)
    std::string s;       // texture file or variable name
};

enum class XfKind { Literal, Translate, Rotate, Scale, ScaleVec };
struct XfExpr {
    XfKind kind = XfKind::Literal;
    float m[16] = {};  // literal, row-major
    V3 v;
    float f = 0;
};

enum class NodeKind { List, Quad, Plane, Obj, Tri, Instance, Bvh, Sphere, Box, Isect, Union, Diff, Xor };
struct Node {
    NodeKind kind = NodeKind::List;
    int mat = -1;      // Doc::mats
    int aux = -1;      // Obj: Doc::strings; Instance: Doc::xforms
    float f[24] = {};  // Tri: p1 p2 p3 n1 n2 n3 t1 t2 t3; Quad: p1..p4; Sphere: c r; Box/Bvh: min max; Plane: abcd
    std::vector<Node> kids;
};

enum class LightKind { Point, Area, Cone, Env };
struct LightExpr {
    LightKind kind = LightKind::Point;
    V3 p, u, v, I;  // Point: p, I; Area: p u v I; Cone: p, u = direction, I; Env: I = coefficient
    float angle = 0;
    std::string file;
};

struct CamExpr {
    V3 eye, look, up;
    float fov = 0, xres = 0, yres = 0;
};

enum class ValKind { Lights, Camera, Intersectable, Material, Ident };
struct Value {
    ValKind kind = ValKind::Ident;
    std::vector<LightExpr> lights;
    CamExpr cam;
    Node node;
    int mat = -1;
    std::string ident;
};

struct Assignment {
    std::string name;
    Value value;
};

struct Doc {
    std::vector<Assignment> as;
    std::vector<MatExpr> mats;
    std::vector<std::string> strings;
    std::vector<std::vector<XfExpr>> xforms;
};

// ---------------------------------------------------------------------------
// Parser (src/sceneparser.cpp:37-125)
// ---------------------------------------------------------------------------
class Parser {
public:
    Parser(const std::string& text, Doc& doc) : s_(text), n_(text.size()), d_(doc) {}

    bool parse(std::string& err)
    {
        for (;;) {  // assignments_rule = *assignment_rule
            size_t at = i_;
            Assignment a;
            if (!assignment(a)) {
                i_ = at;
                break;
            }
            d_.as.push_back(std::move(a));
        }
        skip();  // phrase_parse post-skips
        if (i_ == n_) return true;
        size_t at = std::max(far_, i_);
        int line = 1, col = 1;
        for (size_t k = 0; k < at && k < n_; ++k) {
            if (s_[k] == '\n') {
                ++line;
                col = 1;
            } else {
                ++col;
            }
        }
        std::string near = s_.substr(at, 40);
        for (char& c : near)
            if (c == '\n' || c == '\r' || c == '\t') c = ' ';
        err = "line " + std::to_string(line) + ", column " + std::to_string(col) + ": the grammar stops here: \"" +
              near + "\"";
        if (bad_comment_) err += " (unterminated comment)";
        return false;
    }

private:
    const std::string& s_;
    size_t n_, i_ = 0, far_ = 0;
    bool bad_comment_ = false;
    Doc& d_;

    void reach() { far_ = std::max(far_, i_); }

    // comment_skipper: space | "/*" *(char - "*/") "*/" | "//" *(char - "\n") "\n"
    void skip()
    {
        for (;;) {
            if (i_ < n_ && std::isspace((unsigned char)s_[i_])) {
                ++i_;
            } else if (i_ + 1 < n_ && s_[i_] == '/' && s_[i_ + 1] == '*') {
                size_t e = s_.find("*/", i_ + 2);
                if (e == std::string::npos) {
                    bad_comment_ = true;
                    return;
                }
                i_ = e + 2;
            } else if (i_ + 1 < n_ && s_[i_] == '/' && s_[i_ + 1] == '/') {
                size_t e = s_.find('\n', i_ + 2);
                if (e == std::string::npos) {
                    bad_comment_ = true;
                    return;
                }
                i_ = e + 1;
            } else {
                return;
            }
        }
    }

    bool lit(const char* w)
    {
        skip();
        reach();
        size_t L = std::strlen(w);
        if (n_ - i_ < L || s_.compare(i_, L, w) != 0) return false;
        i_ += L;
        return true;
    }

    // qi::float_ : [sign] (nan | inf[inity] | digits[.digits] | .digits) [(e|E)[sign]digits]
    bool number(float& x)
    {
        skip();
        reach();
        size_t k = i_;
        bool neg = false;
        if (k < n_ && (s_[k] == '+' || s_[k] == '-')) neg = s_[k++] == '-';
        auto ieq = [&](size_t at, const char* w) {
            size_t L = std::strlen(w);
            if (n_ - at < L) return false;
            for (size_t j = 0; j < L; ++j)
                if (std::tolower((unsigned char)s_[at + j]) != w[j]) return false;
            return true;
        };
        if (ieq(k, "nan")) {
            x = std::numeric_limits<float>::quiet_NaN();
            if (neg) x = -x;
            i_ = k + 3;
            return true;
        }
        if (ieq(k, "inf")) {
            x = neg ? -std::numeric_limits<float>::infinity() : std::numeric_limits<float>::infinity();
            i_ = k + (ieq(k, "infinity") ? 8 : 3);
            return true;
        }
        size_t d0 = k;
        while (k < n_ && std::isdigit((unsigned char)s_[k])) ++k;
        size_t ndig = k - d0;
        if (k < n_ && s_[k] == '.') {
            size_t f0 = ++k;
            while (k < n_ && std::isdigit((unsigned char)s_[k])) ++k;
            ndig += k - f0;
        }
        if (ndig == 0) return false;
        if (k < n_ && (s_[k] == 'e' || s_[k] == 'E')) {
            size_t e = k + 1;
            if (e < n_ && (s_[e] == '+' || s_[e] == '-')) ++e;
            size_t e0 = e;
            while (e < n_ && std::isdigit((unsigned char)s_[e])) ++e;
            if (e > e0) k = e;
        }
        std::string tok = s_.substr(i_, k - i_);
        x = std::strtof(tok.c_str(), nullptr);
        i_ = k;
        return true;
    }

    bool string_lit(std::string& out)
    {
        skip();
        reach();
        if (i_ >= n_ || s_[i_] != '"') return false;
        size_t e = s_.find('"', i_ + 1);
        if (e == std::string::npos) return false;
        out = s_.substr(i_ + 1, e - i_ - 1);
        i_ = e + 1;
        return true;
    }

    bool identifier(std::string& out)
    {
        skip();
        reach();
        if (i_ >= n_ || !(std::isalpha((unsigned char)s_[i_]) || s_[i_] == '_')) return false;
        size_t k = i_ + 1;
        while (k < n_ && (std::isalnum((unsigned char)s_[k]) || s_[k] == '_')) ++k;
        out = s_.substr(i_, k - i_);
        i_ = k;
        return true;
    }

    // A sequence either succeeds or leaves the position where it started.
    template <class F>
    bool seq(F&& body)
    {
        size_t at = i_;
        if (body()) return true;
        i_ = at;
        return false;
    }

    bool vec2(float* o)
    {
        return seq([&] { return lit("[") && number(o[0]) && number(o[1]) && lit("]"); });
    }
    bool vec3(V3& v)
    {
        return seq([&] { return lit("[") && number(v.x) && number(v.y) && number(v.z) && lit("]"); });
    }
    bool vec3f(float* o)
    {
        V3 v;
        if (!vec3(v)) return false;
        o[0] = v.x, o[1] = v.y, o[2] = v.z;
        return true;
    }
    bool vec4(float* o)
    {
        return seq([&] { return lit("[") && number(o[0]) && number(o[1]) && number(o[2]) && number(o[3]) && lit("]"); });
    }

    // -- materials -----------------------------------------------------------
    bool material_literal(MatExpr& m)
    {
        if (seq([&] { m = MatExpr(); m.kind = MatKind::Diffuse; return lit("diffuse") && lit("(") && vec3(m.a) && lit(")"); }))
            return true;
        if (seq([&] {
                m = MatExpr();
                m.kind = MatKind::Phong;
                return lit("phong") && lit("(") && vec3(m.a) && lit(",") && vec3(m.b) && lit(",") && number(m.f) && lit(")");
            }))
            return true;
        if (seq([&] {
                m = MatExpr();
                m.kind = MatKind::Blinn;
                return lit("blinn") && lit("(") && vec3(m.a) && lit(",") && vec3(m.b) && lit(",") && number(m.f) && lit(")");
            }))
            return true;
        if (seq([&] { m = MatExpr(); m.kind = MatKind::Mirror; return lit("mirror") && lit("(") && number(m.f) && lit(")"); }))
            return true;
        // texture_material_rule: the rule's attribute (ast_texture_material,
        // src/sceneparser.h:98) keeps the file name only.
        if (seq([&] {
                m = MatExpr();
                m.kind = MatKind::Texture;
                V3 c;
                float g;
                return lit("texture") && lit("(") && string_lit(m.s) && lit(",") && vec3(c) && lit(",") && number(g) &&
                       lit(")");
            }))
            return true;
        if (seq([&] { m = MatExpr(); m.kind = MatKind::Refractive; return lit("refractive") && lit("(") && number(m.f) && lit(")"); }))
            return true;
        return false;
    }

    bool material(int& idx)
    {
        MatExpr m;
        if (material_literal(m)) {
            idx = (int)d_.mats.size();
            d_.mats.push_back(std::move(m));
            return true;
        }
        std::string id;
        size_t at = i_;
        if (identifier(id)) {
            m.kind = MatKind::Ident;
            m.s = id;
            idx = (int)d_.mats.size();
            d_.mats.push_back(std::move(m));
            return true;
        }
        i_ = at;
        return false;
    }

    // -- matrices ------------------------------------------------------------
    bool basic_matrix(XfExpr& x)
    {
        if (seq([&] {
                x = XfExpr();
                x.kind = XfKind::Literal;
                return lit("[") && vec4(x.m) && vec4(x.m + 4) && vec4(x.m + 8) && vec4(x.m + 12) && lit("]");
            }))
            return true;
        if (seq([&] { x = XfExpr(); x.kind = XfKind::Translate; return lit("translate") && lit("(") && vec3(x.v) && lit(")"); }))
            return true;
        if (seq([&] {
                x = XfExpr();
                x.kind = XfKind::Rotate;
                return lit("rotate") && lit("(") && number(x.f) && lit(",") && vec3(x.v) && lit(")");
            }))
            return true;
        if (seq([&] { x = XfExpr(); x.kind = XfKind::Scale; return lit("scale") && lit("(") && number(x.f) && lit(")"); }))
            return true;
        if (seq([&] { x = XfExpr(); x.kind = XfKind::ScaleVec; return lit("scale") && lit("(") && vec3(x.v) && lit(")"); }))
            return true;
        return false;
    }

    bool matrix(std::vector<XfExpr>& out)
    {
        out.clear();
        XfExpr x;
        if (!basic_matrix(x)) return false;
        out.push_back(x);
        for (;;) {
            size_t at = i_;
            if (!(lit("*") && basic_matrix(x))) {
                i_ = at;
                break;
            }
            out.push_back(x);
        }
        return true;
    }

    // -- intersectables ------------------------------------------------------
    bool binary_csg(const char* w, NodeKind k, Node& n)
    {
        return seq([&] {
            n = Node();
            n.kind = k;
            n.kids.resize(2);
            return lit(w) && lit("(") && csg(n.kids[0]) && lit(",") && csg(n.kids[1]) && lit(")");
        });
    }
    bool sphere(Node& n)
    {
        return seq([&] {
            n = Node();
            n.kind = NodeKind::Sphere;
            return lit("sphere") && lit("(") && vec3f(n.f) && lit(",") && number(n.f[3]) && lit(",") && material(n.mat) &&
                   lit(")");
        });
    }
    bool box(Node& n)
    {
        return seq([&] {
            n = Node();
            n.kind = NodeKind::Box;
            return lit("box") && lit("(") && vec3f(n.f) && lit(",") && vec3f(n.f + 3) && lit(",") && material(n.mat) &&
                   lit(")");
        });
    }
    bool csg(Node& n)
    {
        return binary_csg("isect", NodeKind::Isect, n) || binary_csg("union", NodeKind::Union, n) ||
               binary_csg("diff", NodeKind::Diff, n) || binary_csg("xor", NodeKind::Xor, n) || sphere(n) || box(n);
    }

    bool intersectable(Node& n)
    {
        // intersectables{ ... }
        if (seq([&] {
                n = Node();
                n.kind = NodeKind::List;
                if (!(lit("intersectables") && lit("{"))) return false;
                for (;;) {
                    size_t at = i_;
                    Node c;
                    if (!intersectable(c)) {
                        i_ = at;
                        break;
                    }
                    n.kids.push_back(std::move(c));
                }
                return lit("}");
            }))
            return true;
        if (seq([&] {
                n = Node();
                n.kind = NodeKind::Quad;
                return lit("quad") && lit("(") && vec3f(n.f) && lit(",") && vec3f(n.f + 3) && lit(",") && vec3f(n.f + 6) &&
                       lit(",") && vec3f(n.f + 9) && lit(",") && material(n.mat) && lit(")");
            }))
            return true;
        if (seq([&] {
                n = Node();
                n.kind = NodeKind::Plane;
                return lit("plane") && lit("(") && vec4(n.f) && lit(",") && material(n.mat) && lit(")");
            }))
            return true;
        if (seq([&] {
                n = Node();
                n.kind = NodeKind::Obj;
                std::string file;
                if (!(lit("obj") && lit("(") && string_lit(file) && lit(",") && material(n.mat) && lit(")"))) return false;
                n.aux = (int)d_.strings.size();
                d_.strings.push_back(file);
                return true;
            }))
            return true;
        if (seq([&] {
                n = Node();
                n.kind = NodeKind::Tri;
                return lit("t") && lit("(") && vec3f(n.f) && lit(",") && vec3f(n.f + 3) && lit(",") && vec3f(n.f + 6) &&
                       lit(",") && vec3f(n.f + 9) && lit(",") && vec3f(n.f + 12) && lit(",") && vec3f(n.f + 15) &&
                       lit(",") && vec2(n.f + 18) && lit(",") && vec2(n.f + 20) && lit(",") && vec2(n.f + 22) &&
                       lit(",") && material(n.mat) && lit(")");
            }))
            return true;
        if (seq([&] {
                n = Node();
                n.kind = NodeKind::Instance;
                std::vector<XfExpr> xf;
                n.kids.resize(1);
                if (!(lit("instance") && lit("(") && matrix(xf) && lit(",") && intersectable(n.kids[0]) && lit(")")))
                    return false;
                n.aux = (int)d_.xforms.size();
                d_.xforms.push_back(std::move(xf));
                return true;
            }))
            return true;
        if (seq([&] {
                n = Node();
                n.kind = NodeKind::Bvh;
                n.kids.resize(3);
                if (!(lit("b") && lit("(") && intersectable(n.kids[0]) && lit(",") && intersectable(n.kids[1]) &&
                      lit(",") && box(n.kids[2]) && lit(")")))
                    return false;
                std::memcpy(n.f, n.kids[2].f, 6 * sizeof(float));
                n.mat = n.kids[2].mat;
                n.kids.resize(2);
                return true;
            }))
            return true;
        return csg(n);
    }

    // -- camera and lights ---------------------------------------------------
    bool camera(CamExpr& c)
    {
        return seq([&] {
            return lit("camera") && lit("(") && vec3(c.eye) && lit(",") && vec3(c.look) && lit(",") && vec3(c.up) &&
                   lit(",") && number(c.fov) && lit(",") && number(c.xres) && lit(",") && number(c.yres) && lit(")");
        });
    }

    bool light(LightExpr& L)
    {
        if (seq([&] {
                L = LightExpr();
                L.kind = LightKind::Point;
                return lit("pointlight") && lit("(") && vec3(L.p) && lit(",") && vec3(L.I) && lit(")");
            }))
            return true;
        if (seq([&] {
                L = LightExpr();
                L.kind = LightKind::Area;
                return lit("arealight") && lit("(") && vec3(L.p) && lit(",") && vec3(L.u) && lit(",") && vec3(L.v) &&
                       lit(",") && vec3(L.I) && lit(")");
            }))
            return true;
        if (seq([&] {
                L = LightExpr();
                L.kind = LightKind::Cone;
                return lit("conelight") && lit("(") && vec3(L.p) && lit(",") && vec3(L.u) && lit(",") &&
                       number(L.angle) && lit(",") && vec3(L.I) && lit(")");
            }))
            return true;
        if (seq([&] {
                L = LightExpr();
                L.kind = LightKind::Env;
                return lit("environmentmap") && lit("(") && string_lit(L.file) && lit(",") && vec3(L.I) && lit(")");
            }))
            return true;
        return false;
    }

    bool light_list(std::vector<LightExpr>& out)
    {
        return seq([&] {
            out.clear();
            if (!(lit("lights") && lit("{"))) return false;
            for (;;) {
                size_t at = i_;
                LightExpr L;
                if (!light(L)) {
                    i_ = at;
                    break;
                }
                out.push_back(std::move(L));
            }
            return lit("}");
        });
    }

    // value_rule = intersectable | light_list | camera | material_literal | identifier
    bool value(Value& v)
    {
        size_t at = i_;
        v = Value();
        v.kind = ValKind::Intersectable;
        if (intersectable(v.node)) return true;
        i_ = at;
        v = Value();
        v.kind = ValKind::Lights;
        if (light_list(v.lights)) return true;
        v.kind = ValKind::Camera;
        if (camera(v.cam)) return true;
        MatExpr m;
        if (material_literal(m)) {
            v.kind = ValKind::Material;
            v.mat = (int)d_.mats.size();
            d_.mats.push_back(std::move(m));
            return true;
        }
        i_ = at;
        v.kind = ValKind::Ident;
        if (identifier(v.ident)) return true;
        i_ = at;
        return false;
    }

    bool assignment(Assignment& a)
    {
        return seq([&] { return identifier(a.name) && lit("=") && value(a.value) && lit(";"); });
    }
};

// ---------------------------------------------------------------------------
// Paths (QDir::filePath semantics)
// ---------------------------------------------------------------------------
std::string dir_of(const std::string& path)
{
    size_t k = path.find_last_of('/');
    if (k == std::string::npos) return ".";
    if (k == 0) return "/";
    return path.substr(0, k);
}

std::string join_path(const std::string& dir, const std::string& name)
{
    if (!name.empty() && name[0] == '/') return name;
    if (dir.empty()) return name;
    if (dir.back() == '/') return dir + name;
    return dir + "/" + name;
}

bool read_file(const std::string& path, std::string& out)
{
    FILE* fp = std::fopen(path.c_str(), "rb");
    if (!fp) return false;
    out.clear();
    char buf[1 << 16];
    size_t r;
    while ((r = std::fread(buf, 1, sizeof buf, fp)) > 0) out.append(buf, r);
    std::fclose(fp);
    return true;
}

// src/sceneast.cpp:798-826: obj file names relative to the scene directory.
void apply_path(Node& n, Doc& d, const std::string& dir)
{
    switch (n.kind) {
    case NodeKind::List:
    case NodeKind::Bvh:
    case NodeKind::Instance:
        for (Node& c : n.kids) apply_path(c, d, dir);
        break;
    case NodeKind::Obj: d.strings[n.aux] = join_path(dir, d.strings[n.aux]); break;
    default: break;
    }
}

// ---------------------------------------------------------------------------
// OBJ / MTL (src/objreader.cpp:142-366)
// ---------------------------------------------------------------------------

// The std::stringstream extraction rules the reader relies on: a sentry that
// skips whitespace and fails on a stream that is not good, fail + zero on a
// malformed number, eof when a read reaches the end of the line.
struct LineStream {
    const char* p;
    const char* e;
    bool failb = false, eofb = false;
    LineStream(const char* b, const char* end) : p(b), e(end) {}

    bool good() const { return !failb && !eofb; }
    void clear() { failb = eofb = false; }
    bool sentry()
    {
        if (!good()) {
            failb = true;
            return false;
        }
        while (p < e && std::isspace((unsigned char)*p)) ++p;
        if (p == e) {
            eofb = failb = true;
            return false;
        }
        return true;
    }
    int peek()
    {
        if (!good()) {
            failb = true;
            return -1;
        }
        if (p == e) {
            eofb = true;
            return -1;
        }
        return (unsigned char)*p;
    }
    bool token(std::string& out)
    {
        if (!sentry()) return false;
        const char* b = p;
        while (p < e && !std::isspace((unsigned char)*p)) ++p;
        out.assign(b, p);
        if (p == e) eofb = true;
        return true;
    }
    bool get_char(char& c)
    {
        if (!sentry()) return false;
        c = *p++;
        return true;
    }
    // num_get stage 2 for floating point: [sign] digits [. digits] [(e|E) [sign] digits]
    bool get_float(float& v)
    {
        if (!sentry()) return false;
        const char* b = p;
        const char* q = p;
        if (q < e && (*q == '+' || *q == '-')) ++q;
        while (q < e && std::isdigit((unsigned char)*q)) ++q;
        if (q < e && *q == '.') {
            ++q;
            while (q < e && std::isdigit((unsigned char)*q)) ++q;
        }
        if (q < e && (*q == 'e' || *q == 'E')) {
            ++q;
            if (q < e && (*q == '+' || *q == '-')) ++q;
            while (q < e && std::isdigit((unsigned char)*q)) ++q;
        }
        p = q;
        if (p == e) eofb = true;
        std::string tok(b, q);
        char* end = nullptr;
        errno = 0;
        float x = tok.empty() ? 0.f : std::strtof(tok.c_str(), &end);
        if (tok.empty() || end != tok.c_str() + tok.size()) {
            v = 0.f;
            failb = true;
            return false;
        }
        if (errno == ERANGE && std::isinf(x)) {
            v = x > 0 ? std::numeric_limits<float>::max() : -std::numeric_limits<float>::max();
            failb = true;
            return false;
        }
        v = x;
        return true;
    }
    // Returns false on failure; `v` is left alone when the sentry fails and
    // set (0 or the clamped limit) when the digits are malformed.
    bool get_long(long long& v, long long lo, long long hi)
    {
        if (!sentry()) return false;
        const char* q = p;
        bool neg = false;
        if (q < e && (*q == '+' || *q == '-')) neg = *q++ == '-';
        const char* d0 = q;
        long long acc = 0;
        bool over = false;
        while (q < e && std::isdigit((unsigned char)*q)) {
            if (acc > (hi + 1) / 10) over = true;
            acc = acc * 10 + (*q - '0');
            ++q;
        }
        p = q;
        if (p == e) eofb = true;
        if (q == d0) {
            v = 0;
            failb = true;
            return false;
        }
        long long x = neg ? -acc : acc;
        if (over || x > hi || x < lo) {
            v = neg ? lo : hi;
            failb = true;
            return false;
        }
        v = x;
        return true;
    }
    bool get_int(int& v)
    {
        long long x = v;
        bool ok = get_long(x, std::numeric_limits<int>::min(), std::numeric_limits<int>::max());
        v = (int)x;
        return ok;
    }
};

// Lines of a text buffer, as std::getline splits them.
template <class F>
void for_each_line(const std::string& text, F&& f)
{
    size_t i = 0, n = text.size();
    int line = 0;
    while (i < n) {
        size_t e = text.find('\n', i);
        if (e == std::string::npos) e = n;
        ++line;
        if (!f(text.data() + i, text.data() + e, line)) return;
        i = e + 1;
    }
}

void sanitize_material_name(std::string& s)
{
    for (char& c : s)
        if (!(std::isalnum((unsigned char)c) || c == '_')) c = '_';
    if (!s.empty() && std::isdigit((unsigned char)s[0])) s[0] = '_';
}

struct ObjMaterial {
    V3 kd, ks;
    std::string texture;
    float ns = 0, ni = 0;
    int illum = 0;
};

struct ObjReader {
    Doc& d;
    std::string& log;
    std::map<std::string, MatExpr> materials;

    // src/objreader.cpp:275-316
    void create_material(const ObjMaterial& m, const std::string& dir, const std::string& name)
    {
        MatExpr out;
        if (m.illum <= 2) {
            if (!m.texture.empty()) {
                std::string file = join_path(dir, m.texture);
                std::replace(file.begin(), file.end(), '\\', '/');
                out.kind = MatKind::ScaledTexture;
                out.s = file;
                out.a = V3{1, 1, 1};
                out.g = 1;
                if (m.kd.x != 0 || m.kd.y != 0 || m.kd.z != 0) out.a = m.kd;
            } else {
                if (m.kd.x == 0 && m.kd.y == 0 && m.kd.z == 0) log += "Warning: material " + name + " is black\n";
                if ((m.ks.x == 0 && m.ks.y == 0 && m.ks.z == 0) || m.ns == 0) {
                    out.kind = MatKind::Diffuse;
                    out.a = m.kd;
                } else {
                    out.kind = MatKind::Blinn;
                    out.a = m.kd;
                    out.b = m.ks;
                    out.f = m.ns;
                }
            }
        } else if (m.illum == 11) {
            out.kind = MatKind::Mirror;  // the reference leaves the coefficient uninitialised
            out.f = 0;
        } else {
            out.kind = MatKind::Refractive;
            out.f = m.ni;
        }
        materials[name] = out;
    }

    // src/objreader.cpp:318-366
    void load_mtl(const std::string& path)
    {
        std::string text;
        if (!read_file(path, text)) text.clear();
        std::string dir = dir_of(path);
        std::string cur;
        ObjMaterial m;
        for_each_line(text, [&](const char* b, const char* e, int) {
            LineStream ls(b, e);
            std::string type;
            ls.token(type);
            if (type == "newmtl") {
                if (!cur.empty()) {
                    create_material(m, dir, cur);
                    m = ObjMaterial();
                }
                ls.token(cur);
                sanitize_material_name(cur);
            } else if (type == "illum") {
                long long x = m.illum;
                ls.get_long(x, std::numeric_limits<short>::min(), std::numeric_limits<short>::max());
                m.illum = (int)x;
            } else if (type == "Kd") {
                ls.get_float(m.kd.x) && ls.get_float(m.kd.y) && ls.get_float(m.kd.z);
            } else if (type == "map_Kd") {
                ls.token(m.texture);
            } else if (type == "Ks") {
                ls.get_float(m.ks.x) && ls.get_float(m.ks.y) && ls.get_float(m.ks.z);
            } else if (type == "Ns") {
                ls.get_float(m.ns);
            } else if (type == "Ni") {
                ls.get_float(m.ni);
            }
            return true;
        });
        create_material(m, dir, cur);
    }

    struct Face {
        int pos[3], tex[3], nor[3];
        int mat;
    };

    static bool read_index(LineStream& ls, int& dest, int total)
    {
        ls.clear();
        ls.get_int(dest);
        if (ls.failb) return false;
        if (dest < 0) dest += total + 1;
        return true;
    }

    static bool read_indices(Face& f, int i, LineStream& ls, int np, int nt, int nn)
    {
        char c;
        if (!read_index(ls, f.pos[i], np)) return false;
        if (ls.peek() == '/') {
            ls.get_char(c);
            if (ls.peek() == '/') {
                f.tex[i] = 0;
            } else {
                if (!read_index(ls, f.tex[i], nt)) return false;
            }
            if (ls.peek() == '/') {
                ls.get_char(c);
                if (!read_index(ls, f.nor[i], nn)) return false;
            } else {
                f.nor[i] = 0;
            }
        } else {
            f.tex[i] = 0;
            f.nor[i] = 0;
        }
        return true;
    }

    // src/objreader.cpp:142-273.  Returns false (with err) on an index out of
    // range, which the reference reads out of bounds.
    bool load(const std::string& path, int default_mat, Node& out, std::string& err)
    {
        out = Node();
        out.kind = NodeKind::List;
        std::string text;
        if (!read_file(path, text)) {
            log += "Error opening file " + path + "\n";
            return true;
        }
        std::string dir = dir_of(path);
        std::vector<V3> pos, nor;
        std::vector<std::pair<float, float>> tex;
        std::vector<Face> faces;
        int cur = default_mat;
        long num_faces = 0;
        for_each_line(text, [&](const char* b, const char* e, int line) {
            LineStream ls(b, e);
            std::string type;
            ls.token(type);
            if (type == "v") {
                V3 p;
                float w = 0;
                ls.get_float(p.x);
                ls.get_float(p.y);
                ls.get_float(p.z);
                ls.clear();
                bool ok = ls.get_float(w);
                if (ok && w != 0) {
                    p.x /= w;
                    p.y /= w;
                    p.z /= w;
                }
                pos.push_back(p);
            } else if (type == "vt") {
                float u = 0, v = 0;
                ls.get_float(u);
                ls.get_float(v);
                tex.emplace_back(u, v);
            } else if (type == "vn") {
                V3 n;
                ls.get_float(n.x);
                ls.get_float(n.y);
                ls.get_float(n.z);
                nor.push_back(n);
            } else if (type == "f") {
                Face f;
                std::memset(&f, 0, sizeof f);
                f.mat = cur;
                int np = (int)pos.size(), nt = (int)tex.size(), nn = (int)nor.size();
                read_indices(f, 0, ls, np, nt, nn);
                read_indices(f, 2, ls, np, nt, nn);
                while (!ls.eofb) {
                    f.pos[1] = f.pos[2];
                    f.nor[1] = f.nor[2];
                    f.tex[1] = f.tex[2];
                    const char* before = ls.p;
                    if (read_indices(f, 2, ls, np, nt, nn)) {
                        faces.push_back(f);
                    } else if (ls.p == before) {
                        // a token that is not an index: the reference's loop
                        // (src/objreader.cpp:212-220) never ends here
                        log += path + ": line " + std::to_string(line) + ": face entry is not an index\n";
                        break;
                    }
                }
                ++num_faces;
            } else if (type == "usemtl") {
                std::string name;
                ls.token(name);
                sanitize_material_name(name);
                MatExpr m;
                m.kind = MatKind::Ident;
                m.s = name;
                cur = (int)d.mats.size();
                d.mats.push_back(m);
            } else if (type == "mtllib") {
                std::string file;
                ls.token(file);
                load_mtl(join_path(dir, file));
            }
            return true;
        });
        log += "ObjReader: " + std::to_string(faces.size()) + " triangles, " + std::to_string(num_faces) + " faces\n";
        out.kids.resize(faces.size());
        for (size_t i = 0; i < faces.size(); ++i) {
            const Face& f = faces[i];
            Node& t = out.kids[i];
            t.kind = NodeKind::Tri;
            t.mat = f.mat;
            auto bad = [&](int idx, size_t n) { return idx != 0 && (idx < 0 || (size_t)idx > n); };
            for (int k = 0; k < 3; ++k)
                if (bad(f.pos[k], pos.size()) || bad(f.nor[k], nor.size()) || bad(f.tex[k], tex.size())) {
                    err = path + ": face index out of range (triangle " + std::to_string(i) + ")";
                    return false;
                }
            if (f.pos[0] && f.pos[1] && f.pos[2])
                for (int k = 0; k < 3; ++k) {
                    const V3& p = pos[f.pos[k] - 1];
                    t.f[3 * k] = p.x, t.f[3 * k + 1] = p.y, t.f[3 * k + 2] = p.z;
                }
            if (f.nor[0] && f.nor[1] && f.nor[2])
                for (int k = 0; k < 3; ++k) {
                    const V3& n = nor[f.nor[k] - 1];
                    t.f[9 + 3 * k] = n.x, t.f[10 + 3 * k] = n.y, t.f[11 + 3 * k] = n.z;
                }
            if (f.tex[0] && f.tex[1] && f.tex[2])
                for (int k = 0; k < 3; ++k) {
                    t.f[18 + 2 * k] = tex[f.tex[k] - 1].first;
                    t.f[19 + 2 * k] = tex[f.tex[k] - 1].second;
                }
        }
        return true;
    }
};

// src/sceneast.cpp:128-158 (ObjLoader) and 731-747 (resolveVars)
bool resolve_objs(Node& n, Doc& d, ObjReader& rd, std::string& err)
{
    switch (n.kind) {
    case NodeKind::List:
    case NodeKind::Bvh:
    case NodeKind::Instance:
        for (Node& c : n.kids)
            if (!resolve_objs(c, d, rd, err)) return false;
        return true;
    case NodeKind::Obj: {
        ObjReader one{d, rd.log, {}};
        Node list;
        if (!one.load(d.strings[n.aux], n.mat, list, err)) return false;
        rd.materials.insert(one.materials.begin(), one.materials.end());  // the first definition of a name wins
        n = std::move(list);
        return true;
    }
    default: return true;
    }
}

bool resolve_vars(Doc& d, std::string& log, std::string& err)
{
    ObjReader all{d, log, {}};
    for (Assignment& a : d.as)
        if (a.value.kind == ValKind::Intersectable && !resolve_objs(a.value.node, d, all, err)) return false;
    for (auto& kv : all.materials) {  // each inserted at the front: reverse name order
        Assignment a;
        a.name = kv.first;
        a.value.kind = ValKind::Material;
        a.value.mat = (int)d.mats.size();
        d.mats.push_back(kv.second);
        d.as.insert(d.as.begin(), std::move(a));
    }
    return true;
}

// ---------------------------------------------------------------------------
// 4x4 matrices with QMatrix4x4's conventions (rows as written, M *= next,
// map() with the projective divide)
// ---------------------------------------------------------------------------
struct M4 {
    float m[4][4];
    static M4 identity()
    {
        M4 r;
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) r.m[i][j] = i == j ? 1.f : 0.f;
        return r;
    }
    M4 operator*(const M4& b) const
    {
        M4 r;
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) {
                float s = 0;
                for (int k = 0; k < 4; ++k) s += m[i][k] * b.m[k][j];
                r.m[i][j] = s;
            }
        return r;
    }
    V3 map(const V3& p) const
    {
        float x = m[0][0] * p.x + m[0][1] * p.y + m[0][2] * p.z + m[0][3];
        float y = m[1][0] * p.x + m[1][1] * p.y + m[1][2] * p.z + m[1][3];
        float z = m[2][0] * p.x + m[2][1] * p.y + m[2][2] * p.z + m[2][3];
        float w = m[3][0] * p.x + m[3][1] * p.y + m[3][2] * p.z + m[3][3];
        if (w == 1.f) return V3{x, y, z};
        return V3{x / w, y / w, z / w};
    }
    // inverse transpose of the upper 3x3 (identity when singular), computed in double
    void normal_matrix(double n[3][3]) const
    {
        double a[3][3];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) a[i][j] = m[i][j];
        double c[3][3];
        c[0][0] = a[1][1] * a[2][2] - a[1][2] * a[2][1];
        c[0][1] = a[1][2] * a[2][0] - a[1][0] * a[2][2];
        c[0][2] = a[1][0] * a[2][1] - a[1][1] * a[2][0];
        c[1][0] = a[0][2] * a[2][1] - a[0][1] * a[2][2];
        c[1][1] = a[0][0] * a[2][2] - a[0][2] * a[2][0];
        c[1][2] = a[0][1] * a[2][0] - a[0][0] * a[2][1];
        c[2][0] = a[0][1] * a[1][2] - a[0][2] * a[1][1];
        c[2][1] = a[0][2] * a[1][0] - a[0][0] * a[1][2];
        c[2][2] = a[0][0] * a[1][1] - a[0][1] * a[1][0];
        double det = a[0][0] * c[0][0] + a[0][1] * c[0][1] + a[0][2] * c[0][2];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) n[i][j] = det == 0 ? (i == j ? 1.0 : 0.0) : c[i][j] / det;
    }
};

// src/sceneast.cpp:254-300 (matrix_evaluator); rotation by the axis-angle
// formula with the angle in degrees.
M4 eval_xf(const XfExpr& x)
{
    M4 r = M4::identity();
    switch (x.kind) {
    case XfKind::Literal:
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) r.m[i][j] = x.m[4 * i + j];
        break;
    case XfKind::Translate:
        r.m[0][3] = x.v.x, r.m[1][3] = x.v.y, r.m[2][3] = x.v.z;
        break;
    case XfKind::Scale: r.m[0][0] = r.m[1][1] = r.m[2][2] = x.f; break;
    case XfKind::ScaleVec: r.m[0][0] = x.v.x, r.m[1][1] = x.v.y, r.m[2][2] = x.v.z; break;
    case XfKind::Rotate: {
        if (x.f == 0.f) break;
        float s, c;
        if (x.f == 90.f || x.f == -270.f) {
            s = 1, c = 0;
        } else if (x.f == -90.f || x.f == 270.f) {
            s = -1, c = 0;
        } else if (x.f == 180.f || x.f == -180.f) {
            s = 0, c = -1;
        } else {
            double a = (double)x.f * (M_PI / 180.0);
            s = (float)std::sin(a), c = (float)std::cos(a);
        }
        double len = std::sqrt((double)x.v.x * x.v.x + (double)x.v.y * x.v.y + (double)x.v.z * x.v.z);
        if (len == 0) break;
        float ax = (float)(x.v.x / len), ay = (float)(x.v.y / len), az = (float)(x.v.z / len);
        float ic = 1 - c;
        r.m[0][0] = ax * ax * ic + c, r.m[0][1] = ax * ay * ic - az * s, r.m[0][2] = ax * az * ic + ay * s;
        r.m[1][0] = ay * ax * ic + az * s, r.m[1][1] = ay * ay * ic + c, r.m[1][2] = ay * az * ic - ax * s;
        r.m[2][0] = az * ax * ic - ay * s, r.m[2][1] = az * ay * ic + ax * s, r.m[2][2] = az * az * ic + c;
        break;
    }
    }
    return r;
}

M4 eval_matrix(const std::vector<XfExpr>& xs)
{
    M4 r = eval_xf(xs[0]);
    for (size_t i = 1; i < xs.size(); ++i) r = r * eval_xf(xs[i]);
    return r;
}

// ---------------------------------------------------------------------------
// Bounding boxes of AST items (the reference builds the objects with
// DarkMatter and asks them, src/sceneast.cpp:767-776)
// ---------------------------------------------------------------------------
struct Box3 {
    float lo[3] = {INFINITY, INFINITY, INFINITY};
    float hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    void include(const float* p)  // src/axisalignedbox.cpp:90-98
    {
        for (int i = 0; i < 3; ++i) {
            if (p[i] < lo[i]) lo[i] = p[i];
            if (p[i] > hi[i]) hi[i] = p[i];
        }
    }
    void include(const Box3& b)
    {
        include(b.lo);
        include(b.hi);
    }
};

Box3 box_of(const Node& n, const Doc& d)
{
    Box3 b;
    switch (n.kind) {
    case NodeKind::Tri:
        for (int k = 0; k < 3; ++k) b.include(n.f + 3 * k);
        break;
    case NodeKind::Quad: {  // Triangle(p1,p2,p4) and Triangle(p4,p2,p3), src/quad.cpp:36-41
        Box3 t1, t2;
        t1.include(n.f), t1.include(n.f + 3), t1.include(n.f + 9);
        t2.include(n.f + 9), t2.include(n.f + 3), t2.include(n.f + 6);
        b = t1;
        b.include(t2);
        break;
    }
    case NodeKind::Sphere:  // src/sphere.cpp:58-62
        for (int k = 0; k < 3; ++k) {
            b.lo[k] = n.f[k] - n.f[3];
            b.hi[k] = n.f[k] + n.f[3];
        }
        break;
    case NodeKind::Box:
    case NodeKind::Bvh:
        for (int k = 0; k < 3; ++k) b.lo[k] = n.f[k], b.hi[k] = n.f[3 + k];
        break;
    case NodeKind::Plane:
        for (int k = 0; k < 3; ++k) b.lo[k] = -INFINITY, b.hi[k] = INFINITY;
        break;
    case NodeKind::List:
        for (const Node& c : n.kids) b.include(box_of(c, d));
        break;
    case NodeKind::Instance: {  // src/intersectableinstance.cpp:33-47
        Box3 in = box_of(n.kids[0], d);
        M4 M = eval_matrix(d.xforms[n.aux]);
        for (int c = 0; c < 8; ++c) {
            V3 p{(c & 4) ? in.hi[0] : in.lo[0], (c & 2) ? in.hi[1] : in.lo[1], (c & 1) ? in.hi[2] : in.lo[2]};
            V3 q = M.map(p);
            float f[3] = {q.x, q.y, q.z};
            b.include(f);
        }
        break;
    }
    case NodeKind::Isect:
    case NodeKind::Union:
    case NodeKind::Diff:
    case NodeKind::Xor: {  // src/csgoperation.cpp:65-77
        Box3 l = box_of(n.kids[0], d), r = box_of(n.kids[1], d);
        for (int k = 0; k < 3; ++k) {
            b.lo[k] = std::min(l.lo[k], r.lo[k]);
            b.hi[k] = std::max(l.hi[k], r.hi[k]);
        }
        break;
    }
    case NodeKind::Obj: break;
    }
    return b;
}

// src/bvh.cpp:97-133: median split of a list on the longest axis of its box,
// std::sort by box centre, leaves of one item.
Node bvh_create(std::vector<Node> items, const Doc& d, int bb_mat)
{
    if (items.size() == 1) return std::move(items[0]);
    Box3 bb;
    std::vector<Box3> boxes(items.size());
    for (size_t i = 0; i < items.size(); ++i) {
        boxes[i] = box_of(items[i], d);
        bb.include(boxes[i]);
    }
    float dx = bb.hi[0] - bb.lo[0], dy = bb.hi[1] - bb.lo[1], dz = bb.hi[2] - bb.lo[2];
    int axis = dx > dy && dx > dz ? 0 : dy > dz ? 1 : 2;
    struct Key {
        float c;
        size_t i;
    };
    std::vector<Key> keys(items.size());
    for (size_t i = 0; i < items.size(); ++i) keys[i] = Key{boxes[i].hi[axis] + boxes[i].lo[axis], i};
    std::sort(keys.begin(), keys.end(), [](const Key& a, const Key& b) { return a.c < b.c; });
    size_t split = items.size() / 2;
    if (split == 0 || items.size() - split == 0) {
        Node l;
        l.kind = NodeKind::List;
        l.kids = std::move(items);
        return l;
    }
    std::vector<Node> left, right;
    left.reserve(split);
    right.reserve(items.size() - split);
    for (size_t k = 0; k < keys.size(); ++k) (k < split ? left : right).push_back(std::move(items[keys[k].i]));
    items.clear();
    Node n;
    n.kind = NodeKind::Bvh;
    n.mat = bb_mat;
    for (int k = 0; k < 3; ++k) n.f[k] = bb.lo[k], n.f[3 + k] = bb.hi[k];
    n.kids.push_back(bvh_create(std::move(left), d, bb_mat));
    n.kids.push_back(bvh_create(std::move(right), d, bb_mat));
    return n;
}

// src/sceneast.cpp:189-208 (Flattener): nested lists spliced into one.
void flatten(const Node& n, std::vector<Node>& out)
{
    if (n.kind == NodeKind::List) {
        for (const Node& c : n.kids) flatten(c, out);
    } else {
        out.push_back(n);
    }
}

// ---------------------------------------------------------------------------
// Printer (src/scenedumper.cpp): function_name(args,...) with floats as
// operator<< prints them (6 significant digits, %g).
// ---------------------------------------------------------------------------
struct Printer {
    const Doc& d;
    std::string& o;

    void num(float f)
    {
        char buf[64];
        if (std::isnan(f)) {
            std::snprintf(buf, sizeof buf, "%g", (double)f);
            o += buf;
            return;
        }
        // general format, 6 significant digits: the same characters as %g
        auto r = std::to_chars(buf, buf + sizeof buf, (double)f, std::chars_format::general, 6);
        o.append(buf, r.ptr);
    }
    void v2(const float* f)
    {
        o += '[';
        num(f[0]);
        o += ' ';
        num(f[1]);
        o += ']';
    }
    void v3(const float* f)
    {
        o += '[';
        num(f[0]);
        o += ' ';
        num(f[1]);
        o += ' ';
        num(f[2]);
        o += ']';
    }
    void v3(const V3& v)
    {
        float f[3] = {v.x, v.y, v.z};
        v3(f);
    }
    void v4(const float* f)
    {
        o += '[';
        for (int k = 0; k < 4; ++k) {
            if (k) o += ' ';
            num(f[k]);
        }
        o += ']';
    }
    void str(const std::string& s)
    {
        o += '"';
        o += s;
        o += '"';
    }

    void mat(int idx)
    {
        const MatExpr& m = d.mats[idx];
        switch (m.kind) {
        case MatKind::Diffuse: o += "diffuse("; v3(m.a); o += ')'; break;
        case MatKind::Phong: o += "phong("; v3(m.a); o += ','; v3(m.b); o += ','; num(m.f); o += ')'; break;
        case MatKind::Blinn: o += "blinn("; v3(m.a); o += ','; v3(m.b); o += ','; num(m.f); o += ')'; break;
        case MatKind::Mirror: o += "mirror("; num(m.f); o += ')'; break;
        case MatKind::Texture: o += "texture("; str(m.s); o += ')'; break;
        case MatKind::ScaledTexture: o += "texture("; str(m.s); o += ','; v3(m.a); o += ','; num(m.g); o += ')'; break;
        case MatKind::Refractive: o += "refractive("; num(m.f); o += ')'; break;
        case MatKind::Ident: o += m.s; break;
        }
    }

    void xf(const XfExpr& x)
    {
        switch (x.kind) {
        case XfKind::Literal:
            o += '[';
            for (int r = 0; r < 4; ++r) v4(x.m + 4 * r);
            o += ']';
            break;
        case XfKind::Translate: o += "translate("; v3(x.v); o += ')'; break;
        case XfKind::Rotate: o += "rotate("; num(x.f); o += ','; v3(x.v); o += ')'; break;
        case XfKind::Scale: o += "scale("; num(x.f); o += ')'; break;
        case XfKind::ScaleVec: o += "scale("; v3(x.v); o += ')'; break;
        }
    }

    void node(const Node& n)
    {
        switch (n.kind) {
        case NodeKind::List:
            o += "intersectables{";
            for (const Node& c : n.kids) node(c);
            o += '}';
            break;
        case NodeKind::Quad:
            o += "quad(";
            for (int k = 0; k < 4; ++k) {
                v3(n.f + 3 * k);
                o += ',';
            }
            mat(n.mat);
            o += ')';
            break;
        case NodeKind::Plane: o += "plane("; v4(n.f); o += ','; mat(n.mat); o += ')'; break;
        case NodeKind::Obj: o += "obj("; str(d.strings[n.aux]); o += ','; mat(n.mat); o += ')'; break;
        case NodeKind::Tri:
            o += "t(";
            for (int k = 0; k < 6; ++k) {
                v3(n.f + 3 * k);
                o += ',';
            }
            for (int k = 0; k < 3; ++k) {
                v2(n.f + 18 + 2 * k);
                o += ',';
            }
            mat(n.mat);
            o += ')';
            break;
        case NodeKind::Instance: {
            o += "instance(";
            const std::vector<XfExpr>& xs = d.xforms[n.aux];
            for (size_t k = 0; k < xs.size(); ++k) {
                if (k) o += '*';
                xf(xs[k]);
            }
            o += ',';
            node(n.kids[0]);
            o += ')';
            break;
        }
        case NodeKind::Bvh:
            o += "b(";
            node(n.kids[0]);
            o += ',';
            node(n.kids[1]);
            o += ",box(";
            v3(n.f);
            o += ',';
            v3(n.f + 3);
            o += ',';
            mat(n.mat);
            o += "))";
            break;
        case NodeKind::Sphere: o += "sphere("; v3(n.f); o += ','; num(n.f[3]); o += ','; mat(n.mat); o += ')'; break;
        case NodeKind::Box: o += "box("; v3(n.f); o += ','; v3(n.f + 3); o += ','; mat(n.mat); o += ')'; break;
        case NodeKind::Isect:
        case NodeKind::Union:
        case NodeKind::Diff:
        case NodeKind::Xor:
            o += n.kind == NodeKind::Isect ? "isect(" : n.kind == NodeKind::Union ? "union(" : n.kind == NodeKind::Diff ? "diff(" : "xor(";
            node(n.kids[0]);
            o += ',';
            node(n.kids[1]);
            o += ')';
            break;
        }
    }

    void light(const LightExpr& L)
    {
        switch (L.kind) {
        case LightKind::Point: o += "pointlight("; v3(L.p); o += ','; v3(L.I); o += ')'; break;
        // ast_area_light::function_name is "srealight" (src/sceneast.cpp:89)
        case LightKind::Area: o += "srealight("; v3(L.p); o += ','; v3(L.u); o += ','; v3(L.v); o += ','; v3(L.I); o += ')'; break;
        case LightKind::Cone:
            o += "conelight(";
            v3(L.p);
            o += ',';
            v3(L.u);
            o += ',';
            num(L.angle);
            o += ',';
            v3(L.I);
            o += ')';
            break;
        case LightKind::Env: o += "environmentmap("; str(L.file); o += ','; v3(L.I); o += ')'; break;
        }
    }

    void value(const Value& v)
    {
        switch (v.kind) {
        case ValKind::Lights:
            o += "lights{";
            for (const LightExpr& L : v.lights) light(L);
            o += '}';
            break;
        case ValKind::Camera:
            o += "camera(";
            v3(v.cam.eye);
            o += ',';
            v3(v.cam.look);
            o += ',';
            v3(v.cam.up);
            o += ',';
            num(v.cam.fov);
            o += ',';
            num(v.cam.xres);
            o += ',';
            num(v.cam.yres);
            o += ')';
            break;
        case ValKind::Intersectable: node(v.node); break;
        case ValKind::Material: mat(v.mat); break;
        case ValKind::Ident: o += v.ident; break;
        }
    }
};

// ---------------------------------------------------------------------------
// Evaluation (src/sceneast.cpp:589-695)
// ---------------------------------------------------------------------------
struct BuiltMat {
    rt_material m;
    std::string unsupported;
};

struct Geometry {
    std::vector<float> tri_p;   // 9 per triangle
    std::vector<float> tri_n;   // 9 per triangle (valid where tri_has_n)
    std::vector<uint8_t> tri_has_n;
    std::vector<int32_t> tri_mat, tri_grp;
    std::vector<float> tri_tc;  // 6 per triangle: texture coordinates at p1, p2, p3
    std::vector<float> sph;     // 4 per sphere
    std::vector<int32_t> sph_mat, sph_grp;
    std::vector<rt_csg_node> csg;           // postfix programs of planes, boxes, CSG trees
    std::vector<int32_t> obj_first{0}, obj_grp;
    std::vector<float> obj_xf;              // 12 per object: world-to-object [A | b], identity if none
    bool any_obj_xf = false;
    int32_t groups = 0;
    std::string unsupported;
};

void cross(const float* a, const float* b, float* o)
{
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

// Triangle(p1,p2,p3) without normals: n = (p1 - p2) x (p1 - p3) (src/triangle.h:35)
void flat_normal(const float* p1, const float* p2, const float* p3, float* n)
{
    float a[3] = {p1[0] - p2[0], p1[1] - p2[1], p1[2] - p2[2]};
    float b[3] = {p1[0] - p3[0], p1[1] - p3[1], p1[2] - p3[2]};
    cross(a, b, n);
}

class Evaluator {
public:
    Evaluator(const Doc& d, std::string& log) : d_(d), log_(log) {}

    std::map<std::string, CamExpr> cameras;
    std::map<std::string, std::shared_ptr<Geometry>> objects;
    std::map<std::string, std::vector<LightExpr>> lights;
    std::map<std::string, int> materials;  // index into built
    std::vector<BuiltMat> built;

    void run()
    {
        for (const Assignment& a : d_.as) {
            const std::string& name = a.name;
            cameras.erase(name);
            objects.erase(name);
            lights.erase(name);
            materials.erase(name);
            const Value& v = a.value;
            switch (v.kind) {
            case ValKind::Camera: cameras[name] = v.cam; break;
            case ValKind::Lights: lights[name] = v.lights; break;
            case ValKind::Material: materials[name] = build_literal(d_.mats[v.mat]); break;
            case ValKind::Intersectable: {
                auto g = std::make_shared<Geometry>();
                build(v.node, nullptr, *g, -1);
                objects[name] = g;
                break;
            }
            case ValKind::Ident:
                if (cameras.count(v.ident)) {
                    cameras[name] = cameras[v.ident];
                } else if (objects.count(v.ident)) {
                    objects[name] = objects[v.ident];
                } else if (lights.count(v.ident)) {
                    lights[name] = lights[v.ident];
                } else if (materials.count(v.ident)) {
                    materials[name] = materials[v.ident];
                } else {
                    log_ += "Could not assign to " + name + ", no such variable: " + v.ident + "\n";
                }
                break;
            }
        }
    }

    int dark_matter()
    {
        if (dark_ < 0) {
            BuiltMat b;
            std::memset(&b.m, 0, sizeof b.m);
            b.m.type = RT_MAT_DARK;
            dark_ = (int)built.size();
            built.push_back(b);
        }
        return dark_;
    }

private:
    const Doc& d_;
    std::string& log_;
    int dark_ = -1;

    // src/sceneast.cpp:315-385 (material_builder)
    int build_literal(const MatExpr& m)
    {
        BuiltMat b;
        std::memset(&b.m, 0, sizeof b.m);
        switch (m.kind) {
        case MatKind::Diffuse:
            b.m.type = RT_MAT_DIFFUSE;
            b.m.kd[0] = m.a.x, b.m.kd[1] = m.a.y, b.m.kd[2] = m.a.z;
            break;
        case MatKind::Phong:
            b.m.type = RT_MAT_PHONG;
            b.m.kd[0] = m.a.x, b.m.kd[1] = m.a.y, b.m.kd[2] = m.a.z;
            b.m.ks[0] = m.b.x, b.m.ks[1] = m.b.y, b.m.ks[2] = m.b.z;
            b.m.exponent = m.f;
            break;
        case MatKind::Mirror: b.m.type = RT_MAT_MIRROR; break;
        case MatKind::Blinn:
            b.m.type = RT_MAT_BLINN;
            b.m.kd[0] = m.a.x, b.m.kd[1] = m.a.y, b.m.kd[2] = m.a.z;
            b.m.ks[0] = m.b.x, b.m.ks[1] = m.b.y, b.m.ks[2] = m.b.z;
            b.m.exponent = m.f;
            break;
        case MatKind::Refractive:
            b.m.type = RT_MAT_REFRACTIVE;
            b.m.exponent = m.f;
            break;
        case MatKind::Texture:
        case MatKind::ScaledTexture: {
            // TextureMaterial(coefficient, gamma) + load (src/sceneast.cpp:345-365,
            // src/texturematerial.cpp:46-60); an unreadable image is DarkMatter
            int w = 0, h = 0;
            std::vector<float> img;
            std::string err;
            if (!rt_read_image8(m.s, w, h, img, err)) {
                log_ += "Error loading texture \"" + m.s + "\"\n";
                return dark_matter();
            }
            b.m.type = RT_MAT_TEXTURE;
            if (m.kind == MatKind::ScaledTexture) {
                b.m.kd[0] = m.a.x, b.m.kd[1] = m.a.y, b.m.kd[2] = m.a.z;
                b.m.exponent = m.g;
            } else {
                b.m.kd[0] = b.m.kd[1] = b.m.kd[2] = 1.f;
                b.m.exponent = 1.f;
            }
            b.m.texture = (int32_t)tex_images.size();
            tex_images.push_back(std::move(img));
            tex_w.push_back(w);
            tex_h.push_back(h);
            break;
        }
        case MatKind::Ident: return resolve(m);
        }
        int idx = (int)built.size();
        built.push_back(b);
        return idx;
    }

    int resolve(const MatExpr& m)
    {
        if (m.kind != MatKind::Ident) return build_literal(m);
        auto it = materials.find(m.s);
        if (it == materials.end()) {
            if (missing_.insert(m.s).second) log_ += "No such variable: " + m.s + "\n";
            return dark_matter();
        }
        return it->second;
    }

    int material_of(const Node& n, Geometry& g)
    {
        // literal materials used by many primitives (OBJ faces) resolve once
        auto it = cache_.find(n.mat);
        int idx;
        if (it != cache_.end() && d_.mats[n.mat].kind != MatKind::Ident) {
            idx = it->second;
        } else {
            idx = resolve(d_.mats[n.mat]);
            if (d_.mats[n.mat].kind != MatKind::Ident) cache_[n.mat] = idx;
        }
        if (!built[idx].unsupported.empty() && g.unsupported.empty()) g.unsupported = built[idx].unsupported;
        return idx;
    }
    std::map<int, int> cache_;
    std::set<std::string> missing_;
  public:
    std::vector<std::vector<float>> tex_images;  // texture images in material order
    std::vector<int> tex_w, tex_h;
  private:  // the reference prints once per use; here once per name

    void emit_tri(Geometry& g, const float* p1, const float* p2, const float* p3, const float* n, bool has_n, int mat,
                  int grp, const M4* xf, const float* tc = nullptr)
    {
        // the Triangle constructors without texture coordinates use (0, 0), (0, 1),
        // (1, 0) (src/triangle.h:35-36); t(...) and OBJ faces pass theirs
        static const float def_tc[6] = {0, 0, 0, 1, 1, 0};
        const float* t = tc ? tc : def_tc;
        g.tri_tc.insert(g.tri_tc.end(), t, t + 6);
        const float* P[3] = {p1, p2, p3};
        float q[9];
        for (int k = 0; k < 3; ++k) {
            if (xf) {
                V3 r = xf->map(V3{P[k][0], P[k][1], P[k][2]});
                q[3 * k] = r.x, q[3 * k + 1] = r.y, q[3 * k + 2] = r.z;
            } else {
                std::memcpy(q + 3 * k, P[k], 12);
            }
        }
        float nn[9] = {};
        bool hn = has_n;
        if (xf) {
            // the normal matrix applied to the object-space normals (src/hitrecord.h:79-84)
            float src[9];
            if (has_n) {
                std::memcpy(src, n, sizeof src);
            } else {
                flat_normal(p1, p2, p3, src);
                std::memcpy(src + 3, src, 12);
                std::memcpy(src + 6, src, 12);
            }
            double N[3][3];
            xf->normal_matrix(N);
            for (int k = 0; k < 3; ++k)
                for (int i = 0; i < 3; ++i)
                    nn[3 * k + i] = (float)(N[i][0] * src[3 * k] + N[i][1] * src[3 * k + 1] + N[i][2] * src[3 * k + 2]);
            hn = true;
        } else if (has_n) {
            std::memcpy(nn, n, sizeof nn);
        }
        g.tri_p.insert(g.tri_p.end(), q, q + 9);
        g.tri_n.insert(g.tri_n.end(), nn, nn + 9);
        g.tri_has_n.push_back(hn ? 1 : 0);
        g.tri_mat.push_back(mat);
        g.tri_grp.push_back(grp);
    }

    static bool is_null(const float* v) { return v[0] == 0.f && v[1] == 0.f && v[2] == 0.f; }

    // Walks an intersectable in containedIntersectables order
    // (src/intersectablelist.cpp:67-78, src/bvh.cpp:150-160); grp >= 0
    // puts everything under one primitive ID (an instance).
    void build(const Node& n, const M4* xf, Geometry& g, int grp)
    {
        switch (n.kind) {
        case NodeKind::List:
        case NodeKind::Bvh:
            for (const Node& c : n.kids) build(c, xf, g, grp);
            break;
        case NodeKind::Tri: {
            int id = grp >= 0 ? grp : g.groups++;
            int mat = material_of(n, g);
            // t(...) with a null normal is the flat-normal constructor (src/sceneast.cpp:510-520)
            bool has_n = !(is_null(n.f + 9) || is_null(n.f + 12) || is_null(n.f + 15));
            emit_tri(g, n.f, n.f + 3, n.f + 6, n.f + 9, has_n, mat, id, xf, n.f + 18);
            break;
        }
        case NodeKind::Quad: {  // Triangle(p1,p2,p4), Triangle(p4,p2,p3)
            int id = grp >= 0 ? grp : g.groups++;
            int mat = material_of(n, g);
            emit_tri(g, n.f, n.f + 3, n.f + 9, nullptr, false, mat, id, xf);
            emit_tri(g, n.f + 9, n.f + 3, n.f + 6, nullptr, false, mat, id, xf);
            break;
        }
        case NodeKind::Sphere: {
            int id = grp >= 0 ? grp : g.groups++;
            int mat = material_of(n, g);
            float c[3] = {n.f[0], n.f[1], n.f[2]}, r = n.f[3];
            if (xf) {
                float s;
                if (!similarity(*xf, s)) {
                    /* an ellipsoid: the sphere as an analytic object in its
                     * own space, the ray mapped there */
                    float w2o[12];
                    if (!world_to_object(*xf, w2o)) {
                        if (g.unsupported.empty()) g.unsupported = "sphere under a singular or projective instance transform";
                        break;
                    }
                    rt_csg_node nd{};
                    nd.op = RT_CSG_SPHERE;
                    nd.material = mat;
                    nd.p[0] = c[0], nd.p[1] = c[1], nd.p[2] = c[2], nd.p[3] = r;
                    g.csg.push_back(nd);
                    add_object(g, id, w2o);
                    break;
                }
                V3 q = xf->map(V3{c[0], c[1], c[2]});
                c[0] = q.x, c[1] = q.y, c[2] = q.z;
                r *= s;
            }
            g.sph.insert(g.sph.end(), {c[0], c[1], c[2], r});
            g.sph_mat.push_back(mat);
            g.sph_grp.push_back(id);
            break;
        }
        case NodeKind::Instance: {
            int id = grp >= 0 ? grp : g.groups++;
            M4 M = eval_matrix(d_.xforms[n.aux]);
            M4 W = xf ? (*xf) * M : M;
            build(n.kids[0], &W, g, id);
            break;
        }
        case NodeKind::Plane: {  // src/sceneast.cpp plane(...): one object
            int id = grp >= 0 ? grp : g.groups++;
            rt_csg_node nd{};
            nd.op = RT_CSG_PLANE;
            nd.material = material_of(n, g);
            for (int k = 0; k < 4; ++k) nd.p[k] = n.f[k];
            if (xf) {
                // object-space plane P: the world plane is P M^-1 (its normal is
                // the normal matrix applied to P's, as the instance reports it)
                double inv[4][4];
                if (!invert(*xf, inv)) {
                    if (g.unsupported.empty()) g.unsupported = "plane under a singular instance transform";
                    break;
                }
                for (int j = 0; j < 4; ++j) {
                    double v = 0;
                    for (int i = 0; i < 4; ++i) v += (double)n.f[i] * inv[i][j];
                    nd.p[j] = (float)v;
                }
            }
            g.csg.push_back(nd);
            add_object(g, id, nullptr);
            break;
        }
        case NodeKind::Box:
        case NodeKind::Isect:
        case NodeKind::Union:
        case NodeKind::Diff:
        case NodeKind::Xor: {
            int id = grp >= 0 ? grp : g.groups++;
            const size_t start = g.csg.size();
            int leaves = 0;
            std::string why;
            bool ok = csg_program(n, xf, g, leaves, why);
            float w2o[12];
            bool mapped = false;
            if (!ok && xf && why.find("instance transform") != std::string::npos) {
                /* leaves the transform cannot be folded into (a sphere under a
                 * non-similarity, a rotated or sheared box): the program in
                 * object space and the ray mapped into it */
                g.csg.resize(start);
                leaves = 0;
                why.clear();
                if (!world_to_object(*xf, w2o)) {
                    why = "CSG object under a singular or projective instance transform";
                } else {
                    ok = csg_program(n, nullptr, g, leaves, why);
                    mapped = ok;
                }
            }
            if (!ok || leaves > RT_CSG_MAX_LEAVES) {
                g.csg.resize(start);
                if (why.empty()) why = "CSG object with more than 16 spheres and boxes";
                if (g.unsupported.empty()) g.unsupported = why;
                break;
            }
            add_object(g, id, mapped ? w2o : nullptr);
            break;
        }
        case NodeKind::Obj:
            if (g.unsupported.empty()) g.unsupported = "unresolved obj";
            break;
        }
    }

    // Postfix program of a CSG tree (src/sceneast.cpp csg builders): leaves
    // spheres and boxes, then the operation.  Under an instance transform a
    // sphere needs a similarity and a box a positive axis scale plus
    // translation, so that it stays a sphere or an axis-aligned box.
    bool csg_program(const Node& n, const M4* xf, Geometry& g, int& leaves, std::string& why)
    {
        rt_csg_node nd{};
        switch (n.kind) {
        case NodeKind::Sphere: {
            nd.op = RT_CSG_SPHERE;
            nd.material = material_of(n, g);
            float c[3] = {n.f[0], n.f[1], n.f[2]}, r = n.f[3];
            if (xf) {
                float sc;
                if (!similarity(*xf, sc)) {
                    why = "sphere under a non-similarity instance transform";
                    return false;
                }
                V3 q = xf->map(V3{c[0], c[1], c[2]});
                c[0] = q.x, c[1] = q.y, c[2] = q.z;
                r *= sc;
            }
            nd.p[0] = c[0], nd.p[1] = c[1], nd.p[2] = c[2], nd.p[3] = r;
            ++leaves;
            break;
        }
        case NodeKind::Box: {
            nd.op = RT_CSG_BOX;
            nd.material = material_of(n, g);
            for (int k = 0; k < 6; ++k) nd.p[k] = n.f[k];
            if (xf) {
                const M4& M = *xf;
                bool ok = M.m[3][0] == 0 && M.m[3][1] == 0 && M.m[3][2] == 0 && M.m[3][3] == 1;
                for (int i = 0; i < 3 && ok; ++i)
                    for (int j = 0; j < 3 && ok; ++j) ok = i == j ? M.m[i][j] > 0 : M.m[i][j] == 0;
                if (!ok) {
                    why = "box under an instance transform other than scale and translation";
                    return false;
                }
                V3 lo = M.map(V3{n.f[0], n.f[1], n.f[2]}), hi = M.map(V3{n.f[3], n.f[4], n.f[5]});
                nd.p[0] = lo.x, nd.p[1] = lo.y, nd.p[2] = lo.z, nd.p[3] = hi.x, nd.p[4] = hi.y, nd.p[5] = hi.z;
            }
            ++leaves;
            break;
        }
        case NodeKind::Isect:
        case NodeKind::Union:
        case NodeKind::Diff:
        case NodeKind::Xor:
            if (!csg_program(n.kids[0], xf, g, leaves, why) || !csg_program(n.kids[1], xf, g, leaves, why))
                return false;
            nd.op = n.kind == NodeKind::Isect ? RT_CSG_ISECT
                    : n.kind == NodeKind::Union ? RT_CSG_UNION
                    : n.kind == NodeKind::Diff  ? RT_CSG_DIFF
                                                : RT_CSG_XOR;
            break;
        default:
            why = "CSG operand that is not a sphere, box or CSG operation";
            return false;
        }
        g.csg.push_back(nd);
        return true;
    }

    // Close the object just appended to g.csg, with its world-to-object map
    // (nullptr: the object is in world space).
    static void add_object(Geometry& g, int id, const float* w2o)
    {
        static const float identity[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
        g.obj_first.push_back((int32_t)g.csg.size());
        g.obj_grp.push_back(id);
        g.obj_xf.insert(g.obj_xf.end(), w2o ? w2o : identity, (w2o ? w2o : identity) + 12);
        g.any_obj_xf |= w2o != nullptr;
    }

    // The inverse of an affine instance matrix as 3 x 4 rows [A | b]; false
    // for a projective or singular matrix.
    static bool world_to_object(const M4& M, float out[12])
    {
        if (M.m[3][0] != 0 || M.m[3][1] != 0 || M.m[3][2] != 0 || M.m[3][3] != 1) return false;
        double inv[4][4];
        if (!invert(M, inv)) return false;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 4; ++j) out[4 * i + j] = (float)inv[i][j];
        return true;
    }

    static bool invert(const M4& M, double inv[4][4])
    {
        double a[4][8];
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 8; ++j) a[i][j] = j < 4 ? (double)M.m[i][j] : (j - 4 == i ? 1.0 : 0.0);
        for (int c = 0; c < 4; ++c) {
            int piv = c;
            for (int r = c + 1; r < 4; ++r)
                if (std::fabs(a[r][c]) > std::fabs(a[piv][c])) piv = r;
            if (a[piv][c] == 0) return false;
            if (piv != c)
                for (int j = 0; j < 8; ++j) std::swap(a[c][j], a[piv][j]);
            double d = a[c][c];
            for (int j = 0; j < 8; ++j) a[c][j] /= d;
            for (int r = 0; r < 4; ++r)
                if (r != c) {
                    double f = a[r][c];
                    for (int j = 0; j < 8; ++j) a[r][j] -= f * a[c][j];
                }
        }
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) inv[i][j] = a[i][4 + j];
        return true;
    }

    // upper 3x3 = s * rotation, no projective row
    static bool similarity(const M4& M, float& s)
    {
        if (M.m[3][0] != 0 || M.m[3][1] != 0 || M.m[3][2] != 0 || M.m[3][3] != 1) return false;
        double c[3][3], len[3];
        for (int j = 0; j < 3; ++j) {
            for (int i = 0; i < 3; ++i) c[j][i] = M.m[i][j];
            len[j] = std::sqrt(c[j][0] * c[j][0] + c[j][1] * c[j][1] + c[j][2] * c[j][2]);
        }
        if (len[0] == 0) return false;
        for (int j = 1; j < 3; ++j)
            if (std::fabs(len[j] - len[0]) > 1e-5 * len[0]) return false;
        for (int a = 0; a < 3; ++a)
            for (int b = a + 1; b < 3; ++b)
                if (std::fabs(c[a][0] * c[b][0] + c[a][1] * c[b][1] + c[a][2] * c[b][2]) > 1e-5 * len[0] * len[0])
                    return false;
        s = (float)len[0];
        return true;
    }
};

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
struct rt_scene_file {
    Doc doc;
    std::string load_log;  // parse and OBJ warnings
    std::string messages;  // load_log + the warnings of the last evaluation
    std::vector<float> positions, normals, spheres;
    std::vector<int32_t> tri_vertices, tri_material, tri_prim, sph_material, sph_prim;
    std::vector<rt_csg_node> csg_nodes;
    std::vector<int32_t> object_first, object_prim;
    std::vector<float> object_xform;
    std::vector<float> tri_tc;
    std::vector<std::vector<float>> tex_rgb;  // images of texture materials
    std::vector<rt_texture> textures;
    std::vector<rt_material> materials;
    std::vector<rt_light> lights;
    std::vector<float> env_rgb;  /* light probe of an environmentmap(...) light */
    int env_w = 0, env_h = 0;
};

namespace {

int fail(int status, const std::string& msg)
{
    g_last_error = msg;
    return status;
}

int finish_load(const std::string& text, const std::string& base_dir, rt_scene_file** out)
{
    std::unique_ptr<rt_scene_file> f(new rt_scene_file());
    std::string err;
    Parser p(text, f->doc);
    if (!p.parse(err)) return fail(RT_ERR_PARSE, "Failed to parse scene: " + err);
    for (Assignment& a : f->doc.as)
        if (a.value.kind == ValKind::Intersectable) apply_path(a.value.node, f->doc, base_dir);
    if (!resolve_vars(f->doc, f->load_log, err)) return fail(RT_ERR_PARSE, err);
    f->messages = f->load_log;
    *out = f.release();
    g_last_error.clear();
    return RT_OK;
}

std::string dump_text(const rt_scene_file* f)
{
    // createBVH (src/sceneast.cpp:749-766) on a copy of the assignments
    Doc d = f->doc;
    MatExpr black;  // ast_diffuse_material {ast_vector3_literal()} for the node boxes
    black.kind = MatKind::Diffuse;
    int bb_mat = (int)d.mats.size();
    d.mats.push_back(black);
    for (Assignment& a : d.as) {
        if (a.value.kind != ValKind::Intersectable) continue;
        std::vector<Node> leaves;
        flatten(a.value.node, leaves);
        a.value.node = bvh_create(std::move(leaves), d, bb_mat);
    }
    std::string o;
    Printer pr{d, o};
    bool first = true;
    for (const Assignment& a : d.as) {
        if (first && a.name.size() < 10) o.append(10 - a.name.size(), ' ');  // SceneDumper's setw(10)
        first = false;
        o += a.name;
        o += '=';
        pr.value(a.value);
        o += ";\n";
    }
    return o;
}

}  // namespace

extern "C" {

int rt_scene_file_load(const char* path, rt_scene_file** out)
{
    if (!path || !out) return fail(RT_ERR_INVALID, "null argument");
    *out = nullptr;
    std::string text;
    if (!read_file(path, text)) return fail(RT_ERR_IO, std::string("Failed to open file ") + path);
    if (text.empty()) return fail(RT_ERR_IO, std::string("Failed to open file ") + path);  // in.eof() check
    return finish_load(text, dir_of(path), out);
}

int rt_scene_file_parse(const char* text, const char* base_dir, rt_scene_file** out)
{
    if (!text || !out) return fail(RT_ERR_INVALID, "null argument");
    *out = nullptr;
    return finish_load(text, base_dir ? base_dir : ".", out);
}

const char* rt_scene_file_last_error(void) { return g_last_error.c_str(); }

const char* rt_scene_file_messages(const rt_scene_file* f) { return f ? f->messages.c_str() : ""; }

int rt_scene_file_scene(rt_scene_file* f, rt_scene_desc* desc, rt_camera* camera)
{
    if (!f || !desc || !camera) return fail(RT_ERR_INVALID, "null argument");
    f->messages = f->load_log;
    Evaluator ev(f->doc, f->messages);
    ev.run();
    auto cam = ev.cameras.find("camera");
    if (cam == ev.cameras.end()) return fail(RT_ERR_INVALID, "the scene defines no variable \"camera\"");
    auto obj = ev.objects.find("intersectable");
    if (obj == ev.objects.end()) return fail(RT_ERR_INVALID, "the scene defines no variable \"intersectable\"");
    const Geometry& g = *obj->second;
    if (!g.unsupported.empty()) return fail(RT_ERR_UNSUPPORTED, g.unsupported + " is not rendered by the GPU path");

    // lights first (src/sceneast.cpp:676-678)
    f->lights.clear();
    f->env_rgb.clear();
    f->env_w = f->env_h = 0;
    auto li = ev.lights.find("lights");
    if (li != ev.lights.end()) {
        for (const LightExpr& L : li->second) {
            rt_light c;
            std::memset(&c, 0, sizeof c);
            switch (L.kind) {
            case LightKind::Point:
                c.type = RT_LIGHT_POINT;
                c.p[0] = L.p.x, c.p[1] = L.p.y, c.p[2] = L.p.z;
                c.intensity[0] = L.I.x, c.intensity[1] = L.I.y, c.intensity[2] = L.I.z;
                break;
            case LightKind::Area:
                c.type = RT_LIGHT_AREA;
                c.p[0] = L.p.x, c.p[1] = L.p.y, c.p[2] = L.p.z;
                c.u[0] = L.u.x, c.u[1] = L.u.y, c.u[2] = L.u.z;
                c.v[0] = L.v.x, c.v[1] = L.v.y, c.v[2] = L.v.z;
                c.intensity[0] = L.I.x, c.intensity[1] = L.I.y, c.intensity[2] = L.I.z;
                break;
            case LightKind::Cone:
                c.type = RT_LIGHT_CONE;
                c.p[0] = L.p.x, c.p[1] = L.p.y, c.p[2] = L.p.z;
                c.u[0] = L.u.x, c.u[1] = L.u.y, c.u[2] = L.u.z;
                c.v[0] = L.angle;
                c.intensity[0] = L.I.x, c.intensity[1] = L.I.y, c.intensity[2] = L.I.z;
                break;
            case LightKind::Env: {
                // EnvironmentMap::load (src/environmentmap.cpp:20-33): an angular
                // light probe, the path taken as given (relative to the working
                // directory, as imread takes it); failure -> BlackLight
                // (src/sceneast.cpp:574-584): no emission
                std::string err;
                std::vector<float> img;
                int w = 0, h = 0;
                if (!f->env_rgb.empty() || !rt_read_rgbe(L.file, w, h, img, err)) {
                    f->messages += "Failed to load environment map \"" + L.file + "\"" +
                                   (err.empty() ? std::string(" (one environment map per scene)") : ": " + err) + "\n";
                    c.type = RT_LIGHT_POINT;
                    break;
                }
                f->env_rgb.swap(img);
                f->env_w = w;
                f->env_h = h;
                c.type = RT_LIGHT_ENV;
                c.intensity[0] = L.I.x, c.intensity[1] = L.I.y, c.intensity[2] = L.I.z;
                break;
            }
            }
            f->lights.push_back(c);
        }
    }
    const int32_t nL = (int32_t)f->lights.size();

    f->materials.resize(ev.built.size());
    for (size_t i = 0; i < ev.built.size(); ++i) f->materials[i] = ev.built[i].m;

    const size_t nT = g.tri_mat.size();
    bool any_n = false;
    for (uint8_t h : g.tri_has_n) any_n |= h != 0;
    f->positions.assign(g.tri_p.begin(), g.tri_p.end());
    f->normals.clear();
    if (any_n) {
        f->normals.resize(9 * nT);
        for (size_t i = 0; i < nT; ++i) {
            if (g.tri_has_n[i]) {
                std::memcpy(&f->normals[9 * i], &g.tri_n[9 * i], 36);
            } else {
                const float* p = &g.tri_p[9 * i];
                float n[3];
                flat_normal(p, p + 3, p + 6, n);
                for (int k = 0; k < 3; ++k) std::memcpy(&f->normals[9 * i + 3 * k], n, 12);
            }
        }
    }
    f->tri_vertices.resize(3 * nT);
    for (size_t i = 0; i < 3 * nT; ++i) f->tri_vertices[i] = (int32_t)i;
    f->tri_material.assign(g.tri_mat.begin(), g.tri_mat.end());
    f->tri_prim.resize(nT);
    for (size_t i = 0; i < nT; ++i) f->tri_prim[i] = nL + g.tri_grp[i];
    f->spheres.assign(g.sph.begin(), g.sph.end());
    f->sph_material.assign(g.sph_mat.begin(), g.sph_mat.end());
    f->sph_prim.resize(g.sph_grp.size());
    for (size_t i = 0; i < g.sph_grp.size(); ++i) f->sph_prim[i] = nL + g.sph_grp[i];

    f->tri_tc = g.tri_tc;
    f->csg_nodes = g.csg;
    f->object_first = g.obj_first;
    f->object_xform = g.obj_xf;
    f->object_prim.resize(g.obj_grp.size());
    for (size_t i = 0; i < g.obj_grp.size(); ++i) f->object_prim[i] = nL + g.obj_grp[i];

    if (3 * nT > (size_t)std::numeric_limits<int32_t>::max())
        return fail(RT_ERR_INVALID, "too many triangles for one scene description");

    std::memset(desc, 0, sizeof *desc);
    desc->n_vertices = (int32_t)(3 * nT);
    desc->positions = nT ? f->positions.data() : nullptr;
    desc->normals = any_n ? f->normals.data() : nullptr;
    desc->n_triangles = (int32_t)nT;
    desc->tri_vertices = nT ? f->tri_vertices.data() : nullptr;
    desc->tri_material = nT ? f->tri_material.data() : nullptr;
    desc->tri_prim_id = nT ? f->tri_prim.data() : nullptr;
    desc->n_spheres = (int32_t)f->sph_material.size();
    desc->spheres = f->spheres.empty() ? nullptr : f->spheres.data();
    desc->sphere_material = f->sph_material.empty() ? nullptr : f->sph_material.data();
    desc->sphere_prim_id = f->sph_prim.empty() ? nullptr : f->sph_prim.data();
    desc->n_materials = (int32_t)f->materials.size();
    desc->materials = f->materials.empty() ? nullptr : f->materials.data();
    desc->n_lights = nL;
    desc->lights = f->lights.empty() ? nullptr : f->lights.data();
    desc->tri_texcoords = nT ? f->tri_tc.data() : nullptr;
    f->textures.clear();
    for (size_t i = 0; i < ev.tex_images.size(); ++i) {
        rt_texture t{};
        t.width = ev.tex_w[i];
        t.height = ev.tex_h[i];
        t.rgb = ev.tex_images[i].data();
        f->textures.push_back(t);
    }
    f->tex_rgb = std::move(ev.tex_images);
    for (size_t i = 0; i < f->textures.size(); ++i) f->textures[i].rgb = f->tex_rgb[i].data();
    desc->n_textures = (int32_t)f->textures.size();
    desc->textures = f->textures.empty() ? nullptr : f->textures.data();
    if (!f->object_prim.empty()) {
        desc->n_objects = (int32_t)f->object_prim.size();
        desc->csg_nodes = f->csg_nodes.data();
        desc->object_first = f->object_first.data();
        desc->object_prim_id = f->object_prim.data();
        desc->object_xform = g.any_obj_xf ? f->object_xform.data() : nullptr;
    }
    if (!f->env_rgb.empty()) {
        desc->env_width = f->env_w;
        desc->env_height = f->env_h;
        desc->env_rgb = f->env_rgb.data();
        desc->env_kind = 1;
    }

    const CamExpr& c = cam->second;
    std::memset(camera, 0, sizeof *camera);
    camera->eye[0] = c.eye.x, camera->eye[1] = c.eye.y, camera->eye[2] = c.eye.z;
    camera->look_at[0] = c.look.x, camera->look_at[1] = c.look.y, camera->look_at[2] = c.look.z;
    camera->up[0] = c.up.x, camera->up[1] = c.up.y, camera->up[2] = c.up.z;
    camera->fov = c.fov;
    camera->width = (int32_t)c.xres;  // QSize(int, int) from the float arguments
    camera->height = (int32_t)c.yres;
    g_last_error.clear();
    return RT_OK;
}

int64_t rt_scene_file_dump_bvh_text(const rt_scene_file* f, char* buf, int64_t cap)
{
    if (!f) return fail(RT_ERR_INVALID, "null argument");
    std::string t = dump_text(f);
    if (buf && cap > 0) {
        size_t n = std::min((size_t)cap - 1, t.size());
        std::memcpy(buf, t.data(), n);
        buf[n] = 0;
    }
    return (int64_t)t.size();
}

int rt_scene_file_dump_bvh(const rt_scene_file* f, const char* path)
{
    if (!f || !path) return fail(RT_ERR_INVALID, "null argument");
    std::string t = dump_text(f);
    FILE* fp = std::fopen(path, "wb");
    if (!fp) return fail(RT_ERR_IO, std::string("cannot write ") + path);
    size_t w = std::fwrite(t.data(), 1, t.size(), fp);
    int rc = std::fclose(fp);
    if (w != t.size() || rc != 0) return fail(RT_ERR_IO, std::string("cannot write ") + path);
    return RT_OK;
}

void rt_scene_file_free(rt_scene_file* f) { delete f; }

}  // extern "C"
