"""ctypes wrapper of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py, as the checker.  The product never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")


def build(verbose=False):
    res = subprocess.run(["make", "-C", HERE, "-B", "liboracle.so"], capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + res.stdout + res.stderr)
    if verbose:
        print(res.stdout)
    return LIB_PATH


_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    from renderer_amd import _abi  # scene / view / params structs (input format)

    lib = ctypes.CDLL(LIB_PATH)
    P, I64 = ctypes.c_void_p, ctypes.c_int64
    lib.oc_philox4x32_10.argtypes = [P, P, P]
    lib.oc_scene_create.restype = P
    lib.oc_scene_create.argtypes = [ctypes.POINTER(_abi.RtSceneDesc), ctypes.c_int]
    lib.oc_scene_destroy.argtypes = [P]
    lib.oc_trace_closest.argtypes = [P, I64, P, P, P, P, ctypes.c_int]
    lib.oc_trace_any.argtypes = [P, I64, P, P, P, ctypes.c_int]
    lib.oc_flat_trace.argtypes = [P, I64, P, I64, P, P, ctypes.c_int, P, P, P, P, ctypes.c_int]
    lib.oc_hit_normals.argtypes = [P, I64, P, P, P]
    lib.oc_hit_texcoords.argtypes = [P, I64, P, P, P]
    lib.oc_texture_albedo.restype = ctypes.c_int
    lib.oc_texture_albedo.argtypes = [P, ctypes.c_int, ctypes.c_float, ctypes.c_float, P]
    lib.oc_intersect_prim.restype = ctypes.c_float
    lib.oc_intersect_prim.argtypes = [P, ctypes.c_int32, P, P]
    lib.oc_render.restype = I64
    lib.oc_render.argtypes = [P, ctypes.POINTER(_abi.RtView), ctypes.POINTER(_abi.RtParams), ctypes.c_uint32, P,
                              ctypes.c_int, ctypes.c_int, ctypes.c_int]
    lib.oc_integrate_rays.argtypes = [P, I64, P, P, ctypes.POINTER(_abi.RtParams), ctypes.c_uint32, P, ctypes.c_int]
    lib.oc_render_sample.argtypes = [P, ctypes.POINTER(_abi.RtView), ctypes.POINTER(_abi.RtParams),
                                     ctypes.c_uint32, ctypes.c_int, ctypes.c_int, ctypes.c_int, P]
    lib.oc_env_sample.argtypes = [P, ctypes.c_float, ctypes.c_float, P, P]
    lib.oc_env_pdf.restype = ctypes.c_float
    lib.oc_env_pdf.argtypes = [P, P]
    lib.oc_env_value.argtypes = [P, P, P]
    lib.oc_tonemap.argtypes = [P, I64, ctypes.c_float, P]
    lib.oc_transparent_dir.argtypes = [P, P, ctypes.c_float, ctypes.c_double, P]
    lib.oc_bsdf.argtypes = [P, ctypes.c_int, P, P, P, P, P]
    lib.oc_render_mlt.restype = ctypes.c_int
    lib.oc_render_mlt.argtypes = [P, ctypes.POINTER(_abi.RtCamera), ctypes.POINTER(_abi.RtMltParams), P,
                                  ctypes.POINTER(_abi.RtMltResult), ctypes.c_int]
    lib.oc_integrate_samples.argtypes = [P, ctypes.POINTER(_abi.RtCamera), ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                         P, I64, P, ctypes.c_int]
    lib.oc_mlt_perturb.argtypes = [P, P, P, I64, P]
    lib.oc_sym_filter.argtypes = [P, P, P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, P]
    lib.oc_box_blur.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int, P]
    lib.oc_sep3_reflect101.argtypes = [P, ctypes.c_int, ctypes.c_int, P]
    lib.oc_var2.argtypes = [P, P, I64, P, P]
    lib.oc_flt_importance.argtypes = [P, P, ctypes.c_int, ctypes.c_int, P]
    lib.oc_flt_accumulate.argtypes = [ctypes.c_int, ctypes.c_int, P, P, P, P, P, P, I64, P, P, P]
    lib.oc_bsdf_sample.restype = ctypes.c_int
    lib.oc_bsdf_sample.argtypes = [P, ctypes.c_int, P, P, ctypes.c_float, ctypes.c_float, ctypes.c_float, P]
    _lib = lib
    return lib


def philox(ctr, key):
    lib = load()
    c = np.ascontiguousarray(ctr, np.uint32)
    k = np.ascontiguousarray(key, np.uint32)
    out = np.zeros(4, np.uint32)
    lib.oc_philox4x32_10(c.ctypes.data, k.ctypes.data, out.ctypes.data)
    return out


def transparent_dir(d, n, ior, u):
    """TransparentMaterial::outDirection for direction d, normal n."""
    lib = load()
    a, b = np.ascontiguousarray(d, np.float32), np.ascontiguousarray(n, np.float32)
    out = np.zeros(3, np.float32)
    lib.oc_transparent_dir(a.ctypes.data, b.ctypes.data, float(ior), float(u), out.ctypes.data)
    return out


def tonemap(film, gamma=2.2):
    """Tonemapper::map of every channel value: uint8 array of film's shape."""
    lib = load()
    f = np.ascontiguousarray(film, np.float32)
    out = np.zeros(f.shape, np.uint8)
    lib.oc_tonemap(f.ctypes.data, f.size, float(gamma), out.ctypes.data)
    return out


def flat_trace(nodes, records, o_tmin, d_tmax, any_hit=False, threads=8):
    """Traversal of an exported 4-wide BVH in the device traversal's order.
    Closest hit: (g, t, visits, tests) with g the primitive index (-1: miss);
    any hit: (occluded, visits, tests)."""
    lib = load()
    nodes = np.ascontiguousarray(nodes, np.float32)
    records = np.ascontiguousarray(records, np.float32)
    o = np.ascontiguousarray(o_tmin, np.float32)
    d = np.ascontiguousarray(d_tmax, np.float32)
    n = len(o)
    g = np.zeros(n, np.int32)
    t = np.zeros(n, np.float32)
    visits = np.zeros(n, np.int32)
    tests = np.zeros(n, np.int32)
    lib.oc_flat_trace(nodes.ctypes.data, len(nodes), records.ctypes.data, n, o.ctypes.data, d.ctypes.data,
                      int(bool(any_hit)), g.ctypes.data, t.ctypes.data, visits.ctypes.data, tests.ctypes.data,
                      int(threads))
    if any_hit:
        return g.astype(np.uint8), visits, tests
    return g, t, visits, tests


class Oracle:
    """The reference algorithms on the CPU over a flattened scene.
    accel=1: the reference BVH (leaf = 1, both children); 0: linear list."""

    def __init__(self, arrays, accel=1):
        self.lib = load()
        self.arrays = arrays
        self.h = self.lib.oc_scene_create(ctypes.byref(arrays.desc()), int(accel))

    def close(self):
        if self.h:
            self.lib.oc_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def trace_closest(self, o_tmin, d_tmax, threads=8):
        o = np.ascontiguousarray(o_tmin, np.float32)
        d = np.ascontiguousarray(d_tmax, np.float32)
        n = len(o)
        prim = np.empty(n, np.int32)
        t = np.empty(n, np.float32)
        self.lib.oc_trace_closest(self.h, n, o.ctypes.data, d.ctypes.data, prim.ctypes.data, t.ctypes.data, threads)
        return prim, t

    def hit_normals(self, o_tmin, d_tmax):
        o = np.ascontiguousarray(o_tmin, np.float32)
        d = np.ascontiguousarray(d_tmax, np.float32)
        out = np.zeros((len(o), 3), np.float32)
        self.lib.oc_hit_normals(self.h, len(o), o.ctypes.data, d.ctypes.data, out.ctypes.data)
        return out

    def hit_texcoords(self, o_tmin, d_tmax):
        o = np.ascontiguousarray(o_tmin, np.float32)
        d = np.ascontiguousarray(d_tmax, np.float32)
        out = np.zeros((len(o), 2), np.float32)
        self.lib.oc_hit_texcoords(self.h, len(o), o.ctypes.data, d.ctypes.data, out.ctypes.data)
        return out

    def texture_albedo(self, mat, s, t):
        out = np.zeros(3, np.float32)
        if self.lib.oc_texture_albedo(self.h, int(mat), float(s), float(t), out.ctypes.data) != 0:
            raise ValueError("not a texture material")
        return out

    def trace_any(self, o_tmin, d_tmax, threads=8):
        o = np.ascontiguousarray(o_tmin, np.float32)
        d = np.ascontiguousarray(d_tmax, np.float32)
        occ = np.empty(len(o), np.uint8)
        self.lib.oc_trace_any(self.h, len(o), o.ctypes.data, d.ctypes.data, occ.ctypes.data, threads)
        return occ

    def intersect_prim(self, prim, o_tmin, d_tmax):
        o = np.ascontiguousarray(o_tmin, np.float32).reshape(4)
        d = np.ascontiguousarray(d_tmax, np.float32).reshape(4)
        return float(self.lib.oc_intersect_prim(self.h, int(prim), o.ctypes.data, d.ctypes.data))

    @staticmethod
    def _view_params(view, mode, max_depth, rr_start, rr_prob, seed):
        from renderer_amd import _abi

        cam, sx, sy = view
        v = _abi.RtView()
        v.camera = cam.to_c()
        v.spp_x, v.spp_y = int(sx), int(sy)
        p = _abi.RtParams()
        p.mode, p.max_depth, p.rr_start, p.rr_prob, p.seed = int(mode), int(max_depth), int(rr_start), float(rr_prob), int(seed)
        return v, p

    def render(self, view, mode=1, max_depth=8, rr_start=3, rr_prob=0.5, seed=0, view_id=0, rows=None, threads=8,
               film=None):
        """view = (Camera, spp_x, spp_y).  Returns (film (H, W, 3), rays traced)."""
        v, p = self._view_params(view, mode, max_depth, rr_start, rr_prob, seed)
        cam = view[0]
        if film is None:
            film = np.zeros((cam.height, cam.width, 3), np.float32)
        r0, r1 = (0, cam.height) if rows is None else rows
        rays = self.lib.oc_render(self.h, ctypes.byref(v), ctypes.byref(p), int(view_id), film.ctypes.data, r0, r1,
                                  threads)
        return film, int(rays)

    def integrate_rays(self, o_tmin, d_tmax, mode=1, max_depth=8, rr_start=3, rr_prob=0.5, seed=0, key0=0,
                       threads=8):
        """UniDiPathTracingIntegrator::integrate(const Ray&, ...) per ray: (n, 3)."""
        from renderer_amd import _abi
        o = np.ascontiguousarray(o_tmin, np.float32)
        d = np.ascontiguousarray(d_tmax, np.float32)
        p = _abi.RtParams()
        p.mode, p.max_depth, p.rr_start, p.rr_prob, p.seed = int(mode), int(max_depth), int(rr_start), float(rr_prob), int(seed)
        out = np.zeros((len(o), 3), np.float32)
        self.lib.oc_integrate_rays(self.h, len(o), o.ctypes.data, d.ctypes.data, ctypes.byref(p), int(key0) & 0xffffffff,
                                   out.ctypes.data, threads)
        return out

    def render_sample(self, view, px, py, s, mode=1, max_depth=8, rr_start=3, rr_prob=0.5, seed=0, view_id=0):
        v, p = self._view_params(view, mode, max_depth, rr_start, rr_prob, seed)
        out = np.zeros(3, np.float32)
        self.lib.oc_render_sample(self.h, ctypes.byref(v), ctypes.byref(p), int(view_id), px, py, s, out.ctypes.data)
        return out

    def render_mlt(self, camera, threads=8, film=None, **params):
        """Metropolis / ERPT restatement (GpuContext.make_mlt_params keywords).
        Returns (film, result dict)."""
        from renderer_amd.integrator import GpuContext
        p = GpuContext.make_mlt_params(**params)
        res = _abi_mod().RtMltResult()
        if film is None:
            film = np.zeros((camera.height, camera.width, 3), np.float32)
        cam = camera.to_c()
        st = self.lib.oc_render_mlt(self.h, ctypes.byref(cam), ctypes.byref(p), film.ctypes.data, ctypes.byref(res),
                                    threads)
        if st != 0:
            raise ValueError("bad Metropolis parameters")
        return film, {k: getattr(res, k) for k, _ in res._fields_}

    def integrate_samples(self, camera, samples, max_depth=5, rr_start=3, termination_prob=0.0, threads=8):
        x = np.ascontiguousarray(samples, np.float32)
        out = np.zeros((len(x), 3), np.float32)
        cam = camera.to_c()
        self.lib.oc_integrate_samples(self.h, ctypes.byref(cam), int(max_depth), int(rr_start), float(termination_prob),
                                      x.ctypes.data, len(x), out.ctypes.data, threads)
        return out

    def bsdf(self, mat, wo, wi, nf):
        """(value RGB, density) of the unbiased estimators' BSDF of material index mat."""
        a, b, n = (np.ascontiguousarray(v, np.float32) for v in (wo, wi, nf))
        f = np.zeros(3, np.float32)
        pdf = np.zeros(1, np.float32)
        self.lib.oc_bsdf(self.h, int(mat), a.ctypes.data, b.ctypes.data, n.ctypes.data, f.ctypes.data, pdf.ctypes.data)
        return f, float(pdf[0])

    def bsdf_sample(self, mat, wo, nf, u0, u1, u2):
        a, n = (np.ascontiguousarray(v, np.float32) for v in (wo, nf))
        wi = np.zeros(3, np.float32)
        ok = self.lib.oc_bsdf_sample(self.h, int(mat), a.ctypes.data, n.ctypes.data, float(u0), float(u1), float(u2),
                                     wi.ctypes.data)
        return (wi if ok else None)

    def env_value(self, d):
        a = np.ascontiguousarray(d, np.float32)
        out = np.zeros(3, np.float32)
        self.lib.oc_env_value(self.h, a.ctypes.data, out.ctypes.data)
        return out

    def env_sample(self, u0, u1):
        d = np.zeros(3, np.float32)
        pdf = np.zeros(1, np.float32)
        self.lib.oc_env_sample(self.h, float(u0), float(u1), d.ctypes.data, pdf.ctypes.data)
        return d, float(pdf[0])

    def env_pdf(self, d):
        d = np.ascontiguousarray(d, np.float32)
        return float(self.lib.oc_env_pdf(self.h, d.ctypes.data))


def _abi_mod():
    from renderer_amd import _abi
    return _abi


def mlt_perturb(s, u, v):
    """The primary-sample perturbation, elementwise over float32 arrays."""
    lib = load()
    s, u, v = (np.ascontiguousarray(a, np.float32) for a in (s, u, v))
    out = np.zeros_like(s)
    lib.oc_mlt_perturb(s.ctypes.data, u.ctypes.data, v.ctypes.data, len(s), out.ctypes.data)
    return out


def sym_filter(image, guide, pixvar, patch=7, radius=10):
    """SymmetricFilter(patch, radius).filter(image, guide, pixvar) on (H, W, 3) float32 images."""
    lib = load()
    a, g, v = (np.ascontiguousarray(x, np.float32) for x in (image, guide, pixvar))
    out = np.zeros_like(a)
    lib.oc_sym_filter(a.ctypes.data, g.ctypes.data, v.ctypes.data, a.shape[1], a.shape[0], int(patch), int(radius),
                      out.ctypes.data)
    return out


def box_blur(img, p):
    """cv::blur(img, Size(p, p)) with BORDER_REFLECT of an (H, W) float32 image, as the filter restates it."""
    lib = load()
    a = np.ascontiguousarray(img, np.float32)
    out = np.zeros_like(a)
    lib.oc_box_blur(a.ctypes.data, a.shape[1], a.shape[0], int(p), out.ctypes.data)
    return out


def sep3_reflect101(img):
    """sepFilter2D with [.1 .8 .1] both ways and BORDER_REFLECT_101 of an (H, W, 3) float32 image."""
    lib = load()
    a = np.ascontiguousarray(img, np.float32)
    out = np.zeros_like(a)
    lib.oc_sep3_reflect101(a.ctypes.data, a.shape[1], a.shape[0], out.ctypes.data)
    return out


def var2(a, b):
    """mathhelper var() of two float32 arrays: (mean, variance)."""
    lib = load()
    x, y = (np.ascontiguousarray(v, np.float32).ravel() for v in (a, b))
    m, v = np.zeros_like(x), np.zeros_like(x)
    lib.oc_var2(x.ctypes.data, y.ctypes.data, len(x), m.ctypes.data, v.ctypes.data)
    return m.reshape(np.shape(a)), v.reshape(np.shape(a))


def flt_importance(var_filtered, mean_filtered):
    """The filtered renderer's importance map of a later pass from the (H, W, 3)
    variance and mean of the filtered estimate."""
    lib = load()
    v, m = (np.ascontiguousarray(x, np.float32) for x in (var_filtered, mean_filtered))
    out = np.zeros(v.shape[:2], np.float32)
    lib.oc_flt_importance(v.ctypes.data, m.ctypes.data, v.shape[1], v.shape[0], out.ctypes.data)
    return out


def flt_accumulate(state, pos, wt, rgb, sumimp=None, newsum=None):
    """One film's running sums (dict of float32 arrays film, mean, m2 (H, W, 3)
    and sw (H, W)), rescaled to a new pass's importance if sumimp/newsum are
    given, then the deposits (pixel index, weight, rgb) added in order; the
    state is updated in place."""
    lib = load()
    h, w = state["sw"].shape
    pos = np.ascontiguousarray(pos, np.int32)
    wt = np.ascontiguousarray(wt, np.float32)
    rgb = np.ascontiguousarray(rgb, np.float32)
    si = np.ascontiguousarray(sumimp, np.float32) if sumimp is not None else None
    ns = np.ascontiguousarray(newsum, np.float32) if newsum is not None else None
    lib.oc_flt_accumulate(w, h, state["film"].ctypes.data, state["mean"].ctypes.data, state["m2"].ctypes.data,
                          state["sw"].ctypes.data, si.ctypes.data if si is not None else None,
                          ns.ctypes.data if ns is not None else None, len(pos), pos.ctypes.data, wt.ctypes.data,
                          rgb.ctypes.data)
