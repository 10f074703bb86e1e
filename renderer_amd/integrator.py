"""Host interface over librt_hip.so, named after the reference's classes.

GpuContext                      one rt_ctx (device, stream, uploaded scene)
GpuUniDiPathTracingIntegrator   the drop-in Integrator: same construction
                                argument as UniDiPathTracingIntegrator
                                (terminationProb, src/unidipathtracingintegrator.h:30)
PerPixelRenderer                render(scene, spp_x, spp_y, seed) -> film, the
                                batch form of src/perpixelrenderer.cpp:43-101

All work runs in the HIP library; there is no CPU fallback.  Errors raise
RtError carrying the library's status code.
"""
import ctypes

import numpy as np

from . import _abi
from .scene import Camera, Scene


def _check(lib, ctx, status):
    if status != 0:
        msg = lib.rt_last_error(ctx).decode() if ctx else ""
        raise _abi.RtError(status, msg)


def _f32(a, shape_last=4):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if a.ndim != 2 or a.shape[1] != shape_last:
        raise ValueError(f"expected (n, {shape_last}) float32 array")
    return a


def bvh_export(scene):
    """The 4-wide BVH that GpuContext.upload builds for a scene, built on the
    host with no device: (nodes float32 (n, 32), records float32 (m, 12),
    stack_need).  Layout as rt_bvh_export in rt_abi.h states."""
    lib = _abi.load()
    arrays = scene.flatten() if isinstance(scene, Scene) else scene
    desc = arrays.desc()
    nn, nr, need = ctypes.c_int64(0), ctypes.c_int64(0), ctypes.c_int32(0)
    _check(lib, None, lib.rt_bvh_export(ctypes.byref(desc), None, ctypes.byref(nn), None, ctypes.byref(nr),
                                        ctypes.byref(need)))
    nodes = np.zeros((max(nn.value, 1), 32), np.float32)
    recs = np.zeros((max(nr.value, 1), 12), np.float32)
    _check(lib, None, lib.rt_bvh_export(ctypes.byref(desc), nodes.ctypes.data, ctypes.byref(nn), recs.ctypes.data,
                                        ctypes.byref(nr), ctypes.byref(need)))
    return nodes[:nn.value], recs[:nr.value], need.value


class GpuContext:
    def __init__(self, device=0):
        self.lib = _abi.load()
        h = ctypes.c_void_p()
        st = self.lib.rt_create(int(device), ctypes.byref(h))
        if st != 0:
            raise _abi.RtError(st, "rt_create")
        self.h = h
        self.scene_arrays = None

    def close(self):
        if self.h:
            self.lib.rt_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- scene ---------------------------------------------------------------
    def upload(self, scene):
        arrays = scene.flatten() if isinstance(scene, Scene) else scene
        _check(self.lib, self.h, self.lib.rt_scene_upload(self.h, ctypes.byref(arrays.desc())))
        self.scene_arrays = arrays
        return arrays

    # -- ray batches (host arrays) ---------------------------------------------
    def trace_closest(self, o_tmin, d_tmax):
        o, d = _f32(o_tmin), _f32(d_tmax)
        n = len(o)
        prim = np.empty(n, np.int32)
        t = np.empty(n, np.float32)
        _check(self.lib, self.h, self.lib.rt_trace_closest(self.h, o.ctypes.data, d.ctypes.data, n,
                                                           prim.ctypes.data, t.ctypes.data))
        return prim, t

    def trace_any(self, o_tmin, d_tmax):
        o, d = _f32(o_tmin), _f32(d_tmax)
        n = len(o)
        occ = np.empty(n, np.uint8)
        _check(self.lib, self.h, self.lib.rt_trace_any(self.h, o.ctypes.data, d.ctypes.data, n, occ.ctypes.data))
        return occ

    # -- ray batches (device pointers, e.g. torch tensors) ---------------------
    def trace_closest_device(self, o_ptr, d_ptr, n, prim_ptr, t_ptr):
        _check(self.lib, self.h, self.lib.rt_trace_closest_device(self.h, o_ptr, d_ptr, int(n), prim_ptr, t_ptr))

    def trace_any_device(self, o_ptr, d_ptr, n, occ_ptr):
        _check(self.lib, self.h, self.lib.rt_trace_any_device(self.h, o_ptr, d_ptr, int(n), occ_ptr))

    def trace_count_device(self, o_ptr, d_ptr, n, nodes_ptr, prims_ptr):
        _check(self.lib, self.h, self.lib.rt_trace_count_device(self.h, o_ptr, d_ptr, int(n), nodes_ptr, prims_ptr))

    def trace_any_count_device(self, o_ptr, d_ptr, n, nodes_ptr, prims_ptr):
        _check(self.lib, self.h,
               self.lib.rt_trace_any_count_device(self.h, o_ptr, d_ptr, int(n), nodes_ptr, prims_ptr))

    # -- render ------------------------------------------------------------------
    @staticmethod
    def make_views(views):
        """views: list of (Camera, spp_x, spp_y)."""
        arr = (_abi.RtView * len(views))()
        for i, (cam, sx, sy) in enumerate(views):
            arr[i].camera = cam.to_c()
            arr[i].spp_x, arr[i].spp_y = int(sx), int(sy)
        return arr

    @staticmethod
    def make_params(mode=_abi.MODE_UNBIASED, max_depth=8, rr_start=3, rr_prob=0.5, seed=0, guide_train=0,
                    guide_lr=1e-3, guide_batch=1 << 18):
        p = _abi.RtParams()
        p.mode, p.max_depth, p.rr_start, p.rr_prob, p.seed = int(mode), int(max_depth), int(rr_start), float(rr_prob), int(seed)
        p.guide_train, p.guide_lr, p.guide_batch = int(guide_train), float(guide_lr), int(guide_batch)
        return p

    def render(self, views, films=None, **params):
        """Render views (list of (Camera, spp_x, spp_y)) into host films
        (height, width, 3) float32, returned as a list."""
        va = self.make_views(views)
        if films is None:
            films = [np.zeros((cam.height, cam.width, 3), np.float32) for cam, _, _ in views]
        ptrs = (ctypes.c_void_p * len(films))(*[f.ctypes.data for f in films])
        p = self.make_params(**params)
        _check(self.lib, self.h, self.lib.rt_render(self.h, va, len(views), ctypes.byref(p), ptrs))
        return films

    def render_device(self, views, film_ptrs, **params):
        va = self.make_views(views)
        ptrs = (ctypes.c_void_p * len(film_ptrs))(*film_ptrs)
        p = self.make_params(**params)
        _check(self.lib, self.h, self.lib.rt_render_device(self.h, va, len(views), ctypes.byref(p), ptrs))

    def integrate_rays(self, o_tmin, d_tmax, key0=0, **params):
        """Integrator::integrate(const Ray&, ...) for each ray (n x 4 origin +
        range start, n x 4 direction + range end): (n, 3) RGB."""
        o, d = _f32(o_tmin), _f32(d_tmax)
        n = len(o)
        out = np.empty((n, 3), np.float32)
        p = self.make_params(**params)
        _check(self.lib, self.h, self.lib.rt_integrate_rays(self.h, o.ctypes.data, d.ctypes.data, n, ctypes.byref(p),
                                                            int(key0) & 0xffffffff, out.ctypes.data))
        return out

    def integrate_rays_device(self, o_ptr, d_ptr, n, rgb_ptr, key0=0, **params):
        p = self.make_params(**params)
        _check(self.lib, self.h, self.lib.rt_integrate_rays_device(self.h, o_ptr, d_ptr, int(n), ctypes.byref(p),
                                                                   int(key0) & 0xffffffff, rgb_ptr))

    # -- Markov-chain renderers (include/rt_mlt.h) --------------------------------
    @staticmethod
    def make_mlt_params(kind=_abi.MLT_METROPOLIS, max_depth=5, rr_start=3, termination_prob=0.5,
                        large_step_prob=0.1, bootstrap=10000, mutations=16, chains=0, erpt_spp=(2, 2),
                        erpt_ed_termination=0.1, seed=0, flt_passes=3, flt_flags=0):
        p = _abi.RtMltParams()
        p.kind, p.max_depth, p.rr_start = int(kind), int(max_depth), int(rr_start)
        p.termination_prob, p.large_step_prob = float(termination_prob), float(large_step_prob)
        p.bootstrap, p.mutations, p.chains = int(bootstrap), int(mutations), int(chains)
        p.erpt_spp_x, p.erpt_spp_y = int(erpt_spp[0]), int(erpt_spp[1])
        p.erpt_ed_termination, p.seed = float(erpt_ed_termination), int(seed)
        p.flt_passes, p.flt_flags = int(flt_passes), int(flt_flags)
        return p

    @staticmethod
    def _mlt_result(r):
        return {k: getattr(r, k) for k, _ in _abi.RtMltResult._fields_}

    def render_mlt(self, camera, film=None, **params):
        """Metropolis or ERPT render (make_mlt_params keywords) added into a
        host film (height, width, 3); returns (film, result dict)."""
        if film is None:
            film = np.zeros((camera.height, camera.width, 3), np.float32)
        assert film.dtype == np.float32 and film.flags.c_contiguous and film.shape == (camera.height, camera.width, 3)
        cam = camera.to_c()
        p, res = self.make_mlt_params(**params), _abi.RtMltResult()
        _check(self.lib, self.h, self.lib.rt_render_mlt(self.h, ctypes.byref(cam), ctypes.byref(p), film.ctypes.data,
                                                        ctypes.byref(res)))
        return film, self._mlt_result(res)

    def render_mlt_device(self, camera, film_ptr, **params):
        cam = camera.to_c()
        p, res = self.make_mlt_params(**params), _abi.RtMltResult()
        _check(self.lib, self.h, self.lib.rt_render_mlt_device(self.h, ctypes.byref(cam), ctypes.byref(p),
                                                               ctypes.c_void_p(film_ptr), ctypes.byref(res)))
        return self._mlt_result(res)

    def symmetric_filter(self, image, guide, pixvar, patch=7, radius=10):
        """SymmetricFilter::filter on (height, width, 3) float32 arrays."""
        a, g, v = (np.ascontiguousarray(x, np.float32) for x in (image, guide, pixvar))
        assert a.ndim == 3 and a.shape[2] == 3 and g.shape == a.shape and v.shape == a.shape
        out = np.empty_like(a)
        _check(self.lib, self.h, self.lib.rt_symmetric_filter(self.h, a.ctypes.data, g.ctypes.data, v.ctypes.data,
                                                              a.shape[1], a.shape[0], int(patch), int(radius),
                                                              out.ctypes.data))
        return out

    def integrate_samples(self, camera, samples, max_depth=5, rr_start=3, termination_prob=0.0):
        """Radiance of paths given by primary samples (n, 2*max_depth+1, 4) float32."""
        x = np.ascontiguousarray(samples, np.float32)
        assert x.ndim == 3 and x.shape[1:] == (2 * max_depth + 1, 4)
        out = np.zeros((len(x), 3), np.float32)
        cam = camera.to_c()
        _check(self.lib, self.h, self.lib.rt_integrate_samples(self.h, ctypes.byref(cam), int(max_depth), int(rr_start),
                                                               float(termination_prob), x.ctypes.data, len(x),
                                                               out.ctypes.data))
        return out

    def set_launch_config(self, path_slots=0, block=0, blocks_per_cu=0, use_graph=0, count_traversal=0,
                          guide_epochs=0, finish_depth=0, finish_paths=0):
        c = _abi.RtLaunchConfig(int(path_slots), int(block), int(blocks_per_cu), int(use_graph), int(count_traversal),
                                int(guide_epochs), int(finish_depth), int(finish_paths))
        _check(self.lib, self.h, self.lib.rt_set_launch_config(self.h, ctypes.byref(c)))

    def launch_config(self):
        c = _abi.RtLaunchConfig()
        _check(self.lib, self.h, self.lib.rt_get_launch_config(self.h, ctypes.byref(c)))
        return {f: getattr(c, f) for f, _ in c._fields_}

    # -- film outputs (include/rt_film.h) ------------------------------------------
    def tonemap(self, film, gamma=2.2):
        """Tonemapper::tonemap of a (height, width, 3) float film on the GPU:
        uint8 RGB of the same shape."""
        f = np.ascontiguousarray(film, np.float32)
        if f.ndim != 3 or f.shape[2] != 3:
            raise ValueError("film must be (height, width, 3)")
        out = np.empty(f.shape, np.uint8)
        _check(self.lib, self.h, self.lib.rt_film_tonemap(self.h, f.ctypes.data, f.shape[1], f.shape[0],
                                                          float(gamma), out.ctypes.data))
        return out

    def tonemap_device(self, film_ptr, width, height, out_ptr, gamma=2.2):
        """Device pointers: height*width*3 floats in, as many bytes out."""
        _check(self.lib, self.h, self.lib.rt_film_tonemap_device(self.h, film_ptr, int(width), int(height),
                                                                 float(gamma), out_ptr))

    def stats(self):
        s = _abi.RtStats()
        _check(self.lib, self.h, self.lib.rt_get_stats(self.h, ctypes.byref(s)))
        return {f: getattr(s, f) for f, _ in s._fields_}

    def request_stop(self, flag=True):
        _check(self.lib, self.h, self.lib.rt_request_stop(self.h, 1 if flag else 0))

    def set_fault(self, flag=True):
        _check(self.lib, self.h, self.lib.rt_set_fault(self.h, 1 if flag else 0))

    # -- guiding network -----------------------------------------------------------
    def guide_init(self, seed=0):
        _check(self.lib, self.h, self.lib.rt_guide_init(self.h, int(seed)))

    def guide_param_count(self):
        return int(self.lib.rt_guide_param_count())

    def guide_params(self):
        th = np.empty(self.guide_param_count(), np.float32)
        _check(self.lib, self.h, self.lib.rt_guide_get_params(self.h, th.ctypes.data, len(th)))
        return th

    def guide_set_params(self, theta):
        th = np.ascontiguousarray(theta, np.float32)
        _check(self.lib, self.h, self.lib.rt_guide_set_params(self.h, th.ctypes.data, len(th)))

    def guide_set_bounds(self, lo, hi):
        lo = np.ascontiguousarray(lo, np.float32)
        hi = np.ascontiguousarray(hi, np.float32)
        _check(self.lib, self.h, self.lib.rt_guide_set_bounds(self.h, lo.ctypes.data, hi.ctypes.data))

    def guide_forward(self, records):
        rec = _f32(records, 16)
        y = np.empty((len(rec), 48), np.float32)
        _check(self.lib, self.h, self.lib.rt_guide_forward(self.h, rec.ctypes.data, len(rec), y.ctypes.data))
        return y

    def guide_head(self, records):
        """Head outputs (n, 48) as the guided shader receives them."""
        rec = _f32(records, 16)
        out = np.empty((len(rec), 48), np.float32)
        _check(self.lib, self.h, self.lib.rt_guide_head(self.h, rec.ctypes.data, len(rec), out.ctypes.data))
        return out

    def guide_gradient(self, records):
        rec = _f32(records, 16)
        g = np.empty(self.guide_param_count(), np.float32)
        loss = np.zeros(1, np.float32)
        _check(self.lib, self.h, self.lib.rt_guide_gradient(self.h, rec.ctypes.data, len(rec), g.ctypes.data,
                                                            loss.ctypes.data))
        return float(loss[0]), g

    def guide_train(self, records, steps=1, lr=1e-3):
        rec = _f32(records, 16)
        _check(self.lib, self.h, self.lib.rt_guide_train(self.h, rec.ctypes.data, len(rec), int(steps), float(lr)))


class GpuUniDiPathTracingIntegrator:
    """The GPU drop-in for UniDiPathTracingIntegrator(terminationProb)."""

    def __init__(self, termination_prob=0.5, mode=_abi.MODE_COMPAT, max_depth=5, rr_start=3, context=None):
        self.termination_prob = float(termination_prob)
        self.mode, self.max_depth, self.rr_start = int(mode), int(max_depth), int(rr_start)
        self.ctx = context or GpuContext()
        self._scene = None

    def prepare(self, scene):
        if scene is not self._scene:
            self.ctx.upload(scene)
            self._scene = scene

    def render_views(self, scene, views, seed=0):
        self.prepare(scene)
        return self.ctx.render(views, mode=self.mode, max_depth=self.max_depth, rr_start=self.rr_start,
                               rr_prob=self.termination_prob, seed=seed)


class PerPixelRenderer:
    """Batch form of PerPixelRenderer::render: one call per film, pt-x/y-samples
    strata per pixel, --seed."""

    name = "pathtracing"

    def __init__(self, integrator, x_samples=4, y_samples=4, seed=0):
        self.integrator = integrator
        self.x_samples, self.y_samples, self.seed = int(x_samples), int(y_samples), int(seed)

    def render(self, scene):
        cam = scene.camera
        return self.integrator.render_views(scene, [(cam, self.x_samples, self.y_samples)], seed=self.seed)[0]


class MetropolisRenderer:
    """MetropolisRenderer (src/metropolisrenderer.cpp:56-212) on the GPU: the
    met-* options of the reference's command line, chains run side by side."""

    name = "metropolis"

    def __init__(self, large_step_prob=0.1, termination_prob=0.5, bootstrap=10000, mutations=16, chains=0,
                 seed=0, max_depth=5, context=None):
        self.params = dict(kind=_abi.MLT_METROPOLIS, large_step_prob=large_step_prob,
                           termination_prob=termination_prob, bootstrap=bootstrap, mutations=mutations,
                           chains=chains, seed=seed, max_depth=max_depth)
        self.ctx = context or GpuContext()
        self._scene = None
        self.result = None

    def render(self, scene):
        if scene is not self._scene:
            self.ctx.upload(scene)
            self._scene = scene
        film, self.result = self.ctx.render_mlt(scene.camera, **self.params)
        return film


class MetropolisFltRenderer(MetropolisRenderer):
    """MetropolisFltRenderer (src/metropolisfltrenderer.cpp:60-320) on the
    GPU: metflt-num-passes passes of two Metropolis films, an importance map
    from the previous pass's relative variance, and the symmetric
    cross-filter (src/symmetricfilter.cpp:40-137); `flags` combines
    FLT_NO_FILTER, FLT_NO_CROSSFILTERING, FLT_NO_VARIANCE_FILTER and
    FLT_SIMPLE_VARIANCE_FILTER (the --metflt-* switches)."""

    name = "metropolisflt"

    def __init__(self, passes=3, flags=0, large_step_prob=0.1, termination_prob=0.5, bootstrap=10000, mutations=16,
                 chains=0, seed=0, max_depth=5, context=None):
        super().__init__(large_step_prob, termination_prob, bootstrap, mutations, chains, seed, max_depth, context)
        self.params.update(kind=_abi.MLT_FILTERED, flt_passes=passes, flt_flags=flags)


class EnergyRedistributionRenderer(MetropolisRenderer):
    """EnergyRedistributionRenderer (src/energyredistributionrenderer.cpp:38-147)
    on the GPU: erpt-mutations chains steps, erpt-x/y-samples initial samples
    per pixel."""

    name = "energyredist"

    def __init__(self, mutations=100, x_samples=2, y_samples=2, seed=0, max_depth=5, context=None):
        super().__init__(seed=seed, max_depth=max_depth, context=context)
        self.params = dict(kind=_abi.MLT_ERPT, mutations=mutations, erpt_spp=(x_samples, y_samples), seed=seed,
                           max_depth=max_depth)
