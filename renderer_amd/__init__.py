"""MI355X path-tracing hot path: HIP kernels for gfx950 behind a C ABI
(librt_hip.so, include/rt_abi.h), with a host-side scene model and
integrator interface named after the reference renderer's classes."""
from . import _abi
from ._abi import (LIGHT_AREA, LIGHT_CONE, LIGHT_ENV, LIGHT_POINT, MAT_BLINN, MAT_DARK, MAT_DIFFUSE, MAT_MIRROR,
                   MAT_PHONG, MAT_REFRACTIVE, MLT_ERPT, MLT_FILTERED, MLT_METROPOLIS,
                   FLT_NO_CROSSFILTERING, FLT_NO_FILTER, FLT_NO_VARIANCE_FILTER, FLT_SIMPLE_VARIANCE_FILTER,
                   MODE_COMPAT, MODE_GUIDED, MODE_UNBIASED, BVH_MEDIAN, BVH_SAH, RtError)
from .integrator import (EnergyRedistributionRenderer, GpuContext, GpuUniDiPathTracingIntegrator, MetropolisFltRenderer,
                         MetropolisRenderer, PerPixelRenderer)
from .progressive import GuidedProgressive
from .scene import (AreaLight, BlinnMaterial, Camera, ConeLight, DarkMatter, DiffuseMaterial, EnvironmentMap,
                    IntersectableList, Mesh, MirrorMaterial, PhongMaterial, PointLight, Quad, Scene, Sphere,
                    TextureMaterial, TransparentMaterial, Triangle)
from .scene_file import SceneFile

__all__ = [
    "GpuContext", "GpuUniDiPathTracingIntegrator", "PerPixelRenderer", "RtError",
    "AreaLight", "Camera", "DarkMatter", "DiffuseMaterial", "EnvironmentMap", "IntersectableList", "Mesh",
    "MirrorMaterial", "PhongMaterial", "PointLight", "Quad", "Scene", "Sphere", "Triangle", "SceneFile",
    "BlinnMaterial", "TransparentMaterial", "TextureMaterial", "ConeLight", "MetropolisRenderer", "EnergyRedistributionRenderer",
    "MetropolisFltRenderer", "MLT_FILTERED", "FLT_NO_FILTER", "FLT_NO_CROSSFILTERING", "FLT_NO_VARIANCE_FILTER",
    "FLT_SIMPLE_VARIANCE_FILTER",
    "MODE_COMPAT", "MODE_UNBIASED", "MODE_GUIDED", "MLT_METROPOLIS", "MLT_ERPT", "BVH_MEDIAN", "BVH_SAH",
    "GuidedProgressive",
]
