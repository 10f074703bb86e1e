"""Host-side scene model.

Mirrors the reference's scene classes (Camera, DiffuseMaterial,
MirrorMaterial, PhongMaterial, BlinnMaterial, TransparentMaterial, TextureMaterial,
DarkMatter, AreaLight, PointLight, ConeLight, EnvironmentMap, Triangle, Quad,
Sphere, Plane, Box, CSGIntersection, CSGUnion, CSGDifference, CSGXor, IntersectableList,
Scene) with the
same constructor arguments, and flattens a scene into the pointer-free
arrays of ``rt_scene_desc`` (include/rt_scene.h).

Primitive IDs follow the reference's flattened root list
(src/sceneast.cpp:676-678, src/intersectablelist.cpp:67-78): lights first in
file order, then the components of the intersectable tree depth-first.  A
Quad is one primitive made of the triangles (p1,p2,p4), (p4,p2,p3)
(src/quad.h:32); every triangle of a Mesh is its own primitive, as the
triangles of an OBJ file are.
"""
import ctypes

import numpy as np

from . import _abi


def _v3(x):
    a = np.asarray(x, dtype=np.float32).reshape(3)
    return a


# ---------------------------------------------------------------------------
# materials, src/material.h:30-50
# ---------------------------------------------------------------------------
class Material:
    kind = None


class DiffuseMaterial(Material):
    """src/diffusematerial.h: colour, brdf = colour / pi."""
    kind = _abi.MAT_DIFFUSE

    def __init__(self, color):
        self.color = _v3(color)


class MirrorMaterial(Material):
    """src/mirrormaterial.h: perfect specular reflector (the argument is unused there too)."""
    kind = _abi.MAT_MIRROR

    def __init__(self, k=1.0):
        self.k = float(k)


class PhongMaterial(Material):
    """src/phongmaterial.h: diffuse colour, specular colour, exponent."""
    kind = _abi.MAT_PHONG

    def __init__(self, color, specular_color, specular_coefficient):
        self.color = _v3(color)
        self.specular_color = _v3(specular_color)
        self.specular_coefficient = float(specular_coefficient)


class DarkMatter(Material):
    """src/darkmatter.h: absorbs everything."""
    kind = _abi.MAT_DARK


class BlinnMaterial(Material):
    """src/blinnmaterial.h: diffuse colour, specular colour, exponent."""
    kind = _abi.MAT_BLINN

    def __init__(self, color, specular_color, specular_coefficient):
        self.color = _v3(color)
        self.specular_color = _v3(specular_color)
        self.specular_coefficient = float(specular_coefficient)


class TransparentMaterial(Material):
    """src/transparentmaterial.h: refraction coefficient (the scene language's
    refractive(ior)); a specular reflector / refractor."""
    kind = _abi.MAT_REFRACTIVE

    def __init__(self, refraction_coefficient):
        self.refraction_coefficient = float(refraction_coefficient)


class TextureMaterial(Material):
    """src/texturematerial.h: a diffuse surface whose albedo is an image,
    pow(image(texcoord) * coefficient, gamma); image (H, W, 3) float RGB in
    [0, 1] (byte / 255, as the reference converts it), row 0 at the top.
    TextureMaterial.load reads a PNG, PPM or BMP file as the reference's
    imread does."""

    kind = _abi.MAT_TEXTURE

    def __init__(self, image, coefficient=(1.0, 1.0, 1.0), gamma=1.0):
        img = np.ascontiguousarray(image, dtype=np.float32)
        if img.ndim != 3 or img.shape[2] != 3:
            raise ValueError("texture image must be (H, W, 3)")
        self.image, self.coefficient, self.gamma = img, _v3(coefficient), float(gamma)

    @classmethod
    def load(cls, path, coefficient=(1.0, 1.0, 1.0), gamma=1.0):
        lib = _abi.load()
        w, h = ctypes.c_int32(), ctypes.c_int32()
        st = lib.rt_image_read(str(path).encode(), ctypes.byref(w), ctypes.byref(h), None, 0)
        if st != 0:
            raise _abi.RtError(st, lib.rt_film_last_error().decode(errors="replace"))
        img = np.zeros((h.value, w.value, 3), np.float32)
        st = lib.rt_image_read(str(path).encode(), ctypes.byref(w), ctypes.byref(h), img.ctypes.data, img.size)
        if st != 0:
            raise _abi.RtError(st, lib.rt_film_last_error().decode(errors="replace"))
        return cls(img, coefficient, gamma)


# ---------------------------------------------------------------------------
# lights, src/light.h:31-37
# ---------------------------------------------------------------------------
class Light:
    pass


class AreaLight(Light):
    """src/arealight.h: parallelogram origin + s*u + t*v, normal u x v, intensity."""

    def __init__(self, origin, u_direction, v_direction, intensity):
        self.origin, self.u, self.v, self.intensity = _v3(origin), _v3(u_direction), _v3(v_direction), _v3(intensity)


class PointLight(Light):
    """src/pointlight.h: location, intensity."""

    def __init__(self, location, intensity):
        self.location, self.intensity = _v3(location), _v3(intensity)


class ConeLight(Light):
    """src/conelight.h: location, axis direction, opening angle (radians), intensity."""

    def __init__(self, location, direction, opening_angle, intensity):
        self.location, self.direction = _v3(location), _v3(direction)
        self.opening_angle, self.intensity = float(opening_angle), _v3(intensity)


class EnvironmentMap(Light):
    """Environment light: an RGB float image (H, W, 3) times a coefficient
    (src/environmentmap.h).  mapping "latlong": latitude-longitude, row 0 at
    +y; "probe": the reference's angular light probe
    (src/environmentmap.cpp:57-66).  EnvironmentMap.load reads a probe from a
    Radiance .hdr file, as the reference does."""

    def __init__(self, image, coefficient=(1.0, 1.0, 1.0), mapping="latlong"):
        img = np.ascontiguousarray(image, dtype=np.float32)
        if img.ndim != 3 or img.shape[2] != 3:
            raise ValueError("environment image must be (H, W, 3)")
        if mapping not in ("latlong", "probe"):
            raise ValueError("mapping must be 'latlong' or 'probe'")
        self.image, self.coefficient, self.mapping = img, _v3(coefficient), mapping

    @classmethod
    def load(cls, path, coefficient=(1.0, 1.0, 1.0)):
        lib = _abi.load()
        w, h = ctypes.c_int32(), ctypes.c_int32()
        st = lib.rt_image_read_hdr(str(path).encode(), ctypes.byref(w), ctypes.byref(h), None, 0)
        if st != 0:
            raise _abi.RtError(st, lib.rt_film_last_error().decode(errors="replace"))
        img = np.zeros((h.value, w.value, 3), np.float32)
        st = lib.rt_image_read_hdr(str(path).encode(), ctypes.byref(w), ctypes.byref(h), img.ctypes.data, img.size)
        if st != 0:
            raise _abi.RtError(st, lib.rt_film_last_error().decode(errors="replace"))
        return cls(img, coefficient, mapping="probe")


# ---------------------------------------------------------------------------
# intersectables
# ---------------------------------------------------------------------------
class Intersectable:
    pass


class Triangle(Intersectable):
    """src/triangle.h: three points, optional vertex normals, material."""

    def __init__(self, p1, p2, p3, material, normals=None, texcoords=None):
        self.p = np.stack([_v3(p1), _v3(p2), _v3(p3)])
        self.normals = None if normals is None else np.asarray(normals, np.float32).reshape(3, 3)
        # (u, v) at p1, p2, p3; None = the reference's default (0, 0), (0, 1), (1, 0)
        self.texcoords = None if texcoords is None else np.asarray(texcoords, np.float32).reshape(3, 2)
        self.material = material


class Quad(Intersectable):
    """src/quad.h: p1..p4, split into (p1,p2,p4) and (p4,p2,p3)."""

    def __init__(self, p1, p2, p3, p4, material):
        self.p = np.stack([_v3(p1), _v3(p2), _v3(p3), _v3(p4)])
        self.material = material


class Sphere(Intersectable):
    """src/sphere.h: centre, radius, material."""

    def __init__(self, center, radius, material):
        self.center, self.radius, self.material = _v3(center), float(radius), material


class Plane(Intersectable):
    """src/plane.h: the plane a x + b y + c z + d = 0 given as [a b c d], with
    normal (a, b, c)."""

    def __init__(self, plane, material):
        self.plane = np.asarray(plane, np.float32).reshape(4)
        self.material = material


class Box(Intersectable):
    """src/axisalignedbox.h as an intersectable: min and max corners, material."""

    def __init__(self, min_corner, max_corner, material):
        self.min, self.max, self.material = _v3(min_corner), _v3(max_corner), material


class CSGOperation(Intersectable):
    """src/csgoperation.h: a boolean combination of two CSG objects (spheres,
    boxes or operations)."""

    op = None

    def __init__(self, left, right):
        for x in (left, right):
            if not isinstance(x, (Sphere, Box, CSGOperation)):
                raise TypeError("CSG operands are spheres, boxes and CSG operations")
        self.left, self.right = left, right


class CSGIntersection(CSGOperation):
    op = _abi.CSG_ISECT


class CSGUnion(CSGOperation):
    op = _abi.CSG_UNION


class CSGDifference(CSGOperation):
    op = _abi.CSG_DIFF


class CSGXor(CSGOperation):
    op = _abi.CSG_XOR


class Mesh(Intersectable):
    """Indexed triangle mesh (what the reference's obj(...) produces): one
    primitive per triangle, one material for all or one per face."""

    def __init__(self, positions, faces, material, normals=None, face_materials=None, texcoords=None):
        self.texcoords = None if texcoords is None else np.ascontiguousarray(texcoords, dtype=np.float32)  # per vertex
        self.positions = np.ascontiguousarray(positions, dtype=np.float32)
        self.faces = np.ascontiguousarray(faces, dtype=np.int32)
        self.normals = None if normals is None else np.ascontiguousarray(normals, dtype=np.float32)
        self.material = material
        self.face_materials = face_materials  # (list of materials, per-face index array) or None


class IntersectableList(Intersectable):
    """src/intersectablelist.h: ordered list of intersectables."""

    def __init__(self, items):
        self.items = list(items)


class Camera:
    """src/camera.h: eye, look-at, up, vertical fov in radians, resolution."""

    def __init__(self, cop, look_at, up, fov, width, height):
        self.cop, self.look_at, self.up = _v3(cop), _v3(look_at), _v3(up)
        self.fov, self.width, self.height = float(fov), int(width), int(height)

    def to_c(self):
        c = _abi.RtCamera()
        c.eye[:] = self.cop.tolist()
        c.look_at[:] = self.look_at.tolist()
        c.up[:] = self.up.tolist()
        c.fov, c.width, c.height = self.fov, self.width, self.height
        return c


class SceneArrays:
    """Flattened scene: numpy arrays + ctypes descriptor (keeps arrays alive)."""

    def __init__(self, positions, normals, tri_vertices, tri_material, tri_prim, spheres, sphere_material,
                 sphere_prim, materials, lights, env, env_kind=0, csg_nodes=(), object_first=None,
                 object_prim=None, textures=(), tri_texcoords=None):
        self.positions, self.normals = positions, normals
        self.tri_vertices, self.tri_material, self.tri_prim = tri_vertices, tri_material, tri_prim
        self.spheres, self.sphere_material, self.sphere_prim = spheres, sphere_material, sphere_prim
        self.materials, self.lights, self.env = materials, lights, env
        self.env_kind = env_kind
        self.csg_nodes = list(csg_nodes)
        self.object_first = np.zeros(1, np.int32) if object_first is None else np.asarray(object_first, np.int32)
        self.object_prim = np.zeros(0, np.int32) if object_prim is None else np.asarray(object_prim, np.int32)
        self.textures = list(textures)
        self.tri_texcoords = tri_texcoords
        self.bvh_build = 0  # BVH split rule: 0 median (the reference's), 1 surface-area heuristic
        self._desc = None

    @property
    def n_prims(self):
        ids = [len(self.lights) - 1]
        if len(self.tri_prim):
            ids.append(int(self.tri_prim.max()))
        if len(self.sphere_prim):
            ids.append(int(self.sphere_prim.max()))
        return max(ids) + 1

    def desc(self):
        if self._desc is not None:
            self._desc.bvh_build = self.bvh_build
            return self._desc
        d = _abi.RtSceneDesc()

        def ptr(a):
            return a.ctypes.data if a is not None and a.size else None

        d.n_vertices = len(self.positions)
        d.positions = ptr(self.positions)
        d.normals = ptr(self.normals)
        d.n_triangles = len(self.tri_vertices)
        d.tri_vertices, d.tri_material, d.tri_prim_id = ptr(self.tri_vertices), ptr(self.tri_material), ptr(self.tri_prim)
        d.n_spheres = len(self.spheres)
        d.spheres, d.sphere_material, d.sphere_prim_id = ptr(self.spheres), ptr(self.sphere_material), ptr(self.sphere_prim)
        self._mats_c = (_abi.RtMaterial * max(1, len(self.materials)))(*self.materials)
        self._lights_c = (_abi.RtLight * max(1, len(self.lights)))(*self.lights)
        d.n_materials = len(self.materials)
        d.materials = ctypes.addressof(self._mats_c) if self.materials else None
        d.n_lights = len(self.lights)
        d.lights = ctypes.addressof(self._lights_c) if self.lights else None
        if self.env is not None:
            d.env_height, d.env_width = self.env.shape[0], self.env.shape[1]
            d.env_rgb = self.env.ctypes.data
            d.env_kind = self.env_kind
        if len(self.object_prim):
            self._nodes_c = (_abi.RtCsgNode * len(self.csg_nodes))(*self.csg_nodes)
            d.n_objects = len(self.object_prim)
            d.csg_nodes = ctypes.addressof(self._nodes_c)
            d.object_first = self.object_first.ctypes.data
            d.object_prim_id = self.object_prim.ctypes.data
        if self.textures:
            self._tex_c = (_abi.RtTexture * len(self.textures))()
            for i, img in enumerate(self.textures):
                self._tex_c[i].height, self._tex_c[i].width = img.shape[0], img.shape[1]
                self._tex_c[i].rgb = img.ctypes.data
            d.n_textures = len(self.textures)
            d.textures = ctypes.addressof(self._tex_c)
        if self.tri_texcoords is not None:
            d.tri_texcoords = self.tri_texcoords.ctypes.data
        d.bvh_build = self.bvh_build
        self._desc = d
        return d


class Scene:
    """src/scene.h: camera, light list, and the root intersectable."""

    def __init__(self, camera, lights, intersectable):
        self.camera = camera
        self.light = list(lights)
        self.object = intersectable

    def flatten(self):
        mats, mat_index, textures = [], {}, []

        def mid(m):
            if id(m) not in mat_index:
                c = _abi.RtMaterial()
                c.type = m.kind
                if isinstance(m, DiffuseMaterial):
                    c.kd[:] = m.color.tolist()
                elif isinstance(m, (PhongMaterial, BlinnMaterial)):
                    c.kd[:] = m.color.tolist()
                    c.ks[:] = m.specular_color.tolist()
                    c.exponent = m.specular_coefficient
                elif isinstance(m, TransparentMaterial):
                    c.exponent = m.refraction_coefficient
                elif isinstance(m, TextureMaterial):
                    c.kd[:] = m.coefficient.tolist()
                    c.exponent = m.gamma
                    c.texture = len(textures)
                    textures.append(m.image)
                mat_index[id(m)] = len(mats)
                mats.append(c)
            return mat_index[id(m)]

        lights, env, env_kind = [], None, 0
        for L in self.light:
            c = _abi.RtLight()
            if isinstance(L, AreaLight):
                c.type = _abi.LIGHT_AREA
                c.p[:], c.u[:], c.v[:] = L.origin.tolist(), L.u.tolist(), L.v.tolist()
            elif isinstance(L, PointLight):
                c.type = _abi.LIGHT_POINT
                c.p[:] = L.location.tolist()
            elif isinstance(L, ConeLight):
                c.type = _abi.LIGHT_CONE
                c.p[:], c.u[:] = L.location.tolist(), L.direction.tolist()
                c.v[0] = L.opening_angle
            elif isinstance(L, EnvironmentMap):
                if env is not None:
                    raise ValueError("at most one environment map")
                c.type = _abi.LIGHT_ENV
                env = L.image
                env_kind = 1 if L.mapping == "probe" else 0
            else:
                raise TypeError(type(L))
            c.intensity[:] = (L.coefficient if isinstance(L, EnvironmentMap) else L.intensity).tolist()
            lights.append(c)

        pos, nrm, tv, tm, tp, tcs = [], [], [], [], [], []
        sph, sm, sp = [], [], []
        nodes, ofirst, oprim = [], [0], []

        def csg_program(o):
            """postfix nodes of a sphere / box / operation tree"""
            if isinstance(o, CSGOperation):
                csg_program(o.left)
                csg_program(o.right)
                n = _abi.RtCsgNode()
                n.op = o.op
                nodes.append(n)
                return
            n = _abi.RtCsgNode()
            n.material = mid(o.material)
            if isinstance(o, Sphere):
                n.op = _abi.CSG_SPHERE
                n.p[:4] = [*o.center.tolist(), o.radius]
            else:
                n.op = _abi.CSG_BOX
                n.p[:] = [*o.min.tolist(), *o.max.tolist()]
            nodes.append(n)
        nv = [0]
        has_normals = [False]
        next_id = [len(lights)]

        def add_tris(P, F, N, mats_per_face, ids, tc=None):
            tcs.append(tc)
            pos.append(P)
            nrm.append(N)
            if N is not None:
                has_normals[0] = True
            tv.append(F + nv[0])
            tm.append(mats_per_face)
            tp.append(ids)
            nv[0] += len(P)

        def walk(o):
            if isinstance(o, IntersectableList):
                for it in o.items:
                    walk(it)
            elif isinstance(o, Triangle):
                add_tris(o.p, np.array([[0, 1, 2]], np.int32), o.normals, np.array([mid(o.material)], np.int32),
                         np.array([next_id[0]], np.int32), None if o.texcoords is None else o.texcoords.reshape(1, 6))
                next_id[0] += 1
            elif isinstance(o, Quad):
                add_tris(o.p, np.array([[0, 1, 3], [3, 1, 2]], np.int32), None,
                         np.full(2, mid(o.material), np.int32), np.full(2, next_id[0], np.int32))
                next_id[0] += 1
            elif isinstance(o, (Box, CSGOperation)):
                csg_program(o)
                ofirst.append(len(nodes))
                oprim.append(next_id[0])
                next_id[0] += 1
            elif isinstance(o, Plane):
                n = _abi.RtCsgNode()
                n.op, n.material = _abi.CSG_PLANE, mid(o.material)
                n.p[:4] = o.plane.tolist()
                nodes.append(n)
                ofirst.append(len(nodes))
                oprim.append(next_id[0])
                next_id[0] += 1
            elif isinstance(o, Sphere):
                sph.append(np.array([*o.center.tolist(), o.radius], np.float32))
                sm.append(mid(o.material))
                sp.append(next_id[0])
                next_id[0] += 1
            elif isinstance(o, Mesh):
                n = len(o.faces)
                if o.face_materials is not None:
                    mlist, fidx = o.face_materials
                    lut = np.array([mid(m) for m in mlist], np.int32)
                    fm = lut[np.asarray(fidx)]
                else:
                    fm = np.full(n, mid(o.material), np.int32)
                add_tris(o.positions, o.faces, o.normals, fm.astype(np.int32),
                         np.arange(next_id[0], next_id[0] + n, dtype=np.int32),
                         None if o.texcoords is None else o.texcoords[o.faces].reshape(n, 6))
                next_id[0] += n
            else:
                raise TypeError(type(o))

        if self.object is not None:
            walk(self.object)
        if has_normals[0]:
            nrm = [n if n is not None else _flat_vertex_normals(P, F - off)
                   for n, P, F, off in zip(nrm, pos, tv, np.cumsum([0] + [len(p) for p in pos])[:-1])]
        positions = np.ascontiguousarray(np.concatenate(pos) if pos else np.zeros((0, 3), np.float32), np.float32)
        normals = np.ascontiguousarray(np.concatenate(nrm), np.float32) if has_normals[0] else None
        tri_vertices = np.ascontiguousarray(np.concatenate(tv) if tv else np.zeros((0, 3), np.int32), np.int32)
        tri_material = np.ascontiguousarray(np.concatenate(tm) if tm else np.zeros(0, np.int32), np.int32)
        tri_prim = np.ascontiguousarray(np.concatenate(tp) if tp else np.zeros(0, np.int32), np.int32)
        spheres = np.ascontiguousarray(np.stack(sph) if sph else np.zeros((0, 4), np.float32), np.float32)
        tri_tc = None
        if any(t is not None for t in tcs):
            default = np.array([0, 0, 0, 1, 1, 0], np.float32)
            tri_tc = np.ascontiguousarray(np.concatenate(
                [t if t is not None else np.tile(default, (len(F), 1)) for t, F in zip(tcs, tv)]), np.float32)
        return SceneArrays(positions, normals, tri_vertices, tri_material, tri_prim, spheres,
                           np.asarray(sm, np.int32), np.asarray(sp, np.int32), mats, lights, env, env_kind,
                           nodes, np.asarray(ofirst, np.int32), np.asarray(oprim, np.int32), textures, tri_tc)


def _flat_vertex_normals(P, F):
    """Per-vertex copies of flat normals for a triangle set without normals,
    used only when another part of the scene has vertex normals."""
    raise ValueError("mixing meshes with and without vertex normals is not supported")
