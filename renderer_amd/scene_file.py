"""Scene files in the reference's scene language, read by the library's
front end (include/rt_scene_file.h, renderer_amd/csrc/scene_file.cpp).

    sf = SceneFile.load("box.scene")      # SceneGrammar::parse + resolveVars
    arrays = sf.arrays()                   # buildScene -> rt_scene_desc + camera
    ctx.upload(arrays)
    sf.dump_bvh("box.bvh")                 # --dump-bvh (src/main.cpp:166-180)

The reference's flow is src/main.cpp:150-182.  Errors raise RtError with the
library's status (RT_ERR_PARSE, RT_ERR_IO, RT_ERR_UNSUPPORTED, ...).
"""
import ctypes

import numpy as np

from . import _abi
from .scene import Camera


def _err(lib, status):
    return _abi.RtError(status, lib.rt_scene_file_last_error().decode(errors="replace"))


class FileSceneArrays:
    """The evaluated scene: the rt_scene_desc the front end filled (its arrays
    belong to the SceneFile, which this object keeps alive) and numpy views of
    it, with the same desc() interface as scene.SceneArrays."""

    def __init__(self, owner, desc, camera):
        self._owner, self._desc = owner, desc
        d = desc

        def view(ptr, n, dtype, width=None):
            if not ptr or n == 0:
                shape = (0, width) if width else (0,)
                return np.zeros(shape, dtype)
            ct = ctypes.c_float if dtype == np.float32 else ctypes.c_int32
            count = n * (width or 1)
            a = np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(ct)), shape=(count,))
            return a.reshape((n, width)) if width else a

        self.positions = view(d.positions, d.n_vertices, np.float32, 3)
        self.normals = view(d.normals, d.n_vertices, np.float32, 3) if d.normals else None
        self.tri_vertices = view(d.tri_vertices, d.n_triangles, np.int32, 3)
        self.tri_material = view(d.tri_material, d.n_triangles, np.int32)
        self.tri_prim = view(d.tri_prim_id, d.n_triangles, np.int32)
        self.spheres = view(d.spheres, d.n_spheres, np.float32, 4)
        self.sphere_material = view(d.sphere_material, d.n_spheres, np.int32)
        self.sphere_prim = view(d.sphere_prim_id, d.n_spheres, np.int32)
        mats = ctypes.cast(d.materials, ctypes.POINTER(_abi.RtMaterial)) if d.materials else None
        self.materials = [mats[i] for i in range(d.n_materials)] if mats else []
        lights = ctypes.cast(d.lights, ctypes.POINTER(_abi.RtLight)) if d.lights else None
        self.lights = [lights[i] for i in range(d.n_lights)] if lights else []
        self.env = (np.ctypeslib.as_array(ctypes.cast(d.env_rgb, ctypes.POINTER(ctypes.c_float)),
                                          shape=(d.env_height * d.env_width * 3,)).reshape(d.env_height, d.env_width, 3)
                    if d.env_rgb else None)
        self.env_kind = d.env_kind
        nodes = ctypes.cast(d.csg_nodes, ctypes.POINTER(_abi.RtCsgNode)) if d.csg_nodes else None
        self.object_first = view(d.object_first, d.n_objects + 1, np.int32) if d.n_objects else np.zeros(1, np.int32)
        self.object_prim = view(d.object_prim_id, d.n_objects, np.int32)
        self.csg_nodes = [nodes[i] for i in range(int(self.object_first[-1]))] if nodes else []
        # world-to-object maps of instanced objects (n_objects x 3 x 4), or None
        self.object_xform = (view(d.object_xform, d.n_objects, np.float32, 12).reshape(-1, 3, 4)
                             if d.object_xform else None)
        self.camera = Camera(list(camera.eye), list(camera.look_at), list(camera.up), camera.fov, camera.width,
                             camera.height)

    @property
    def n_prims(self):
        ids = [len(self.lights) - 1]
        if len(self.object_prim):
            ids.append(int(self.object_prim.max()))
        if len(self.tri_prim):
            ids.append(int(self.tri_prim.max()))
        if len(self.sphere_prim):
            ids.append(int(self.sphere_prim.max()))
        return max(ids) + 1

    def desc(self):
        return self._desc


class SceneFile:
    """A parsed scene file (src/sceneparser.h SceneGrammar + resolveVars)."""

    def __init__(self, handle):
        self.lib = _abi.load()
        self.h = handle

    @classmethod
    def load(cls, path):
        lib = _abi.load()
        h = ctypes.c_void_p()
        st = lib.rt_scene_file_load(str(path).encode(), ctypes.byref(h))
        if st != 0:
            raise _err(lib, st)
        return cls(h)

    @classmethod
    def parse(cls, text, base_dir=None):
        lib = _abi.load()
        h = ctypes.c_void_p()
        st = lib.rt_scene_file_parse(text.encode(), None if base_dir is None else str(base_dir).encode(),
                                     ctypes.byref(h))
        if st != 0:
            raise _err(lib, st)
        return cls(h)

    @property
    def messages(self):
        return self.lib.rt_scene_file_messages(self.h).decode(errors="replace")

    def arrays(self):
        """Evaluate the assignments into a scene description and the camera."""
        d, cam = _abi.RtSceneDesc(), _abi.RtCamera()
        st = self.lib.rt_scene_file_scene(self.h, ctypes.byref(d), ctypes.byref(cam))
        if st != 0:
            raise _err(self.lib, st)
        return FileSceneArrays(self, d, cam)

    def dump_bvh(self, path):
        st = self.lib.rt_scene_file_dump_bvh(self.h, str(path).encode())
        if st != 0:
            raise _err(self.lib, st)

    def dump_bvh_text(self):
        n = self.lib.rt_scene_file_dump_bvh_text(self.h, None, 0)
        if n < 0:
            raise _err(self.lib, int(n))
        buf = ctypes.create_string_buffer(n + 1)
        self.lib.rt_scene_file_dump_bvh_text(self.h, buf, n + 1)
        return buf.raw[:n].decode()

    def close(self):
        if self.h:
            self.lib.rt_scene_file_free(self.h)
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
