"""Scenes read by the scene-language front end, rendered and traced on the
GPU: the box written as a scene file renders bit-identically to the box
built by the host scene model, and an OBJ mesh loaded through obj(...)
traces like the oracle and renders like the same mesh handed over as an
indexed Mesh."""
import numpy as np
import pytest

from oracle.oracle import Oracle
from renderer_amd import MODE_UNBIASED, scenes
from renderer_amd.scene import DiffuseMaterial, IntersectableList, Mesh, PhongMaterial, PointLight, Quad, Scene
from renderer_amd.scene_file import SceneFile
from tests.helpers import compare_hits, random_rays

pytestmark = pytest.mark.gpu

BOX = __import__("os").path.join(__import__("os").path.dirname(__file__), "golden", "box.scene")


def test_box_scene_file_renders_like_the_host_model(gpu_ctx):
    a = SceneFile.load(BOX).arrays()
    b = scenes.box_scene(256, 256).flatten()
    views = [(a.camera, 2, 2)]
    print("box.scene: 256x256, 4 spp, unbiased, seed 5", flush=True)
    gpu_ctx.upload(a)
    img_file = gpu_ctx.render(views, mode=MODE_UNBIASED, max_depth=8, seed=5)[0]
    gpu_ctx.upload(b)
    img_host = gpu_ctx.render(views, mode=MODE_UNBIASED, max_depth=8, seed=5)[0]
    assert img_file.mean() > 0
    assert np.array_equal(img_file, img_host)


def _write_obj(path, P, F):
    with open(path, "w") as f:
        f.write("# displaced sphere\n")
        f.write("".join(f"v {x:.9g} {y:.9g} {z:.9g}\n" for x, y, z in P))
        f.write("".join(f"f {a + 1} {b + 1} {c + 1}\n" for a, b, c in F))


def test_obj_mesh_traces_like_the_oracle_and_renders_like_a_mesh(gpu_ctx, tmp_path):
    P, F = scenes.displaced_sphere("tetra", 32, 0.1, 3)  # 4096 triangles
    _write_obj(tmp_path / "sphere.obj", P, F)
    (tmp_path / "mesh.scene").write_text(
        "camera = camera([0 0.9 3.6], [0 -0.05 0], [0 1 0], 0.6981317008, 96, 96);\n"
        "lights = lights{ pointlight([2.5 2.5 2], [8 8 8]) };\n"
        "intersectable = intersectables{\n"
        '  obj("sphere.obj", phong([0.6 0.6 0.6], [0.3 0.3 0.3], 50))\n'
        "  quad([-6 -1.15 -6], [-6 -1.15 6], [6 -1.15 6], [6 -1.15 -6], diffuse([0.5 0.5 0.5]))\n"
        "};\n")
    sf = SceneFile.load(tmp_path / "mesh.scene")
    a = sf.arrays()
    assert len(a.tri_prim) == len(F) + 2 and a.n_prims == 1 + len(F) + 1
    print(f"obj scene: {len(F)} triangles; trace parity on 50000 rays", flush=True)
    gpu_ctx.upload(a)
    orc = Oracle(a, accel=1)
    o4, d4 = random_rays(50000, (-1.5, -1.2, -1.5), (1.5, 1.5, 1.5), seed=4)
    pg, tg = gpu_ctx.trace_closest(o4, d4)
    pr, tr = orc.trace_closest(o4, d4)
    exact, ties, bad = compare_hits(orc, o4, d4, pg, tg, pr, tr)
    print(f"exact {exact}, ties {ties}, bad {len(bad)}", flush=True)
    assert len(bad) == 0 and exact > 10000
    view = (a.camera, 2, 2)
    img_file = gpu_ctx.render([view], mode=MODE_UNBIASED, max_depth=6, seed=9)[0]
    host = Scene(a.camera, [PointLight((2.5, 2.5, 2), (8, 8, 8))],
                 IntersectableList([Mesh(P, F, PhongMaterial((0.6, 0.6, 0.6), (0.3, 0.3, 0.3), 50.0)),
                                    Quad((-6, -1.15, -6), (-6, -1.15, 6), (6, -1.15, 6), (6, -1.15, -6),
                                         DiffuseMaterial((0.5, 0.5, 0.5)))]))
    gpu_ctx.upload(host)
    img_host = gpu_ctx.render([view], mode=MODE_UNBIASED, max_depth=6, seed=9)[0]
    assert img_file.mean() > 0
    assert np.array_equal(img_file, img_host)


CSG_SCENE = """
camera = camera([0.3 0.8 4.5], [0 -0.3 0], [0 1 0], 0.75, 48, 48);
lights = lights{ arealight([-0.5 2.9 -0.5], [1 0 0], [0 0 1], [9 9 8]) pointlight([2 2 2], [2 2 2]) };
grey = diffuse([0.7 0.7 0.7]);
red = diffuse([0.1 0.2 0.7]);
intersectable = intersectables{
  plane([0 1 0 1], grey)
  box([-2 -1 -1.5], [-0.9 0.2 -0.4], red)
  diff(sphere([0.6 -0.2 -0.6], 0.8, phong([0.2 0.4 0.4], [0.3 0.3 0.3], 30)), box([0.1 -0.3 -0.1], [1.5 1 0.5], grey))
  instance(translate([0 0.5 0]) * scale([0.5 0.5 0.5]),
           xor(sphere([-3 1 1], 0.9, grey), sphere([-2.4 1 1], 0.9, red)))
};
"""


def test_csg_scene_file_renders_like_the_host_model_and_the_oracle(gpu_ctx):
    from renderer_amd import MODE_COMPAT
    from renderer_amd.scene import (AreaLight, Box, Camera, CSGDifference, CSGXor, Plane, Sphere)
    a = SceneFile.parse(CSG_SCENE).arrays()
    grey, red = DiffuseMaterial((0.7, 0.7, 0.7)), DiffuseMaterial((0.1, 0.2, 0.7))
    shiny = PhongMaterial((0.2, 0.4, 0.4), (0.3, 0.3, 0.3), 30.0)
    host = Scene(Camera((0.3, 0.8, 4.5), (0, -0.3, 0), (0, 1, 0), 0.75, 48, 48),
                 [AreaLight((-0.5, 2.9, -0.5), (1, 0, 0), (0, 0, 1), (9, 9, 8)), PointLight((2, 2, 2), (2, 2, 2))],
                 IntersectableList([
                     Plane([0, 1, 0, 1], grey),
                     Box((-2, -1, -1.5), (-0.9, 0.2, -0.4), red),
                     CSGDifference(Sphere((0.6, -0.2, -0.6), 0.8, shiny), Box((0.1, -0.3, -0.1), (1.5, 1, 0.5), grey)),
                     CSGXor(Sphere((-1.5, 1.0, 0.5), 0.45, grey), Sphere((-1.2, 1.0, 0.5), 0.45, red)),
                 ])).flatten()
    views = [(a.camera, 2, 2)]
    print("CSG scene file: 48x48, 4 spp", flush=True)
    gpu_ctx.upload(a)
    img_file = gpu_ctx.render(views, mode=MODE_UNBIASED, max_depth=6, seed=9)[0]
    gpu_ctx.upload(host)
    img_host = gpu_ctx.render(views, mode=MODE_UNBIASED, max_depth=6, seed=9)[0]
    assert img_file.mean() > 0
    assert np.array_equal(img_file, img_host)
    gpu_ctx.upload(a)
    img = gpu_ctx.render(views, mode=MODE_COMPAT, max_depth=5, seed=4)[0]
    ref, _ = Oracle(a).render(views[0], mode=MODE_COMPAT, max_depth=5, seed=4)
    agree = np.all(np.abs(img - ref) <= 1e-3 * np.abs(ref) + 1e-4, axis=2).mean()
    print(f"scene file vs oracle agreement {agree:.4f}", flush=True)
    assert agree > 0.97
