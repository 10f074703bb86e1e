"""The scene-language front end (include/rt_scene_file.h) on the CPU: grammar,
variables, OBJ/MTL import, evaluation into rt_scene_desc, and the --dump-bvh
text.  The reference has no tests or scene files for it, so the pins are:
the box workload written in the scene language against the host scene
model that builds it in Python; hand-derived --dump-bvh texts (the median
split and the printing rules of src/bvh.cpp:97-133 and src/scenedumper.cpp
worked out by hand for two small scenes); closed forms for the transforms;
and idempotence of dump -> parse -> dump."""
import math
import os

import numpy as np
import pytest

from renderer_amd import _abi, scenes
from renderer_amd.scene_file import SceneFile

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def soup(a):
    """(n_tri, 3, 3) vertex positions of a scene description."""
    return np.asarray(a.positions)[np.asarray(a.tri_vertices)]


def mat_tuple(m):
    return (m.type, tuple(m.kd), tuple(m.ks), m.exponent)


def light_tuple(L):
    return (L.type, tuple(L.p), tuple(L.u), tuple(L.v), tuple(L.intensity))


def test_box_scene_file_equals_the_host_model():
    sf = SceneFile.load(os.path.join(GOLDEN, "box.scene"))
    a = sf.arrays()
    b = scenes.box_scene(256, 256).flatten()
    assert np.array_equal(soup(a), soup(b))
    assert np.array_equal(a.tri_prim, b.tri_prim)
    assert [mat_tuple(a.materials[i]) for i in a.tri_material] == [mat_tuple(b.materials[i]) for i in b.tri_material]
    assert np.array_equal(a.spheres, b.spheres)
    assert np.array_equal(a.sphere_prim, b.sphere_prim)
    assert ([mat_tuple(a.materials[i]) for i in a.sphere_material] ==
            [mat_tuple(b.materials[i]) for i in b.sphere_material])
    assert [light_tuple(L) for L in a.lights] == [light_tuple(L) for L in b.lights]
    assert a.normals is None
    cam = scenes.box_scene(256, 256).camera
    assert np.array_equal(a.camera.cop, cam.cop) and np.array_equal(a.camera.look_at, cam.look_at)
    assert np.float32(a.camera.fov) == np.float32(cam.fov)
    assert (a.camera.width, a.camera.height) == (256, 256)
    assert a.n_prims == b.n_prims == 8


def test_dump_bvh_matches_hand_derived_texts(tmp_path):
    for name in ("dump_spheres", "dump_obj"):
        sf = SceneFile.load(os.path.join(GOLDEN, name + ".scene"))
        expected = open(os.path.join(GOLDEN, name + ".bvh")).read()
        assert sf.dump_bvh_text() == expected
        out = tmp_path / (name + ".bvh")
        sf.dump_bvh(out)
        assert out.read_text() == expected


def test_dump_parse_dump_is_idempotent():
    sf = SceneFile.load(os.path.join(GOLDEN, "dump_spheres.scene"))
    text = sf.dump_bvh_text()
    again = SceneFile.parse(text)
    assert again.dump_bvh_text() == text
    a, b = sf.arrays(), again.arrays()
    # same primitives; the dumped tree lists them in its leaf order
    assert sorted(map(tuple, a.spheres.tolist())) == sorted(map(tuple, b.spheres.tolist()))
    assert sorted(soup(a).reshape(-1).tolist()) == sorted(soup(b).reshape(-1).tolist())


def test_area_lights_dump_as_the_reference_spells_them():
    # ast_area_light::function_name is "srealight" (src/sceneast.cpp:89): the
    # dump is not readable again, as in the reference
    sf = SceneFile.load(os.path.join(GOLDEN, "box.scene"))
    text = sf.dump_bvh_text()
    assert "lights=lights{srealight([0.375 0.999 0.375],[0.25 0 0],[0 0 0.25],[17 12 4])};" in text
    with pytest.raises(_abi.RtError) as e:
        SceneFile.parse(text)
    assert e.value.status == _abi.ERR_PARSE


def test_obj_import_faces_materials_and_overrides():
    sf = SceneFile.load(os.path.join(GOLDEN, "dump_obj.scene"))
    assert "ObjReader: 3 triangles, 2 faces" in sf.messages
    # b_mat is the MTL's Blinn material (Ks and Ns given)
    a = sf.arrays()
    m = a.materials[a.tri_material[0]]
    assert (m.type, tuple(np.float32(m.kd)), tuple(np.float32(m.ks)), m.exponent) == (
        _abi.MAT_BLINN, tuple(np.float32([0.2, 0.4, 0.6])), tuple(np.float32([0.1, 0.1, 0.1])), 20.0)
    assert a.materials[a.tri_material[2]].type == _abi.MAT_DIFFUSE
    # a later assignment overrides the MTL material of the same name
    base = os.path.join(GOLDEN, "")
    text = open(os.path.join(GOLDEN, "dump_obj.scene")).read()
    text = text.replace("intersectable =", "b_mat = diffuse([0.25 0.5 1]);\nintersectable =")
    a = SceneFile.parse(text, base_dir=base).arrays()
    P = soup(a)
    assert np.array_equal(P, np.array([[[0, 0, 0], [1, 0, 0], [1, 1, 0]],      # fan of f 1 2 3 4
                                       [[0, 0, 0], [1, 1, 0], [0, 1, 0]],
                                       [[0, 0, 0], [1, 1, 0], [0, 1, 0]]],     # f -4 -2 -1
                                      np.float32))
    assert list(a.tri_prim) == [0, 1, 2]  # one primitive per OBJ triangle, no lights
    kd = [tuple(a.materials[i].kd) for i in a.tri_material]
    assert kd == [(0.25, 0.5, 1.0), (0.25, 0.5, 1.0), (0.5, 0.5, 0.5)]
    # the first two triangles carry vertex normals, the third a flat one
    # (p1 - p2) x (p1 - p3), normalised at upload
    N = np.asarray(a.normals).reshape(3, 3, 3)
    assert np.array_equal(N[0], np.tile([0, 0, 1], (3, 1))) and np.array_equal(N[1], np.tile([0, 0, 1], (3, 1)))
    assert np.array_equal(N[2], np.tile([0, 0, 1], (3, 1)))


def test_obj_forms(tmp_path):
    (tmp_path / "m.obj").write_text(
        "v 0 0 0\nv 2 0 0 2\nv 0 1 0\nv 0 0 1\nvt 0 0\nvt 1 0\nvt 0 1\nvn 1 0 0\n"
        "f 1/1 2/2 3/3\n"            # v/t
        "f 1/1/1 3/3/1 4/1/1\r\n"    # v/t/n with a CRLF ending
        "f 1 2\n"                    # fewer than three vertices: no triangle
        "f 2 3 4 1 2\n"              # pentagon: three triangles
        "# comment\n"
        "usemtl 9lives! x\n"         # the first token only; '!' and a leading digit become '_'
        "f 1 2 3\n")
    (tmp_path / "s.scene").write_text('camera = camera([0 0 5],[0 0 0],[0 1 0],0.5,8,8);\n'
                                      'intersectable = obj("m.obj", diffuse([1 1 1]));\n'
                                      '_lives_ = diffuse([0.1 0.2 0.3]);\n'
                                      'intersectable = obj("m.obj", diffuse([1 1 1]));\n')
    sf = SceneFile.load(tmp_path / "s.scene")
    a = sf.arrays()
    assert "ObjReader: 6 triangles, 5 faces" in sf.messages
    P = soup(a)
    assert len(P) == 6
    assert np.array_equal(P[0], [[0, 0, 0], [1, 0, 0], [0, 1, 0]])  # w = 2 divides
    assert np.array_equal(P[2], [[1, 0, 0], [0, 1, 0], [0, 0, 1]])
    assert np.array_equal(P[4], [[1, 0, 0], [0, 0, 0], [1, 0, 0]])
    kd = [tuple(np.float32(a.materials[i].kd)) for i in a.tri_material]
    assert kd[:5] == [(1, 1, 1)] * 5 and kd[5] == tuple(np.float32([0.1, 0.2, 0.3]))


def test_obj_index_out_of_range_and_missing_files(tmp_path):
    (tmp_path / "bad.obj").write_text("v 0 0 0\nf 1 2 3\n")
    (tmp_path / "bad.scene").write_text('intersectable = obj("bad.obj", diffuse([1 1 1]));\n')
    with pytest.raises(_abi.RtError) as e:
        SceneFile.load(tmp_path / "bad.scene")
    assert e.value.status == _abi.ERR_PARSE and "out of range" in str(e.value)
    with pytest.raises(_abi.RtError) as e:
        SceneFile.load(tmp_path / "none.scene")
    assert e.value.status == _abi.ERR_IO
    # an unreadable OBJ is an empty list with a warning, as in the reference
    sf = SceneFile.parse('camera = camera([0 0 5],[0 0 0],[0 1 0],0.5,8,8);\n'
                         'intersectable = obj("missing.obj", diffuse([1 1 1]));', base_dir=str(tmp_path))
    assert "Error opening file" in sf.messages
    assert sf.arrays().n_prims == 0


def test_grammar_forms_and_variables():
    text = """
    // comments of both kinds /* even nested-looking */
    /* block
       comment */ camera = camera([0 0 +5], [0 0 0], [0 1 0], .5, 64.9, 48.);
    m1 = phong([.6 .6 .6], [.3 .3 .3], 5e1);
    m2 = m1;                        // a copy of a material variable
    m1 = diffuse([1 0 0]);          // redefinition does not change m2
    cam2 = camera;
    lights = lights{pointlight([0 5 0],[1 1 1])arealight([0 1 0],[1 0 0],[0 0 1],[2 2 2])};
    scene = intersectables{ quad([0 0 0],[1 0 0],[1 0 1],[0 0 1],m2) sphere([0 0 0],-1e-1,m1) };
    intersectable = scene;
    """
    sf = SceneFile.parse(text)
    a = sf.arrays()
    assert (a.camera.width, a.camera.height) == (64, 48)  # QSize truncates
    assert np.float32(a.camera.fov) == np.float32(0.5) and a.camera.cop[2] == 5
    assert [light_tuple(L)[0] for L in a.lights] == [_abi.LIGHT_POINT, _abi.LIGHT_AREA]
    assert mat_tuple(a.materials[a.tri_material[0]])[0] == _abi.MAT_PHONG
    assert a.materials[a.tri_material[0]].exponent == 50
    assert mat_tuple(a.materials[a.sphere_material[0]])[:2] == (_abi.MAT_DIFFUSE, (1.0, 0.0, 0.0))
    assert np.float32(a.spheres[0, 3]) == np.float32(-0.1)
    assert list(a.tri_prim) == [2, 2] and list(a.sphere_prim) == [3]


def test_later_definitions_win_and_missing_material_is_dark_matter():
    sf = SceneFile.parse("""
    camera = camera([0 0 5],[0 0 0],[0 1 0],0.5,8,8);
    intersectable = intersectables{ sphere([0 0 0], 1, nosuch) };
    x = lights;
    """)
    a = sf.arrays()
    assert a.materials[a.sphere_material[0]].type == _abi.MAT_DARK
    assert "No such variable: nosuch" in sf.messages
    assert "Could not assign to x, no such variable: lights" in sf.messages


def test_parse_errors_name_the_position():
    for text, where in [("camera = camera([0 0 0], [0 0 1]);", "line 1"),
                        ("a = diffuse([1 1 1]);\nb = diffuse([1 1 1])\nc = a;", "line 3"),
                        ("a = diffuse([1 1 1]); /* open", "unterminated comment"),
                        ("a = diffuse([1 1 1]); // no newline at the end", "unterminated comment"),
                        ("a = frobnicate(1);", "line 1, column 15")]:
        with pytest.raises(_abi.RtError) as e:
            SceneFile.parse(text)
        assert e.value.status == _abi.ERR_PARSE
        assert where in str(e.value), (text, str(e.value))
    with pytest.raises(_abi.RtError) as e:
        SceneFile.parse("intersectable = intersectables{};").arrays()
    assert e.value.status == _abi.ERR_INVALID and "camera" in str(e.value)


def test_unsupported_constructs_parse_and_dump_but_do_not_render():
    """Singular instance transforms have no object space to map rays into."""
    cases = {
        "instance(scale([1 0 1]), plane([0 1 0 1], diffuse([1 1 1])))": "singular",
        "instance(scale([1 0 1]), sphere([0 0 0], 1, diffuse([1 1 1])))": "singular",
        "instance(scale([0 1 1]), diff(box([0 0 0],[1 1 1], diffuse([1 1 1])), sphere([0 0 0],1,mirror(1))))":
            "singular",
    }
    for obj, word in cases.items():
        sf = SceneFile.parse("camera = camera([0 0 5],[0 0 0],[0 1 0],0.5,8,8);\n"
                             f"intersectable = {obj};")
        assert obj.split("(")[0] in sf.dump_bvh_text()
        with pytest.raises(_abi.RtError) as e:
            sf.arrays()
        assert e.value.status == _abi.ERR_UNSUPPORTED and word in str(e.value), str(e.value)


def test_planes_boxes_and_csg_map_to_object_programs():
    sf = SceneFile.parse(
        "camera = camera([0 0 5],[0 0 0],[0 1 0],0.5,8,8);\n"
        "lights = lights{pointlight([0 5 0],[1 1 1])};\n"
        "intersectable = intersectables{ plane([0 1 0 1], diffuse([.5 .5 .5]))"
        " box([0 0 0],[1 1 1], mirror(1))"
        " diff(sphere([0 0 0],1,diffuse([1 0 0])), box([0 0 0],[2 2 2],diffuse([0 1 0])))"
        " instance(translate([1 2 3]) * scale([2 2 2]),"
        "          union(sphere([0 0 0],1,diffuse([1 1 1])), box([-1 -1 -1],[0 0 0],diffuse([1 1 1]))))"
        " instance(translate([0 1 0]), plane([0 1 0 0], diffuse([1 1 1]))) };")
    a = sf.arrays()
    from renderer_amd._abi import CSG_BOX, CSG_DIFF, CSG_PLANE, CSG_SPHERE, CSG_UNION
    assert list(a.object_first) == [0, 1, 2, 5, 8, 9]
    assert list(a.object_prim) == [1, 2, 3, 4, 5]          # the point light is primitive 0
    ops = [n.op for n in a.csg_nodes]
    assert ops == [CSG_PLANE, CSG_BOX, CSG_SPHERE, CSG_BOX, CSG_DIFF, CSG_SPHERE, CSG_BOX, CSG_UNION, CSG_PLANE]
    p = [list(n.p) for n in a.csg_nodes]
    assert p[0][:4] == [0, 1, 0, 1]
    assert p[5][:4] == [1, 2, 3, 2]                         # centre moved, radius scaled
    assert p[6] == [-1, 0, 1, 1, 2, 3]                      # box corners mapped
    assert p[8][:4] == [0, 1, 0, -1]                        # the plane y = 1
    mats = a.materials
    assert mats[a.csg_nodes[1].material].type == _abi.MAT_MIRROR


def test_cone_light_and_refractive_material_map_to_the_gpu_kinds():
    sf = SceneFile.parse("camera = camera([0 0 5],[0 0 0],[0 1 0],0.5,8,8);\n"
                         "lights = lights{conelight([0 1 0],[0 -1 0],0.5,[1 2 3])};\n"
                         "intersectable = intersectables{ sphere([0 0 0], 1, refractive(1.5)) };")
    a = sf.arrays()
    L = a.lights[0]
    assert (L.type, tuple(L.p), tuple(L.u), np.float32(L.v[0]), tuple(L.intensity)) == (
        _abi.LIGHT_CONE, (0, 1, 0), (0, -1, 0), np.float32(0.5), (1, 2, 3))
    m = a.materials[a.sphere_material[0]]
    assert (m.type, m.exponent) == (_abi.MAT_REFRACTIVE, 1.5)


def test_missing_environment_map_is_a_black_light():
    sf = SceneFile.parse("camera = camera([0 0 5],[0 0 0],[0 1 0],0.5,8,8);\n"
                         'lights = lights{environmentmap("no/such/probe.hdr", [1 1 1]) pointlight([0 1 0],[1 1 1])};\n'
                         "intersectable = intersectables{ sphere([0 0 0], 1, diffuse([1 1 1])) };")
    a = sf.arrays()
    assert "Failed to load environment map" in sf.messages
    assert [light_tuple(L) for L in a.lights][0] == (_abi.LIGHT_POINT, (0, 0, 0), (0, 0, 0), (0, 0, 0), (0, 0, 0))
    assert list(a.sphere_prim) == [2]


def test_instance_transforms_closed_forms():
    sf = SceneFile.parse("""
    camera = camera([0 0 5],[0 0 0],[0 1 0],0.5,8,8);
    intersectable = intersectables{
      instance(translate([1 2 3]) * scale(2), quad([0 0 0],[1 0 0],[1 0 1],[0 0 1], diffuse([1 1 1])))
      instance(rotate(90, [0 0 1]), t([1 0 0],[2 0 0],[1 1 0],[0 0 0],[0 0 0],[0 0 0],[0 0],[0 0],[0 0],mirror(1)))
      instance([[1 0 0 5][0 1 0 0][0 0 1 0][0 0 0 1]] * rotate(180, [0 1 0]) * scale(3),
               sphere([1 0 0], 0.5, diffuse([1 1 1])))
    };
    """)
    a = sf.arrays()
    P = soup(a)
    q = np.array([[0, 0, 0], [1, 0, 0], [0, 0, 1]], np.float32)  # first quad triangle (p1, p2, p4)
    assert np.array_equal(P[0], 2 * q + [1, 2, 3])
    assert np.array_equal(P[2], [[0, 1, 0], [0, 2, 0], [-1, 1, 0]])  # +90 deg about z: x -> y
    # both triangles of the instanced quad are one primitive, the instance one ID
    assert list(a.tri_prim) == [0, 0, 1]
    # translate(5) * rotate(180 about y) * scale(3): centre (1,0,0) -> (5 - 3, 0, 0), radius 1.5
    assert np.array_equal(a.spheres[0], np.float32([2, 0, 0, 1.5]))
    assert list(a.sphere_prim) == [2]
    # instanced triangles carry the normal matrix applied to their normals
    N = np.asarray(a.normals).reshape(-1, 3, 3)
    assert np.allclose(N[0] / np.linalg.norm(N[0], axis=1, keepdims=True), [[0, -1, 0]] * 3)
    # rotate(45) about z by the axis-angle formula
    sf = SceneFile.parse("camera = camera([0 0 5],[0 0 0],[0 1 0],0.5,8,8);\n"
                         "intersectable = instance(rotate(45, [0 0 2]), "
                         "t([1 0 0],[0 1 0],[0 0 1],[0 0 0],[0 0 0],[0 0 0],[0 0],[0 0],[0 0],mirror(1)));")
    P = soup(sf.arrays())
    c = math.sqrt(0.5)
    assert np.allclose(P[0], [[c, c, 0], [-c, c, 0], [0, 0, 1]], atol=1e-7)


def test_bvh_nodes_in_input_are_lists_of_their_children():
    text = ("camera = camera([0 0 5],[0 0 0],[0 1 0],0.5,8,8);\n"
            "intersectable = b(sphere([0 0 0],1,diffuse([1 1 1])), b(sphere([3 0 0],1,mirror(1)),"
            "sphere([6 0 0],1,mirror(1)), box([2 -1 -1],[7 1 1],diffuse([0 0 0]))),"
            " box([-1 -1 -1],[7 1 1],diffuse([0 0 0])));")
    a = SceneFile.parse(text).arrays()
    assert a.spheres[:, 0].tolist() == [0, 3, 6]
    assert list(a.sphere_prim) == [0, 1, 2]
