"""The host-side export of the 4-wide BVH (rt_bvh_export) and the oracle's
traversal of it in the device's visiting order (oc_flat_trace), on the CPU.

The flat traversal is the oracle that counts what the device traversal
fetches (interior nodes) and tests (primitive records) per ray.  It is
pinned here by
  * its hits: closest hit and any hit equal to the oracle's reference BVH
    (leaf = 1 primitive, both children visited, src/bvh.cpp), which shares
    no code with it;
  * its counts against what pruned nearest-first traversal fixes
    independently of visiting order: every node whose box is entered no
    later than the final closest hit must be fetched (and every record of a
    leaf entered that early tested), and nothing outside the boxes the ray
    enters within its range can be; a ray missing the root box fetches the
    root alone; an empty range fetches nothing;
  * the export itself: every record in exactly one leaf, each child box
    enclosing what lies below it, records equal to the scene's primitives.
The device counts are compared with it in test_gpu_flat_counts.py."""
import numpy as np
import pytest

from oracle.oracle import Oracle, flat_trace
from renderer_amd import BVH_MEDIAN, BVH_SAH, _abi, scenes
from renderer_amd.integrator import bvh_export
from renderer_amd.scene import DiffuseMaterial, IntersectableList, Scene, Sphere
from tests.helpers import camera_rays, compare_hits, random_rays

SLACK = 2.0 ** -20


def _words(nodes):
    return nodes[:, 24:28].copy().view(np.int32)


def _leaf(word):
    x = (~np.int64(word)) & 0xFFFFFFFF
    return int(x >> 5), int(x & 7), int((x >> 3) & 3)


def _g_to_prim(arr):
    area = [i for i, l in enumerate(arr.lights) if l.type == _abi.LIGHT_AREA]
    return np.concatenate([np.asarray(arr.tri_prim, np.int64), np.asarray(arr.sphere_prim, np.int64),
                           np.asarray(area, np.int64)])


def _sphere_cloud_scene(n=300, seed=5):
    """The mesh scene (16,384 triangles) plus a cloud of spheres, so that
    leaves of all three kinds occur."""
    base = scenes.mesh_scene(64, 64, n=64, seed=1)
    rng = np.random.default_rng(seed)
    grey = DiffuseMaterial((0.5, 0.5, 0.5))
    balls = [Sphere(tuple(rng.uniform(-2.5, 2.5, 3)), float(rng.uniform(0.02, 0.2)), grey) for _ in range(n)]
    return Scene(base.camera, base.light, IntersectableList(list(base.object.items) + balls))


@pytest.fixture(scope="module", params=[BVH_MEDIAN, BVH_SAH], ids=["median", "sah"])
def cloud(request):
    arr = _sphere_cloud_scene().flatten()
    arr.bvh_build = request.param
    nodes, recs, need = bvh_export(arr)
    return arr, nodes, recs, need


def _child_box(nodes, i, k):
    nd = nodes[i]
    return np.array([nd[k], nd[8 + k], nd[16 + k]]), np.array([nd[4 + k], nd[12 + k], nd[20 + k]])


def test_export_structure(cloud):
    arr, nodes, recs, need = cloud
    nT, nS = len(arr.tri_prim), len(arr.sphere_prim)
    nP = sum(1 for l in arr.lights if l.type == _abi.LIGHT_AREA)
    assert len(recs) == nT + nS + nP
    assert 0 < need <= 32
    words = _words(nodes)
    g = recs[:, 3].copy().view(np.int32)
    assert sorted(g.tolist()) == list(range(len(recs)))
    seen = np.zeros(len(recs), np.int32)
    reached = np.zeros(len(nodes), np.int32)
    reached[0] = 1
    P = np.asarray(arr.positions, np.float32)
    TV = np.asarray(arr.tri_vertices).reshape(-1, 3)
    S = np.asarray(arr.spheres, np.float32).reshape(-1, 4)
    for i in range(len(nodes)):
        for k in range(4):
            w = int(words[i, k])
            lo, hi = _child_box(nodes, i, k)
            if w >= 0:
                reached[w] += 1
                # the child box encloses the boxes of the child node
                clo = np.array([nodes[w, 0:4].min(), nodes[w, 8:12].min(), nodes[w, 16:20].min()])
                chi = np.array([nodes[w, 4:8].max(), nodes[w, 12:16].max(), nodes[w, 20:24].max()])
                assert np.all(lo <= clo) and np.all(hi >= chi)
                continue
            first, count, kind = _leaf(w)
            if count == 0:
                assert np.all(lo > hi)  # unused slot: empty box
                continue
            for r in range(first, first + count):
                seen[r] += 1
                gi = int(g[r])
                if kind == 0:
                    assert gi < nT
                    p = P[TV[gi]]
                    assert np.array_equal(recs[r, 0:3], p[0])
                    assert np.array_equal(recs[r, 4:7], p[1] - p[0])
                    assert np.array_equal(recs[r, 8:11], p[2] - p[0])
                    pts = p
                elif kind == 1:
                    assert nT <= gi < nT + nS
                    s = S[gi - nT]
                    assert np.array_equal(recs[r, 0:3], s[:3]) and recs[r, 4] == s[3]
                    pts = np.stack([s[:3] - s[3], s[:3] + s[3]])
                else:
                    assert kind == 2 and gi >= nT + nS
                    o, u, v = recs[r, 0:3], recs[r, 4:7], recs[r, 8:11]
                    pts = np.stack([o, o + u, o + v, o + u + v])
                assert np.all(lo <= pts.min(0)) and np.all(hi >= pts.max(0))
    assert np.all(seen == 1)
    assert np.all(reached == 1)  # a tree: every node has one parent (the root none)


def test_flat_closest_matches_reference_bvh(cloud):
    arr, nodes, recs, _ = cloud
    orc = Oracle(arr)
    o4, d4 = random_rays(20000, (-2.5, -1.2, -2.5), (2.5, 3.0, 2.5), seed=11)
    g, t, visits, tests = flat_trace(nodes, recs, o4, d4)
    prim = np.where(g >= 0, _g_to_prim(arr)[np.maximum(g, 0)], -1).astype(np.int32)
    ref_prim, ref_t = orc.trace_closest(o4, d4)
    exact, ties, bad = compare_hits(orc, o4, d4, prim, t, ref_prim, ref_t, rtol=0)
    assert len(bad) == 0, bad[:10]
    assert exact + ties == len(o4)
    assert (g >= 0).mean() > 0.3
    assert np.all(visits >= 1) and np.all(tests >= 0)


def test_flat_any_matches_reference_bvh(cloud):
    arr, nodes, recs, _ = cloud
    orc = Oracle(arr)
    o4, d4 = random_rays(20000, (-2.5, -1.2, -2.5), (2.5, 3.0, 2.5), seed=12)
    d4[:, 3] = np.random.default_rng(13).uniform(0.05, 3.0, len(d4))  # shadow-ray ranges
    occ, visits, tests = flat_trace(nodes, recs, o4, d4, any_hit=True)
    assert np.array_equal(occ, orc.trace_any(o4, d4))
    # an any-hit traversal never fetches more than the closest-hit one
    _, _, cv, ct = flat_trace(nodes, recs, o4, d4)
    assert 0.2 < occ.mean() < 0.9
    assert visits.sum() <= cv.sum() and tests.sum() <= ct.sum()


def _entry(nodes, i, k, o, inv, tmin, tmax):
    """Entry distance of child box k of node i, with the device's relative
    slack, in float32; inf when the ray misses it within [tmin, tmax]."""
    lo, hi = _child_box(nodes, i, k)
    with np.errstate(invalid="ignore", over="ignore"):
        a = (lo - o) * inv
        b = (hi - o) * inv
        near = np.where(inv >= 0, a, b)
        far = np.where(inv >= 0, b, a)
        tl = np.float32(max(np.max(near), tmin))
        th = np.float32(min(np.min(far), tmax))
        if np.isnan(tl) or np.isnan(th) or not tl <= th + abs(th) * np.float32(SLACK):
            return np.float32(np.inf)
    return tl


def _bounds(nodes, recs, o, d, tmin, tmax, t_hit):
    """(must-fetch nodes, must-test records, may-fetch nodes, may-test
    records) of a nearest-first traversal pruned by the closest hit: a child
    entered at or before the final hit distance is reached whatever the order,
    a child not entered within the range never is."""
    inv = np.where(np.abs(d) < 1e-30, np.copysign(np.float32(1e-30), d), d)
    inv = (np.float32(1) / inv).astype(np.float32)
    words = _words(nodes)
    bound = np.float32(t_hit) * np.float32(1 + 4 * SLACK) if np.isfinite(t_hit) else np.float32(np.inf)
    cap = np.nextafter(np.float32(tmax), np.float32(np.inf))
    out = [1, 0, 1, 0]
    for lim, base in ((bound, 0), (cap, 2)):
        todo = [0]
        while todo:
            i = todo.pop()
            for k in range(4):
                key = _entry(nodes, i, k, o, inv, np.float32(tmin), cap)
                if not (np.isfinite(key) and key <= lim):
                    continue
                w = int(words[i, k])
                if w >= 0:
                    out[base] += 1
                    todo.append(w)
                else:
                    out[base + 1] += _leaf(w)[1]
    return out


def test_flat_counts_within_order_free_bounds(cloud):
    arr, nodes, recs, _ = cloud
    o4, d4 = random_rays(300, (-2.5, -1.2, -2.5), (2.5, 3.0, 2.5), seed=14)
    g, t, visits, tests = flat_trace(nodes, recs, o4, d4)
    lo_n = lo_t = 0
    slack_used = slack_room = 0
    for i in range(len(o4)):
        must_n, must_t, may_n, may_t = _bounds(nodes, recs, o4[i, :3], d4[i, :3], o4[i, 3], d4[i, 3],
                                               t[i] if g[i] >= 0 else np.inf)
        if g[i] >= 0:
            slack_used += visits[i] - must_n
            slack_room += may_n - must_n
        # the must-sets use the hit with a few ulps of margin, so a tie at the
        # boundary can lift them by at most the boxes entered exactly there
        assert must_n - 2 <= visits[i] <= may_n, (i, must_n, visits[i], may_n)
        assert must_t - 8 <= tests[i] <= may_t, (i, must_t, tests[i], may_t)
        lo_n += visits[i] >= must_n
        lo_t += tests[i] >= must_t
    assert lo_n >= 0.97 * len(o4) and lo_t >= 0.97 * len(o4)
    # nearest first: hit rays fetch little beyond the must-set (an unordered
    # or unpruned traversal would fetch most of the may-set)
    assert slack_room > 200 and slack_used <= 0.3 * slack_room, (slack_used, slack_room)


def test_flat_counts_special_cases(cloud):
    _, nodes, recs, _ = cloud
    # far outside the scene box, pointing away: only the root is fetched
    o4 = np.array([[50, 50, 50, 0.001], [0, 0, 0, 2.0]], np.float32)
    d4 = np.array([[1, 0, 0, np.inf], [0, 1, 0, 1.0]], np.float32)
    g, t, visits, tests = flat_trace(nodes, recs, o4, d4)
    assert g[0] == -1 and visits[0] == 1 and tests[0] == 0
    assert g[1] == -1 and visits[1] == 0 and tests[1] == 0  # tmin > tmax: empty range
    occ, v2, t2 = flat_trace(nodes, recs, o4, d4, any_hit=True)
    assert occ.tolist() == [0, 0] and v2.tolist() == [1, 0] and t2.tolist() == [0, 0]


def test_flat_closest_box_scene():
    """Quads, spheres and the area light (parallelogram leaves) of the box
    scene, rays from inside it: the flat traversal's hits equal the
    reference BVH's."""
    arr = scenes.box_scene(32, 32).flatten()
    nodes, recs, _ = bvh_export(arr)
    orc = Oracle(arr)
    o4, d4 = random_rays(20000, (0.02, 0.02, 0.02), (0.98, 0.98, 1.5), seed=15)
    g, t, _, _ = flat_trace(nodes, recs, o4, d4)
    prim = np.where(g >= 0, _g_to_prim(arr)[np.maximum(g, 0)], -1).astype(np.int32)
    ref_prim, ref_t = orc.trace_closest(o4, d4)
    exact, ties, bad = compare_hits(orc, o4, d4, prim, t, ref_prim, ref_t, rtol=0)
    assert len(bad) == 0, bad[:10]
    light = [i for i, l in enumerate(arr.lights) if l.type == _abi.LIGHT_AREA][0]
    assert (prim == light).sum() > 0  # the parallelogram is hit


def test_export_capacity_and_errors():
    import ctypes
    lib = _abi.load()
    arr = scenes.box_scene(8, 8).flatten()
    nn, nr, need = ctypes.c_int64(0), ctypes.c_int64(0), ctypes.c_int32(0)
    assert lib.rt_bvh_export(ctypes.byref(arr.desc()), None, ctypes.byref(nn), None, ctypes.byref(nr),
                             ctypes.byref(need)) == 0
    assert nn.value >= 1 and nr.value == len(arr.tri_prim) + len(arr.sphere_prim) + 1
    small = np.zeros((nn.value, 32), np.float32)
    recs = np.zeros((nr.value - 1, 12), np.float32)
    cap_n, cap_r = ctypes.c_int64(nn.value), ctypes.c_int64(nr.value - 1)
    st = lib.rt_bvh_export(ctypes.byref(arr.desc()), small.ctypes.data, ctypes.byref(cap_n), recs.ctypes.data,
                           ctypes.byref(cap_r), None)
    assert st == _abi.ERR_INVALID and cap_r.value == nr.value
    assert lib.rt_bvh_export(None, None, ctypes.byref(nn), None, ctypes.byref(nr), None) == _abi.ERR_INVALID


def test_sah_tree_cuts_traversal_work():
    """The surface-area-heuristic tree of the mesh scene (16,384 triangles on
    a ground quad) fetches fewer nodes and tests fewer records per ray than
    the median tree, for camera rays and for rays from inside the scene, with
    the same hits."""
    sc = scenes.mesh_scene(64, 64, n=64, seed=1)
    arr = sc.flatten()
    rays = [camera_rays(sc, 1, seed=5), random_rays(20000, (-1.3, -1.2, -1.3), (1.3, 1.3, 1.3), seed=6)]
    work = {}
    hits = {}
    for rule in (BVH_MEDIAN, BVH_SAH):
        arr.bvh_build = rule
        nodes, recs, need = bvh_export(arr)
        assert 0 < need < 48
        work[rule], hits[rule] = [], []
        for o4, d4 in rays:
            g, t, v, te = flat_trace(nodes, recs, o4, d4)
            work[rule].append((v.mean(), te.mean()))
            hits[rule].append((g, t))
    for k in range(len(rays)):
        (vm, tm), (vs, ts) = work[BVH_MEDIAN][k], work[BVH_SAH][k]
        assert vs < 0.85 * vm and ts < tm, (k, work)
        gm, tmd = hits[BVH_MEDIAN][k]
        gs, tsa = hits[BVH_SAH][k]
        same = (gm == gs) & ((gm < 0) | (tmd == tsa))
        assert same.mean() > 0.999
