"""Synthetic benchmark scenes (seeded, deterministic).

box_scene        plain PT box: 5 axis-aligned quads, one 0.25 x 0.25 area
                 light under the ceiling (radiance 17,12,4), a diffuse and a
                 mirror sphere (r = 0.2), 40 degree pinhole camera, 256^2.
mesh_scene       a displaced sphere of 1,048,576 triangles with a Phong
                 material (kd .6, ks .3, n 50) on a diffuse ground quad, one
                 area light and one point light, 1024^2.
batched_views    32 seeded camera views of the mesh scene with mixed
                 resolutions and sample counts.
env_scene        the mesh scene lit only by a procedural lat-long sky (512 x 256)
                 with a sun disk.
room_scene       a closed room of 12 quads whose only emitter sits outside
                 behind a door gap of 5% of the wall width, holding 8
                 displaced meshes of 524,288 triangles each.

Tessellation: 20 * 4^k triangles (an icosphere) never equals 2^20, so the
spheres here subdivide a tetrahedron (4 n^2 triangles; n = 512 gives
1,048,576) or an octahedron (8 n^2; n = 256 gives 524,288) with n segments
per edge and project the vertices onto the unit sphere before the radial
displacement by 4-octave value noise.
"""
import math

import numpy as np

from .scene import (AreaLight, Camera, DiffuseMaterial, EnvironmentMap, IntersectableList, Mesh,
                    MirrorMaterial, PhongMaterial, PointLight, Quad, Scene, Sphere)

# ---------------------------------------------------------------------------
# geometry helpers
# ---------------------------------------------------------------------------
_TETRA_V = np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]], np.float64) / math.sqrt(3.0)
_TETRA_F = np.array([[0, 1, 2], [0, 3, 1], [0, 2, 3], [1, 3, 2]], np.int64)
_OCTA_V = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]], np.float64)
_OCTA_F = np.array([[0, 2, 4], [2, 1, 4], [1, 3, 4], [3, 0, 4], [2, 0, 5], [1, 2, 5], [3, 1, 5], [0, 3, 5]],
                   np.int64)


def subdivided_sphere(base, n):
    """Unit sphere from a subdivided platonic solid: `base` 'tetra' (4 n^2
    triangles) or 'octa' (8 n^2).  Returns (vertices float64 (V,3), faces
    int32 (F,3)) with vertices shared along face edges."""
    V, Fb = (_TETRA_V, _TETRA_F) if base == "tetra" else (_OCTA_V, _OCTA_F)
    ii, jj = np.meshgrid(np.arange(n + 1), np.arange(n + 1), indexing="ij")
    keep = ii + jj <= n
    ii, jj = ii[keep], jj[keep]
    idx = -np.ones((n + 1, n + 1), np.int64)
    idx[ii, jj] = np.arange(len(ii))
    # local triangles of one face grid
    a_i, a_j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    up = a_i + a_j <= n - 1
    t1 = np.stack([idx[a_i[up], a_j[up]], idx[a_i[up] + 1, a_j[up]], idx[a_i[up], a_j[up] + 1]], 1)
    dn = a_i + a_j <= n - 2
    t2 = np.stack([idx[a_i[dn] + 1, a_j[dn]], idx[a_i[dn] + 1, a_j[dn] + 1], idx[a_i[dn], a_j[dn] + 1]], 1)
    local = np.concatenate([t1, t2])
    pts, faces = [], []
    nloc = len(ii)
    for f, (A, B, C) in enumerate(Fb):
        # integer barycentric weights, so points on shared edges are exact
        # functions of the two edge endpoints and the integer k
        wa, wb, wc = (n - ii - jj), ii, jj
        P = (np.outer(wa, V[A]) + np.outer(wb, V[B]) + np.outer(wc, V[C])) / n
        pts.append(P)
        faces.append(local + f * nloc)
    P = np.concatenate(pts)
    F = np.concatenate(faces)
    key = np.round(P * (1 << 30)).astype(np.int64)
    _, first, inv = np.unique(key, axis=0, return_index=True, return_inverse=True)
    P = P[first]
    F = inv.reshape(-1)[F]
    P /= np.linalg.norm(P, axis=1, keepdims=True)
    return P, F.astype(np.int32)


def value_noise(p, seed, octaves=4, base_freq=2.0):
    """Fractal value noise in [-1, 1]: lattice values from a seeded table,
    smoothstep-trilinear interpolation, octaves at doubling frequency and
    halving amplitude, normalised by the amplitude sum."""
    rng = np.random.default_rng(seed)
    table = rng.uniform(-1.0, 1.0, 4096)
    perm = rng.permutation(4096)

    def lattice(ix, iy, iz):
        h = perm[(ix * 73856093 ^ iy * 19349663 ^ iz * 83492791) & 4095]
        return table[h]

    total = np.zeros(len(p))
    amp, norm = 1.0, 0.0
    for o in range(octaves):
        q = p * (base_freq * (2 ** o)) + 17.0 * o
        i0 = np.floor(q).astype(np.int64)
        f = q - i0
        s = f * f * (3.0 - 2.0 * f)
        acc = np.zeros(len(p))
        for dx in (0, 1):
            for dy in (0, 1):
                for dz in (0, 1):
                    w = (np.where(dx, s[:, 0], 1 - s[:, 0]) * np.where(dy, s[:, 1], 1 - s[:, 1]) *
                         np.where(dz, s[:, 2], 1 - s[:, 2]))
                    acc += w * lattice(i0[:, 0] + dx, i0[:, 1] + dy, i0[:, 2] + dz)
        total += amp * acc
        norm += amp
        amp *= 0.5
    return total / norm


def displaced_sphere(base, n, amplitude, seed, center=(0, 0, 0), radius=1.0):
    P, F = subdivided_sphere(base, n)
    r = radius * (1.0 + amplitude * value_noise(P, seed))
    P = P * r[:, None] + np.asarray(center, np.float64)
    return P.astype(np.float32), F


# ---------------------------------------------------------------------------
# scenes
# ---------------------------------------------------------------------------
def box_scene(width=256, height=256):
    grey = DiffuseMaterial((0.75, 0.75, 0.75))
    red = DiffuseMaterial((0.63, 0.06, 0.05))
    green = DiffuseMaterial((0.14, 0.45, 0.09))
    walls = [
        Quad((0, 0, 0), (1, 0, 0), (1, 0, 1), (0, 0, 1), grey),   # floor y = 0
        Quad((0, 1, 0), (0, 1, 1), (1, 1, 1), (1, 1, 0), grey),   # ceiling y = 1
        Quad((0, 0, 0), (0, 1, 0), (1, 1, 0), (1, 0, 0), grey),   # back z = 0
        Quad((0, 0, 0), (0, 0, 1), (0, 1, 1), (0, 1, 0), red),    # left x = 0
        Quad((1, 0, 0), (1, 1, 0), (1, 1, 1), (1, 0, 1), green),  # right x = 1
    ]
    spheres = [Sphere((0.3, 0.2, 0.62), 0.2, grey), Sphere((0.7, 0.2, 0.35), 0.2, MirrorMaterial(1))]
    light = AreaLight((0.375, 0.999, 0.375), (0.25, 0, 0), (0, 0, 0.25), (17, 12, 4))
    cam = Camera((0.5, 0.5, 2.4), (0.5, 0.5, 0.0), (0, 1, 0), math.radians(40.0), width, height)
    return Scene(cam, [light], IntersectableList(walls + spheres))


def mesh_scene(width=1024, height=1024, n=512, seed=1, env=False):
    """Displaced sphere of 4 n^2 triangles (n = 512: 1,048,576), Phong
    (kd .6, ks .3, n 50), on a diffuse ground quad.  Lights: one 1.5 x 1.5
    area light above (radiance 6) and one point light (intensity 8), or with
    env=True only the procedural sky."""
    P, F = displaced_sphere("tetra", n, 0.1, seed)
    phong = PhongMaterial((0.6, 0.6, 0.6), (0.3, 0.3, 0.3), 50.0)
    ground = Quad((-6, -1.15, -6), (-6, -1.15, 6), (6, -1.15, 6), (6, -1.15, -6), DiffuseMaterial((0.5, 0.5, 0.5)))
    if env:
        lights = [EnvironmentMap(sky_image(seed=seed))]
    else:
        lights = [AreaLight((-0.75, 3.0, -0.75), (1.5, 0, 0), (0, 0, 1.5), (6, 6, 6)),
                  PointLight((2.5, 2.5, 2.0), (8, 8, 8))]
    cam = Camera((0.0, 0.9, 3.6), (0.0, -0.05, 0.0), (0, 1, 0), math.radians(40.0), width, height)
    return Scene(cam, lights, IntersectableList([Mesh(P, F, phong), ground]))


def sky_image(width=512, height=256, seed=1, sun_radiance=5.0e4, sun_diameter_deg=0.5):
    """Lat-long sky: vertical gradient (zenith blue to horizon white, dim
    ground) plus a sun disk of the given angular diameter.  Row 0 is +y;
    column u maps to phi = (u - 1/2) 2 pi with direction
    (sin t sin phi, cos t, -sin t cos phi)."""
    rng = np.random.default_rng(seed)
    v = (np.arange(height) + 0.5) / height
    u = (np.arange(width) + 0.5) / width
    theta = v * math.pi
    phi = (u - 0.5) * 2 * math.pi
    T, PH = np.meshgrid(theta, phi, indexing="ij")
    dirs = np.stack([np.sin(T) * np.sin(PH), np.cos(T), -np.sin(T) * np.cos(PH)], -1)
    up = np.clip(dirs[..., 1], 0.0, 1.0)
    zen = np.array([0.25, 0.45, 1.0])
    hor = np.array([1.0, 1.0, 1.0])
    sky = hor[None, None, :] * (1 - up[..., None]) + zen[None, None, :] * up[..., None]
    sky = np.where(dirs[..., 1:2] < 0, 0.15 * np.ones(3)[None, None, :], sky)
    sun_theta = math.radians(rng.uniform(35.0, 55.0))
    sun_phi = rng.uniform(-math.pi, math.pi)
    sd = np.array([math.sin(sun_theta) * math.sin(sun_phi), math.cos(sun_theta), -math.sin(sun_theta) * math.cos(sun_phi)])
    cosang = dirs @ sd
    half = math.radians(sun_diameter_deg / 2)
    img = sky.copy()
    disk = cosang >= math.cos(half)
    # the disk is smaller than a texel (0.70 x 0.70 deg at 512 x 256): put its
    # power into the texel containing the sun direction
    if not disk.any():
        iy = min(int(sun_theta / math.pi * height), height - 1)
        ix = min(int((sun_phi / (2 * math.pi) + 0.5) * width), width - 1)
        texel_sa = (2 * math.pi / width) * (math.pi / height) * math.sin((iy + 0.5) / height * math.pi)
        sun_sa = 2 * math.pi * (1 - math.cos(half))
        img[iy, ix] += sun_radiance * sun_sa / texel_sa
    else:
        img[disk] += sun_radiance
    return img.astype(np.float32)


def room_scene(width=1920, height=1080, n=256, seed=3):
    """Closed room [-3,3] x [0,3] x [-3,3] of 8 quads (floor, ceiling, three
    walls, and the +x wall as two panels around a door gap of 5% of the wall
    width plus a lintel) and an outside antechamber of 4 quads (outer wall,
    two sides, ceiling) holding the only emitter, a 1 x 1 area light of
    radiance 100 facing the door.  Inside: 8 displaced octahedron spheres
    (8 n^2 triangles each; n = 256 gives 524,288, 4,194,304 in total),
    diffuse albedo per object uniform in [0.3, 0.8] per channel."""
    rng = np.random.default_rng(seed)
    wall = DiffuseMaterial((0.75, 0.75, 0.75))
    gap = 0.05 * 6.0 / 2  # half width of the door gap
    door_h = 2.2
    quads = [
        Quad((-3, 0, -3), (4.5, 0, -3), (4.5, 0, 3), (-3, 0, 3), wall),            # floor (room + antechamber)
        Quad((-3, 3, -3), (-3, 3, 3), (3, 3, 3), (3, 3, -3), wall),                # room ceiling
        Quad((-3, 0, -3), (-3, 3, -3), (3, 3, -3), (3, 0, -3), wall),              # z = -3
        Quad((-3, 0, 3), (3, 0, 3), (3, 3, 3), (-3, 3, 3), wall),                  # z = +3
        Quad((-3, 0, -3), (-3, 0, 3), (-3, 3, 3), (-3, 3, -3), wall),              # x = -3
        Quad((3, 0, -3), (3, 3, -3), (3, 3, -gap), (3, 0, -gap), wall),            # x = +3, left panel
        Quad((3, 0, gap), (3, 3, gap), (3, 3, 3), (3, 0, 3), wall),                # x = +3, right panel
        Quad((3, door_h, -gap), (3, 3, -gap), (3, 3, gap), (3, door_h, gap), wall),  # lintel
        Quad((4.5, 0, -1.5), (4.5, 0, 1.5), (4.5, 3, 1.5), (4.5, 3, -1.5), wall),  # antechamber outer wall
        Quad((3, 0, -1.5), (4.5, 0, -1.5), (4.5, 3, -1.5), (3, 3, -1.5), wall),    # antechamber side
        Quad((3, 0, 1.5), (3, 3, 1.5), (4.5, 3, 1.5), (4.5, 0, 1.5), wall),        # antechamber side
        Quad((3, 3, -1.5), (4.5, 3, -1.5), (4.5, 3, 1.5), (3, 3, 1.5), wall),      # antechamber ceiling
    ]
    # light on the antechamber's outer wall, facing -x: u x v must point to -x
    light = AreaLight((4.49, 0.8, -0.5), (0, 0, 1), (0, 1, 0), (100, 100, 100))
    objs = []
    centers = []
    while len(centers) < 8:
        c = np.array([rng.uniform(-2.2, 2.0), 0.0, rng.uniform(-2.2, 2.2)])
        r = rng.uniform(0.35, 0.6)
        c[1] = r * 1.1
        if all(np.linalg.norm(c[[0, 2]] - q[0][[0, 2]]) > r + q[1] + 0.15 for q in centers):
            centers.append((c, r))
    for k, (c, r) in enumerate(centers):
        P, F = displaced_sphere("octa", n, 0.1, seed * 100 + k, center=c, radius=r)
        albedo = rng.uniform(0.3, 0.8, 3)
        objs.append(Mesh(P, F, DiffuseMaterial(albedo)))
    cam = Camera((-2.7, 1.6, 2.6), (0.5, 0.8, -0.5), (0, 1, 0), math.radians(60.0), width, height)
    return Scene(cam, [light], IntersectableList(quads + objs))


def batched_views(count=32, seed=7, resolutions=((256, 256), (512, 384), (1024, 768), (1920, 1080)),
                  spp_range=(8, 128), radius=3.0):
    """Seeded views of the mesh scene: resolution uniform over `resolutions`,
    samples per pixel uniform in spp_range (split into spp_x * spp_y strata,
    spp_x the largest divisor <= sqrt(spp)), eyes uniform on a sphere of the
    given radius (upper hemisphere) looking at the origin."""
    rng = np.random.default_rng(seed)
    views = []
    for _ in range(count):
        w, h = resolutions[rng.integers(len(resolutions))]
        spp = int(rng.integers(spp_range[0], spp_range[1] + 1))
        sx = max(d for d in range(1, int(math.isqrt(spp)) + 1) if spp % d == 0)
        z = rng.uniform(0.05, 0.9)
        ph = rng.uniform(0, 2 * math.pi)
        s = math.sqrt(1 - z * z)
        eye = (radius * s * math.cos(ph), radius * z, radius * s * math.sin(ph))
        cam = Camera(eye, (0, 0, 0), (0, 1, 0), math.radians(40.0), w, h)
        views.append((cam, sx, spp // sx))
    return views
