"""Seeded scene generators for the benchmark and test workloads, as TRIANGLE_DTYPE arrays.

Colours cross the ABI sRGB-encoded and the library linearises them (Native/renderer.cu:347),
so a wanted *linear* albedo a is passed as srgb_encode(a).

  cornell()                 36 triangles: 5 walls (white .73, red (.65,.05,.05), green
                            (.12,.45,.15)), a one-sided 2-triangle ceiling light of radiance
                            15 and two boxes (24 triangles)
  triangle_soup(n, seed)    n triangles, centres U[0,1]^3, two edge lengths
                            lognormal(ln .01, .5), uniform random orientation, albedo U(.2,.8)
  displaced_icosphere(...)  icosphere (20*4^subdiv triangles) displaced along the normal by
                            6-octave seeded value-noise fBm (amplitude .05, lacunarity 2,
                            gain .5), Lambertian or metal (p = .3, roughness U(0,1)), inside a
                            20-triangle room with two emissive ceiling panels
  displaced_geodesic(...)   the same surface, materials and room on a Class I geodesic sphere
                            of frequency f (20*f^2 triangles; f = 592 gives 7,009,280)
"""
import math

import numpy as np

from native_api import TRIANGLE_DTYPE, Camera, Settings, Vector


def srgb_encode(linear):
    x = np.asarray(linear, dtype=np.float64)
    return np.where(x <= 0.0031308, 12.92 * x, 1.055 * np.power(np.maximum(x, 0), 1 / 2.4) - 0.055)


def _set3(arr, field, values):
    values = np.asarray(values, dtype=np.float32)
    arr[field]["x"], arr[field]["y"], arr[field]["z"] = values[..., 0], values[..., 1], values[..., 2]


def triangles_from_vertices(v0, v1, v2, color=(1, 1, 1), emission=0.0, one_sided=0, normals=None,
                            specular=(1, 1, 1), smoothness=0.0, specular_probability=0.0, mesh=0, srgb=True):
    """Vectorised triangle records; colours are linear unless srgb=False."""
    v0 = np.asarray(v0, np.float64).reshape(-1, 3)
    v1 = np.asarray(v1, np.float64).reshape(-1, 3)
    v2 = np.asarray(v2, np.float64).reshape(-1, 3)
    n = len(v0)
    t = np.zeros(n, dtype=TRIANGLE_DTYPE)
    o = v0.astype(np.float32)
    _set3(t, "origin", o)
    _set3(t, "edge1", (v1.astype(np.float32) - o))
    _set3(t, "edge2", (v2.astype(np.float32) - o))
    col = np.broadcast_to(np.asarray(color, np.float64), (n, 3))
    spec = np.broadcast_to(np.asarray(specular, np.float64), (n, 3))
    _set3(t, "color", srgb_encode(col) if srgb else col)
    _set3(t, "specularColor", srgb_encode(spec) if srgb else spec)
    t["emission"] = emission
    t["oneSided"] = one_sided
    if normals is None:
        g = np.cross(v1 - v0, v2 - v0)
        g /= np.maximum(np.linalg.norm(g, axis=1, keepdims=True), 1e-30)
        normals = (g, g, g)
    for k, name in enumerate(("normal0", "normal1", "normal2")):
        _set3(t, name, np.broadcast_to(normals[k], (n, 3)))
    t["mesh"] = mesh
    t["smoothness"] = smoothness
    t["specularProbability"] = specular_probability
    return t


def quad(a, b, c, d, **kw):
    """Two triangles (a, b, c) and (a, c, d); winding sets the one-sided facing."""
    return triangles_from_vertices([a, a], [b, c], [c, d], **kw)


def box(center, size, angle_deg, **kw):
    cx, cy, cz = center
    sx, sy, sz = size
    a = math.radians(angle_deg)
    ca, sa = math.cos(a), math.sin(a)

    def p(x, y, z):
        x, z = x * sx / 2, z * sz / 2
        return (cx + ca * x - sa * z, cy + y * sy, cz + sa * x + ca * z)

    faces = [
        (p(-1, 1, -1), p(-1, 1, 1), p(1, 1, 1), p(1, 1, -1)),     # top
        (p(-1, 0, -1), p(1, 0, -1), p(1, 0, 1), p(-1, 0, 1)),     # bottom
        (p(-1, 0, -1), p(-1, 1, -1), p(1, 1, -1), p(1, 0, -1)),   # front
        (p(1, 0, 1), p(1, 1, 1), p(-1, 1, 1), p(-1, 0, 1)),       # back
        (p(-1, 0, 1), p(-1, 1, 1), p(-1, 1, -1), p(-1, 0, -1)),   # left
        (p(1, 0, -1), p(1, 1, -1), p(1, 1, 1), p(1, 0, 1)),       # right
    ]
    return np.concatenate([quad(*f, **kw) for f in faces])


def cornell():
    white, red, green = (.73, .73, .73), (.65, .05, .05), (.12, .45, .15)
    parts = [
        quad((-1, 0, -1), (-1, 0, 1), (1, 0, 1), (1, 0, -1), color=white),      # floor
        quad((-1, 2, -1), (1, 2, -1), (1, 2, 1), (-1, 2, 1), color=white),      # ceiling
        quad((-1, 0, 1), (-1, 2, 1), (1, 2, 1), (1, 0, 1), color=white),        # back
        quad((-1, 0, -1), (-1, 2, -1), (-1, 2, 1), (-1, 0, 1), color=red),      # left
        quad((1, 0, 1), (1, 2, 1), (1, 2, -1), (1, 0, -1), color=green),        # right
        # light: winding gives a downward geometric normal, one-sided
        quad((-.25, 1.99, -.25), (.25, 1.99, -.25), (.25, 1.99, .25), (-.25, 1.99, .25),
             color=(1, 1, 1), emission=15.0, one_sided=1),
        box((-.35, 0, .3), (.6, 1.2, .6), 18, color=white),
        box((.4, 0, -.35), (.6, .6, .6), -17, color=white),
    ]
    tris = np.concatenate(parts)
    assert len(tris) == 36
    return tris


def cornell_camera(width=256, height=256):
    return make_camera((0, 1, -3.4), (0, 1, 1), fov_deg=40, aspect=width / height)


def make_camera(origin, target, fov_deg=60.0, aspect=1.0, up=(0, 1, 0), orthographic=0, ortho_size=1.0):
    o = np.asarray(origin, np.float64)
    f = np.asarray(target, np.float64) - o
    f /= np.linalg.norm(f)
    r = np.cross(np.asarray(up, np.float64), f)      # left-handed: right = up x forward
    r /= np.linalg.norm(r)
    u = np.cross(f, r)
    return Camera(Vector(*o), Vector(*r), Vector(*u), Vector(*f), math.tan(math.radians(fov_deg) / 2), aspect,
                  ortho_size, orthographic)


def make_settings(max_bounces, rpp=1, seed=42, sky=0.0, exposure=1.0, aperture=0.0, focus=1.0):
    return Settings(max_bounces, exposure, aperture, focus, sky, seed, rpp)


def random_rotations(rng, n):
    """Uniform random rotation matrices (n, 3, 3) from unit quaternions."""
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    return np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)], -1),
        np.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], -1),
        np.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], -1),
    ], 1)


def triangle_soup(n, seed=1):
    rng = np.random.default_rng(seed)
    centers = rng.uniform(0, 1, size=(n, 3))
    l1 = rng.lognormal(math.log(0.01), 0.5, size=n)
    l2 = rng.lognormal(math.log(0.01), 0.5, size=n)
    theta = rng.uniform(math.pi / 6, 5 * math.pi / 6, size=n)
    local = np.zeros((n, 3, 3))
    local[:, 1, 0] = l1
    local[:, 2, 0] = l2 * np.cos(theta)
    local[:, 2, 1] = l2 * np.sin(theta)
    local -= local.mean(axis=1, keepdims=True)
    rot = random_rotations(rng, n)
    world = np.einsum("nij,nkj->nki", rot, local) + centers[:, None, :]
    albedo = rng.uniform(0.2, 0.8, size=(n, 3))
    return triangles_from_vertices(world[:, 0], world[:, 1], world[:, 2], color=albedo)


def soup_camera(width=1920, height=1080):
    return make_camera((0.5, 0.5, -1.1), (0.5, 0.5, 0.5), fov_deg=55, aspect=width / height)


# ----------------------------------------------------------------------------- icosphere
def _icosahedron():
    t = (1 + 5 ** 0.5) / 2
    v = np.array([(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
                  (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)], dtype=np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
                  (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5),
                  (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)], dtype=np.int64)
    return v, f


def icosphere(subdiv):
    """Unit icosphere with shared vertices: (vertices (V,3) float64, faces (F,3) int64)."""
    v, f = _icosahedron()
    for _ in range(subdiv):
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        e.sort(axis=1)
        key = e[:, 0] * (len(v) + 1) + e[:, 1]
        uniq, inv = np.unique(key, return_inverse=True)
        a, b = uniq // (len(v) + 1), uniq % (len(v) + 1)
        mid = v[a] + v[b]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        m = inv.reshape(3, -1) + len(v)
        v = np.concatenate([v, mid])
        m01, m12, m20 = m
        f = np.concatenate([np.stack([f[:, 0], m01, m20], 1), np.stack([f[:, 1], m12, m01], 1),
                            np.stack([f[:, 2], m20, m12], 1), np.stack([m01, m12, m20], 1)])
    return v, f


def _lattice_hash(ix, iy, iz, seed):
    h = (ix * 73856093) ^ (iy * 19349663) ^ (iz * 83492791) ^ (seed * 2654435761)
    h = h & 0xFFFFFFFF
    h ^= h >> 16
    h = (h * 0x7FEB352D) & 0xFFFFFFFF
    h ^= h >> 15
    h = (h * 0x846CA68B) & 0xFFFFFFFF
    h ^= h >> 16
    return h.astype(np.float64) / 4294967295.0 * 2 - 1


def value_noise(p, seed):
    """Trilinear value noise in [-1, 1] with smoothstep weights."""
    fl = np.floor(p)
    i = fl.astype(np.int64)
    fr = p - fl
    w = fr * fr * (3 - 2 * fr)
    out = np.zeros(len(p))
    for dx in (0, 1):
        wx = w[:, 0] if dx else 1 - w[:, 0]
        for dy in (0, 1):
            wy = w[:, 1] if dy else 1 - w[:, 1]
            for dz in (0, 1):
                wz = w[:, 2] if dz else 1 - w[:, 2]
                out += wx * wy * wz * _lattice_hash(i[:, 0] + dx, i[:, 1] + dy, i[:, 2] + dz, seed)
    return out


def fbm(p, seed, octaves=6, lacunarity=2.0, gain=0.5, frequency=4.0):
    total, amp, freq, norm = np.zeros(len(p)), 1.0, frequency, 0.0
    for o in range(octaves):
        total += amp * value_noise(p * freq, seed + o)
        norm += amp
        amp *= gain
        freq *= lacunarity
    return total / norm


def room(size=4.0, height=4.0, light=1.0, light_emission=12.0):
    s, h, l = size, height, light
    white = (.7, .7, .7)
    parts = [
        quad((-s, 0, -s), (-s, 0, s), (s, 0, s), (s, 0, -s), color=white),
        quad((-s, h, -s), (s, h, -s), (s, h, s), (-s, h, s), color=white),
        quad((-s, 0, s), (-s, h, s), (s, h, s), (s, 0, s), color=(.6, .6, .75)),
        quad((s, 0, -s), (s, h, -s), (-s, h, -s), (-s, 0, -s), color=white),
        quad((-s, 0, -s), (-s, h, -s), (-s, h, s), (-s, 0, s), color=(.65, .2, .15)),
        quad((s, 0, s), (s, h, s), (s, h, -s), (s, 0, -s), color=(.15, .45, .2)),
        quad((-l, h - .01, -l), (l, h - .01, -l), (l, h - .01, l), (-l, h - .01, l), color=(1, 1, 1),
             emission=light_emission, one_sided=1),
        quad((-s + .2, h - .01, -.3), (-s + .6, h - .01, -.3), (-s + .6, h - .01, .3), (-s + .2, h - .01, .3),
             color=(1, .9, .8), emission=light_emission, one_sided=1),
        quad((-1.5, 0, -1.5), (-1.5, .25, -1.5), (1.5, .25, -1.5), (1.5, 0, -1.5), color=white),
        quad((-1.5, .25, -1.5), (-1.5, .25, 1.5), (1.5, .25, 1.5), (1.5, .25, -1.5), color=white),
    ]
    t = np.concatenate(parts)
    assert len(t) == 20
    return t


def geodesic_sphere(frequency):
    """Class I geodesic sphere: every icosahedron face cut into frequency^2 triangles on its
    barycentric grid, grid points projected to the unit sphere.  Vertices on the icosahedron's
    edges and corners are shared by the faces that meet there, so the mesh is closed.
    Returns (vertices (10 f^2 + 2, 3) float64, faces (20 f^2, 3) int64), faces wound as the
    icosahedron's (outward)."""
    nu = int(frequency)
    if nu < 1:
        raise ValueError("frequency must be >= 1")
    cv, cf = _icosahedron()
    # Edges, each once with its lower corner first; interior edge points k = 1..nu-1 from that corner.
    pairs = np.sort(np.concatenate([cf[:, [0, 1]], cf[:, [1, 2]], cf[:, [2, 0]]]), axis=1)
    edges = np.unique(pairs, axis=0)
    edge_id = {(int(a), int(b)): e for e, (a, b) in enumerate(edges)}
    k = np.arange(1, nu, dtype=np.float64) / nu
    ep = cv[edges[:, 0], None, :] + (cv[edges[:, 1]] - cv[edges[:, 0]])[:, None, :] * k[None, :, None]
    n_edge = nu - 1
    edge_base = len(cv)
    # Interior grid points of each face: i >= 1, j >= 1, i + j <= nu - 1.
    ii, jj = np.meshgrid(np.arange(nu + 1), np.arange(nu + 1), indexing="ij")
    inner = (ii >= 1) & (jj >= 1) & (ii + jj <= nu - 1)
    gi, gj = ii[inner], jj[inner]
    n_inner = len(gi)
    face_base = edge_base + len(edges) * n_edge
    fp = (cv[cf[:, 0], None, :] + (cv[cf[:, 1]] - cv[cf[:, 0]])[:, None, :] * (gi / nu)[None, :, None]
          + (cv[cf[:, 2]] - cv[cf[:, 0]])[:, None, :] * (gj / nu)[None, :, None])
    verts = np.concatenate([cv, ep.reshape(-1, 3), fp.reshape(-1, 3)])
    verts /= np.linalg.norm(verts, axis=1, keepdims=True)
    inner_index = np.full((nu + 1, nu + 1), -1, np.int64)
    inner_index[gi, gj] = np.arange(n_inner)

    def edge_points(a, b, t):
        """Vertex ids of the points at steps t (1..nu-1) from corner a toward corner b."""
        e = edge_id[(min(a, b), max(a, b))]
        step = t if a < b else nu - t
        return edge_base + e * n_edge + step - 1

    faces = []
    for fi, (c0, c1, c2) in enumerate(cf):
        c0, c1, c2 = int(c0), int(c1), int(c2)
        ids = np.full((nu + 1, nu + 1), -1, np.int64)
        ids[0, 0], ids[nu, 0], ids[0, nu] = c0, c1, c2
        t = np.arange(1, nu)
        ids[t, 0] = edge_points(c0, c1, t)
        ids[0, t] = edge_points(c0, c2, t)
        ids[nu - t, t] = edge_points(c1, c2, t)
        ids[gi, gj] = face_base + fi * n_inner + inner_index[gi, gj]
        ui, uj = np.nonzero(np.add.outer(np.arange(nu + 1), np.arange(nu + 1)) <= nu - 1)
        up = np.stack([ids[ui, uj], ids[ui + 1, uj], ids[ui, uj + 1]], 1)
        di, dj = np.nonzero(np.add.outer(np.arange(nu), np.arange(nu)) <= nu - 2)
        down = np.stack([ids[di + 1, dj], ids[di + 1, dj + 1], ids[di, dj + 1]], 1)
        faces.append(np.concatenate([up, down]))
    return verts, np.concatenate(faces)


def displaced_icosphere(subdiv=9, seed=7, amplitude=0.05, metal_probability=0.3, center=(0, 1.3, 0), radius=1.0):
    return _displaced_mesh(*icosphere(subdiv), seed, amplitude, metal_probability, center, radius)


def displaced_geodesic(frequency=592, seed=7, amplitude=0.05, metal_probability=0.3, center=(0, 1.3, 0), radius=1.0):
    return _displaced_mesh(*geodesic_sphere(frequency), seed, amplitude, metal_probability, center, radius)


def _displaced_mesh(v, f, seed, amplitude, metal_probability, center, radius):
    rng = np.random.default_rng(seed)
    disp = fbm(v, seed) * amplitude
    pos = v * (radius + disp)[:, None] + np.asarray(center)
    # Area-weighted vertex normals of the displaced surface, for smooth shading.
    fn = np.cross(pos[f[:, 1]] - pos[f[:, 0]], pos[f[:, 2]] - pos[f[:, 0]])
    vn = np.zeros_like(pos)
    for k in range(3):
        for c in range(3):
            vn[:, c] += np.bincount(f[:, k], weights=fn[:, c], minlength=len(pos))
    vn /= np.maximum(np.linalg.norm(vn, axis=1, keepdims=True), 1e-30)
    n = len(f)
    metal = rng.uniform(size=n) < metal_probability
    albedo = rng.uniform(0.2, 0.8, size=(n, 3))
    rough = rng.uniform(0, 1, size=n)
    t = triangles_from_vertices(pos[f[:, 0]], pos[f[:, 1]], pos[f[:, 2]], color=albedo,
                                normals=(vn[f[:, 0]], vn[f[:, 1]], vn[f[:, 2]]), specular=albedo,
                                smoothness=np.where(metal, 1 - rough, 0.0).astype(np.float32),
                                specular_probability=np.where(metal, 1.0, 0.0).astype(np.float32))
    return np.concatenate([t, room()])


def icosphere_camera(width=1920, height=1080):
    return make_camera((0, 1.6, -3.6), (0, 1.25, 0), fov_deg=50, aspect=width / height)


# ----------------------------------------------------------------------------- batched views
def batch_views(count, seed=2024):
    """Batched views: scenes from the soup / displaced-sphere generators with triangle counts
    log-uniform in [1e3, 2e6], image sides log-uniform in [256, 1024], 2..8 bounces."""
    rng = np.random.default_rng(seed)
    out = []
    for v in range(count):
        n = int(round(10 ** rng.uniform(3, np.log10(2e6))))
        w, h = (int(round(2 ** rng.uniform(8, 10))) for _ in range(2))
        bounces = int(rng.integers(2, 9))
        if v % 2 == 0:
            tris = triangle_soup(n, seed=seed + v)
            cam, st = soup_camera(w, h), make_settings(bounces, sky=1.0, seed=seed + v)
        else:
            sub = int(np.clip(round(np.log(max(n - 20, 20) / 20) / np.log(4)), 1, 8))
            tris = displaced_icosphere(subdiv=sub, seed=seed + v)
            cam, st = icosphere_camera(w, h), make_settings(bounces, seed=seed + v)
        out.append((tris, w, h, cam, st))
    return out
