"""Oracle intersection, BVH build and traversal against closed forms, FP64, brute force and
the structural invariants of Tools/verify_bvh.cu:42-66."""
import numpy as np
import pytest

import oracle
from helpers import check_layout, vbvh_rays_and_scenes, vbvh_triangle


def tri(a, b, d, sided=0):
    return vbvh_triangle(a, b, d, sided)


def test_intersect_closed_form():
    t = tri((0, 0, 5), (2, 0, 5), (0, 2, 5))
    d, u, v = oracle.intersect32(t, (0.5, 0.25, 0), (0, 0, 1))
    assert d == pytest.approx(5.0, rel=1e-7) and u == pytest.approx(0.25) and v == pytest.approx(0.125)
    d64, u64, v64 = oracle.intersect64(t, (0.5, 0.25, 0), (0, 0, 1))
    assert (d64, u64, v64) == pytest.approx((5.0, 0.25, 0.125), rel=1e-12)


def test_intersect_rejections():
    t = tri((0, 0, 5), (2, 0, 5), (0, 2, 5))
    assert oracle.intersect32(t, (0.5, 0.25, 0), (1, 0, 0))[0] == -1          # parallel, |det| < 1e-7
    assert oracle.intersect32(t, (3, 3, 0), (0, 0, 1))[0] == -1               # outside (u + v > 1)
    assert oracle.intersect32(t, (0.5, 0.25, 5), (0, 0, 1))[0] == -1          # t = 0 <= 1e-4
    assert oracle.intersect32(t, (0.5, 0.25, 5 - 5e-5), (0, 0, 1))[0] == -1   # t = 5e-5 <= 1e-4
    assert oracle.intersect32(t, (0.5, 0.25, 10), (0, 0, -1))[0] == pytest.approx(5)
    one = tri((0, 0, 5), (2, 0, 5), (0, 2, 5), sided=1)
    # e1 x e2 points to +z; det = -d.(e1 x e2): accepted from +z looking down (d.z < 0) only.
    assert oracle.intersect32(one, (0.5, 0.25, 0), (0, 0, 1))[0] == -1
    assert oracle.intersect32(one, (0.5, 0.25, 10), (0, 0, -1))[0] == pytest.approx(5)
    zero = tri((0, 0, 0), (0, 0, 0), (0, 0, 0))
    assert oracle.intersect32(zero, (0, 0, -1), (0, 0, 1))[0] == -1


def test_fp32_vs_fp64_distance():
    rng = np.random.default_rng(5)
    for _ in range(3000):
        a, b, d = rng.uniform(-1, 1, (3, 3))
        t = tri(a, b, d)
        o = rng.uniform(-3, 3, 3).astype(np.float32)
        target = (a + rng.uniform(0, 1) * (b - a) * 0.5 + rng.uniform(0, 1) * (d - a) * 0.5).astype(np.float32)
        dirv = (target - o).astype(np.float32)
        dirv /= np.linalg.norm(dirv)
        t32, _, _ = oracle.intersect32(t, o, dirv)
        t64, u64, v64 = oracle.intersect64(t, o.astype(np.float64), dirv.astype(np.float64))
        # FP32 Moller-Trumbore loses accuracy as det -> 0 (grazing rays); the 1e-5 relative
        # bound is asserted for rays at least ~11.5 degrees off the triangle plane.
        n = np.cross(b - a, d - a)
        cosine = abs(float(np.dot(n, dirv))) / float(np.linalg.norm(n))
        if t32 > 0 and t64 > 0 and min(u64, v64, 1 - u64 - v64) > 1e-4 and cosine > 0.2:
            assert abs(t32 - t64) <= 1e-5 * max(1.0, abs(t64))


@pytest.fixture(scope="module")
def vbvh():
    return vbvh_rays_and_scenes()


@pytest.mark.parametrize("name", ["edges", "scattered", "overlap", "unbalanced"])
@pytest.mark.parametrize("bins", [12, 32])
def test_bvh_layout_and_brute_force(vbvh, name, bins):
    rays, scenes = vbvh
    tris = scenes[name]
    bvh = oracle.Bvh(tris, bins=bins)
    nodes, order = bvh.export()
    check_layout(nodes, order, tris)
    for any_hit in (False, True):
        a = bvh.trace(rays, any_hit)
        b = oracle.brute32(tris, rays, any_hit)
        if any_hit:
            assert np.array_equal(a["index"] >= 0, b["index"] >= 0)
        else:
            assert np.array_equal(a["index"], b["index"])
            hit = a["index"] >= 0
            assert np.array_equal(a["distance"][hit], b["distance"][hit])
            assert np.array_equal(a["u"][hit], b["u"][hit]) and np.array_equal(a["v"][hit], b["v"][hit])


def test_depth_limit_forced_chain(vbvh):
    """Tools/verify_bvh.cu:70-83: a hand-built depth-32 chain fills every stack slot."""
    rays, scenes = vbvh
    tris = np.concatenate([scenes["edges"][:1]] * 33)
    base = oracle.Bvh(tris).export()[0][0]
    nodes = [base.copy()]
    parent = 0
    for depth in range(32):
        child = len(nodes)
        nodes += [base.copy(), base.copy()]
        nodes[parent]["first"], nodes[parent]["count"] = child, 0
        nodes[child + 1]["first"], nodes[child + 1]["count"] = depth, 1
        parent = child
    nodes[parent]["first"], nodes[parent]["count"] = 32, 1
    nodes = np.array(nodes, dtype=oracle.NODE)
    order = np.arange(33, dtype=np.int32)
    check_layout(nodes, order, tris)
    bvh = oracle.Bvh(tris, nodes=nodes, order=order)
    for any_hit in (False, True):
        a = bvh.trace(rays, any_hit)
        b = oracle.brute32(tris, rays, any_hit)
        assert np.array_equal(a["index"] >= 0, b["index"] >= 0)
        if not any_hit:
            assert np.array_equal(a["index"], b["index"])


def test_reference_tie_rule():
    # Two identical triangles: every hit must report the lower index (Native/renderer.cu:152-153).
    t = tri((-1, -1, 0), (1, -1, 0), (1, 1, 0))
    tris = np.concatenate([t] * 6)
    rays = np.zeros(50, dtype=oracle.RAY)
    rng = np.random.default_rng(1)
    rays["origin"][:, :2] = rng.uniform(-.2, .8, (50, 2)) * np.array([1, -1]) + np.array([0, 0.0])
    rays["origin"][:, 0] = rng.uniform(0.1, 0.9, 50)
    rays["origin"][:, 1] = rng.uniform(-0.9, -0.1, 50) * 0.5
    rays["origin"][:, 2] = -2
    rays["direction"][:, 2] = 1
    rays["limit"] = 1e30
    hits = oracle.Bvh(tris).trace(rays)
    assert (hits["index"] == 0).all()


def test_sah_cost_known_tree():
    # Root box 2x2x2 (half area 12) with two unit-cube leaves (half area 3) of 3 triangles:
    # cost = 12/12 + 2 * 3/12 * 3 = 2.5.
    nodes = np.zeros(3, dtype=oracle.NODE)
    nodes[0]["low"], nodes[0]["high"], nodes[0]["first"], nodes[0]["count"] = (0, 0, 0), (2, 2, 2), 1, 0
    nodes[1]["low"], nodes[1]["high"], nodes[1]["count"] = (0, 0, 0), (1, 1, 1), 3
    nodes[2]["low"], nodes[2]["high"], nodes[2]["count"] = (1, 1, 1), (2, 2, 2), 3
    assert oracle.sah_cost(nodes) == pytest.approx(2.5)


def test_12_bin_build_matches_recorded_dragon_shape():
    """Benchmarks/embree/native.json:128-130 records 659 nodes, depth 17 for the 982-triangle
    Suzanne room; bvh.cuh is unchanged since that record (Benchmarks/embree/build.json:7)."""
    from helpers import suzanne
    tris = suzanne()[0]
    bvh = oracle.Bvh(tris, bins=12)
    nodes, _ = bvh.export()
    assert len(nodes) == 659 and bvh.depth == 17


# ---------------------------------------------------------------- watertight test (Woop et al. 2013)
def _fan(n=7, r=1.0, center=(0.1, -0.2), z=0.0):
    """n triangles around a shared centre vertex; every triangle's origin is the centre, so the
    shared vertices reconstruct to the same floats in both triangles of each edge."""
    from helpers import vbvh_triangle
    c = np.array([center[0], center[1], z], np.float32)
    ang = np.linspace(0, 2 * np.pi, n, endpoint=False) + 0.3
    pts = [np.array([center[0] + r * np.cos(a), center[1] + r * np.sin(a), z], np.float32) for a in ang]
    return np.concatenate([vbvh_triangle(c, pts[i], pts[(i + 1) % n]) for i in range(n)])


def _rays_at_fan_seams(count, seed, n=7, center=(0.1, -0.2)):
    """Rays from both sides aimed at the centre vertex and at points on the shared spokes."""
    rng = np.random.default_rng(seed)
    o = np.column_stack([rng.uniform(-2, 2, count), rng.uniform(-2, 2, count),
                         rng.uniform(0.5, 3, count) * np.where(rng.uniform(size=count) < .5, 1, -1)])
    ang = 0.3 + 2 * np.pi * rng.integers(0, n, count) / n
    s = np.where(rng.uniform(size=count) < 0.2, 0.0, rng.uniform(0, 0.9, count))
    tgt = np.column_stack([center[0] + s * np.cos(ang), center[1] + s * np.sin(ang), np.zeros(count)])
    d = tgt - o
    rays = np.zeros(count, oracle.RAY)
    rays["origin"] = o.astype(np.float32)
    rays["direction"] = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    rays["limit"] = 1e30
    return rays


def test_watertight_closed_form_cases():
    from helpers import vbvh_triangle
    tri = vbvh_triangle((0, 0, 0), (2, 0, 0), (0, 2, 0))
    t, u, v = oracle.intersect_wt32(tri, (0.5, 0.25, 3.0), (0, 0, -1))
    assert (t, u, v) == (3.0, 0.25, 0.125)                       # exact: powers of two throughout
    assert oracle.intersect_wt32(tri, (0.5, 0.25, 3.0), (0, 0, 1))[0] == -1          # behind the origin
    assert oracle.intersect_wt32(tri, (0.5, 0.25, 1e-5), (0, 0, -1))[0] == -1        # t <= 1e-4
    assert oracle.intersect_wt32(tri, (1.5, 1.5, 3.0), (0, 0, -1))[0] == -1          # outside
    assert oracle.intersect_wt32(tri, (0.5, 0.5, 0.0), (1, 0, 0))[0] == -1           # parallel, in plane
    one = vbvh_triangle((0, 0, 0), (2, 0, 0), (0, 2, 0), sided=1)                   # normal +z
    assert oracle.intersect_wt32(one, (0.5, 0.25, 3.0), (0, 0, -1))[0] == 3.0        # front
    assert oracle.intersect_wt32(one, (0.5, 0.25, -3.0), (0, 0, 1))[0] == -1         # back face culled
    zero = vbvh_triangle((1, 1, 0), (1, 1, 0), (1, 1, 0))
    assert oracle.intersect_wt32(zero, (1, 1, 3.0), (0, 0, -1))[0] == -1             # zero area: det 0


def test_watertight_agrees_with_fp64_moller_trumbore():
    """Away from the triangle edges the watertight test accepts exactly the rays FP64
    Moller-Trumbore accepts, with t, u, v within 1e-5 of it."""
    from scenes import triangles_from_vertices
    rng = np.random.default_rng(8)
    tris = triangles_from_vertices(*(rng.uniform(-1, 1, (400, 3)) for _ in range(3)))
    checked = 0
    for k in range(3000):
        i = int(rng.integers(len(tris)))
        v0 = np.array([tris[i]["origin"][a] for a in "xyz"], np.float64)
        e1 = np.array([tris[i]["edge1"][a] for a in "xyz"], np.float64)
        e2 = np.array([tris[i]["edge2"][a] for a in "xyz"], np.float64)
        b = rng.uniform(-0.3, 1.0, 2)
        target = v0 + b[0] * e1 + b[1] * e2
        o = (target + rng.normal(size=3)).astype(np.float32)
        d = ((target - o) / np.linalg.norm(target - o)).astype(np.float32)
        t64, u64, v64 = oracle.intersect64(tris[i], o.astype(np.float64), d.astype(np.float64))
        n = np.cross(e1, e2)
        if abs(np.dot(n, d)) / np.linalg.norm(n) < 0.2:
            continue                                             # grazing: both lose digits
        inside = b[0] > 1e-3 and b[1] > 1e-3 and b[0] + b[1] < 1 - 1e-3
        outside = b[0] < -1e-3 or b[1] < -1e-3 or b[0] + b[1] > 1 + 1e-3
        t, u, v = oracle.intersect_wt32(tris[i], o, d)
        if inside:
            assert t > 0 and t64 > 0
            assert abs(t - t64) <= 1e-5 * max(1, t64) and abs(u - u64) <= 1e-5 and abs(v - v64) <= 1e-5
            checked += 1
        elif outside:
            assert t == -1
    assert checked > 500


def test_watertight_fan_has_no_gaps():
    """Rays through the shared vertex and spokes of a triangle fan: the watertight test hits a
    triangle every time; the reference's test lets some through (gaps at the seams)."""
    tris = _fan()
    rays = _rays_at_fan_seams(50000, 4)
    wt = oracle.brute_wt32(tris, rays)
    mt = oracle.brute32(tris, rays)
    assert (wt["index"] >= 0).all()
    assert (mt["index"] < 0).sum() > 100
    both = (mt["index"] >= 0) & (wt["index"] == mt["index"])
    assert np.abs(wt["distance"][both] - mt["distance"][both]).max() <= 1e-5 * 3
