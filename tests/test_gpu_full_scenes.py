"""The C3 scenes at full size: the displaced subdivision-9 icosphere (5,242,880 triangles) and
the displaced frequency-592 geodesic sphere (7,009,280 triangles), each in its 20-triangle
room, built by the GPU with the benchmark's launch options.

- Hit records: camera rays of the 1920x1080 frame (jittered, every 6th pixel), first-bounce
  rays from their hits (cosine-distributed about the geometric normal, origin offset 5e-4 as
  Native/renderer.cu:244) and shadow rays toward points on the emitters (limit dist - .001,
  Native/renderer.cu:261) are traced by RtTraceRays and by the oracle's 12-bin reference
  build (Native/bvh.cuh, Native/renderer.cu:136-178): closest-hit index, t, u and v are
  identical bit for bit; any-hit queries agree on occlusion.
- The full 1080p frame of the 7.0M scene (8 bounces + RR, 1 spp) against the oracle: HDR mean
  relative error <= 1e-3, the display gate, covered counts exact, ray counts within 1e-3."""
import numpy as np
import pytest

import oracle
import scenes
from bench import TUNED
from native_api import Session
from test_gpu_configs import assert_frame_matches, gpu_frames, oracle_frames

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(900)]

W, H = 1920, 1080
DEFAULT_BINS = 16   # the library's build_bins default


def _f3(rec, name):
    return np.stack([rec[name][a] for a in "xyz"], -1).astype(np.float32)


def frame_rays(tris, cam, oscene, stride=6, seed=3):
    """(primary, bounce, shadow) RAY arrays derived from the frame's camera; hits from the oracle."""
    rng = np.random.default_rng(seed)
    ys, xs = np.mgrid[0:H:stride, 0:W:stride]
    xs = xs.ravel() + rng.uniform(0, 1, xs.size)
    ys = ys.ravel() + rng.uniform(0, 1, ys.size)
    o = np.array([cam.origin.x, cam.origin.y, cam.origin.z], np.float32)
    r = np.array([cam.right.x, cam.right.y, cam.right.z], np.float32)
    u = np.array([cam.up.x, cam.up.y, cam.up.z], np.float32)
    f = np.array([cam.forward.x, cam.forward.y, cam.forward.z], np.float32)
    px = (2 * xs / W - 1).astype(np.float32)[:, None]
    py = (2 * ys / H - 1).astype(np.float32)[:, None]
    d = f + r * (px * np.float32(cam.aspect * cam.tanHalfFov)) + u * (py * np.float32(cam.tanHalfFov))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    prim = np.zeros(len(d), oracle.RAY)
    prim["origin"], prim["direction"], prim["limit"] = o, d, 1e30
    hits = oscene.trace(prim)
    ok = hits["index"] >= 0
    assert ok.mean() > 0.99   # the room encloses the camera
    idx = hits["index"][ok]
    e1, e2 = _f3(tris[idx], "edge1"), _f3(tris[idx], "edge2")
    ng = np.cross(e1, e2)
    ng /= np.linalg.norm(ng, axis=1, keepdims=True)
    ng *= np.where((ng * d[ok]).sum(1) > 0, -1, 1)[:, None]
    p = (o + d[ok] * hits["distance"][ok][:, None] + ng * 5e-4).astype(np.float32)
    # cosine-distributed directions about ng
    a, b = rng.uniform(0, 1, len(p)), rng.uniform(0, 2 * np.pi, len(p))
    helper = np.where(np.abs(ng[:, 1:2]) < .999, [[0, 1, 0]], [[1, 0, 0]])
    t = np.cross(helper, ng)
    t /= np.linalg.norm(t, axis=1, keepdims=True)
    s = np.cross(ng, t)
    dirs = ng * np.sqrt(a)[:, None] + (t * np.cos(b)[:, None] + s * np.sin(b)[:, None]) * np.sqrt(1 - a)[:, None]
    bounce = np.zeros(len(p), oracle.RAY)
    bounce["origin"], bounce["limit"] = p, 1e30
    bounce["direction"] = (dirs / np.linalg.norm(dirs, axis=1, keepdims=True)).astype(np.float32)
    lights = np.nonzero(tris["emission"] > 0)[0]
    li = lights[rng.integers(0, len(lights), len(p))]
    su, sv = rng.uniform(0, 1, len(p)), rng.uniform(0, 1, len(p))
    flip = su + sv > 1
    su, sv = np.where(flip, 1 - su, su), np.where(flip, 1 - sv, sv)
    q = _f3(tris[li], "origin") + _f3(tris[li], "edge1") * su[:, None] + _f3(tris[li], "edge2") * sv[:, None]
    to = q - p
    dist = np.linalg.norm(to, axis=1)
    shadow = np.zeros(len(p), oracle.RAY)
    shadow["origin"], shadow["direction"] = p, (to / dist[:, None]).astype(np.float32)
    shadow["limit"] = (dist - .001).astype(np.float32)
    return prim, bounce, shadow


@pytest.fixture(scope="module", params=["icosphere9", "geodesic592"])
def full_scene(request):
    if request.param == "icosphere9":
        tris = scenes.displaced_icosphere(subdiv=9, seed=7)
        assert len(tris) == 5_242_880 + 20
    else:
        tris = scenes.displaced_geodesic(frequency=592, seed=7)
        assert len(tris) == 7_009_280 + 20
    print(f"{request.param}: {len(tris)} triangles generated", flush=True)
    yield request.param, tris, oracle.OracleScene(tris)


def test_full_scene_hit_records_match_oracle(lib, full_scene):
    name, tris, osc = full_scene
    cam = scenes.icosphere_camera(W, H)
    prim, bounce, shadow = frame_rays(tris, cam, osc)
    closest = np.concatenate([prim, bounce])
    assert len(closest) + len(shadow) >= 64 * 1024
    with Session(lib) as s:
        for k, v in TUNED["icosphere"].items():
            assert s.set_option(k, v), k
        s.initialize(tris, 1, 1)
        got_c = s.trace(closest)
        got_s = s.trace(shadow, any_hit=True)
        assert s.set_option("intersect_watertight", 1)
        wt_c = s.trace(closest)
    changed = wt_c["index"] != got_c["index"]
    print(f"{name}: the watertight test changes {changed.sum()} of {len(closest)} closest hits "
          f"({(got_c['index'][changed] < 0).sum()} were misses of the reference test)", flush=True)
    assert changed.mean() < 1e-3
    ref_c = osc.trace(closest)
    ref_s = osc.trace(shadow, any_hit=True)
    print(f"{name}: {len(closest)} closest-hit rays ({(ref_c['index'] >= 0).mean():.3f} hit), {len(shadow)} "
          f"shadow rays ({(ref_s['index'] >= 0).mean():.3f} occluded)", flush=True)
    assert np.array_equal(got_c["index"], ref_c["index"])
    hit = ref_c["index"] >= 0
    for k in ("distance", "u", "v"):
        assert np.array_equal(got_c[k][hit].view(np.uint32), ref_c[k][hit].view(np.uint32)), k
    assert np.array_equal(got_s["index"] >= 0, ref_s["index"] >= 0)
    # Hit distance against FP64 Moller-Trumbore on the hit triangle: within 1e-5 relative for
    # every hit more than 11.5 degrees off the triangle's plane; the grazing rest is reported.
    t64, cosine = mt64(tris, got_c["index"][hit], closest[hit])
    rel = np.abs(got_c["distance"][hit] - t64) / np.maximum(1, np.abs(t64))
    steep = cosine > np.sin(np.radians(11.5))
    over = rel > 1e-5
    print(f"{name}: {hit.sum()} hits vs FP64: {over.sum()} beyond 1e-5 relative ({over.mean():.4%}), worst "
          f"{rel.max():.2e}; steeper than 11.5 deg: worst {rel[steep].max():.2e}", flush=True)
    assert (rel[steep] <= 1e-5).all()


def mt64(tris, index, rays):
    """FP64 Moller-Trumbore distance of each ray to triangle index[i], and |cos| of incidence."""
    t = tris[index]
    v0, e1, e2 = (np.stack([t[f][a] for a in "xyz"], 1).astype(np.float64) for f in ("origin", "edge1", "edge2"))
    o, d = rays["origin"].astype(np.float64), rays["direction"].astype(np.float64)
    p = np.cross(d, e2)
    inv = 1 / (e1 * p).sum(1)
    q = np.cross(o - v0, e1)
    n = np.cross(e1, e2)
    cosine = np.abs((n * d).sum(1)) / (np.linalg.norm(n, axis=1) * np.linalg.norm(d, axis=1))
    return (e2 * q).sum(1) * inv, cosine


@pytest.mark.parametrize("bins", [12, 16, 32])
def test_full_scene_gpu_build_sah_cost(lib, full_scene, bins):
    """The GPU's binned-SAH build (leaf <= 4, the reference's rule) of the full scene against the
    reference's 12-bin build (Native/bvh.cuh:56-105): at 12 bins it makes the same tree (equal
    SAH cost); at the library's default bin count its SAH cost is within 2% (or lower); every
    build keeps each triangle in exactly one leaf."""
    name, tris, osc = full_scene
    with Session(lib) as s:
        assert s.set_option("build_bins", bins)
        s.initialize(tris, 1, 1)
        nodes, order = s.bvh(), s.bvh_order(len(tris))
        n_nodes, depth, build_ms = s.bvh_info()
    gpu, ref = oracle.sah_cost(nodes), osc.sah_cost
    print(f"{name}: GPU build {bins} bins {build_ms:.1f} ms, {n_nodes} nodes, depth {depth}; SAH cost {gpu:.3f} vs "
          f"reference 12-bin build {ref:.3f} ({gpu / ref:.4f}x)", flush=True)
    if bins == 12:
        assert gpu == pytest.approx(ref, rel=1e-9)
    if bins == DEFAULT_BINS:
        assert gpu <= 1.02 * ref
    assert np.array_equal(np.sort(order), np.arange(len(tris)))
    leaves = nodes[nodes["count"] > 0]
    assert leaves["count"].sum() == len(tris) and (leaves["count"] <= 4).all() or depth >= 32


def test_full_scene_frame_matches_oracle(lib, full_scene):
    name, tris, osc = full_scene
    if name != "geodesic592":
        pytest.skip("the full frame is checked on the 7.0M scene; the 5.24M scene is covered by its hit records")
    cam, st = scenes.icosphere_camera(W, H), scenes.make_settings(8, rpp=1)
    got = gpu_frames(lib, tris, cam, st, W, H, 1, TUNED["icosphere"])
    assert_frame_matches(*got, *oracle_frames(tris, cam, st, W, H, 1, scene=osc), f"C3 {name} 1080p 1 spp")
