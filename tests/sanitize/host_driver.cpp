// Host-only driver for the sanitizer build (AddressSanitizer + UndefinedBehaviorSanitizer).
//
// It runs the code the CPU backends share with the gfx950 kernels - rt_core.h (sampler,
// Moller-Trumbore, slab tests, integrator, display) and traverse.h (W-wide traversal over
// full-precision and quantized nodes) - together with the oracle (oracle/oracle.cpp, linked
// in), on seeded scenes that reach the edge cases: zero-area and one-sided triangles, axis-
// parallel and denormal ray directions, a 200-level unbalanced tree (deep stacks), leaves of
// many triangles, emitters, glossy and mirror materials.  Checks, besides the sanitizers:
//   - closest-hit records bit-identical to the oracle's FP32 brute force, any-hit occlusion equal,
//     for W = 2, 4, 8 full-precision nodes and W = 4 quantized nodes;
//   - an image rendered by the rt_core.h integrator over the wide BVH equal to the oracle's
//     render within 1e-3 HDR mean relative error.
// Wide nodes come from the oracle's binary reference build, collapsed here as the GPU
// collapse does (largest-area inner child opened first).  Exit status 0 when every check
// passes.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>
#include "../../custom-ray-tracing-shader_amd/csrc/traverse.h"

using namespace rt;

extern "C" {
void* or_bvh_build(const Triangle* tris, int n, int bins, int maxLeaf);
void or_bvh_free(void* h);
int or_bvh_node_count(void* h);
void or_bvh_export(void* h, RtBvhNode* nodes, int* order);
void or_brute32(const Triangle* tris, int n, const RtRay* rays, int count, int any, RtHit* hits);
void* or_scene_create(const Triangle* tris, int n, const unsigned* dirs);
void or_scene_free(void* h);
void or_render(void* h, const CameraData* cam, const RenderSettings* rs, int w, int hgt, unsigned sample, float* sum,
               unsigned char* rgba, int threads, int y0, int y1, unsigned long long* rays);
}

static const unsigned kDirs[kSobolDims * 32] = {
#include "../../custom-ray-tracing-shader_amd/csrc/sobol_directions.inc"
};

static int failures = 0;
#define CHECK(cond, ...) do { if (!(cond)) { failures++; std::printf("FAIL: " __VA_ARGS__); std::printf("\n"); } } while (0)

static Triangle tri(f3 a, f3 b, f3 c, f3 color, float emission = 0, int oneSided = 0, float smooth = 0, float spec = 0) {
    Triangle t{};
    t.origin = RtVec3{a.x, a.y, a.z};
    const f3 e1 = b - a, e2 = c - a;
    t.edge1 = RtVec3{e1.x, e1.y, e1.z};
    t.edge2 = RtVec3{e2.x, e2.y, e2.z};
    const f3 n = unitv(crossv(e1, e2));
    t.normal0 = t.normal1 = t.normal2 = RtVec3{n.x, n.y, n.z};
    t.color = RtVec3{color.x, color.y, color.z};
    t.specularColor = t.color;
    t.emission = emission; t.oneSided = oneSided; t.smoothness = smooth; t.specularProbability = spec;
    return t;
}

// Random soup in a closed box with two lights, a mirror and a few degenerate triangles.
static std::vector<Triangle> soupScene(int n, unsigned seed) {
    std::mt19937 rng(seed);
    std::uniform_real_distribution<float> u(0, 1);
    std::vector<Triangle> t;
    auto quad = [&](f3 a, f3 b, f3 c, f3 d, f3 col, float em = 0, int one = 0, float sm = 0, float sp = 0) {
        t.push_back(tri(a, b, c, col, em, one, sm, sp));
        t.push_back(tri(a, c, d, col, em, one, sm, sp));
    };
    const float s = 2;
    quad(F3(-s, -s, -s), F3(-s, -s, s), F3(s, -s, s), F3(s, -s, -s), F3(.6f, .6f, .6f));            // floor
    quad(F3(-s, s, -s), F3(s, s, -s), F3(s, s, s), F3(-s, s, s), F3(.6f, .6f, .6f));                 // ceiling
    quad(F3(-s, -s, s), F3(-s, s, s), F3(s, s, s), F3(s, -s, s), F3(.5f, .5f, .7f));                 // back
    quad(F3(-s, -s, -s), F3(-s, s, -s), F3(-s, s, s), F3(-s, -s, s), F3(.7f, .2f, .2f));             // left
    quad(F3(s, -s, s), F3(s, s, s), F3(s, s, -s), F3(s, -s, -s), F3(.2f, .7f, .2f), 0, 0, 1, 1);     // mirror
    quad(F3(-.5f, s - .01f, -.5f), F3(.5f, s - .01f, -.5f), F3(.5f, s - .01f, .5f), F3(-.5f, s - .01f, .5f),
         F3(1, 1, 1), 10, 1);                                                                          // light
    quad(F3(-1.9f, .5f, -1), F3(-1.9f, .9f, -1), F3(-1.9f, .9f, 1), F3(-1.9f, .5f, 1), F3(1, .8f, .6f), 4, 0);
    for (int i = 0; i < n; i++) {
        const f3 c = F3(u(rng) * 3 - 1.5f, u(rng) * 3 - 1.5f, u(rng) * 3 - 1.5f);
        const f3 a = c + F3(u(rng) * .2f - .1f, u(rng) * .2f - .1f, u(rng) * .2f - .1f);
        const f3 b = c + F3(u(rng) * .2f - .1f, u(rng) * .2f - .1f, u(rng) * .2f - .1f);
        const f3 d = (i % 97 == 0) ? a : c + F3(u(rng) * .2f - .1f, u(rng) * .2f - .1f, u(rng) * .2f - .1f);
        t.push_back(tri(a, b, d, F3(u(rng) * .6f + .2f, u(rng) * .6f + .2f, u(rng) * .6f + .2f), 0, i % 5 == 0,
                        u(rng), i % 3 == 0 ? .3f : 0));
    }
    return t;
}

// Tools/verify_bvh.cu:124-127-style chain: widths grow geometrically, so the SAH tree is deep.
static std::vector<Triangle> chainScene() {
    std::vector<Triangle> t;
    for (int i = 0; i < 200; i++) {
        const float x = std::pow(1.5f, (float)i);
        if (!std::isfinite(x)) break;
        t.push_back(tri(F3(x, -1, 0), F3(x, 1, 0), F3(x, 0, 1), F3(.5f, .5f, .5f)));
    }
    return t;
}

struct WideBvh {
    std::vector<unsigned char> nodes;
    std::vector<TriRecord> recs;
};

static float halfArea(const RtBvhNode& n) {
    const float x = n.high[0] - n.low[0], y = n.high[1] - n.low[1], z = n.high[2] - n.low[2];
    return x * y + x * z + y * z;
}

// The binary reference tree collapsed into N nodes, breadth first, as the GPU collapse.
template <class N>
static WideBvh collapse(const std::vector<Triangle>& tris, const std::vector<RtBvhNode>& bin, const std::vector<int>& order) {
    constexpr int W = NodeTraits<N>::kWidth;
    WideBvh out;
    std::vector<N> wide(1);
    std::vector<std::pair<int, int>> tasks{{0, 0}};   // (binary node, wide slot)
    for (size_t q = 0; q < tasks.size(); q++) {
        const int bi = tasks[q].first, slot = tasks[q].second;
        int kids[W], nk = 0;
        if (bin[bi].count > 0) kids[nk++] = bi;
        else { kids[nk++] = bin[bi].first; kids[nk++] = bin[bi].first + 1; }
        while (nk < W) {
            int pick = -1;
            float area = -1;
            for (int k = 0; k < nk; k++)
                if (bin[kids[k]].count == 0 && halfArea(bin[kids[k]]) > area) { area = halfArea(bin[kids[k]]); pick = k; }
            if (pick < 0) break;
            const int f = bin[kids[pick]].first;
            kids[pick] = f;
            kids[nk++] = f + 1;
        }
        int refs[W];
        float lo[W][3], hi[W][3];
        for (int c = 0; c < nk; c++) {
            const RtBvhNode& k = bin[kids[c]];
            for (int a = 0; a < 3; a++) { lo[c][a] = k.low[a]; hi[c][a] = k.high[a]; }
            if (k.count > 0) refs[c] = makeLeafRef(k.first, k.count);
            else { refs[c] = (int)wide.size(); wide.emplace_back(); tasks.push_back({kids[c], refs[c]}); }
        }
        N& node = wide[slot];
        if constexpr (std::is_same<N, WideNode<W>>::value) {
            for (int c = 0; c < W; c++) {
                node.lox[c] = node.loy[c] = node.loz[c] = node.hix[c] = node.hiy[c] = node.hiz[c] = INFINITY;
                node.child[c] = kEmptyChild; node.count[c] = 0;
            }
            for (int c = 0; c < nk; c++) {
                node.lox[c] = lo[c][0]; node.loy[c] = lo[c][1]; node.loz[c] = lo[c][2];
                node.hix[c] = hi[c][0]; node.hiy[c] = hi[c][1]; node.hiz[c] = hi[c][2];
                node.child[c] = refs[c]; node.count[c] = bin[kids[c]].count;
            }
        } else {
            quantizeNode<W>(node, lo, hi, refs, nk);
        }
    }
    out.nodes.resize(wide.size() * sizeof(N));
    std::memcpy(out.nodes.data(), wide.data(), out.nodes.size());
    std::vector<int> last(order.size(), 0);
    for (const RtBvhNode& n : bin)
        if (n.count > 0) last[n.first + n.count - 1] = 1;
    out.recs.resize(order.size());
    for (size_t j = 0; j < order.size(); j++) {
        const Triangle& t = tris[order[j]];
        TriRecord& r = out.recs[j];
        r.v0x = t.origin.x; r.v0y = t.origin.y; r.v0z = t.origin.z; r.index = order[j];
        r.e1x = t.edge1.x; r.e1y = t.edge1.y; r.e1z = t.edge1.z;
        r.flags = (t.oneSided ? kTriOneSided : 0) | (last[j] ? kTriLastInLeaf : 0);
        r.e2x = t.edge2.x; r.e2y = t.edge2.y; r.e2z = t.edge2.z; r.pad = 0;
    }
    return out;
}

static std::vector<RtRay> randomRays(int count, unsigned seed, float lo, float hi) {
    std::mt19937 rng(seed);
    std::uniform_real_distribution<float> u(0, 1);
    std::vector<RtRay> rays(count);
    for (int i = 0; i < count; i++) {
        RtRay& r = rays[i];
        for (int a = 0; a < 3; a++) r.origin[a] = lo + (hi - lo) * u(rng);
        f3 d = unitv(F3(u(rng) * 2 - 1, u(rng) * 2 - 1, u(rng) * 2 - 1));
        if (i % 4 == 0) d = F3(i % 3 == 0, i % 3 == 1, i % 3 == 2);
        if (i % 5 == 0) d.x = 1e-40f;
        r.direction[0] = d.x; r.direction[1] = d.y; r.direction[2] = d.z;
        r.limit = (i % 2) ? 1e30f : 0.01f + 3 * u(rng);
        r.flags = 0;
    }
    return rays;
}

template <class N>
static void checkQueries(const char* name, const std::vector<Triangle>& tris, const std::vector<RtBvhNode>& bin,
                         const std::vector<int>& order, const std::vector<RtRay>& rays) {
    const WideBvh w = collapse<N>(tris, bin, order);
    const N* nodes = reinterpret_cast<const N*>(w.nodes.data());
    for (int any = 0; any < 2; any++) {
        std::vector<RtHit> ref(rays.size());
        or_brute32(tris.data(), (int)tris.size(), rays.data(), (int)rays.size(), any, ref.data());
        int bad = 0;
        for (size_t i = 0; i < rays.size(); i++) {
            const RtRay& r = rays[i];
            const f3 o = F3(r.origin[0], r.origin[1], r.origin[2]), d = F3(r.direction[0], r.direction[1], r.direction[2]);
            const RtHit h = any ? traverseHost<N, true>(nodes, w.recs.data(), o, d, r.limit)
                                : traverseHost<N, false>(nodes, w.recs.data(), o, d, r.limit);
            const bool same = any ? ((h.index >= 0) == (ref[i].index >= 0))
                                  : (h.index == ref[i].index &&
                                     (h.index < 0 || std::memcmp(&h, &ref[i], sizeof(RtHit)) == 0));
            bad += !same;
        }
        CHECK(bad == 0, "%s %s: %d of %zu rays differ from brute force", name, any ? "any-hit" : "closest-hit", bad,
              rays.size());
    }
}

// rt_core.h's integrator over the 4-wide BVH, rows of one call as the library's CPU backend.
static void renderHost(const std::vector<Triangle>& abiTris, const WideBvh& w, const CameraData& cam,
                       const RenderSettings& rs, int width, int height, std::vector<float>& sum) {
    std::vector<Triangle> tris = abiTris;
    std::vector<Light> lights;
    float area = 0;
    for (size_t i = 0; i < tris.size(); i++) {
        Triangle& t = tris[i];
        t.color = RtVec3{srgbToLinear(t.color.x), srgbToLinear(t.color.y), srgbToLinear(t.color.z)};
        t.specularColor = RtVec3{srgbToLinear(t.specularColor.x), srgbToLinear(t.specularColor.y), srgbToLinear(t.specularColor.z)};
        if (t.emission > 0) {
            const f3 an = crossv(V(t.edge1), V(t.edge2));
            area += .5f * sqrtf(dotv(an, an));
            lights.push_back(Light{(int)i, area});
        }
    }
    SceneView sc{tris.data(), lights.data(), (int)lights.size(), area};
    const auto* nodes = reinterpret_cast<const WideNode<4>*>(w.nodes.data());
    sum.assign((size_t)width * height * 10, 0.0f);
    for (int y = 0; y < height; y++)
        for (int x = 0; x < width; x++) {
            float* s = &sum[((size_t)y * width + x) * 10];
            for (int r = 0; r < rs.raysPerPixel; r++) {
                const unsigned pix = (unsigned)(y * width + x), smp = (unsigned)r;
                unsigned base = 0;
                auto rng = [&](int k) { return sobolSample(kDirs, pix, smp, rs.seed, base + (unsigned)k); };
                f3 o, d;
                cameraRay(cam, rs, x, y, width, height, rng, o, d);
                PathState ps; ps.thr = F3(1, 1, 1); ps.prevPdf = 0; ps.rad = F3(0, 0, 0);
                FirstHit fh; fh.covered = false; fh.albedo = fh.normal = F3(0, 0, 0); fh.depth = 0;
                for (int b = 0; b < rs.maxBounces; b++) {
                    base = 4 + 8 * (unsigned)b;
                    const RtHit hit = traverseHost<WideNode<4>, false>(nodes, w.recs.data(), o, d, 1e30f);
                    ShadeResult res;
                    shadeBounce(sc, rs, b, rng, o, d, hit, ps, &fh, res);
                    if (res.shadow && traverseHost<WideNode<4>, true>(nodes, w.recs.data(), res.nextO, res.sdir, res.slimit).index < 0)
                        ps.rad = ps.rad + res.contrib;
                    if (!res.cont) break;
                    o = res.nextO; d = res.nextD;
                }
                s[0] += ps.rad.x; s[1] += ps.rad.y; s[2] += ps.rad.z;
                s[3] += fh.albedo.x; s[4] += fh.albedo.y; s[5] += fh.albedo.z;
                s[6] += fh.normal.x; s[7] += fh.normal.y; s[8] += fh.normal.z;
                s[9] += fh.covered ? 1.0f : 0.0f;
                (void)displayByte(s[0] / (r + 1), rs.exposure);
            }
        }
}

int main() {
    const std::vector<Triangle> soup = soupScene(3000, 7), chain = chainScene();
    struct Case { const char* name; const std::vector<Triangle>* tris; float lo, hi; };
    const Case cases[] = {{"soup", &soup, -2, 2}, {"chain", &chain, -3, 40}};
    for (const Case& c : cases) {
        for (int leaf : {1, 4, 16}) {
            void* h = or_bvh_build(c.tris->data(), (int)c.tris->size(), 12, leaf);
            std::vector<RtBvhNode> bin(or_bvh_node_count(h));
            std::vector<int> order(c.tris->size());
            or_bvh_export(h, bin.data(), order.data());
            or_bvh_free(h);
            const std::vector<RtRay> rays = randomRays(6000, 11 + leaf, c.lo, c.hi);
            char name[64];
            std::snprintf(name, sizeof(name), "%s leaf<=%d W2", c.name, leaf);
            checkQueries<WideNode<2>>(name, *c.tris, bin, order, rays);
            std::snprintf(name, sizeof(name), "%s leaf<=%d W4", c.name, leaf);
            checkQueries<WideNode<4>>(name, *c.tris, bin, order, rays);
            std::snprintf(name, sizeof(name), "%s leaf<=%d W8", c.name, leaf);
            checkQueries<WideNode<8>>(name, *c.tris, bin, order, rays);
            std::snprintf(name, sizeof(name), "%s leaf<=%d Q4", c.name, leaf);
            checkQueries<QNode<4>>(name, *c.tris, bin, order, rays);
        }
    }
    // Integrator: 48x40, 4 samples, 6 bounces, thin lens, sky, against the oracle's render.
    void* h = or_bvh_build(soup.data(), (int)soup.size(), 12, 4);
    std::vector<RtBvhNode> bin(or_bvh_node_count(h));
    std::vector<int> order(soup.size());
    or_bvh_export(h, bin.data(), order.data());
    or_bvh_free(h);
    const WideBvh w4 = collapse<WideNode<4>>(soup, bin, order);
    CameraData cam{};
    cam.origin = RtVec3{0, 0, -1.9f}; cam.right = RtVec3{1, 0, 0}; cam.up = RtVec3{0, 1, 0}; cam.forward = RtVec3{0, 0, 1};
    cam.tanHalfFov = .6f; cam.aspect = 48.0f / 40.0f; cam.orthographicSize = 1; cam.orthographic = 0;
    RenderSettings rs{6, 1.0f, .02f, 2.0f, .3f, 42u, 4};
    const int width = 48, height = 40;
    std::vector<float> mine, ref((size_t)width * height * 10, 0.0f);
    renderHost(soup, w4, cam, rs, width, height, mine);
    void* sc = or_scene_create(soup.data(), (int)soup.size(), kDirs);
    unsigned long long rays[2];
    or_render(sc, &cam, &rs, width, height, 0, ref.data(), nullptr, 1, 0, height, rays);
    or_scene_free(sc);
    double rel = 0;
    int coveredDiff = 0;
    for (int i = 0; i < width * height; i++) {
        for (int k = 0; k < 3; k++) {
            const double a = mine[i * 10 + k], b = ref[i * 10 + k];
            rel += std::fabs(a - b) / std::fmax(std::fabs(b), 1e-2);
        }
        coveredDiff += mine[i * 10 + 9] != ref[i * 10 + 9];
    }
    rel /= 3.0 * width * height;
    CHECK(rel <= 1e-3, "integrator: HDR mean relative error %.3g vs the oracle", rel);
    CHECK(coveredDiff == 0, "integrator: %d covered counts differ", coveredDiff);
    std::printf("host driver: %d failures; integrator HDR mean relative error %.3g (%llu closest, %llu shadow rays)\n",
                failures, rel, rays[0], rays[1]);
    return failures ? 1 : 0;
}
