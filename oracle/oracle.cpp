// CPU oracle for the path-tracing hot path.  TEST INFRASTRUCTURE ONLY: tests/, the smoke
// check and the CPU-baseline leg of bench.py call it as the checker.  The product library
// (libCudaRayTracer.so) never links, loads or calls anything here.
//
// This is a plain restatement of the reference's algorithms, compiled with
// -ffp-contract=off so every float operation rounds like the reference build
// (Native/build.sh:11 uses --fmad=false).  Each function cites what it follows.
//
// Pinning (what checks each function against something other than itself):
//   or_hash                 - hand-computed known answers (tests/test_oracle_sampler.py)
//   or_sobol_*              - rocRAND JOEKUO6 table equality + stratification at 16/64/256
//                             samples (Tools/verify_sampling.cu:16-25) + golden PNGs
//   or_intersect32/64       - FP64 vs FP32 agreement, closed-form special cases
//   or_bvh_* / or_trace_rays- structural invariants (Tools/verify_bvh.cu:42-66) + brute-force
//                             differential (Tools/verify_bvh.cu:9-34)
//   or_sah_cost             - hand-built trees with known areas
//   or_render               - verify_native known answers (Tools/verify_native.py:74-213)
//                             and the reference's golden CUDA PNGs (Benchmarks/quality/*)
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>
#include <omp.h>
#include "../include/rt_abi.h"

namespace {

struct V { float x, y, z; };
inline V mk(float x, float y, float z) { return V{x, y, z}; }
inline V fromAbi(const RtVec3& a) { return V{a.x, a.y, a.z}; }
inline V operator+(V a, V b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V operator-(V a, V b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
inline V operator*(V a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
inline V had(V a, V b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
inline float dot3(V a, V b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V cross3(V a, V b) { return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
inline float comp(V a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }
inline void setc(V& a, int k, float v) { if (k == 0) a.x = v; else if (k == 1) a.y = v; else a.z = v; }
// Native/renderer.cu:82 - multiply by the reciprocal square root of the clamped length.
inline V normalize3(V a) { return a * (1 / sqrtf(fmaxf(dot3(a, a), 1e-20f))); }
const float kPi = 3.14159265358979323846f;

// ---------------------------------------------------------------- sampler (sampling.cuh)
// Native/sampling.cuh:21-25: the "lowbias32" integer hash.
inline unsigned lowbias(unsigned x) {
    x ^= x >> 16; x *= 0x7feb352du;
    x ^= x >> 15; x *= 0x846ca68bu;
    return x ^ (x >> 16);
}
// Native/sampling.cuh:26-35: digital shift hash(pixel ^ hash(seed + dim * golden)), then the
// XOR of the direction numbers selected by the Gray code of the sample index.
inline float sobolValue(const unsigned* dirs, unsigned pixel, unsigned sample, unsigned seed, unsigned dim) {
    unsigned bits = lowbias(pixel ^ lowbias(seed + dim * 0x9e3779b9u));
    unsigned gray = sample ^ (sample >> 1);
    for (unsigned k = 0; gray != 0; ++k, gray >>= 1)
        if (gray & 1u) bits ^= dirs[dim * 32 + k];
    return (float)(bits >> 8) * (1.0f / 16777216.0f);
}

// ---------------------------------------------------------------- geometry
// Native/renderer.cu:99-112, Moller-Trumbore with the reference's acceptance rules.
inline float mt32(V o, V d, const Triangle& tri, float& u, float& v) {
    V e1 = fromAbi(tri.edge1), e2 = fromAbi(tri.edge2);
    V p = cross3(d, e2);
    float det = dot3(e1, p);
    if (fabsf(det) < 1e-7f || (tri.oneSided && det < 0)) return -1;
    float inv = 1 / det;
    V s = o - fromAbi(tri.origin);
    u = dot3(s, p) * inv;
    if (u < 0 || u > 1) return -1;
    V q = cross3(s, e1);
    v = dot3(d, q) * inv;
    if (v < 0 || u + v > 1) return -1;
    float t = dot3(e2, q) * inv;
    return t > 1e-4f ? t : -1;
}
// Watertight ray/triangle test, restated from Woop, Benthin and Wald, "Watertight
// Ray/Triangle Intersection" (JCGT 2(1), 2013, Fig. 3-5): translate the vertices to the ray
// origin, shear and scale so the ray runs along +z (axis kz = largest |d|, kx/ky swapped when
// d[kz] < 0), 2D edge functions U, V, W of the projected triangle, reject mixed signs, t =
// (U Az + V Bz + W Cz) / det.  The edge functions use Kahan's FMA difference of products in
// FP32 (exact sign) where the paper falls back to FP64; the acceptance rules are the
// reference's: one-sided rejects det < 0, t > 1e-4 (Native/renderer.cu:99-112), only det == 0
// is parallel.  Vertices are origin, origin + edge1, origin + edge2 in FP32.
inline float diffProd(float a, float b, float c, float d) {
    float cd = c * d;
    float err = fmaf(-c, d, cd);
    return fmaf(a, b, -cd) + err;
}
inline float wt32(V o, V d, const Triangle& tri, float& u, float& v) {
    int kz = 0;
    if (fabsf(d.y) > fabsf(comp(d, kz))) kz = 1;
    if (fabsf(d.z) > fabsf(comp(d, kz))) kz = 2;
    int kx = (kz + 1) % 3, ky = (kx + 1) % 3;
    if (comp(d, kz) < 0) std::swap(kx, ky);
    float sx = comp(d, kx) / comp(d, kz), sy = comp(d, ky) / comp(d, kz), sz = 1.0f / comp(d, kz);
    V p0 = fromAbi(tri.origin);
    V a = p0 - o, b = (p0 + fromAbi(tri.edge1)) - o, c = (p0 + fromAbi(tri.edge2)) - o;
    float ax = comp(a, kx) - sx * comp(a, kz), ay = comp(a, ky) - sy * comp(a, kz);
    float bx = comp(b, kx) - sx * comp(b, kz), by = comp(b, ky) - sy * comp(b, kz);
    float cx = comp(c, kx) - sx * comp(c, kz), cy = comp(c, ky) - sy * comp(c, kz);
    float U = diffProd(cx, by, cy, bx), Vv = diffProd(ax, cy, ay, cx), W = diffProd(bx, ay, by, ax);
    if ((U < 0 || Vv < 0 || W < 0) && (U > 0 || Vv > 0 || W > 0)) return -1;
    float det = U + Vv + W;
    if (det == 0 || (tri.oneSided && det < 0)) return -1;
    float T = U * (sz * comp(a, kz)) + Vv * (sz * comp(b, kz)) + W * (sz * comp(c, kz));
    float t = T / det;
    if (!(t > 1e-4f)) return -1;
    u = Vv / det;
    v = W / det;
    return t;
}
// The same test in double precision (inputs widened exactly), for hit-distance parity.
inline double mt64(const double o[3], const double d[3], const Triangle& tri, double& u, double& v) {
    double e1[3] = {tri.edge1.x, tri.edge1.y, tri.edge1.z}, e2[3] = {tri.edge2.x, tri.edge2.y, tri.edge2.z};
    double p[3] = {d[1] * e2[2] - d[2] * e2[1], d[2] * e2[0] - d[0] * e2[2], d[0] * e2[1] - d[1] * e2[0]};
    double det = e1[0] * p[0] + e1[1] * p[1] + e1[2] * p[2];
    if (std::fabs(det) < 1e-7 || (tri.oneSided && det < 0)) return -1;
    double s[3] = {o[0] - tri.origin.x, o[1] - tri.origin.y, o[2] - tri.origin.z};
    u = (s[0] * p[0] + s[1] * p[1] + s[2] * p[2]) / det;
    if (u < 0 || u > 1) return -1;
    double q[3] = {s[1] * e1[2] - s[2] * e1[1], s[2] * e1[0] - s[0] * e1[2], s[0] * e1[1] - s[1] * e1[0]};
    v = (d[0] * q[0] + d[1] * q[1] + d[2] * q[2]) / det;
    if (v < 0 || u + v > 1) return -1;
    double t = (e2[0] * q[0] + e2[1] * q[1] + e2[2] * q[2]) / det;
    return t > 1e-4 ? t : -1;
}

// ---------------------------------------------------------------- BVH (bvh.cuh)
struct Box {
    V lo = mk(FLT_MAX, FLT_MAX, FLT_MAX), hi = mk(-FLT_MAX, -FLT_MAX, -FLT_MAX);
    void add(V a, V b) {
        lo = mk(fminf(lo.x, a.x), fminf(lo.y, a.y), fminf(lo.z, a.z));
        hi = mk(fmaxf(hi.x, b.x), fmaxf(hi.y, b.y), fmaxf(hi.z, b.z));
    }
    void add(const Box& o) { add(o.lo, o.hi); }
    // Native/bvh.cuh:20-23: half surface area of the clamped extent.
    float half() const {
        float x = fmaxf(0, hi.x - lo.x), y = fmaxf(0, hi.y - lo.y), z = fmaxf(0, hi.z - lo.z);
        return x * y + x * z + y * z;
    }
};

struct Bvh {
    std::vector<RtBvhNode> nodes;
    std::vector<int> order;
    std::vector<Box> primBox;
    std::vector<V> centre;
    int depth = 0;
    int bins = 12;
    int maxLeaf = 4;
    int depthLimit = 32;

    int binOf(float c, float lo, float scale) const {
        return std::min(bins - 1, std::max(0, (int)((c - lo) * scale)));
    }
    // Native/bvh.cuh:35-52: padded primitive boxes and vertex-mean centroids.
    void build(const Triangle* tris, int n) {
        primBox.resize(n); centre.resize(n); order.resize(n);
        std::iota(order.begin(), order.end(), 0);
        for (int i = 0; i < n; i++) {
            const Triangle& t = tris[i];
            V a = fromAbi(t.origin);
            V b = mk(t.origin.x + t.edge1.x, t.origin.y + t.edge1.y, t.origin.z + t.edge1.z);
            V c = mk(t.origin.x + t.edge2.x, t.origin.y + t.edge2.y, t.origin.z + t.edge2.z);
            Box bx; bx.add(a, a); bx.add(b, b); bx.add(c, c);
            V ct;
            for (int k = 0; k < 3; k++) {
                setc(ct, k, (comp(a, k) + comp(b, k) + comp(c, k)) / 3);
                float pad = 1e-5f + 1e-6f * fmaxf(fabsf(comp(bx.lo, k)), fabsf(comp(bx.hi, k)));
                setc(bx.lo, k, comp(bx.lo, k) - pad);
                setc(bx.hi, k, comp(bx.hi, k) + pad);
            }
            primBox[i] = bx; centre[i] = ct;
        }
        nodes.clear(); nodes.reserve((size_t)n * 2); nodes.resize(1);
        split(0, 0, n, 0);
    }
    // Native/bvh.cuh:56-105: binned SAH split, adjacent children, pre-order layout.
    void split(int ni, int first, int count, int d) {
        depth = std::max(depth, d);
        Box bounds, cb;
        for (int i = first; i < first + count; i++) { bounds.add(primBox[order[i]]); cb.add(centre[order[i]], centre[order[i]]); }
        RtBvhNode& node = nodes[ni];
        node.low[0] = bounds.lo.x; node.low[1] = bounds.lo.y; node.low[2] = bounds.lo.z;
        node.high[0] = bounds.hi.x; node.high[1] = bounds.hi.y; node.high[2] = bounds.hi.z;
        node.first = first; node.count = count;
        if (count <= maxLeaf || d == depthLimit) return;
        int bestAxis = -1, bestBin = -1;
        float best = bounds.half() * count;
        std::vector<Box> binBox(bins), suffixBox(bins);
        std::vector<int> binCount(bins), suffixCount(bins);
        for (int axis = 0; axis < 3; axis++) {
            float ext = comp(cb.hi, axis) - comp(cb.lo, axis);
            if (ext <= 1e-12f) continue;
            float scale = bins / ext, lo = comp(cb.lo, axis);
            std::fill(binBox.begin(), binBox.end(), Box());
            std::fill(binCount.begin(), binCount.end(), 0);
            for (int i = first; i < first + count; i++) {
                int b = binOf(comp(centre[order[i]], axis), lo, scale);
                binCount[b]++; binBox[b].add(primBox[order[i]]);
            }
            Box acc; int accN = 0;
            for (int b = bins - 1; b >= 0; b--) {
                if (binCount[b]) acc.add(binBox[b]);
                accN += binCount[b]; suffixBox[b] = acc; suffixCount[b] = accN;
            }
            Box left; int leftN = 0;
            for (int b = 0; b < bins - 1; b++) {
                if (binCount[b]) left.add(binBox[b]);
                leftN += binCount[b];
                if (!leftN || !suffixCount[b + 1]) continue;
                float cost = bounds.half() + left.half() * leftN + suffixBox[b + 1].half() * suffixCount[b + 1];
                if (cost < best) { best = cost; bestAxis = axis; bestBin = b; }
            }
        }
        if (bestAxis < 0) return;
        float lo = comp(cb.lo, bestAxis), scale = bins / (comp(cb.hi, bestAxis) - lo);
        auto mid = std::partition(order.begin() + first, order.begin() + first + count,
                                  [&](int i) { return binOf(comp(centre[i], bestAxis), lo, scale) <= bestBin; });
        int nl = (int)(mid - (order.begin() + first));
        if (nl == 0 || nl == count) return;
        int child = (int)nodes.size();
        nodes.resize(nodes.size() + 2);
        nodes[ni].first = child; nodes[ni].count = 0;
        split(child, first, nl, d + 1);
        split(child + 1, first + nl, count - nl, d + 1);
    }
};

// Native/renderer.cu:115-135 (host branch): slab entry distance or FLT_MAX.
inline float slab(V o, V inv, const RtBvhNode& b, float limit) {
    float nearT = 0, farT = limit;
    for (int k = 0; k < 3; k++) {
        float oc = comp(o, k), ic = comp(inv, k);
        if (ic == 0) {
            if (oc < b.low[k] || oc > b.high[k]) return FLT_MAX;
        } else {
            float a = (b.low[k] - oc) * ic, c = (b.high[k] - oc) * ic;
            if (a > c) std::swap(a, c);
            if (a > nearT) nearT = a;
            if (c < farT) farT = c;
            if (nearT > farT) return FLT_MAX;
        }
    }
    return nearT;
}

struct Counters { uint64_t nodes = 0, tris = 0; };

// Native/renderer.cu:136-178: near-first traversal, 32-entry stack, index tie rule.
RtHit traverse(const Bvh& bvh, const Triangle* tris, V o, V d, float limit, bool any, Counters* ctr = nullptr) {
    RtHit hit{-1, limit, 0, 0};
    V inv = mk(fabsf(d.x) < FLT_MIN ? 0 : 1 / d.x, fabsf(d.y) < FLT_MIN ? 0 : 1 / d.y, fabsf(d.z) < FLT_MIN ? 0 : 1 / d.z);
    const RtBvhNode* nodes = bvh.nodes.data();
    if (slab(o, inv, nodes[0], limit) == FLT_MAX) return hit;
    int stack[64], top = 0, cur = 0;
    for (;;) {
        const RtBvhNode& n = nodes[cur];
        if (ctr) ctr->nodes++;
        if (n.count > 0) {
            for (int j = n.first; j < n.first + n.count; j++) {
                int i = bvh.order[j];
                float u, v, t = mt32(o, d, tris[i], u, v);
                if (ctr) ctr->tris++;
                if (t > 0 && (t < hit.distance || (t == hit.distance && i < hit.index))) {
                    hit.index = i; hit.distance = t; hit.u = u; hit.v = v;
                    if (any) return hit;
                }
            }
        } else {
            int a = n.first, b = a + 1;
            float ta = slab(o, inv, nodes[a], hit.distance), tb = slab(o, inv, nodes[b], hit.distance);
            if (tb < ta) { std::swap(a, b); std::swap(ta, tb); }
            if (ta != FLT_MAX) {
                if (tb != FLT_MAX) stack[top++] = b;
                cur = a;
                continue;
            }
        }
        bool more = false;
        while (top > 0) {
            cur = stack[--top];
            if (slab(o, inv, nodes[cur], hit.distance) != FLT_MAX) { more = true; break; }
        }
        if (!more) break;
    }
    return hit;
}

// ---------------------------------------------------------------- integrator (renderer.cu)
struct Light { int tri; float cum; };
struct Sum { V rad, alb, nrm; unsigned covered; };

struct Scene {
    std::vector<Triangle> tris;    // colours linearised
    std::vector<Light> lights;
    float lightArea = 0;
    Bvh bvh;
    const unsigned* dirs = nullptr;
};

// Native/renderer.cu:84: sRGB decode of one channel.
inline float srgbDecode(float c) { return c <= .04045f ? c / 12.92f : powf((c + .055f) / 1.055f, 2.4f); }
// Native/renderer.cu:86-90: exposure tone map, sRGB encode, round to a byte.
inline unsigned char toByte(float c, float exposure) {
    c = 1 - expf(-fmaxf(c, 0) * exposure);
    c = c <= .0031308f ? 12.92f * c : 1.055f * powf(c, 1 / 2.4f) - .055f;
    return (unsigned char)(255 * fminf(c, 1) + .5f);
}
// Native/renderer.cu:91-95: direction at angle/cosine about an axis.
inline V rotateAbout(V axis, float cosine, float angle) {
    V helper = fabsf(axis.y) < .999f ? mk(0, 1, 0) : mk(1, 0, 0);
    V tangent = normalize3(cross3(helper, axis));
    float sine = sqrtf(fmaxf(0, 1 - cosine * cosine));
    return normalize3(axis * cosine + (tangent * (sine * cosf(angle)) + cross3(axis, tangent) * (sine * sinf(angle))));
}
inline float power2(float a, float b) { return a * a / fmaxf(a * a + b * b, 1e-30f); }
// Native/renderer.cu:181: Phong exponent from smoothness.
inline float phongExp(const Triangle& t) { float r = fmaxf(1 - t.smoothness, .015f); return fmaxf(0, 2 / (r * r) - 2); }
// Native/renderer.cu:182-192: diffuse + normalised Phong value and pdf.
inline V brdf(const Triangle& t, V n, V refl, V dir, float& pdf) {
    float c = fmaxf(0, dot3(n, dir)), p = t.specularProbability;
    pdf = (1 - p) * c / kPi;
    V val = fromAbi(t.color) * ((1 - p) / kPi);
    if (p > 0 && t.smoothness < .999f) {
        float e = phongExp(t), lobe = powf(fmaxf(0, dot3(refl, dir)), e);
        pdf += p * (e + 1) * lobe / (2 * kPi);
        val = val + fromAbi(t.specularColor) * (p * (e + 2) * lobe / (2 * kPi));
    }
    return val;
}
// Native/renderer.cu:193-196: vertical sky gradient.
inline V sky(V d, float strength) {
    float t = fminf(1, fmaxf(0, d.y));
    return (mk(.65f, .72f, .82f) * (1 - t) + mk(.15f, .32f, .65f) * t) * strength;
}

struct RayTally { uint64_t closest = 0, shadow = 0; };

// Native/renderer.cu:197-288: one path.
Sum tracePath(const Scene& sc, int x, int y, const CameraData& cam, const RenderSettings& rs, int w, int h,
              unsigned sample, RayTally& tally) {
    unsigned pix = (unsigned)(y * w + x);
    unsigned base = 0;
    auto rnd = [&](unsigned dim) { return sobolValue(sc.dirs, pix, sample, rs.seed, dim + base); };
    float px = 2 * (x + rnd(0)) / w - 1, py = 2 * (y + rnd(1)) / h - 1;
    V right = fromAbi(cam.right), up = fromAbi(cam.up), fwd = fromAbi(cam.forward);
    V o = fromAbi(cam.origin);
    V d = normalize3(fwd + (right * (px * cam.aspect * cam.tanHalfFov) + up * (py * cam.tanHalfFov)));
    if (cam.orthographic) {
        o = o + (right * (px * cam.aspect * cam.orthographicSize) + up * (py * cam.orthographicSize));
        d = fwd;
    } else if (rs.aperture > 0) {
        float ang = 2 * kPi * rnd(2), rad = .5f * rs.aperture * sqrtf(rnd(3));
        V lens = right * (rad * cosf(ang)) + up * (rad * sinf(ang));
        V focus = o + d * (rs.focusDistance / dot3(d, fwd));
        o = o + lens; d = normalize3(focus - o);
    }
    Sum out{};
    V thr = mk(1, 1, 1);
    float prevPdf = 0;
    for (int b = 0; b < rs.maxBounces; b++) {
        base = 4 + 8 * b;
        RtHit hit = traverse(sc.bvh, sc.tris.data(), o, d, 1e30f, false);
        tally.closest++;
        if (hit.index < 0) {
            if (b > 0) out.rad = out.rad + had(thr, sky(d, rs.skyStrength));
            break;
        }
        const Triangle& t = sc.tris[hit.index];
        V g = normalize3(cross3(fromAbi(t.edge1), fromAbi(t.edge2)));
        if (dot3(g, d) > 0) g = g * -1;
        if (b == 0) {
            out.covered = 1;
            out.alb = t.emission > 0 ? mk(1, 1, 1)
                                     : fromAbi(t.color) * (1 - t.specularProbability) + fromAbi(t.specularColor) * t.specularProbability;
            out.nrm = g;
        }
        if (t.emission > 0) {
            float lp = sc.lightArea > 0 ? hit.distance * hit.distance / (sc.lightArea * fmaxf(1e-6f, -dot3(g, d))) : 0;
            float wgt = prevPdf > 0 ? power2(prevPdf, lp) : 1;
            out.rad = out.rad + had(thr, fromAbi(t.color)) * (t.emission * wgt);
            break;
        }
        V n = normalize3(fromAbi(t.normal0) * (1 - hit.u - hit.v) + (fromAbi(t.normal1) * hit.u + fromAbi(t.normal2) * hit.v));
        if (dot3(n, g) < 0) n = n * -1;
        if (dot3(n, d) >= 0) n = g;
        if (b == 0) out.nrm = n;
        V p = o + d * hit.distance;
        V refl = d - n * (2 * dot3(d, n));
        o = p + g * .0005f;
        if (!sc.lights.empty()) {
            float sel = rnd(0) * sc.lightArea;
            int l = 0, nl = (int)sc.lights.size();
            while (l < nl - 1 && sel >= sc.lights[l].cum) l++;
            const Triangle& L = sc.tris[sc.lights[l].tri];
            float r = sqrtf(rnd(1)), vv = rnd(2);
            V target = fromAbi(L.origin) + (fromAbi(L.edge1) * (r * (1 - vv)) + fromAbi(L.edge2) * (r * vv));
            V off = target - o;
            float dist2 = dot3(off, off), dist = sqrtf(dist2);
            V in = off * (1 / dist), ln = normalize3(cross3(fromAbi(L.edge1), fromAbi(L.edge2)));
            float lc = -dot3(ln, in);
            if (!L.oneSided) lc = fabsf(lc);
            float sc2 = fmaxf(0, dot3(n, in));
            if (lc > 0 && sc2 > 0 && dot3(g, in) > 0) {
                tally.shadow++;
                if (traverse(sc.bvh, sc.tris.data(), o, in, dist - .001f, true).index < 0) {
                    float pdf, lp = dist2 / (sc.lightArea * lc);
                    V f = brdf(t, n, refl, in, pdf);
                    float wgt = b + 1 == rs.maxBounces ? 1 : power2(lp, pdf);
                    out.rad = out.rad + had(had(thr, f), fromAbi(L.color)) * (L.emission * sc2 * wgt / lp);
                }
            }
        }
        if (b + 1 == rs.maxBounces) break;
        bool spec = rnd(3) < t.specularProbability;
        if (spec && t.smoothness >= .999f) {
            d = refl; thr = had(thr, fromAbi(t.specularColor)); prevPdf = 0;
        } else {
            float cosine = spec ? powf(rnd(4), 1 / (phongExp(t) + 1)) : sqrtf(rnd(4));
            d = rotateAbout(spec ? refl : n, cosine, 2 * kPi * rnd(5));
            if (dot3(g, d) <= 0 || dot3(n, d) <= 0) break;
            float pdf;
            V f = brdf(t, n, refl, d, pdf);
            if (pdf < 1e-12f) break;
            thr = had(thr, f * (dot3(n, d) / pdf)); prevPdf = pdf;
        }
        if (b >= 3) {
            float surv = fminf(.95f, fmaxf(thr.x, fmaxf(thr.y, thr.z)));
            if (rnd(6) >= surv) break;
            thr = thr * (1 / surv);
        }
    }
    return out;
}

}  // namespace

// ==================================================================== C interface
extern "C" {

unsigned or_hash(unsigned x) { return lowbias(x); }

void or_sobol_batch(const unsigned* dirs, const unsigned* pixels, const unsigned* samples, int count,
                    unsigned seed, int dims, float* out) {
    for (int i = 0; i < count; i++)
        for (int d = 0; d < dims; d++) out[(size_t)i * dims + d] = sobolValue(dirs, pixels[i], samples[i], seed, (unsigned)d);
}

float or_intersect32(const Triangle* t, const float* o, const float* d, float* u, float* v) {
    float uu = 0, vv = 0;
    float r = mt32(mk(o[0], o[1], o[2]), mk(d[0], d[1], d[2]), *t, uu, vv);
    *u = uu; *v = vv; return r;
}
double or_intersect64(const Triangle* t, const double* o, const double* d, double* u, double* v) {
    double uu = 0, vv = 0;
    double r = mt64(o, d, *t, uu, vv);
    *u = uu; *v = vv; return r;
}

// Brute-force loops (Tools/verify_bvh.cu:9-19): lowest index wins exact ties.
// Brute force with the watertight test; ties go to the lower index as Native/renderer.cu:152-153.
void or_brute_wt32(const Triangle* tris, int n, const RtRay* rays, int count, int any, RtHit* hits) {
    #pragma omp parallel for schedule(dynamic, 64)
    for (int r = 0; r < count; r++) {
        V o = mk(rays[r].origin[0], rays[r].origin[1], rays[r].origin[2]);
        V d = mk(rays[r].direction[0], rays[r].direction[1], rays[r].direction[2]);
        RtHit h{-1, rays[r].limit, 0, 0};
        for (int i = 0; i < n; i++) {
            float u, v, t = wt32(o, d, tris[i], u, v);
            if (t > 0 && t < h.distance) { h.index = i; h.distance = t; h.u = u; h.v = v; if (any) break; }
        }
        hits[r] = h;
    }
}
float or_intersect_wt32(const Triangle* tri, const float* o, const float* d, float* u, float* v) {
    return wt32(mk(o[0], o[1], o[2]), mk(d[0], d[1], d[2]), *tri, *u, *v);
}
void or_brute32(const Triangle* tris, int n, const RtRay* rays, int count, int any, RtHit* hits) {
    #pragma omp parallel for schedule(dynamic, 64)
    for (int r = 0; r < count; r++) {
        V o = mk(rays[r].origin[0], rays[r].origin[1], rays[r].origin[2]);
        V d = mk(rays[r].direction[0], rays[r].direction[1], rays[r].direction[2]);
        RtHit h{-1, rays[r].limit, 0, 0};
        for (int i = 0; i < n; i++) {
            float u, v, t = mt32(o, d, tris[i], u, v);
            if (t > 0 && t < h.distance) { h.index = i; h.distance = t; h.u = u; h.v = v; if (any) break; }
        }
        hits[r] = h;
    }
}
// FP64 brute force: index and distance of the nearest double-precision hit below the limit.
void or_brute64(const Triangle* tris, int n, const RtRay* rays, int count, int* index, double* dist) {
    #pragma omp parallel for schedule(dynamic, 64)
    for (int r = 0; r < count; r++) {
        double o[3] = {rays[r].origin[0], rays[r].origin[1], rays[r].origin[2]};
        double d[3] = {rays[r].direction[0], rays[r].direction[1], rays[r].direction[2]};
        int best = -1; double bt = rays[r].limit;
        for (int i = 0; i < n; i++) {
            double u, v, t = mt64(o, d, tris[i], u, v);
            if (t > 0 && t < bt) { best = i; bt = t; }
        }
        index[r] = best; dist[r] = bt;
    }
}

void* or_bvh_build(const Triangle* tris, int n, int bins, int maxLeaf) {
    Bvh* b = new Bvh();
    b->bins = bins; b->maxLeaf = maxLeaf;
    b->build(tris, n);
    return b;
}
void or_bvh_free(void* h) { delete static_cast<Bvh*>(h); }
int or_bvh_node_count(void* h) { return (int)static_cast<Bvh*>(h)->nodes.size(); }
int or_bvh_depth(void* h) { return static_cast<Bvh*>(h)->depth; }
void or_bvh_export(void* h, RtBvhNode* nodes, int* order) {
    Bvh* b = static_cast<Bvh*>(h);
    std::memcpy(nodes, b->nodes.data(), b->nodes.size() * sizeof(RtBvhNode));
    std::memcpy(order, b->order.data(), b->order.size() * sizeof(int));
}
// Import an externally built binary BVH (e.g. the GPU build) for traversal checks.
void* or_bvh_import(const RtBvhNode* nodes, int nodeCount, const int* order, int n) {
    Bvh* b = new Bvh();
    b->nodes.assign(nodes, nodes + nodeCount);
    b->order.assign(order, order + n);
    return b;
}
void or_trace_rays(void* h, const Triangle* tris, const RtRay* rays, int count, int any, RtHit* hits,
                   unsigned long long* nodeVisits, unsigned long long* triTests) {
    const Bvh& b = *static_cast<Bvh*>(h);
    unsigned long long nv = 0, nt = 0;
    #pragma omp parallel for schedule(dynamic, 64) reduction(+ : nv, nt)
    for (int r = 0; r < count; r++) {
        Counters c;
        hits[r] = traverse(b, tris, mk(rays[r].origin[0], rays[r].origin[1], rays[r].origin[2]),
                           mk(rays[r].direction[0], rays[r].direction[1], rays[r].direction[2]), rays[r].limit, any != 0, &c);
        nv += c.nodes; nt += c.tris;
    }
    if (nodeVisits) *nodeVisits = nv;
    if (triTests) *triTests = nt;
}

// SAH cost of a binary tree in the reference layout, normalised by the root area, with
// traversal and intersection constants 1 and 1 (Native/bvh.cuh:66,89).
double or_sah_cost(const RtBvhNode* nodes, int count) {
    auto half = [](const RtBvhNode& n) {
        double x = std::max(0.0, (double)n.high[0] - n.low[0]), y = std::max(0.0, (double)n.high[1] - n.low[1]),
               z = std::max(0.0, (double)n.high[2] - n.low[2]);
        return x * y + x * z + y * z;
    };
    double root = half(nodes[0]), cost = 0;
    if (root <= 0) return 0;
    for (int i = 0; i < count; i++) cost += nodes[i].count > 0 ? half(nodes[i]) * nodes[i].count / root : half(nodes[i]) / root;
    return cost;
}

// Native/renderer.cu:340-357: copy, linearise colours, light CDF, 12-bin BVH.
void* or_scene_create(const Triangle* tris, int n, const unsigned* dirs) {
    Scene* s = new Scene();
    s->tris.assign(tris, tris + n);
    s->dirs = dirs;
    for (int i = 0; i < n; i++) {
        Triangle& t = s->tris[i];
        t.color = RtVec3{srgbDecode(t.color.x), srgbDecode(t.color.y), srgbDecode(t.color.z)};
        t.specularColor = RtVec3{srgbDecode(t.specularColor.x), srgbDecode(t.specularColor.y), srgbDecode(t.specularColor.z)};
        if (t.emission > 0) {
            V an = cross3(fromAbi(t.edge1), fromAbi(t.edge2));
            s->lightArea += .5f * sqrtf(dot3(an, an));
            s->lights.push_back({i, s->lightArea});
        }
    }
    s->bvh.build(s->tris.data(), n);
    return s;
}
void or_scene_free(void* h) { delete static_cast<Scene*>(h); }
// Ray queries against the scene's own 12-bin BVH (the one or_render traverses).
void or_scene_trace_rays(void* h, const RtRay* rays, int count, int any, RtHit* hits) {
    const Scene& sc = *static_cast<Scene*>(h);
    or_trace_rays((void*)&sc.bvh, sc.tris.data(), rays, count, any, hits, nullptr, nullptr);
}
int or_scene_bvh_nodes(void* h) { return (int)static_cast<Scene*>(h)->bvh.nodes.size(); }
// SAH cost (or_sah_cost) of the scene's own 12-bin reference build.
double or_scene_sah_cost(void* h) {
    const Scene& sc = *static_cast<Scene*>(h);
    return or_sah_cost(sc.bvh.nodes.data(), (int)sc.bvh.nodes.size());
}
int or_scene_bvh_depth(void* h) { return static_cast<Scene*>(h)->bvh.depth; }
float or_scene_light_area(void* h) { return static_cast<Scene*>(h)->lightArea; }

// Native/renderer.cu:289-302 + 405-416: rows [y0, y1) of one RtRender call on the CPU.
// sum holds 10 floats per pixel (radiance, albedo, normal, covered); sample = samples so
// far; rgba may be null.  threads <= 1 runs single-threaded.  Returns ray counts.
void or_render(void* h, const CameraData* cam, const RenderSettings* rs, int w, int hgt, unsigned sample,
               float* sum, unsigned char* rgba, int threads, int y0, int y1, unsigned long long* rays) {
    const Scene& sc = *static_cast<Scene*>(h);
    if (y1 > hgt) y1 = hgt;
    unsigned long long closest = 0, shadow = 0;
    int nt = threads < 1 ? 1 : threads;
    #pragma omp parallel for schedule(static) num_threads(nt) reduction(+ : closest, shadow) if (nt > 1)
    for (int y = y0; y < y1; y++) {
        RayTally tally;
        for (int x = 0; x < w; x++) {
            size_t i = (size_t)y * w + x;
            float* s = sum + i * 10;
            Sum tot{};
            if (sample) {
                tot.rad = mk(s[0], s[1], s[2]); tot.alb = mk(s[3], s[4], s[5]); tot.nrm = mk(s[6], s[7], s[8]);
                tot.covered = (unsigned)s[9];
            }
            for (int r = 0; r < rs->raysPerPixel; r++) {
                Sum p = tracePath(sc, x, y, *cam, *rs, w, hgt, sample + r, tally);
                tot.rad = tot.rad + p.rad; tot.alb = tot.alb + p.alb; tot.nrm = tot.nrm + p.nrm;
                tot.covered += p.covered;
            }
            s[0] = tot.rad.x; s[1] = tot.rad.y; s[2] = tot.rad.z;
            s[3] = tot.alb.x; s[4] = tot.alb.y; s[5] = tot.alb.z;
            s[6] = tot.nrm.x; s[7] = tot.nrm.y; s[8] = tot.nrm.z; s[9] = (float)tot.covered;
            if (rgba) {
                V avg = tot.rad * (1.0f / (sample + rs->raysPerPixel));
                unsigned char* px = rgba + i * 4;
                px[0] = toByte(avg.x, rs->exposure); px[1] = toByte(avg.y, rs->exposure);
                px[2] = toByte(avg.z, rs->exposure); px[3] = 255;
            }
        }
        closest += tally.closest; shadow += tally.shadow;
    }
    if (rays) { rays[0] = closest; rays[1] = shadow; }
}

unsigned char or_display(float c, float exposure) { return toByte(c, exposure); }
float or_srgb_decode(float c) { return srgbDecode(c); }

}  // extern "C"
