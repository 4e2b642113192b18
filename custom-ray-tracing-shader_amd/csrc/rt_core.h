// rt_core.h - arithmetic shared by the gfx950 kernels and the host (CPU backend) code.
//
// Every function here is written so that each float operation is the same IEEE operation,
// in the same order, as the reference integrator (Native/renderer.cu:74-302); the library
// is compiled with -ffp-contract=off (the reference uses --fmad=false, Native/build.sh:11),
// correctly rounded fp32 division/sqrt and denormals kept, so the GPU and the CPU paths
// produce identical hit records and differ only where libm transcendentals (powf, expf,
// sinf, cosf) differ by ulps.
#pragma once
#include <hip/hip_runtime.h>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include "../../include/rt_abi.h"

#define RT_HD __host__ __device__ __forceinline__

namespace rt {

constexpr float kPi = 3.14159265358979323846f;
constexpr int kSobolDims = 4 + 8 * 32;   // Native/sampling.cuh:6
constexpr int kMaxBounces = 32;           // RtRender bound (Native/renderer.cu:379)
constexpr int kMaxRaysPerPixel = 16;      // RtRender bound (Native/renderer.cu:379)
constexpr int kBvhDepthLimit = 32;        // forced-leaf depth (Native/bvh.cuh:9,64)

struct f3 { float x, y, z; };
RT_HD f3 F3(float x, float y, float z) { f3 r; r.x = x; r.y = y; r.z = z; return r; }
RT_HD f3 V(const RtVec3& a) { return F3(a.x, a.y, a.z); }
RT_HD f3 operator+(f3 a, f3 b) { return F3(a.x + b.x, a.y + b.y, a.z + b.z); }
RT_HD f3 operator-(f3 a, f3 b) { return F3(a.x - b.x, a.y - b.y, a.z - b.z); }
RT_HD f3 operator*(f3 a, float s) { return F3(a.x * s, a.y * s, a.z * s); }
RT_HD f3 mulv(f3 a, f3 b) { return F3(a.x * b.x, a.y * b.y, a.z * b.z); }
RT_HD float dotv(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
RT_HD f3 crossv(f3 a, f3 b) { return F3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
// Native/renderer.cu:82
RT_HD f3 unitv(f3 a) { return a * (1 / sqrtf(fmaxf(dotv(a, a), 1e-20f))); }

// ------------------------------------------------------------------ Sobol sampler
// Native/sampling.cuh:21-25 (integer hash) and :26-35 (digital-shifted Gray-code Sobol).
RT_HD unsigned hash32(unsigned x) {
    x ^= x >> 16; x *= 0x7feb352du;
    x ^= x >> 15; x *= 0x846ca68bu;
    return x ^ (x >> 16);
}
// Per-(seed, dimension) part of the digital shift; uniform across a launch.
RT_HD unsigned dimShift(unsigned seed, unsigned dim) { return hash32(seed + dim * 0x9e3779b9u); }
// XOR of the direction numbers picked by gray(sample); uniform across pixels.
RT_HD unsigned grayTerm(const unsigned* dirs, unsigned dim, unsigned sample) {
    unsigned acc = 0, g = sample ^ (sample >> 1);
    for (unsigned k = 0; g != 0; ++k, g >>= 1)
        if (g & 1u) acc ^= dirs[dim * 32 + k];
    return acc;
}
// XOR is associative, so hash(pixel ^ shift) ^ term equals the reference's incremental XOR.
RT_HD float sobolFromParts(unsigned pixel, unsigned shift, unsigned term) {
    return (float)((hash32(pixel ^ shift) ^ term) >> 8) * (1.0f / 16777216.0f);
}
RT_HD float sobolSample(const unsigned* dirs, unsigned pixel, unsigned sample, unsigned seed, unsigned dim) {
    return sobolFromParts(pixel, dimShift(seed, dim), grayTerm(dirs, dim, sample));
}

// ------------------------------------------------------------------ geometry
// Native/renderer.cu:99-112: Moller-Trumbore with the reference's acceptance rules.
RT_HD float intersectTri(f3 o, f3 d, f3 v0, f3 e1, f3 e2, bool oneSided, float& u, float& v) {
    f3 p = crossv(d, e2);
    float det = dotv(e1, p);
    if (fabsf(det) < 1e-7f || (oneSided && det < 0)) return -1;
    float inv = 1 / det;
    f3 s = o - v0;
    u = dotv(s, p) * inv;
    if (u < 0 || u > 1) return -1;
    f3 q = crossv(s, e1);
    v = dotv(d, q) * inv;
    if (v < 0 || u + v > 1) return -1;
    float t = dotv(e2, q) * inv;
    return t > 1e-4f ? t : -1;
}
// a*b - c*d with the rounding error of c*d compensated by fused multiply-adds (Kahan's
// difference of products, as pbrt-v4's DifferenceOfProducts): within 1.5 ulp of the exact
// value, so its sign is the exact value's sign whenever that is not zero.
RT_HD float diffProducts(float a, float b, float c, float d) {
    const float cd = c * d;
    const float err = fmaf(-c, d, cd);
    return fmaf(a, b, -cd) + err;
}
// Watertight ray/triangle test (Woop, Benthin and Wald, "Watertight Ray/Triangle
// Intersection", JCGT 2(1), 2013): vertices translated to the ray origin, sheared and scaled so
// the ray runs along +z, then the three 2D edge functions of the projected triangle.  An edge
// shared by two triangles gets the same edge function up to sign in both (the ray's shear is
// the same), so a ray through the edge cannot miss both: no gaps between triangles that share
// vertices.  The paper's double-precision fallback for zero edge functions is replaced by
// diffProducts, whose sign is exact.  Acceptance keeps the reference's other rules
// (Native/renderer.cu:99-112): one-sided triangles reject det < 0 (det has the sign of the
// Moller-Trumbore determinant), t > 1e-4; only an exactly zero det counts as parallel.
// Vertices are v0, v0 + e1, v0 + e2 in FP32, as the ABI's triangles define them.  u, v are the
// barycentric weights of v1 and v2.
RT_HD float intersectTriWatertight(f3 o, f3 d, f3 v0, f3 e1, f3 e2, bool oneSided, float& u, float& v) {
    const float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    const int kz = (ax >= ay && ax >= az) ? 0 : (ay >= az ? 1 : 2);
    int kx = kz == 2 ? 0 : kz + 1, ky = kx == 2 ? 0 : kx + 1;
    auto comp = [](f3 q, int k) { return k == 0 ? q.x : (k == 1 ? q.y : q.z); };
    const float dz = comp(d, kz);
    if (dz < 0) { const int t = kx; kx = ky; ky = t; }
    const float sx = comp(d, kx) / dz, sy = comp(d, ky) / dz, sz = 1.0f / dz;
    const f3 A = v0 - o, B = (v0 + e1) - o, C = (v0 + e2) - o;
    const float Ax = comp(A, kx) - sx * comp(A, kz), Ay = comp(A, ky) - sy * comp(A, kz);
    const float Bx = comp(B, kx) - sx * comp(B, kz), By = comp(B, ky) - sy * comp(B, kz);
    const float Cx = comp(C, kx) - sx * comp(C, kz), Cy = comp(C, ky) - sy * comp(C, kz);
    const float U = diffProducts(Cx, By, Cy, Bx), V = diffProducts(Ax, Cy, Ay, Cx), W = diffProducts(Bx, Ay, By, Ax);
    if ((U < 0 || V < 0 || W < 0) && (U > 0 || V > 0 || W > 0)) return -1;
    const float det = U + V + W;
    if (det == 0 || (oneSided && det < 0)) return -1;
    const float T = U * (sz * comp(A, kz)) + V * (sz * comp(B, kz)) + W * (sz * comp(C, kz));
    const float t = T / det;
    if (!(t > 1e-4f)) return -1;
    u = V / det;
    v = W / det;
    return t;
}
// The reference's test or the watertight one.
RT_HD float intersectTriMode(bool watertight, f3 o, f3 d, f3 v0, f3 e1, f3 e2, bool oneSided, float& u, float& v) {
    return watertight ? intersectTriWatertight(o, d, v0, e1, e2, oneSided, u, v) : intersectTri(o, d, v0, e1, e2, oneSided, u, v);
}
// Native/renderer.cu:142-143: components below FLT_MIN become a containment test.
RT_HD f3 safeInverse(f3 d) {
    return F3(fabsf(d.x) < FLT_MIN ? 0 : 1 / d.x, fabsf(d.y) < FLT_MIN ? 0 : 1 / d.y, fabsf(d.z) < FLT_MIN ? 0 : 1 / d.z);
}
// Slab interval of one axis; an inverse of 0 turns into [-inf, inf] inside the slab and an
// empty interval outside it (Native/renderer.cu:120).
RT_HD void slabAxis(float o, float inv, float lo, float hi, float& tmin, float& tmax) {
    float a = inv != 0 ? (lo - o) * inv : (o < lo ? INFINITY : -INFINITY);
    float b = inv != 0 ? (hi - o) * inv : (o > hi ? -INFINITY : INFINITY);
    tmin = fminf(a, b); tmax = fmaxf(a, b);
}
// Entry distance of the box for a ray limited to [0, limit], or +inf on a miss.
RT_HD float boxEntry(f3 o, f3 inv, float lx, float ly, float lz, float hx, float hy, float hz, float limit) {
    float ax, bx, ay, by, az, bz;
    slabAxis(o.x, inv.x, lx, hx, ax, bx);
    slabAxis(o.y, inv.y, ly, hy, ay, by);
    slabAxis(o.z, inv.z, lz, hz, az, bz);
    float nearT = fmaxf(fmaxf(ax, ay), fmaxf(az, 0.0f));
    float farT = fminf(fminf(bx, by), fminf(bz, limit));
    return nearT <= farT ? nearT : INFINITY;
}
// Slab test for a ray whose inverse direction has no zero component (the slab selects of
// boxEntry drop out), with each slab distance as one fused multiply-add, lo * inv - o * inv,
// where oi = o * inv is computed once per ray.  Its rounding error, like that of
// (lo - o) * inv, is a few units in the last place of max(|o|, |lo|) in position, inside the
// padding every primitive box carries, so the boxes it enters still contain every hit that
// can win.
RT_HD float boxEntryFma(f3 oi, f3 inv, float lx, float ly, float lz, float hx, float hy, float hz, float limit) {
    const float ax = fmaf(lx, inv.x, -oi.x), bx = fmaf(hx, inv.x, -oi.x);
    const float ay = fmaf(ly, inv.y, -oi.y), by = fmaf(hy, inv.y, -oi.y);
    const float az = fmaf(lz, inv.z, -oi.z), bz = fmaf(hz, inv.z, -oi.z);
    float nearT = fmaxf(fmaxf(fminf(ax, bx), fminf(ay, by)), fmaxf(fminf(az, bz), 0.0f));
    float farT = fminf(fminf(fmaxf(ax, bx), fmaxf(ay, by)), fminf(fmaxf(az, bz), limit));
    return nearT <= farT ? nearT : INFINITY;
}

// ------------------------------------------------------------------ wide BVH layout
// Child references: >= 0 is an inner node index; < 0 is a leaf, ~ref = first record << 3 |
// code, where code = count - 1 for leaves of 1..7 triangles and 7 for larger ones (whose end
// is found from the last-in-leaf record flag).  kEmptyChild marks an unused slot.
constexpr int kEmptyChild = INT32_MIN;
constexpr int kLeafFirstLimit = 1 << 28;   // records addressable by a leaf reference
RT_HD int makeLeafRef(int first, int count) { return ~((first << 3) | (count < 8 ? count - 1 : 7)); }
RT_HD int leafFirst(int ref) { return (~ref) >> 3; }
// Records certainly in the leaf (8 means "8 or more").
RT_HD int leafBatch(int ref) { return ((~ref) & 7) + 1; }

RT_HD float bitsToFloat(unsigned u) { union { unsigned u; float f; } c; c.u = u; return c.f; }

// Full-precision W-wide node, W*32 bytes (64 B for W=2), 64-byte aligned: six W-float planes
// of child bounds, then W child references and W triangle counts.  Unused slots have +inf
// bounds, so every slab test misses them.
template <int W> struct alignas(64) WideNode {
    float lox[W], hix[W], loy[W], hiy[W], loz[W], hiz[W];
    int child[W];
    int count[W];
    RT_HD float entry(int c, f3 o, f3 inv, float limit) const {
        return boxEntry(o, inv, lox[c], loy[c], loz[c], hix[c], hiy[c], hiz[c], limit);
    }
    RT_HD float entryFma(int c, f3 oi, f3 inv, float limit) const {
        return boxEntryFma(oi, inv, lox[c], loy[c], loz[c], hix[c], hiy[c], hiz[c], limit);
    }
};
static_assert(sizeof(WideNode<2>) == 64 && sizeof(WideNode<4>) == 128 && sizeof(WideNode<8>) == 256, "node sizes");

// Quantized W-wide node: child bounds as 8-bit offsets from the node's own box corner in
// per-axis power-of-two steps, rounded outwards, so every decoded child box contains the
// exact one (the build checks each with the decode arithmetic below).  A looser box can only
// add node visits, never drop a triangle whose hit could win, so hit records are unchanged.
// 56 B (W=4) or 96 B (W=8) are read per visit instead of 128 B or 256 B.
template <int W> struct alignas(64) QNode {
    float px, py, pz;        // corner: minimum of the child boxes
    unsigned steps;          // biased FP32 exponents of the x, y, z steps in bytes 0, 1, 2
    int child[W];
    unsigned char qlox[W], qhix[W], qloy[W], qhiy[W], qloz[W], qhiz[W];
    RT_HD float step(int axis) const { return bitsToFloat(((steps >> (8 * axis)) & 255u) << 23); }
    RT_HD float entry(int c, f3 o, f3 inv, float limit) const {
        if (child[c] == kEmptyChild) return INFINITY;
        const float sx = step(0), sy = step(1), sz = step(2);
        return boxEntry(o, inv, px + (float)qlox[c] * sx, py + (float)qloy[c] * sy, pz + (float)qloz[c] * sz,
                        px + (float)qhix[c] * sx, py + (float)qhiy[c] * sy, pz + (float)qhiz[c] * sz, limit);
    }
    RT_HD float entryFma(int c, f3 oi, f3 inv, float limit) const {
        if (child[c] == kEmptyChild) return INFINITY;
        const float sx = step(0), sy = step(1), sz = step(2);
        return boxEntryFma(oi, inv, px + (float)qlox[c] * sx, py + (float)qloy[c] * sy, pz + (float)qloz[c] * sz,
                           px + (float)qhix[c] * sx, py + (float)qhiy[c] * sy, pz + (float)qhiz[c] * sz, limit);
    }
};
static_assert(sizeof(QNode<2>) == 64 && sizeof(QNode<4>) == 64 && sizeof(QNode<8>) == 128, "quantized node sizes");

// Bytes of a node a traversal step reads (16-byte vectors; unused tail padding skipped).
template <class N> struct NodeTraits;
template <int W> struct NodeTraits<WideNode<W>> { static constexpr int kWidth = W, kVectors = 2 * W; };
template <int W> struct NodeTraits<QNode<W>> { static constexpr int kWidth = W, kVectors = (16 + 10 * W + 15) / 16; };

// Quantizes child boxes lo/hi[c][axis] (c < nk) and references into a QNode.  Exponents
// are clamped to [27, 254] (steps 2^-100 .. 2^127), where q * step is exact for q <= 255.
RT_HD float quantDecode(float corner, float step, int q) { return corner + (float)q * step; }
template <int W>
RT_HD void quantizeNode(QNode<W>& n, const float (*lo)[3], const float (*hi)[3], const int* refs, int nk) {
    float corner[3], step[3];
    unsigned steps = 0;
    for (int a = 0; a < 3; a++) {
        float l = INFINITY, h = -INFINITY;
        for (int c = 0; c < nk; c++) { l = fminf(l, lo[c][a]); h = fmaxf(h, hi[c][a]); }
        if (!(l <= h)) { l = 0; h = 0; }
        const float ext = (h - l) / 255.0f;
        union { float f; unsigned u; } e; e.f = ext;
        int be = (int)((e.u >> 23) & 255u) + 1;
        be = be < 27 ? 27 : (be > 254 ? 254 : be);
        while (be > 27 && quantDecode(l, bitsToFloat((unsigned)(be - 1) << 23), 255) >= h) be--;
        while (be < 254 && !(quantDecode(l, bitsToFloat((unsigned)be << 23), 255) >= h)) be++;
        corner[a] = l;
        step[a] = bitsToFloat((unsigned)be << 23);
        steps |= (unsigned)be << (8 * a);
    }
    n.px = corner[0]; n.py = corner[1]; n.pz = corner[2]; n.steps = steps;
    unsigned char* qlo[3] = {n.qlox, n.qloy, n.qloz};
    unsigned char* qhi[3] = {n.qhix, n.qhiy, n.qhiz};
    for (int c = 0; c < W; c++) {
        n.child[c] = c < nk ? refs[c] : kEmptyChild;
        for (int a = 0; a < 3; a++) {
            if (c >= nk) { qlo[a][c] = 255; qhi[a][c] = 0; continue; }
            const float l = corner[a], s = step[a];
            int ql = (int)fminf(fmaxf(floorf((lo[c][a] - l) / s), 0.0f), 255.0f);
            while (ql > 0 && quantDecode(l, s, ql) > lo[c][a]) ql--;
            while (ql < 255 && quantDecode(l, s, ql + 1) <= lo[c][a]) ql++;
            int qh = (int)fminf(fmaxf(ceilf((hi[c][a] - l) / s), 0.0f), 255.0f);
            while (qh < 255 && quantDecode(l, s, qh) < hi[c][a]) qh++;
            while (qh > 0 && quantDecode(l, s, qh - 1) >= hi[c][a]) qh--;
            qlo[a][c] = (unsigned char)ql;
            qhi[a][c] = (unsigned char)qh;
        }
    }
}

// Intersection record, 48 bytes: origin/edge1/edge2 exactly as the caller passed them (so the
// Moller-Trumbore arithmetic is the reference's), original triangle index and flags.
constexpr int kTriOneSided = 1, kTriLastInLeaf = 2;
struct alignas(16) TriRecord {
    float v0x, v0y, v0z; int index;
    float e1x, e1y, e1z; int flags;
    float e2x, e2y, e2z; int pad;
};
static_assert(sizeof(TriRecord) == 48, "tri record");

// ------------------------------------------------------------------ integrator pieces
struct Light { int triangle; float cumulativeArea; };   // Native/renderer.cu:32

struct SceneView {
    const Triangle* tris;   // colours already linear
    const Light* lights;
    int lightCount;
    float lightArea;
};

// Native/renderer.cu:84,86-90
RT_HD float srgbToLinear(float c) { return c <= .04045f ? c / 12.92f : powf((c + .055f) / 1.055f, 2.4f); }
RT_HD unsigned char displayByte(float c, float exposure) {
    c = 1 - expf(-fmaxf(c, 0) * exposure);
    c = c <= .0031308f ? 12.92f * c : 1.055f * powf(c, 1 / 2.4f) - .055f;
    return (unsigned char)(255 * fminf(c, 1) + .5f);
}
// Native/renderer.cu:91-95
RT_HD f3 aroundAxis(f3 axis, float cosine, float angle) {
    f3 helper = fabsf(axis.y) < .999f ? F3(0, 1, 0) : F3(1, 0, 0);
    f3 tangent = unitv(crossv(helper, axis));
    float sine = sqrtf(fmaxf(0, 1 - cosine * cosine));
    return unitv(axis * cosine + (tangent * (sine * cosf(angle)) + crossv(axis, tangent) * (sine * sinf(angle))));
}
RT_HD float powerHeuristic(float a, float b) { return a * a / fmaxf(a * a + b * b, 1e-30f); }
// Native/renderer.cu:181-192
RT_HD float glossExponent(const Triangle& t) { float r = fmaxf(1 - t.smoothness, .015f); return fmaxf(0, 2 / (r * r) - 2); }
RT_HD f3 evalBrdf(const Triangle& t, f3 n, f3 refl, f3 dir, float& pdf) {
    float c = fmaxf(0, dotv(n, dir)), p = t.specularProbability;
    pdf = (1 - p) * c / kPi;
    f3 val = V(t.color) * ((1 - p) / kPi);
    if (p > 0 && t.smoothness < .999f) {
        float e = glossExponent(t), lobe = powf(fmaxf(0, dotv(refl, dir)), e);
        pdf += p * (e + 1) * lobe / (2 * kPi);
        val = val + V(t.specularColor) * (p * (e + 2) * lobe / (2 * kPi));
    }
    return val;
}
// Native/renderer.cu:193-196
RT_HD f3 skyRadiance(f3 d, float strength) {
    float t = fminf(1, fmaxf(0, d.y));
    return (F3(.65f, .72f, .82f) * (1 - t) + F3(.15f, .32f, .65f) * t) * strength;
}

// Native/renderer.cu:198-210: primary ray from camera dimensions 0..3.
template <class Rng>
RT_HD void cameraRay(const CameraData& cam, const RenderSettings& rs, int x, int y, int w, int h, const Rng& rng, f3& o, f3& d) {
    float px = 2 * (x + rng(0)) / w - 1, py = 2 * (y + rng(1)) / h - 1;
    f3 right = V(cam.right), up = V(cam.up), fwd = V(cam.forward);
    o = V(cam.origin);
    d = unitv(fwd + (right * (px * cam.aspect * cam.tanHalfFov) + up * (py * cam.tanHalfFov)));
    if (cam.orthographic) {
        o = o + (right * (px * cam.aspect * cam.orthographicSize) + up * (py * cam.orthographicSize));
        d = fwd;
    } else if (rs.aperture > 0) {
        float ang = 2 * kPi * rng(2), rad = .5f * rs.aperture * sqrtf(rng(3));
        f3 lens = right * (rad * cosf(ang)) + up * (rad * sinf(ang));
        f3 focus = o + d * (rs.focusDistance / dotv(d, fwd));
        o = o + lens;
        d = unitv(focus - o);
    }
}

struct PathState {
    f3 thr; float prevPdf;
    f3 rad;
};
struct FirstHit { f3 albedo; f3 normal; float depth; bool covered; };
struct ShadeResult {
    bool cont;        // a next ray (nextO, nextD) follows
    bool shadow;      // a shadow query (o=nextO, dir=sdir, limit) guards contribution
    f3 nextO, nextD;
    f3 sdir; float slimit;
    f3 contrib;       // added to the path radiance when the shadow query misses
};

// One bounce of Native/renderer.cu:215-286 after the closest-hit query.  rng(k) returns the
// sample of dimension 4 + 8*bounce + k.  The next-event contribution is returned instead of
// added, and the caller adds it when the shadow query reports no occluder; radiance terms
// are still added in the reference's order because the next bounce starts after that.
template <class Rng>
RT_HD void shadeBounce(const SceneView& sc, const RenderSettings& rs, int bounce, const Rng& rng,
                       f3 o, f3 d, const RtHit& hit, PathState& ps, FirstHit* first, ShadeResult& res) {
    res.cont = false; res.shadow = false;
    if (hit.index < 0) {
        if (bounce > 0) ps.rad = ps.rad + mulv(ps.thr, skyRadiance(d, rs.skyStrength));
        return;
    }
    const Triangle& t = sc.tris[hit.index];
    f3 g = unitv(crossv(V(t.edge1), V(t.edge2)));
    if (dotv(g, d) > 0) g = g * -1;
    if (bounce == 0) {
        first->covered = true;
        first->albedo = t.emission > 0 ? F3(1, 1, 1)
                                       : V(t.color) * (1 - t.specularProbability) + V(t.specularColor) * t.specularProbability;
        first->normal = g;
        first->depth = hit.distance;
    }
    if (t.emission > 0) {
        float lp = sc.lightArea > 0 ? hit.distance * hit.distance / (sc.lightArea * fmaxf(1e-6f, -dotv(g, d))) : 0;
        float w = ps.prevPdf > 0 ? powerHeuristic(ps.prevPdf, lp) : 1;
        ps.rad = ps.rad + mulv(ps.thr, V(t.color)) * (t.emission * w);
        return;
    }
    f3 n = unitv(V(t.normal0) * (1 - hit.u - hit.v) + (V(t.normal1) * hit.u + V(t.normal2) * hit.v));
    if (dotv(n, g) < 0) n = n * -1;
    if (dotv(n, d) >= 0) n = g;
    if (bounce == 0) first->normal = n;
    f3 p = o + d * hit.distance;
    f3 refl = d - n * (2 * dotv(d, n));
    f3 no = p + g * .0005f;
    res.nextO = no;
    if (sc.lightCount > 0) {
        float sel = rng(0) * sc.lightArea;
        // First light whose cumulative area exceeds sel: the reference's linear scan
        // (Native/renderer.cu:250) written as a binary search with the same result.
        int lo = 0, hi = sc.lightCount - 1;
        while (lo < hi) {
            int mid = (lo + hi) >> 1;
            if (sel >= sc.lights[mid].cumulativeArea) lo = mid + 1; else hi = mid;
        }
        const Triangle& L = sc.tris[sc.lights[lo].triangle];
        float r = sqrtf(rng(1)), vv = rng(2);
        f3 target = V(L.origin) + (V(L.edge1) * (r * (1 - vv)) + V(L.edge2) * (r * vv));
        f3 off = target - no;
        float dist2 = dotv(off, off), dist = sqrtf(dist2);
        f3 in = off * (1 / dist), ln = unitv(crossv(V(L.edge1), V(L.edge2)));
        float lc = -dotv(ln, in);
        if (!L.oneSided) lc = fabsf(lc);
        float sc2 = fmaxf(0, dotv(n, in));
        if (lc > 0 && sc2 > 0 && dotv(g, in) > 0) {
            float pdf, lp = dist2 / (sc.lightArea * lc);
            f3 f = evalBrdf(t, n, refl, in, pdf);
            float w = bounce + 1 == rs.maxBounces ? 1 : powerHeuristic(lp, pdf);
            res.shadow = true;
            res.sdir = in;
            res.slimit = dist - .001f;
            res.contrib = mulv(mulv(ps.thr, f), V(L.color)) * (L.emission * sc2 * w / lp);
        }
    }
    if (bounce + 1 == rs.maxBounces) return;
    bool spec = rng(3) < t.specularProbability;
    f3 nd;
    if (spec && t.smoothness >= .999f) {
        nd = refl; ps.thr = mulv(ps.thr, V(t.specularColor)); ps.prevPdf = 0;
    } else {
        float cosine = spec ? powf(rng(4), 1 / (glossExponent(t) + 1)) : sqrtf(rng(4));
        nd = aroundAxis(spec ? refl : n, cosine, 2 * kPi * rng(5));
        if (dotv(g, nd) <= 0 || dotv(n, nd) <= 0) return;
        float pdf;
        f3 f = evalBrdf(t, n, refl, nd, pdf);
        if (pdf < 1e-12f) return;
        ps.thr = mulv(ps.thr, f * (dotv(n, nd) / pdf)); ps.prevPdf = pdf;
    }
    if (bounce >= 3) {
        float surv = fminf(.95f, fmaxf(ps.thr.x, fmaxf(ps.thr.y, ps.thr.z)));
        if (rng(6) >= surv) return;
        ps.thr = ps.thr * (1 / surv);
    }
    res.cont = true;
    res.nextD = nd;
}

}  // namespace rt
