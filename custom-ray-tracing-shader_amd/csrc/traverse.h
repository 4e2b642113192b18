// traverse.h - W-wide BVH traversal over 64-byte aligned nodes (full precision or
// quantized), for the gfx950 kernels
// (LDS short stack with a global spill ring) and for the host backends (array stack).
//
// Result semantics follow Native/renderer.cu:136-178: the accepted hit is the smallest t in
// (1e-4, limit), ties broken by the lower original triangle index; any-hit queries stop at
// the first accepted triangle.  Child boxes are visited near-first; a deferred child is
// skipped when its entry distance exceeds the current best hit, exactly when the reference's
// re-test on pop (renderer.cu:170-175) would reject it.
#pragma once
#include "rt_core.h"

namespace rt {

// Sorts (dist, ref) pairs ascending by dist with a fixed compare-exchange network.
template <int W> RT_HD void sortChildren(float* d, int* r);
template <> RT_HD void sortChildren<2>(float* d, int* r) {
    if (d[1] < d[0]) { float t = d[0]; d[0] = d[1]; d[1] = t; int q = r[0]; r[0] = r[1]; r[1] = q; }
}
#define RT_CX(a, b) if (d[b] < d[a]) { float t = d[a]; d[a] = d[b]; d[b] = t; int q = r[a]; r[a] = r[b]; r[b] = q; }
template <> RT_HD void sortChildren<4>(float* d, int* r) {
    RT_CX(0, 1) RT_CX(2, 3) RT_CX(0, 2) RT_CX(1, 3) RT_CX(1, 2)
}
template <> RT_HD void sortChildren<8>(float* d, int* r) {
    RT_CX(0, 2) RT_CX(1, 3) RT_CX(4, 6) RT_CX(5, 7) RT_CX(0, 4) RT_CX(1, 5) RT_CX(2, 6) RT_CX(3, 7)
    RT_CX(0, 1) RT_CX(2, 3) RT_CX(4, 5) RT_CX(6, 7) RT_CX(2, 4) RT_CX(3, 5) RT_CX(1, 4) RT_CX(3, 6)
    RT_CX(1, 2) RT_CX(3, 4) RT_CX(5, 6)
}
#undef RT_CX

// Tests the triangles of one leaf (records first.. up to the last-in-leaf flag).
template <bool ANY>
RT_HD bool leafTest(const TriRecord* recs, int first, f3 o, f3 d, float& best, int& bestIdx, float& bu, float& bv,
                    unsigned long long& triTests, bool watertight) {
    for (int k = first;; k++) {
        const TriRecord& r = recs[k];
        float u = 0, v = 0;
        float t = intersectTriMode(watertight, o, d, F3(r.v0x, r.v0y, r.v0z), F3(r.e1x, r.e1y, r.e1z),
                                   F3(r.e2x, r.e2y, r.e2z), (r.flags & kTriOneSided) != 0, u, v);
        triTests++;
        if (t > 0 && (t < best || (t == best && r.index < bestIdx))) {
            best = t; bestIdx = r.index; bu = u; bv = v;
            if (ANY) return true;
        }
        if (r.flags & kTriLastInLeaf) return false;
    }
}

// Host traversal with a plain array stack; the same visiting rules as the device kernel.
// N is WideNode<W> or QNode<W>.
template <class N, bool ANY>
inline RtHit traverseHost(const N* nodes, const TriRecord* recs, f3 o, f3 d, float limit, bool watertight = false) {
    constexpr int W = NodeTraits<N>::kWidth;
    f3 inv = safeInverse(d);
    float best = limit, bu = 0, bv = 0;
    int bestIdx = -1;
    int sref[36 * (W - 1) + 1];
    float sdist[36 * (W - 1) + 1];
    int sp = 0, ref = 0;
    unsigned long long tt = 0;
    for (;;) {
        if (ref >= 0) {
            const N& n = nodes[ref];
            float dist[W]; int cr[W];
            for (int c = 0; c < W; c++) {
                dist[c] = n.entry(c, o, inv, best);
                cr[c] = n.child[c];
            }
            sortChildren<W>(dist, cr);
            for (int c = W - 1; c >= 1; c--)
                if (dist[c] != INFINITY) { sref[sp] = cr[c]; sdist[sp] = dist[c]; sp++; }
            if (dist[0] != INFINITY) { ref = cr[0]; continue; }
        } else {
            if (leafTest<ANY>(recs, leafFirst(ref), o, d, best, bestIdx, bu, bv, tt, watertight)) break;
        }
        bool found = false;
        while (sp > 0) {
            sp--;
            if (sdist[sp] <= best) { ref = sref[sp]; found = true; break; }
        }
        if (!found) break;
    }
    RtHit h; h.index = bestIdx; h.distance = best; h.u = bu; h.v = bv;
    return h;
}

}  // namespace rt
