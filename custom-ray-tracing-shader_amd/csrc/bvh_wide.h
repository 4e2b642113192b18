// bvh_wide.h - collapse of a binary BVH (reference node layout) into the W-wide, 64-byte
// aligned traversal layout, and the leaf-ordered 48-byte triangle records.
//
// The binary tree follows the reference's conventions (Native/bvh.cuh:7,101-104): inner
// nodes have count 0 and two adjacent children at `first`, leaves index `order`.  A wide
// node is formed by repeatedly opening the inner child with the largest surface area until
// W children are gathered (the usual greedy collapse); leaves keep their exact triangle
// sets, so closest-hit results are the same as on the binary tree.
#pragma once
#include <vector>
#include <cstring>
#include "rt_core.h"

namespace rt {

inline float halfArea(const RtBvhNode& n) {
    float x = fmaxf(0, n.high[0] - n.low[0]), y = fmaxf(0, n.high[1] - n.low[1]), z = fmaxf(0, n.high[2] - n.low[2]);
    return x * y + x * z + y * z;
}

template <int W>
struct WideBuilder {
    const RtBvhNode* bin;
    const int* order;
    const Triangle* tris;
    std::vector<WideNode<W>> nodes;
    std::vector<TriRecord> recs;
    int maxDepth = 0;

    int leafRecords(const RtBvhNode& leaf) {
        int first = (int)recs.size();
        for (int j = 0; j < leaf.count; j++) {
            const Triangle& t = tris[order[leaf.first + j]];
            TriRecord r;
            r.v0x = t.origin.x; r.v0y = t.origin.y; r.v0z = t.origin.z; r.index = order[leaf.first + j];
            r.e1x = t.edge1.x; r.e1y = t.edge1.y; r.e1z = t.edge1.z;
            r.flags = (t.oneSided ? kTriOneSided : 0) | (j + 1 == leaf.count ? kTriLastInLeaf : 0);
            r.e2x = t.edge2.x; r.e2y = t.edge2.y; r.e2z = t.edge2.z; r.pad = 0;
            recs.push_back(r);
        }
        return first;
    }
    // Emits the wide node whose children are gathered below binary node b; returns its index.
    int emit(int b, int depth) {
        if (depth > maxDepth) maxDepth = depth;
        int idx = (int)nodes.size();
        nodes.emplace_back();
        int kids[W], nk = 0;
        if (bin[b].count > 0) { kids[nk++] = b; }
        else { kids[nk++] = bin[b].first; kids[nk++] = bin[b].first + 1; }
        while (nk < W) {
            int pick = -1; float area = -1;
            for (int k = 0; k < nk; k++)
                if (bin[kids[k]].count == 0 && halfArea(bin[kids[k]]) > area) { area = halfArea(bin[kids[k]]); pick = k; }
            if (pick < 0) break;
            int open = kids[pick];
            kids[pick] = bin[open].first;
            kids[nk++] = bin[open].first + 1;
        }
        WideNode<W> node;
        for (int c = 0; c < W; c++) {
            node.lox[c] = node.loy[c] = node.loz[c] = INFINITY;
            node.hix[c] = node.hiy[c] = node.hiz[c] = INFINITY;
            node.child[c] = kEmptyChild; node.count[c] = 0;
        }
        for (int c = 0; c < nk; c++) {
            const RtBvhNode& k = bin[kids[c]];
            node.lox[c] = k.low[0]; node.loy[c] = k.low[1]; node.loz[c] = k.low[2];
            node.hix[c] = k.high[0]; node.hiy[c] = k.high[1]; node.hiz[c] = k.high[2];
            if (k.count > 0) { node.child[c] = ~leafRecords(k); node.count[c] = k.count; }
            else { node.child[c] = emit(kids[c], depth + 1); node.count[c] = 0; }
        }
        nodes[idx] = node;
        return idx;
    }
    void build(const RtBvhNode* b, const int* o, const Triangle* t) {
        bin = b; order = o; tris = t;
        nodes.clear(); recs.clear(); maxDepth = 0;
        emit(0, 0);
    }
};

// Host binned-SAH build in the reference's binary layout: `bins` bins per axis over the
// centroid extent, leaf at <= maxLeaf triangles or at depth 32, cost A + A_L n_L + A_R n_R
// against the leaf cost A n (Native/bvh.cuh:35-105).
struct HostBinaryBvh {
    struct Box { float lo[3], hi[3]; };
    std::vector<RtBvhNode> nodes;
    std::vector<int> order;
    std::vector<Box> pbox;
    std::vector<float> cen;   // 3 per primitive
    int depth = 0, bins = 32, maxLeaf = 4;

    static void empty(Box& b) { for (int k = 0; k < 3; k++) { b.lo[k] = FLT_MAX; b.hi[k] = -FLT_MAX; } }
    static void grow(Box& b, const Box& o) { for (int k = 0; k < 3; k++) { b.lo[k] = fminf(b.lo[k], o.lo[k]); b.hi[k] = fmaxf(b.hi[k], o.hi[k]); } }
    static float area(const Box& b) {
        float x = fmaxf(0, b.hi[0] - b.lo[0]), y = fmaxf(0, b.hi[1] - b.lo[1]), z = fmaxf(0, b.hi[2] - b.lo[2]);
        return x * y + x * z + y * z;
    }
    int binOf(float c, float lo, float scale) const {
        int b = (int)((c - lo) * scale);
        return b < 0 ? 0 : (b > bins - 1 ? bins - 1 : b);
    }
    void build(const Triangle* t, int n) {
        pbox.resize(n); cen.resize((size_t)n * 3); order.resize(n);
        for (int i = 0; i < n; i++) {
            order[i] = i;
            float v[3][3] = {{t[i].origin.x, t[i].origin.y, t[i].origin.z},
                             {t[i].origin.x + t[i].edge1.x, t[i].origin.y + t[i].edge1.y, t[i].origin.z + t[i].edge1.z},
                             {t[i].origin.x + t[i].edge2.x, t[i].origin.y + t[i].edge2.y, t[i].origin.z + t[i].edge2.z}};
            Box& b = pbox[i];
            for (int k = 0; k < 3; k++) {
                b.lo[k] = fminf(fminf(fminf(FLT_MAX, v[0][k]), v[1][k]), v[2][k]);
                b.hi[k] = fmaxf(fmaxf(fmaxf(-FLT_MAX, v[0][k]), v[1][k]), v[2][k]);
                cen[(size_t)i * 3 + k] = (v[0][k] + v[1][k] + v[2][k]) / 3;
                float pad = 1e-5f + 1e-6f * fmaxf(fabsf(b.lo[k]), fabsf(b.hi[k]));
                b.lo[k] -= pad; b.hi[k] += pad;
            }
        }
        nodes.clear(); nodes.reserve((size_t)n * 2); nodes.resize(1); depth = 0;
        split(0, 0, n, 0);
    }
    void split(int ni, int first, int count, int d) {
        if (d > depth) depth = d;
        Box bb, cb; empty(bb); empty(cb);
        for (int i = first; i < first + count; i++) {
            grow(bb, pbox[order[i]]);
            const float* c = &cen[(size_t)order[i] * 3];
            for (int k = 0; k < 3; k++) { cb.lo[k] = fminf(cb.lo[k], c[k]); cb.hi[k] = fmaxf(cb.hi[k], c[k]); }
        }
        RtBvhNode& nd = nodes[ni];
        for (int k = 0; k < 3; k++) { nd.low[k] = bb.lo[k]; nd.high[k] = bb.hi[k]; }
        nd.first = first; nd.count = count;
        if (count <= maxLeaf || d == kBvhDepthLimit) return;
        int bestAxis = -1, bestBin = -1;
        float best = area(bb) * count;
        std::vector<Box> bbox(bins), suf(bins);
        std::vector<int> bcnt(bins), sufn(bins);
        for (int axis = 0; axis < 3; axis++) {
            float ext = cb.hi[axis] - cb.lo[axis];
            if (ext <= 1e-12f) continue;
            float scale = bins / ext, lo = cb.lo[axis];
            for (int b = 0; b < bins; b++) { empty(bbox[b]); bcnt[b] = 0; }
            for (int i = first; i < first + count; i++) {
                int b = binOf(cen[(size_t)order[i] * 3 + axis], lo, scale);
                bcnt[b]++; grow(bbox[b], pbox[order[i]]);
            }
            Box acc; empty(acc); int an = 0;
            for (int b = bins - 1; b >= 0; b--) {
                if (bcnt[b]) grow(acc, bbox[b]);
                an += bcnt[b]; suf[b] = acc; sufn[b] = an;
            }
            Box left; empty(left); int ln = 0;
            for (int b = 0; b < bins - 1; b++) {
                if (bcnt[b]) grow(left, bbox[b]);
                ln += bcnt[b];
                if (!ln || !sufn[b + 1]) continue;
                float cost = area(bb) + area(left) * ln + area(suf[b + 1]) * sufn[b + 1];
                if (cost < best) { best = cost; bestAxis = axis; bestBin = b; }
            }
        }
        if (bestAxis < 0) return;
        float lo = cb.lo[bestAxis], scale = bins / (cb.hi[bestAxis] - lo);
        int i = first, j = first + count - 1;
        while (i <= j) {
            if (binOf(cen[(size_t)order[i] * 3 + bestAxis], lo, scale) <= bestBin) i++;
            else { int tmp = order[i]; order[i] = order[j]; order[j] = tmp; j--; }
        }
        int nl = i - first;
        if (nl == 0 || nl == count) return;
        int child = (int)nodes.size();
        nodes.resize(nodes.size() + 2);
        nodes[ni].first = child; nodes[ni].count = 0;
        split(child, first, nl, d + 1);
        split(child + 1, first + nl, count - nl, d + 1);
    }
};

}  // namespace rt
