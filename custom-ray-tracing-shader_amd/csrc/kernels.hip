// kernels.hip - gfx950 kernels of the wavefront path tracer.
//
// One RtRender call runs, per chunk of R sample groups (R*npix path slots) and per path
// pipeline (a part of the chunk's paths on its own stream):
//   reset counters -> raygen -> for each bounce b: one traversal launch over bounce b's
//   closest-hit queue and bounce b-1's shadow queue (persistent lanes; a shadow ray adds its
//   NEE term when unoccluded) -> shade (NEE shadow request + next ray, compacted with
//   workgroup-aggregated atomics) -> traversal of the last bounce's shadow queue
//   [or, from a chosen bounce on, one path-completion launch for the remaining bounces]
//   -> accumulate (adds the R groups to the FP32 sums in sample order, tone-maps).
// Queue sizes live in device memory, so no host synchronisation happens inside a frame.
#include "kernels.h"
#include "traverse.h"
#include <type_traits>

namespace rt {

__device__ const unsigned gSobolDirs[kSobolDims * 32] = {
#include "sobol_directions.inc"
};
static const unsigned kSobolDirsHost[kSobolDims * 32] = {
#include "sobol_directions.inc"
};
const unsigned* hostSobolDirections() { return kSobolDirsHost; }

// ------------------------------------------------------------------ traversal
// Leaf records fetched together before testing (one memory round trip per batch instead of
// one per triangle).
#ifndef RT_LEAF_BATCH
#define RT_LEAF_BATCH 2
#endif
constexpr int kLeafBatch = RT_LEAF_BATCH;

template <class N>
__device__ __forceinline__ void loadNode(const N* nodes, int ref, N& n) {
    constexpr int V = NodeTraits<N>::kVectors;
    union { float4 v[sizeof(N) / 16]; N node; } u;
    const float4* p = reinterpret_cast<const float4*>(nodes + ref);
#pragma unroll
    for (int j = 0; j < V; j++) u.v[j] = p[j];
    n = u.node;
}

// Per-lane traversal state, kept in registers across refill rounds.
struct LaneRay {
    f3 o, d, inv;
    f3 oi;          // o * inv
    bool nz;        // no zero component in inv: slab tests take boxEntryFma
    float best, bu, bv;
    int bestIdx, sp, ref;
    int leaf;       // postponed leaf reference (< 0), or >= 0 for none
    int index;      // position of the ray in its queue, -1 when the lane is idle
    int slot;       // path slot
    bool any;       // any-hit query: answered by the first accepted hit
};

// Per-lane node visits and triangle tests by query kind (measurement launches only).
struct TraceCounters {
    unsigned long long nodesC = 0, trisC = 0, nodesA = 0, trisA = 0;
    __device__ __forceinline__ void node(bool any) { if (any) nodesA++; else nodesC++; }
    __device__ __forceinline__ void tri(bool any) { if (any) trisA++; else trisC++; }
};

template <int S>
struct LdsStack {
    int* ref;
    float* dist;
    int2* spill;
    int spillEntries, lanes, gl, tid;
    // Ring slot of stack position p (S = 8, 12 or 16 entries).
    static __device__ __forceinline__ int ringSlot(int p) {
        return (S & (S - 1)) == 0 ? (p & (S - 1)) : (int)((unsigned)p % (unsigned)S);
    }
    __device__ __forceinline__ void push(int& sp, int r, float dd, bool& overflow) {
        const int slot = ringSlot(sp);
        if (sp >= S) {
            const int e = sp - S;
            if (e < spillEntries)
                spill[(size_t)e * lanes + gl] = make_int2(ref[slot * kBlock + tid], __float_as_int(dist[slot * kBlock + tid]));
            else
                overflow = true;
        }
        ref[slot * kBlock + tid] = r;
        dist[slot * kBlock + tid] = dd;
        sp++;
    }
    // Before n entries are written at sp..sp+n-1: the ring entries they displace (index
    // p - S for each p >= S) move to the spill ring, as push() does one at a time.
    __device__ __forceinline__ void spillBelow(int sp, int n, bool& overflow) {
        if (sp + n <= S) return;
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const int p = sp + k;
            if (k < n && p >= S) {
                const int e = p - S, slot = ringSlot(p);
                if (e < spillEntries)
                    spill[(size_t)e * lanes + gl] = make_int2(ref[slot * kBlock + tid], __float_as_int(dist[slot * kBlock + tid]));
                else
                    overflow = true;
            }
        }
    }
    __device__ __forceinline__ void put(int p, int r, float dd) {
        const int slot = ringSlot(p);
        ref[slot * kBlock + tid] = r;
        dist[slot * kBlock + tid] = dd;
    }
    // Pushes the entered children 3, 2, 1 of a distance-sorted 4-wide node, far to near, with
    // the stack pointer and spill test computed once instead of per entry: entered children
    // form a prefix of the sorted order (INFINITY sorts last), so child j <= n goes to
    // sp + n - j.
    __device__ __forceinline__ void pushSorted3(int& sp, const float* d, const int* r, bool& overflow) {
        const bool v1 = d[1] != INFINITY, v2 = d[2] != INFINITY, v3 = d[3] != INFINITY;
        const int n = (int)v1 + (int)v2 + (int)v3;
        spillBelow(sp, n, overflow);
        if (v1) put(sp + n - 1, r[1], d[1]);
        if (v2) put(sp + n - 2, r[2], d[2]);
        if (v3) put(sp + n - 3, r[3], d[3]);
        sp += n;
    }
    // Any-hit visit of a 4-wide node: descends into the first entered child in slot order and
    // pushes the other entered children in slot order (as push() calls would).  Child 0 is
    // never pushed; child c > 0 is pushed when entered after an earlier entered child, at
    // sp + (children pushed before it).
    __device__ __forceinline__ bool descendAny4(int& sp, const float* d, const int* r, bool& overflow, int* next) {
        const bool v0 = d[0] != INFINITY, v1 = d[1] != INFINITY, v2 = d[2] != INFINITY, v3 = d[3] != INFINITY;
        *next = v0 ? r[0] : v1 ? r[1] : v2 ? r[2] : r[3];
        const bool u1 = v0 && v1, u2 = (v0 || v1) && v2, u3 = (v0 || v1 || v2) && v3;
        const int b2 = (int)u1, b3 = b2 + (int)u2, n = b3 + (int)u3;
        spillBelow(sp, n, overflow);
        if (u1) put(sp, r[1], d[1]);
        if (u2) put(sp + b2, r[2], d[2]);
        if (u3) put(sp + b3, r[3], d[3]);
        sp += n;
        return v0 || v1 || v2 || v3;
    }
    __device__ __forceinline__ void pop(int& sp, int& r, float& dd) {
        sp--;
        const int slot = ringSlot(sp);
        r = ref[slot * kBlock + tid];
        dd = dist[slot * kBlock + tid];
        if (sp >= S) {
            const int e = sp - S;
            if (e < spillEntries) {
                const int2 back = spill[(size_t)e * lanes + gl];
                ref[slot * kBlock + tid] = back.x;
                dist[slot * kBlock + tid] = __int_as_float(back.y);
            }
        }
    }
};

// Entry distances and references of a node's children.
template <class N>
__device__ __forceinline__ void childEntries(const N* nodes, const LaneRay& L, float* dist, int* cr) {
    constexpr int W = NodeTraits<N>::kWidth;
    N n;
    loadNode<N>(nodes, L.ref, n);
    if (L.nz) {
#pragma unroll
        for (int c = 0; c < W; c++) dist[c] = n.entryFma(c, L.oi, L.inv, L.best);
    } else {
#pragma unroll
        for (int c = 0; c < W; c++) dist[c] = n.entry(c, L.o, L.inv, L.best);
    }
#pragma unroll
    for (int c = 0; c < W; c++) cr[c] = n.child[c];
}
// 4-wide nodes, rays with no zero inverse component: the entry and exit planes of each axis
// are loaded directly (offsets chosen by the direction's signs, the far plane 16 bytes away
// from the near one), so each slab distance needs no min/max pair.  fma(x, inv, -oi) is
// monotonic in x, so the plane chosen by the sign gives exactly min / max of the two slab
// distances that boxEntryFma takes: the same entries, bit for bit.
template <>
__device__ __forceinline__ void childEntries<WideNode<4>>(const WideNode<4>* nodes, const LaneRay& L, float* dist,
                                                         int* cr) {
    // A 4-wide node is six float4 planes, lox hix loy hiy loz hiz (16 B each), then the child
    // references.  A ray enters along an axis through the low plane when its inverse
    // direction's sign bit is clear, else through the high plane: byte offset 16 * sign.
    const unsigned sx = (__float_as_uint(L.inv.x) >> 27) & 16u, sy = (__float_as_uint(L.inv.y) >> 27) & 16u,
                   sz = (__float_as_uint(L.inv.z) >> 27) & 16u;
    // 32-bit byte offsets from the node array (builds refuse more than kMaxWide4Nodes nodes).
    const char* arr = reinterpret_cast<const char*>(nodes);
    const unsigned node = (unsigned)L.ref << 7;
    const float4 nx = *reinterpret_cast<const float4*>(arr + (node | sx));
    const float4 fx = *reinterpret_cast<const float4*>(arr + (node | (sx ^ 16u)));
    const float4 ny = *reinterpret_cast<const float4*>(arr + (node | (32u + sy)));
    const float4 fy = *reinterpret_cast<const float4*>(arr + (node | (32u + (sy ^ 16u))));
    const float4 nz = *reinterpret_cast<const float4*>(arr + (node | (64u + sz)));
    const float4 fz = *reinterpret_cast<const float4*>(arr + (node | (64u + (sz ^ 16u))));
    const int4 ch = *reinterpret_cast<const int4*>(arr + (node | 96u));
    const float nxa[4] = {nx.x, nx.y, nx.z, nx.w}, fxa[4] = {fx.x, fx.y, fx.z, fx.w};
    const float nya[4] = {ny.x, ny.y, ny.z, ny.w}, fya[4] = {fy.x, fy.y, fy.z, fy.w};
    const float nza[4] = {nz.x, nz.y, nz.z, nz.w}, fza[4] = {fz.x, fz.y, fz.z, fz.w};
    if (L.nz) {
#pragma unroll
        for (int c = 0; c < 4; c++) {
            const float tnx = fmaf(nxa[c], L.inv.x, -L.oi.x), tfx = fmaf(fxa[c], L.inv.x, -L.oi.x);
            const float tny = fmaf(nya[c], L.inv.y, -L.oi.y), tfy = fmaf(fya[c], L.inv.y, -L.oi.y);
            const float tnz = fmaf(nza[c], L.inv.z, -L.oi.z), tfz = fmaf(fza[c], L.inv.z, -L.oi.z);
            const float nearT = fmaxf(fmaxf(tnx, tny), fmaxf(tnz, 0.0f));
            const float farT = fminf(fminf(tfx, tfy), fminf(tfz, L.best));
            dist[c] = nearT <= farT ? nearT : INFINITY;
        }
    } else {
        // A zero inverse component: the full slab test on the planes put back in order.
#pragma unroll
        for (int c = 0; c < 4; c++)
            dist[c] = boxEntry(L.o, L.inv, sx ? fxa[c] : nxa[c], sy ? fya[c] : nya[c], sz ? fza[c] : nza[c],
                               sx ? nxa[c] : fxa[c], sy ? nya[c] : fya[c], sz ? nza[c] : fza[c], L.best);
    }
    cr[0] = ch.x; cr[1] = ch.y; cr[2] = ch.z; cr[3] = ch.w;
}

// The 256-byte node in two halves of four children: the second half's loads are issued
// after the first half's slab tests (they hit in L1: both halves span the same two cache
// lines), so only half the node's registers are live at once.
template <>
__device__ __forceinline__ void childEntries<WideNode<8>>(const WideNode<8>* nodes, const LaneRay& L, float* dist,
                                                         int* cr) {
    const float4* p = reinterpret_cast<const float4*>(nodes + L.ref);
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const float4 lx = p[h], hx = p[2 + h], ly = p[4 + h], hy = p[6 + h], lz = p[8 + h], hz = p[10 + h];
        const int4 ch = reinterpret_cast<const int4*>(p)[12 + h];
        const float lxa[4] = {lx.x, lx.y, lx.z, lx.w}, hxa[4] = {hx.x, hx.y, hx.z, hx.w};
        const float lya[4] = {ly.x, ly.y, ly.z, ly.w}, hya[4] = {hy.x, hy.y, hy.z, hy.w};
        const float lza[4] = {lz.x, lz.y, lz.z, lz.w}, hza[4] = {hz.x, hz.y, hz.z, hz.w};
        if (L.nz) {
#pragma unroll
            for (int c = 0; c < 4; c++)
                dist[4 * h + c] = boxEntryFma(L.oi, L.inv, lxa[c], lya[c], lza[c], hxa[c], hya[c], hza[c], L.best);
        } else {
#pragma unroll
            for (int c = 0; c < 4; c++)
                dist[4 * h + c] = boxEntry(L.o, L.inv, lxa[c], lya[c], lza[c], hxa[c], hya[c], hza[c], L.best);
        }
        cr[4 * h] = ch.x; cr[4 * h + 1] = ch.y; cr[4 * h + 2] = ch.z; cr[4 * h + 3] = ch.w;
        __builtin_amdgcn_sched_barrier(0);
    }
}

// sortChildren<4>'s network (the same comparisons, so the same order, ties included) as
// selects: every lane runs the five exchanges without branching.
__device__ __forceinline__ void exchangeSelect(float& da, float& db, int& ra, int& rb) {
    const bool s = db < da;
    const float lo = s ? db : da, hi = s ? da : db;
    const int rl = s ? rb : ra, rh = s ? ra : rb;
    da = lo; db = hi; ra = rl; rb = rh;
}
__device__ __forceinline__ void sortChildrenSelect4(float* d, int* r) {
    // Named values rather than array slots: selects between array elements would otherwise be
    // rewritten into selects of indices.
    float d0 = d[0], d1 = d[1], d2 = d[2], d3 = d[3];
    int r0 = r[0], r1 = r[1], r2 = r[2], r3 = r[3];
    exchangeSelect(d0, d1, r0, r1); exchangeSelect(d2, d3, r2, r3); exchangeSelect(d0, d2, r0, r2);
    exchangeSelect(d1, d3, r1, r3); exchangeSelect(d1, d2, r1, r2);
    d[0] = d0; d[1] = d1; d[2] = d2; d[3] = d3;
    r[0] = r0; r[1] = r1; r[2] = r2; r[3] = r3;
}

// Pushes the entered children that are not descended into; returns whether the lane
// descends (into *next).  Closest-hit queries go to the nearest child and push the others
// far to near.  Any-hit queries take the first entered child in slot order and skip the
// sort: whether the ray is occluded does not depend on the order.  MIXED: the launch has
// any-hit rays (otherwise every query is closest-hit and the any-hit paths compile away).
template <int W, int S, bool MIXED>
__device__ __forceinline__ bool descend(LaneRay& L, LdsStack<S>& st, float* dist, int* cr, bool& overflow, int* next) {
    if constexpr (W == 4) {
        if (MIXED && L.any) return st.descendAny4(L.sp, dist, cr, overflow, next);
    }
    if (MIXED && L.any) {
        bool taken = false;
#pragma unroll
        for (int c = 0; c < W; c++) {
            const bool in = dist[c] != INFINITY;
            if (in && taken) st.push(L.sp, cr[c], dist[c], overflow);
            if (in && !taken) { *next = cr[c]; taken = true; }
        }
        return taken;
    }
    if constexpr (W == 4) {
        sortChildrenSelect4(dist, cr);
        st.pushSorted3(L.sp, dist, cr, overflow);
        *next = cr[0];
        return dist[0] != INFINITY;
    }
    sortChildren<W>(dist, cr);
#pragma unroll
    for (int c = W - 1; c >= 1; c--)
        if (dist[c] != INFINITY) st.push(L.sp, cr[c], dist[c], overflow);
    *next = cr[0];
    return dist[0] != INFINITY;
}

// One traversal step: visit an inner node (push far children, descend to the nearest) or
// test a leaf, then pop until a deferred entry survives the current best distance.
// Returns true when the ray is finished.
template <class N, bool MIXED, bool STATS, int S, bool WT = false>
__device__ __forceinline__ bool traceStep(LaneRay& L, const N* nodes, const float4* recs, LdsStack<S>& st,
                                          bool& overflow, TraceCounters& tc) {
    constexpr int W = NodeTraits<N>::kWidth;
    if (L.ref >= 0) {
        if (STATS) tc.node(L.any);
        float dist[W]; int cr[W];
        childEntries<N>(nodes, L, dist, cr);
        int next;
        if (descend<W, S, MIXED>(L, st, dist, cr, overflow, &next)) { L.ref = next; return false; }
    } else {
        // Records k.. of the leaf, kLeafBatch at a time; the reference count bounds the first
        // batches exactly, larger leaves continue until the last-in-leaf flag.
        int k = leafFirst(L.ref), known = leafBatch(L.ref);
        for (;;) {
            const int m = known < kLeafBatch ? known : kLeafBatch;
            // Unconditional loads (slots past m repeat the last record) so the whole batch is
            // in flight at once.
            float4 A[kLeafBatch], B[kLeafBatch], C[kLeafBatch];
#pragma unroll
            for (int j = 0; j < kLeafBatch; j++) {
                const int r = 3 * (k + (j < m ? j : m - 1));
                A[j] = recs[r]; B[j] = recs[r + 1]; C[j] = recs[r + 2];
            }
            bool last = false;
#pragma unroll
            for (int j = 0; j < kLeafBatch; j++) {
                if (j < m && !last) {
                    const int flags = __float_as_int(B[j].w), index = __float_as_int(A[j].w);
                    float u = 0, v = 0;
                    const float t = intersectTriMode(WT, L.o, L.d, F3(A[j].x, A[j].y, A[j].z), F3(B[j].x, B[j].y, B[j].z),
                                                     F3(C[j].x, C[j].y, C[j].z), (flags & kTriOneSided) != 0, u, v);
                    if (STATS) tc.tri(L.any);
                    if (t > 0 && (t < L.best || (t == L.best && index < L.bestIdx))) {
                        L.best = t; L.bestIdx = index; L.bu = u; L.bv = v;
                        if (MIXED && L.any) return true;
                    }
                    last = (flags & kTriLastInLeaf) != 0;
                }
            }
            if (last) break;
            k += m;
            known = known > m ? known - m : kLeafBatch;
        }
    }
    while (L.sp > 0) {
        int r; float dd;
        st.pop(L.sp, r, dd);
        if (dd <= L.best) { L.ref = r; return false; }
    }
    return true;
}

// ---- speculative while-while traversal (inner-node and leaf phases run separately, so the
// lanes of a wave execute the same phase; a lane that reaches its first leaf postpones it and
// keeps descending until every lane of the wave holds a leaf)
constexpr int kDone = INT32_MAX;   // reference after the last deferred entry

// Next deferred entry that can still hold a hit nearer than best, or kDone.
template <int S>
__device__ __forceinline__ int popNext(LaneRay& L, LdsStack<S>& st) {
    while (L.sp > 0) {
        int r; float dd;
        st.pop(L.sp, r, dd);
        if (dd <= L.best) return r;
    }
    return kDone;
}

template <class N, bool MIXED, bool STATS, int S>
__device__ __forceinline__ void visitNode(LaneRay& L, const N* nodes, LdsStack<S>& st, bool& overflow,
                                          TraceCounters& tc) {
    constexpr int W = NodeTraits<N>::kWidth;
    if (STATS) tc.node(L.any);
    float dist[W]; int cr[W];
    childEntries<N>(nodes, L, dist, cr);
    int next;
    L.ref = descend<W, S, MIXED>(L, st, dist, cr, overflow, &next) ? next : popNext<S>(L, st);
}

// Tests the triangles of leaf reference ref; true when an any-hit query is answered.
template <bool MIXED, bool STATS>
__device__ __forceinline__ bool testLeaf(LaneRay& L, int ref, const float4* recs, TraceCounters& tc) {
    int k = leafFirst(ref), known = leafBatch(ref);
    for (;;) {
        const int m = known < kLeafBatch ? known : kLeafBatch;
        float4 A[kLeafBatch], B[kLeafBatch], C[kLeafBatch];
#pragma unroll
        for (int j = 0; j < kLeafBatch; j++) {
            const int r = 3 * (k + (j < m ? j : m - 1));
            A[j] = recs[r]; B[j] = recs[r + 1]; C[j] = recs[r + 2];
        }
        bool last = false;
#pragma unroll
        for (int j = 0; j < kLeafBatch; j++) {
            if (j < m && !last) {
                const int flags = __float_as_int(B[j].w), index = __float_as_int(A[j].w);
                float u = 0, v = 0;
                const float t = intersectTri(L.o, L.d, F3(A[j].x, A[j].y, A[j].z), F3(B[j].x, B[j].y, B[j].z),
                                             F3(C[j].x, C[j].y, C[j].z), (flags & kTriOneSided) != 0, u, v);
                if (STATS) tc.tri(L.any);
                if (t > 0 && (t < L.best || (t == L.best && index < L.bestIdx))) {
                    L.best = t; L.bestIdx = index; L.bu = u; L.bv = v;
                    if (MIXED && L.any) return true;
                }
                last = (flags & kTriLastInLeaf) != 0;
            }
        }
        if (last) return false;
        k += m;
        known = known > m ? known - m : kLeafBatch;
    }
}

__device__ __forceinline__ void finishRay(const TraceArgs& a, const LaneRay& L) {
    int4* hits = L.any ? a.anyHits : a.hits;
    if (hits) {
        if (a.nontemporal) {
            typedef int v4i __attribute__((ext_vector_type(4)));
            const v4i hv = {L.bestIdx, __float_as_int(L.best), __float_as_int(L.bu), __float_as_int(L.bv)};
            __builtin_nontemporal_store(hv, reinterpret_cast<v4i*>(hits + L.index));
        } else {
            hits[L.index] = make_int4(L.bestIdx, __float_as_int(L.best), __float_as_int(L.bu), __float_as_int(L.bv));
        }
    } else if (L.bestIdx < 0) {
        // Shadow query found no occluder: add the stored next-event term.
        float4 acc = a.rad[L.slot];
        const float4 c = a.contrib[L.slot];
        acc.x = acc.x + c.x; acc.y = acc.y + c.y; acc.z = acc.z + c.z;
        a.rad[L.slot] = acc;
    }
}

// View owning a path slot: last v with viewBase[v] <= slot.
__device__ __forceinline__ int findView(const int* viewBase, int views, int slot) {
    int lo = 0, hi = views - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (viewBase[mid] <= slot) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Ray i of the launch: closest-hit queue entry i, or any-hit entry i - closest when i >= closest.
__device__ __forceinline__ void startRay(const TraceArgs& a, int i, int closest, LaneRay& L) {
    L.any = i >= closest;
    if (L.any) i -= closest;
    const float4* rays = L.any ? a.anyRays : a.rays;
    // Rays are read once: keep them from displacing BVH nodes and records in the caches.
    float4 r0, r1;
    if (a.nontemporal) {
        typedef float v4f __attribute__((ext_vector_type(4)));
        const v4f* rp = reinterpret_cast<const v4f*>(rays + 2 * i);
        const v4f q0 = __builtin_nontemporal_load(rp), q1 = __builtin_nontemporal_load(rp + 1);
        r0 = make_float4(q0.x, q0.y, q0.z, q0.w);
        r1 = make_float4(q1.x, q1.y, q1.z, q1.w);
    } else {
        r0 = rays[2 * i];
        r1 = rays[2 * i + 1];
    }
    L.o = F3(r0.x, r0.y, r0.z);
    L.d = F3(r1.x, r1.y, r1.z);
    L.inv = safeInverse(L.d);
    L.nz = L.inv.x != 0 && L.inv.y != 0 && L.inv.z != 0;
    L.oi = F3(L.o.x * L.inv.x, L.o.y * L.inv.y, L.o.z * L.inv.z);
    L.best = r0.w; L.bu = 0; L.bv = 0;
    L.bestIdx = -1; L.sp = 0;
    L.index = i;
    L.slot = __float_as_int(r1.w);
    L.ref = a.roots ? a.roots[findView(a.viewBase, a.views, L.slot)] : 0;
    L.leaf = 0;
}

// First index of segment k over [0, total) cut into n = 1 or kXcds segments.
__device__ __forceinline__ int segmentBound(int total, int k, int n) {
    static_assert(kXcds == 8, "segment bounds shift by 3");
    return n == 1 ? (k ? total : 0) : (int)(((long long)total * k) >> 3);
}

// Hands out a persistent launch's ray indices [0, total).  Each lane's first index is fixed
// by its grid position; after that a wave takes indices from a cursor (counting past the
// first batch), so the launch does not open with one atomic per wave on one address.
// With segment cursors (xcdWork), group g = blockIdx.x % 8 (the workgroups of one XCD)
// starts in segment g of the index space, so consecutive rays (neighbouring pixels, nearby
// paths) share that XCD's L2, and moves on to the other segments once its own is drained.
// A launch with fewer rays than lanes keeps one cursor: its waves would spend more on the
// atomics that find drained segments than they gain.  All state is wave-uniform.
#ifndef RT_SEGMENT_RAYS_PER_LANE
#define RT_SEGMENT_RAYS_PER_LANE 1
#endif
struct RayDealer {
    int X, firstRays, total, seg, hops, lo, hi;
    int *cursor, *segCursors;
    bool exhausted;   // every segment's cursor has passed its end
    __device__ __forceinline__ void init(int* xcdWork, int* work, int n) {
        total = n;
        X = (xcdWork && gridDim.x % kXcds == 0 && total >= RT_SEGMENT_RAYS_PER_LANE * (int)gridDim.x * kBlock) ? kXcds : 1;
        firstRays = (int)gridDim.x / X * kBlock;   // per segment: one ray per lane of its group
        seg = (int)blockIdx.x % X; hops = 0;
        segCursors = xcdWork;
        lo = segmentBound(total, seg, X); hi = segmentBound(total, seg + 1, X);
        cursor = X > 1 ? segCursors + seg * kXcdCursorStride : work;
        exhausted = false;
    }
    // This lane's first index, or -1.  Call once, before take().
    __device__ __forceinline__ int firstIndex(int tid) {
        const int i = lo + (int)blockIdx.x / X * kBlock + tid, end = hi;
        if (lo + firstRays >= hi) next();
        return i < end ? i : -1;
    }
    __device__ __forceinline__ void next() {
        if (++hops >= X) { exhausted = true; return; }
        seg = seg + 1 == X ? 0 : seg + 1;
        lo = segmentBound(total, seg, X); hi = segmentBound(total, seg + 1, X);
        cursor = segCursors + seg * kXcdCursorStride;
    }
    // First index of the idle lanes' batch from the current segment (indices >= hi are none).
    __device__ __forceinline__ int take(unsigned long long idle, int lane) {
        const int leader = __ffsll((long long)idle) - 1;
        int base = 0;
        if (lane == leader) base = lo + firstRays + atomicAdd(cursor, __popcll(idle));
        return __builtin_amdgcn_readlane(base, leader);
    }
    // After a take(): moves on when the batch reached the segment's end.
    __device__ __forceinline__ void after(int base, unsigned long long idle) {
        if (base + __popcll(idle) >= hi) next();
    }
};

// Persistent traversal of a closest-hit queue followed by an any-hit queue, in one ray index
// space, so the two overlap in one launch.  Each wave's first 64 rays are fixed by its
// position in the grid; after that it takes rays from a cursor (which counts the rays past
// the first batch), so the launch does not start with one atomic per wave on one address.
// The cursor is shared by the whole grid, or (a.xcdWork) one per XCD segment.
// With REFILL, lanes whose ray finished take new rays as soon as fewer than refillBelow
// lanes are busy, instead of waiting for the wave's slowest ray.
// Waves per SIMD the register allocation aims for.  The traversal is latency-bound, so its
// throughput follows the resident waves: the 8-entry stack's 16 KB of LDS lets ten
// workgroups share a CU, and both loops fit 72 VGPRs (7 waves) without spilling (64 VGPRs, 8
// waves, spills ~25 values of the step loop and measured 8% slower); the 12- and 16-entry
// stacks' LDS allows six and five.  8-wide nodes and the instrumented variants keep the
// compiler's own choice (a tighter budget would spill).
#ifndef RT_STEP_WAVES
#define RT_STEP_WAVES 7
#endif
#ifndef RT_WW_WAVES
#define RT_WW_WAVES 7
#endif
template <class N, int S, bool WW, bool STATS> constexpr int traceWavesPerSimd() {
    return (STATS || NodeTraits<N>::kWidth == 8) ? 1 : S == 16 ? 5 : S == 12 ? 6 : WW ? RT_WW_WAVES : RT_STEP_WAVES;
}
// WT: the watertight ray/triangle test (instantiated with the step loop and 16 entries only).
template <class N, bool MIXED, bool STATS, int S, bool REFILL, bool WW, bool WT = false>
__global__ __launch_bounds__(kBlock, (traceWavesPerSimd<N, S, WW, STATS>())) void traceKernel(TraceArgs a) {
    static_assert(!(WT && WW), "the watertight test is built into the step loop");
    __shared__ int sRef[S * kBlock];
    __shared__ float sDist[S * kBlock];
    const int tid = threadIdx.x, lane = tid & 63;
    const int closest = a.countPtr ? *a.countPtr : a.count;
    const int total = closest + (a.anyCountPtr ? *a.anyCountPtr : a.anyCount);
    const N* nodes = static_cast<const N*>(a.nodes);
    const float4* recs = reinterpret_cast<const float4*>(a.recs);
    LdsStack<S> st{sRef, sDist, reinterpret_cast<int2*>(a.spill), a.spillEntries, a.totalLanes,
                   (int)(blockIdx.x * kBlock + tid), tid};
    TraceCounters tc;
    bool overflow = false;
    LaneRay L;
    L.index = -1;
    RayDealer deal;
    deal.init(a.xcdWork, a.work, total);
    {
        const int i = deal.firstIndex(tid);
        if (i >= 0) startRay(a, i, closest, L);
    }
    for (;;) {
        const unsigned long long idle = __ballot(L.index < 0);
        if (idle && !deal.exhausted) {
            const int base = deal.take(idle, lane);
            if (L.index < 0) {
                const int i = base + (int)__builtin_amdgcn_mbcnt_hi((unsigned)(idle >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)idle, 0u));
                if (i < deal.hi) startRay(a, i, closest, L);
            }
            deal.after(base, idle);
        }
        if (__ballot(L.index >= 0) == 0) {
            if (deal.exhausted) break;
            continue;   // the segment ran dry before any lane got a ray: take the next one
        }
        if (WW && L.index >= 0) {
            for (;;) {
                bool searching = true;
                while ((unsigned)L.ref < (unsigned)kDone) {
                    visitNode<N, MIXED, STATS, S>(L, nodes, st, overflow, tc);
                    if (L.ref < 0 && L.leaf >= 0) { searching = false; L.leaf = L.ref; L.ref = popNext<S>(L, st); }
                    if (!__any(searching)) break;
                }
                bool answered = false;
                while (L.leaf < 0) {
                    if (testLeaf<MIXED, STATS>(L, L.leaf, recs, tc)) { answered = true; break; }
                    L.leaf = L.ref;
                    if (L.ref < 0) L.ref = popNext<S>(L, st);
                }
                if (answered) { L.ref = kDone; L.leaf = 0; }
                if (L.ref == kDone && L.leaf >= 0) {
                    finishRay(a, L);
                    L.index = -1;
                    break;
                }
                if (REFILL && !deal.exhausted && __popcll(__ballot(true)) < a.refillBelow) break;
            }
        }
        if (!WW && L.index >= 0) {
            for (;;) {
                if (traceStep<N, MIXED, STATS, S, WT>(L, nodes, recs, st, overflow, tc)) {
                    finishRay(a, L);
                    L.index = -1;
                    break;
                }
                if (REFILL && !deal.exhausted && __popcll(__ballot(true)) < a.refillBelow) break;
            }
        }
    }
    if (STATS) {
        atomicAdd(&a.stats[0], tc.nodesC);
        atomicAdd(&a.stats[1], tc.trisC);
        atomicAdd(&a.stats[2], tc.nodesA);
        atomicAdd(&a.stats[3], tc.trisA);
    }
    if (overflow) atomicOr(a.overflow, 1);
}

template <class N, int S, bool R, bool WW, bool WT = false>
static void launchTraceNS(bool stats, int blocks, const TraceArgs& a, hipStream_t s) {
    dim3 grid(blocks), block(kBlock);
    const bool mixed = a.anyRays != nullptr;
    if (mixed) {
        if (stats) hipLaunchKernelGGL((traceKernel<N, true, true, S, R, WW, WT>), grid, block, 0, s, a);
        else hipLaunchKernelGGL((traceKernel<N, true, false, S, R, WW, WT>), grid, block, 0, s, a);
    } else {
        if (stats) hipLaunchKernelGGL((traceKernel<N, false, true, S, R, WW, WT>), grid, block, 0, s, a);
        else hipLaunchKernelGGL((traceKernel<N, false, false, S, R, WW, WT>), grid, block, 0, s, a);
    }
}
template <class N>
static hipError_t launchTraceN(bool stats, int blocks, const TraceLaunch& l, const TraceArgs& a, hipStream_t s) {
    // The 12-entry ring exists for 4-wide nodes; other layouts take 16 entries for it.
    if (a.watertight) {   // one loop shape: the loop and stack change the work, never the hits
        launchTraceNS<N, 16, false, false, true>(stats, blocks, a, s);
        return hipGetLastError();
    }
    const int stack = (l.stack == 12 && !std::is_same<N, WideNode<4>>::value) ? 16 : l.stack;
    const int key = (stack == 16 ? 4 : stack == 8 ? 0 : stack == 12 ? 8 : -100) + (l.refill ? 2 : 0) + (l.whileWhile ? 1 : 0);
    switch (key) {
        case 0: launchTraceNS<N, 8, false, false>(stats, blocks, a, s); break;
        case 1: launchTraceNS<N, 8, false, true>(stats, blocks, a, s); break;
        case 2: launchTraceNS<N, 8, true, false>(stats, blocks, a, s); break;
        case 3: launchTraceNS<N, 8, true, true>(stats, blocks, a, s); break;
        case 4: launchTraceNS<N, 16, false, false>(stats, blocks, a, s); break;
        case 5: launchTraceNS<N, 16, false, true>(stats, blocks, a, s); break;
        case 6: launchTraceNS<N, 16, true, false>(stats, blocks, a, s); break;
        case 7: launchTraceNS<N, 16, true, true>(stats, blocks, a, s); break;
        case 8: launchTraceNS<WideNode<4>, 12, false, false>(stats, blocks, a, s); break;
        case 9: launchTraceNS<WideNode<4>, 12, false, true>(stats, blocks, a, s); break;
        case 10: launchTraceNS<WideNode<4>, 12, true, false>(stats, blocks, a, s); break;
        case 11: launchTraceNS<WideNode<4>, 12, true, true>(stats, blocks, a, s); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
hipError_t launchTrace(const TraceLaunch& l, bool stats, int blocks, const TraceArgs& args, hipStream_t s) {
    if (l.refillBelow < 1 || l.refillBelow > 64) return hipErrorInvalidValue;
    if ((!args.rays && (args.countPtr || args.count)) || (!args.anyRays && (args.anyCountPtr || args.anyCount)))
        return hipErrorInvalidValue;
    TraceArgs a = args;
    a.refillBelow = l.refillBelow;
    if (l.quantized) {
        switch (l.width) {
            case 2: return launchTraceN<QNode<2>>(stats, blocks, l, a, s);
            case 4: return launchTraceN<QNode<4>>(stats, blocks, l, a, s);
            case 8: return launchTraceN<QNode<8>>(stats, blocks, l, a, s);
            default: return hipErrorInvalidValue;
        }
    }
    switch (l.width) {
        case 2: return launchTraceN<WideNode<2>>(stats, blocks, l, a, s);
        case 4: return launchTraceN<WideNode<4>>(stats, blocks, l, a, s);
        case 8: return launchTraceN<WideNode<8>>(stats, blocks, l, a, s);
        default: return hipErrorInvalidValue;
    }
}

// ------------------------------------------------------------------ frame kernels
__global__ void resetCountersKernel(int* c, int initialRays, int* xcdWork) {
    for (int k = threadIdx.x; k < kCounterStride; k += blockDim.x) c[k] = 0;
    if (xcdWork)
        for (int k = threadIdx.x; k < kXcdSlotInts; k += blockDim.x) xcdWork[k] = 0;
    __syncthreads();
    if (threadIdx.x == 0) c[kCntRays] = initialRays;
}
hipError_t launchResetCounters(int* counters, int initialRays, int* xcdWork, hipStream_t s) {
    hipLaunchKernelGGL(resetCountersKernel, dim3(1), dim3(kBlock), 0, s, counters, initialRays, xcdWork);
    return hipGetLastError();
}

// Sobol terms shared by a launch: shift[k] depends on (seed, dim), term[r][k] on
// (sample, dim); a lane combines them with its pixel hash (Native/sampling.cuh:26-35).
template <int K>
__device__ __forceinline__ void loadSobolTerms(unsigned* sShift, unsigned* sTerm, unsigned seed, unsigned dimBase,
                                               unsigned sample0, int R) {
    const int tid = threadIdx.x;
    if (tid < K) sShift[tid] = dimShift(seed, dimBase + tid);
    if (tid < R * K) sTerm[tid] = grayTerm(gSobolDirs, dimBase + (tid % K), sample0 + tid / K);
    __syncthreads();
}
// The same for a number of dimensions known at launch: shifts of dimensions [dimBase,
// dimBase + dims), terms of samples sample0 .. sample0 + R - 1 at sTerm[r * dims + k].
__device__ __forceinline__ void loadSobolTermsN(unsigned* sShift, unsigned* sTerm, unsigned seed, unsigned dimBase,
                                                int dims, unsigned sample0, int R) {
    for (int i = threadIdx.x; i < dims; i += blockDim.x) sShift[i] = dimShift(seed, dimBase + (unsigned)i);
    for (int i = threadIdx.x; i < R * dims; i += blockDim.x)
        sTerm[i] = grayTerm(gSobolDirs, dimBase + (unsigned)(i % dims), sample0 + (unsigned)(i / dims));
    __syncthreads();
}
// Direction-number rows [dimBase, dimBase + rows) of the table in LDS, for launches whose lanes
// sample different (seed, sample) pairs (batched views): a lane forms its own shift and term
// from them.
__device__ __forceinline__ void loadSobolRows(unsigned* sDirs, unsigned dimBase, int rows) {
    for (int i = threadIdx.x; i < rows * 32; i += blockDim.x) sDirs[i] = gSobolDirs[dimBase * 32 + (unsigned)i];
    __syncthreads();
}

// Bijection on [0, w*h): queue position -> pixel, visiting 8x8 tiles first (each wave of
// primary rays covers one compact tile), then the right and bottom remainder strips.
__device__ __forceinline__ int tiledPixel(int q, int w, int h) {
    const int tw = w / 8, th = h / 8;
    const int nA = tw * th * 64;
    if (q < nA) {
        const int t = q >> 6, k = q & 63;
        return ((t / tw) * 8 + (k >> 3)) * w + (t % tw) * 8 + (k & 7);
    }
    q -= nA;
    const int rw = w - tw * 8, nR = rw * th * 8;
    if (q < nR) return (q / rw) * w + tw * 8 + q % rw;
    q -= nR;
    return (th * 8 + q / w) * w + q % w;
}

__global__ __launch_bounds__(kBlock) void raygenKernel(FrameArgs f) {
    __shared__ unsigned sShift[4], sTerm[kMaxRaysPerPixel * 4];
    loadSobolTerms<4>(sShift, sTerm, f.rs.seed, 0, f.sample0, f.R);
    const int l = blockIdx.x * kBlock + threadIdx.x;   // position in this pipeline's queue
    if (l >= f.qCount) return;
    const int q = f.qBase + l;                          // position in the chunk's order
    const int r = q / f.npix;
    const unsigned pixel = (unsigned)(f.tiled ? tiledPixel(q - r * f.npix, f.w, f.h) : q - r * f.npix);
    const int slot = r * f.npix + (int)pixel;
    const int x = (int)(pixel % (unsigned)f.w), y = (int)(pixel / (unsigned)f.w);
    auto rng = [&](int k) { return sobolFromParts(pixel, sShift[k], sTerm[r * 4 + k]); };
    f3 o, d;
    cameraRay(f.cam, f.rs, x, y, f.w, f.h, rng, o, d);
    f.thr[slot] = make_float4(1, 1, 1, 0);
    f.rad[slot] = make_float4(0, 0, 0, 0);
    f.alb[slot] = make_float4(0, 0, 0, 0);
    f.nrm[slot] = make_float4(0, 0, 0, 0);
    f.rayOut[2 * l] = make_float4(o.x, o.y, o.z, 1e30f);
    f.rayOut[2 * l + 1] = make_float4(d.x, d.y, d.z, __int_as_float(slot));
}
hipError_t launchRaygen(const FrameArgs& f, hipStream_t s) {
    const int n = f.qCount;
    if (n < 1 || f.qBase < 0 || f.qBase + n > f.R * f.npix) return hipErrorInvalidValue;
    hipLaunchKernelGGL(raygenKernel, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, f);
    return hipGetLastError();
}

// Appends one 32-byte ray per predicated lane to each of two queues, with one atomic per
// workgroup and queue.  Every thread of the workgroup calls it once per round of the
// caller's grid-stride loop.  The per-wave counts alternate between two LDS buffers by
// round: a wave can only write round k+2's counts after every wave has passed round k+1's
// first barrier, which each wave reaches after it has read round k's counts.
struct AppendLds {
    int count[2][2][kBlock / 64];   // [round & 1][queue][wave]
    int base[2];
};
__device__ __forceinline__ void appendRays2(AppendLds& sh, int round, bool p0, int* c0, float4* q0, float4 a0,
                                            float4 b0, bool p1, int* c1, float4* q1, float4 a1, float4 b1) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const unsigned long long m0 = __ballot(p0), m1 = __ballot(p1);
    int(*cnt)[kBlock / 64] = sh.count[round & 1];
    if (lane == 0) {
        cnt[0][wave] = __popcll(m0);
        cnt[1][wave] = __popcll(m1);
    }
    __syncthreads();
    if (tid < 2) {
        int t = 0;
        for (int w = 0; w < kBlock / 64; w++) t += cnt[tid][w];
        sh.base[tid] = t ? atomicAdd(tid ? c1 : c0, t) : 0;
    }
    __syncthreads();
    int o0 = sh.base[0], o1 = sh.base[1];
    for (int w = 0; w < wave; w++) {
        o0 += cnt[0][w];
        o1 += cnt[1][w];
    }
    if (p0) {
        const int off = o0 + (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m0 >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m0, 0u));
        q0[2 * off] = a0;
        q0[2 * off + 1] = b0;
    }
    if (p1) {
        const int off = o1 + (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m1 >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m1, 0u));
        q1[2 * off] = a1;
        q1[2 * off + 1] = b1;
    }
}

// Grid-stride over the bounce's queue, so the grid is sized to the GPU, not to the largest
// possible queue; every lane of a wave reaches the appends in each round.  Capped at 96
// VGPRs so five workgroups fit on a compute unit without spilling.
#ifndef RT_SHADE_WAVES
#define RT_SHADE_WAVES 5
#endif
constexpr int kShadeBlocksPerCU = RT_SHADE_WAVES;
__global__ __launch_bounds__(kBlock, kShadeBlocksPerCU) void shadeKernel(FrameArgs f) {
    __shared__ unsigned sShift[8], sTerm[kMaxRaysPerPixel * 8];
    __shared__ AppendLds sApp;
    loadSobolTerms<8>(sShift, sTerm, f.rs.seed, 4 + 8 * f.bounce, f.sample0, f.R);
    const int count = f.counters[kCntRays + f.bounce];
    int round = 0;
    for (int base = blockIdx.x * kBlock; base < count; base += gridDim.x * kBlock, round++) {
        const int i = base + threadIdx.x;
        bool wantShadow = false, wantNext = false;
        float4 sa = make_float4(0, 0, 0, 0), sb = sa, na = sa, nb = sa;
        if (i < count) {
            const float4 r0 = f.rayIn[2 * i], r1 = f.rayIn[2 * i + 1];
            const int slot = __float_as_int(r1.w);
            const int r = slot / f.npix;
            const unsigned pixel = (unsigned)(slot - r * f.npix);
            const int4 hv = f.hits[i];
            RtHit hit;
            hit.index = hv.x; hit.distance = __int_as_float(hv.y); hit.u = __int_as_float(hv.z); hit.v = __int_as_float(hv.w);
            const float4 t4 = f.thr[slot], q4 = f.rad[slot];
            PathState ps;
            ps.thr = F3(t4.x, t4.y, t4.z); ps.prevPdf = t4.w; ps.rad = F3(q4.x, q4.y, q4.z);
            FirstHit fh; fh.covered = false;
            auto rng = [&](int k) { return sobolFromParts(pixel, sShift[k], sTerm[r * 8 + k]); };
            ShadeResult res;
            shadeBounce(f.sc, f.rs, f.bounce, rng, F3(r0.x, r0.y, r0.z), F3(r1.x, r1.y, r1.z), hit, ps, &fh, res);
            f.thr[slot] = make_float4(ps.thr.x, ps.thr.y, ps.thr.z, ps.prevPdf);
            f.rad[slot] = make_float4(ps.rad.x, ps.rad.y, ps.rad.z, q4.w);
            if (fh.covered) {
                f.alb[slot] = make_float4(fh.albedo.x, fh.albedo.y, fh.albedo.z, 1.0f);
                f.nrm[slot] = make_float4(fh.normal.x, fh.normal.y, fh.normal.z, fh.depth);
            }
            if (res.shadow) {
                wantShadow = true;
                f.nee[slot] = make_float4(res.contrib.x, res.contrib.y, res.contrib.z, 0);
                sa = make_float4(res.nextO.x, res.nextO.y, res.nextO.z, res.slimit);
                sb = make_float4(res.sdir.x, res.sdir.y, res.sdir.z, __int_as_float(slot));
            }
            if (res.cont) {
                wantNext = true;
                na = make_float4(res.nextO.x, res.nextO.y, res.nextO.z, 1e30f);
                nb = make_float4(res.nextD.x, res.nextD.y, res.nextD.z, __int_as_float(slot));
            }
        }
        appendRays2(sApp, round, wantShadow, &f.counters[kCntShadow + f.bounce], f.shadow, sa, sb, wantNext,
                    &f.counters[kCntRays + f.bounce + 1], f.rayOut, na, nb);
    }
}
#ifndef RT_SHADE_GRID_PER_CU
#define RT_SHADE_GRID_PER_CU kShadeBlocksPerCU
#endif
int shadeBlocksPerCU() { return RT_SHADE_GRID_PER_CU; }
hipError_t launchShade(const FrameArgs& f, int blocks, hipStream_t s) {
    const int n = f.R * f.npix;
    const int need = (n + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(shadeKernel, dim3(need < blocks ? need : blocks), dim3(kBlock), 0, s, f);
    return hipGetLastError();
}

// ------------------------------------------------------------------ path completion
// Late bounces hold few paths, and a launch per bounce ends only when its slowest ray does,
// so a frame's tail is a chain of latency-bound launches.  From bounce firstBounce on, this
// kernel carries each path through all its remaining bounces inside one launch: a lane
// traces the closest hit, shades it (shadeBounce), traces the shadow ray and adds its term,
// and continues with the next ray.  The lanes of a wave move through these phases together.
// Each path sees the same operations in the same order as in the per-bounce kernels (its
// emission, then its next-event term, bounce by bounce), so results are bit-identical.
__device__ __forceinline__ void beginRay(LaneRay& L, f3 o, f3 d, float limit, bool any) {
    L.o = o; L.d = d;
    L.inv = safeInverse(d);
    L.nz = L.inv.x != 0 && L.inv.y != 0 && L.inv.z != 0;
    L.oi = F3(o.x * L.inv.x, o.y * L.inv.y, o.z * L.inv.z);
    L.best = limit; L.bu = 0; L.bv = 0;
    L.bestIdx = -1; L.sp = 0; L.ref = 0; L.leaf = 0;
    L.any = any;
}

// With the 8- and 12-entry stacks, capped at 96 VGPRs (five workgroups per CU; a few spilled
// values) instead of 109 (four): +2% on the icosphere.  The 16-entry stack's LDS allows four.
#ifndef RT_TAIL_WAVES
#define RT_TAIL_WAVES 5
#endif
template <int S> constexpr int tailBlocksPerCU() { return S == 16 ? 4 : RT_TAIL_WAVES; }
template <class N, bool STATS, int S>
__global__ __launch_bounds__(kBlock, tailBlocksPerCU<S>()) void pathTailKernel(FrameArgs f, TraceArgs a, int firstBounce) {
    __shared__ int sRef[S * kBlock];
    __shared__ float sDist[S * kBlock];
    const int tid = threadIdx.x, lane = tid & 63;
    const int total = *a.countPtr;
    const N* nodes = static_cast<const N*>(a.nodes);
    const float4* recs = reinterpret_cast<const float4*>(a.recs);
    LdsStack<S> st{sRef, sDist, reinterpret_cast<int2*>(a.spill), a.spillEntries, a.totalLanes,
                   (int)(blockIdx.x * kBlock + tid), tid};
    // Sobol shifts of the launch's dimensions (bounces firstBounce .. maxBounces - 1, 8 each),
    // then their terms for the chunk's R samples; sized at launch (tailSobolBytes).
    extern __shared__ unsigned sSobol[];
    const int tailDims = 8 * (f.rs.maxBounces - firstBounce);
    loadSobolTermsN(sSobol, sSobol + tailDims, f.rs.seed, 4 + 8 * (unsigned)firstBounce, tailDims, f.sample0, f.R);
    TraceCounters tc;
    bool overflow = false;
    LaneRay L;
    PathState ps;
    unsigned moreClosest = 0, shadowRays = 0;   // queries beyond the input queue, for the frame counters
    int bounce = 0, slot = 0, r = 0;
    unsigned pixel = 0;
    float radW = 0;
    bool active = false;
    auto startPath = [&](int i) {
        const float4 r0 = a.rays[2 * i], r1 = a.rays[2 * i + 1];
        slot = __float_as_int(r1.w);
        r = slot / f.npix;
        pixel = (unsigned)(slot - r * f.npix);
        const float4 t4 = f.thr[slot], q4 = f.rad[slot];
        ps.thr = F3(t4.x, t4.y, t4.z); ps.prevPdf = t4.w; ps.rad = F3(q4.x, q4.y, q4.z);
        radW = q4.w;
        bounce = firstBounce;
        beginRay(L, F3(r0.x, r0.y, r0.z), F3(r1.x, r1.y, r1.z), r0.w, false);
        active = true;
    };
    RayDealer deal;
    deal.init(a.xcdWork, a.work, total);
    {
        const int i = deal.firstIndex(tid);
        if (i >= 0) startPath(i);
    }
    for (;;) {
        const unsigned long long idle = __ballot(!active);
        if (idle && !deal.exhausted) {
            const int base = deal.take(idle, lane);
            if (!active) {
                const int i = base + (int)__builtin_amdgcn_mbcnt_hi((unsigned)(idle >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)idle, 0u));
                if (i < deal.hi) startPath(i);
            }
            deal.after(base, idle);
        }
        if (__ballot(active) == 0) {
            if (deal.exhausted) break;
            continue;
        }
        // closest hit of this bounce
        if (active) while (!traceStep<N, true, STATS, S>(L, nodes, recs, st, overflow, tc)) {}
        ShadeResult res;
        res.cont = false; res.shadow = false;
        if (active) {
            RtHit hit;
            hit.index = L.bestIdx; hit.distance = L.best; hit.u = L.bu; hit.v = L.bv;
            const int kb = 8 * (bounce - firstBounce);
            const unsigned* terms = sSobol + tailDims + r * tailDims + kb;
            auto rng = [&](int k) { return sobolFromParts(pixel, sSobol[kb + k], terms[k]); };
            FirstHit unused;   // first-hit outputs belong to bounce 0
            shadeBounce(f.sc, f.rs, bounce, rng, L.o, L.d, hit, ps, &unused, res);
        }
        // next-event shadow ray
        if (active && res.shadow) {
            shadowRays++;
            beginRay(L, res.nextO, res.sdir, res.slimit, true);
            while (!traceStep<N, true, STATS, S>(L, nodes, recs, st, overflow, tc)) {}
            if (L.bestIdx < 0) { ps.rad.x = ps.rad.x + res.contrib.x; ps.rad.y = ps.rad.y + res.contrib.y; ps.rad.z = ps.rad.z + res.contrib.z; }
        }
        if (active) {
            if (res.cont) {
                moreClosest++;
                beginRay(L, res.nextO, res.nextD, 1e30f, false);
                bounce++;
            } else {
                f.thr[slot] = make_float4(ps.thr.x, ps.thr.y, ps.thr.z, ps.prevPdf);
                f.rad[slot] = make_float4(ps.rad.x, ps.rad.y, ps.rad.z, radW);
                active = false;
            }
        }
    }
    // Query counts into the frame counters (their per-bounce split is not kept): one atomic
    // per workgroup and counter.
    __shared__ unsigned sCount[2];
    if (tid < 2) sCount[tid] = 0;
    __syncthreads();
    for (int o = 32; o > 0; o >>= 1) {
        moreClosest += __shfl_xor(moreClosest, o, 64);
        shadowRays += __shfl_xor(shadowRays, o, 64);
    }
    if (lane == 0) {
        atomicAdd(&sCount[0], moreClosest);
        atomicAdd(&sCount[1], shadowRays);
    }
    __syncthreads();
    if (tid == 0 && sCount[0]) atomicAdd(&f.counters[kCntRays + firstBounce + 1], (int)sCount[0]);
    if (tid == 1 && sCount[1]) atomicAdd(&f.counters[kCntShadow + firstBounce], (int)sCount[1]);
    if (STATS) {
        atomicAdd(&a.stats[0], tc.nodesC);
        atomicAdd(&a.stats[1], tc.trisC);
        atomicAdd(&a.stats[2], tc.nodesA);
        atomicAdd(&a.stats[3], tc.trisA);
    }
    if (overflow) atomicOr(a.overflow, 1);
}

// The grid is capped at the workgroups that fit on the GPU at once: a lane's first path is
// fixed by its grid position, so a workgroup that had to wait for another to finish would
// start its paths late.
template <class N, int S>
static hipError_t launchTailNS(bool stats, int blocks, int cus, const FrameArgs& f, const TraceArgs& a, int firstBounce,
                               hipStream_t s) {
    const void* k = stats ? reinterpret_cast<const void*>(&pathTailKernel<N, true, S>)
                          : reinterpret_cast<const void*>(&pathTailKernel<N, false, S>);
    if (firstBounce < 1 || firstBounce >= f.rs.maxBounces || f.R < 1 || f.R > kMaxRaysPerPixel) return hipErrorInvalidValue;
    const size_t sobolBytes = (size_t)8 * (f.rs.maxBounces - firstBounce) * (1 + f.R) * sizeof(unsigned);
    int perCU = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, k, kBlock, sobolBytes);
    if (e != hipSuccess) return e;
    blocks = std::max(1, std::min(blocks, perCU * cus));
    if (stats) hipLaunchKernelGGL((pathTailKernel<N, true, S>), dim3(blocks), dim3(kBlock), sobolBytes, s, f, a, firstBounce);
    else hipLaunchKernelGGL((pathTailKernel<N, false, S>), dim3(blocks), dim3(kBlock), sobolBytes, s, f, a, firstBounce);
    return hipGetLastError();
}
template <class N>
static hipError_t launchTailN(bool stats, int blocks, int cus, const TraceLaunch& l, const FrameArgs& f,
                              const TraceArgs& a, int firstBounce, hipStream_t s) {
    if (l.stack == 16) return launchTailNS<N, 16>(stats, blocks, cus, f, a, firstBounce, s);
    if (l.stack == 8) return launchTailNS<N, 8>(stats, blocks, cus, f, a, firstBounce, s);
    if (l.stack == 12) {
        if constexpr (std::is_same<N, WideNode<4>>::value) return launchTailNS<N, 12>(stats, blocks, cus, f, a, firstBounce, s);
        return launchTailNS<N, 16>(stats, blocks, cus, f, a, firstBounce, s);   // the 12-entry ring is 4-wide only
    }
    return hipErrorInvalidValue;
}
hipError_t launchPathTail(const TraceLaunch& l, bool stats, int blocks, int cus, const FrameArgs& f, const TraceArgs& a,
                          int firstBounce, hipStream_t s) {
    if (!a.rays || !a.countPtr || a.anyRays || firstBounce < 1 || firstBounce >= f.rs.maxBounces || a.roots)
        return hipErrorInvalidValue;
    if (l.quantized) {
        switch (l.width) {
            case 2: return launchTailN<QNode<2>>(stats, blocks, cus, l, f, a, firstBounce, s);
            case 4: return launchTailN<QNode<4>>(stats, blocks, cus, l, f, a, firstBounce, s);
            case 8: return launchTailN<QNode<8>>(stats, blocks, cus, l, f, a, firstBounce, s);
            default: return hipErrorInvalidValue;
        }
    }
    switch (l.width) {
        case 2: return launchTailN<WideNode<2>>(stats, blocks, cus, l, f, a, firstBounce, s);
        case 4: return launchTailN<WideNode<4>>(stats, blocks, cus, l, f, a, firstBounce, s);
        case 8: return launchTailN<WideNode<8>>(stats, blocks, cus, l, f, a, firstBounce, s);
        default: return hipErrorInvalidValue;
    }
}

// Native/renderer.cu:292-301: add the R paths of each pixel in sample order, then tone-map.
__global__ __launch_bounds__(kBlock) void accumulateKernel(AccumArgs a) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= a.npix) return;
    float4 sr = make_float4(0, 0, 0, 0), sa = sr, sn = sr;
    if (!a.startFromZero) { sr = a.sumRad[i]; sa = a.sumAlb[i]; sn = a.sumNrm[i]; }
    for (int r = 0; r < a.R; r++) {
        const size_t slot = (size_t)r * a.npix + i;
        const float4 pr = a.rad[slot], pa = a.alb[slot], pn = a.nrm[slot];
        sr.x = sr.x + pr.x; sr.y = sr.y + pr.y; sr.z = sr.z + pr.z;
        sa.x = sa.x + pa.x; sa.y = sa.y + pa.y; sa.z = sa.z + pa.z;
        sn.x = sn.x + pn.x; sn.y = sn.y + pn.y; sn.z = sn.z + pn.z;
        sr.w = sr.w + pa.w;   // covered count
        sn.w = sn.w + pn.w;   // first-hit depth sum
        const float lum = 0.2126f * pr.x + 0.7152f * pr.y + 0.0722f * pr.z;
        sa.w = sa.w + lum * lum;   // sum of squared path luminance (variance estimate)
    }
    a.sumRad[i] = sr; a.sumAlb[i] = sa; a.sumNrm[i] = sn;
    if (a.writeDisplay) {
        const float e = a.exposure;
        a.pixels[i] = make_uchar4(displayByte(sr.x * a.invSamples, e), displayByte(sr.y * a.invSamples, e),
                                  displayByte(sr.z * a.invSamples, e), 255);
    }
}
hipError_t launchAccumulate(const AccumArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(accumulateKernel, dim3((a.npix + kBlock - 1) / kBlock), dim3(kBlock), 0, s, a);
    return hipGetLastError();
}

// ------------------------------------------------------------------ batched views
// The same stages as the single-view frame, with each path slot's view looked up from its
// slot; Sobol values are formed per lane (views have their own seeds and sample counts).
__global__ __launch_bounds__(kBlock) void batchRaygenKernel(BatchArgs b) {
    __shared__ unsigned sDirs[4 * 32];   // direction numbers of the camera dimensions 0..3
    loadSobolRows(sDirs, 0, 4);
    const int l = blockIdx.x * kBlock + threadIdx.x;   // position in this pipeline's queue
    if (l >= b.qCount) return;
    const int q = b.qBase + l;                          // position in the frame's order
    const int v = findView(b.viewBase, b.viewCount, q);
    const BatchView& V = b.views[v];
    const unsigned pixel = (unsigned)tiledPixel(q - V.slotBase, V.w, V.h);
    const int slot = V.slotBase + (int)pixel;
    const int x = (int)(pixel % (unsigned)V.w), y = (int)(pixel / (unsigned)V.w);
    const unsigned sample = b.samples[v];
    auto rng = [&](int k) { return sobolFromParts(pixel, dimShift(V.rs.seed, (unsigned)k), grayTerm(sDirs, (unsigned)k, sample)); };
    f3 o, d;
    cameraRay(V.cam, V.rs, x, y, V.w, V.h, rng, o, d);
    b.thr[slot] = make_float4(1, 1, 1, 0);
    b.rad[slot] = make_float4(0, 0, 0, 0);
    b.alb[slot] = make_float4(0, 0, 0, 0);
    b.nrm[slot] = make_float4(0, 0, 0, 0);
    b.rayOut[2 * l] = make_float4(o.x, o.y, o.z, 1e30f);
    b.rayOut[2 * l + 1] = make_float4(d.x, d.y, d.z, __int_as_float(slot));
}
hipError_t launchBatchRaygen(const BatchArgs& b, hipStream_t s) {
    if (b.qCount < 1 || b.qBase < 0 || b.qBase + b.qCount > b.totalSlots) return hipErrorInvalidValue;
    hipLaunchKernelGGL(batchRaygenKernel, dim3((b.qCount + kBlock - 1) / kBlock), dim3(kBlock), 0, s, b);
    return hipGetLastError();
}

__global__ __launch_bounds__(kBlock, kShadeBlocksPerCU) void batchShadeKernel(BatchArgs b) {
    __shared__ AppendLds sApp;
    __shared__ unsigned sDirs[8 * 32];   // direction numbers of this bounce's dimensions
    const unsigned dimBase = 4 + 8 * (unsigned)b.bounce;
    loadSobolRows(sDirs, dimBase, 8);
    const int count = b.counters[kCntRays + b.bounce];
    int round = 0;
    for (int base = blockIdx.x * kBlock; base < count; base += gridDim.x * kBlock, round++) {
        const int i = base + threadIdx.x;
        bool wantShadow = false, wantNext = false;
        float4 sa = make_float4(0, 0, 0, 0), sb = sa, na = sa, nb = sa;
        if (i < count) {
            const float4 r0 = b.rayIn[2 * i], r1 = b.rayIn[2 * i + 1];
            const int slot = __float_as_int(r1.w);
            const int v = findView(b.viewBase, b.viewCount, slot);
            const BatchView& V = b.views[v];
            const unsigned pixel = (unsigned)(slot - V.slotBase);
            const unsigned sample = b.samples[v];
            const int4 hv = b.hits[i];
            RtHit hit;
            hit.index = hv.x; hit.distance = __int_as_float(hv.y); hit.u = __int_as_float(hv.z); hit.v = __int_as_float(hv.w);
            const float4 t4 = b.thr[slot], q4 = b.rad[slot];
            PathState ps;
            ps.thr = F3(t4.x, t4.y, t4.z); ps.prevPdf = t4.w; ps.rad = F3(q4.x, q4.y, q4.z);
            FirstHit fh; fh.covered = false;
            auto rng = [&](int k) {
                return sobolFromParts(pixel, dimShift(V.rs.seed, dimBase + (unsigned)k), grayTerm(sDirs, (unsigned)k, sample));
            };
            ShadeResult res;
            shadeBounce(V.sc, V.rs, b.bounce, rng, F3(r0.x, r0.y, r0.z), F3(r1.x, r1.y, r1.z), hit, ps, &fh, res);
            b.thr[slot] = make_float4(ps.thr.x, ps.thr.y, ps.thr.z, ps.prevPdf);
            b.rad[slot] = make_float4(ps.rad.x, ps.rad.y, ps.rad.z, q4.w);
            if (fh.covered) {
                b.alb[slot] = make_float4(fh.albedo.x, fh.albedo.y, fh.albedo.z, 1.0f);
                b.nrm[slot] = make_float4(fh.normal.x, fh.normal.y, fh.normal.z, fh.depth);
            }
            if (res.shadow) {
                wantShadow = true;
                b.nee[slot] = make_float4(res.contrib.x, res.contrib.y, res.contrib.z, 0);
                sa = make_float4(res.nextO.x, res.nextO.y, res.nextO.z, res.slimit);
                sb = make_float4(res.sdir.x, res.sdir.y, res.sdir.z, __int_as_float(slot));
            }
            if (res.cont) {
                wantNext = true;
                na = make_float4(res.nextO.x, res.nextO.y, res.nextO.z, 1e30f);
                nb = make_float4(res.nextD.x, res.nextD.y, res.nextD.z, __int_as_float(slot));
            }
        }
        appendRays2(sApp, round, wantShadow, &b.counters[kCntShadow + b.bounce], b.shadow, sa, sb, wantNext,
                    &b.counters[kCntRays + b.bounce + 1], b.rayOut, na, nb);
    }
}
hipError_t launchBatchShade(const BatchArgs& b, int blocks, hipStream_t s) {
    const int need = (b.totalSlots + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(batchShadeKernel, dim3(need < blocks ? need : blocks), dim3(kBlock), 0, s, b);
    return hipGetLastError();
}

__global__ __launch_bounds__(kBlock) void batchAccumulateKernel(BatchArgs b) {
    const int slot = blockIdx.x * kBlock + threadIdx.x;
    if (slot >= b.totalSlots) return;
    const int v = findView(b.viewBase, b.viewCount, slot);
    const BatchView& V = b.views[v];
    const int i = slot - V.slotBase;
    const unsigned sample = b.samples[v];
    float4 sr = make_float4(0, 0, 0, 0), sa = sr, sn = sr;
    if (sample) { sr = V.sumRad[i]; sa = V.sumAlb[i]; sn = V.sumNrm[i]; }
    const float4 pr = b.rad[slot], pa = b.alb[slot], pn = b.nrm[slot];
    sr.x = sr.x + pr.x; sr.y = sr.y + pr.y; sr.z = sr.z + pr.z;
    sa.x = sa.x + pa.x; sa.y = sa.y + pa.y; sa.z = sa.z + pa.z;
    sn.x = sn.x + pn.x; sn.y = sn.y + pn.y; sn.z = sn.z + pn.z;
    sr.w = sr.w + pa.w;
    sn.w = sn.w + pn.w;
    const float lum = 0.2126f * pr.x + 0.7152f * pr.y + 0.0722f * pr.z;
    sa.w = sa.w + lum * lum;
    V.sumRad[i] = sr; V.sumAlb[i] = sa; V.sumNrm[i] = sn;
    const float inv = 1.0f / (float)(sample + 1u);
    const float e = V.rs.exposure;
    V.pixels[i] = make_uchar4(displayByte(sr.x * inv, e), displayByte(sr.y * inv, e), displayByte(sr.z * inv, e), 255);
}
hipError_t launchBatchAccumulate(const BatchArgs& b, hipStream_t s) {
    hipLaunchKernelGGL(batchAccumulateKernel, dim3((b.totalSlots + kBlock - 1) / kBlock), dim3(kBlock), 0, s, b);
    return hipGetLastError();
}

__global__ void batchAdvanceKernel(unsigned* samples, int views) {
    const int v = blockIdx.x * kBlock + threadIdx.x;
    if (v < views) samples[v] += 1u;
}
hipError_t launchBatchAdvance(const BatchArgs& b, hipStream_t s) {
    hipLaunchKernelGGL(batchAdvanceKernel, dim3((b.viewCount + kBlock - 1) / kBlock), dim3(kBlock), 0, s, b.samples,
                       b.viewCount);
    return hipGetLastError();
}

// ------------------------------------------------------------------ denoiser stages
// Native/renderer.cu:422-428 (HDR colour and albedo averages, unit normal sum) plus the
// mean first-hit depth; the image is edge-extended to the padded size.
__global__ __launch_bounds__(kBlock) void denoisePrepareKernel(DenoiseArgs a) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= a.wp * a.hp) return;
    const int x = min(i % a.wp, a.w - 1), y = min(i / a.wp, a.h - 1);
    const int src = y * a.w + x;
    const float inv = 1.0f / a.samples;
    const float4 r = a.sumRad[src], al = a.sumAlb[src], n = a.sumNrm[src];
    const f3 nu = unitv(F3(n.x, n.y, n.z));
    float2* o = reinterpret_cast<float2*>(a.planes + (size_t)i * 10);   // 40 bytes: five 8-byte stores
    o[0] = make_float2(r.x * inv, r.y * inv);
    o[1] = make_float2(r.z * inv, al.x * inv);
    o[2] = make_float2(al.y * inv, al.z * inv);
    o[3] = make_float2(nu.x, nu.y);
    o[4] = make_float2(nu.z, r.w > 0 ? n.w / r.w : 0.0f);
}
__global__ void mirrorPlanesKernel(const float* in, float* out, int hp, int wp, int channels, int flip) {
    const long long i = blockIdx.x * (long long)kBlock + threadIdx.x;
    if (i >= (long long)hp * wp * channels) return;
    const int c = (int)(i % channels);
    const long long p = i / channels;
    const int x = (int)(p % wp), y = (int)(p / wp);
    const int sx = (flip & 1) ? wp - 1 - x : x, sy = (flip & 2) ? hp - 1 - y : y;
    out[i] = in[((size_t)sy * wp + sx) * channels + c];
}
hipError_t launchMirrorPlanes(const float* in, float* out, int hp, int wp, int channels, int flip, hipStream_t s) {
    const long long n = (long long)hp * wp * channels;
    hipLaunchKernelGGL(mirrorPlanesKernel, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, in, out, hp, wp,
                       channels, flip);
    return hipGetLastError();
}
__global__ void addMirroredKernel(float* acc, const float* img, int hp, int wp, int channels, int flip, float scale) {
    const long long i = blockIdx.x * (long long)kBlock + threadIdx.x;
    if (i >= (long long)hp * wp * channels) return;
    const int c = (int)(i % channels);
    const long long p = i / channels;
    const int x = (int)(p % wp), y = (int)(p / wp);
    const int sx = (flip & 1) ? wp - 1 - x : x, sy = (flip & 2) ? hp - 1 - y : y;
    acc[i] = (acc[i] + img[((size_t)sy * wp + sx) * channels + c]) * scale;
}
hipError_t launchAddMirrored(float* acc, const float* img, int hp, int wp, int channels, int flip, float scale,
                             hipStream_t s) {
    const long long n = (long long)hp * wp * channels;
    hipLaunchKernelGGL(addMirroredKernel, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, acc, img, hp, wp,
                       channels, flip, scale);
    return hipGetLastError();
}

hipError_t launchDenoisePrepare(const DenoiseArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(denoisePrepareKernel, dim3((a.wp * a.hp + kBlock - 1) / kBlock), dim3(kBlock), 0, s, a);
    return hipGetLastError();
}

// One workgroup per 16x16 tile: accept the network's tile when its mean luminance stays
// within 4 standard errors (plus 10%) of the accumulated mean, else show plain
// accumulation; then mask pixels with no primary hit and tone-map
// (Native/renderer.cu:429-435).
__device__ __forceinline__ float waveSum(float v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}
__global__ __launch_bounds__(kBlock) void denoisePresentKernel(DenoiseArgs a) {
    constexpr int kWaves = kBlock / 64;
    __shared__ float sNoisy[kWaves], sVar[kWaves], sDen[kWaves];
    __shared__ int sBad[kWaves];
    __shared__ int sReject;
    const int lx = threadIdx.x & 15, ly = threadIdx.x >> 4;
    const int x = blockIdx.x * 16 + lx, y = blockIdx.y * 16 + ly;
    const bool valid = x < a.w && y < a.h;
    const float inv = 1.0f / a.samples;
    f3 noisy = F3(0, 0, 0), den = F3(0, 0, 0);
    float covered = 0, mu = 0, var = 0, dl = 0;
    bool bad = false;
    if (valid) {
        const int i = y * a.w + x;
        const float4 r = a.sumRad[i], al = a.sumAlb[i];
        noisy = F3(r.x * inv, r.y * inv, r.z * inv);
        covered = r.w;
        const float4 d = *reinterpret_cast<const float4*>(a.denoised + ((size_t)y * a.wp + x) * 4);
        den = F3(d.x, d.y, d.z);
        mu = 0.2126f * noisy.x + 0.7152f * noisy.y + 0.0722f * noisy.z;
        var = fmaxf(al.w * inv - mu * mu, 0.0f) * inv;
        dl = 0.2126f * den.x + 0.7152f * den.y + 0.0722f * den.z;
        bad = !(isfinite(den.x) && isfinite(den.y) && isfinite(den.z));
    }
    // Tile sums: within each wave by lane exchanges, then the four waves' sums in LDS.
    const int wave = threadIdx.x >> 6;
    mu = waveSum(mu);
    var = waveSum(var);
    dl = waveSum(dl);
    const bool anyBad = __any(bad);
    if ((threadIdx.x & 63) == 0) {
        sNoisy[wave] = mu; sVar[wave] = var; sDen[wave] = dl; sBad[wave] = anyBad;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        float tn = 0, tv = 0, td = 0;
        int tb = 0;
#pragma unroll
        for (int k = 0; k < kWaves; k++) { tn += sNoisy[k]; tv += sVar[k]; td += sDen[k]; tb |= sBad[k]; }
        const int nx = min(16, a.w - blockIdx.x * 16), ny = min(16, a.h - blockIdx.y * 16);
        const float cnt = (float)(nx * ny);
        const float mn = tn / cnt, md = td / cnt, se = sqrtf(tv) / cnt;
        int reject = 0;
        if (a.fallback == 1) reject = 1;
        else if (a.fallback == 0) reject = tb || fabsf(md - mn) > 4.0f * se + 0.1f * mn + 1e-4f;
        sReject = reject;
        // rejected tiles, spread over kTileStatSlots counters (one address for every tile
        // would serialise the atomics of thousands of workgroups)
        if (reject) atomicAdd(&a.tileStats[(blockIdx.y * gridDim.x + blockIdx.x) % kTileStatSlots], 1);
    }
    __syncthreads();
    if (!valid) return;
    f3 c = sReject ? noisy : den;
    if (!(isfinite(c.x) && isfinite(c.y) && isfinite(c.z))) c = noisy;
    if (covered == 0) c = F3(0, 0, 0);
    a.pixels[y * a.w + x] = make_uchar4(displayByte(c.x, a.exposure), displayByte(c.y, a.exposure),
                                        displayByte(c.z, a.exposure), 255);
}
hipError_t launchDenoisePresent(const DenoiseArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(denoisePresentKernel, dim3((a.w + 15) / 16, (a.h + 15) / 16), dim3(kBlock), 0, s, a);
    return hipGetLastError();
}

// ------------------------------------------------------------------ bandwidth probe
__global__ __launch_bounds__(kBlock) void copyKernel(const float4* __restrict__ in, float4* __restrict__ out, long long n) {
    for (long long i = blockIdx.x * (long long)kBlock + threadIdx.x; i < n; i += (long long)gridDim.x * kBlock)
        out[i] = in[i];
}
__global__ void zeroIntsKernel(int* p, int n) {
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) p[k] = 0;
}
hipError_t launchZeroInts(int* p, int n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(zeroIntsKernel, dim3(std::min((n + kBlock - 1) / kBlock, 1024)), dim3(kBlock), 0, s, p, n);
    return hipGetLastError();
}

hipError_t launchCopy(const float4* in, float4* out, long long n, int blocks, hipStream_t s) {
    hipLaunchKernelGGL(copyKernel, dim3(blocks), dim3(kBlock), 0, s, in, out, n);
    return hipGetLastError();
}

// ------------------------------------------------------------------ Sobol batch
__global__ __launch_bounds__(kBlock) void sobolBatchKernel(const unsigned* pixels, const unsigned* samples, int count,
                                                           unsigned seed, int dims, float* out) {
    __shared__ unsigned sDirs[kSobolDims * 32];   // the requested dimensions' direction numbers
    loadSobolRows(sDirs, 0, dims);
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= count) return;
    const unsigned p = pixels[i], s = samples[i];
    for (int d = 0; d < dims; d++) out[(size_t)i * dims + d] = sobolSample(sDirs, p, s, seed, (unsigned)d);
}
hipError_t launchSobolBatch(const unsigned* pixels, const unsigned* samples, int count, unsigned seed, int dims,
                            float* out, hipStream_t s) {
    hipLaunchKernelGGL(sobolBatchKernel, dim3((count + kBlock - 1) / kBlock), dim3(kBlock), 0, s, pixels, samples,
                       count, seed, dims, out);
    return hipGetLastError();
}

}  // namespace rt
