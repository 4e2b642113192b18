// bvh_build.hip - binned-SAH BVH construction on the GPU and its collapse to the W-wide,
// 64-byte aligned traversal layout.
//
// The binary tree follows the reference builder's rules (Native/bvh.cuh:35-105): padded
// primitive boxes, vertex-mean centroids, `bins` bins per axis over the centroid extent,
// leaf at <= maxLeaf primitives or depth 32, cost A + A_L n_L + A_R n_R against the leaf cost
// A n, first strictly better (axis, bin) in x, y, z and ascending-bin order, children
// adjacent.  Bounds are exact min/max (order independent) and the costs are evaluated with the
// reference's float expression, so with the same bin count the GPU tree has the same
// partitions as the reference's; only node numbering and the order inside a leaf differ.
//
// Construction is level-synchronous over "tasks" (nodes still to split), each owning a
// contiguous range of the primitive order:
//   bounds   per-task node box and centroid box (ordered-int atomics, LDS pre-reduction
//            when a workgroup's positions all belong to one task)
//   bin      per-task 3 x bins bins of counts and boxes (same pre-reduction)
//   sah      one lane per task: leaf decision or best split; allocates the two children
//   partition scatter to the left/right child range (wave-aggregated cursors)
// then a top-down collapse writes WideNode<W> records and the leaf-ordered triangle records.
#include <vector>
#include "bvh_gpu.h"
#include "rt_core.h"

namespace rt {

namespace {

constexpr int kB = 256;         // threads per workgroup
constexpr int kPer = 4;         // positions per thread in the bounds / bin passes
constexpr int kMaxBins = 32;
constexpr int kBinInts = 7;     // count, lo xyz, hi xyz (ordered ints)

__device__ __forceinline__ int ordKey(float f) {
    const int i = __float_as_int(f);
    return i >= 0 ? i : i ^ 0x7FFFFFFF;
}
__device__ __forceinline__ float ordVal(int k) { return __int_as_float(k >= 0 ? k : k ^ 0x7FFFFFFF); }

struct Task { int node, first, count, depth, binSlot; };

struct SplitInfo { int axis, bin, leftCount, childTask; };

__device__ __forceinline__ float halfAreaOf(float lx, float ly, float lz, float hx, float hy, float hz) {
    const float x = fmaxf(0, hx - lx), y = fmaxf(0, hy - ly), z = fmaxf(0, hz - lz);
    return x * y + x * z + y * z;
}

// Primitive boxes with the reference's padding (Native/bvh.cuh:38-48).
__global__ void primPrepKernel(const Triangle* tris, int n, float4* plo, float4* phi, float4* cen, int* order,
                               int* posTask) {
    const int i = blockIdx.x * kB + threadIdx.x;
    if (i >= n) return;
    const Triangle t = tris[i];
    const float v[3][3] = {{t.origin.x, t.origin.y, t.origin.z},
                           {t.origin.x + t.edge1.x, t.origin.y + t.edge1.y, t.origin.z + t.edge1.z},
                           {t.origin.x + t.edge2.x, t.origin.y + t.edge2.y, t.origin.z + t.edge2.z}};
    float lo[3], hi[3], c[3];
    for (int k = 0; k < 3; k++) {
        lo[k] = fminf(fminf(fminf(FLT_MAX, v[0][k]), v[1][k]), v[2][k]);
        hi[k] = fmaxf(fmaxf(fmaxf(-FLT_MAX, v[0][k]), v[1][k]), v[2][k]);
        c[k] = (v[0][k] + v[1][k] + v[2][k]) / 3;
        const float pad = 1e-5f + 1e-6f * fmaxf(fabsf(lo[k]), fabsf(hi[k]));
        lo[k] -= pad; hi[k] += pad;
    }
    plo[i] = make_float4(lo[0], lo[1], lo[2], 0);
    phi[i] = make_float4(hi[0], hi[1], hi[2], 0);
    cen[i] = make_float4(c[0], c[1], c[2], 0);
    order[i] = i;
    posTask[i] = 0;
}

__global__ void initBoxKernel(int* box, int count12) {
    const int i = blockIdx.x * kB + threadIdx.x;
    if (i >= count12) return;
    const int f = i % 12;   // 0..2 node lo, 3..5 node hi, 6..8 centroid lo, 9..11 centroid hi
    box[i] = (f < 3 || (f >= 6 && f < 9)) ? 0x7F800000 : (int)(0xFF800000u ^ 0x7FFFFFFFu);
}

// Per-task node box and centroid box.
__global__ __launch_bounds__(kB) void taskBoundsKernel(const int* order, const int* posTask, const float4* plo,
                                                       const float4* phi, const float4* cen, int* box, int n) {
    __shared__ int sBox[12];
    __shared__ int sUniform;
    const int base = blockIdx.x * kB * kPer;
    const int last = min(base + kB * kPer, n) - 1;
    if (threadIdx.x == 0) {
        const int t0 = posTask[base], t1 = posTask[last];
        sUniform = (t0 >= 0 && t0 == t1) ? t0 : -1;
    }
    if (threadIdx.x < 12) sBox[threadIdx.x] = (threadIdx.x < 3 || (threadIdx.x >= 6 && threadIdx.x < 9)) ? 0x7F800000 : (int)(0xFF800000u ^ 0x7FFFFFFFu);
    __syncthreads();
    const int uni = sUniform;
    for (int k = 0; k < kPer; k++) {
        const int j = base + k * kB + threadIdx.x;
        if (j > last) break;
        const int t = uni >= 0 ? uni : posTask[j];
        if (t < 0) continue;
        const int p = order[j];
        const float4 l = plo[p], h = phi[p], c = cen[p];
        int* dst = uni >= 0 ? sBox : box + (size_t)t * 12;
        atomicMin(&dst[0], ordKey(l.x)); atomicMin(&dst[1], ordKey(l.y)); atomicMin(&dst[2], ordKey(l.z));
        atomicMax(&dst[3], ordKey(h.x)); atomicMax(&dst[4], ordKey(h.y)); atomicMax(&dst[5], ordKey(h.z));
        atomicMin(&dst[6], ordKey(c.x)); atomicMin(&dst[7], ordKey(c.y)); atomicMin(&dst[8], ordKey(c.z));
        atomicMax(&dst[9], ordKey(c.x)); atomicMax(&dst[10], ordKey(c.y)); atomicMax(&dst[11], ordKey(c.z));
    }
    __syncthreads();
    if (uni >= 0 && threadIdx.x < 12) {
        int* g = box + (size_t)uni * 12 + threadIdx.x;
        const int f = threadIdx.x;
        if (f < 3 || (f >= 6 && f < 9)) atomicMin(g, sBox[f]);
        else atomicMax(g, sBox[f]);
    }
}

__global__ void initBinsKernel(int* bins, long long slotsTimesBins) {
    const long long i = blockIdx.x * (long long)kB + threadIdx.x;
    if (i >= slotsTimesBins) return;
    int* b = bins + i * kBinInts;
    b[0] = 0;
    b[1] = b[2] = b[3] = 0x7F800000;
    b[4] = b[5] = b[6] = (int)(0xFF800000u ^ 0x7FFFFFFFu);
}

__device__ __forceinline__ int binIndex(float c, float lo, float scale, int bins) {
    const int b = (int)((c - lo) * scale);
    return b < 0 ? 0 : (b > bins - 1 ? bins - 1 : b);
}

// Bins of every task that may split (binSlot >= 0); 3 axes x bins records of kBinInts.
__global__ __launch_bounds__(kB) void binKernel(const int* order, const int* posTask, const Task* tasks,
                                                const float4* plo, const float4* phi, const float4* cen,
                                                const int* box, int* bins, int nbins, int n) {
    __shared__ int sBins[3 * kMaxBins * kBinInts];
    __shared__ int sUniform;
    __shared__ float sLo[3], sScale[3];
    const int base = blockIdx.x * kB * kPer;
    const int last = min(base + kB * kPer, n) - 1;
    if (threadIdx.x == 0) {
        const int t0 = posTask[base], t1 = posTask[last];
        sUniform = (t0 >= 0 && t0 == t1 && tasks[t0].binSlot >= 0) ? t0 : -1;
        if (sUniform >= 0) {
            const int* bx = box + (size_t)t0 * 12;
            for (int a = 0; a < 3; a++) {
                const float lo = ordVal(bx[6 + a]), ext = ordVal(bx[9 + a]) - lo;
                sLo[a] = lo;
                sScale[a] = ext > 1e-12f ? nbins / ext : 0.0f;
            }
        }
    }
    for (int k = threadIdx.x; k < 3 * nbins; k += kB) {
        int* b = sBins + k * kBinInts;
        b[0] = 0; b[1] = b[2] = b[3] = 0x7F800000; b[4] = b[5] = b[6] = (int)(0xFF800000u ^ 0x7FFFFFFFu);
    }
    __syncthreads();
    const int uni = sUniform;
    for (int k = 0; k < kPer; k++) {
        const int j = base + k * kB + threadIdx.x;
        if (j > last) break;
        const int t = uni >= 0 ? uni : posTask[j];
        if (t < 0) continue;
        int* dstBase;
        float lo3[3], sc3[3];
        if (uni >= 0) {
            dstBase = sBins;
            for (int a = 0; a < 3; a++) { lo3[a] = sLo[a]; sc3[a] = sScale[a]; }
        } else {
            const int slot = tasks[t].binSlot;
            if (slot < 0) continue;
            dstBase = bins + (size_t)slot * 3 * nbins * kBinInts;
            const int* bx = box + (size_t)t * 12;
            for (int a = 0; a < 3; a++) {
                const float lo = ordVal(bx[6 + a]), ext = ordVal(bx[9 + a]) - lo;
                lo3[a] = lo;
                sc3[a] = ext > 1e-12f ? nbins / ext : 0.0f;
            }
        }
        const int p = order[j];
        const float4 l = plo[p], h = phi[p], c = cen[p];
        const float cc[3] = {c.x, c.y, c.z};
        for (int a = 0; a < 3; a++) {
            if (sc3[a] == 0.0f) continue;
            int* b = dstBase + (a * nbins + binIndex(cc[a], lo3[a], sc3[a], nbins)) * kBinInts;
            atomicAdd(&b[0], 1);
            atomicMin(&b[1], ordKey(l.x)); atomicMin(&b[2], ordKey(l.y)); atomicMin(&b[3], ordKey(l.z));
            atomicMax(&b[4], ordKey(h.x)); atomicMax(&b[5], ordKey(h.y)); atomicMax(&b[6], ordKey(h.z));
        }
    }
    __syncthreads();
    if (uni >= 0) {
        int* g = bins + (size_t)tasks[uni].binSlot * 3 * nbins * kBinInts;
        for (int k = threadIdx.x; k < 3 * nbins; k += kB) {
            const int* s = sBins + k * kBinInts;
            if (s[0] == 0) continue;
            int* d = g + k * kBinInts;
            atomicAdd(&d[0], s[0]);
            atomicMin(&d[1], s[1]); atomicMin(&d[2], s[2]); atomicMin(&d[3], s[3]);
            atomicMax(&d[4], s[4]); atomicMax(&d[5], s[5]); atomicMax(&d[6], s[6]);
        }
    }
}

struct LevelArgs {
    const Task* tasks; int taskCount;
    const int* box; const int* bins; int nbins; int maxLeaf;
    RtBvhNode* nodes; int* counters;   // [0] node count, [1] next task count, [2] next bin slots, [3] max depth
    Task* nextTasks; SplitInfo* split;
};

// One lane per task: the reference's leaf test and SAH sweep (Native/bvh.cuh:56-100).
__global__ void sahKernel(LevelArgs a) {
    const int i = blockIdx.x * kB + threadIdx.x;
    if (i >= a.taskCount) return;
    const Task t = a.tasks[i];
    atomicMax(&a.counters[3], t.depth);
    const int* bx = a.box + (size_t)i * 12;
    const float blx = ordVal(bx[0]), bly = ordVal(bx[1]), blz = ordVal(bx[2]);
    const float bhx = ordVal(bx[3]), bhy = ordVal(bx[4]), bhz = ordVal(bx[5]);
    RtBvhNode nd;
    nd.low[0] = blx; nd.low[1] = bly; nd.low[2] = blz;
    nd.high[0] = bhx; nd.high[1] = bhy; nd.high[2] = bhz;
    nd.first = t.first; nd.count = t.count;
    SplitInfo si; si.axis = -1; si.bin = -1; si.leftCount = 0; si.childTask = -1;
    if (t.binSlot >= 0) {
        const float area = halfAreaOf(blx, bly, blz, bhx, bhy, bhz);
        float best = area * t.count;
        const int* tb = a.bins + (size_t)t.binSlot * 3 * a.nbins * kBinInts;
        for (int axis = 0; axis < 3; axis++) {
            const float ext = ordVal(bx[9 + axis]) - ordVal(bx[6 + axis]);
            if (ext <= 1e-12f) continue;
            const int* ab = tb + axis * a.nbins * kBinInts;
            // suffix sums from the right, as in the reference (bins 1..nbins-1 are needed)
            float rlx = FLT_MAX, rly = FLT_MAX, rlz = FLT_MAX, rhx = -FLT_MAX, rhy = -FLT_MAX, rhz = -FLT_MAX;
            float sufArea[kMaxBins];
            int sufCount[kMaxBins];
            int rc = 0;
            for (int b = a.nbins - 1; b >= 0; b--) {
                const int* e = ab + b * kBinInts;
                if (e[0]) {
                    rlx = fminf(rlx, ordVal(e[1])); rly = fminf(rly, ordVal(e[2])); rlz = fminf(rlz, ordVal(e[3]));
                    rhx = fmaxf(rhx, ordVal(e[4])); rhy = fmaxf(rhy, ordVal(e[5])); rhz = fmaxf(rhz, ordVal(e[6]));
                }
                rc += e[0];
                sufArea[b] = halfAreaOf(rlx, rly, rlz, rhx, rhy, rhz);
                sufCount[b] = rc;
            }
            float llx = FLT_MAX, lly = FLT_MAX, llz = FLT_MAX, lhx = -FLT_MAX, lhy = -FLT_MAX, lhz = -FLT_MAX;
            int lc = 0;
            for (int b = 0; b < a.nbins - 1; b++) {
                const int* e = ab + b * kBinInts;
                if (e[0]) {
                    llx = fminf(llx, ordVal(e[1])); lly = fminf(lly, ordVal(e[2])); llz = fminf(llz, ordVal(e[3]));
                    lhx = fmaxf(lhx, ordVal(e[4])); lhy = fmaxf(lhy, ordVal(e[5])); lhz = fmaxf(lhz, ordVal(e[6]));
                }
                lc += e[0];
                if (!lc || !sufCount[b + 1]) continue;
                const float cost = area + halfAreaOf(llx, lly, llz, lhx, lhy, lhz) * lc + sufArea[b + 1] * sufCount[b + 1];
                if (cost < best) { best = cost; si.axis = axis; si.bin = b; si.leftCount = lc; }
            }
        }
    }
    if (si.axis >= 0 && si.leftCount > 0 && si.leftCount < t.count) {
        const int child = atomicAdd(&a.counters[0], 2);
        nd.first = child; nd.count = 0;
        const int slot = atomicAdd(&a.counters[1], 2);
        si.childTask = slot;
        const int counts[2] = {si.leftCount, t.count - si.leftCount};
        const int firsts[2] = {t.first, t.first + si.leftCount};
        for (int c = 0; c < 2; c++) {
            Task ct;
            ct.node = child + c; ct.first = firsts[c]; ct.count = counts[c]; ct.depth = t.depth + 1;
            const bool mayBeLeaf = ct.count <= a.maxLeaf || ct.depth == kBvhDepthLimit;
            ct.binSlot = mayBeLeaf ? -1 : atomicAdd(&a.counters[2], 1);
            a.nextTasks[slot + c] = ct;
        }
    } else {
        si.axis = -1;
    }
    a.nodes[t.node] = nd;
    a.split[i] = si;
}

// Scatter each split task's positions to its left / right child ranges.  When a whole wave
// works on one task (the common case near the root), each side takes one atomic per wave.
__global__ __launch_bounds__(kB) void partitionKernel(const int* order, int* orderOut, const int* posTask,
                                                      int* posTaskOut, const float4* cen, const Task* tasks,
                                                      const SplitInfo* split, const int* box, int* cursors,
                                                      int nbins, int n, int* error) {
    const int j = blockIdx.x * kB + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const int t = j < n ? posTask[j] : -2;
    const SplitInfo si = t >= 0 ? split[t] : SplitInfo{-1, -1, 0, -1};
    const bool moving = t >= 0 && si.axis >= 0;
    if (j < n && !moving) {
        orderOut[j] = order[j];
        posTaskOut[j] = -1;
    }
    int right = 0, p = 0;
    if (moving) {
        p = order[j];
        const float4 c = cen[p];
        const float cc = si.axis == 0 ? c.x : (si.axis == 1 ? c.y : c.z);
        const int* bx = box + (size_t)t * 12;
        const float lo = ordVal(bx[6 + si.axis]);
        const float scale = nbins / (ordVal(bx[9 + si.axis]) - lo);
        right = binIndex(cc, lo, scale, nbins) <= si.bin ? 0 : 1;
    }
    // Wave-uniform task?  (all lanes of the wave reach this point)
    const int t0 = __shfl(t, 0, 64);
    const unsigned long long all = __ballot(true);
    const bool uniform = __ballot(moving && t == t0) == all;
    int k;
    if (uniform) {
        const unsigned long long mr = __ballot(right == 1), ml = all & ~mr;
        int baseL = 0, baseR = 0;
        if (lane == 0) {
            if (ml) baseL = atomicAdd(&cursors[2 * t0], __popcll(ml));
            if (mr) baseR = atomicAdd(&cursors[2 * t0 + 1], __popcll(mr));
        }
        baseL = __shfl(baseL, 0, 64);
        baseR = __shfl(baseR, 0, 64);
        const unsigned long long mine = right ? mr : ml;
        k = (right ? baseR : baseL) +
            (int)__builtin_amdgcn_mbcnt_hi((unsigned)(mine >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)mine, 0u));
    } else {
        k = moving ? atomicAdd(&cursors[2 * t + right], 1) : 0;
    }
    if (!moving) return;
    const int limit = right ? tasks[t].count - si.leftCount : si.leftCount;
    if (k >= limit) { atomicOr(error, 1); return; }   // bins and partition disagree: refuse
    const int dst = tasks[t].first + (right ? si.leftCount : 0) + k;
    orderOut[dst] = p;
    posTaskOut[dst] = si.childTask + right;
}

__global__ void leafMarkKernel(const RtBvhNode* nodes, int count, int* lastFlag) {
    const int i = blockIdx.x * kB + threadIdx.x;
    if (i >= count) return;
    const RtBvhNode nd = nodes[i];
    if (nd.count > 0) lastFlag[nd.first + nd.count - 1] = 1;
}

__global__ void recordKernel(const Triangle* tris, const int* order, const int* lastFlag, TriRecord* recs, int n) {
    const int j = blockIdx.x * kB + threadIdx.x;
    if (j >= n) return;
    const int i = order[j];
    const Triangle& t = tris[i];
    TriRecord r;
    r.v0x = t.origin.x; r.v0y = t.origin.y; r.v0z = t.origin.z; r.index = i;
    r.e1x = t.edge1.x; r.e1y = t.edge1.y; r.e1z = t.edge1.z;
    r.flags = (t.oneSided ? kTriOneSided : 0) | (lastFlag[j] ? kTriLastInLeaf : 0);
    r.e2x = t.edge2.x; r.e2y = t.edge2.y; r.e2z = t.edge2.z; r.pad = 0;
    recs[j] = r;
}

struct WideTask { int bin, out, depth; };

__device__ __forceinline__ float nodeHalfArea(const RtBvhNode& n) {
    return halfAreaOf(n.low[0], n.low[1], n.low[2], n.high[0], n.high[1], n.high[2]);
}

// Writes one wide node from its nk children (binary node boxes and wide references).
template <int W>
__device__ __forceinline__ void writeWide(WideNode<W>& node, const RtBvhNode* kids, const int* refs, int nk) {
    for (int c = 0; c < W; c++) {
        node.lox[c] = node.loy[c] = node.loz[c] = INFINITY;
        node.hix[c] = node.hiy[c] = node.hiz[c] = INFINITY;
        node.child[c] = kEmptyChild; node.count[c] = 0;
    }
    for (int c = 0; c < nk; c++) {
        const RtBvhNode& k = kids[c];
        node.lox[c] = k.low[0]; node.loy[c] = k.low[1]; node.loz[c] = k.low[2];
        node.hix[c] = k.high[0]; node.hiy[c] = k.high[1]; node.hiz[c] = k.high[2];
        node.child[c] = refs[c]; node.count[c] = k.count;
    }
}
template <int W>
__device__ __forceinline__ void writeWide(QNode<W>& node, const RtBvhNode* kids, const int* refs, int nk) {
    float lo[W][3], hi[W][3];
    for (int c = 0; c < nk; c++)
        for (int a = 0; a < 3; a++) { lo[c][a] = kids[c].low[a]; hi[c][a] = kids[c].high[a]; }
    quantizeNode<W>(node, lo, hi, refs, nk);
}

// Greedy collapse (largest-area inner child opened first) of the binary tree into N nodes
// (WideNode<W> or QNode<W>); leaves keep their binary-tree record ranges.
template <class N>
__global__ void collapseKernel(const WideTask* tasks, int taskCount, const RtBvhNode* bin, N* out, int* counters,
                               WideTask* next) {
    constexpr int W = NodeTraits<N>::kWidth;
    const int i = blockIdx.x * kB + threadIdx.x;
    if (i >= taskCount) return;
    const WideTask t = tasks[i];
    atomicMax(&counters[2], t.depth);
    int kids[W], nk = 0;
    const RtBvhNode b = bin[t.bin];
    if (b.count > 0) kids[nk++] = t.bin;
    else { kids[nk++] = b.first; kids[nk++] = b.first + 1; }
    while (nk < W) {
        int pick = -1;
        float area = -1;
        for (int k = 0; k < nk; k++) {
            const RtBvhNode kn = bin[kids[k]];
            if (kn.count == 0 && nodeHalfArea(kn) > area) { area = nodeHalfArea(kn); pick = k; }
        }
        if (pick < 0) break;
        const int open = kids[pick];
        const int f = bin[open].first;
        kids[pick] = f;
        kids[nk++] = f + 1;
    }
    RtBvhNode kn[W];
    int refs[W];
    for (int c = 0; c < nk; c++) {
        kn[c] = bin[kids[c]];
        if (kn[c].count > 0) {
            refs[c] = makeLeafRef(kn[c].first, kn[c].count);
        } else {
            const int idx = atomicAdd(&counters[0], 1);
            refs[c] = idx;
            next[atomicAdd(&counters[1], 1)] = WideTask{kids[c], idx, t.depth + 1};
        }
    }
    N node;
    writeWide(node, kn, refs, nk);
    out[t.out] = node;
}

inline int blocksFor(long long n, int per = 1) { return (int)((n + (long long)kB * per - 1) / ((long long)kB * per)); }

template <class T> hipError_t grow(T*& p, size_t& cap, size_t need) {
    if (need <= cap) return hipSuccess;
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = need + need / 2 + 64;
    return hipMalloc(reinterpret_cast<void**>(&p), cap * sizeof(T));
}

#define RT_TRY(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { err = e_; goto done; } } while (0)

}  // namespace

template <class N>
static hipError_t collapseW(const RtBvhNode* dBin, int binCount, void** wideOut, int* wideCount, int* wideDepth,
                            hipStream_t s) {
    hipError_t err = hipSuccess;
    N* out = nullptr;
    WideTask *ta = nullptr, *tb = nullptr;
    int* ctr = nullptr;
    int h[3] = {1, 1, 0};   // wide node count, next task count, max depth
    int taskCount = 1;
    const WideTask root{0, 0, 0};
    // At most one wide node per binary inner node, plus the root.
    RT_TRY(hipMalloc(&out, (size_t)(binCount + 1) * sizeof(N)));
    RT_TRY(hipMalloc(&ta, (size_t)(binCount + 1) * sizeof(WideTask)));
    RT_TRY(hipMalloc(&tb, (size_t)(binCount + 1) * sizeof(WideTask)));
    RT_TRY(hipMalloc(&ctr, 3 * sizeof(int)));
    RT_TRY(hipStreamSynchronize(s));
    RT_TRY(hipMemcpy(ta, &root, sizeof(root), hipMemcpyHostToDevice));
    RT_TRY(hipMemcpy(ctr, h, sizeof(h), hipMemcpyHostToDevice));
    while (taskCount > 0) {
        RT_TRY(hipMemsetAsync(ctr + 1, 0, sizeof(int), s));
        hipLaunchKernelGGL((collapseKernel<N>), dim3(blocksFor(taskCount)), dim3(kB), 0, s, ta, taskCount, dBin, out,
                           ctr, tb);
        RT_TRY(hipGetLastError());
        RT_TRY(hipMemcpyAsync(h, ctr, sizeof(h), hipMemcpyDeviceToHost, s));
        RT_TRY(hipStreamSynchronize(s));
        taskCount = h[1];
        std::swap(ta, tb);
    }
    *wideOut = out; out = nullptr;
    *wideCount = h[0];
    *wideDepth = h[2];
done:
    if (out) (void)hipFree(out);
    if (ta) (void)hipFree(ta);
    if (tb) (void)hipFree(tb);
    if (ctr) (void)hipFree(ctr);
    return err;
}

hipError_t buildBvhGpu(const Triangle* dTris, int n, const GpuBuildParams& prm, GpuBuildResult& res, hipStream_t s) {
    hipError_t err = hipSuccess;
    const int nbins = prm.bins;
    float4 *plo = nullptr, *phi = nullptr, *cen = nullptr;
    int *order = nullptr, *order2 = nullptr, *pos = nullptr, *pos2 = nullptr, *lastFlag = nullptr;
    Task *tasks = nullptr, *next = nullptr;
    SplitInfo* split = nullptr;
    int *box = nullptr, *bins = nullptr, *cursors = nullptr, *ctr = nullptr;
    RtBvhNode* nodes = nullptr;
    size_t taskCap = 0, nextCap = 0, splitCap = 0, boxCap = 0, binCap = 0, curCap = 0;
    int taskCount = 1, binSlots = 0;
    int h[5] = {1, 0, 0, 0, 0};   // nodes, next tasks, next bin slots, max depth, partition error
    if (nbins < 2 || nbins > kMaxBins || n < 1 || n >= kLeafFirstLimit) return hipErrorInvalidValue;
    RT_TRY(hipMalloc(&plo, (size_t)n * 16));
    RT_TRY(hipMalloc(&phi, (size_t)n * 16));
    RT_TRY(hipMalloc(&cen, (size_t)n * 16));
    RT_TRY(hipMalloc(&order, (size_t)n * 4));
    RT_TRY(hipMalloc(&order2, (size_t)n * 4));
    RT_TRY(hipMalloc(&pos, (size_t)n * 4));
    RT_TRY(hipMalloc(&pos2, (size_t)n * 4));
    RT_TRY(hipMalloc(&nodes, (size_t)(2 * n) * sizeof(RtBvhNode)));
    RT_TRY(hipMalloc(&ctr, 5 * sizeof(int)));
    hipLaunchKernelGGL(primPrepKernel, dim3(blocksFor(n)), dim3(kB), 0, s, dTris, n, plo, phi, cen, order, pos);
    RT_TRY(hipGetLastError());
    {
        RT_TRY(grow(tasks, taskCap, 1));
        Task root{0, 0, n, 0, (n <= prm.maxLeaf) ? -1 : 0};
        binSlots = root.binSlot >= 0 ? 1 : 0;
        RT_TRY(hipStreamSynchronize(s));
        RT_TRY(hipMemcpy(tasks, &root, sizeof(root), hipMemcpyHostToDevice));
        RT_TRY(hipMemcpy(ctr, h, sizeof(h), hipMemcpyHostToDevice));
    }
    while (taskCount > 0) {
        RT_TRY(grow(box, boxCap, (size_t)taskCount * 12));
        RT_TRY(grow(split, splitCap, (size_t)taskCount));
        RT_TRY(grow(next, nextCap, (size_t)taskCount * 2));
        RT_TRY(grow(cursors, curCap, (size_t)taskCount * 2));
        RT_TRY(grow(bins, binCap, (size_t)(binSlots > 0 ? binSlots : 1) * 3 * nbins * kBinInts));
        hipLaunchKernelGGL(initBoxKernel, dim3(blocksFor((long long)taskCount * 12)), dim3(kB), 0, s, box, taskCount * 12);
        RT_TRY(hipMemsetAsync(cursors, 0, (size_t)taskCount * 2 * sizeof(int), s));
        hipLaunchKernelGGL(taskBoundsKernel, dim3(blocksFor(n, kPer)), dim3(kB), 0, s, order, pos, plo, phi, cen, box, n);
        RT_TRY(hipGetLastError());
        if (binSlots > 0) {
            hipLaunchKernelGGL(initBinsKernel, dim3(blocksFor((long long)binSlots * 3 * nbins)), dim3(kB), 0, s, bins,
                               (long long)binSlots * 3 * nbins);
            hipLaunchKernelGGL(binKernel, dim3(blocksFor(n, kPer)), dim3(kB), 0, s, order, pos, tasks, plo, phi, cen,
                               box, bins, nbins, n);
            RT_TRY(hipGetLastError());
        }
        {
            RT_TRY(hipMemsetAsync(ctr + 1, 0, 2 * sizeof(int), s));
            LevelArgs la;
            la.tasks = tasks; la.taskCount = taskCount; la.box = box; la.bins = bins; la.nbins = nbins;
            la.maxLeaf = prm.maxLeaf; la.nodes = nodes; la.counters = ctr; la.nextTasks = next; la.split = split;
            hipLaunchKernelGGL(sahKernel, dim3(blocksFor(taskCount)), dim3(kB), 0, s, la);
            RT_TRY(hipGetLastError());
        }
        hipLaunchKernelGGL(partitionKernel, dim3(blocksFor(n)), dim3(kB), 0, s, order, order2, pos, pos2, cen, tasks,
                           split, box, cursors, nbins, n, ctr + 4);
        RT_TRY(hipGetLastError());
        std::swap(order, order2);
        std::swap(pos, pos2);
        RT_TRY(hipMemcpyAsync(h, ctr, sizeof(h), hipMemcpyDeviceToHost, s));
        RT_TRY(hipStreamSynchronize(s));
        if (h[4]) { err = hipErrorUnknown; goto done; }
        taskCount = h[1];
        binSlots = h[2];
        std::swap(tasks, next);
        std::swap(taskCap, nextCap);
    }
    res.binNodeCount = h[0];
    res.depth = h[3];
    // leaf-ordered triangle records with last-in-leaf flags
    RT_TRY(hipMalloc(&lastFlag, (size_t)n * 4));
    RT_TRY(hipMemsetAsync(lastFlag, 0, (size_t)n * 4, s));
    hipLaunchKernelGGL(leafMarkKernel, dim3(blocksFor(res.binNodeCount)), dim3(kB), 0, s, nodes, res.binNodeCount, lastFlag);
    RT_TRY(hipGetLastError());
    RT_TRY(hipMalloc(&res.recs, (size_t)(n + kRecordPad) * sizeof(TriRecord)));
    RT_TRY(hipMemsetAsync(res.recs + n, 0, kRecordPad * sizeof(TriRecord), s));
    hipLaunchKernelGGL(recordKernel, dim3(blocksFor(n)), dim3(kB), 0, s, dTris, order, lastFlag, res.recs, n);
    RT_TRY(hipGetLastError());
    if (prm.quantized) {
        switch (prm.width) {
            case 2: RT_TRY(collapseW<QNode<2>>(nodes, res.binNodeCount, &res.wideNodes, &res.wideCount, &res.wideDepth, s)); break;
            case 8: RT_TRY(collapseW<QNode<8>>(nodes, res.binNodeCount, &res.wideNodes, &res.wideCount, &res.wideDepth, s)); break;
            default: RT_TRY(collapseW<QNode<4>>(nodes, res.binNodeCount, &res.wideNodes, &res.wideCount, &res.wideDepth, s)); break;
        }
    } else {
        switch (prm.width) {
            case 2: RT_TRY(collapseW<WideNode<2>>(nodes, res.binNodeCount, &res.wideNodes, &res.wideCount, &res.wideDepth, s)); break;
            case 8: RT_TRY(collapseW<WideNode<8>>(nodes, res.binNodeCount, &res.wideNodes, &res.wideCount, &res.wideDepth, s)); break;
            default: RT_TRY(collapseW<WideNode<4>>(nodes, res.binNodeCount, &res.wideNodes, &res.wideCount, &res.wideDepth, s)); break;
        }
    }
    RT_TRY(hipStreamSynchronize(s));
    res.binNodes = nodes; nodes = nullptr;
    res.order = order; order = nullptr;
done:
    for (void* p : {(void*)plo, (void*)phi, (void*)cen, (void*)order, (void*)order2, (void*)pos, (void*)pos2,
                    (void*)lastFlag, (void*)tasks, (void*)next, (void*)split, (void*)box, (void*)bins, (void*)cursors,
                    (void*)ctr, (void*)nodes})
        if (p) (void)hipFree(p);
    return err;
}

}  // namespace rt
