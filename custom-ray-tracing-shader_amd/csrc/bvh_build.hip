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
// Construction is level-synchronous over "tasks" (nodes still to decide), each owning a
// contiguous range of the primitive order; primitive records (box, centroid, index) move with
// their positions, so every pass streams them.  Per level:
//   large tasks (> 2048 primitives): position-parallel bounds and bins.  Threads fold their
//            positions in registers (or, at task boundaries, runs of lanes fold across the
//            wavefront) and each workgroup adds its task's box / bins with a few atomics.
//   small tasks: a group of lanes per task (a wavefront, or 16 / 32 lanes for tasks of at most
//            64 primitives) reduces the bounds, bins in LDS and runs the SAH sweep across the
//            lanes; its records move by ballot compaction.
//   decide   one lane per large task: the SAH sweep over its bins; leaf boxes.
//   emit     an exclusive scan of the split flags places the children (nodes and tasks), then
//            one lane per task writes its node and classifies its children.
//   partition the large tasks' records to their child ranges (workgroup-aggregated cursors).
// Integer min / max / add over ordered-int boxes are exact, so no result depends on the order
// of the atomics.  A top-down collapse then writes the W-wide nodes, and the leaf-ordered
// triangle records are gathered.
#include <algorithm>
#include <mutex>
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

// A task is a node still to be decided, owning positions [first, first + count) of the
// primitive order.  binSlot >= 0: a large task binned by the position-parallel passes (its
// slot in the global bins); kSmallTask: binned by one wavefront in LDS; kLeafTask: a leaf by
// the leaf rule (count <= maxLeaf or depth limit).
struct Task { int node, first, count, depth, binSlot; };
constexpr int kLeafTask = -1, kSmallTask = -2;
// Tasks up to this many primitives are binned by one wavefront; larger ones by the
// position-parallel passes (at most n / kSmallMax of them per level).
constexpr int kSmallMax = 2048;
// Small tasks up to this many primitives are "tiny": a group of 16 or 32 lanes decides each,
// several to a wavefront.
constexpr int kTinyMax = 64;

struct SplitInfo { int axis, bin, leftCount, childTask; };

__device__ __forceinline__ float halfAreaOf(float lx, float ly, float lz, float hx, float hy, float hz) {
    const float x = fmaxf(0, hx - lx), y = fmaxf(0, hy - ly), z = fmaxf(0, hz - lz);
    return x * y + x * z + y * z;
}

// Primitive records in position order (the order array's current permutation), so that every
// pass reads them contiguously and the partition moves them with their positions:
//   pa = (box lo xyz, centroid x), pb = (box hi xyz, centroid y), pc = (centroid z, index bits).
// Primitive boxes carry the reference's padding (Native/bvh.cuh:38-48).
__global__ void primPrepKernel(const Triangle* tris, int n, float4* pa, float4* pb, float2* pc, int* posTask) {
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
    pa[i] = make_float4(lo[0], lo[1], lo[2], c[0]);
    pb[i] = make_float4(hi[0], hi[1], hi[2], c[1]);
    pc[i] = make_float2(c[2], __int_as_float(i));
    posTask[i] = 0;
}

__device__ __forceinline__ int boxInit(int f) {
    return (f < 3 || (f >= 6 && f < 9)) ? 0x7F800000 : (int)(0xFF800000u ^ 0x7FFFFFFFu);
}

__global__ void initBoxKernel(int* box, int count12) {
    const int i = blockIdx.x * kB + threadIdx.x;
    if (i >= count12) return;
    box[i] = boxInit(i % 12);   // 0..2 node lo, 3..5 node hi, 6..8 centroid lo, 9..11 centroid hi
}

// Runs of lanes: maximal blocks of consecutive lanes of a wavefront holding one key.  A run's
// values fold into its first lane in log2(64) shuffle steps, so a run costs one set of atomics
// instead of one per lane.  Every lane of the wave takes part in these calls.
__device__ __forceinline__ int runLast(long long key, int lane) {
    const long long next = __shfl_down(key, 1, 64);
    const unsigned long long tails = __ballot(lane == 63 || next != key);
    return lane + (int)__builtin_ctzll(tails >> lane);
}
__device__ __forceinline__ bool runFirst(long long key, int lane) {
    const long long prev = __shfl_up(key, 1, 64);
    return lane == 0 || prev != key;
}

// Box fields as ordered ints: 0..2 node lo, 3..5 node hi, 6..8 centroid lo, 9..11 centroid hi.
__device__ __forceinline__ bool boxFieldIsMin(int f) { return f < 3 || (f >= 6 && f < 9); }
__device__ __forceinline__ void boxFold(int* b, float4 l, float4 h, float cz) {
    b[0] = min(b[0], ordKey(l.x)); b[1] = min(b[1], ordKey(l.y)); b[2] = min(b[2], ordKey(l.z));
    b[3] = max(b[3], ordKey(h.x)); b[4] = max(b[4], ordKey(h.y)); b[5] = max(b[5], ordKey(h.z));
    b[6] = min(b[6], ordKey(l.w)); b[7] = min(b[7], ordKey(h.w)); b[8] = min(b[8], ordKey(cz));
    b[9] = max(b[9], ordKey(l.w)); b[10] = max(b[10], ordKey(h.w)); b[11] = max(b[11], ordKey(cz));
}
// The box over the whole wavefront, in every lane.
__device__ __forceinline__ void waveBox(int* b) {
    for (int f = 0; f < 12; f++) {
        const bool isMin = boxFieldIsMin(f);
        for (int o = 32; o > 0; o >>= 1) {
            const int v = __shfl_xor(b[f], o, 64);
            b[f] = isMin ? min(b[f], v) : max(b[f], v);
        }
    }
}
// The box of each run, in the run's first lane.
__device__ __forceinline__ void runBox(int* b, int lastLane, int lane) {
    for (int o = 1; o < 64; o <<= 1) {
        const bool take = lane + o <= lastLane;
        for (int f = 0; f < 12; f++) {
            const int v = __shfl_down(b[f], o, 64);
            if (take) b[f] = boxFieldIsMin(f) ? min(b[f], v) : max(b[f], v);
        }
    }
}
__device__ __forceinline__ void boxAtomics(int* dst, const int* b) {
    for (int f = 0; f < 12; f++) {
        if (boxFieldIsMin(f)) atomicMin(&dst[f], b[f]);
        else atomicMax(&dst[f], b[f]);
    }
}

// Node and centroid boxes of the large tasks.  A workgroup walks kChunks consecutive chunks of
// 1024 positions.  While its chunks belong to one task, each thread folds its positions into
// registers; when the task changes (or the walk ends) the registers reduce across the
// wavefront, then the workgroup (LDS), and the task's global box takes 12 atomics.  A chunk
// holding a task boundary reduces each wavefront's runs of one task and issues 12 global
// atomics per run.  Min / max over ordered ints is exact, so the boxes do not depend on order.
constexpr int kChunks = 4;
__device__ __forceinline__ int chunkTask(const int* posTask, const Task* tasks, int base, int last) {
    const int t0 = posTask[base], t1 = posTask[last];
    return (t0 >= 0 && t0 == t1) ? (tasks[t0].binSlot >= 0 ? t0 : -2) : -1;   // -2: one small/leaf task
}
__device__ __forceinline__ void flushBoxRun(int* acc, int* sBox, int* box, int task, int lane) {
    waveBox(acc);
    if (lane == 0)
        for (int f = 0; f < 12; f++) {
            if (boxFieldIsMin(f)) atomicMin(&sBox[f], acc[f]);
            else atomicMax(&sBox[f], acc[f]);
        }
    __syncthreads();
    if (threadIdx.x < 12) {
        int* g = box + (size_t)task * 12 + threadIdx.x;
        if (boxFieldIsMin(threadIdx.x)) atomicMin(g, sBox[threadIdx.x]);
        else atomicMax(g, sBox[threadIdx.x]);
        sBox[threadIdx.x] = boxInit(threadIdx.x);
    }
    __syncthreads();
    for (int f = 0; f < 12; f++) acc[f] = boxInit(f);
}
__global__ __launch_bounds__(kB) void bigBoundsKernel(const int* posTask, const Task* tasks, const float4* pa,
                                                      const float4* pb, const float2* pc, int* box, int n) {
    __shared__ int sBox[12];
    __shared__ int sUniform;
    const int lane = threadIdx.x & 63;
    int acc[12];
    for (int f = 0; f < 12; f++) acc[f] = boxInit(f);
    if (threadIdx.x < 12) sBox[threadIdx.x] = boxInit(threadIdx.x);
    int run = -1;   // the task the registers accumulate (the same in every thread)
    for (int ch = 0; ch < kChunks; ch++) {
        const int base = (blockIdx.x * kChunks + ch) * kB * kPer;
        if (base >= n) break;
        const int last = min(base + kB * kPer, n) - 1;
        __syncthreads();   // the previous chunk's reads of sUniform are done
        if (threadIdx.x == 0) sUniform = chunkTask(posTask, tasks, base, last);
        __syncthreads();
        const int uni = sUniform;
        if (uni != run) {
            if (run >= 0) flushBoxRun(acc, sBox, box, run, lane);
            run = uni >= 0 ? uni : -1;
        }
        if (uni >= 0) {
            for (int k = 0; k < kPer; k++) {
                const int j = base + k * kB + threadIdx.x;
                if (j <= last) boxFold(acc, pa[j], pb[j], pc[j].x);
            }
        } else if (uni == -1) {
            for (int k = 0; k < kPer; k++) {
                const int j = base + k * kB + threadIdx.x;
                int t = j <= last ? posTask[j] : -1;
                if (t >= 0 && tasks[t].binSlot < 0) t = -1;
                int b[12];
                for (int f = 0; f < 12; f++) b[f] = boxInit(f);
                if (t >= 0) boxFold(b, pa[j], pb[j], pc[j].x);
                runBox(b, runLast(t, lane), lane);
                const bool first = runFirst(t, lane);
                if (t >= 0 && first) boxAtomics(box + (size_t)t * 12, b);
            }
        }
    }
    if (run >= 0) flushBoxRun(acc, sBox, box, run, lane);
}

__device__ __forceinline__ void binInit(int* b) {
    b[0] = 0;
    b[1] = b[2] = b[3] = 0x7F800000;
    b[4] = b[5] = b[6] = (int)(0xFF800000u ^ 0x7FFFFFFFu);
}

__global__ void initBinsKernel(int* bins, long long slotsTimesBins) {
    const long long i = blockIdx.x * (long long)kB + threadIdx.x;
    if (i >= slotsTimesBins) return;
    binInit(bins + i * kBinInts);
}

__device__ __forceinline__ int binIndex(float c, float lo, float scale, int bins) {
    const int b = (int)((c - lo) * scale);
    return b < 0 ? 0 : (b > bins - 1 ? bins - 1 : b);
}

// Centroid bin origin and scale per axis of a task's box (scale 0: no extent, axis skipped).
__device__ __forceinline__ void binFrame(const int* bx, int nbins, float* lo3, float* sc3) {
    for (int a = 0; a < 3; a++) {
        const float lo = ordVal(bx[6 + a]), ext = ordVal(bx[9 + a]) - lo;
        lo3[a] = lo;
        sc3[a] = ext > 1e-12f ? nbins / ext : 0.0f;
    }
}

// Adds one primitive to a task's 3 x nbins bins (counts and boxes as ordered ints).
__device__ __forceinline__ void binAdd(int* binsOfTask, int nbins, const float* lo3, const float* sc3, float4 l, float4 h,
                                       float cz) {
    const float cc[3] = {l.w, h.w, cz};
    for (int a = 0; a < 3; a++) {
        if (sc3[a] == 0.0f) continue;
        int* b = binsOfTask + (a * nbins + binIndex(cc[a], lo3[a], sc3[a], nbins)) * kBinInts;
        atomicAdd(&b[0], 1);
        atomicMin(&b[1], ordKey(l.x)); atomicMin(&b[2], ordKey(l.y)); atomicMin(&b[3], ordKey(l.z));
        atomicMax(&b[4], ordKey(h.x)); atomicMax(&b[5], ordKey(h.y)); atomicMax(&b[6], ordKey(h.z));
    }
}

// Adds each live lane's primitive to its task's 3 x nbins bins (counts and boxes as ordered
// ints).  `group` names the lane's bin set (lanes of one group and one bin form runs), and each
// run adds its folded count and box with one set of atomics.  Every lane of the wave calls it.
__device__ __forceinline__ void binAddRuns(int* binsOfTask, int nbins, const float* lo3, const float* sc3, float4 l,
                                           float4 h, float cz, bool live, long long group, int lane) {
    const float cc[3] = {l.w, h.w, cz};
    for (int a = 0; a < 3; a++) {
        const long long key = (live && sc3[a] != 0.0f) ? (group << 6) | binIndex(cc[a], lo3[a], sc3[a], nbins) : -1;
        int v[kBinInts] = {live ? 1 : 0, ordKey(l.x), ordKey(l.y), ordKey(l.z), ordKey(h.x), ordKey(h.y), ordKey(h.z)};
        const int lastLane = runLast(key, lane);
        for (int o = 1; o < 64; o <<= 1) {
            const bool take = lane + o <= lastLane;
            int u[kBinInts];
            for (int q = 0; q < kBinInts; q++) u[q] = __shfl_down(v[q], o, 64);
            if (take) {
                v[0] += u[0];
                for (int q = 1; q < 4; q++) v[q] = min(v[q], u[q]);
                for (int q = 4; q < 7; q++) v[q] = max(v[q], u[q]);
            }
        }
        const bool first = runFirst(key, lane);
        if (key >= 0 && first) {
            int* b = binsOfTask + (a * nbins + (int)(key & 63)) * kBinInts;
            atomicAdd(&b[0], v[0]);
            atomicMin(&b[1], v[1]); atomicMin(&b[2], v[2]); atomicMin(&b[3], v[3]);
            atomicMax(&b[4], v[4]); atomicMax(&b[5], v[5]); atomicMax(&b[6], v[6]);
        }
    }
}

// Bins of the large tasks.  A workgroup takes kChunks consecutive chunks of 1024 positions
// (4096 positions, so parts of at most 3 large tasks, each longer than 2048) and bins them into
// LDS bin sets, one per large task it meets, claimed (compare-and-swap on a tag) by the first
// run of lanes that meets the task; runs of lanes of one task and bin add to a set with one
// atomic set.  At the end each set is added to its task's global bins.  A chunk inside one
// small or leaf task is skipped.
constexpr int kBinSets = 4;
__device__ __forceinline__ int claimBinSet(int* sTag, int task) {
    for (int k = 0; k < kBinSets; k++) {
        const int old = atomicCAS(&sTag[k], -1, task);
        if (old == -1 || old == task) return k;
    }
    return -1;   // more tasks than sets (not possible by the sizes above): global atomics
}
__global__ __launch_bounds__(kB) void bigBinKernel(const int* posTask, const Task* tasks, const float4* pa,
                                                   const float4* pb, const float2* pc, const int* box, int* bins,
                                                   int nbins, int n) {
    __shared__ int sBins[kBinSets][3 * kMaxBins * kBinInts];
    __shared__ int sTag[kBinSets];
    __shared__ int sSkip;
    const int lane = threadIdx.x & 63, setInts = 3 * nbins * kBinInts;
    for (int k = threadIdx.x; k < kBinSets * 3 * nbins; k += kB) binInit(&sBins[k / (3 * nbins)][(k % (3 * nbins)) * kBinInts]);
    if (threadIdx.x < kBinSets) sTag[threadIdx.x] = -1;
    for (int ch = 0; ch < kChunks; ch++) {
        const int base = (blockIdx.x * kChunks + ch) * kB * kPer;
        if (base >= n) break;
        const int last = min(base + kB * kPer, n) - 1;
        __syncthreads();   // the previous chunk's reads of sSkip are done (and the sets initialised)
        if (threadIdx.x == 0) sSkip = chunkTask(posTask, tasks, base, last) == -2;
        __syncthreads();
        if (sSkip) continue;
        for (int k = 0; k < kPer; k++) {
            const int j = base + k * kB + threadIdx.x;
            int t = j <= last ? posTask[j] : -1;
            if (t >= 0 && tasks[t].binSlot < 0) t = -1;
            // the bin set of each run of one task, claimed by the run's first lane
            const bool head = runFirst(t, lane);
            int set = (t >= 0 && head) ? claimBinSet(sTag, t) : -1;
            const unsigned long long heads = __ballot(head);
            const int headLane = 63 - __clzll(heads & (~0ull >> (63 - lane)));
            set = __shfl(set, headLane, 64);
            const bool live = t >= 0;
            float lo3[3] = {0, 0, 0}, sc3[3] = {0, 0, 0};
            float4 l = make_float4(0, 0, 0, 0), hh = l;
            float cz = 0;
            int* dst = &sBins[0][0];
            if (live) {
                binFrame(box + (size_t)t * 12, nbins, lo3, sc3);
                l = pa[j]; hh = pb[j]; cz = pc[j].x;
                dst = set >= 0 ? &sBins[set][0] : bins + (size_t)tasks[t].binSlot * setInts;
            }
            binAddRuns(dst, nbins, lo3, sc3, l, hh, cz, live, live ? t : 0, lane);
        }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < kBinSets * 3 * nbins; k += kB) {
        const int set = k / (3 * nbins), task = sTag[set];
        const int* sb = &sBins[set][(k % (3 * nbins)) * kBinInts];
        if (task < 0 || sb[0] == 0) continue;
        int* d = bins + (size_t)tasks[task].binSlot * setInts + (k % (3 * nbins)) * kBinInts;
        atomicAdd(&d[0], sb[0]);
        atomicMin(&d[1], sb[1]); atomicMin(&d[2], sb[2]); atomicMin(&d[3], sb[3]);
        atomicMax(&d[4], sb[4]); atomicMax(&d[5], sb[5]); atomicMax(&d[6], sb[6]);
    }
}

struct LevelArgs {
    const Task* tasks; int taskCount;
    int* box; const int* bins; int nbins; int maxLeaf;
    const float4 *pa, *pb;
    RtBvhNode* nodes;
    int* counters;         // [2] next bin slots, [4] error, [5] next small, [6] splits, [7] next tiny
    Task* nextTasks; SplitInfo* split;
    int* splitFlag;        // 1 per task that splits (scanned into its child slots)
    const int* childBase;  // exclusive scan of splitFlag: split k of the level -> slots 2k, 2k + 1
    const int* blockPrefix;   // scan of the per-1024-task block sums (added to childBase)
    int nodeBase;          // nodes allocated before this level's children
    int* nextSmallList;    // next level's small tasks above kTinyMax primitives
    int* nextTinyList;     // next level's small tasks of at most kTinyMax primitives
};

// The reference's SAH sweep (Native/bvh.cuh:56-100) over one task's bins: the first strictly
// better (axis, bin) in x, y, z and ascending-bin order against the leaf cost area * count.
__device__ void sahSweep(const int* tb, const int* bx, int count, int nbins, SplitInfo& si) {
    const float blx = ordVal(bx[0]), bly = ordVal(bx[1]), blz = ordVal(bx[2]);
    const float bhx = ordVal(bx[3]), bhy = ordVal(bx[4]), bhz = ordVal(bx[5]);
    const float area = halfAreaOf(blx, bly, blz, bhx, bhy, bhz);
    float best = area * count;
    for (int axis = 0; axis < 3; axis++) {
        const float ext = ordVal(bx[9 + axis]) - ordVal(bx[6 + axis]);
        if (ext <= 1e-12f) continue;
        const int* ab = tb + axis * nbins * kBinInts;
        // suffix sums from the right, as in the reference (bins 1..nbins-1 are needed)
        float rlx = FLT_MAX, rly = FLT_MAX, rlz = FLT_MAX, rhx = -FLT_MAX, rhy = -FLT_MAX, rhz = -FLT_MAX;
        float sufArea[kMaxBins];
        int sufCount[kMaxBins];
        int rc = 0;
        for (int b = nbins - 1; b >= 0; b--) {
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
        for (int b = 0; b < nbins - 1; b++) {
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

__device__ __forceinline__ bool splits(const SplitInfo& si, int count) {
    return si.axis >= 0 && si.leftCount > 0 && si.leftCount < count;
}

// One lane per task: the box of each leaf task (from its few primitives) and the SAH choice of
// each large task (from the global bins); small tasks are decided by smallTaskKernel.
__global__ void decideKernel(LevelArgs a) {
    const int i = blockIdx.x * kB + threadIdx.x;
    if (i >= a.taskCount) return;
    const Task t = a.tasks[i];
    if (t.binSlot == kSmallTask) return;
    SplitInfo si; si.axis = -1; si.bin = -1; si.leftCount = 0; si.childTask = -1;
    if (t.binSlot == kLeafTask) {
        int b[6];
        for (int f = 0; f < 6; f++) b[f] = boxInit(f);
        for (int j = t.first; j < t.first + t.count; j++) {
            const float4 l = a.pa[j], h = a.pb[j];
            b[0] = min(b[0], ordKey(l.x)); b[1] = min(b[1], ordKey(l.y)); b[2] = min(b[2], ordKey(l.z));
            b[3] = max(b[3], ordKey(h.x)); b[4] = max(b[4], ordKey(h.y)); b[5] = max(b[5], ordKey(h.z));
        }
        int* bx = a.box + (size_t)i * 12;
        for (int f = 0; f < 6; f++) bx[f] = b[f];
    } else {
        sahSweep(a.bins + (size_t)t.binSlot * 3 * a.nbins * kBinInts, a.box + (size_t)i * 12, t.count, a.nbins, si);
        if (!splits(si, t.count)) si.axis = -1;
    }
    a.split[i] = si;
    a.splitFlag[i] = si.axis >= 0 ? 1 : 0;
}

// Exclusive scan of the split flags: per 1024 tasks a local scan (written to childBase) and the
// block's sum; then one workgroup scans the block sums (blockPrefix).
constexpr int kScanBlock = 1024;
__global__ __launch_bounds__(kScanBlock) void scanBlocksKernel(const int* flags, int* local, int* blockSums, int n) {
    __shared__ int sWave[kScanBlock / 64];
    const int i = blockIdx.x * kScanBlock + threadIdx.x, lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int f = i < n ? flags[i] : 0;
    int v = f;   // inclusive scan within the wave
    for (int o = 1; o < 64; o <<= 1) { const int u = __shfl_up(v, o, 64); if (lane >= o) v += u; }
    if (lane == 63) sWave[wave] = v;
    __syncthreads();
    if (threadIdx.x < kScanBlock / 64) {
        int w = sWave[threadIdx.x];
        for (int o = 1; o < kScanBlock / 64; o <<= 1) {
            const int u = __shfl_up(w, o, 64);
            if ((int)threadIdx.x >= o) w += u;
        }
        sWave[threadIdx.x] = w;
    }
    __syncthreads();
    const int before = (wave ? sWave[wave - 1] : 0) + v - f;
    if (i < n) local[i] = before;
    if (threadIdx.x == kScanBlock - 1) blockSums[blockIdx.x] = sWave[kScanBlock / 64 - 1];
}
__global__ __launch_bounds__(kScanBlock) void scanTopKernel(const int* blockSums, int* blockPrefix, int blocks, int* total) {
    __shared__ int sCarry;
    if (threadIdx.x == 0) sCarry = 0;
    __syncthreads();
    for (int base = 0; base < blocks; base += kScanBlock) {
        const int i = base + threadIdx.x, lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        __shared__ int sWave[kScanBlock / 64];
        const int f = i < blocks ? blockSums[i] : 0;
        int v = f;
        for (int o = 1; o < 64; o <<= 1) { const int u = __shfl_up(v, o, 64); if (lane >= o) v += u; }
        if (lane == 63) sWave[wave] = v;
        __syncthreads();
        if (threadIdx.x < kScanBlock / 64) {
            int w = sWave[threadIdx.x];
            for (int o = 1; o < kScanBlock / 64; o <<= 1) {
                const int u = __shfl_up(w, o, 64);
                if ((int)threadIdx.x >= o) w += u;
            }
            sWave[threadIdx.x] = w;
        }
        __syncthreads();
        const int carry = sCarry;
        if (i < blocks) blockPrefix[i] = carry + (wave ? sWave[wave - 1] : 0) + v - f;
        __syncthreads();
        if (threadIdx.x == 0) sCarry = carry + sWave[kScanBlock / 64 - 1];
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = sCarry;
}

// Distinct slots from `counter` for the lanes with `want`, one atomic per wavefront (the
// lanes' ranks in the ballot); -1 for the others.  Every active lane of the wave calls it.
__device__ __forceinline__ int waveAlloc(int* counter, bool want) {
    const unsigned long long m = __ballot(want);
    if (!m) return -1;
    const int leader = __ffsll((unsigned long long)m) - 1;
    const int rank = (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
    int base = 0;
    if ((int)(threadIdx.x & 63) == leader) base = atomicAdd(counter, __popcll(m));
    base = __shfl(base, leader, 64);
    return want ? base + rank : -1;
}

// One lane per task: the task's node and, for a split, its two children in the slots the scan
// gave it (nodes nodeBase + 2k, 2k + 1), classified for the next level: large children take
// bin slots, small ones join the small or tiny list.
__global__ void emitKernel(LevelArgs a) {
    const int i = blockIdx.x * kB + threadIdx.x;
    if (i >= a.taskCount) return;
    const Task t = a.tasks[i];
    SplitInfo si = a.split[i];
    const int* bx = a.box + (size_t)i * 12;
    RtBvhNode nd;
    nd.low[0] = ordVal(bx[0]); nd.low[1] = ordVal(bx[1]); nd.low[2] = ordVal(bx[2]);
    nd.high[0] = ordVal(bx[3]); nd.high[1] = ordVal(bx[4]); nd.high[2] = ordVal(bx[5]);
    nd.first = t.first; nd.count = t.count;
    const bool isSplit = si.axis >= 0;
    const int k = isSplit ? a.childBase[i] + a.blockPrefix[i / kScanBlock] : 0;
    const int child = a.nodeBase + 2 * k;
    const int counts[2] = {si.leftCount, t.count - si.leftCount};
    const int firsts[2] = {t.first, t.first + si.leftCount};
    for (int c = 0; c < 2; c++) {   // convergent: the slot allocations are per wavefront
        Task ct;
        ct.node = child + c; ct.first = firsts[c]; ct.count = counts[c]; ct.depth = t.depth + 1;
        const bool leaf = ct.count <= a.maxLeaf || ct.depth == kBvhDepthLimit;
        const bool large = isSplit && !leaf && ct.count > kSmallMax;
        const bool tiny = isSplit && !leaf && !large && ct.count <= kTinyMax;
        const bool small = isSplit && !leaf && !large && !tiny;
        const int slot = waveAlloc(&a.counters[2], large);
        const int tinyAt = waveAlloc(&a.counters[7], tiny);
        const int smallAt = waveAlloc(&a.counters[5], small);
        if (isSplit) {
            ct.binSlot = leaf ? kLeafTask : (large ? slot : kSmallTask);
            a.nextTasks[2 * k + c] = ct;
            if (tiny) a.nextTinyList[tinyAt] = 2 * k + c;
            if (small) a.nextSmallList[smallAt] = 2 * k + c;
        }
    }
    if (isSplit) {
        nd.first = child; nd.count = 0;
        si.childTask = 2 * k;
        a.split[i] = si;
    }
    a.nodes[t.node] = nd;
}

// Inclusive scans inside groups of G lanes (lane = the lane's index in its group), over the
// first min(G, 32) lanes: the up scans combine lanes 0..lane, the down scans lane..n-1.
template <int G> __device__ __forceinline__ int scanMinUp(int v, int lane) {
    for (int o = 1; o < (G < 32 ? G : 32); o <<= 1) { const int u = __shfl_up(v, o, G); if (lane >= o) v = min(v, u); }
    return v;
}
template <int G> __device__ __forceinline__ int scanMaxUp(int v, int lane) {
    for (int o = 1; o < (G < 32 ? G : 32); o <<= 1) { const int u = __shfl_up(v, o, G); if (lane >= o) v = max(v, u); }
    return v;
}
template <int G> __device__ __forceinline__ int scanAddUp(int v, int lane) {
    for (int o = 1; o < (G < 32 ? G : 32); o <<= 1) { const int u = __shfl_up(v, o, G); if (lane >= o) v += u; }
    return v;
}
template <int G> __device__ __forceinline__ int scanMinDown(int v, int lane, int n) {
    for (int o = 1; o < (G < 32 ? G : 32); o <<= 1) { const int u = __shfl_down(v, o, G); if (lane + o < n) v = min(v, u); }
    return v;
}
template <int G> __device__ __forceinline__ int scanMaxDown(int v, int lane, int n) {
    for (int o = 1; o < (G < 32 ? G : 32); o <<= 1) { const int u = __shfl_down(v, o, G); if (lane + o < n) v = max(v, u); }
    return v;
}
template <int G> __device__ __forceinline__ int scanAddDown(int v, int lane, int n) {
    for (int o = 1; o < (G < 32 ? G : 32); o <<= 1) { const int u = __shfl_down(v, o, G); if (lane + o < n) v += u; }
    return v;
}

// sahSweep across a group of G >= nbins lanes, lane b = bin b: the left boxes and counts of
// bins 0..b and the right ones of b+1..nbins-1 by prefix / suffix scans (min, max and sums are
// exact, so every cost is the sequential sweep's float), then per axis the smallest cost at
// the lowest bin, and across axes the first strictly better than the best so far: the
// sequential sweep's choice.  The result is the same in every lane of the group.
template <int G>
__device__ void sahSweepGroup(const int* tb, const int* bx, int count, int nbins, int lane, SplitInfo& si) {
    const float area = halfAreaOf(ordVal(bx[0]), ordVal(bx[1]), ordVal(bx[2]), ordVal(bx[3]), ordVal(bx[4]), ordVal(bx[5]));
    float best = area * count;
    for (int axis = 0; axis < 3; axis++) {
        const float ext = ordVal(bx[9 + axis]) - ordVal(bx[6 + axis]);
        if (ext <= 1e-12f) continue;
        const int* e = tb + (axis * nbins + (lane < nbins ? lane : 0)) * kBinInts;
        const bool in = lane < nbins;
        const int c = in ? e[0] : 0;
        int lo[3], hi[3];
        for (int q = 0; q < 3; q++) {
            lo[q] = in ? e[1 + q] : 0x7F800000;
            hi[q] = in ? e[4 + q] : (int)(0xFF800000u ^ 0x7FFFFFFFu);
        }
        const int lc = scanAddUp<G>(c, lane), rc = scanAddDown<G>(c, lane, nbins);
        int llo[3], lhi[3], rlo[3], rhi[3];
        for (int q = 0; q < 3; q++) {
            llo[q] = scanMinUp<G>(lo[q], lane); lhi[q] = scanMaxUp<G>(hi[q], lane);
            rlo[q] = scanMinDown<G>(lo[q], lane, nbins); rhi[q] = scanMaxDown<G>(hi[q], lane, nbins);
        }
        // the right side of split b is the suffix from bin b + 1
        const int rcNext = __shfl_down(rc, 1, G);
        int rloN[3], rhiN[3];
        for (int q = 0; q < 3; q++) { rloN[q] = __shfl_down(rlo[q], 1, G); rhiN[q] = __shfl_down(rhi[q], 1, G); }
        const bool valid = lane < nbins - 1 && lc > 0 && rcNext > 0;
        float cost = INFINITY;
        if (valid) {
            const float al = halfAreaOf(ordVal(llo[0]), ordVal(llo[1]), ordVal(llo[2]), ordVal(lhi[0]), ordVal(lhi[1]),
                                        ordVal(lhi[2]));
            const float ar = halfAreaOf(ordVal(rloN[0]), ordVal(rloN[1]), ordVal(rloN[2]), ordVal(rhiN[0]),
                                        ordVal(rhiN[1]), ordVal(rhiN[2]));
            cost = area + al * lc + ar * rcNext;
        }
        // smallest cost, lowest bin on ties
        float mc = cost;
        int mb = valid ? lane : G;
        for (int o = G / 2; o > 0; o >>= 1) {
            const float oc = __shfl_xor(mc, o, G);
            const int ob = __shfl_xor(mb, o, G);
            if (oc < mc || (oc == mc && ob < mb)) { mc = oc; mb = ob; }
        }
        if (mb < G && mc < best) {
            best = mc;
            si.axis = axis; si.bin = mb; si.leftCount = __shfl(lc, mb, G);
        }
    }
}

// Small tasks are decided without global atomics by a group of G lanes each: a wavefront
// (G = 64) for tasks above kTinyMax primitives, 16 or 32 lanes (at least the bin count) for
// tiny ones, several groups to a wavefront.  The group reduces the task's bounds, bins its
// primitives in its own LDS bins (one atomic set per run of lanes in one bin when a whole
// wavefront bins a large range) and runs the SAH sweep with lane b holding bin b; the task's
// records move in smallMoveKernel.
template <int G> __device__ __forceinline__ void groupBox(int* b) {
    for (int f = 0; f < 12; f++) {
        const bool isMin = boxFieldIsMin(f);
        for (int o = G / 2; o > 0; o >>= 1) {
            const int v = __shfl_xor(b[f], o, G);
            b[f] = isMin ? min(b[f], v) : max(b[f], v);
        }
    }
}
constexpr int kRunBinMin = 256;   // a wavefront bins tasks this large by runs of lanes
template <int G>
__global__ __launch_bounds__(kB) void smallTaskKernel(LevelArgs a, const float2* pc, const int* list, int count) {
    constexpr int kGroups = kB / G, kGroupBins = G < kMaxBins ? G : kMaxBins;
    __shared__ int sBins[kGroups][3 * kGroupBins * kBinInts];
    const int group = threadIdx.x / G, lane = threadIdx.x % G;
    const int w = blockIdx.x * kGroups + group;
    if (w >= count) return;   // the whole group leaves; no workgroup barrier below
    const int i = list[w];
    const Task t = a.tasks[i];
    int b[12];
    for (int f = 0; f < 12; f++) b[f] = boxInit(f);
    const int end = t.first + t.count;
    for (int j = t.first + lane; j < end; j += G) boxFold(b, a.pa[j], a.pb[j], pc[j].x);
    groupBox<G>(b);
    int* bx = a.box + (size_t)i * 12;
    if (lane < 12) {
        int v = b[0];
        for (int f = 1; f < 12; f++) if (lane == f) v = b[f];
        bx[lane] = v;
    }
    int* bins = sBins[group];
    for (int k = lane; k < 3 * a.nbins; k += G) binInit(bins + k * kBinInts);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    float lo3[3], sc3[3];
    for (int q = 0; q < 3; q++) {
        const float lo = ordVal(b[6 + q]), ext = ordVal(b[9 + q]) - lo;
        lo3[q] = lo;
        sc3[q] = ext > 1e-12f ? a.nbins / ext : 0.0f;
    }
    if (G == 64 && t.count >= kRunBinMin) {
        for (int j0 = t.first; j0 < end; j0 += 64) {
            const int j = j0 + lane;
            const bool live = j < end;
            float4 l = make_float4(0, 0, 0, 0), h = l;
            float cz = 0;
            if (live) { l = a.pa[j]; h = a.pb[j]; cz = pc[j].x; }
            binAddRuns(bins, a.nbins, lo3, sc3, l, h, cz, live, 0, lane);
        }
    } else {
        for (int j = t.first + lane; j < end; j += G) binAdd(bins, a.nbins, lo3, sc3, a.pa[j], a.pb[j], pc[j].x);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
    SplitInfo si; si.axis = -1; si.bin = -1; si.leftCount = 0; si.childTask = -1;
    sahSweepGroup<G>(bins, b, t.count, a.nbins, lane, si);
    if (!splits(si, t.count)) si.axis = -1;
    if (lane == 0) {
        a.split[i] = si;
        a.splitFlag[i] = si.axis >= 0 ? 1 : 0;
    }
}

// One group of G lanes per small task (as smallTaskKernel), after emitKernel: its records to
// the left / right child ranges of the other buffer (ballot compaction in rounds of G, order
// kept), or, for a leaf, copied once and marked finished in both buffers.
struct MoveArgs {
    const float4 *pa, *pb;
    const float2* pc;
    float4 *paOut, *pbOut;
    float2* pcOut;
    int *posTask, *posTaskOut;
    const Task* tasks; const SplitInfo* split; const int* box;
    const int* list; int count; int nbins;
    int* error;
};
template <int G>
__global__ __launch_bounds__(kB) void smallMoveKernel(MoveArgs m) {
    const int group = threadIdx.x / G, lane = threadIdx.x % G;
    const int w = blockIdx.x * (kB / G) + group;
    if (w >= m.count) return;
    // this group's lanes in the wavefront's ballots
    const unsigned long long groupMask =
        G == 64 ? ~0ull : (((1ull << (G & 63)) - 1) << (((threadIdx.x & 63) / G) * G));
    const int i = m.list[w];
    const Task t = m.tasks[i];
    const SplitInfo si = m.split[i];
    const bool split = si.axis >= 0;
    const int end = t.first + t.count;
    float lo = 0, scale = 0;
    if (split) {
        const int* bx = m.box + (size_t)i * 12;
        lo = ordVal(bx[6 + si.axis]);
        scale = m.nbins / (ordVal(bx[9 + si.axis]) - lo);
    }
    int nl = 0, nr = 0;
    for (int j0 = t.first; j0 < end; j0 += G) {
        const int j = j0 + lane;
        const bool live = j < end;
        float4 l = make_float4(0, 0, 0, 0), h = l;
        float2 c = make_float2(0, 0);
        if (live) { l = m.pa[j]; h = m.pb[j]; c = m.pc[j]; }
        int right = 0;
        if (live && split) {
            const float cc = si.axis == 0 ? l.w : (si.axis == 1 ? h.w : c.x);
            right = binIndex(cc, lo, scale, m.nbins) <= si.bin ? 0 : 1;
        }
        const unsigned long long liveMask = __ballot(live) & groupMask;
        const unsigned long long mr = __ballot(live && right) & groupMask, ml = liveMask & ~mr;
        const unsigned long long mine = right ? mr : ml;
        const int rank = (int)__builtin_amdgcn_mbcnt_hi((unsigned)(mine >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)mine, 0u));
        if (live) {
            int dst = j, task = -1;
            if (split) {
                dst = right ? t.first + si.leftCount + nr + rank : t.first + nl + rank;
                task = si.childTask + right;
                if (dst >= (right ? end : t.first + si.leftCount)) {   // bins and partition disagree: refuse
                    atomicOr(m.error, 1);
                    dst = -1;
                }
            } else {
                m.posTask[j] = -1;   // a leaf: finished in both buffers
            }
            if (dst >= 0) {
                m.paOut[dst] = l; m.pbOut[dst] = h; m.pcOut[dst] = c;
                m.posTaskOut[dst] = task;
            }
        }
        nl += __popcll(ml);
        nr += __popcll(mr);
    }
}

// Moves the records of large and leaf tasks (small tasks move in smallMoveKernel): a split
// task's records go to its left / right child range in the other buffer; a record whose task
// became a leaf this level is copied once to the other buffer and marked finished in both (its
// position never changes again); finished records are skipped.  A workgroup takes 1024
// positions: when they all belong to one splitting task it places them with one cursor atomic
// per side (ranks from a workgroup prefix sum), else per position (one atomic per wave and side
// when the wave is uniform).
__global__ __launch_bounds__(kB) void partitionKernel(const float4* pa, const float4* pb, const float2* pc, int* posTask,
                                                      float4* paOut, float4* pbOut, float2* pcOut, int* posTaskOut,
                                                      const Task* tasks, const SplitInfo* split, const int* box,
                                                      int* cursors, int nbins, int n, int* error) {
    __shared__ int sUniform, sWaveL[kB / 64], sWaveR[kB / 64], sBaseL, sBaseR;
    const int base = blockIdx.x * kB * kPer;
    const int last = min(base + kB * kPer, n) - 1;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (threadIdx.x == 0) {
        const int t0 = posTask[base], t1 = posTask[last];
        sUniform = (t0 >= 0 && t0 == t1 && tasks[t0].binSlot >= 0 && split[t0].axis >= 0) ? t0 : -1;
    }
    __syncthreads();
    const int uni = sUniform;
    if (uni >= 0) {
        const SplitInfo si = split[uni];
        const int* bx = box + (size_t)uni * 12;
        const float lo = ordVal(bx[6 + si.axis]);
        const float scale = nbins / (ordVal(bx[9 + si.axis]) - lo);
        float4 l[kPer], h[kPer];
        float2 c[kPer];
        int right[kPer], nl = 0, nr = 0;
        for (int k = 0; k < kPer; k++) {
            const int j = base + k * kB + threadIdx.x;
            right[k] = -1;
            if (j > last) continue;
            l[k] = pa[j]; h[k] = pb[j]; c[k] = pc[j];
            const float cc = si.axis == 0 ? l[k].w : (si.axis == 1 ? h[k].w : c[k].x);
            right[k] = binIndex(cc, lo, scale, nbins) <= si.bin ? 0 : 1;
            nl += right[k] == 0; nr += right[k] == 1;
        }
        // exclusive prefix of (nl, nr) over the workgroup's threads
        int il = nl, ir = nr;
        for (int o = 1; o < 64; o <<= 1) {
            const int ul = __shfl_up(il, o, 64), ur = __shfl_up(ir, o, 64);
            if (lane >= o) { il += ul; ir += ur; }
        }
        if (lane == 63) { sWaveL[wave] = il; sWaveR[wave] = ir; }
        __syncthreads();
        if (threadIdx.x == 0) {
            int tl = 0, tr = 0;
            for (int w = 0; w < kB / 64; w++) {
                const int a = sWaveL[w], b = sWaveR[w];
                sWaveL[w] = tl; sWaveR[w] = tr;
                tl += a; tr += b;
            }
            sBaseL = tl ? atomicAdd(&cursors[2 * uni], tl) : 0;
            sBaseR = tr ? atomicAdd(&cursors[2 * uni + 1], tr) : 0;
            if (sBaseL + tl > si.leftCount || sBaseR + tr > tasks[uni].count - si.leftCount) atomicOr(error, 1);
        }
        __syncthreads();
        int offL = sBaseL + sWaveL[wave] + il - nl, offR = sBaseR + sWaveR[wave] + ir - nr;
        const int first = tasks[uni].first;
        for (int k = 0; k < kPer; k++) {
            if (right[k] < 0) continue;
            const int kk = right[k] ? offR++ : offL++;
            const int limit = right[k] ? tasks[uni].count - si.leftCount : si.leftCount;
            if (kk >= limit) continue;   // reported above: bins and partition disagree
            const int dst = first + (right[k] ? si.leftCount : 0) + kk;
            paOut[dst] = l[k]; pbOut[dst] = h[k]; pcOut[dst] = c[k];
            posTaskOut[dst] = si.childTask + right[k];
        }
        return;
    }
    for (int kq = 0; kq < kPer; kq++) {
        const int j = base + kq * kB + threadIdx.x;
        int t = j <= last ? posTask[j] : -2;
        if (t >= 0 && tasks[t].binSlot == kSmallTask) t = -2;   // moved by its task's wavefront
        const SplitInfo si = t >= 0 ? split[t] : SplitInfo{-1, -1, 0, -1};
        const bool moving = t >= 0 && si.axis >= 0;
        float4 l = make_float4(0, 0, 0, 0), h = l;
        float2 c = make_float2(0, 0);
        if (t >= 0) { l = pa[j]; h = pb[j]; c = pc[j]; }
        if (t >= 0 && !moving) {   // finishes here: into the other buffer once, finished in both
            paOut[j] = l; pbOut[j] = h; pcOut[j] = c;
            posTaskOut[j] = -1;
            posTask[j] = -1;
        }
        int right = 0;
        if (moving) {
            const float cc = si.axis == 0 ? l.w : (si.axis == 1 ? h.w : c.x);
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
        if (!moving) continue;
        const int limit = right ? tasks[t].count - si.leftCount : si.leftCount;
        if (k >= limit) { atomicOr(error, 1); continue; }   // bins and partition disagree: refuse
        const int dst = tasks[t].first + (right ? si.leftCount : 0) + k;
        paOut[dst] = l; pbOut[dst] = h; pcOut[dst] = c;
        posTaskOut[dst] = si.childTask + right;
    }
}

__global__ void orderKernel(const float2* pc, int* order, int n) {
    const int j = blockIdx.x * kB + threadIdx.x;
    if (j < n) order[j] = __float_as_int(pc[j].y);
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

// A build's temporaries come from one device buffer per device, kept for later builds (grown
// when a larger scene comes, never shrunk, held until the process ends) so that a build does
// not allocate and free about 230 bytes per triangle each time; builds on one device take
// turns on it.
struct BuildWorkspace { char* dev = nullptr; size_t bytes = 0; int* pinned = nullptr; };
constexpr int kMaxBuildDevices = 64;
std::mutex gWorkspaceMutex[kMaxBuildDevices];
BuildWorkspace gWorkspace[kMaxBuildDevices];
constexpr int kPinnedInts = 16;   // [0..7] level counters, [8..10] collapse counters

hipError_t reserveWorkspace(BuildWorkspace& ws, size_t bytes) {
    if (!ws.pinned) {
        const hipError_t e = hipHostMalloc(&ws.pinned, kPinnedInts * sizeof(int), hipHostMallocDefault);
        if (e != hipSuccess) { ws.pinned = nullptr; return e; }
    }
    if (bytes <= ws.bytes) return hipSuccess;
    if (ws.dev) (void)hipFree(ws.dev);
    ws.dev = nullptr;
    ws.bytes = 0;
    const hipError_t e = hipMalloc(&ws.dev, bytes);
    if (e == hipSuccess) ws.bytes = bytes;
    return e;
}

// Takes 256-byte aligned arrays from a buffer in order (with a null base it only adds up the
// bytes they need).
struct Carver {
    char* base;
    size_t off = 0;
    template <class T> T* take(size_t count) {
        T* p = base ? reinterpret_cast<T*>(base + off) : nullptr;
        off += (count * sizeof(T) + 255) & ~(size_t)255;
        return p;
    }
};

#define RT_TRY(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { err = e_; goto done; } } while (0)

}  // namespace

// Top-down collapse, one level of wide nodes per launch; its task lists come from the build
// workspace `ws` (free by now), the wide nodes are the caller's.
template <class N>
static hipError_t collapseW(const RtBvhNode* dBin, int binCount, char* ws, int* h, void** wideOut, int* wideCount,
                            int* wideDepth, hipStream_t s) {
    hipError_t err = hipSuccess;
    N* out = nullptr;
    Carver c{ws};
    // At most one wide node per binary inner node, plus the root.
    WideTask* ta = c.take<WideTask>((size_t)binCount + 1);
    WideTask* tb = c.take<WideTask>((size_t)binCount + 1);
    int* ctr = c.take<int>(3);
    int taskCount = 1;
    const WideTask root{0, 0, 0};
    RT_TRY(hipMalloc(&out, (size_t)(binCount + 1) * sizeof(N)));
    h[0] = 1; h[1] = 1; h[2] = 0;   // wide node count, next task count, max depth
    RT_TRY(hipMemcpyAsync(ta, &root, sizeof(root), hipMemcpyHostToDevice, s));
    RT_TRY(hipMemcpyAsync(ctr, h, 3 * sizeof(int), hipMemcpyHostToDevice, s));
    RT_TRY(hipStreamSynchronize(s));
    while (taskCount > 0) {
        RT_TRY(hipMemsetAsync(ctr + 1, 0, sizeof(int), s));
        hipLaunchKernelGGL((collapseKernel<N>), dim3(blocksFor(taskCount)), dim3(kB), 0, s, ta, taskCount, dBin, out,
                           ctr, tb);
        RT_TRY(hipGetLastError());
        RT_TRY(hipMemcpyAsync(h, ctr, 3 * sizeof(int), hipMemcpyDeviceToHost, s));
        RT_TRY(hipStreamSynchronize(s));
        taskCount = h[1];
        std::swap(ta, tb);
    }
    *wideOut = out; out = nullptr;
    *wideCount = h[0];
    *wideDepth = h[2];
done:
    if (out) (void)hipFree(out);
    return err;
}

// Temporaries of one build, carved from the workspace.  A level holds at most n tasks (each
// owns at least one primitive), at most n / (kSmallMax + 1) of them large.
struct BuildBuffers {
    float4 *pa, *pb, *pa2, *pb2;
    float2 *pc, *pc2;
    int *pos, *pos2, *lastFlag;
    Task *tasks, *next;
    SplitInfo* split;
    int *box, *bins, *cursors, *ctr, *small, *small2, *tiny, *tiny2, *flags, *childBase, *blockSums, *blockPrefix;
    size_t maxTasks, maxSlots;
    void carve(Carver& c, int n, int nbins) {
        maxTasks = (size_t)n;
        maxSlots = (size_t)n / (kSmallMax + 1) + 1;
        const size_t maxScan = (maxTasks + kScanBlock - 1) / kScanBlock;
        pa = c.take<float4>(n); pb = c.take<float4>(n); pa2 = c.take<float4>(n); pb2 = c.take<float4>(n);
        pc = c.take<float2>(n); pc2 = c.take<float2>(n);
        pos = c.take<int>(n); pos2 = c.take<int>(n); lastFlag = c.take<int>(n);
        tasks = c.take<Task>(maxTasks); next = c.take<Task>(maxTasks); split = c.take<SplitInfo>(maxTasks);
        box = c.take<int>(maxTasks * 12); bins = c.take<int>(maxSlots * 3 * nbins * kBinInts);
        cursors = c.take<int>(maxTasks * 2); ctr = c.take<int>(8);
        small = c.take<int>(maxTasks); small2 = c.take<int>(maxTasks);
        tiny = c.take<int>(maxTasks); tiny2 = c.take<int>(maxTasks);
        flags = c.take<int>(maxTasks); childBase = c.take<int>(maxTasks);
        blockSums = c.take<int>(maxScan); blockPrefix = c.take<int>(maxScan);
    }
};

template <int G> static void launchSmall(const LevelArgs& la, const float2* pc, const int* list, int count,
                                         const MoveArgs& mv, bool move, hipStream_t s) {
    const int blocks = (count + kB / G - 1) / (kB / G);
    if (!move) {
        hipLaunchKernelGGL(smallTaskKernel<G>, dim3(blocks), dim3(kB), 0, s, la, pc, list, count);
    } else {
        MoveArgs m = mv;
        m.list = list; m.count = count;
        hipLaunchKernelGGL(smallMoveKernel<G>, dim3(blocks), dim3(kB), 0, s, m);
    }
}

hipError_t buildBvhGpu(const Triangle* dTris, int n, const GpuBuildParams& prm, GpuBuildResult& res, hipStream_t s) {
    const int nbins = prm.bins;
    if (nbins < 2 || nbins > kMaxBins || n < 1 || n >= kLeafFirstLimit) return hipErrorInvalidValue;
    int device = 0;
    hipError_t err = hipGetDevice(&device);
    if (err != hipSuccess) return err;
    if (device < 0 || device >= kMaxBuildDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(gWorkspaceMutex[device]);
    BuildWorkspace& ws = gWorkspace[device];
    BuildBuffers bb;
    int *order = nullptr, *h = nullptr;
    RtBvhNode* nodes = nullptr;
    int taskCount = 1, binSlots = 0, smallCount = 0, tinyCount = 0, nodeCount = 1, level = 0, lastLevel = 0;
    const int tinyG = nbins <= 16 ? 16 : 32;   // lanes per tiny task: at least one per bin
    res.recs = nullptr;
    {
        Carver sizing{nullptr};
        bb.carve(sizing, n, nbins);
        Carver collapseSizing{nullptr};
        collapseSizing.take<WideTask>(2 * (size_t)n); collapseSizing.take<WideTask>(2 * (size_t)n); collapseSizing.take<int>(3);
        RT_TRY(reserveWorkspace(ws, std::max(sizing.off, collapseSizing.off)));
        Carver c{ws.dev};
        bb.carve(c, n, nbins);
        h = ws.pinned;   // [2] next bin slots, [4] partition error, [5] next small, [6] splits, [7] next tiny
    }
    RT_TRY(hipMalloc(&nodes, (size_t)(2 * n) * sizeof(RtBvhNode)));
    hipLaunchKernelGGL(primPrepKernel, dim3(blocksFor(n)), dim3(kB), 0, s, dTris, n, bb.pa, bb.pb, bb.pc, bb.pos);
    RT_TRY(hipGetLastError());
    {
        Task root{0, 0, n, 0, n <= prm.maxLeaf ? kLeafTask : (n <= kSmallMax ? kSmallTask : 0)};
        binSlots = root.binSlot >= 0 ? 1 : 0;
        const bool rootTiny = root.binSlot == kSmallTask && n <= kTinyMax;
        smallCount = root.binSlot == kSmallTask && !rootTiny ? 1 : 0;
        tinyCount = rootTiny ? 1 : 0;
        for (int k = 0; k < 8; k++) h[k] = 0;
        static const int zero = 0;
        RT_TRY(hipMemcpyAsync(bb.tasks, &root, sizeof(root), hipMemcpyHostToDevice, s));
        RT_TRY(hipMemcpyAsync(bb.small, &zero, sizeof(int), hipMemcpyHostToDevice, s));
        RT_TRY(hipMemcpyAsync(bb.tiny, &zero, sizeof(int), hipMemcpyHostToDevice, s));
        RT_TRY(hipMemcpyAsync(bb.ctr, h, 8 * sizeof(int), hipMemcpyHostToDevice, s));
        RT_TRY(hipStreamSynchronize(s));   // the root task is a stack value
    }
    // One level of tasks per round: bounds and bins of the large tasks (position-parallel),
    // every task's decision (leaf box, SAH split), an exclusive scan of the splits that places
    // the children (nodes nodeCount + 2k, tasks 2k, 2k + 1), then the records' move.
    while (taskCount > 0) {
        const int scanBlocks = (taskCount + kScanBlock - 1) / kScanBlock;
        if ((size_t)taskCount > bb.maxTasks || (size_t)binSlots > bb.maxSlots) {   // the bounds of BuildBuffers
            err = hipErrorUnknown;
            goto done;
        }
        RT_TRY(hipMemsetAsync(bb.cursors, 0, (size_t)taskCount * 2 * sizeof(int), s));
        if (binSlots > 0) {   // large tasks: position-parallel bounds and bins
            hipLaunchKernelGGL(initBoxKernel, dim3(blocksFor((long long)taskCount * 12)), dim3(kB), 0, s, bb.box,
                               taskCount * 12);
            hipLaunchKernelGGL(bigBoundsKernel, dim3(blocksFor(n, kPer * kChunks)), dim3(kB), 0, s, bb.pos, bb.tasks,
                               bb.pa, bb.pb, bb.pc, bb.box, n);
            RT_TRY(hipGetLastError());
            hipLaunchKernelGGL(initBinsKernel, dim3(blocksFor((long long)binSlots * 3 * nbins)), dim3(kB), 0, s, bb.bins,
                               (long long)binSlots * 3 * nbins);
            hipLaunchKernelGGL(bigBinKernel, dim3(blocksFor(n, kPer * kChunks)), dim3(kB), 0, s, bb.pos, bb.tasks, bb.pa,
                               bb.pb, bb.pc, bb.box, bb.bins, nbins, n);
            RT_TRY(hipGetLastError());
        }
        RT_TRY(hipMemsetAsync(bb.ctr + 2, 0, sizeof(int), s));
        RT_TRY(hipMemsetAsync(bb.ctr + 5, 0, 3 * sizeof(int), s));
        LevelArgs la;
        la.tasks = bb.tasks; la.taskCount = taskCount; la.box = bb.box; la.bins = bb.bins; la.nbins = nbins;
        la.maxLeaf = prm.maxLeaf; la.pa = bb.pa; la.pb = bb.pb; la.nodes = nodes; la.counters = bb.ctr;
        la.nextTasks = bb.next; la.split = bb.split; la.splitFlag = bb.flags; la.childBase = bb.childBase;
        la.blockPrefix = bb.blockPrefix; la.nodeBase = nodeCount; la.nextSmallList = bb.small2;
        la.nextTinyList = bb.tiny2;
        MoveArgs mv{bb.pa, bb.pb, bb.pc, bb.pa2, bb.pb2, bb.pc2, bb.pos, bb.pos2, bb.tasks, bb.split, bb.box,
                    nullptr, 0, nbins, bb.ctr + 4};
        hipLaunchKernelGGL(decideKernel, dim3(blocksFor(taskCount)), dim3(kB), 0, s, la);
        RT_TRY(hipGetLastError());
        if (smallCount > 0) launchSmall<64>(la, bb.pc, bb.small, smallCount, mv, false, s);
        if (tinyCount > 0) {
            if (tinyG == 16) launchSmall<16>(la, bb.pc, bb.tiny, tinyCount, mv, false, s);
            else launchSmall<32>(la, bb.pc, bb.tiny, tinyCount, mv, false, s);
        }
        RT_TRY(hipGetLastError());
        hipLaunchKernelGGL(scanBlocksKernel, dim3(scanBlocks), dim3(kScanBlock), 0, s, bb.flags, bb.childBase,
                           bb.blockSums, taskCount);
        hipLaunchKernelGGL(scanTopKernel, dim3(1), dim3(kScanBlock), 0, s, bb.blockSums, bb.blockPrefix, scanBlocks,
                           bb.ctr + 6);
        hipLaunchKernelGGL(emitKernel, dim3(blocksFor(taskCount)), dim3(kB), 0, s, la);
        RT_TRY(hipGetLastError());
        if (smallCount > 0) launchSmall<64>(la, bb.pc, bb.small, smallCount, mv, true, s);
        if (tinyCount > 0) {
            if (tinyG == 16) launchSmall<16>(la, bb.pc, bb.tiny, tinyCount, mv, true, s);
            else launchSmall<32>(la, bb.pc, bb.tiny, tinyCount, mv, true, s);
        }
        RT_TRY(hipGetLastError());
        hipLaunchKernelGGL(partitionKernel, dim3(blocksFor(n, kPer)), dim3(kB), 0, s, bb.pa, bb.pb, bb.pc, bb.pos, bb.pa2,
                           bb.pb2, bb.pc2, bb.pos2, bb.tasks, bb.split, bb.box, bb.cursors, nbins, n, bb.ctr + 4);
        RT_TRY(hipGetLastError());
        std::swap(bb.pa, bb.pa2); std::swap(bb.pb, bb.pb2); std::swap(bb.pc, bb.pc2); std::swap(bb.pos, bb.pos2);
        RT_TRY(hipMemcpyAsync(h, bb.ctr, 8 * sizeof(int), hipMemcpyDeviceToHost, s));
        RT_TRY(hipStreamSynchronize(s));
        if (h[4]) { err = hipErrorUnknown; goto done; }
        lastLevel = level++;
        nodeCount += 2 * h[6];
        taskCount = 2 * h[6];
        binSlots = h[2];
        smallCount = h[5];
        tinyCount = h[7];
        std::swap(bb.tasks, bb.next);
        std::swap(bb.small, bb.small2);
        std::swap(bb.tiny, bb.tiny2);
    }
    RT_TRY(hipMalloc(&order, (size_t)n * 4));
    hipLaunchKernelGGL(orderKernel, dim3(blocksFor(n)), dim3(kB), 0, s, bb.pc, order, n);
    RT_TRY(hipGetLastError());
    res.binNodeCount = nodeCount;
    res.depth = lastLevel;
    // leaf-ordered triangle records with last-in-leaf flags
    RT_TRY(hipMemsetAsync(bb.lastFlag, 0, (size_t)n * 4, s));
    hipLaunchKernelGGL(leafMarkKernel, dim3(blocksFor(res.binNodeCount)), dim3(kB), 0, s, nodes, res.binNodeCount,
                       bb.lastFlag);
    RT_TRY(hipGetLastError());
    RT_TRY(hipMalloc(&res.recs, (size_t)(n + kRecordPad) * sizeof(TriRecord)));
    RT_TRY(hipMemsetAsync(res.recs + n, 0, kRecordPad * sizeof(TriRecord), s));
    hipLaunchKernelGGL(recordKernel, dim3(blocksFor(n)), dim3(kB), 0, s, dTris, order, bb.lastFlag, res.recs, n);
    RT_TRY(hipGetLastError());
    // the collapse reuses the workspace from its start (stream order: the kernels above are done
    // with it before the collapse's first copy)
    {
        int* hc = h + 8;
        if (prm.quantized) {
            switch (prm.width) {
                case 2: RT_TRY(collapseW<QNode<2>>(nodes, res.binNodeCount, ws.dev, hc, &res.wideNodes, &res.wideCount, &res.wideDepth, s)); break;
                case 8: RT_TRY(collapseW<QNode<8>>(nodes, res.binNodeCount, ws.dev, hc, &res.wideNodes, &res.wideCount, &res.wideDepth, s)); break;
                default: RT_TRY(collapseW<QNode<4>>(nodes, res.binNodeCount, ws.dev, hc, &res.wideNodes, &res.wideCount, &res.wideDepth, s)); break;
            }
        } else {
            switch (prm.width) {
                case 2: RT_TRY(collapseW<WideNode<2>>(nodes, res.binNodeCount, ws.dev, hc, &res.wideNodes, &res.wideCount, &res.wideDepth, s)); break;
                case 8: RT_TRY(collapseW<WideNode<8>>(nodes, res.binNodeCount, ws.dev, hc, &res.wideNodes, &res.wideCount, &res.wideDepth, s)); break;
                default: RT_TRY(collapseW<WideNode<4>>(nodes, res.binNodeCount, ws.dev, hc, &res.wideNodes, &res.wideCount, &res.wideDepth, s)); break;
            }
        }
    }
    RT_TRY(hipStreamSynchronize(s));
    res.binNodes = nodes; nodes = nullptr;
    res.order = order; order = nullptr;
done:
    if (err != hipSuccess) {
        (void)hipStreamSynchronize(s);   // nothing in flight uses the workspace or the buffers below
        if (res.recs) { (void)hipFree(res.recs); res.recs = nullptr; }
    }
    if (order) (void)hipFree(order);
    if (nodes) (void)hipFree(nodes);
    return err;
}

}  // namespace rt
