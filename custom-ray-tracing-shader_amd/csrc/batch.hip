// batch.hip - many scenes and cameras rendered together: one launch per stage over a
// shared path queue, replayed as a captured hipGraph per frame.
//
// Every scene keeps its own GPU-built BVH (same build as a session's), rebased into global
// node / record / triangle arrays; a path slot's view is found from the slot, and the view
// gives the camera, settings, lights and root node.  Per-view sample counters live in device
// memory and advance on the device, so the captured frame replays unchanged.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <string>
#include <vector>
#include "bvh_gpu.h"
#include "tuning.h"
#include "kernels.h"

using namespace rt;

namespace {

// Moves one view's wide nodes into the global arrays: inner references by the node offset,
// leaf references by the record offset (the first record sits above the 3-bit count code).
template <class N>
__global__ void rebaseNodesKernel(N* nodes, int count, int nodeOffset, int recordOffset) {
    constexpr int W = NodeTraits<N>::kWidth;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= count) return;
    N& n = nodes[i];
    for (int c = 0; c < W; c++) {
        const int r = n.child[c];
        if (r == kEmptyChild) continue;
        n.child[c] = r >= 0 ? r + nodeOffset : ~(~r + (recordOffset << 3));
    }
}
template <class N>
void launchRebase(const GpuBuildResult& r, int nodeOffset, int recordOffset, hipStream_t s) {
    const int nv = r.wideCount;
    hipLaunchKernelGGL(rebaseNodesKernel<N>, dim3((nv + kBlock - 1) / kBlock), dim3(kBlock), 0, s, (N*)r.wideNodes, nv,
                       nodeOffset, recordOffset);
}

__global__ void rebaseRecordsKernel(TriRecord* recs, int count, int triangleOffset) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < count) recs[i].index += triangleOffset;
}

struct PendingView {
    std::vector<Triangle> tris;
    int w, h;
    CameraData cam;
    RenderSettings rs;
};

struct Batch {
    std::vector<PendingView> pending;
    std::string error;
    int width = 4, quantized = 0, blocksPerCU = 3, shortStack = 12, refill = 1, refillBelow = 40, whileWhile = 1,
        useGraph = 1, pipelines = 3, xcdSegments = 1;
    unsigned userSet = 0;             // tuned options the caller set (bit k: kTunedFields[k])
    int autotuneCache = 1;            // 1: RtBatchBuild applies cached launch options the caller did not set
    bool built = false;
    int cus = 0, views = 0, totalSlots = 0, maxBounces = 0;
    long long totalTris = 0;
    std::vector<int> viewBase;        // views + 1
    std::vector<int> pixelBase;       // offsets of each view's sums / pixels
    // device
    Triangle* dTris = nullptr;
    Light* dLights = nullptr;
    void* dNodes = nullptr;
    TriRecord* dRecs = nullptr;
    BatchView* dViews = nullptr;
    int *dViewBase = nullptr, *dRoots = nullptr;
    unsigned* dSamples = nullptr;
    float4 *dSumRad = nullptr, *dSumAlb = nullptr, *dSumNrm = nullptr;
    uchar4* dPixels = nullptr;
    float4 *dThr = nullptr, *dRad = nullptr, *dAlb = nullptr, *dNrm = nullptr, *dNee = nullptr;
    float4 *dRayA = nullptr, *dRayB = nullptr, *dShadow = nullptr;
    int4* dHits = nullptr;
    int *dCounters = nullptr, *dSpill = nullptr, *dOverflow = nullptr;   // counters: kBatchPipelines x kCounterStride
    int* dXcdWork = nullptr;          // kBatchPipelines x kXcdSlotInts segment cursors
    int spillEntries = 0, traceBlocks = 0;
    hipStream_t stream = nullptr;
    hipStream_t pipeStream[3] = {};   // path pipelines 1..3 (default priority, as stream)
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipEvent_t evFork = nullptr, evJoin[3] = {};
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    float buildMs = 0;
    unsigned long long lastClosest = 0, lastShadow = 0;
};

constexpr int kBatchPipelines = 4;   // path pipelines per frame at most

template <class T> bool dalloc(Batch* b, T*& p, size_t count) {
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&p), std::max<size_t>(count * sizeof(T), 16));
    if (e != hipSuccess) { p = nullptr; b->error = std::string("device allocation: ") + hipGetErrorString(e); return false; }
    return true;
}
template <class T> void dfree(T*& p) { if (p) (void)hipFree(p); p = nullptr; }

void release(Batch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr; b->graph = nullptr;
    dfree(b->dTris); dfree(b->dLights); dfree(b->dNodes); dfree(b->dRecs); dfree(b->dViews); dfree(b->dViewBase);
    dfree(b->dRoots); dfree(b->dSamples); dfree(b->dSumRad); dfree(b->dSumAlb); dfree(b->dSumNrm); dfree(b->dPixels);
    dfree(b->dThr); dfree(b->dRad); dfree(b->dAlb); dfree(b->dNrm); dfree(b->dNee); dfree(b->dRayA); dfree(b->dRayB);
    dfree(b->dShadow); dfree(b->dHits); dfree(b->dCounters); dfree(b->dXcdWork); dfree(b->dSpill); dfree(b->dOverflow);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->evFork) (void)hipEventDestroy(b->evFork);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    for (int k = 0; k < kBatchPipelines - 1; k++) {
        if (b->evJoin[k]) (void)hipEventDestroy(b->evJoin[k]);
        if (b->pipeStream[k]) (void)hipStreamDestroy(b->pipeStream[k]);
        b->evJoin[k] = nullptr; b->pipeStream[k] = nullptr;
    }
    b->ev0 = b->ev1 = b->evFork = nullptr; b->stream = nullptr;
    b->built = false;
}

#define B_TRY(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { b->error = std::string(#x) + ": " + hipGetErrorString(e_); return false; } } while (0)

BatchArgs batchArgs(Batch* b) {
    BatchArgs a{};
    a.views = b->dViews; a.viewBase = b->dViewBase; a.viewCount = b->views; a.totalSlots = b->totalSlots;
    a.samples = b->dSamples;
    a.thr = b->dThr; a.rad = b->dRad; a.alb = b->dAlb; a.nrm = b->dNrm; a.nee = b->dNee;
    a.shadow = b->dShadow; a.hits = b->dHits; a.counters = b->dCounters;
    return a;
}

// The launches of one frame (recorded into the graph, or issued directly).
TraceLaunch launchOf(const Batch* b) {
    TraceLaunch l;
    l.width = b->width; l.quantized = b->quantized != 0; l.stack = b->shortStack;
    l.refill = b->refill != 0; l.whileWhile = b->whileWhile != 0; l.refillBelow = b->refillBelow;
    return l;
}

// Path pipelines of a frame: its queue positions split in P parts (>= 64K paths each), each
// on its own stream with its own counters, queue regions and spill ring, as in a session.
int framePipelines(const Batch* b) { return std::max(1, std::min(b->pipelines, b->totalSlots >> 16)); }

// Segment cursors of pipeline p's launch k, or null when the option is off.
int* xcdCursors(Batch* b, int p, int k) {
    return b->xcdSegments ? b->dXcdWork + (size_t)p * kXcdSlotInts + (size_t)k * kXcdCursorInts : nullptr;
}

bool enqueueFrame(Batch* b) {
    hipStream_t st = b->stream;
    const int P = framePipelines(b);
    // A frame starts with a clear overflow flag (a fill kernel: the frame may be a captured graph).
    B_TRY(launchZeroInts(b->dOverflow, 1, st));
    hipStream_t pst[kBatchPipelines] = {st, b->pipeStream[0], b->pipeStream[1], b->pipeStream[2]};
    if (P > 1) {
        B_TRY(hipEventRecord(b->evFork, st));
        for (int p = 1; p < P; p++) B_TRY(hipStreamWaitEvent(pst[p], b->evFork, 0));
    }
    const int part = (b->totalSlots / P + 63) / 64 * 64;
    BatchArgs a[kBatchPipelines];
    for (int p = 0; p < P; p++) {
        const int qBase = p * part, qCount = p + 1 < P ? part : b->totalSlots - qBase;
        a[p] = batchArgs(b);
        a[p].counters = b->dCounters + (size_t)p * kCounterStride;
        a[p].shadow = b->dShadow + 2 * (size_t)qBase; a[p].hits = b->dHits + qBase;
        a[p].qBase = qBase; a[p].qCount = qCount;
        B_TRY(launchResetCounters(a[p].counters, qCount, xcdCursors(b, p, 0), pst[p]));
        a[p].rayOut = b->dRayA + 2 * (size_t)qBase;
        B_TRY(launchBatchRaygen(a[p], pst[p]));
    }
    // As a session frame: one launch traces bounce k's closest-hit rays and bounce k-1's
    // shadow rays; the last launch only the last bounce's shadow rays.
    for (int bounce = 0; bounce <= b->maxBounces; bounce++) {
        const bool rays = bounce < b->maxBounces;
        for (int p = 0; p < P; p++) {
            const size_t off = 2 * (size_t)a[p].qBase;
            float4* in = ((bounce & 1) ? b->dRayB : b->dRayA) + off;
            float4* out = ((bounce & 1) ? b->dRayA : b->dRayB) + off;
            TraceArgs ta{};
            ta.nodes = b->dNodes; ta.recs = b->dRecs;
            ta.spill = b->dSpill + (size_t)p * b->spillEntries * b->traceBlocks * kBlock * 2;
            ta.spillEntries = b->spillEntries;
            ta.totalLanes = b->traceBlocks * kBlock; ta.overflow = b->dOverflow;
            ta.roots = b->dRoots; ta.viewBase = b->dViewBase; ta.views = b->views; ta.nontemporal = 1;
            ta.work = a[p].counters + kWorkRays + bounce;
            ta.xcdWork = xcdCursors(b, p, bounce);
            if (rays) { ta.rays = in; ta.countPtr = a[p].counters + kCntRays + bounce; ta.hits = a[p].hits; }
            if (bounce > 0) {
                ta.anyRays = a[p].shadow; ta.anyCountPtr = a[p].counters + kCntShadow + (bounce - 1);
                ta.rad = b->dRad; ta.contrib = b->dNee;
            }
            B_TRY(launchTrace(launchOf(b), false, b->traceBlocks, ta, pst[p]));
            if (!rays) continue;
            a[p].bounce = bounce; a[p].rayIn = in; a[p].rayOut = out;
            B_TRY(launchBatchShade(a[p], b->cus * shadeBlocksPerCU(), pst[p]));
        }
    }
    for (int p = 1; p < P; p++) {
        B_TRY(hipEventRecord(b->evJoin[p - 1], pst[p]));
        B_TRY(hipStreamWaitEvent(st, b->evJoin[p - 1], 0));
    }
    B_TRY(launchBatchAccumulate(a[0], st));
    B_TRY(launchBatchAdvance(a[0], st));
    return true;
}

// The traversal grid and spill rings for the current launch options, and the frame's graph.
bool prepareLaunch(Batch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr; b->graph = nullptr;
    dfree(b->dSpill);
    b->traceBlocks = b->cus * b->blocksPerCU;
    b->spillEntries = 36 * (b->width - 1);
    if (!dalloc(b, b->dSpill, (size_t)kBatchPipelines * b->spillEntries * b->traceBlocks * kBlock * 2)) return false;
    if (b->useGraph) {
        B_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        const bool ok = enqueueFrame(b);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(b->stream, &g);
        if (!ok) { if (g) (void)hipGraphDestroy(g); return false; }
        B_TRY(e);
        b->graph = g;
        B_TRY(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    }
    return true;
}

// Scene class of the pending views (the autotuner's cache key).
std::string batchKey(const Batch* b) {
    long long tris = 0, px = 0;
    int bounces = 0;
    for (const PendingView& pv : b->pending) {
        tris += (long long)pv.tris.size();
        px += (long long)pv.w * pv.h;
        bounces = std::max(bounces, pv.rs.maxBounces);
    }
    return batchTuneKey((int)b->pending.size(), tris, px, bounces);
}

int setBatchOption(Batch* b, const char* key, int value) {
    if (!std::strcmp(key, "bvh_width") && (value == 2 || value == 4 || value == 8)) { b->width = value; return 1; }
    if (!std::strcmp(key, "bvh_quantized") && (value == 0 || value == 1)) { b->quantized = value; return 1; }
    if (!std::strcmp(key, "blocks_per_cu") && value >= 1 && value <= 32) { b->blocksPerCU = value; return 1; }
    if (!std::strcmp(key, "short_stack") && (value == 8 || value == 12 || value == 16)) { b->shortStack = value; return 1; }
    if (!std::strcmp(key, "refill") && (value == 0 || value == 1)) { b->refill = value; return 1; }
    if (!std::strcmp(key, "refill_below") && value >= 1 && value <= 64) { b->refillBelow = value; return 1; }
    if (!std::strcmp(key, "while_while") && (value == 0 || value == 1)) { b->whileWhile = value; return 1; }
    if (!std::strcmp(key, "use_graph") && (value == 0 || value == 1)) { b->useGraph = value; return 1; }
    if (!std::strcmp(key, "pipelines") && value >= 1 && value <= kBatchPipelines) { b->pipelines = value; return 1; }
    if (!std::strcmp(key, "xcd_segments") && (value == 0 || value == 1)) { b->xcdSegments = value; return 1; }
    if (!std::strcmp(key, "autotune_cache") && (value == 0 || value == 1)) { b->autotuneCache = value; return 1; }
    return 0;
}

// Options of o (path completion and leaf size have no batch counterpart) that the caller
// did not set.
void applyTuned(Batch* b, const LaunchOptions& o) {
    for (int k = 0; k < 8; k++) {
        const int v = o.*(kTunedFields[k].field);
        if (v >= 0 && !(b->userSet & (1u << k))) (void)setBatchOption(b, kTunedFields[k].name, v);
    }
}

bool build(Batch* b) {
    if (b->pending.empty()) { b->error = "batch has no views"; return false; }
    auto t0 = std::chrono::steady_clock::now();
    int device = 0;
    B_TRY(hipGetDevice(&device));
    B_TRY(hipDeviceGetAttribute(&b->cus, hipDeviceAttributeMultiprocessorCount, device));
    B_TRY(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    B_TRY(hipEventCreate(&b->ev0));
    B_TRY(hipEventCreate(&b->ev1));
    B_TRY(hipEventCreateWithFlags(&b->evFork, hipEventDisableTiming));
    for (int k = 0; k < kBatchPipelines - 1; k++) {
        B_TRY(hipStreamCreateWithFlags(&b->pipeStream[k], hipStreamNonBlocking));
        B_TRY(hipEventCreateWithFlags(&b->evJoin[k], hipEventDisableTiming));
    }
    b->views = (int)b->pending.size();
    // Host preprocessing per scene, as RtInitialize (Native/renderer.cu:345-353).
    std::vector<Triangle> tris;
    std::vector<Light> lights;
    std::vector<int> triOff, lightOff, lightCount;
    std::vector<float> lightArea;
    for (const PendingView& pv : b->pending) {
        triOff.push_back((int)tris.size());
        lightOff.push_back((int)lights.size());
        float area = 0;
        for (size_t i = 0; i < pv.tris.size(); i++) {
            Triangle t = pv.tris[i];
            t.color = RtVec3{srgbToLinear(t.color.x), srgbToLinear(t.color.y), srgbToLinear(t.color.z)};
            t.specularColor = RtVec3{srgbToLinear(t.specularColor.x), srgbToLinear(t.specularColor.y), srgbToLinear(t.specularColor.z)};
            if (t.emission > 0) {
                f3 an = crossv(V(t.edge1), V(t.edge2));
                area += .5f * sqrtf(dotv(an, an));
                lights.push_back(Light{(int)tris.size(), area});
            }
            tris.push_back(t);
        }
        lightCount.push_back((int)lights.size() - lightOff.back());
        lightArea.push_back(area);
    }
    b->totalTris = (long long)tris.size();
    if (!dalloc(b, b->dTris, tris.size()) || !dalloc(b, b->dLights, lights.size())) return false;
    B_TRY(hipMemcpy(b->dTris, tris.data(), tris.size() * sizeof(Triangle), hipMemcpyHostToDevice));
    if (!lights.empty()) B_TRY(hipMemcpy(b->dLights, lights.data(), lights.size() * sizeof(Light), hipMemcpyHostToDevice));
    // Per-scene GPU builds, then one global node / record array.
    std::vector<GpuBuildResult> res(b->views);
    GpuBuildParams prm;
    prm.width = b->width;
    prm.quantized = b->quantized;
    if ((long long)tris.size() >= kLeafFirstLimit) {
        b->error = "batch holds too many triangles for leaf references";
        return false;
    }
    long long nodeTotal = 0;
    for (int v = 0; v < b->views; v++) {
        B_TRY(buildBvhGpu(b->dTris + triOff[v], (int)b->pending[v].tris.size(), prm, res[v], b->stream));
        nodeTotal += res[v].wideCount;
    }
    if (b->width == 4 && !b->quantized && nodeTotal >= kMaxWide4Nodes) {
        b->error = "batch needs more than 2^25 4-wide nodes (4 GB); use bvh_width 8 or bvh_quantized";
        return false;
    }
    const size_t nodeBytes = wideNodeBytes(b->width, b->quantized);
    unsigned char* nodes = nullptr;
    if (!dalloc(b, nodes, (size_t)nodeTotal * nodeBytes) || !dalloc(b, b->dRecs, tris.size() + kRecordPad)) return false;
    B_TRY(hipMemsetAsync(b->dRecs + tris.size(), 0, kRecordPad * sizeof(TriRecord), b->stream));
    b->dNodes = nodes;
    std::vector<int> roots(b->views);
    long long nodeOff = 0;
    for (int v = 0; v < b->views; v++) {
        const int nv = res[v].wideCount, nt = (int)b->pending[v].tris.size();
        const int no = (int)nodeOff, ro = triOff[v];
        if (b->quantized) {
            if (b->width == 2) launchRebase<QNode<2>>(res[v], no, ro, b->stream);
            else if (b->width == 8) launchRebase<QNode<8>>(res[v], no, ro, b->stream);
            else launchRebase<QNode<4>>(res[v], no, ro, b->stream);
        } else {
            if (b->width == 2) launchRebase<WideNode<2>>(res[v], no, ro, b->stream);
            else if (b->width == 8) launchRebase<WideNode<8>>(res[v], no, ro, b->stream);
            else launchRebase<WideNode<4>>(res[v], no, ro, b->stream);
        }
        hipLaunchKernelGGL(rebaseRecordsKernel, dim3((nt + kBlock - 1) / kBlock), dim3(kBlock), 0, b->stream, res[v].recs, nt, triOff[v]);
        B_TRY(hipGetLastError());
        B_TRY(hipMemcpyAsync(nodes + nodeOff * nodeBytes, res[v].wideNodes, (size_t)nv * nodeBytes, hipMemcpyDeviceToDevice, b->stream));
        B_TRY(hipMemcpyAsync(b->dRecs + triOff[v], res[v].recs, (size_t)nt * sizeof(TriRecord), hipMemcpyDeviceToDevice, b->stream));
        roots[v] = (int)nodeOff;
        nodeOff += nv;
    }
    B_TRY(hipStreamSynchronize(b->stream));
    for (GpuBuildResult& r : res) {
        dfree(r.binNodes); dfree(r.order); dfree(r.recs);
        if (r.wideNodes) (void)hipFree(r.wideNodes);
        r.wideNodes = nullptr;
    }
    // Views, slots and accumulation buffers.
    b->viewBase.assign(1, 0);
    long long slots = 0;
    for (const PendingView& pv : b->pending) {
        slots += (long long)pv.w * pv.h;
        b->viewBase.push_back((int)slots);
        b->maxBounces = std::max(b->maxBounces, pv.rs.maxBounces);
    }
    if (slots > (1LL << 28)) { b->error = "batch exceeds 2^28 pixels"; return false; }
    b->totalSlots = (int)slots;
    const size_t n = (size_t)slots;
    if (!dalloc(b, b->dSumRad, n) || !dalloc(b, b->dSumAlb, n) || !dalloc(b, b->dSumNrm, n) || !dalloc(b, b->dPixels, n) ||
        !dalloc(b, b->dThr, n) || !dalloc(b, b->dRad, n) || !dalloc(b, b->dAlb, n) || !dalloc(b, b->dNrm, n) ||
        !dalloc(b, b->dNee, n) || !dalloc(b, b->dRayA, 2 * n) || !dalloc(b, b->dRayB, 2 * n) ||
        !dalloc(b, b->dShadow, 2 * n) || !dalloc(b, b->dHits, n) || !dalloc(b, b->dCounters, (size_t)kBatchPipelines * kCounterStride) ||
        !dalloc(b, b->dXcdWork, (size_t)kBatchPipelines * kXcdSlotInts) ||
        !dalloc(b, b->dOverflow, 1) || !dalloc(b, b->dViews, b->views) || !dalloc(b, b->dViewBase, b->views + 1) ||
        !dalloc(b, b->dRoots, b->views) || !dalloc(b, b->dSamples, b->views))
        return false;
    std::vector<BatchView> hv(b->views);
    for (int v = 0; v < b->views; v++) {
        const PendingView& pv = b->pending[v];
        BatchView& V = hv[v];
        V.cam = pv.cam; V.rs = pv.rs; V.rs.raysPerPixel = 1;
        V.sc.tris = b->dTris; V.sc.lights = b->dLights + lightOff[v]; V.sc.lightCount = lightCount[v];
        V.sc.lightArea = lightArea[v];
        V.w = pv.w; V.h = pv.h; V.slotBase = b->viewBase[v];
        V.sumRad = b->dSumRad + b->viewBase[v]; V.sumAlb = b->dSumAlb + b->viewBase[v];
        V.sumNrm = b->dSumNrm + b->viewBase[v]; V.pixels = b->dPixels + b->viewBase[v];
    }
    B_TRY(hipMemcpy(b->dViews, hv.data(), hv.size() * sizeof(BatchView), hipMemcpyHostToDevice));
    B_TRY(hipMemcpy(b->dViewBase, b->viewBase.data(), b->viewBase.size() * sizeof(int), hipMemcpyHostToDevice));
    B_TRY(hipMemcpy(b->dRoots, roots.data(), roots.size() * sizeof(int), hipMemcpyHostToDevice));
    B_TRY(hipMemset(b->dSamples, 0, b->views * sizeof(unsigned)));
    B_TRY(hipMemset(b->dOverflow, 0, sizeof(int)));
    if (!prepareLaunch(b)) return false;
    b->buildMs = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
    b->built = true;
    return true;
}

}  // namespace

extern "C" {

void* RtBatchCreate(void) { return new Batch(); }

void RtBatchDestroy(void* p) {
    Batch* b = static_cast<Batch*>(p);
    if (!b) return;
    release(b);
    delete b;
}

const char* RtBatchLastError(void* p) {
    Batch* b = static_cast<Batch*>(p);
    return b ? b->error.c_str() : "invalid argument";
}

int RtBatchAddView(void* p, const Triangle* triangles, int count, int width, int height, const CameraData* camera,
                   const RenderSettings* settings) {
    Batch* b = static_cast<Batch*>(p);
    if (!b) return 0;
    if (b->built) { b->error = "batch already built"; return 0; }
    if (!triangles || count <= 0 || width <= 0 || height <= 0 || width > 4096 || height > 4096 || !camera || !settings ||
        settings->maxBounces < 1 || settings->maxBounces > kMaxBounces || !(settings->exposure > 0) ||
        !(settings->aperture >= 0) || !(settings->focusDistance > 0) || !(settings->skyStrength >= 0)) {
        b->error = "invalid view";
        return 0;
    }
    PendingView pv;
    pv.tris.assign(triangles, triangles + count);
    pv.w = width; pv.h = height; pv.cam = *camera; pv.rs = *settings;
    b->pending.push_back(std::move(pv));
    return (int)b->pending.size();
}

int RtBatchSetOption(void* p, const char* key, int value) {
    Batch* b = static_cast<Batch*>(p);
    if (!b || !key || b->built) return 0;
    if (!setBatchOption(b, key, value)) return 0;
    for (int k = 0; k < 8; k++)
        if (!std::strcmp(key, kTunedFields[k].name)) b->userSet |= 1u << k;
    return 1;
}

int RtBatchGetOption(void* p, const char* key) {
    Batch* b = static_cast<Batch*>(p);
    if (!b || !key) return -1;
    if (!std::strcmp(key, "bvh_width")) return b->width;
    if (!std::strcmp(key, "bvh_quantized")) return b->quantized;
    if (!std::strcmp(key, "blocks_per_cu")) return b->blocksPerCU;
    if (!std::strcmp(key, "short_stack")) return b->shortStack;
    if (!std::strcmp(key, "refill")) return b->refill;
    if (!std::strcmp(key, "refill_below")) return b->refillBelow;
    if (!std::strcmp(key, "while_while")) return b->whileWhile;
    if (!std::strcmp(key, "use_graph")) return b->useGraph;
    if (!std::strcmp(key, "pipelines")) return b->pipelines;
    if (!std::strcmp(key, "xcd_segments")) return b->xcdSegments;
    if (!std::strcmp(key, "autotune_cache")) return b->autotuneCache;
    return -1;
}

int RtBatchBuild(void* p, float* buildMs) {
    Batch* b = static_cast<Batch*>(p);
    if (!b) return 0;
    if (b->built) return 1;
    LaunchOptions tuned;
    if (b->autotuneCache && !b->pending.empty() && tuneLookup(batchKey(b), false, tuned)) applyTuned(b, tuned);
    if (!build(b)) { release(b); return 0; }
    if (buildMs) *buildMs = b->buildMs;
    return 1;
}

int RtBatchAutotune(void* p, int frames, float* mraysPerSecond, int* trialCount) {
    Batch* b = static_cast<Batch*>(p);
    if (!b || frames < 1 || b->pending.empty()) { if (b) b->error = "invalid argument"; return 0; }
    const int graph = b->useGraph;
    TuneSpace space;
    space.widths = (b->userSet & 1u) ? std::vector<int>{b->width} : std::vector<int>{4, 8};
    space.leaves = {-1};     // the batch build keeps the reference's leaf rule
    space.tails.clear();     // no path completion in batches
    int maxBounces = 0;
    for (const PendingView& pv : b->pending) maxBounces = std::max(maxBounces, pv.rs.maxBounces);
    bool failed = false;
    auto measure = [&](const LaunchOptions& c) -> double {
        if (failed) return -1;
        const bool rebuild = !b->built || c.bvhWidth != b->width;
        (void)setBatchOption(b, "bvh_width", c.bvhWidth);
        (void)setBatchOption(b, "blocks_per_cu", c.blocksPerCU);
        (void)setBatchOption(b, "short_stack", c.shortStack);
        (void)setBatchOption(b, "while_while", c.whileWhile);
        (void)setBatchOption(b, "refill", c.refill);
        (void)setBatchOption(b, "pipelines", c.pipelines);
        if (rebuild) {
            release(b);
            if (!build(b)) { release(b); failed = true; return -1; }
        } else if (!prepareLaunch(b)) {
            failed = true;
            return -1;
        }
        float ms = 0;
        if (!RtBatchRender(b, 1, &ms)) { failed = true; return -1; }   // warm
        double best = 0;
        for (int rep = 0; rep < 2; rep++) {
            unsigned long long rays = 0;
            const auto t0 = std::chrono::steady_clock::now();
            for (int f = 0; f < frames; f++) {
                if (!RtBatchRender(b, 1, &ms)) { failed = true; return -1; }
                rays += b->lastClosest + b->lastShadow;
            }
            const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            best = std::max(best, rays / s / 1e6);
        }
        return best;
    };
    int trials = 0;
    const LaunchOptions best = tuneSearch(space, maxBounces, measure, &trials);
    b->useGraph = graph;
    if (failed || best.mraysPerSecond <= 0) { if (b->error.empty()) b->error = "autotune trial failed"; return 0; }
    const bool rebuild = best.bvhWidth != b->width;
    (void)setBatchOption(b, "bvh_width", best.bvhWidth);
    (void)setBatchOption(b, "blocks_per_cu", best.blocksPerCU);
    (void)setBatchOption(b, "short_stack", best.shortStack);
    (void)setBatchOption(b, "while_while", best.whileWhile);
    (void)setBatchOption(b, "refill", best.refill);
    (void)setBatchOption(b, "pipelines", best.pipelines);
    if (rebuild) { release(b); if (!build(b)) { release(b); return 0; } }
    else if (!prepareLaunch(b)) return 0;
    if (!RtBatchReset(b)) return 0;
    (void)tuneStore(batchKey(b), best);
    if (mraysPerSecond) *mraysPerSecond = (float)best.mraysPerSecond;
    if (trialCount) *trialCount = trials;
    return 1;
}

int RtBatchRender(void* p, int frames, float* deviceMs) {
    Batch* b = static_cast<Batch*>(p);
    if (!b || !b->built || frames < 1 || !deviceMs) return 0;
    B_TRY(hipEventRecord(b->ev0, b->stream));
    for (int f = 0; f < frames; f++) {
        if (b->exec) B_TRY(hipGraphLaunch(b->exec, b->stream));
        else if (!enqueueFrame(b)) return 0;
    }
    B_TRY(hipEventRecord(b->ev1, b->stream));
    B_TRY(hipEventSynchronize(b->ev1));
    B_TRY(hipEventElapsedTime(deviceMs, b->ev0, b->ev1));
    int ctr[kBatchPipelines * kCounterStride];
    int overflow = 0;
    B_TRY(hipMemcpy(ctr, b->dCounters, sizeof(ctr), hipMemcpyDeviceToHost));
    B_TRY(hipMemcpy(&overflow, b->dOverflow, sizeof(int), hipMemcpyDeviceToHost));
    if (overflow) { b->error = "traversal stack exceeded its spill ring"; return 0; }
    b->lastClosest = b->lastShadow = 0;
    for (int p = 0; p < framePipelines(b); p++)
        for (int k = 0; k < b->maxBounces; k++) {
            b->lastClosest += (unsigned)ctr[p * kCounterStride + kCntRays + k];
            b->lastShadow += (unsigned)ctr[p * kCounterStride + kCntShadow + k];
        }
    return 1;
}

int RtBatchReset(void* p) {
    Batch* b = static_cast<Batch*>(p);
    if (!b || !b->built) return 0;
    B_TRY(hipMemsetAsync(b->dSamples, 0, b->views * sizeof(unsigned), b->stream));
    B_TRY(hipStreamSynchronize(b->stream));
    return 1;
}

int RtBatchRayCounts(void* p, unsigned long long* closest, unsigned long long* shadow) {
    Batch* b = static_cast<Batch*>(p);
    if (!b || !closest || !shadow) return 0;
    *closest = b->lastClosest; *shadow = b->lastShadow;
    return 1;
}

int RtBatchViewInfo(void* p, int view, int* width, int* height, unsigned* samples) {
    Batch* b = static_cast<Batch*>(p);
    if (!b || view < 0 || view >= (int)b->pending.size()) return 0;
    if (width) *width = b->pending[view].w;
    if (height) *height = b->pending[view].h;
    if (samples) {
        *samples = 0;
        if (b->built) B_TRY(hipMemcpy(samples, b->dSamples + view, sizeof(unsigned), hipMemcpyDeviceToHost));
    }
    return 1;
}

int RtBatchReadView(void* p, int view, unsigned char* rgba, int bytes, float* accumulation, long long floatCount) {
    Batch* b = static_cast<Batch*>(p);
    if (!b || !b->built || view < 0 || view >= b->views) return 0;
    const size_t npix = (size_t)b->pending[view].w * b->pending[view].h;
    const int base = b->viewBase[view];
    if (rgba) {
        if (bytes < 0 || (size_t)bytes < npix * 4) { b->error = "output buffer too small"; return 0; }
        B_TRY(hipMemcpy(rgba, b->dPixels + base, npix * 4, hipMemcpyDeviceToHost));
    }
    if (accumulation) {
        if (floatCount < (long long)(npix * 10)) { b->error = "accumulation buffer too small"; return 0; }
        std::vector<float4> r(npix), a(npix), n(npix);
        B_TRY(hipMemcpy(r.data(), b->dSumRad + base, npix * 16, hipMemcpyDeviceToHost));
        B_TRY(hipMemcpy(a.data(), b->dSumAlb + base, npix * 16, hipMemcpyDeviceToHost));
        B_TRY(hipMemcpy(n.data(), b->dSumNrm + base, npix * 16, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < npix; i++) {
            float* o = accumulation + i * 10;
            o[0] = r[i].x; o[1] = r[i].y; o[2] = r[i].z; o[3] = a[i].x; o[4] = a[i].y; o[5] = a[i].z;
            o[6] = n[i].x; o[7] = n[i].y; o[8] = n[i].z; o[9] = r[i].w;
        }
    }
    return 1;
}

}  // extern "C"
