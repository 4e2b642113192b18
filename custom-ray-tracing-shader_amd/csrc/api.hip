// api.hip - the plugin's C ABI: session state, validation, the GPU frame driver and the
// CPU backends.  Entry points keep the reference plugin's names, argument meaning and
// 1 = success / 0 = failure contract (Native/renderer.cu:309-458).
#include <algorithm>
#include <array>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#include <mutex>
#include <dlfcn.h>
#include <omp.h>
#include "kernels.h"
#include "traverse.h"
#include "bvh_gpu.h"
#include "denoiser.h"
#include "embree_cpu.h"
#include "tuning.h"

using namespace rt;

// Device scratch of RtTraceRaysNoAlloc, one copy per GPU for the life of the process: the
// traversal's work cursor (int 0) and overflow flag (int 1) in the first 256 bytes, then the
// stack spill ring.  Host code never uses the array's name as an address: the per-device
// address comes from hipGetSymbolAddress (scratchAddress below).
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
constexpr size_t kNoAllocHeaderBytes = 256;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex gScratchMutex;
void* gScratch[kMaxDevices] = {};

// The scratch array's address on `device`, which must be the current device; resolved once per
// device.
hipError_t scratchAddress(int device, void** out) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(gScratchMutex);
    if (!gScratch[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        gScratch[device] = p;
    }
    *out = gScratch[device];
    return hipSuccess;
}

struct HostSum { f3 rad, alb, nrm; float covered, depth, lum2; };

struct Session {
    std::string hostError;
    hipError_t error = hipSuccess;
    // scene (colours linearised), light CDF, BVH
    std::vector<Triangle> tris;
    EmbreeScene embree;          // backend 3, built on its first render
    std::vector<Light> lights;
    float lightArea = 0;
    std::vector<RtBvhNode> binNodes;   // host copies, fetched on demand
    std::vector<int> binOrder;
    int binNodeCount = 0, wideNodeCount = 0;
    int bvhDepth = 0;
    int wideDepth = 0;           // deepest wide node (root 0): bounds a traversal stack by (W-1)(wideDepth+1)
    float buildMs = 0;
    int width = 0, height = 0;
    int bvhWidth = 4;            // option, used at the next RtInitialize
    int buildBins = 16;          // option: SAH bins per axis of the GPU build
    int buildMaxLeaf = 4;        // option: leaf size limit of the GPU build
    int builtWidth = 4;
    int bvhQuantized = 0;        // option: quantized (QNode) wide nodes at the next RtInitialize
    int builtQuantized = 0;
    int blocksPerCU = 3;         // persistent traversal workgroups per CU (per pipeline)
    int shortStack = 12;         // LDS stack entries per traversal lane (8, 12 or 16)
    int refill = 1;              // per-lane ray refill in the traversal kernel
    int refillBelow = 40;        // busy lanes below which a refilling wave fetches rays
    int whileWhile = 1;          // speculative while-while traversal loop
    int tiledPrimary = 1;        // primary rays queued in 8x8 pixel tiles
    int pipelines = 4;           // independent path pipelines per chunk, each on its own stream
    int tailBounce = 0;          // >= 1: paths complete from this bounce on in one launch (0: off)
    int nontemporal = 1;         // rays / hits bypass cache retention in traversal
    int xcdSegments = 1;         // per-XCD segment cursors in the frame's traversal launches
    int watertight = 0;          // 1: watertight ray/triangle test (Woop et al. 2013) instead of the reference's
    unsigned userSet = 0;        // tuned options the caller set (bit k: kTunedFields[k])
    int autotuneCache = 1;       // 1: apply the autotuner's cached launch options the caller did not set
    int tunedBounces = -1;       // bounce count the cached launch options were last looked up for
    std::vector<unsigned char> hostNodes;
    std::vector<TriRecord> hostRecs;
    std::vector<HostSum> hostSum;
    // device
    Triangle* dTris = nullptr;
    Light* dLights = nullptr;
    void* dNodes = nullptr;          // WideNode<builtWidth>[] or QNode<builtWidth>[]
    TriRecord* dRecs = nullptr;
    RtBvhNode* dBinNodes = nullptr;  // binary tree of the GPU build
    int* dOrder = nullptr;
    float4 *dSumRad = nullptr, *dSumAlb = nullptr, *dSumNrm = nullptr;
    uchar4* dPixels = nullptr;
    float4 *dThr = nullptr, *dRad = nullptr, *dAlb = nullptr, *dNrm = nullptr, *dNee = nullptr;
    float4 *dRayA = nullptr, *dRayB = nullptr, *dShadow = nullptr;
    int4* dHits = nullptr;
    int* dCounters = nullptr;     // kMaxRaysPerPixel chunks x kPipelines x kCounterStride
    int* dXcdWork = nullptr;      // kPipelines x kXcdSlotInts segment cursors
    int* dSpill = nullptr;
    int* dOverflow = nullptr;
    int* hCounters = nullptr;        // pinned host copy of the counters, then the overflow flag
    unsigned long long* dStats = nullptr;
    float4* dQueryRays = nullptr;    // RtTraceRays buffers, grown on demand and kept
    int4* dQueryHits = nullptr;
    int* dQueryWork = nullptr;
    size_t queryCap = 0;
    long long slotCap = 0;
    int spillEntries = 0, traceBlocks = 0, traceLanes = 0, queryBlocks = 0, cus = 0;
    size_t spillLanes = 0;           // lanes each pipeline's spill ring is sized for
    hipStream_t stream = nullptr;
    hipStream_t pipeStream[3] = {};  // streams of path pipelines 1..3 (default priority, as stream)
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipEvent_t evFork = nullptr, evJoin[3] = {};
    // progressive state (Native/renderer.cu:65-68)
    int device = 0;              // GPU current at RtInitialize
    int cpuThreads = 0, lastBackend = -1;
    unsigned samples = 0, version = 0;
    CameraData lastCamera{};
    RenderSettings lastSettings{};
    bool ready = false;
    unsigned long long lastClosest = 0, lastShadow = 0;
    // measurement hooks
    bool stageTiming = false, traceStats = false;
    std::vector<hipEvent_t> eventPool;
    std::vector<std::pair<int, int>> timedLaunches;   // (stage, first event index)
    float stageMs[RT_STAGE_COUNT] = {};
    int stageLaunches[RT_STAGE_COUNT] = {};
    std::vector<std::array<float, 3>> launchSpans;   // last frame: (stage, start ms, end ms) from frame start
    unsigned long long traceCounts[4] = {};
    bool hasFrame = false;
    // denoiser
    DenoiserNet* denoiser = nullptr;   // borrowed (RtSetDenoiser) or owned
    bool ownDenoiser = false;
    int denoiseFallback = 0;           // 0 per-tile check, 1 always plain accumulation, 2 never
    float* dDenoiseIn = nullptr;       // [hp][wp][10]
    float* dDenoiseOut = nullptr;      // [hp][wp][4]
    float* dDenoiseAuxIn = nullptr;    // [hp][wp][10] mirrored input (self-ensemble)
    float* dDenoiseAuxOut = nullptr;   // [hp][wp][4] its network output
    int denoiseEnsemble = 1;           // network passes per RtDenoise: 1, 2 (x mirror) or 4 (x, y, xy)
    int* dTileStats = nullptr;         // [kTileStatSlots] rejected-tile counters
    int lastTiles = 0, lastRejected = 0;
    float lastInferMs = 0;
};

constexpr int kPipelines = 4;   // path pipelines per chunk at most

// Brackets one launch with pooled events when stage timing is on.
struct StageScope {
    Session* s;
    int stage, ev;
    hipStream_t st;
    StageScope(Session* s_, int stage_, hipStream_t st_) : s(s_), stage(stage_), ev(-1), st(st_) {
        if (!s->stageTiming) return;
        const int used = (int)s->timedLaunches.size() * 2;
        if (used + 2 > (int)s->eventPool.size()) return;
        ev = used;
        (void)hipEventRecord(s->eventPool[ev], st);
    }
    ~StageScope() {
        if (ev < 0) return;
        (void)hipEventRecord(s->eventPool[ev + 1], st);
        s->timedLaunches.emplace_back(stage, ev);
    }
};

template <class T> bool devAlloc(Session* s, T*& p, size_t count) {
    s->error = hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T) > 0 ? count * sizeof(T) : 16);
    if (s->error != hipSuccess) { p = nullptr; return false; }
    return true;
}
template <class T> bool upload(Session* s, T*& p, const T* host, size_t count) {
    if (!devAlloc(s, p, count)) return false;
    if (count == 0) return true;
    s->error = hipMemcpy(p, host, count * sizeof(T), hipMemcpyHostToDevice);
    return s->error == hipSuccess;
}
template <class T> void devFree(T*& p) { if (p) (void)hipFree(p); p = nullptr; }

void freeDevice(Session* s) {
    devFree(s->dTris); devFree(s->dLights); devFree(s->dNodes); devFree(s->dRecs);
    devFree(s->dBinNodes); devFree(s->dOrder);
    devFree(s->dSumRad); devFree(s->dSumAlb); devFree(s->dSumNrm); devFree(s->dPixels);
    devFree(s->dThr); devFree(s->dRad); devFree(s->dAlb); devFree(s->dNrm); devFree(s->dNee);
    devFree(s->dRayA); devFree(s->dRayB); devFree(s->dShadow); devFree(s->dHits);
    devFree(s->dCounters); devFree(s->dXcdWork); devFree(s->dSpill); devFree(s->dOverflow); devFree(s->dStats);
    if (s->hCounters) (void)hipHostFree(s->hCounters);
    s->hCounters = nullptr;
    devFree(s->dDenoiseIn); devFree(s->dDenoiseOut); devFree(s->dTileStats);
    devFree(s->dDenoiseAuxIn); devFree(s->dDenoiseAuxOut);
    devFree(s->dQueryRays); devFree(s->dQueryHits); devFree(s->dQueryWork);
    s->queryCap = 0;
    for (hipEvent_t e : s->eventPool)
        if (e) (void)hipEventDestroy(e);
    s->eventPool.clear();
    if (s->ev0) (void)hipEventDestroy(s->ev0);
    if (s->ev1) (void)hipEventDestroy(s->ev1);
    if (s->evFork) (void)hipEventDestroy(s->evFork);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    for (int k = 0; k < kPipelines - 1; k++) {
        if (s->evJoin[k]) (void)hipEventDestroy(s->evJoin[k]);
        if (s->pipeStream[k]) (void)hipStreamDestroy(s->pipeStream[k]);
        s->evJoin[k] = nullptr; s->pipeStream[k] = nullptr;
    }
    s->ev0 = s->ev1 = s->evFork = nullptr; s->stream = nullptr;
}

SceneView deviceScene(const Session* s) {
    SceneView v; v.tris = s->dTris; v.lights = s->dLights; v.lightCount = (int)s->lights.size(); v.lightArea = s->lightArea;
    return v;
}

// GPU binned-SAH build (32 bins, leaf <= 4) and collapse to the session's wide layout.
bool buildAccel(Session* s) {
    auto t0 = std::chrono::steady_clock::now();
    GpuBuildParams prm;
    prm.bins = s->buildBins; prm.maxLeaf = s->buildMaxLeaf; prm.width = s->bvhWidth; prm.quantized = s->bvhQuantized;
    GpuBuildResult r;
    if ((s->error = buildBvhGpu(s->dTris, (int)s->tris.size(), prm, r, s->stream)) != hipSuccess) {
        s->hostError = std::string("GPU BVH build failed: ") + hipGetErrorString(s->error);
        return false;
    }
    s->dBinNodes = r.binNodes; s->dOrder = r.order; s->dRecs = r.recs; s->dNodes = r.wideNodes;
    s->binNodeCount = r.binNodeCount; s->wideNodeCount = r.wideCount; s->bvhDepth = r.depth; s->wideDepth = r.wideDepth;
    if (s->bvhWidth == 4 && !s->bvhQuantized && r.wideCount >= kMaxWide4Nodes) {
        s->hostError = "scene needs more than 2^25 4-wide nodes (4 GB); use bvh_width 8 or bvh_quantized";
        return false;
    }
    s->builtWidth = s->bvhWidth;
    s->builtQuantized = s->bvhQuantized;
    s->buildMs = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return true;
}

// Host copies of the wide nodes and records for the CPU backends (same tree as the GPU).
bool fetchHostAccel(Session* s) {
    if (!s->hostRecs.empty()) return true;
    const size_t nodeBytes = (size_t)s->wideNodeCount * wideNodeBytes(s->builtWidth, s->builtQuantized);
    s->hostNodes.resize(nodeBytes);
    s->hostRecs.resize(s->tris.size());
    if ((s->error = hipMemcpy(s->hostNodes.data(), s->dNodes, nodeBytes, hipMemcpyDeviceToHost)) != hipSuccess ||
        (s->error = hipMemcpy(s->hostRecs.data(), s->dRecs, s->hostRecs.size() * sizeof(TriRecord),
                              hipMemcpyDeviceToHost)) != hipSuccess) {
        s->hostRecs.clear();
        return false;
    }
    return true;
}

bool fetchBinaryTree(Session* s) {
    if (!s->binNodes.empty()) return true;
    s->binNodes.resize(s->binNodeCount);
    s->binOrder.resize(s->tris.size());
    if ((s->error = hipMemcpy(s->binNodes.data(), s->dBinNodes, s->binNodes.size() * sizeof(RtBvhNode),
                              hipMemcpyDeviceToHost)) != hipSuccess ||
        (s->error = hipMemcpy(s->binOrder.data(), s->dOrder, s->binOrder.size() * sizeof(int), hipMemcpyDeviceToHost)) !=
            hipSuccess) {
        s->binNodes.clear();
        return false;
    }
    return true;
}

TraceLaunch traceLaunch(const Session* s) {
    TraceLaunch l;
    l.width = s->builtWidth; l.quantized = s->builtQuantized != 0; l.stack = s->shortStack;
    l.refill = s->refill != 0; l.whileWhile = s->whileWhile != 0; l.refillBelow = s->refillBelow;
    return l;
}

// Sizes the persistent traversal grid from the current options; grows the spill ring.
// RtTraceRays runs one launch with nothing beside it, so it takes at least this many
// persistent workgroups per CU whatever blocks_per_cu (sized for concurrent pipelines) says.
constexpr int kQueryBlocksPerCU = 6;

bool prepareTraceGrid(Session* s) {
    s->traceBlocks = s->cus * s->blocksPerCU;
    s->traceLanes = s->traceBlocks * kBlock;
    s->queryBlocks = s->cus * std::max(s->blocksPerCU, kQueryBlocksPerCU);
    s->spillEntries = 36 * (s->builtWidth - 1);   // >= (W-1) * (wide depth <= 33) entries
    const size_t lanes = (size_t)s->queryBlocks * kBlock;   // >= traceLanes
    if (lanes > s->spillLanes) {
        devFree(s->dSpill);
        s->spillLanes = 0;
        if (!devAlloc(s, s->dSpill, (size_t)kPipelines * s->spillEntries * lanes * 2)) return false;
        s->spillLanes = lanes;
    }
    return true;
}

// Traversal arguments of path pipeline p (each concurrent pipeline has its own spill ring).
TraceArgs baseTraceArgs(Session* s, int p = 0) {
    TraceArgs a{};
    a.nodes = s->dNodes; a.recs = s->dRecs;
    a.spill = s->dSpill + (size_t)p * s->spillEntries * s->spillLanes * 2;
    a.spillEntries = s->spillEntries; a.totalLanes = s->traceLanes;
    a.overflow = s->dOverflow;
    a.nontemporal = s->nontemporal;
    a.watertight = s->watertight;
    return a;
}

// Segment cursors of pipeline p's work slot k (the launch that uses work cursor
// kWorkRays + k), or null when the option is off.
static int* xcdWork(Session* s, int p, int k) {
    return s->xcdSegments ? s->dXcdWork + (size_t)p * kXcdSlotInts + (size_t)k * kXcdCursorInts : nullptr;
}

// Shadow rays of bounce b-1 wait to be traced in launch b.
static bool shadowAt(bool nee, int b) { return nee && b > 0; }

// One GPU RtRender: chunks of R sample groups; see kernels.hip for the stage order.
bool renderGpu(Session* s, const CameraData& cam, const RenderSettings& rs, float* ms) {
    const int npix = s->width * s->height;
    const int rpp = rs.raysPerPixel;
    int chunkR = (int)std::min<long long>(rpp, s->slotCap / npix);
    if (chunkR < 1) chunkR = 1;
    hipStream_t st = s->stream;
    if (!prepareTraceGrid(s)) return false;
    s->timedLaunches.clear();
    if (s->stageTiming && s->eventPool.empty()) {
        // Timing-only events: no system-scope fence (cache writeback and invalidation) at
        // each record, which would otherwise slow the frame being measured.  They are read
        // after the frame's closing event has completed.
        s->eventPool.resize(4096);
        for (auto& e : s->eventPool)
            if ((s->error = hipEventCreateWithFlags(&e, hipEventDisableSystemFence)) != hipSuccess) return false;
    }
    if (s->traceStats && (s->error = hipMemsetAsync(s->dStats, 0, 4 * sizeof(unsigned long long), st)) != hipSuccess)
        return false;
    // Each frame starts with a clear overflow flag, so one failed frame does not fail the next.
    if ((s->error = launchZeroInts(s->dOverflow, 1, st)) != hipSuccess) return false;
    if ((s->error = hipEventRecord(s->ev0, st)) != hipSuccess) return false;
    // Path pipelines: a chunk's queue positions are split in P parts, each traced and shaded
    // on its own stream with its own counters, queues and spill ring, so one part's kernels
    // fill the GPU while another's launch drains its last rays.  Frames with trace
    // statistics (instrumented kernels) run one pipeline.
    const bool nee = !s->lights.empty();
    const bool measure = s->traceStats;
    std::vector<int> chunkPipelines;
    int chunk = 0;
    for (int r0 = 0; r0 < rpp; r0 += chunkR, chunk++) {
        const int R = std::min(chunkR, rpp - r0);
        const int Q = R * npix;
        const int P = measure ? 1 : std::max(1, std::min(s->pipelines, Q >> 16));   // >= 64K paths each
        chunkPipelines.push_back(P);
        hipStream_t pst[kPipelines] = {st, s->pipeStream[0], s->pipeStream[1], s->pipeStream[2]};
        if (P > 1) {
            if ((s->error = hipEventRecord(s->evFork, st)) != hipSuccess) return false;
            for (int p = 1; p < P; p++)
                if ((s->error = hipStreamWaitEvent(pst[p], s->evFork, 0)) != hipSuccess) return false;
        }
        int* ctr[kPipelines];
        FrameArgs f[kPipelines];
        const int part = (Q / P + 63) / 64 * 64;   // queue positions per pipeline, the last takes the rest
        for (int p = 0; p < P; p++) {
            const int qBase = p * part, qCount = p + 1 < P ? part : Q - qBase;
            ctr[p] = s->dCounters + ((size_t)chunk * kPipelines + p) * kCounterStride;
            FrameArgs& fp = f[p];
            fp = FrameArgs{};
            fp.cam = cam; fp.rs = rs; fp.sc = deviceScene(s);
            fp.w = s->width; fp.h = s->height; fp.npix = npix; fp.R = R; fp.sample0 = s->samples + (unsigned)r0;
            fp.thr = s->dThr; fp.rad = s->dRad; fp.alb = s->dAlb; fp.nrm = s->dNrm; fp.nee = s->dNee;
            // Queue buffers: pipeline p owns positions [qBase, qBase + qCount) of each.
            fp.shadow = s->dShadow + 2 * (size_t)qBase; fp.hits = s->dHits + qBase;
            fp.counters = ctr[p]; fp.tiled = s->tiledPrimary;
            fp.qBase = qBase; fp.qCount = qCount;
            {
                StageScope sc(s, RT_STAGE_RESET, pst[p]);
                if ((s->error = launchResetCounters(ctr[p], qCount, xcdWork(s, p, 0), pst[p])) != hipSuccess) return false;
            }
            fp.rayOut = s->dRayA + 2 * (size_t)qBase;
            StageScope sc(s, RT_STAGE_RAYGEN, pst[p]);
            if ((s->error = launchRaygen(fp, pst[p])) != hipSuccess) return false;
        }
        // Bounce b's closest-hit rays and bounce b-1's shadow rays are independent, so one
        // launch traces both queues; the shadow rays' radiance is still added before shade
        // b adds bounce b's emission, the reference's order (Native/renderer.cu:232-267).
        // With tail_bounce = k, bounce k-1's shadow rays get a launch of their own and the
        // path-completion kernel takes every path from bounce k to its end.
        // Path completion is built for the reference's intersection test: with the watertight
        // test every bounce gets its own launches (images do not depend on the split).
        const int tail = (s->tailBounce >= 1 && s->tailBounce < rs.maxBounces && !s->watertight) ? s->tailBounce : -1;
        for (int b = 0; b <= rs.maxBounces; b++) {
            if (b == tail) {
                for (int p = 0; p < P; p++) {
                    const size_t off = 2 * (size_t)f[p].qBase;
                    if (shadowAt(nee, b)) {
                        TraceArgs sa = baseTraceArgs(s, p);
                        sa.work = ctr[p] + kWorkRays + rs.maxBounces;   // unused by the per-bounce launches
                        sa.xcdWork = xcdWork(s, p, rs.maxBounces);
                        sa.stats = s->dStats;
                        sa.anyRays = f[p].shadow; sa.anyCountPtr = ctr[p] + kCntShadow + (b - 1);
                        sa.rad = s->dRad; sa.contrib = s->dNee;
                        StageScope sc(s, RT_STAGE_SHADOW, pst[p]);
                        if ((s->error = launchTrace(traceLaunch(s), s->traceStats, s->traceBlocks, sa, pst[p])) != hipSuccess)
                            return false;
                    }
                    TraceArgs ta = baseTraceArgs(s, p);
                    ta.rays = ((b & 1) ? s->dRayB : s->dRayA) + off;
                    ta.countPtr = ctr[p] + kCntRays + b;
                    ta.work = ctr[p] + kWorkRays + b;
                    ta.xcdWork = xcdWork(s, p, b);
                    ta.stats = s->dStats;
                    StageScope sc(s, RT_STAGE_TAIL, pst[p]);
                    if ((s->error = launchPathTail(traceLaunch(s), s->traceStats, s->traceBlocks, s->cus, f[p], ta, b, pst[p])) !=
                        hipSuccess)
                        return false;
                }
                break;
            }
            const bool rays = b < rs.maxBounces, shadow = shadowAt(nee, b);
            if (!rays && !shadow) break;
            for (int p = 0; p < P; p++) {
                const size_t off = 2 * (size_t)f[p].qBase;
                float4* in = ((b & 1) ? s->dRayB : s->dRayA) + off;
                float4* out = ((b & 1) ? s->dRayA : s->dRayB) + off;
                TraceArgs ta = baseTraceArgs(s, p);
                ta.work = ctr[p] + kWorkRays + b;
                ta.xcdWork = xcdWork(s, p, b);
                ta.stats = s->dStats;
                if (rays) { ta.rays = in; ta.countPtr = ctr[p] + kCntRays + b; ta.hits = f[p].hits; }
                if (shadow) {
                    ta.anyRays = f[p].shadow; ta.anyCountPtr = ctr[p] + kCntShadow + (b - 1);
                    ta.rad = s->dRad; ta.contrib = s->dNee;
                }
                {
                    StageScope sc(s, rays ? RT_STAGE_CLOSEST : RT_STAGE_SHADOW, pst[p]);
                    if ((s->error = launchTrace(traceLaunch(s), s->traceStats, s->traceBlocks, ta, pst[p])) != hipSuccess)
                        return false;
                }
                if (!rays) continue;
                f[p].bounce = b; f[p].rayIn = in; f[p].rayOut = out;
                StageScope sc(s, RT_STAGE_SHADE, pst[p]);
                if ((s->error = launchShade(f[p], s->cus * shadeBlocksPerCU(), pst[p])) != hipSuccess) return false;
            }
        }
        for (int p = 1; p < P; p++) {
            if ((s->error = hipEventRecord(s->evJoin[p - 1], pst[p])) != hipSuccess) return false;
            if ((s->error = hipStreamWaitEvent(st, s->evJoin[p - 1], 0)) != hipSuccess) return false;
        }
        AccumArgs a{};
        a.thr = s->dThr; a.rad = s->dRad; a.alb = s->dAlb; a.nrm = s->dNrm;
        a.sumRad = s->dSumRad; a.sumAlb = s->dSumAlb; a.sumNrm = s->dSumNrm; a.pixels = s->dPixels;
        a.npix = npix; a.R = R;
        a.startFromZero = (s->samples == 0 && r0 == 0) ? 1 : 0;
        a.writeDisplay = (r0 + R >= rpp) ? 1 : 0;
        a.invSamples = 1.0f / (float)(s->samples + (unsigned)rpp);
        a.exposure = rs.exposure;
        StageScope sc(s, RT_STAGE_ACCUMULATE, st);
        if ((s->error = launchAccumulate(a, st)) != hipSuccess) return false;
    }
    if ((s->error = hipEventRecord(s->ev1, st)) != hipSuccess) return false;
    // Counters and the overflow flag come back with the frame, in one wait.
    const size_t nCtr = (size_t)chunk * kPipelines * kCounterStride;
    const int* ctrs = s->hCounters;
    if ((s->error = hipMemcpyAsync(s->hCounters, s->dCounters, nCtr * sizeof(int), hipMemcpyDeviceToHost, st)) !=
            hipSuccess ||
        (s->error = hipMemcpyAsync(s->hCounters + nCtr, s->dOverflow, sizeof(int), hipMemcpyDeviceToHost, st)) !=
            hipSuccess ||
        (s->error = hipStreamSynchronize(st)) != hipSuccess)
        return false;
    if ((s->error = hipEventElapsedTime(ms, s->ev0, s->ev1)) != hipSuccess) return false;
    if (s->hCounters[nCtr]) {
        // The frame's accumulate kernel has already added its samples: restart the
        // accumulation so the next frame does not build on an incomplete one.
        RtReset(s);
        s->hostError = "traversal stack exceeded its spill ring";
        return false;
    }
    unsigned long long closest = 0, shadow = 0;
    for (int c = 0; c < chunk; c++)
        for (int p = 0; p < chunkPipelines[c]; p++)
            for (int b = 0; b < rs.maxBounces; b++) {
                const size_t base = ((size_t)c * kPipelines + p) * kCounterStride;
                closest += (unsigned)ctrs[base + kCntRays + b];
                shadow += (unsigned)ctrs[base + kCntShadow + b];
            }
    s->lastClosest = closest; s->lastShadow = shadow;
    for (int k = 0; k < RT_STAGE_COUNT; k++) { s->stageMs[k] = 0; s->stageLaunches[k] = 0; }
    s->launchSpans.clear();
    for (const auto& tl : s->timedLaunches) {
        float t0 = 0, t1 = 0;
        if ((s->error = hipEventElapsedTime(&t0, s->ev0, s->eventPool[tl.second])) != hipSuccess ||
            (s->error = hipEventElapsedTime(&t1, s->ev0, s->eventPool[tl.second + 1])) != hipSuccess)
            return false;
        s->stageMs[tl.first] += t1 - t0;
        s->stageLaunches[tl.first]++;
        s->launchSpans.push_back({(float)tl.first, t0, t1});
    }
    if (s->traceStats &&
        (s->error = hipMemcpy(s->traceCounts, s->dStats, sizeof(s->traceCounts), hipMemcpyDeviceToHost)) != hipSuccess)
        return false;
    s->hasFrame = true;
    return true;
}

int setOption(Session* s, const char* key, int value);

// Options of o that the caller did not set (with build = false, only the launch options;
// the BVH layout is fixed at RtInitialize).
void applyTuned(Session* s, const LaunchOptions& o, bool build) {
    for (int k = build ? 0 : 2; k < 8; k++) {
        const int v = o.*(kTunedFields[k].field);
        if (v >= 0 && !(s->userSet & (1u << k))) (void)setOption(s, kTunedFields[k].name, v);
    }
}

// Before a frame: the cached launch options of this scene class and bounce count, once per
// bounce count.
void applyTunedLaunch(Session* s, int maxBounces) {
    if (!s->autotuneCache || s->tunedBounces == maxBounces) return;
    s->tunedBounces = maxBounces;
    LaunchOptions o;
    if (tuneLookup(sessionTuneKey((long long)s->tris.size(), s->width, s->height, maxBounces), false, o))
        applyTuned(s, o, false);
}

// Rebuilds the session's BVH with another width or leaf size (the autotuner's layout trials).
bool rebuildAccel(Session* s, int width, int maxLeaf) {
    devFree(s->dNodes); devFree(s->dRecs); devFree(s->dBinNodes); devFree(s->dOrder); devFree(s->dSpill);
    s->spillLanes = 0;
    s->hostNodes.clear(); s->hostRecs.clear(); s->binNodes.clear(); s->binOrder.clear();
    s->bvhWidth = width; s->buildMaxLeaf = maxLeaf;
    return buildAccel(s) && prepareTraceGrid(s);
}

bool validRender(Session* s, const CameraData* camera, const RenderSettings* settings) {
    return camera && settings && s->ready && settings->raysPerPixel >= 1 && settings->raysPerPixel <= kMaxRaysPerPixel &&
           settings->maxBounces >= 1 && settings->maxBounces <= kMaxBounces && settings->exposure > 0 &&
           settings->aperture >= 0 && settings->focusDistance > 0 && settings->skyStrength >= 0;
}

// Native/renderer.cu:289-302 on the host, over the same wide BVH as the GPU.
// trace(o, d, limit, anyHit) answers the queries: the session's BVH (backends 0, 1) or
// Embree (backend 3).
template <class Tracer>
void renderCpuRows(Session* s, const CameraData& cam, const RenderSettings& rs, unsigned char* out, bool multi,
                   const Tracer& trace, unsigned long long& closest, unsigned long long& shadow) {
    SceneView sc; sc.tris = s->tris.data(); sc.lights = s->lights.data(); sc.lightCount = (int)s->lights.size(); sc.lightArea = s->lightArea;
    const int w = s->width, h = s->height;
    const unsigned* dirs = hostSobolDirections();
    const unsigned sample = s->samples;
    unsigned long long cc = 0, ss = 0;
    #pragma omp parallel if (multi) reduction(+ : cc, ss)
    {
        #pragma omp single
        s->cpuThreads = omp_get_num_threads();
        #pragma omp for schedule(static)
        for (int y = 0; y < h; y++)
            for (int x = 0; x < w; x++) {
                const size_t i = (size_t)y * w + x;
                HostSum tot{};
                if (sample) tot = s->hostSum[i];
                for (int r = 0; r < rs.raysPerPixel; r++) {
                    const unsigned smp = sample + (unsigned)r, pix = (unsigned)(y * w + x);
                    unsigned base = 0;
                    auto rng = [&](int k) { return sobolSample(dirs, pix, smp, rs.seed, base + (unsigned)k); };
                    f3 o, d;
                    cameraRay(cam, rs, x, y, w, h, rng, o, d);
                    PathState ps; ps.thr = F3(1, 1, 1); ps.prevPdf = 0; ps.rad = F3(0, 0, 0);
                    FirstHit fh; fh.covered = false; fh.albedo = fh.normal = F3(0, 0, 0); fh.depth = 0;
                    for (int b = 0; b < rs.maxBounces; b++) {
                        base = 4 + 8 * (unsigned)b;
                        RtHit hit = trace(o, d, 1e30f, false);
                        cc++;
                        ShadeResult res;
                        shadeBounce(sc, rs, b, rng, o, d, hit, ps, &fh, res);
                        if (res.shadow) {
                            ss++;
                            if (trace(res.nextO, res.sdir, res.slimit, true).index < 0)
                                ps.rad = ps.rad + res.contrib;
                        }
                        if (!res.cont) break;
                        o = res.nextO; d = res.nextD;
                    }
                    tot.rad = tot.rad + ps.rad;
                    tot.alb = tot.alb + fh.albedo;
                    tot.nrm = tot.nrm + fh.normal;
                    tot.covered = tot.covered + (fh.covered ? 1.0f : 0.0f);
                    tot.depth = tot.depth + fh.depth;
                    const float lum = 0.2126f * ps.rad.x + 0.7152f * ps.rad.y + 0.0722f * ps.rad.z;
                    tot.lum2 = tot.lum2 + lum * lum;
                }
                s->hostSum[i] = tot;
                const float inv = 1.0f / (float)(sample + (unsigned)rs.raysPerPixel);
                unsigned char* px = out + i * 4;
                px[0] = displayByte(tot.rad.x * inv, rs.exposure);
                px[1] = displayByte(tot.rad.y * inv, rs.exposure);
                px[2] = displayByte(tot.rad.z * inv, rs.exposure);
                px[3] = 255;
            }
    }
    closest = cc; shadow = ss;
}

template <class N>
void renderBvhRows(Session* s, const CameraData& cam, const RenderSettings& rs, unsigned char* out, bool multi,
                   unsigned long long& closest, unsigned long long& shadow) {
    const N* nodes = reinterpret_cast<const N*>(s->hostNodes.data());
    const TriRecord* recs = s->hostRecs.data();
    const bool wt = s->watertight != 0;
    auto trace = [nodes, recs, wt](f3 o, f3 d, float limit, bool any) {
        return any ? traverseHost<N, true>(nodes, recs, o, d, limit, wt) : traverseHost<N, false>(nodes, recs, o, d, limit, wt);
    };
    renderCpuRows(s, cam, rs, out, multi, trace, closest, shadow);
}

}  // namespace

// ======================================================================== C ABI
extern "C" {

void* RtCreate(void) { return new Session(); }

void RtReset(void* p) {
    Session* s = static_cast<Session*>(p);
    if (s) { s->samples = 0; s->version++; }
}
unsigned RtAccumulationVersion(void* p) { Session* s = static_cast<Session*>(p); return s ? s->version : 0; }
unsigned RtSampleCount(void* p) { Session* s = static_cast<Session*>(p); return s ? s->samples : 0; }
float RtEmbreeBuildMilliseconds(void* p) { Session* s = static_cast<Session*>(p); return s ? s->embree.buildMs : 0; }
int RtCpuThreadCount(void* p) { Session* s = static_cast<Session*>(p); return s ? s->cpuThreads : 0; }

void RtBvhInfo(void* p, int* nodes, int* depth, float* buildMs) {
    Session* s = static_cast<Session*>(p);
    *nodes = s ? s->binNodeCount : 0;
    *depth = s ? s->bvhDepth : 0;
    *buildMs = s ? s->buildMs : 0;
}

const char* RtLastError(void* p) {
    Session* s = static_cast<Session*>(p);
    if (s && !s->hostError.empty()) return s->hostError.c_str();
    return hipGetErrorString(s ? s->error : hipErrorInvalidValue);
}

void RtShutdown(void* p) {
    Session* s = static_cast<Session*>(p);
    if (!s) return;
    s->hostError.clear();
    s->embree.release();   // before the triangles its filter reads go away
    freeDevice(s);
    if (s->ownDenoiser) denoiserDestroy(s->denoiser);
    if (s->ownDenoiser) s->denoiser = nullptr;
    s->ownDenoiser = false;
    s->ready = false;
    s->tris.clear(); s->lights.clear(); s->binNodes.clear(); s->binOrder.clear();
    s->hostNodes.clear(); s->hostRecs.clear(); s->hostSum.clear();
    s->binNodeCount = s->wideNodeCount = 0;
    s->width = s->height = 0; s->samples = 0; s->lightArea = 0; s->buildMs = 0; s->bvhDepth = 0; s->lastBackend = -1;
    s->hasFrame = false;
}

void RtDestroy(void* p) {
    Session* s = static_cast<Session*>(p);
    if (s) { RtShutdown(s); delete s; }
}

int RtInitialize(void* p, const Triangle* triangles, int count, int w, int h) {
    Session* s = static_cast<Session*>(p);
    if (!s) return 0;
    RtShutdown(s);
    if (!triangles || count <= 0 || w <= 0 || h <= 0 || w > 4096 || h > 4096) { s->error = hipErrorInvalidValue; return 0; }
    int device = 0, cus = 0;
    if ((s->error = hipGetDevice(&device)) != hipSuccess) return 0;
    if ((s->error = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device)) != hipSuccess) return 0;
    void* scratchPtr = nullptr;   // set-up of RtTraceRaysNoAlloc's buffer on this device
    if ((s->error = scratchAddress(device, &scratchPtr)) != hipSuccess) {
        s->hostError = std::string("no-allocation scratch buffer: ") + hipGetErrorString(s->error);
        return 0;
    }
    s->device = device;
    s->tris.assign(triangles, triangles + count);
    s->width = w; s->height = h;
    s->tunedBounces = -1;
    if (s->autotuneCache) {   // the autotuner's choice for this scene class (any bounce count)
        LaunchOptions o;
        if (tuneLookup(sessionTunePrefix(count, w, h), true, o)) applyTuned(s, o, true);
    }
    s->hostSum.assign((size_t)w * h, HostSum{});
    // Native/renderer.cu:345-353: linear colours, emitter CDF by FP32 running area sum.
    for (int i = 0; i < count; i++) {
        Triangle& t = s->tris[i];
        t.color = RtVec3{srgbToLinear(t.color.x), srgbToLinear(t.color.y), srgbToLinear(t.color.z)};
        t.specularColor = RtVec3{srgbToLinear(t.specularColor.x), srgbToLinear(t.specularColor.y), srgbToLinear(t.specularColor.z)};
        if (t.emission > 0) {
            f3 an = crossv(V(t.edge1), V(t.edge2));
            s->lightArea += .5f * sqrtf(dotv(an, an));
            s->lights.push_back(Light{i, s->lightArea});
        }
    }
    if ((s->error = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking)) != hipSuccess) return 0;
    for (int k = 0; k < kPipelines - 1; k++) {
        if ((s->error = hipStreamCreateWithFlags(&s->pipeStream[k], hipStreamNonBlocking)) != hipSuccess) return 0;
        if ((s->error = hipEventCreateWithFlags(&s->evJoin[k], hipEventDisableTiming)) != hipSuccess) return 0;
    }
    if ((s->error = hipEventCreate(&s->ev0)) != hipSuccess) return 0;
    if ((s->error = hipEventCreate(&s->ev1)) != hipSuccess) return 0;
    if ((s->error = hipEventCreateWithFlags(&s->evFork, hipEventDisableTiming)) != hipSuccess) return 0;

    if (!upload(s, s->dTris, s->tris.data(), s->tris.size())) return 0;
    if (!buildAccel(s)) return 0;
    if (!upload(s, s->dLights, s->lights.data(), s->lights.size())) return 0;
    const size_t npix = (size_t)w * h;
    s->slotCap = std::max<long long>((long long)npix, std::min<long long>((long long)npix * kMaxRaysPerPixel, 1LL << 24));
    const size_t slots = (size_t)s->slotCap;
    if (!devAlloc(s, s->dSumRad, npix) || !devAlloc(s, s->dSumAlb, npix) || !devAlloc(s, s->dSumNrm, npix) ||
        !devAlloc(s, s->dPixels, npix) || !devAlloc(s, s->dThr, slots) || !devAlloc(s, s->dRad, slots) ||
        !devAlloc(s, s->dAlb, slots) || !devAlloc(s, s->dNrm, slots) || !devAlloc(s, s->dNee, slots) ||
        !devAlloc(s, s->dRayA, 2 * slots) || !devAlloc(s, s->dRayB, 2 * slots) || !devAlloc(s, s->dShadow, 2 * slots) ||
        !devAlloc(s, s->dHits, slots) ||
        !devAlloc(s, s->dCounters, (size_t)kMaxRaysPerPixel * kPipelines * kCounterStride) ||
        !devAlloc(s, s->dXcdWork, (size_t)kPipelines * kXcdSlotInts) ||
        !devAlloc(s, s->dOverflow, 1) || !devAlloc(s, s->dStats, 4))
        return 0;
    if ((s->error = hipHostMalloc(&s->hCounters, ((size_t)kMaxRaysPerPixel * kPipelines * kCounterStride + 1) * sizeof(int),
                                  hipHostMallocDefault)) != hipSuccess) {
        s->hCounters = nullptr;
        return 0;
    }
    s->cus = cus;
    s->spillLanes = 0;
    if (!prepareTraceGrid(s)) return 0;
    if ((s->error = hipMemset(s->dOverflow, 0, sizeof(int))) != hipSuccess) return 0;
    if ((s->error = hipMemset(s->dSumRad, 0, npix * sizeof(float4))) != hipSuccess) return 0;
    s->ready = true;
    return 1;
}

int RtRender(void* p, const CameraData* camera, const RenderSettings* settings, unsigned char* output, int outputBytes,
             int backend, float* computeMs) {
    Session* s = static_cast<Session*>(p);
    if (!s) return 0;
    if (!camera || !settings || !output || !computeMs || !s->ready || backend < 0 || backend > 3 || outputBytes < 0 ||
        (size_t)outputBytes < (size_t)s->width * s->height * 4 || settings->raysPerPixel < 1 ||
        settings->raysPerPixel > kMaxRaysPerPixel || settings->maxBounces < 1 || settings->maxBounces > kMaxBounces ||
        !(settings->exposure > 0) || !(settings->aperture >= 0) || !(settings->focusDistance > 0) ||
        !(settings->skyStrength >= 0)) {
        s->error = hipErrorInvalidValue;
        return 0;
    }
    s->hostError.clear();
    if (backend == RT_BACKEND_EMBREE && !s->embree.scene) {
        // Built on first use, as the reference's backend 3 (Native/renderer.cu:391-395); fails
        // with the loader's message when libembree3.so.3 is not available.
        if (!s->embree.build(s->tris, s->hostError)) { s->embree.release(); return 0; }
    }
    if (s->lastBackend != backend || std::memcmp(camera, &s->lastCamera, sizeof(CameraData)) ||
        std::memcmp(settings, &s->lastSettings, sizeof(RenderSettings)))
        RtReset(s);
    s->lastBackend = backend; s->lastCamera = *camera; s->lastSettings = *settings; s->cpuThreads = 0;
    if (backend == RT_BACKEND_GPU) {
        applyTunedLaunch(s, settings->maxBounces);
        if (!renderGpu(s, *camera, *settings, computeMs)) return 0;
        s->error = hipMemcpy(output, s->dPixels, (size_t)s->width * s->height * 4, hipMemcpyDeviceToHost);
        if (s->error != hipSuccess) return 0;
    } else if (backend == RT_BACKEND_EMBREE) {
        auto t0 = std::chrono::steady_clock::now();
        const EmbreeScene* em = &s->embree;
        auto trace = [em](f3 o, f3 d, float limit, bool any) { return em->query(o, d, limit, any); };
        unsigned long long c = 0, sh = 0;
        renderCpuRows(s, *camera, *settings, output, true, trace, c, sh);
        s->lastClosest = c; s->lastShadow = sh;
        *computeMs = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
    } else {
        if (!fetchHostAccel(s)) return 0;
        auto t0 = std::chrono::steady_clock::now();
        const bool multi = backend == RT_BACKEND_CPU_MULTI;
        unsigned long long c = 0, sh = 0;
        if (s->builtQuantized) {
            if (s->builtWidth == 2) renderBvhRows<QNode<2>>(s, *camera, *settings, output, multi, c, sh);
            else if (s->builtWidth == 8) renderBvhRows<QNode<8>>(s, *camera, *settings, output, multi, c, sh);
            else renderBvhRows<QNode<4>>(s, *camera, *settings, output, multi, c, sh);
        } else {
            if (s->builtWidth == 2) renderBvhRows<WideNode<2>>(s, *camera, *settings, output, multi, c, sh);
            else if (s->builtWidth == 8) renderBvhRows<WideNode<8>>(s, *camera, *settings, output, multi, c, sh);
            else renderBvhRows<WideNode<4>>(s, *camera, *settings, output, multi, c, sh);
        }
        s->lastClosest = c; s->lastShadow = sh;
        *computeMs = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    s->samples += (unsigned)settings->raysPerPixel;
    s->error = hipSuccess;
    return 1;
}

int RtRenderResident(void* p, const CameraData* camera, const RenderSettings* settings, float* computeMs) {
    Session* s = static_cast<Session*>(p);
    if (!s) return 0;
    if (!computeMs || !validRender(s, camera, settings)) { s->error = hipErrorInvalidValue; return 0; }
    s->hostError.clear();
    if (s->lastBackend != RT_BACKEND_GPU || std::memcmp(camera, &s->lastCamera, sizeof(CameraData)) ||
        std::memcmp(settings, &s->lastSettings, sizeof(RenderSettings)))
        RtReset(s);
    s->lastBackend = RT_BACKEND_GPU; s->lastCamera = *camera; s->lastSettings = *settings; s->cpuThreads = 0;
    applyTunedLaunch(s, settings->maxBounces);
    if (!renderGpu(s, *camera, *settings, computeMs)) return 0;
    s->samples += (unsigned)settings->raysPerPixel;
    s->error = hipSuccess;
    return 1;
}

int RtAutotune(void* p, const CameraData* camera, const RenderSettings* settings, int frames, float* mraysPerSecond,
               int* trialCount) {
    Session* s = static_cast<Session*>(p);
    if (!s) return 0;
    if (frames < 1 || !validRender(s, camera, settings)) { s->error = hipErrorInvalidValue; return 0; }
    s->hostError.clear();
    TuneSpace space;
    // Options the caller set stay fixed: the search covers the others.
    auto fixed = [&](int k, int v) { return (s->userSet & (1u << k)) ? std::vector<int>{v} : std::vector<int>(); };
    if (!fixed(0, s->builtWidth).empty()) space.widths = fixed(0, s->builtWidth);
    if (!fixed(1, s->buildMaxLeaf).empty()) space.leaves = fixed(1, s->buildMaxLeaf);
    if (!fixed(2, s->blocksPerCU).empty()) space.blocks = fixed(2, s->blocksPerCU);
    if (!fixed(3, s->shortStack).empty()) space.stacks = fixed(3, s->shortStack);
    if (s->userSet & (3u << 4)) space.loops = {{s->whileWhile, s->refill}};
    if (!fixed(6, s->pipelines).empty()) space.pipelines = fixed(6, s->pipelines);
    if (!fixed(7, s->tailBounce).empty()) space.tails = fixed(7, s->tailBounce);
    bool failed = false;
    auto apply = [&](const LaunchOptions& c) {
        if (c.bvhWidth != s->builtWidth || c.buildMaxLeaf != s->buildMaxLeaf || !s->dNodes)
            if (!rebuildAccel(s, c.bvhWidth, c.buildMaxLeaf)) return false;
        (void)setOption(s, "blocks_per_cu", c.blocksPerCU);
        (void)setOption(s, "short_stack", c.shortStack);
        (void)setOption(s, "while_while", c.whileWhile);
        (void)setOption(s, "refill", c.refill);
        (void)setOption(s, "pipelines", c.pipelines);
        (void)setOption(s, "tail_bounce", c.tailBounce);
        return true;
    };
    auto measure = [&](const LaunchOptions& c) -> double {
        if (failed) return -1;
        if (!apply(c)) { failed = true; return -1; }
        float ms = 0;
        RtReset(s);
        if (!renderGpu(s, *camera, *settings, &ms)) { failed = true; return -1; }   // warm: buffers, caches
        double best = 0;
        for (int rep = 0; rep < 2; rep++) {
            unsigned long long rays = 0;
            const auto t0 = std::chrono::steady_clock::now();
            for (int f = 0; f < frames; f++) {
                if (!renderGpu(s, *camera, *settings, &ms)) { failed = true; return -1; }
                rays += s->lastClosest + s->lastShadow;
            }
            best = std::max(best, rays / std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() / 1e6);
        }
        return best;
    };
    int trials = 0;
    const LaunchOptions best = tuneSearch(space, settings->maxBounces, measure, &trials);
    RtReset(s);
    s->hasFrame = false;
    if (failed || best.mraysPerSecond <= 0 || !apply(best)) {
        if (s->hostError.empty()) s->hostError = "autotune trial failed";
        return 0;
    }
    s->tunedBounces = settings->maxBounces;
    if (!tuneStore(sessionTuneKey((long long)s->tris.size(), s->width, s->height, settings->maxBounces), best))
        s->hostError = "autotune: could not write " + tuneCachePath();
    if (mraysPerSecond) *mraysPerSecond = (float)best.mraysPerSecond;
    if (trialCount) *trialCount = trials;
    return 1;
}

int RtCopyOutput(void* p, unsigned char* output, int outputBytes) {
    Session* s = static_cast<Session*>(p);
    if (!s) return 0;
    if (!s->ready || !s->hasFrame || !output || outputBytes < 0 || (size_t)outputBytes < (size_t)s->width * s->height * 4) {
        s->error = hipErrorInvalidValue;
        return 0;
    }
    s->error = hipMemcpy(output, s->dPixels, (size_t)s->width * s->height * 4, hipMemcpyDeviceToHost);
    return s->error == hipSuccess ? 1 : 0;
}

int RtStageTimes(void* p, float* milliseconds, int* launches, int count) {
    Session* s = static_cast<Session*>(p);
    if (!s || count < 0) return 0;
    for (int k = 0; k < count && k < RT_STAGE_COUNT; k++) {
        if (milliseconds) milliseconds[k] = s->stageMs[k];
        if (launches) launches[k] = s->stageLaunches[k];
    }
    return 1;
}

int RtStageBusy(void* p, unsigned stageMask, float* milliseconds) {
    Session* s = static_cast<Session*>(p);
    if (!s || !milliseconds) return 0;
    std::vector<std::pair<float, float>> spans;
    for (const auto& l : s->launchSpans)
        if (stageMask & (1u << (int)l[0])) spans.emplace_back(l[1], l[2]);
    std::sort(spans.begin(), spans.end());
    float busy = 0, end = -1e30f;
    for (const auto& sp : spans) {
        if (sp.first > end) { busy += sp.second - sp.first; end = sp.second; }
        else if (sp.second > end) { busy += sp.second - end; end = sp.second; }
    }
    *milliseconds = busy;
    return 1;
}

int RtTraceStats(void* p, unsigned long long* counts4) {
    Session* s = static_cast<Session*>(p);
    if (!s || !counts4) return 0;
    for (int k = 0; k < 4; k++) counts4[k] = s->traceCounts[k];
    return 1;
}

int RtBandwidthProbe(long long bytes, int iterations, float* milliseconds) {
    if (bytes < 16 || iterations < 1 || !milliseconds) return 0;
    const long long n = bytes / 16;
    float4 *a = nullptr, *b = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int device = 0, cus = 0;
    bool ok = hipGetDevice(&device) == hipSuccess &&
              hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess &&
              hipMalloc(&a, n * 16) == hipSuccess && hipMalloc(&b, n * 16) == hipSuccess &&
              hipMemset(a, 0, n * 16) == hipSuccess && hipEventCreate(&e0) == hipSuccess &&
              hipEventCreate(&e1) == hipSuccess;
    if (ok) ok = launchCopy(a, b, n, cus * 8, nullptr) == hipSuccess && hipDeviceSynchronize() == hipSuccess;
    if (ok) ok = hipEventRecord(e0, nullptr) == hipSuccess;
    for (int k = 0; ok && k < iterations; k++) ok = launchCopy(a, b, n, cus * 8, nullptr) == hipSuccess;
    if (ok) ok = hipEventRecord(e1, nullptr) == hipSuccess && hipEventSynchronize(e1) == hipSuccess &&
                 hipEventElapsedTime(milliseconds, e0, e1) == hipSuccess;
    if (ok) *milliseconds /= iterations;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return ok ? 1 : 0;
}

// Default weights: $RT_DENOISER_WEIGHTS, else denoiser_weights.bin next to this library.
static std::string defaultWeightsPath() {
    if (const char* env = std::getenv("RT_DENOISER_WEIGHTS")) return env;
    Dl_info info;
    if (dladdr(reinterpret_cast<void*>(&RtCreate), &info) && info.dli_fname) {
        std::string so = info.dli_fname;
        const size_t slash = so.rfind('/');
        return (slash == std::string::npos ? std::string(".") : so.substr(0, slash)) + "/denoiser_weights.bin";
    }
    return "denoiser_weights.bin";
}

static bool ensureDenoiser(Session* s) {
    if (s->denoiser) return true;
    DenoiserNet* net = denoiserCreate();
    if (!net) { s->hostError = "denoiser: out of device memory"; return false; }
    if (!denoiserLoad(net, defaultWeightsPath().c_str(), s->hostError)) { denoiserDestroy(net); return false; }
    s->denoiser = net;
    s->ownDenoiser = true;
    return true;
}

// Uploads the CPU backends' accumulation so the GPU stages see the same sums.
static bool uploadHostSums(Session* s) {
    const size_t npix = (size_t)s->width * s->height;
    std::vector<float4> r(npix), a(npix), n(npix);
    for (size_t i = 0; i < npix; i++) {
        const HostSum& h = s->hostSum[i];
        r[i] = make_float4(h.rad.x, h.rad.y, h.rad.z, h.covered);
        a[i] = make_float4(h.alb.x, h.alb.y, h.alb.z, h.lum2);
        n[i] = make_float4(h.nrm.x, h.nrm.y, h.nrm.z, h.depth);
    }
    return (s->error = hipMemcpy(s->dSumRad, r.data(), npix * 16, hipMemcpyHostToDevice)) == hipSuccess &&
           (s->error = hipMemcpy(s->dSumAlb, a.data(), npix * 16, hipMemcpyHostToDevice)) == hipSuccess &&
           (s->error = hipMemcpy(s->dSumNrm, n.data(), npix * 16, hipMemcpyHostToDevice)) == hipSuccess;
}

static bool denoiseBuffers(Session* s, int& hp, int& wp) {
    hp = (s->height + 15) / 16 * 16;
    wp = (s->width + 15) / 16 * 16;
    if (!s->dDenoiseIn && !devAlloc(s, s->dDenoiseIn, (size_t)hp * wp * kDenoiseInputChannels)) return false;
    if (!s->dDenoiseOut && !devAlloc(s, s->dDenoiseOut, (size_t)hp * wp * 4)) return false;
    if (!s->dTileStats && !devAlloc(s, s->dTileStats, kTileStatSlots)) return false;
    if (s->denoiseEnsemble > 1) {
        if (!s->dDenoiseAuxIn && !devAlloc(s, s->dDenoiseAuxIn, (size_t)hp * wp * kDenoiseInputChannels)) return false;
        if (!s->dDenoiseAuxOut && !devAlloc(s, s->dDenoiseAuxOut, (size_t)hp * wp * 4)) return false;
    }
    return true;
}

int RtDenoise(void* p, unsigned char* output, int outputBytes, float* milliseconds) {
    Session* s = static_cast<Session*>(p);
    if (!s) return 0;
    if (!s->ready || !s->samples || !output || !milliseconds || outputBytes < 0 ||
        (size_t)outputBytes < (size_t)s->width * s->height * 4) {
        s->error = hipErrorInvalidValue;
        return 0;
    }
    auto t0 = std::chrono::steady_clock::now();
    s->hostError.clear();
    if (!ensureDenoiser(s)) return 0;
    if (s->lastBackend != RT_BACKEND_GPU && !uploadHostSums(s)) return 0;
    int hp, wp;
    if (!denoiseBuffers(s, hp, wp)) return 0;
    hipStream_t st = s->stream;
    DenoiseArgs a{};
    a.sumRad = s->dSumRad; a.sumAlb = s->dSumAlb; a.sumNrm = s->dSumNrm;
    a.w = s->width; a.h = s->height; a.wp = wp; a.hp = hp; a.samples = s->samples;
    a.planes = s->dDenoiseIn; a.denoised = s->dDenoiseOut; a.pixels = s->dPixels; a.tileStats = s->dTileStats;
    a.exposure = s->lastSettings.exposure; a.fallback = s->denoiseFallback;
    if ((s->error = launchDenoisePrepare(a, st)) != hipSuccess) return 0;
    if ((s->error = denoiserInferDevice(s->denoiser, s->dDenoiseIn, 1, hp, wp, s->dDenoiseOut, st)) != hipSuccess) return 0;
    s->lastInferMs = denoiserLastForwardMs(s->denoiser);
    // Self-ensemble: the network also sees the input mirrored (in x; or in x, y and both),
    // its outputs mirrored back and averaged with the first (training flips crops the same way).
    for (int e = 1; e < s->denoiseEnsemble; e++) {
        const int flip = s->denoiseEnsemble == 2 ? 1 : e;
        if ((s->error = launchMirrorPlanes(s->dDenoiseIn, s->dDenoiseAuxIn, hp, wp, kDenoiseInputChannels, flip, st)) !=
                hipSuccess ||
            (s->error = denoiserInferDevice(s->denoiser, s->dDenoiseAuxIn, 1, hp, wp, s->dDenoiseAuxOut, st)) != hipSuccess ||
            (s->error = launchAddMirrored(s->dDenoiseOut, s->dDenoiseAuxOut, hp, wp, 4, flip,
                                          e + 1 == s->denoiseEnsemble ? 1.0f / s->denoiseEnsemble : 1.0f, st)) != hipSuccess)
            return 0;
        s->lastInferMs += denoiserLastForwardMs(s->denoiser);
    }
    if ((s->error = hipMemsetAsync(s->dTileStats, 0, kTileStatSlots * sizeof(int), st)) != hipSuccess) return 0;
    if ((s->error = launchDenoisePresent(a, st)) != hipSuccess) return 0;
    if ((s->error = hipStreamSynchronize(st)) != hipSuccess) return 0;
    int stats[kTileStatSlots];
    if ((s->error = hipMemcpy(stats, s->dTileStats, sizeof(stats), hipMemcpyDeviceToHost)) != hipSuccess) return 0;
    s->lastTiles = ((s->width + 15) / 16) * ((s->height + 15) / 16);   // one workgroup per 16x16 tile
    s->lastRejected = 0;
    for (int k = 0; k < kTileStatSlots; k++) s->lastRejected += stats[k];
    s->error = hipMemcpy(output, s->dPixels, (size_t)s->width * s->height * 4, hipMemcpyDeviceToHost);
    if (s->error != hipSuccess) return 0;
    *milliseconds = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
    s->error = hipSuccess;
    return 1;
}

// ------------------------------------------------------------------ extensions
static_assert(sizeof(RtRay) == 2 * sizeof(float4) && sizeof(RtHit) == sizeof(int4), "query records match the queues");

int RtTraceRays(void* p, const RtRay* rays, int count, int anyHit, RtHit* hits) {
    Session* s = static_cast<Session*>(p);
    if (!s) return 0;
    // The traversal's int ray cursor runs past the count by up to 64 rays per wave.
    if (!s->ready || !rays || !hits || count < 0 || count > (1 << 30)) { s->error = hipErrorInvalidValue; return 0; }
    if (count == 0) return 1;
    if (!prepareTraceGrid(s)) return 0;
    if ((size_t)count > s->queryCap) {
        devFree(s->dQueryRays); devFree(s->dQueryHits);
        s->queryCap = 0;
        if (!devAlloc(s, s->dQueryRays, 2 * (size_t)count) || !devAlloc(s, s->dQueryHits, (size_t)count)) return 0;
        s->queryCap = (size_t)count;
    }
    if (!s->dQueryWork && !devAlloc(s, s->dQueryWork, 1)) return 0;
    float4* dr = s->dQueryRays; int4* dh = s->dQueryHits; int* work = s->dQueryWork;
    bool ok = (s->error = hipMemcpyAsync(dr, rays, (size_t)count * sizeof(RtRay), hipMemcpyHostToDevice, s->stream)) ==
              hipSuccess;
    if (ok) ok = (s->error = hipMemsetAsync(work, 0, sizeof(int), s->stream)) == hipSuccess;
    if (ok) ok = (s->error = launchZeroInts(s->dOverflow, 1, s->stream)) == hipSuccess;
    if (ok) {
        TraceArgs a = baseTraceArgs(s);
        a.work = work;
        if (anyHit) { a.anyRays = dr; a.anyCount = count; a.anyHits = dh; }
        else { a.rays = dr; a.count = count; a.hits = dh; }
        a.totalLanes = s->queryBlocks * kBlock;
        ok = (s->error = launchTrace(traceLaunch(s), false, s->queryBlocks, a, s->stream)) == hipSuccess;
    }
    if (ok) ok = (s->error = hipStreamSynchronize(s->stream)) == hipSuccess;
    if (ok) ok = (s->error = hipMemcpy(hits, dh, (size_t)count * sizeof(RtHit), hipMemcpyDeviceToHost)) == hipSuccess;
    int overflow = 0;
    if (ok) ok = (s->error = hipMemcpy(&overflow, s->dOverflow, sizeof(int), hipMemcpyDeviceToHost)) == hipSuccess;
    if (ok && overflow) { s->hostError = "traversal stack exceeded its spill ring"; ok = false; }
    return ok ? 1 : 0;
}

int RtTraceRaysNoAlloc(void* p, const RtRay* deviceRays, int count, int anyHit, RtHit* deviceHits) {
    Session* s = static_cast<Session*>(p);
    if (!s) return 0;
    s->hostError.clear();
    if (!s->ready || !deviceRays || !deviceHits || count < 0 || count > (1 << 30)) {
        s->error = hipErrorInvalidValue;
        return 0;
    }
    if (count == 0) return 1;
    void* base = nullptr;
    {
        std::lock_guard<std::mutex> lock(gScratchMutex);
        base = (s->device >= 0 && s->device < kMaxDevices) ? gScratch[s->device] : nullptr;
    }
    if (!base) { s->hostError = "no-allocation scratch buffer not set up on this device"; s->error = hipErrorInvalidValue; return 0; }
    // Stack entries a lane can hold beyond its LDS ring: (W-1) per inner node on the deepest path.
    const int ldsEntries = (s->shortStack == 12 && !(s->builtWidth == 4 && !s->builtQuantized)) ? 16 : s->shortStack;
    const int spillEntries = std::max(1, (s->builtWidth - 1) * (s->wideDepth + 1) - ldsEntries);
    const size_t laneBytes = (size_t)spillEntries * 8;
    const long long fitBlocks = (long long)((kNoAllocScratchBytes - kNoAllocHeaderBytes) / laneBytes / kBlock);
    const int blocks = (int)std::min<long long>((long long)s->cus * std::max(s->blocksPerCU, kQueryBlocksPerCU), fitBlocks);
    if (blocks < s->cus) {
        s->hostError = "BVH too deep for the no-allocation scratch buffer (" + std::to_string(kNoAllocScratchBytes >> 20) +
                       " MiB): needs " + std::to_string(spillEntries) + " spill entries per lane";
        s->error = hipErrorInvalidValue;
        return 0;
    }
    int* header = static_cast<int*>(base);
    TraceArgs a = baseTraceArgs(s);
    a.spill = reinterpret_cast<int*>(static_cast<unsigned char*>(base) + kNoAllocHeaderBytes);
    a.spillEntries = spillEntries;
    a.totalLanes = blocks * kBlock;
    a.work = header;
    a.overflow = header + 1;
    const float4* dr = reinterpret_cast<const float4*>(deviceRays);
    int4* dh = reinterpret_cast<int4*>(deviceHits);
    if (anyHit) { a.anyRays = dr; a.anyCount = count; a.anyHits = dh; }
    else { a.rays = dr; a.count = count; a.hits = dh; }
    int overflow = 0;
    bool ok = (s->error = launchZeroInts(header, 2, s->stream)) == hipSuccess;
    if (ok) ok = (s->error = launchTrace(traceLaunch(s), false, blocks, a, s->stream)) == hipSuccess;
    if (ok) ok = (s->error = hipMemcpyAsync(&overflow, header + 1, sizeof(int), hipMemcpyDeviceToHost, s->stream)) ==
                 hipSuccess;
    if (ok) ok = (s->error = hipStreamSynchronize(s->stream)) == hipSuccess;
    if (ok && overflow) { s->hostError = "traversal stack exceeded its spill ring"; ok = false; }
    return ok ? 1 : 0;
}

int RtSobolBatch(const unsigned* pixels, const unsigned* samples, int count, unsigned seed, int dimensions, float* out) {
    if (!pixels || !samples || !out || count < 0 || dimensions < 1 || dimensions > kSobolDims) return 0;
    if (count == 0) return 1;
    unsigned *dp = nullptr, *ds = nullptr;
    float* dout = nullptr;
    bool ok = hipMalloc(&dp, (size_t)count * 4) == hipSuccess && hipMalloc(&ds, (size_t)count * 4) == hipSuccess &&
              hipMalloc(&dout, (size_t)count * dimensions * 4) == hipSuccess;
    if (ok) ok = hipMemcpy(dp, pixels, (size_t)count * 4, hipMemcpyHostToDevice) == hipSuccess &&
                 hipMemcpy(ds, samples, (size_t)count * 4, hipMemcpyHostToDevice) == hipSuccess;
    if (ok) ok = launchSobolBatch(dp, ds, count, seed, dimensions, dout, nullptr) == hipSuccess;
    if (ok) ok = hipMemcpy(out, dout, (size_t)count * dimensions * 4, hipMemcpyDeviceToHost) == hipSuccess;
    if (dp) (void)hipFree(dp);
    if (ds) (void)hipFree(ds);
    if (dout) (void)hipFree(dout);
    return ok ? 1 : 0;
}

int RtBvhExport(void* p, RtBvhNode* nodes, int nodeCapacity, int* order, int orderCapacity) {
    Session* s = static_cast<Session*>(p);
    if (!s || !s->ready) return 0;
    if ((nodes || order) && !fetchBinaryTree(s)) return 0;
    if (nodes && nodeCapacity > 0)
        std::memcpy(nodes, s->binNodes.data(), sizeof(RtBvhNode) * std::min<size_t>(nodeCapacity, s->binNodes.size()));
    if (order && orderCapacity > 0)
        std::memcpy(order, s->binOrder.data(), sizeof(int) * std::min<size_t>(orderCapacity, s->binOrder.size()));
    return s->binNodeCount;
}

int RtReadAccumulation(void* p, float* out, long long floatCount) {
    Session* s = static_cast<Session*>(p);
    if (!s || !s->ready || !out) return 0;
    const size_t npix = (size_t)s->width * s->height;
    if (floatCount < (long long)(npix * 10)) { s->error = hipErrorInvalidValue; return 0; }
    if (s->lastBackend == RT_BACKEND_GPU) {
        std::vector<float4> r(npix), a(npix), n(npix);
        if ((s->error = hipMemcpy(r.data(), s->dSumRad, npix * 16, hipMemcpyDeviceToHost)) != hipSuccess ||
            (s->error = hipMemcpy(a.data(), s->dSumAlb, npix * 16, hipMemcpyDeviceToHost)) != hipSuccess ||
            (s->error = hipMemcpy(n.data(), s->dSumNrm, npix * 16, hipMemcpyDeviceToHost)) != hipSuccess)
            return 0;
        for (size_t i = 0; i < npix; i++) {
            float* o = out + i * 10;
            o[0] = r[i].x; o[1] = r[i].y; o[2] = r[i].z; o[3] = a[i].x; o[4] = a[i].y; o[5] = a[i].z;
            o[6] = n[i].x; o[7] = n[i].y; o[8] = n[i].z; o[9] = r[i].w;
        }
    } else {
        for (size_t i = 0; i < npix; i++) {
            const HostSum& h = s->hostSum[i];
            float* o = out + i * 10;
            o[0] = h.rad.x; o[1] = h.rad.y; o[2] = h.rad.z; o[3] = h.alb.x; o[4] = h.alb.y; o[5] = h.alb.z;
            o[6] = h.nrm.x; o[7] = h.nrm.y; o[8] = h.nrm.z; o[9] = h.covered;
        }
    }
    return 1;
}

int RtRayCounts(void* p, unsigned long long* closest, unsigned long long* shadow) {
    Session* s = static_cast<Session*>(p);
    if (!s || !closest || !shadow) return 0;
    *closest = s->lastClosest; *shadow = s->lastShadow;
    return 1;
}

int RtSetDenoiser(void* p, void* net) {
    Session* s = static_cast<Session*>(p);
    if (!s) return 0;
    if (s->ownDenoiser) denoiserDestroy(s->denoiser);
    s->denoiser = static_cast<DenoiserNet*>(net);
    s->ownDenoiser = false;
    return 1;
}

int RtDenoiseStats(void* p, int* tiles, int* rejected, float* inferMs) {
    Session* s = static_cast<Session*>(p);
    if (!s) return 0;
    if (tiles) *tiles = s->lastTiles;
    if (rejected) *rejected = s->lastRejected;
    if (inferMs) *inferMs = s->lastInferMs;
    return 1;
}

int RtReadDenoiserInputs(void* p, float* out, long long floatCount) {
    Session* s = static_cast<Session*>(p);
    if (!s || !s->ready || !s->samples || !out) return 0;
    const size_t npix = (size_t)s->width * s->height;
    if (floatCount < (long long)(npix * kDenoiseInputChannels)) { s->error = hipErrorInvalidValue; return 0; }
    if (s->lastBackend != RT_BACKEND_GPU && !uploadHostSums(s)) return 0;
    int hp, wp;
    if (!denoiseBuffers(s, hp, wp)) return 0;
    DenoiseArgs a{};
    a.sumRad = s->dSumRad; a.sumAlb = s->dSumAlb; a.sumNrm = s->dSumNrm;
    a.w = s->width; a.h = s->height; a.wp = wp; a.hp = hp; a.samples = s->samples; a.planes = s->dDenoiseIn;
    if ((s->error = launchDenoisePrepare(a, s->stream)) != hipSuccess) return 0;
    if ((s->error = hipStreamSynchronize(s->stream)) != hipSuccess) return 0;
    std::vector<float> planes((size_t)hp * wp * kDenoiseInputChannels);
    if ((s->error = hipMemcpy(planes.data(), s->dDenoiseIn, planes.size() * 4, hipMemcpyDeviceToHost)) != hipSuccess)
        return 0;
    for (int y = 0; y < s->height; y++)
        std::memcpy(out + (size_t)y * s->width * kDenoiseInputChannels, planes.data() + (size_t)y * wp * kDenoiseInputChannels,
                    (size_t)s->width * kDenoiseInputChannels * 4);
    return 1;
}

// ------------------------------------------------------------------ denoiser network
void* RtDenoiserCreate(void) { return denoiserCreate(); }
void RtDenoiserDestroy(void* net) { denoiserDestroy(static_cast<DenoiserNet*>(net)); }
long long RtDenoiserParamCount(void* net) { return denoiserParamCount(static_cast<DenoiserNet*>(net)); }
int RtDenoiserInit(void* net, unsigned seed) { return net && denoiserInit(static_cast<DenoiserNet*>(net), seed) == hipSuccess; }
int RtDenoiserGetParams(void* net, float* params) {
    return net && params && denoiserGetParams(static_cast<DenoiserNet*>(net), params) == hipSuccess;
}
int RtDenoiserSetParams(void* net, const float* params) {
    return net && params && denoiserSetParams(static_cast<DenoiserNet*>(net), params) == hipSuccess;
}
int RtDenoiserSave(void* net, const char* path) {
    std::string err;
    return net && path && denoiserSave(static_cast<DenoiserNet*>(net), path, err);
}
int RtDenoiserLoad(void* net, const char* path) {
    std::string err;
    return net && path && denoiserLoad(static_cast<DenoiserNet*>(net), path, err);
}
int RtDenoiserInfer(void* net, const float* input, int n, int h, int w, float* output) {
    if (!net || !input || !output || n < 1 || h < 16 || w < 16 || h % 16 || w % 16) return 0;
    return denoiserInferHost(static_cast<DenoiserNet*>(net), input, n, h, w, output) == hipSuccess;
}
int RtDenoiserGradients(void* net, const float* input, const float* target, int n, int h, int w, float* loss,
                        float* grads) {
    if (!net || !input || !target || !loss || !grads || n < 1 || h != w || h < 16 || h % 16) return 0;
    return denoiserGradients(static_cast<DenoiserNet*>(net), input, target, n, h, w, loss, grads) == hipSuccess;
}
int RtDenoiserSetDataset(void* net, const float* inputs, const float* targets, int count, int h, int w) {
    if (!net || !inputs || !targets || count < 1 || h < 16 || w < 16) return 0;
    return denoiserSetDataset(static_cast<DenoiserNet*>(net), inputs, targets, count, h, w) == hipSuccess;
}
int RtDenoiserTrain(void* net, int steps, int batch, int crop, float learningRate, unsigned seed, float* losses) {
    if (!net || steps < 1 || batch < 1 || crop < 16 || crop % 16 || !(learningRate > 0)) return 0;
    return denoiserTrain(static_cast<DenoiserNet*>(net), steps, batch, crop, learningRate, seed, losses) == hipSuccess;
}
int RtDenoiserSetDisplayLoss(void* net, float weight) {
    return denoiserSetDisplayWeight(static_cast<DenoiserNet*>(net), weight) ? 1 : 0;
}
int RtDenoiserTimings(void* net, float* forwardMs, float* trainStepMs) {
    if (!net) return 0;
    if (forwardMs) *forwardMs = denoiserLastForwardMs(static_cast<DenoiserNet*>(net));
    if (trainStepMs) *trainStepMs = denoiserLastTrainStepMs(static_cast<DenoiserNet*>(net));
    return 1;
}
double RtDenoiserForwardFlops(int n, int h, int w) { return denoiserForwardFlops(n, h, w); }
int RtConv3x3Forward(const float* in, int n, int h, int w, int cin, const float* weight, const float* bias, int cout,
                     int relu, float* out) {
    if (!in || !weight || !out || n < 1 || h < 1 || w < 1 || cin < 1 || cout < 1 || cin > 256 || cout > 256) return 0;
    return convForwardHost(in, n, h, w, cin, weight, bias, cout, relu, out) == hipSuccess;
}
int RtConv3x3Backward(const float* in, int n, int h, int w, int cin, const float* weight, const float* gradOut, int cout,
                      float* gradIn, float* gradWeight, float* gradBias) {
    if (!in || !weight || !gradOut || !gradIn || !gradWeight || !gradBias || n < 1 || h < 1 || w < 1 || cin < 1 ||
        cout < 1 || cin > 256 || cout > 256)
        return 0;
    return convBackwardHost(in, n, h, w, cin, weight, gradOut, cout, gradIn, gradWeight, gradBias) == hipSuccess;
}

}  // extern "C"

namespace {
int setOption(Session* s, const char* key, int value) {
    if (!std::strcmp(key, "autotune_cache")) {
        if (value != 0 && value != 1) return 0;
        s->autotuneCache = value; return 1;
    }
    if (!std::strcmp(key, "bvh_width")) {
        if (value != 2 && value != 4 && value != 8) return 0;
        s->bvhWidth = value; return 1;
    }
    if (!std::strcmp(key, "bvh_quantized")) {
        if (value != 0 && value != 1) return 0;
        s->bvhQuantized = value; return 1;
    }
    if (!std::strcmp(key, "build_bins")) {
        if (value < 2 || value > 32) return 0;
        s->buildBins = value; return 1;
    }
    if (!std::strcmp(key, "build_max_leaf")) {
        if (value < 1 || value > 16) return 0;
        s->buildMaxLeaf = value; return 1;
    }
    if (!std::strcmp(key, "blocks_per_cu")) {
        if (value < 1 || value > 32) return 0;
        s->blocksPerCU = value; return 1;
    }
    if (!std::strcmp(key, "denoise_fallback")) {
        if (value < 0 || value > 2) return 0;
        s->denoiseFallback = value; return 1;
    }
    if (!std::strcmp(key, "short_stack")) {
        if (value != 8 && value != 12 && value != 16) return 0;
        s->shortStack = value; return 1;
    }
    if (!std::strcmp(key, "refill")) {
        if (value != 0 && value != 1) return 0;
        s->refill = value; return 1;
    }
    if (!std::strcmp(key, "refill_below")) {
        if (value < 1 || value > 64) return 0;
        s->refillBelow = value; return 1;
    }
    if (!std::strcmp(key, "pipelines")) {
        if (value < 1 || value > kPipelines) return 0;
        s->pipelines = value; return 1;
    }
    if (!std::strcmp(key, "tail_bounce")) {
        if (value < 0 || value > kMaxBounces) return 0;
        s->tailBounce = value; return 1;
    }
    if (!std::strcmp(key, "denoise_ensemble")) {
        if (value != 1 && value != 2 && value != 4) return 0;
        s->denoiseEnsemble = value; return 1;
    }
    if (!std::strcmp(key, "while_while")) {
        if (value != 0 && value != 1) return 0;
        s->whileWhile = value; return 1;
    }
    if (!std::strcmp(key, "tiled_primary")) {
        if (value != 0 && value != 1) return 0;
        s->tiledPrimary = value; return 1;
    }
    if (!std::strcmp(key, "nontemporal")) {
        if (value != 0 && value != 1) return 0;
        s->nontemporal = value; return 1;
    }
    if (!std::strcmp(key, "xcd_segments")) {
        if (value != 0 && value != 1) return 0;
        s->xcdSegments = value; return 1;
    }
    if (!std::strcmp(key, "intersect_watertight")) {
        if (value != 0 && value != 1) return 0;
        s->watertight = value; return 1;
    }
    if (!std::strcmp(key, "stage_timing")) {
        if (value != 0 && value != 1) return 0;
        s->stageTiming = value != 0; return 1;
    }
    if (!std::strcmp(key, "trace_stats")) {
        if (value != 0 && value != 1) return 0;
        s->traceStats = value != 0; return 1;
    }
    return 0;
}
}  // namespace

extern "C" {

int RtSetOption(void* p, const char* key, int value) {
    Session* s = static_cast<Session*>(p);
    if (!s || !key) return 0;
    if (!setOption(s, key, value)) return 0;
    for (int k = 0; k < 8; k++)
        if (!std::strcmp(key, kTunedFields[k].name)) s->userSet |= 1u << k;
    return 1;
}

int RtGetOption(void* p, const char* key) {
    Session* s = static_cast<Session*>(p);
    if (!s || !key) return -1;
    if (!std::strcmp(key, "bvh_width")) return s->ready ? s->builtWidth : s->bvhWidth;
    if (!std::strcmp(key, "bvh_quantized")) return s->ready ? s->builtQuantized : s->bvhQuantized;
    if (!std::strcmp(key, "node_bytes")) return (int)wideNodeBytes(s->ready ? s->builtWidth : s->bvhWidth,
                                                                    s->ready ? s->builtQuantized : s->bvhQuantized);
    if (!std::strcmp(key, "blocks_per_cu")) return s->blocksPerCU;
    if (!std::strcmp(key, "build_bins")) return s->buildBins;
    if (!std::strcmp(key, "build_max_leaf")) return s->buildMaxLeaf;
    if (!std::strcmp(key, "short_stack")) return s->shortStack;
    if (!std::strcmp(key, "denoise_fallback")) return s->denoiseFallback;
    if (!std::strcmp(key, "refill")) return s->refill;
    if (!std::strcmp(key, "refill_below")) return s->refillBelow;
    if (!std::strcmp(key, "pipelines")) return s->pipelines;
    if (!std::strcmp(key, "tail_bounce")) return s->tailBounce;
    if (!std::strcmp(key, "denoise_ensemble")) return s->denoiseEnsemble;
    if (!std::strcmp(key, "while_while")) return s->whileWhile;
    if (!std::strcmp(key, "tiled_primary")) return s->tiledPrimary;
    if (!std::strcmp(key, "xcd_segments")) return s->xcdSegments;
    if (!std::strcmp(key, "nontemporal")) return s->nontemporal;
    if (!std::strcmp(key, "stage_timing")) return s->stageTiming ? 1 : 0;
    if (!std::strcmp(key, "trace_stats")) return s->traceStats ? 1 : 0;
    if (!std::strcmp(key, "autotune_cache")) return s->autotuneCache;
    if (!std::strcmp(key, "intersect_watertight")) return s->watertight;
    return -1;
}

}  // extern "C"
