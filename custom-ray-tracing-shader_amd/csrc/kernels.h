// kernels.h - launch interface of the gfx950 kernels (defined in kernels.hip).
#pragma once
#include "rt_core.h"

namespace rt {

constexpr int kBlock = 256;       // threads per workgroup for every kernel here
constexpr int kCounterStride = 4 * (kMaxBounces + 1);   // ints per chunk of frame counters
// Frame counter layout within one chunk: [0,33) closest-ray queue sizes per bounce,
// [33,66) shadow queue sizes, [66,99) work cursors of the traversal launches (launch b
// traces bounce b's closest-hit rays and bounce b-1's shadow rays), [99,132) unused.
constexpr int kCntRays = 0, kCntShadow = kMaxBounces + 1, kWorkRays = 2 * (kMaxBounces + 1);
// Segmented work cursors: a launch's ray index space is cut into kXcds contiguous segments,
// one per group of workgroups that share an XCD (workgroups are dispatched to the eight XCDs
// round-robin, so blockIdx.x % kXcds picks the group), each with its own cursor, kept
// kXcdCursorStride ints apart.  A pipeline holds one set per work cursor slot above.
// 4-wide full-precision nodes are addressed by 32-bit byte offsets in the traversal.
constexpr long long kMaxWide4Nodes = 1LL << 25;
constexpr int kXcds = 8, kXcdCursorStride = 32, kXcdCursorInts = kXcds * kXcdCursorStride;
constexpr int kTileStatSlots = 64;   // counters of the denoiser's rejected tiles
constexpr int kXcdSlotInts = (kMaxBounces + 1) * kXcdCursorInts;   // per pipeline

// One traversal launch covers two queues: closest-hit rays, then any-hit rays (ray indices
// past the closest-hit count).  Either count may be zero.
struct TraceArgs {
    const float4* rays;      // closest-hit queue, 2 float4 per ray: (origin, limit), (direction, path slot bits)
    const int* countPtr;     // device-resident ray count, or null to use count
    int count;
    const float4* anyRays;   // any-hit queue, same layout
    const int* anyCountPtr;  // device-resident any-hit count, or null to use anyCount
    int anyCount;
    int* work;               // work cursor, starts at 0
    const void* nodes;       // WideNode<W>[] or QNode<W>[]
    const TriRecord* recs;
    int4* hits;              // RtHit per closest-hit ray
    int4* anyHits;           // query mode: RtHit per any-hit ray; null in frame mode
    float4* rad;             // frame mode: radiance += contrib of an unoccluded any-hit ray's slot
    const float4* contrib;
    int* spill;              // global spill ring: 2 ints per entry, entry-major
    int spillEntries;        // entries per lane beyond the LDS part
    int totalLanes;          // launched lanes (spill stride)
    unsigned long long* stats;   // optional: node visits, triangle tests of closest-hit rays, then of any-hit rays
    int* overflow;           // set to 1 if a lane's stack exceeded the spill ring
    int nontemporal;         // 1: stream rays in and hits out past the caches
    const int* roots;        // batched views: root node per view (null: root 0)
    const int* viewBase;     // batched views: first path slot per view, views + 1 entries
    int views;
    int refillBelow;         // set by launchTrace from TraceLaunch::refillBelow
    int* xcdWork;            // optional: kXcds segment cursors (kXcdCursorStride apart), all 0 at
                             // launch; null: one cursor (work) for the whole index space
    int watertight;          // 1: watertight ray/triangle test instead of the reference's
};

struct FrameArgs {
    CameraData cam;
    RenderSettings rs;
    SceneView sc;
    int w, h, npix, R;
    unsigned sample0;        // sample index of slot group r = 0
    float4 *thr, *rad, *alb, *nrm, *nee;
    float4 *rayIn, *rayOut, *shadow;
    int4* hits;
    int* counters;           // this pipeline's counters
    int qBase, qCount;       // raygen: queue positions [qBase, qBase + qCount) of the chunk
    int bounce;
    int tiled;               // 1: primary rays queued in 8x8 pixel tiles
};

struct AccumArgs {
    float4 *thr, *rad, *alb, *nrm;
    float4 *sumRad, *sumAlb, *sumNrm;
    uchar4* pixels;
    int npix, R;
    int startFromZero;       // first chunk of a call with no prior samples
    int writeDisplay;        // last chunk: tone-map into pixels
    float invSamples, exposure;
};

struct DenoiseArgs {
    const float4 *sumRad, *sumAlb, *sumNrm;   // radiance+covered, albedo+lum^2, normal+depth
    int w, h, wp, hp;                          // image and padded (multiple of 16) sizes
    unsigned samples;
    float* planes;                             // [hp][wp][10] network input
    const float* denoised;                     // [hp][wp][4] network output radiance
    uchar4* pixels;
    int* tileStats;                            // [kTileStatSlots] counts of tiles that fell back
    float exposure;
    int fallback;                              // 0 per-tile check, 1 always fall back, 2 never
};
hipError_t launchDenoisePrepare(const DenoiseArgs& a, hipStream_t s);
hipError_t launchDenoisePresent(const DenoiseArgs& a, hipStream_t s);
// Self-ensemble of the denoiser: out = the [hp][wp][C] image mirrored in x (flip & 1) and/or
// y (flip & 2); acc = (acc + mirrored img) * scale.
hipError_t launchMirrorPlanes(const float* in, float* out, int hp, int wp, int channels, int flip, hipStream_t s);
hipError_t launchAddMirrored(float* acc, const float* img, int hp, int wp, int channels, int flip, float scale,
                             hipStream_t s);

// Variant of the persistent traversal kernel.
struct TraceLaunch {
    int width = 4;          // children per node: 2, 4 or 8
    bool quantized = false; // QNode (8-bit child bounds) instead of WideNode
    int stack = 8;          // LDS short-stack entries per lane: 8 or 16
    bool refill = false;    // lanes take new rays while the wave still traces others
    int refillBelow = 40;   // with refill: a wave fetches when fewer lanes than this are busy
    bool whileWhile = true; // speculative while-while loop (inner-node and leaf phases)
};
hipError_t launchTrace(const TraceLaunch& l, bool stats, int blocks, const TraceArgs& a, hipStream_t s);
// Zeroes one chunk's counters (closest-ray count of bounce 0 = initialRays) and, when
// xcdWork is not null, kXcdSlotInts segment cursors.
hipError_t launchResetCounters(int* counters, int initialRays, int* xcdWork, hipStream_t s);
hipError_t launchRaygen(const FrameArgs& f, hipStream_t s);
hipError_t launchShade(const FrameArgs& f, int blocks, hipStream_t s);
// Completes every path of the closest-hit queue a.rays (bounce firstBounce, >= 1) through its
// remaining bounces in one launch; a carries the traversal resources (no any-hit queue).
hipError_t launchPathTail(const TraceLaunch& l, bool stats, int blocks, int cus, const FrameArgs& f, const TraceArgs& a,
                          int firstBounce, hipStream_t s);
int shadeBlocksPerCU();   // resident shade workgroups per compute unit
hipError_t launchAccumulate(const AccumArgs& a, hipStream_t s);
hipError_t launchSobolBatch(const unsigned* pixels, const unsigned* samples, int count, unsigned seed, int dims,
                            float* out, hipStream_t s);
// ---- batched views: many scenes and cameras in one launch per stage ------------------
struct BatchView {
    CameraData cam;
    RenderSettings rs;         // raysPerPixel is 1: one sample per view per frame
    SceneView sc;              // triangles of the whole batch, this view's lights
    int w, h, slotBase;
    float4 *sumRad, *sumAlb, *sumNrm;
    uchar4* pixels;
};
struct BatchArgs {
    const BatchView* views;
    const int* viewBase;       // views + 1 entries
    int viewCount, totalSlots;
    unsigned* samples;         // per view samples so far (advanced on the device each frame)
    float4 *thr, *rad, *alb, *nrm, *nee;
    float4 *rayIn, *rayOut, *shadow;
    int4* hits;
    int* counters;             // this pipeline's counters
    int qBase, qCount;         // raygen: queue positions [qBase, qBase + qCount) of the frame
    int bounce;
};
hipError_t launchBatchRaygen(const BatchArgs& b, hipStream_t s);
hipError_t launchBatchShade(const BatchArgs& b, int blocks, hipStream_t s);
hipError_t launchBatchAccumulate(const BatchArgs& b, hipStream_t s);
hipError_t launchBatchAdvance(const BatchArgs& b, hipStream_t s);

hipError_t launchCopy(const float4* in, float4* out, long long n, int blocks, hipStream_t s);
// Sets n ints to zero with a kernel (no memset node, so the launch may sit in a captured graph).
hipError_t launchZeroInts(int* p, int n, hipStream_t s);
const unsigned* hostSobolDirections();

}  // namespace rt
