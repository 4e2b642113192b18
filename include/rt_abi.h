/*
 * rt_abi.h - the plain C ABI of libCudaRayTracer.so, the MI355X (gfx950) path tracer.
 *
 * The first part is the drop-in boundary: the same 13 exports, argument meaning, struct
 * layouts and 1 = success / 0 = failure contract that the reference plugin gives Unity
 * (P/Invoke, UnityRayTracer/Assets/RayTracing/RayTraceSession.cs:36-63) and the ctypes tools
 * (Tools/native_api.py:28-43).  Each declaration names the reference interface it replaces.
 *
 * The second part holds additive extension exports (ray-query batches, Sobol batches, BVH
 * dumps, HDR readback, ray counters, tuning knobs).  They have no counterpart in the
 * reference; the reference's own C++ tests reached the same internals by including
 * Native/renderer.cu directly (Tools/verify_bvh.cu:3,69-85).
 *
 * All structs are packed to 4 bytes and sequential, exactly as the reference's
 * (Native/renderer.cu:11-31, Tools/native_api.py:6-26).  Colours arrive sRGB-encoded.
 * Output images are RGBA8, row 0 = bottom row, alpha = 255.
 */
#ifndef RT_ABI_H
#define RT_ABI_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct RtVec3 { float x, y, z; } RtVec3;

/* Native/renderer.cu:11-18; 116 bytes.  vertices = origin, origin+edge1, origin+edge2. */
typedef struct Triangle {
    RtVec3 origin, edge1, edge2, color;  /* color: sRGB-encoded albedo */
    float emission;
    int oneSided;
    RtVec3 normal0, normal1, normal2;    /* per-vertex shading normals */
    int mesh;                            /* carried, unused by the tracer */
    RtVec3 specularColor;                /* sRGB-encoded */
    float smoothness, specularProbability;
} Triangle;

/* Native/renderer.cu:21-24; 64 bytes. */
typedef struct CameraData {
    RtVec3 origin, right, up, forward;
    float tanHalfFov, aspect, orthographicSize;
    int orthographic;
} CameraData;

/* Native/renderer.cu:25-30; 28 bytes. */
typedef struct RenderSettings {
    int maxBounces;
    float exposure, aperture, focusDistance, skyStrength;
    unsigned seed;
    int raysPerPixel;
} RenderSettings;

#ifdef __cplusplus
static_assert(sizeof(Triangle) == 116 && sizeof(CameraData) == 64 && sizeof(RenderSettings) == 28,
              "ABI struct sizes must match the reference plugin");
#endif

/* Backends accepted by RtRender (UnityRayTracer/Assets/RayTracing/CudaRenderer.cs:9-15). */
enum { RT_BACKEND_CPU_SINGLE = 0, RT_BACKEND_CPU_MULTI = 1, RT_BACKEND_GPU = 2, RT_BACKEND_EMBREE = 3 };

/* ---------------------------------------------------------------------------------------
 * Part 1: the reference's 13 exports (Native/renderer.cu:309-458)
 * ------------------------------------------------------------------------------------- */

/* Native/renderer.cu:309 - new, empty session. */
void* RtCreate(void);
/* Native/renderer.cu:332 - RtShutdown then delete. */
void RtDestroy(void* session);
/* Native/renderer.cu:319-331 - free every host/device buffer, keep the object. */
void RtShutdown(void* session);
/* Native/renderer.cu:310 - samples = 0, accumulation version + 1. */
void RtReset(void* session);
/* Native/renderer.cu:311 */
unsigned RtAccumulationVersion(void* session);
/* Native/renderer.cu:312 */
unsigned RtSampleCount(void* session);
/* Native/renderer.cu:313 - backend 3 scene build time; 0 when Embree is absent. */
float RtEmbreeBuildMilliseconds(void* session);
/* Native/renderer.cu:314 - 1 / OpenMP threads / 0 / Embree threads for backends 0/1/2/3. */
int RtCpuThreadCount(void* session);
/* Native/renderer.cu:315-317 - binary BVH node count, depth and GPU build milliseconds. */
void RtBvhInfo(void* session, int* nodes, int* depth, float* buildMs);
/* Native/renderer.cu:318 - host error text, else the HIP runtime's error string. */
const char* RtLastError(void* session);
/* Native/renderer.cu:340-373 - copy, linearise colours, build light CDF and BVH, upload. */
int RtInitialize(void* session, const Triangle* triangles, int count, int width, int height);
/* Native/renderer.cu:374-418 - raysPerPixel progressive samples per pixel, RGBA8 out.
 * computeMs = device time of the frame's kernels (backend 2) or CPU render loop time. */
int RtRender(void* session, const CameraData* camera, const RenderSettings* settings,
             unsigned char* output, int outputBytes, int backend, float* computeMs);
/* Native/renderer.cu:436-458 - learned denoiser over the accumulated HDR image and its
 * albedo/normal/depth guides; never feeds back into the accumulation. */
int RtDenoise(void* session, unsigned char* output, int outputBytes, float* milliseconds);

/* ---------------------------------------------------------------------------------------
 * Part 2: extension exports (no counterpart in the reference)
 * ------------------------------------------------------------------------------------- */

/* One ray query: 32 bytes. Hit accepted when 1e-4 < t < limit (Native/renderer.cu:111,153). */
typedef struct RtRay { float origin[3]; float limit; float direction[3]; int flags; } RtRay;
/* One hit record: 16 bytes, the reference's Hit (Native/renderer.cu:43); index -1 = miss. */
typedef struct RtHit { int index; float distance, u, v; } RtHit;
/* Reference BvhNode layout (Native/bvh.cuh:7): inner nodes have count 0, first = left child. */
typedef struct RtBvhNode { float low[3]; int first; float high[3]; int count; } RtBvhNode;

/* Closest-hit (anyHit = 0) or any-hit queries against the session's GPU BVH on the GPU;
 * at most 2^30 rays per call. */
int RtTraceRays(void* session, const RtRay* rays, int count, int anyHit, RtHit* hits);
/* No counterpart in the reference.  For callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): the
 * RtTraceRays query over rays and hits that already live in device memory the caller owns.
 * The traversal's work cursor, overflow flag and stack spill ring live in a library-static
 * 64 MiB device buffer, one per GPU, resolved by RtInitialize and held until the process
 * ends; the call makes no hipMalloc, hipFree or other allocation.  It launches as many
 * persistent workgroups as that buffer's spill ring serves at the session's BVH depth and
 * fails (RtLastError says why) when that is fewer than one per compute unit.  Every call on a
 * GPU shares that GPU's buffer, so calls that overlap on one GPU race.  Synchronous on the
 * session's stream (the caller's writes of the rays must be complete); at most 2^30 rays. */
int RtTraceRaysNoAlloc(void* session, const RtRay* deviceRays, int count, int anyHit, RtHit* deviceHits);
/* Scrambled Sobol values computed on the GPU: out[i*dimensions + d] = sample(pixels[i],
 * samples[i], seed, d) for d < dimensions <= 260 (Native/sampling.cuh:26-35). */
int RtSobolBatch(const unsigned* pixels, const unsigned* samples, int count, unsigned seed,
                 int dimensions, float* out);
/* Copy the binary GPU BVH back in the reference's node layout, plus the primitive order.
 * Returns the node count (0 on failure); copies min(capacity, size) entries. */
int RtBvhExport(void* session, RtBvhNode* nodes, int nodeCapacity, int* order, int orderCapacity);
/* Accumulated HDR sums: per pixel 10 floats = radiance xyz, albedo xyz, normal xyz, covered. */
int RtReadAccumulation(void* session, float* out, long long floatCount);
/* Ray queries issued by the last RtRender (closest-hit and shadow). */
int RtRayCounts(void* session, unsigned long long* closest, unsigned long long* shadow);
/* Tuning knobs: "bvh_width" (2, 4 or 8), "bvh_quantized" (0/1: 8-bit child bounds, rounded
 * outwards, so hits are unchanged), "build_bins" (SAH bins per axis of the GPU build, 2..32,
 * default 16) and "build_max_leaf" (leaf size limit, 1..16, default 4 as the reference
 * builder) take effect at the next RtInitialize; "blocks_per_cu" (persistent traversal
 * workgroups per compute unit and pipeline, 1..32, default 3), "short_stack" (LDS stack entries per traversal lane,
 * 8, 12 or 16, default 12; layouts other than 4-wide full-precision nodes use 16 for 12), "while_while" (0/1: speculative while-while traversal loop, default 1),
 * "refill" (0/1: lanes take new rays while the wave traces others, default 1),
 * "refill_below" (1..64: with refill, a wave fetches new rays once fewer lanes than this
 * are busy, default 40), "pipelines" (1..4: a frame's paths split over that many streams
 * whose launches overlap, at least 65536 paths each, default 4; frames with trace
 * statistics use one), "tail_bounce" (0 = off, default; k >= 1: from bounce k on, each path
 * runs to its end inside one launch instead of one launch per bounce; images unchanged),
 * "denoise_ensemble" (1, 2 or 4 network passes per RtDenoise: inputs mirrored in x, or in x, y
 * and both, outputs mirrored back and averaged; default 1), "xcd_segments" (0/1, default 1:
 * a frame's traversal launches with at least two rays per lane hand out rays from eight
 * contiguous segments, one per XCD group of workgroups, instead of one cursor; images
 * unchanged),
 * "intersect_watertight" (0/1, default 0: the reference's Moller-Trumbore test and its
 * acceptance rules, Native/renderer.cu:99-112; 1: the watertight test of Woop, Benthin and Wald
 * 2013 with the same one-sided and t > 1e-4 rules, so a ray through an edge or vertex shared
 * by two triangles cannot pass between them),
 * "tiled_primary", "nontemporal", "stage_timing", "trace_stats" and "denoise_fallback" at the
 * next render or query.  Returns 1 if the key is known and the value valid.  RtGetOption
 * returns -1 for unknown keys; it also reports "node_bytes", the size of one wide node. */
int RtSetOption(void* session, const char* key, int value);
int RtGetOption(void* session, const char* key);

/* No counterpart in the reference, which launches a fixed 16x16 grid (Native/renderer.cu:397-398).
 * The launch autotuner: times `frames` device-resident frames of camera/settings for each
 * launch configuration of a staged search (BVH width {4, 8} x leaf size {2, 4} x persistent
 * workgroups per CU x LDS stack x traversal loop at two pipelines, then pipelines, workgroups,
 * the path-completion bounce and stack x loop again; options the caller set with RtSetOption
 * stay fixed), leaves the session on the fastest (rebuilding its BVH if the layout changed;
 * samples reset) and stores it in the autotune cache ($RT_AUTOTUNE_CACHE, else
 * ~/.cache/rt_amd/autotune.json) under the scene class (triangle count to the nearest power of
 * two, image size, bounces).  RtInitialize applies a cached entry of the scene class, and each
 * render the entry of its bounce count, to every tuned option the caller has not set, unless
 * option "autotune_cache" is 0.  Reports the best Mrays/s and the number of trials. */
int RtAutotune(void* session, const CameraData* camera, const RenderSettings* settings, int frames,
               float* mraysPerSecond, int* trialCount);

/* RtRender on the GPU (backend 2) without the RGBA readback: the frame stays in device
 * memory for GPU-resident consumers; RtCopyOutput copies the last frame to the host. */
int RtRenderResident(void* session, const CameraData* camera, const RenderSettings* settings,
                     float* computeMs);
int RtCopyOutput(void* session, unsigned char* output, int outputBytes);

/* Measurement hooks.  With option "stage_timing" = 1 each GPU kernel launch of RtRender is
 * bracketed by HIP events on the session's stream; RtStageTimes returns, per stage
 * (0 reset, 1 raygen, 2 traversal of a bounce's closest-hit rays together with the previous
 * bounce's shadow rays, 3 shade, 4 traversal of shadow rays alone (the last bounce's, or those
 * before the path-completion launch), 5 accumulate, 6 path completion: every remaining
 * bounce of the paths from bounce "tail_bounce" on, traversal and shading in one launch),
 * the summed milliseconds and launch counts of the last frame.  With option "trace_stats"
 * = 1 the traversal kernels also count node visits and triangle tests; RtTraceStats returns
 * {closest nodes, closest triangles, shadow nodes, shadow triangles} of the last frame. */
enum { RT_STAGE_RESET = 0, RT_STAGE_RAYGEN, RT_STAGE_CLOSEST, RT_STAGE_SHADE, RT_STAGE_SHADOW, RT_STAGE_ACCUMULATE,
       RT_STAGE_TAIL, RT_STAGE_COUNT };
int RtStageTimes(void* session, float* milliseconds, int* launches, int count);
/* With stage timing on: the milliseconds of the last frame during which at least one launch
 * of the stages in stageMask (bit k = stage k) was running, the union of their launch
 * intervals.  Launches of concurrent path pipelines overlap, so this is the figure to price
 * a stage's throughput on; RtStageTimes sums launch durations. */
int RtStageBusy(void* session, unsigned stageMask, float* milliseconds);
int RtTraceStats(void* session, unsigned long long* counts4);
/* Plain device-to-device copy of `bytes` (float4 grid-stride kernel), `iterations` times;
 * milliseconds = mean kernel time.  Calibrates HBM byte counters against a known size. */
int RtBandwidthProbe(long long bytes, int iterations, float* milliseconds);

/* Learned denoiser used by RtDenoise.  A session loads $RT_DENOISER_WEIGHTS or
 * denoiser_weights.bin next to the library on first use, unless RtSetDenoiser attached a
 * network (the session borrows it; pass null to return to the default weights).  Option
 * "denoise_fallback": 0 per-16x16-tile variance check (default), 1 always show plain
 * accumulation, 2 never fall back.  RtDenoiseStats reports the last call's tile count,
 * tiles that fell back and network milliseconds.  RtReadDenoiserInputs returns the
 * h*w*10 network input planes (radiance, albedo, normal averages, mean first-hit depth). */
int RtSetDenoiser(void* session, void* denoiser);
int RtDenoiseStats(void* session, int* tiles, int* rejected, float* inferMs);
int RtReadDenoiserInputs(void* session, float* out, long long floatCount);

/* The network: 5-level U-Net of 3x3 convolutions (BF16 MFMA, FP32 accumulation).
 * Parameters are FP32, per layer [9][coutPad][cinPad] weights then [coutPad] biases.
 * Images are [n][h][w][10] raw planes in, [n][h][w][3] radiance out; h and w multiples of 16
 * (RtDenoiserGradients also needs h == w).  Weight files carry a versioned header. */
void* RtDenoiserCreate(void);
void RtDenoiserDestroy(void* denoiser);
long long RtDenoiserParamCount(void* denoiser);
int RtDenoiserInit(void* denoiser, unsigned seed);
int RtDenoiserGetParams(void* denoiser, float* params);
int RtDenoiserSetParams(void* denoiser, const float* params);
int RtDenoiserSave(void* denoiser, const char* path);
int RtDenoiserLoad(void* denoiser, const char* path);
int RtDenoiserInfer(void* denoiser, const float* input, int n, int h, int w, float* output);
int RtDenoiserGradients(void* denoiser, const float* input, const float* target, int n, int h, int w, float* loss,
                        float* grads);
int RtDenoiserSetDataset(void* denoiser, const float* inputs, const float* targets, int count, int h, int w);
/* steps Adam steps on random crop x crop windows (flips as augmentation) of the dataset;
 * losses[steps] receives each step's L1 loss in log space. */
int RtDenoiserTrain(void* denoiser, int steps, int batch, int crop, float learningRate, unsigned seed, float* losses);
int RtDenoiserTimings(void* denoiser, float* forwardMs, float* trainStepMs);
/* Adds weight * |D(out) - D(target)| to the per-channel training loss (RtDenoiserTrain and
 * RtDenoiserGradients), D being the display transform of Native/renderer.cu:86-90 at
 * exposure 1 before 8-bit rounding; 0 (the default) trains on the log-space L1 alone. */
int RtDenoiserSetDisplayLoss(void* denoiser, float weight);
double RtDenoiserForwardFlops(int n, int h, int w);
/* One 3x3 convolution (stride 1, zero padding, NHWC, weights [9][cout][cin]) and its
 * gradients on the GPU, inputs rounded to BF16: the kernels the network is built from. */
int RtConv3x3Forward(const float* in, int n, int h, int w, int cin, const float* weight, const float* bias, int cout,
                     int relu, float* out);
int RtConv3x3Backward(const float* in, int n, int h, int w, int cin, const float* weight, const float* gradOut,
                      int cout, float* gradIn, float* gradWeight, float* gradBias);

/* Batched views: many scenes and cameras rendered together, one launch per stage over a
 * shared path queue, each frame replayed as a captured hipGraph ("use_graph", default 1).
 * Each view takes one progressive sample per frame (raysPerPixel is ignored) and equals a
 * session rendering its scene with raysPerPixel 1.  Options ("bvh_width", "bvh_quantized",
 * "blocks_per_cu", "short_stack", "while_while", "refill", "refill_below", "pipelines", "xcd_segments", "use_graph") must be set before
 * RtBatchBuild.  RtBatchReadView
 * copies a view's RGBA8 image and/or its 10-float-per-pixel accumulation. */
void* RtBatchCreate(void);
void RtBatchDestroy(void* batch);
const char* RtBatchLastError(void* batch);
int RtBatchAddView(void* batch, const Triangle* triangles, int count, int width, int height,
                   const CameraData* camera, const RenderSettings* settings);
int RtBatchSetOption(void* batch, const char* key, int value);
int RtBatchBuild(void* batch, float* buildMs);
int RtBatchGetOption(void* batch, const char* key);
/* The launch autotuner for a batch (no counterpart in the reference): BVH width x workgroups x
 * stack x loop, then pipelines, workgroups and stack x loop again, each trial re-capturing the
 * frame graph (a rebuild when the width changes); the batch is left built with the fastest,
 * samples reset, and the choice stored in the autotune cache under the batch's class (views,
 * triangles and pixels to the nearest power of two, largest bounce count), which RtBatchBuild
 * applies to options the caller did not set unless "autotune_cache" is 0. */
int RtBatchAutotune(void* batch, int frames, float* mraysPerSecond, int* trialCount);
int RtBatchRender(void* batch, int frames, float* deviceMs);
int RtBatchReset(void* batch);
int RtBatchRayCounts(void* batch, unsigned long long* closest, unsigned long long* shadow);
int RtBatchViewInfo(void* batch, int view, int* width, int* height, unsigned* samples);
int RtBatchReadView(void* batch, int view, unsigned char* rgba, int bytes, float* accumulation, long long floatCount);

#ifdef __cplusplus
}
#endif
#endif /* RT_ABI_H */
