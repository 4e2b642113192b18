/* unity_harness.c - drives libCudaRayTracer.so the way the Unity plugin does.
 *
 * The Unity side binds the 13 exports with [DllImport("CudaRayTracer")], Cdecl, and passes
 * sequential pack-4 structs (UnityRayTracer/Assets/RayTracing/RayTraceSession.cs:10-63).  This
 * program declares the same structs independently of the library's own header, resolves
 * every export by name with dlsym, and replays RayTraceSession's call order
 * (RayTraceSession.cs:135-238): UploadScene (RtCreate, RtInitialize, RtBvhInfo), Calculate
 * (RtRender, RtAccumulationVersion, RtCpuThreadCount), Render with denoise (RtDenoise or
 * RtLastError), and Dispose (RtDestroy).  It prints one JSON object of what it observed.
 *
 *   unity_harness LIB bind
 *       resolve all exports and check the null-session contract (no device work)
 *   unity_harness LIB session TRIS.bin COUNT WIDTH HEIGHT OUT.rgba
 *       full call sequence on a scene of COUNT packed triangles; writes the first GPU frame
 *       followed by the CameraData and Settings bytes it rendered with
 */
#include <dlfcn.h>
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#pragma pack(push, 4)
typedef struct { float x, y, z; } Vector3;
typedef struct {
    Vector3 origin, edge1, edge2, color;
    float emission;
    int oneSided;
    Vector3 normal0, normal1, normal2;
    int mesh;
    Vector3 specularColor;
    float smoothness, specularProbability;
} Triangle;
typedef struct {
    Vector3 origin, right, up, forward;
    float tanHalfFov, aspect, orthographicSize;
    int orthographic;
} CameraData;
typedef struct {
    int maxBounces;
    float exposure, aperture, focusDistance, skyStrength;
    unsigned seed;
    int raysPerPixel;
} Settings;
#pragma pack(pop)

_Static_assert(sizeof(Triangle) == 116, "Triangle is 116 bytes");
_Static_assert(sizeof(CameraData) == 64, "CameraData is 64 bytes");
_Static_assert(sizeof(Settings) == 28, "Settings is 28 bytes");
_Static_assert(offsetof(Triangle, color) == 36 && offsetof(Triangle, emission) == 48 &&
               offsetof(Triangle, normal0) == 56 && offsetof(Triangle, mesh) == 92 &&
               offsetof(Triangle, specularColor) == 96 && offsetof(Triangle, specularProbability) == 112,
               "Triangle field offsets");
_Static_assert(offsetof(CameraData, tanHalfFov) == 48 && offsetof(CameraData, orthographic) == 60,
               "CameraData field offsets");
_Static_assert(offsetof(Settings, seed) == 20 && offsetof(Settings, raysPerPixel) == 24, "Settings field offsets");

/* The Cdecl signatures RayTraceSession.cs:36-63 declares. */
typedef void* (*PFN_RtCreate)(void);
typedef void (*PFN_RtSession)(void*);
typedef unsigned (*PFN_RtUnsigned)(void*);
typedef float (*PFN_RtFloat)(void*);
typedef int (*PFN_RtInt)(void*);
typedef void (*PFN_RtBvhInfo)(void*, int*, int*, float*);
typedef const char* (*PFN_RtLastError)(void*);
typedef int (*PFN_RtInitialize)(void*, const Triangle*, int, int, int);
typedef int (*PFN_RtRender)(void*, const CameraData*, const Settings*, unsigned char*, int, int, float*);
typedef int (*PFN_RtDenoise)(void*, unsigned char*, int, float*);

static struct {
    PFN_RtCreate RtCreate;
    PFN_RtSession RtDestroy, RtShutdown, RtReset;
    PFN_RtUnsigned RtAccumulationVersion, RtSampleCount;
    PFN_RtFloat RtEmbreeBuildMilliseconds;
    PFN_RtInt RtCpuThreadCount;
    PFN_RtBvhInfo RtBvhInfo;
    PFN_RtLastError RtLastError;
    PFN_RtInitialize RtInitialize;
    PFN_RtRender RtRender;
    PFN_RtDenoise RtDenoise;
} api;

static int failures = 0;
#define EXPECT(cond)                                                                  \
    do {                                                                              \
        if (!(cond)) {                                                                \
            fprintf(stderr, "unity_harness: expectation failed: %s (line %d)\n", #cond, __LINE__); \
            failures++;                                                               \
        }                                                                             \
    } while (0)

static int bind(const char* path) {
    void* h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
    if (!h) {
        fprintf(stderr, "dlopen: %s\n", dlerror());
        return 0;
    }
#define BIND(name, type)                                      \
    api.name = (type)dlsym(h, #name);                         \
    if (!api.name) {                                          \
        fprintf(stderr, "missing export %s\n", #name);        \
        return 0;                                             \
    }
    BIND(RtCreate, PFN_RtCreate)
    BIND(RtDestroy, PFN_RtSession)
    BIND(RtShutdown, PFN_RtSession)
    BIND(RtReset, PFN_RtSession)
    BIND(RtAccumulationVersion, PFN_RtUnsigned)
    BIND(RtSampleCount, PFN_RtUnsigned)
    BIND(RtEmbreeBuildMilliseconds, PFN_RtFloat)
    BIND(RtCpuThreadCount, PFN_RtInt)
    BIND(RtBvhInfo, PFN_RtBvhInfo)
    BIND(RtLastError, PFN_RtLastError)
    BIND(RtInitialize, PFN_RtInitialize)
    BIND(RtRender, PFN_RtRender)
    BIND(RtDenoise, PFN_RtDenoise)
#undef BIND
    return 1;
}

/* Native/renderer.cu:310-318: every export accepts a null session. */
static void null_session_contract(void) {
    int nodes = -1, depth = -1;
    float ms = -1;
    unsigned char px[4];
    Triangle t;
    CameraData cam;
    Settings st;
    memset(&t, 0, sizeof t);
    memset(&cam, 0, sizeof cam);
    memset(&st, 0, sizeof st);
    const char* e = api.RtLastError(NULL);
    EXPECT(e != NULL && e[0] != 0);
    EXPECT(api.RtSampleCount(NULL) == 0);
    EXPECT(api.RtAccumulationVersion(NULL) == 0);
    EXPECT(api.RtCpuThreadCount(NULL) == 0);
    EXPECT(api.RtEmbreeBuildMilliseconds(NULL) == 0.0f);
    api.RtBvhInfo(NULL, &nodes, &depth, &ms);
    EXPECT(nodes == 0 && depth == 0 && ms == 0.0f);
    EXPECT(api.RtInitialize(NULL, &t, 1, 1, 1) == 0);
    EXPECT(api.RtRender(NULL, &cam, &st, px, 4, 2, &ms) == 0);
    EXPECT(api.RtDenoise(NULL, px, 4, &ms) == 0);
    api.RtReset(NULL);
    api.RtShutdown(NULL);
    api.RtDestroy(NULL);
}

/* RayTraceSession.GetCameraData (RayTraceSession.cs:156-167) for a Unity camera at eye
 * looking along +z with the given vertical field of view in degrees. */
static CameraData unity_camera(Vector3 eye, float fovDegrees, int width, int height) {
    CameraData c;
    memset(&c, 0, sizeof c);
    c.origin = eye;
    c.right = (Vector3){1, 0, 0};
    c.up = (Vector3){0, 1, 0};
    c.forward = (Vector3){0, 0, 1};
    c.tanHalfFov = tanf(fovDegrees * (3.14159265358979f / 180.0f) * .5f);
    c.aspect = (float)width / height;
    c.orthographic = 0;
    c.orthographicSize = 5;
    return c;
}

static unsigned long long fnv1a(const unsigned char* p, size_t n) {
    unsigned long long h = 1469598103934665603ULL;
    for (size_t i = 0; i < n; i++) h = (h ^ p[i]) * 1099511628211ULL;
    return h;
}

static int session(int argc, char** argv) {
    if (argc < 8) return 2;
    int count = atoi(argv[4]), width = atoi(argv[5]), height = atoi(argv[6]);
    Triangle* tris = (Triangle*)malloc(sizeof(Triangle) * (size_t)count);
    FILE* f = fopen(argv[3], "rb");
    if (!f || !tris || fread(tris, sizeof(Triangle), (size_t)count, f) != (size_t)count) {
        fprintf(stderr, "cannot read %d triangles from %s\n", count, argv[3]);
        return 2;
    }
    fclose(f);
    size_t bytes = (size_t)width * height * 4;
    unsigned char* pixels = (unsigned char*)malloc(bytes);
    unsigned char* first = (unsigned char*)malloc(bytes);
    Settings st = {4, 1.0f, 0.0f, 1.0f, 0.0f, 42u, 2};
    CameraData cam = unity_camera((Vector3){0, 1, -3.4f}, 40, width, height);
    float ms = 0;

    /* Calculate before UploadScene: the library refuses a session that has no scene. */
    void* native = api.RtCreate();
    EXPECT(native != NULL);
    EXPECT(api.RtRender(native, &cam, &st, pixels, (int)bytes, 2, &ms) == 0);
    const char* early = api.RtLastError(native);
    EXPECT(early != NULL && early[0] != 0);

    /* UploadScene (RayTraceSession.cs:135-155). */
    EXPECT(api.RtInitialize(native, tris, count, width, height) == 1);
    int nodes = 0, depth = 0;
    float buildMs = -1;
    api.RtBvhInfo(native, &nodes, &depth, &buildMs);
    EXPECT(nodes > 0 && depth > 0 && buildMs >= 0);
    EXPECT(api.RtInitialize(native, tris, count, 4097, height) == 0);  /* size outside 1..4096 */
    EXPECT(api.RtInitialize(native, tris, count, width, height) == 1);

    /* Calculate x4 on the GPU backend (RayTraceSession.cs:185-199). */
    unsigned version = api.RtAccumulationVersion(native);
    int gpuFrames = 4, okSamples = 1;
    for (int k = 0; k < gpuFrames; k++) {
        EXPECT(api.RtRender(native, &cam, &st, pixels, (int)bytes, 2, &ms) == 1);
        if (k == 0) memcpy(first, pixels, bytes);
        EXPECT(ms > 0);
        if (api.RtSampleCount(native) != (unsigned)((k + 1) * st.raysPerPixel)) okSamples = 0;
    }
    EXPECT(okSamples);
    unsigned versionAfterFrames = api.RtAccumulationVersion(native);
    EXPECT(api.RtCpuThreadCount(native) == 0);
    int alphaOk = 1;
    for (size_t i = 3; i < bytes; i += 4) alphaOk &= pixels[i] == 255;
    EXPECT(alphaOk);

    /* Render(denoise: true) (RayTraceSession.cs:201-226): samples and version unchanged. */
    unsigned samplesBefore = api.RtSampleCount(native);
    float denoiseMs = 0;
    int denoised = api.RtDenoise(native, pixels, (int)bytes, &denoiseMs);
    const char* denoiseError = denoised ? "" : api.RtLastError(native);
    EXPECT(api.RtSampleCount(native) == samplesBefore);
    EXPECT(api.RtAccumulationVersion(native) == versionAfterFrames);
    EXPECT(api.RtDenoise(native, pixels, (int)bytes - 1, &denoiseMs) == 0);  /* undersized output */

    /* A moved camera restarts accumulation: the version changes, samples restart. */
    CameraData moved = cam;
    moved.origin.x += 0.05f;
    EXPECT(api.RtRender(native, &moved, &st, pixels, (int)bytes, 2, &ms) == 1);
    unsigned versionMoved = api.RtAccumulationVersion(native);
    EXPECT(versionMoved != versionAfterFrames);
    EXPECT(api.RtSampleCount(native) == (unsigned)st.raysPerPixel);

    /* Backend switches: CPU multi, CPU single (RtCpuThreadCount 1 / N). */
    EXPECT(api.RtRender(native, &moved, &st, pixels, (int)bytes, 1, &ms) == 1);
    int multiThreads = api.RtCpuThreadCount(native);
    EXPECT(multiThreads >= 1);
    EXPECT(api.RtAccumulationVersion(native) != versionMoved);
    EXPECT(api.RtRender(native, &moved, &st, pixels, (int)bytes, 0, &ms) == 1);
    EXPECT(api.RtCpuThreadCount(native) == 1);
    /* Backend 3 either renders or fails with a message. */
    int embree = api.RtRender(native, &moved, &st, pixels, (int)bytes, 3, &ms);
    const char* embreeError = embree ? "" : api.RtLastError(native);
    EXPECT(embree || (embreeError && embreeError[0]));
    /* Invalid settings are refused. */
    Settings bad = st;
    bad.raysPerPixel = 17;
    EXPECT(api.RtRender(native, &moved, &bad, pixels, (int)bytes, 2, &ms) == 0);

    /* Dispose. */
    api.RtDestroy(native);

    FILE* out = fopen(argv[7], "wb");
    if (!out || fwrite(first, 1, bytes, out) != bytes || fwrite(&cam, sizeof cam, 1, out) != 1 ||
        fwrite(&st, sizeof st, 1, out) != 1) {
        fprintf(stderr, "cannot write %s\n", argv[7]);
        return 2;
    }
    fclose(out);
    printf("{\"failures\": %d, \"nodes\": %d, \"depth\": %d, \"build_ms\": %.3f, \"version_initial\": %u, "
           "\"version_after_frames\": %u, \"version_moved\": %u, \"samples_after_frames\": %u, "
           "\"denoised\": %d, \"denoise_ms\": %.3f, \"denoise_error\": \"%s\", \"cpu_multi_threads\": %d, "
           "\"embree\": %d, \"first_frame_fnv\": \"%016llx\"}\n",
           failures, nodes, depth, buildMs, version, versionAfterFrames, versionMoved, samplesBefore, denoised,
           denoiseMs, denoiseError, multiThreads, embree, fnv1a(first, bytes));
    free(tris);
    free(pixels);
    free(first);
    return failures ? 1 : 0;
}

int main(int argc, char** argv) {
    if (argc < 3) {
        fprintf(stderr, "usage: %s LIB bind | LIB session TRIS.bin COUNT WIDTH HEIGHT OUT.rgba\n", argv[0]);
        return 2;
    }
    if (!bind(argv[1])) return 2;
    if (strcmp(argv[2], "bind") == 0) {
        null_session_contract();
        printf("{\"failures\": %d, \"exports\": 13}\n", failures);
        return failures ? 1 : 0;
    }
    if (strcmp(argv[2], "session") == 0) return session(argc, argv);
    return 2;
}
