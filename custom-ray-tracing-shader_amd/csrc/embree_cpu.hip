// embree_cpu.hip - backend 3 (host code only): Embree 3 loaded with dlopen.
#include "embree_cpu.h"

#include <dlfcn.h>
#include <chrono>
#include <cmath>

namespace rt {

namespace {

// ---- the Embree 3 API subset this backend calls (rtcore_*.h of Embree 3.x)
typedef void* EmDevice;
typedef void* EmScene;
typedef void* EmGeometry;
constexpr int kGeometryTriangle = 0;     // RTC_GEOMETRY_TYPE_TRIANGLE
constexpr int kBufferIndex = 0;          // RTC_BUFFER_TYPE_INDEX
constexpr int kBufferVertex = 1;         // RTC_BUFFER_TYPE_VERTEX
constexpr int kFormatUInt3 = 0x5003;     // RTC_FORMAT_UINT3
constexpr int kFormatFloat3 = 0x9003;    // RTC_FORMAT_FLOAT3
constexpr int kSceneFlagRobust = 1 << 2; // RTC_SCENE_FLAG_ROBUST
constexpr int kBuildQualityHigh = 2;     // RTC_BUILD_QUALITY_HIGH
constexpr int kErrorNone = 0;            // RTC_ERROR_NONE
constexpr unsigned kInvalidId = 0xFFFFFFFFu;   // RTC_INVALID_GEOMETRY_ID

struct alignas(16) EmRay {
    float orgX, orgY, orgZ, tnear;
    float dirX, dirY, dirZ, time;
    float tfar;
    unsigned mask, id, flags;
};
struct alignas(16) EmHit {
    float ngX, ngY, ngZ, u, v;
    unsigned primID, geomID, instID[1];
};
struct alignas(16) EmRayHit { EmRay ray; EmHit hit; };
static_assert(sizeof(EmRay) == 48 && sizeof(EmHit) == 32 && sizeof(EmRayHit) == 80, "Embree ray layouts");

struct EmFilterArgs;
typedef void (*EmFilterFn)(const EmFilterArgs*);
struct EmContext { int flags; EmFilterFn filter; unsigned instID[1]; };
// Filter arguments: ray and hit are N-wide structure-of-arrays blocks.
struct EmFilterArgs { int* valid; void* userPtr; const EmContext* context; float* ray; float* hit; unsigned n; };

struct EmbreeApi {
    EmDevice (*newDevice)(const char*) = nullptr;
    void (*releaseDevice)(EmDevice) = nullptr;
    int (*getDeviceError)(EmDevice) = nullptr;
    EmScene (*newScene)(EmDevice) = nullptr;
    void (*setSceneFlags)(EmScene, int) = nullptr;
    void (*setSceneBuildQuality)(EmScene, int) = nullptr;
    EmGeometry (*newGeometry)(EmDevice, int) = nullptr;
    void* (*setNewGeometryBuffer)(EmGeometry, int, unsigned, int, size_t, size_t) = nullptr;
    void (*setGeometryUserData)(EmGeometry, void*) = nullptr;
    void (*setIntersectFilter)(EmGeometry, EmFilterFn) = nullptr;
    void (*setOccludedFilter)(EmGeometry, EmFilterFn) = nullptr;
    void (*commitGeometry)(EmGeometry) = nullptr;
    unsigned (*attachGeometry)(EmScene, EmGeometry) = nullptr;
    void (*releaseGeometry)(EmGeometry) = nullptr;
    void (*commitScene)(EmScene) = nullptr;
    void (*releaseScene)(EmScene) = nullptr;
    void (*intersect1)(EmScene, EmContext*, EmRayHit*) = nullptr;
    void (*occluded1)(EmScene, EmContext*, EmRay*) = nullptr;
    std::string error;

    EmbreeApi() {
        // Resident for the process lifetime; deep binding keeps another Embree in the
        // process (a host application's own copy) from satisfying these symbols.
        void* lib = dlopen("libembree3.so.3", RTLD_NOW | RTLD_LOCAL | RTLD_DEEPBIND);
        if (!lib) {
            const char* e = dlerror();
            error = std::string("Embree backend unavailable: ") + (e ? e : "libembree3.so.3 not found");
            return;
        }
        bind(lib, newDevice, "rtcNewDevice");
        bind(lib, releaseDevice, "rtcReleaseDevice");
        bind(lib, getDeviceError, "rtcGetDeviceError");
        bind(lib, newScene, "rtcNewScene");
        bind(lib, setSceneFlags, "rtcSetSceneFlags");
        bind(lib, setSceneBuildQuality, "rtcSetSceneBuildQuality");
        bind(lib, newGeometry, "rtcNewGeometry");
        bind(lib, setNewGeometryBuffer, "rtcSetNewGeometryBuffer");
        bind(lib, setGeometryUserData, "rtcSetGeometryUserData");
        bind(lib, setIntersectFilter, "rtcSetGeometryIntersectFilterFunction");
        bind(lib, setOccludedFilter, "rtcSetGeometryOccludedFilterFunction");
        bind(lib, commitGeometry, "rtcCommitGeometry");
        bind(lib, attachGeometry, "rtcAttachGeometry");
        bind(lib, releaseGeometry, "rtcReleaseGeometry");
        bind(lib, commitScene, "rtcCommitScene");
        bind(lib, releaseScene, "rtcReleaseScene");
        bind(lib, intersect1, "rtcIntersect1");
        bind(lib, occluded1, "rtcOccluded1");
    }
    template <class F> void bind(void* lib, F& fn, const char* name) {
        if (!error.empty()) return;
        fn = reinterpret_cast<F>(dlsym(lib, name));
        if (!fn) error = std::string("Embree backend unavailable: missing symbol ") + name;
    }
    static EmbreeApi& get() {
        static EmbreeApi api;
        return api;
    }
};

// Native/embree.cuh:45-55: the reference's determinant test on Embree's geometric normal
// (det = -(dir . Ng)); rejects near-parallel hits and back faces of one-sided triangles.
void acceptanceFilter(const EmFilterArgs* a) {
    const Triangle* tris = static_cast<const Triangle*>(a->userPtr);
    const unsigned n = a->n;
    const unsigned* primID = reinterpret_cast<const unsigned*>(a->hit) + 5 * n;
    for (unsigned i = 0; i < n; i++) {
        if (!a->valid[i]) continue;
        const Triangle& t = tris[primID[i]];
        const float det = -(a->ray[4 * n + i] * a->hit[0 * n + i] + a->ray[5 * n + i] * a->hit[1 * n + i] +
                            a->ray[6 * n + i] * a->hit[2 * n + i]);
        if (fabsf(det) < 1e-7f || (t.oneSided && det < 0)) a->valid[i] = 0;
    }
}

void initContext(EmContext& c) {
    c.flags = 0;   // RTC_INTERSECT_CONTEXT_FLAG_INCOHERENT
    c.filter = nullptr;
    c.instID[0] = kInvalidId;
}

}  // namespace

const std::string& embreeLoadError() { return EmbreeApi::get().error; }

void EmbreeScene::release() {
    if (scene || device) {   // only a built scene needs the library (and loads it)
        EmbreeApi& api = EmbreeApi::get();
        if (scene) api.releaseScene(scene);
        if (device) api.releaseDevice(device);
    }
    scene = device = nullptr;
    buildMs = 0;
}

bool EmbreeScene::build(const std::vector<Triangle>& tris, std::string& error) {
    EmbreeApi& api = EmbreeApi::get();
    if (!api.error.empty()) { error = api.error; return false; }
    release();
    const auto t0 = std::chrono::steady_clock::now();
    device = api.newDevice(nullptr);
    if (!device) { error = "rtcNewDevice returned no device"; return false; }
    scene = api.newScene(device);
    api.setSceneFlags(scene, kSceneFlagRobust);
    api.setSceneBuildQuality(scene, kBuildQualityHigh);
    EmGeometry geom = api.newGeometry(device, kGeometryTriangle);
    float* v = static_cast<float*>(api.setNewGeometryBuffer(geom, kBufferVertex, 0, kFormatFloat3, 12, tris.size() * 3));
    unsigned* idx = static_cast<unsigned*>(api.setNewGeometryBuffer(geom, kBufferIndex, 0, kFormatUInt3, 12, tris.size()));
    if (!v || !idx) {
        api.releaseGeometry(geom);
        error = "rtcNewGeometry returned no triangle geometry";
        return false;
    }
    // Vertices origin, origin + edge1, origin + edge2 (Native/embree.cuh:76-79).
    for (size_t i = 0; i < tris.size(); i++) {
        const Triangle& t = tris[i];
        const float p[9] = {t.origin.x, t.origin.y, t.origin.z,
                            t.origin.x + t.edge1.x, t.origin.y + t.edge1.y, t.origin.z + t.edge1.z,
                            t.origin.x + t.edge2.x, t.origin.y + t.edge2.y, t.origin.z + t.edge2.z};
        for (int k = 0; k < 9; k++) v[i * 9 + k] = p[k];
        for (int k = 0; k < 3; k++) idx[i * 3 + k] = (unsigned)(i * 3 + k);
    }
    api.setGeometryUserData(geom, const_cast<Triangle*>(tris.data()));
    api.setIntersectFilter(geom, acceptanceFilter);
    api.setOccludedFilter(geom, acceptanceFilter);
    api.commitGeometry(geom);
    api.attachGeometry(scene, geom);
    api.releaseGeometry(geom);
    api.commitScene(scene);
    if (api.getDeviceError(device) != kErrorNone) {
        error = "rtcCommitScene reported an error while building the scene";
        return false;
    }
    buildMs = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return true;
}

// Native/embree.cuh:92-112.
RtHit EmbreeScene::query(f3 o, f3 d, float limit, bool any) const {
    EmbreeApi& api = EmbreeApi::get();
    EmRayHit q{};
    q.ray.orgX = o.x; q.ray.orgY = o.y; q.ray.orgZ = o.z;
    q.ray.dirX = d.x; q.ray.dirY = d.y; q.ray.dirZ = d.z;
    q.ray.tnear = nextafterf(1e-4f, INFINITY);
    q.ray.tfar = nextafterf(limit, 0);
    q.ray.mask = ~0u;
    q.hit.geomID = kInvalidId;
    EmContext ctx;
    initContext(ctx);
    RtHit h;
    h.index = -1; h.distance = limit; h.u = 0; h.v = 0;
    if (any) {
        api.occluded1(scene, &ctx, &q.ray);
        if (q.ray.tfar < 0) h.index = 0;
    } else {
        api.intersect1(scene, &ctx, &q);
        if (q.hit.geomID != kInvalidId) {
            h.index = (int)q.hit.primID; h.distance = q.ray.tfar; h.u = q.hit.u; h.v = q.hit.v;
        }
    }
    return h;
}

}  // namespace rt
