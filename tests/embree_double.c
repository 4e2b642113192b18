/* embree_double.c - a brute-force test double with the slice of the Embree 3 C ABI that the
 * plugin's backend 3 calls (rtcNewDevice ... rtcOccluded1).  Built as libembree3.so.3 in a
 * temporary directory and put on LD_LIBRARY_PATH, it lets the test suite run backend 3's
 * dlopen, geometry upload, filter callback and result mapping where no Embree library is
 * installed.  Every query tests every triangle with FP32 Moller-Trumbore; the geometric
 * normal handed to the filter is edge1 x edge2, as Embree's triangle intersector reports it,
 * and the filter receives one-wide structure-of-arrays ray/hit blocks.  Among accepted hits
 * with tnear <= t <= tfar the first nearest in primitive order wins. */
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct { int* valid; void* userPtr; const void* context; float* ray; float* hit; unsigned n; } FilterArgs;
typedef void (*FilterFn)(const FilterArgs*);

typedef struct {
    float* vertices;
    unsigned* indices;
    size_t triangles;
    void* user;
    FilterFn intersectFilter, occludedFilter;
    int attached, released;
} Geometry;

typedef struct { Geometry* geometry; } Scene;
typedef struct { int unused; } Device;

typedef struct { float ox, oy, oz, tnear, dx, dy, dz, time, tfar; unsigned mask, id, flags; } Ray;
typedef struct { float ngx, ngy, ngz, u, v; unsigned primID, geomID, instID; } Hit;
typedef struct { Ray ray; Hit hit; } RayHit;

void* rtcNewDevice(const char* config) { (void)config; return calloc(1, sizeof(Device)); }
void rtcReleaseDevice(void* d) { free(d); }
int rtcGetDeviceError(void* d) { (void)d; return 0; }
void* rtcNewScene(void* d) { (void)d; return calloc(1, sizeof(Scene)); }
void rtcSetSceneFlags(void* s, int f) { (void)s; (void)f; }
void rtcSetSceneBuildQuality(void* s, int q) { (void)s; (void)q; }
void* rtcNewGeometry(void* d, int type) { (void)d; return type == 0 ? calloc(1, sizeof(Geometry)) : NULL; }

void* rtcSetNewGeometryBuffer(void* g, int type, unsigned slot, int format, size_t stride, size_t count) {
    Geometry* geo = (Geometry*)g;
    (void)slot;
    if (type == 1 && format == 0x9003 && stride == 12) {
        geo->vertices = (float*)malloc(stride * count);
        return geo->vertices;
    }
    if (type == 0 && format == 0x5003 && stride == 12) {
        geo->indices = (unsigned*)malloc(stride * count);
        geo->triangles = count;
        return geo->indices;
    }
    return NULL;
}
void rtcSetGeometryUserData(void* g, void* p) { ((Geometry*)g)->user = p; }
void rtcSetGeometryIntersectFilterFunction(void* g, FilterFn f) { ((Geometry*)g)->intersectFilter = f; }
void rtcSetGeometryOccludedFilterFunction(void* g, FilterFn f) { ((Geometry*)g)->occludedFilter = f; }
void rtcCommitGeometry(void* g) { (void)g; }
static void freeGeometry(Geometry* g) {
    free(g->vertices);
    free(g->indices);
    free(g);
}
unsigned rtcAttachGeometry(void* s, void* g) {
    ((Scene*)s)->geometry = (Geometry*)g;
    ((Geometry*)g)->attached = 1;
    return 0;
}
void rtcReleaseGeometry(void* g) {
    Geometry* geo = (Geometry*)g;
    if (geo->attached) geo->released = 1;
    else freeGeometry(geo);
}
void rtcCommitScene(void* s) { (void)s; }
void rtcReleaseScene(void* s) {
    Scene* sc = (Scene*)s;
    if (sc->geometry) freeGeometry(sc->geometry);
    free(sc);
}

/* Candidate hit of triangle i: returns 1 and t, u, v, Ng when Moller-Trumbore hits. */
static int candidate(const Geometry* g, size_t i, const Ray* r, float* t, float* u, float* v, float ng[3]) {
    const float* a = g->vertices + 3 * g->indices[3 * i];
    const float* b = g->vertices + 3 * g->indices[3 * i + 1];
    const float* c = g->vertices + 3 * g->indices[3 * i + 2];
    const float e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    const float e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    const float p[3] = {r->dy * e2[2] - r->dz * e2[1], r->dz * e2[0] - r->dx * e2[2], r->dx * e2[1] - r->dy * e2[0]};
    const float det = e1[0] * p[0] + e1[1] * p[1] + e1[2] * p[2];
    if (det == 0.0f) return 0;
    const float inv = 1.0f / det;
    const float s[3] = {r->ox - a[0], r->oy - a[1], r->oz - a[2]};
    *u = (s[0] * p[0] + s[1] * p[1] + s[2] * p[2]) * inv;
    if (*u < 0 || *u > 1) return 0;
    const float q[3] = {s[1] * e1[2] - s[2] * e1[1], s[2] * e1[0] - s[0] * e1[2], s[0] * e1[1] - s[1] * e1[0]};
    *v = (r->dx * q[0] + r->dy * q[1] + r->dz * q[2]) * inv;
    if (*v < 0 || *u + *v > 1) return 0;
    *t = (e2[0] * q[0] + e2[1] * q[1] + e2[2] * q[2]) * inv;
    ng[0] = e1[1] * e2[2] - e1[2] * e2[1];
    ng[1] = e1[2] * e2[0] - e1[0] * e2[2];
    ng[2] = e1[0] * e2[1] - e1[1] * e2[0];
    return *t >= r->tnear && *t <= r->tfar;
}

static int accepted(const Geometry* g, FilterFn f, const Ray* r, float t, float u, float v, const float ng[3], size_t i) {
    if (!f) return 1;
    float rayBlock[12] = {r->ox, r->oy, r->oz, r->tnear, r->dx, r->dy, r->dz, r->time, t, 0, 0, 0};
    float hitBlock[8];
    unsigned prim = (unsigned)i, geomID = 0;
    hitBlock[0] = ng[0]; hitBlock[1] = ng[1]; hitBlock[2] = ng[2]; hitBlock[3] = u; hitBlock[4] = v;
    memcpy(&hitBlock[5], &prim, 4);
    memcpy(&hitBlock[6], &geomID, 4);
    memcpy(&hitBlock[7], &geomID, 4);
    int valid = -1;
    FilterArgs args = {&valid, g->user, NULL, rayBlock, hitBlock, 1};
    f(&args);
    return valid != 0;
}

void rtcIntersect1(void* s, void* ctx, RayHit* rh) {
    const Geometry* g = ((Scene*)s)->geometry;
    (void)ctx;
    for (size_t i = 0; g && i < g->triangles; i++) {
        float t, u, v, ng[3];
        if (!candidate(g, i, &rh->ray, &t, &u, &v, ng)) continue;
        if (rh->hit.geomID != 0xFFFFFFFFu && t >= rh->ray.tfar) continue;
        if (!accepted(g, g->intersectFilter, &rh->ray, t, u, v, ng, i)) continue;
        rh->ray.tfar = t;
        rh->hit.u = u; rh->hit.v = v;
        rh->hit.ngx = ng[0]; rh->hit.ngy = ng[1]; rh->hit.ngz = ng[2];
        rh->hit.primID = (unsigned)i;
        rh->hit.geomID = 0;
    }
}

void rtcOccluded1(void* s, void* ctx, Ray* r) {
    const Geometry* g = ((Scene*)s)->geometry;
    (void)ctx;
    for (size_t i = 0; g && i < g->triangles; i++) {
        float t, u, v, ng[3];
        if (candidate(g, i, r, &t, &u, &v, ng) && accepted(g, g->occludedFilter, r, t, u, v, ng, i)) {
            r->tfar = -INFINITY;
            return;
        }
    }
}
