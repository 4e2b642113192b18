// embree_cpu.h - backend 3: Intel Embree 3 as the host BVH, loaded at run time.
//
// Native/embree.cuh:8-31 resolves libembree3.so.3 with dlopen (deep binding, so symbols of
// another Embree already in the process are not picked up) and keeps the library resident;
// each session builds and releases its own device and scene.  The integrator is unchanged:
// only the closest-hit and any-hit queries go to Embree, with the reference's acceptance
// rules applied in a filter (|det| >= 1e-7, one-sided culling) and t limited to
// (1e-4, limit) exclusive (Native/embree.cuh:45-55,92-112).
//
// Embree's headers are not part of this build: the few types, constants and functions used
// are declared here from the Embree 3 API documentation.  When the library cannot be loaded,
// backend 3 reports itself unavailable through RtLastError and returns 0.
#pragma once
#include <string>
#include <vector>
#include "rt_core.h"

namespace rt {

struct EmbreeScene {
    void* device = nullptr;
    void* scene = nullptr;
    float buildMs = 0;
    ~EmbreeScene() { release(); }
    void release();
    // Builds the scene from the session's triangles; false with error set on failure.
    bool build(const std::vector<Triangle>& tris, std::string& error);
    // Closest hit (any == false) or occlusion (any == true) for a ray limited to limit.
    RtHit query(f3 o, f3 d, float limit, bool any) const;
};

// Empty when libembree3.so.3 and its symbols resolved, else the reason.
const std::string& embreeLoadError();

}  // namespace rt
