// bvh_gpu.h - GPU binned-SAH build (bvh_build.hip).
#pragma once
#include "rt_core.h"

namespace rt {

struct GpuBuildParams {
    int bins = 32;      // bins per axis (2..32)
    int maxLeaf = 4;    // leaf at <= maxLeaf primitives (or at depth 32)
    int width = 4;      // wide layout: 2, 4 or 8 children per node
};

// Device buffers owned by the caller after a successful build (free with hipFree).
struct GpuBuildResult {
    RtBvhNode* binNodes = nullptr;   // binary tree, reference layout, root at 0
    int* order = nullptr;            // leaf order of triangle indices
    TriRecord* recs = nullptr;       // records[j] = triangle order[j]
    void* wideNodes = nullptr;       // WideNode<width>[wideCount], root at 0
    int binNodeCount = 0, depth = 0, wideCount = 0, wideDepth = 0;
};

hipError_t buildBvhGpu(const Triangle* dTris, int n, const GpuBuildParams& prm, GpuBuildResult& res, hipStream_t s);

}  // namespace rt
