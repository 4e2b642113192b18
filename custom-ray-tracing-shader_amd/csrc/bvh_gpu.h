// bvh_gpu.h - GPU binned-SAH build (bvh_build.hip).
#pragma once
#include "rt_core.h"

namespace rt {

struct GpuBuildParams {
    int bins = 16;      // bins per axis (2..32)
    int maxLeaf = 4;    // leaf at <= maxLeaf primitives (or at depth 32)
    int width = 4;      // wide layout: 2, 4 or 8 children per node
    int quantized = 0;  // 1: QNode<width> (8-bit child bounds), 0: WideNode<width>
};

// Device buffers owned by the caller after a successful build (free with hipFree).
struct GpuBuildResult {
    RtBvhNode* binNodes = nullptr;   // binary tree, reference layout, root at 0
    int* order = nullptr;            // leaf order of triangle indices
    TriRecord* recs = nullptr;       // records[j] = triangle order[j], plus kRecordPad zeroed records
    void* wideNodes = nullptr;       // WideNode<width>[] or QNode<width>[wideCount], root at 0
    int binNodeCount = 0, depth = 0, wideCount = 0, wideDepth = 0;
};

// Zeroed records after the last one: leaf batches of the traversal may read (never test)
// records past the end of a large final leaf.
constexpr int kRecordPad = 8;

// Bytes per wide node of the given layout.
inline size_t wideNodeBytes(int width, int quantized) {
    if (quantized) return width == 8 ? sizeof(QNode<8>) : width == 2 ? sizeof(QNode<2>) : sizeof(QNode<4>);
    return (size_t)32 * width;
}

hipError_t buildBvhGpu(const Triangle* dTris, int n, const GpuBuildParams& prm, GpuBuildResult& res, hipStream_t s);

}  // namespace rt
