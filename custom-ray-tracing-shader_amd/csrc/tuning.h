// tuning.h - the launch autotuner's search and its persistent cache (tuning.hip).
//
// A launch configuration is the BVH layout (width, leaf size) plus the traversal launch
// (persistent workgroups per CU, LDS stack entries, loop variant, path pipelines, the bounce
// from which paths complete in one launch).  The search is staged: the full grid of
// workgroups x stack x loop for every layout at two pipelines, then pipelines, workgroups,
// the path-completion bounce and stack x loop again around the best so far.  Winners persist
// in a JSON file ($RT_AUTOTUNE_CACHE, else $HOME/.cache/rt_amd/autotune.json) keyed by scene
// class, which RtInitialize, RtRender and RtBatchBuild consult for options their caller did
// not set.
#pragma once
#include <functional>
#include <string>
#include <vector>

namespace rt {

struct LaunchOptions {
    int bvhWidth = -1, buildMaxLeaf = -1, blocksPerCU = -1, shortStack = -1, whileWhile = -1, refill = -1,
        pipelines = -1, tailBounce = -1;   // -1: not part of this configuration
    double mraysPerSecond = 0;
};

// Option names as RtSetOption takes them, paired with the field of each.
struct TunedField { const char* name; int LaunchOptions::*field; };
extern const TunedField kTunedFields[8];

std::string tuneCachePath();
// Scene class of a session: triangle count to the nearest power of two, image size, bounces.
std::string sessionTuneKey(long long triangles, int width, int height, int maxBounces);
// The same without the bounce count (matches every bounce count of the class).
std::string sessionTunePrefix(long long triangles, int width, int height);
// Scene class of a batch: views, total triangles and pixels to the nearest power of two,
// largest bounce count.
std::string batchTuneKey(int views, long long triangles, long long pixels, int maxBounces);

// Entry for key (or, with prefix, the first entry whose key starts with it).
bool tuneLookup(const std::string& key, bool prefix, LaunchOptions& out);
// Adds or replaces an entry; false if the file cannot be written.
bool tuneStore(const std::string& key, const LaunchOptions& options);

struct TuneSpace {
    std::vector<int> widths{4, 8}, leaves{2, 4}, blocks{2, 3, 4, 6, 8}, stacks{8, 12, 16}, pipelines{1, 2, 3, 4},
        tails{0, 3, 4, 5};
    std::vector<std::pair<int, int>> loops{{1, 1}, {1, 0}, {0, 0}};   // (while_while, refill)
};
// Staged search; measure(cfg) returns the configuration's Mrays/s (<= 0 when it fails).
// Tails >= maxBounces are skipped; an empty tails list leaves tailBounce at -1.
LaunchOptions tuneSearch(const TuneSpace& space, int maxBounces, const std::function<double(const LaunchOptions&)>& measure,
                         int* trials);

}  // namespace rt
