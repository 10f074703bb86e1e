/*
 * bvh_build.h — host BVH builder interface (see bvh_build.cpp).
 */
#ifndef RT_BVH_BUILD_H
#define RT_BVH_BUILD_H

#include <cstdint>
#include <vector>

#include "rt_layout.h"

namespace rtb {

// One primitive to place: its box, kind (RT_KIND_*) and shading index g.
// build() reorders the array into leaf order; the primitive records are
// then written in that order by the caller.
struct PrimRef {
    float lo[3], hi[3];
    float c[3];
    int32_t kind;
    int32_t g;
};

struct BuildResult {
    std::vector<float> nodes;  // RT_NODE_FLOATS floats (128 B) per 4-wide node
    int depth = 0;             // levels of 4-wide nodes
    int stack_need = 0;        // most traversal-stack entries a ray can hold at once
};

// Split rule of the interior nodes: median of the box centres along the
// longest axis, or the surface-area heuristic (binned) down to binary level
// sah_levels and the median below it (the median keeps subtrees balanced,
// which bounds the depth and so the traversal stack).
enum { BUILD_MEDIAN = 0, BUILD_SAH = 1 };

// Returns 0 on success, -1 if there are too many primitives for the leaf
// encoding, -2 if a traversal could need more stack than RT_STACK_DEPTH.
int build(std::vector<PrimRef>& refs, BuildResult& out, int method = BUILD_MEDIAN, int sah_levels = 64);

// The SAH tree with the most SAH levels (64, 32, 24, 16, 8) whose traversal
// fits the stack, else the median tree.  Same return values as build().
int build_best(std::vector<PrimRef>& refs, BuildResult& out);

}  // namespace rtb

#endif
