/*
 * bvh_build.cpp — host construction of the flattened BVH (layout in
 * rt_layout.h).
 *
 * The reference builds a pointer tree with one primitive per leaf by
 * sorting the whole list at every level (BVHNode::create, src/bvh.cpp:67-94:
 * longest axis of the list's box, std::sort by box centre, split at n/2).
 * This builder keeps the median split but:
 *   - partitions with nth_element (O(n) per level instead of a full sort),
 *   - uses the longest axis of the box-centre bounds,
 *   - stops at leaves of at most 4 primitives of one kind,
 *   - runs the top levels on several threads; the result does not depend
 *     on the thread schedule (each range is partitioned by one thread and
 *     the layout pass below is sequential),
 *   - writes nodes depth-first with both child boxes in the parent.
 * Closest-hit results do not depend on the tree shape except for exact
 * ties in t.
 */
#include "bvh_build.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <thread>

namespace rtb {

namespace {

struct Box3 {
    float lo[3] = {INFINITY, INFINITY, INFINITY};
    float hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    void grow(const float* l, const float* h)
    {
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], l[k]);
            hi[k] = std::max(hi[k], h[k]);
        }
    }
};

bool is_leaf(const std::vector<PrimRef>& refs, size_t b, size_t e)
{
    if (e - b > RT_LEAF_MAX) return false;
    for (size_t i = b + 1; i < e; ++i)
        if (refs[i].kind != refs[b].kind) return false;
    return true;
}

void partition(std::vector<PrimRef>& refs, size_t b, size_t e, int level)
{
    if (e - b <= 1 || is_leaf(refs, b, e)) return;
    float clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (size_t i = b; i < e; ++i)
        for (int k = 0; k < 3; ++k) {
            clo[k] = std::min(clo[k], refs[i].c[k]);
            chi[k] = std::max(chi[k], refs[i].c[k]);
        }
    float dx = chi[0] - clo[0], dy = chi[1] - clo[1], dz = chi[2] - clo[2];
    int axis = dx > dy && dx > dz ? 0 : (dy > dz ? 1 : 2);
    size_t mid = b + (e - b) / 2;
    std::nth_element(refs.begin() + b, refs.begin() + mid, refs.begin() + e,
                     [axis](const PrimRef& a, const PrimRef& c) {
                         if (a.c[axis] != c.c[axis]) return a.c[axis] < c.c[axis];
                         return a.g < c.g;
                     });
    if (level < 4 && e - b > 65536) {
        std::thread t([&refs, b, mid, level] { partition(refs, b, mid, level + 1); });
        partition(refs, mid, e, level + 1);
        t.join();
    } else {
        partition(refs, b, mid, level + 1);
        partition(refs, mid, e, level + 1);
    }
}

struct Layout {
    const std::vector<PrimRef>& refs;
    std::vector<float>& nodes;
    int max_depth = 0;

    // Returns the child word for range [b, e) and its box.
    int32_t emit(size_t b, size_t e, int depth, Box3& box)
    {
        max_depth = std::max(max_depth, depth);
        if (is_leaf(refs, b, e)) {
            for (size_t i = b; i < e; ++i) box.grow(refs[i].lo, refs[i].hi);
            uint32_t x = ((uint32_t)b << 5) | ((uint32_t)refs[b].kind << 3) | (uint32_t)(e - b);
            return (int32_t)~x;
        }
        size_t idx = nodes.size() / 16;
        nodes.resize(nodes.size() + 16);
        size_t mid = b + (e - b) / 2;
        Box3 l, r;
        int32_t cl = emit(b, mid, depth + 1, l);
        int32_t cr = emit(mid, e, depth + 1, r);
        write(idx, l, r, cl, cr);
        box.grow(l.lo, l.hi);
        box.grow(r.lo, r.hi);
        return (int32_t)idx;
    }
    void write(size_t idx, const Box3& l, const Box3& r, int32_t cl, int32_t cr)
    {
        float* n = &nodes[idx * 16];
        n[0] = l.lo[0]; n[1] = l.hi[0]; n[2] = l.lo[1]; n[3] = l.hi[1];
        n[4] = l.lo[2]; n[5] = l.hi[2]; n[6] = r.lo[0]; n[7] = r.hi[0];
        n[8] = r.lo[1]; n[9] = r.hi[1]; n[10] = r.lo[2]; n[11] = r.hi[2];
        std::memcpy(&n[12], &cl, 4);
        std::memcpy(&n[13], &cr, 4);
        int32_t z = 0;
        std::memcpy(&n[14], &z, 4);
        std::memcpy(&n[15], &z, 4);
    }
};

}  // namespace

int build(std::vector<PrimRef>& refs, BuildResult& out)
{
    out.nodes.clear();
    out.depth = 0;
    if (refs.size() >= (size_t(1) << 26)) return -1;
    for (auto& r : refs)
        for (int k = 0; k < 3; ++k) r.c[k] = 0.5f * (r.lo[k] + r.hi[k]);
    if (refs.empty()) return 0;
    partition(refs, 0, refs.size(), 0);
    Layout lay{refs, out.nodes};
    if (is_leaf(refs, 0, refs.size())) {
        // A root that is a leaf still gets an interior node: child 1 is an
        // empty leaf whose box is empty (never hit).
        out.nodes.resize(16);
        Box3 l, r;
        int32_t cl = lay.emit(0, refs.size(), 1, l);
        int32_t cr = (int32_t)~(uint32_t)0;
        lay.write(0, l, r, cl, cr);
        out.depth = 1;
        return 0;
    }
    Box3 root;
    lay.emit(0, refs.size(), 0, root);
    out.depth = lay.max_depth;
    if (out.depth >= RT_STACK_DEPTH) return -2;
    return 0;
}

}  // namespace rtb
