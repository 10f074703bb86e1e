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

/* Binary tree of the median splits (intermediate form). */
struct BinNode {
    Box3 box[2];
    int32_t child[2];  // >= 0: binary interior node, < 0: leaf word
};

struct Binary {
    const std::vector<PrimRef>& refs;
    std::vector<BinNode> nodes;

    // Returns the child word for range [b, e) and its box.
    int32_t emit(size_t b, size_t e, Box3& box)
    {
        if (is_leaf(refs, b, e)) {
            for (size_t i = b; i < e; ++i) box.grow(refs[i].lo, refs[i].hi);
            uint32_t x = ((uint32_t)b << 5) | ((uint32_t)refs[b].kind << 3) | (uint32_t)(e - b);
            return (int32_t)~x;
        }
        size_t idx = nodes.size();
        nodes.emplace_back();
        size_t mid = b + (e - b) / 2;
        BinNode n;
        n.child[0] = emit(b, mid, n.box[0]);
        n.child[1] = emit(mid, e, n.box[1]);
        nodes[idx] = n;
        box.grow(n.box[0].lo, n.box[0].hi);
        box.grow(n.box[1].lo, n.box[1].hi);
        return (int32_t)idx;
    }
};

/* 4-wide nodes: each binary node absorbs its interior children, so a node
 * holds its grandchildren (2 to 4 children) and the tree is half as deep. */
struct Wide {
    const std::vector<BinNode>& bin;
    std::vector<float>& nodes;
    int max_depth = 0;

    // Returns the 4-wide node index for binary interior node b, and the
    // stack entries a traversal below it can need at once (it continues
    // with one hit interior child and defers the others).
    int32_t collapse(int32_t b, int depth, int& need)
    {
        max_depth = std::max(max_depth, depth);
        int32_t word[4];
        const Box3* box[4];
        int m = 0;
        for (int s = 0; s < 2; ++s) {
            int32_t c = bin[b].child[s];
            if (c >= 0) {
                for (int t = 0; t < 2; ++t) {
                    word[m] = bin[c].child[t];
                    box[m] = &bin[c].box[t];
                    ++m;
                }
            } else {
                word[m] = c;
                box[m] = &bin[b].box[s];
                ++m;
            }
        }
        size_t idx = nodes.size() / RT_NODE_FLOATS;
        nodes.resize(nodes.size() + RT_NODE_FLOATS, 0.f);
        int interior = 0, deepest = 0;
        for (int k = 0; k < m; ++k)
            if (word[k] >= 0) {
                int sub = 0;
                word[k] = collapse(word[k], depth + 1, sub);
                deepest = std::max(deepest, sub);
                ++interior;
            }
        need = std::max(0, interior - 1) + deepest;
        float* n = &nodes[idx * RT_NODE_FLOATS];
        for (int k = 0; k < 4; ++k) {
            bool used = k < m;
            for (int a = 0; a < 3; ++a) {
                n[8 * a + k] = used ? box[k]->lo[a] : INFINITY;       // lo of axis a
                n[8 * a + 4 + k] = used ? box[k]->hi[a] : -INFINITY;  // hi of axis a
            }
            int32_t w = used ? word[k] : (int32_t)~(uint32_t)0;  // empty leaf, never hit
            std::memcpy(&n[24 + k], &w, 4);
        }
        return (int32_t)idx;
    }
};

}  // namespace

int build(std::vector<PrimRef>& refs, BuildResult& out)
{
    out.nodes.clear();
    out.depth = 0;
    out.stack_need = 0;
    if (refs.size() >= (size_t(1) << 26)) return -1;
    for (auto& r : refs)
        for (int k = 0; k < 3; ++k) r.c[k] = 0.5f * (r.lo[k] + r.hi[k]);
    if (refs.empty()) return 0;
    partition(refs, 0, refs.size(), 0);
    Binary bt{refs, {}};
    bt.nodes.reserve(refs.size() / 2 + 1);
    Box3 root;
    int32_t r = bt.emit(0, refs.size(), root);
    if (r < 0) {
        // A root that is a leaf still gets an interior node.
        BinNode n;
        n.box[0] = root;
        n.child[0] = r;
        n.child[1] = (int32_t)~(uint32_t)0;  // empty leaf with an empty box
        bt.nodes.push_back(n);
        r = 0;
    }
    Wide wt{bt.nodes, out.nodes};
    int need = 0;
    wt.collapse(r, 0, need);
    out.depth = wt.max_depth + 1;
    out.stack_need = need;
    if (need + 1 > RT_STACK_DEPTH) return -2;
    return 0;
}

}  // namespace rtb
