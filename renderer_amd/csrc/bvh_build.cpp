/*
 * bvh_build.cpp — host construction of the flattened BVH (layout in
 * rt_layout.h).
 *
 * The reference builds a pointer tree with one primitive per leaf by
 * sorting the whole list at every level (BVHNode::create, src/bvh.cpp:67-94:
 * longest axis of the list's box, std::sort by box centre, split at n/2).
 * This builder:
 *   - splits by the surface-area heuristic over 32 bins of the box centres
 *     per axis (Wald 2007), falling back to the centre median where no plane
 *     separates the centres, and to the median for every node below a
 *     binary level cap when the SAH tree would need more traversal stack
 *     than RT_STACK_DEPTH (build_best lowers the cap until it fits);
 *   - can also keep the reference's median split (nth_element, O(n) per
 *     level instead of a full sort, longest axis of the centre bounds);
 *   - stops at leaves of at most 4 primitives of one kind;
 *   - builds the top levels on several threads; the result does not depend
 *     on the thread schedule (each subtree is built by one thread into its
 *     own list, appended in a fixed order);
 *   - collapses the binary tree into 4-wide nodes written depth-first.
 * Closest-hit results do not depend on the tree shape except for exact
 * ties in t; only the work per ray does.
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

/* Binary tree (intermediate form). */
struct BinNode {
    Box3 box[2];
    int32_t child[2];  // >= 0: binary interior node, < 0: leaf word
};

float area(const Box3& b)
{
    const float x = b.hi[0] - b.lo[0], y = b.hi[1] - b.lo[1], z = b.hi[2] - b.lo[2];
    return x < 0.f ? 0.f : x * y + y * z + z * x;
}

/* Split point of [b, e) by the median of the box centres along the
 * longest axis of their bounds (ties by shading index). */
size_t split_median(std::vector<PrimRef>& refs, size_t b, size_t e, const float* clo, const float* chi)
{
    const float dx = chi[0] - clo[0], dy = chi[1] - clo[1], dz = chi[2] - clo[2];
    const int axis = dx > dy && dx > dz ? 0 : (dy > dz ? 1 : 2);
    const size_t mid = b + (e - b) / 2;
    std::nth_element(refs.begin() + b, refs.begin() + mid, refs.begin() + e,
                     [axis](const PrimRef& a, const PrimRef& c) {
                         if (a.c[axis] != c.c[axis]) return a.c[axis] < c.c[axis];
                         return a.g < c.g;
                     });
    return mid;
}

/* Split point of [b, e) by the surface-area heuristic over SAH_BINS bins of
 * the centre bounds per axis (Wald 2007): the plane minimising
 * area(left) * ceil(n_left / 4) + area(right) * ceil(n_right / 4) (leaves
 * hold up to four records, so a record count costs in steps of four).
 * Records keep their relative order within each side.  Returns b when no
 * plane separates the centres. */
constexpr int SAH_BINS = 32;
size_t split_sah(std::vector<PrimRef>& refs, size_t b, size_t e, const float* clo, const float* chi)
{
    float best = INFINITY;
    int best_axis = -1, best_plane = 0;
    for (int axis = 0; axis < 3; ++axis) {
        const float ext = chi[axis] - clo[axis];
        if (!(ext > 0.f)) continue;
        const float scale = SAH_BINS / ext;
        Box3 bin[SAH_BINS];
        size_t cnt[SAH_BINS] = {};
        for (size_t i = b; i < e; ++i) {
            int k = std::min(SAH_BINS - 1, (int)((refs[i].c[axis] - clo[axis]) * scale));
            bin[k].grow(refs[i].lo, refs[i].hi);
            ++cnt[k];
        }
        float right_area[SAH_BINS];
        size_t right_cnt[SAH_BINS];
        Box3 acc;
        size_t n = 0;
        for (int k = SAH_BINS - 1; k > 0; --k) {
            acc.grow(bin[k].lo, bin[k].hi);
            n += cnt[k];
            right_area[k] = area(acc);
            right_cnt[k] = n;
        }
        Box3 left;
        size_t nl = 0;
        for (int k = 0; k < SAH_BINS - 1; ++k) {
            left.grow(bin[k].lo, bin[k].hi);
            nl += cnt[k];
            const size_t nr = right_cnt[k + 1];
            if (nl == 0 || nr == 0) continue;
            const float cost = area(left) * (float)((nl + 3) / 4) + right_area[k + 1] * (float)((nr + 3) / 4);
            if (cost < best) {
                best = cost;
                best_axis = axis;
                best_plane = k + 1;
            }
        }
    }
    if (best_axis < 0) return b;
    const int axis = best_axis;
    const float lo = clo[axis], scale = SAH_BINS / (chi[axis] - clo[axis]);
    auto it = std::stable_partition(refs.begin() + b, refs.begin() + e, [&](const PrimRef& r) {
        return std::min(SAH_BINS - 1, (int)((r.c[axis] - lo) * scale)) < best_plane;
    });
    return (size_t)(it - refs.begin());
}

/* Builds the binary tree of [b, e) into `out` (depth-first, parent first)
 * and returns its child word and box.  The first levels of large ranges
 * build their second half on another thread into a list of their own,
 * appended after the first half's nodes, so the result does not depend on
 * the thread schedule. */
struct Builder {
    std::vector<PrimRef>& refs;
    int method, sah_levels;

    int32_t emit(size_t b, size_t e, int level, std::vector<BinNode>& out, Box3& box)
    {
        if (is_leaf(refs, b, e)) {
            for (size_t i = b; i < e; ++i) box.grow(refs[i].lo, refs[i].hi);
            uint32_t x = ((uint32_t)b << 5) | ((uint32_t)refs[b].kind << 3) | (uint32_t)(e - b);
            return (int32_t)~x;
        }
        float clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (size_t i = b; i < e; ++i)
            for (int k = 0; k < 3; ++k) {
                clo[k] = std::min(clo[k], refs[i].c[k]);
                chi[k] = std::max(chi[k], refs[i].c[k]);
            }
        size_t mid = b;
        if (e - b <= RT_LEAF_MAX) {
            /* a small range of mixed kinds: one leaf per kind */
            std::stable_sort(refs.begin() + b, refs.begin() + e,
                             [](const PrimRef& x, const PrimRef& y) { return x.kind < y.kind; });
            mid = b + 1;
            while (refs[mid].kind == refs[b].kind) ++mid;
        } else {
            if (method == BUILD_SAH && level < sah_levels) mid = split_sah(refs, b, e, clo, chi);
            if (mid == b || mid == e) mid = split_median(refs, b, e, clo, chi);
        }
        const size_t idx = out.size();
        out.emplace_back();
        BinNode n;
        if (level < 4 && e - b > 65536) {
            std::vector<BinNode> second;
            int32_t c1 = 0;
            std::thread t([&] { c1 = emit(mid, e, level + 1, second, n.box[1]); });
            n.child[0] = emit(b, mid, level + 1, out, n.box[0]);
            t.join();
            const int32_t off = (int32_t)out.size();
            for (BinNode m : second) {
                for (int s = 0; s < 2; ++s)
                    if (m.child[s] >= 0) m.child[s] += off;
                out.push_back(m);
            }
            n.child[1] = c1 >= 0 ? c1 + off : c1;
        } else {
            n.child[0] = emit(b, mid, level + 1, out, n.box[0]);
            n.child[1] = emit(mid, e, level + 1, out, n.box[1]);
        }
        out[idx] = n;
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

/* Renumbers the 4-wide nodes so that the first levels come first, breadth
 * first (RT_TOP_NODES of them at most), and the others keep the depth-first
 * order of the collapse.  Child words follow
 * the new numbering; a traversal visits the same nodes in the same order. */
void top_levels_first(std::vector<float>& nodes)
{
    const size_t n = nodes.size() / RT_NODE_FLOATS;
    auto word = [&](const std::vector<float>& v, size_t node, int k) {
        int32_t w;
        std::memcpy(&w, &v[node * RT_NODE_FLOATS + 24 + k], 4);
        return w;
    };
    std::vector<int32_t> order{0};
    for (size_t i = 0; i < order.size() && order.size() < RT_TOP_NODES; ++i)
        for (int k = 0; k < 4 && order.size() < RT_TOP_NODES; ++k) {
            const int32_t w = word(nodes, (size_t)order[i], k);
            if (w >= 0) order.push_back(w);
        }
    std::vector<int32_t> idx(n, -1);
    for (size_t j = 0; j < order.size(); ++j) idx[(size_t)order[j]] = (int32_t)j;
    int32_t next = (int32_t)order.size();
    for (size_t i = 0; i < n; ++i)
        if (idx[i] < 0) idx[i] = next++;
    std::vector<float> out(nodes.size());
    for (size_t i = 0; i < n; ++i) {
        float* dst = &out[(size_t)idx[i] * RT_NODE_FLOATS];
        std::memcpy(dst, &nodes[i * RT_NODE_FLOATS], RT_NODE_FLOATS * sizeof(float));
        for (int k = 0; k < 4; ++k) {
            const int32_t w = word(nodes, i, k);
            if (w >= 0) std::memcpy(&dst[24 + k], &idx[(size_t)w], 4);
        }
    }
    nodes.swap(out);
}

}  // namespace

int build(std::vector<PrimRef>& refs, BuildResult& out, int method, int sah_levels)
{
    out.nodes.clear();
    out.depth = 0;
    out.stack_need = 0;
    if (refs.size() >= (size_t(1) << 26)) return -1;
    for (auto& r : refs)
        for (int k = 0; k < 3; ++k) r.c[k] = 0.5f * (r.lo[k] + r.hi[k]);
    if (refs.empty()) return 0;
    Builder builder{refs, method, sah_levels};
    std::vector<BinNode> bin;
    bin.reserve(refs.size() / 2 + 1);
    Box3 root;
    int32_t r = builder.emit(0, refs.size(), 0, bin, root);
    if (r < 0) {
        // A root that is a leaf still gets an interior node.
        BinNode n;
        n.box[0] = root;
        n.child[0] = r;
        n.child[1] = (int32_t)~(uint32_t)0;  // empty leaf with an empty box
        bin.push_back(n);
        r = 0;
    }
    Wide wt{bin, out.nodes};
    int need = 0;
    wt.collapse(r, 0, need);
    top_levels_first(out.nodes);
    out.depth = wt.max_depth + 1;
    out.stack_need = need;
    if (need + 1 > RT_STACK_DEPTH) return -2;
    return 0;
}

int build_best(std::vector<PrimRef>& refs, BuildResult& out)
{
    for (int levels : {64, 32, 24, 16, 8}) {
        const int r = build(refs, out, BUILD_SAH, levels);
        if (r != -2) return r;
    }
    return build(refs, out, BUILD_MEDIAN);
}

}  // namespace rtb
