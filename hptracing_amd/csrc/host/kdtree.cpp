// KD-tree build and flatten with the behaviour of hp::KDTree
// (src/hp/scene/kdtree.cpp).  The tree must be identical to the reference's,
// node for node and bit for bit, because the traversal visits nodes in
// their pre-order and keeps only the first 512 non-empty leaves: a different
// tree gives different hit triangles on long rays.
//
// Kept from the reference:
//   * candidates are every triangle min and max along an axis, scanned in
//     sorted order with the float cost 1 + 3*(L*l/len + R*(1-l/len))
//     (findBestSplit, :47-89), first strict minimum wins, axes tie to the
//     lowest index;
//   * no split below 5 triangles, or when the best cost exceeds 3N, or when
//     the position falls on the box border (split, :93-130);
//   * children inherit the parent box cut at the split; only the root box is
//     fitted (+-1e-3) to the triangles (:36-41, :133-137, :240-241);
//   * a triangle goes left unless all its vertices are > pos, right unless
//     all are < pos (:145-156);
//   * one pass of empty-leaf collapse, children first (:171-206);
//   * pre-order numbering, sibling links on left children, leaf lists
//     [count, ids...] (getData, :248-292).
// Changed, without effect on the output: explicit stacks instead of the
// recursion, the per-node sorted bound arrays released as soon as the
// split is chosen, and the two subtrees of large nodes built on separate
// threads.  count_kdtree walks the same splits keeping only counts.
#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cstring>
#include <functional>
#include <future>
#include <thread>
#include <unordered_map>

#include "host.h"

namespace hpt {
namespace {

struct Node {
    float box_start[3], box_end[3];
    std::vector<int32_t> geo;
    std::vector<float> mins[3], maxs[3];  // sorted per-axis triangle bounds
    std::unique_ptr<Node> left, right;
    Node *parent = nullptr;
};

struct Builder {
    const std::vector<float> &points;  // 4 floats per point
    const std::vector<int32_t> &mesh;  // int4 per triangle

    inline float coord(int32_t p, int d) const { return points[4 * (int64_t)p + d]; }

    // triangleMinMax + calcMinMaxVals (kdtree.cpp:14-34).
    void calc_bounds(Node &n) const {
        for (int d = 0; d < 3; ++d) {
            n.mins[d].resize(n.geo.size());
            n.maxs[d].resize(n.geo.size());
            for (size_t i = 0; i < n.geo.size(); ++i) {
                const int32_t *t = &mesh[4 * (int64_t)n.geo[i]];
                const float a = coord(t[0], d), b = coord(t[1], d), c = coord(t[2], d);
                // std::min_element / max_element: first extreme wins.
                float lo = a, hi = a;
                if (b < lo) lo = b;
                if (c < lo) lo = c;
                if (hi < b) hi = b;
                if (hi < c) hi = c;
                n.mins[d][i] = lo;
                n.maxs[d][i] = hi;
            }
            std::sort(n.mins[d].begin(), n.mins[d].end());
            std::sort(n.maxs[d].begin(), n.maxs[d].end());
        }
    }

    // findBestSplit (kdtree.cpp:47-89).  The reference sorts the
    // concatenation of the two sorted arrays; merging them yields the same
    // sequence of values.
    void best_split(const Node &n, int d, float &best_pos, float &best_cost) const {
        const std::vector<float> &mn = n.mins[d], &mx = n.maxs[d];
        const size_t N = mn.size();
        std::vector<float> all(2 * N);
        std::merge(mn.begin(), mn.end(), mx.begin(), mx.end(), all.begin());
        best_cost = FLT_MAX;
        best_pos = 0;
        size_t imin = 0, imax = 0;
        int left_count = 0, right_count = (int)N;
        const float lo = n.box_start[d], hi = n.box_end[d];
        for (float val : all) {
            if (val <= lo || val >= hi) continue;
            while (imin < N && mn[imin] <= val) {
                ++imin;
                ++left_count;
            }
            while (imax < N && mx[imax] < val) {
                ++imax;
                --right_count;
            }
            const float area_length = hi - lo;
            const float left_length = val - lo;
            const float cost = 1.f + 3.f * ((float)left_count * left_length / area_length +
                                            (float)right_count * (1.f - left_length / area_length));
            if (cost < best_cost) {
                best_cost = cost;
                best_pos = val;
            }
        }
    }

    // The split decision of split() (kdtree.cpp:93-130): false for a leaf
    // (under 5 triangles, best cost above 3N, or a position on the border),
    // else the axis of the first strictly smallest cost and its position.
    bool choose_split(const Node &n, int &axis, float &pos) const {
        if (n.geo.size() < 5) return false;
        const float no_split_cost = 3.f * (float)n.geo.size();
        float p[3], cost[3];
        for (int d = 0; d < 3; ++d) best_split(n, d, p[d], cost[d]);
        int bd = 0;
        for (int d = 1; d < 3; ++d)
            if (cost[d] < cost[bd]) bd = d;
        if (cost[bd] > no_split_cost) return false;
        if (p[bd] <= n.box_start[bd] || p[bd] >= n.box_end[bd]) return false;
        axis = bd;
        pos = p[bd];
        return true;
    }

    // The children of a split (kdtree.cpp:132-156): the parent's box cut at
    // `pos`; a triangle goes left unless all its vertices are > pos, right
    // unless all are < pos.  The parent keeps its box and no triangles.
    void partition(Node &n, int axis, float pos) const {
        n.left.reset(new Node());
        n.right.reset(new Node());
        Node &L = *n.left, &R = *n.right;
        L.parent = R.parent = &n;
        std::memcpy(L.box_start, n.box_start, sizeof(n.box_start));
        std::memcpy(R.box_start, n.box_start, sizeof(n.box_start));
        std::memcpy(L.box_end, n.box_end, sizeof(n.box_end));
        std::memcpy(R.box_end, n.box_end, sizeof(n.box_end));
        L.box_end[axis] = pos;
        R.box_start[axis] = pos;
        for (int32_t g : n.geo) {
            const int32_t *t = &mesh[4 * (int64_t)g];
            const float a = coord(t[0], axis), b = coord(t[1], axis), c = coord(t[2], axis);
            if (!(a > pos && b > pos && c > pos)) L.geo.push_back(g);
            if (!(a < pos && b < pos && c < pos)) R.geo.push_back(g);
        }
        std::vector<int32_t>().swap(n.geo);
    }

    // Bounds, decision and (if split) children of one node; the sorted
    // bounds are released before the children exist.
    bool expand(Node &n) const {
        calc_bounds(n);
        int axis = 0;
        float pos = 0.f;
        const bool split = choose_split(n, axis, pos);
        for (int d = 0; d < 3; ++d) {
            std::vector<float>().swap(n.mins[d]);
            std::vector<float>().swap(n.maxs[d]);
        }
        if (split) partition(n, axis, pos);
        return split;
    }
};

// A node with more triangles than this in its two children hands the right
// child's subtree to a new thread while the thread budget lasts.
constexpr size_t kThreadMin = 20000;

// Builds the subtree under `root` (box and triangles set) with an explicit
// stack instead of the reference's recursion (split, kdtree.cpp:93-169):
// nodes are expanded in the recursion's order (left subtree first), and
// each node's result depends only on its own box and triangles, so the
// tree is the same for every thread count.  `threads`: extra threads this
// subtree may still start.
void build_subtree(const Builder &b, Node *root, int threads) {
    struct Item {
        Node *n;
        int threads;
    };
    std::vector<Item> stack{{root, threads}};
    std::vector<std::thread> workers;
    while (!stack.empty()) {
        const Item it = stack.back();
        stack.pop_back();
        if (!b.expand(*it.n)) continue;
        Node *L = it.n->left.get(), *R = it.n->right.get();
        if (it.threads > 0 && L->geo.size() + R->geo.size() > kThreadMin) {
            const int rest = it.threads - 1;
            workers.emplace_back(build_subtree, std::cref(b), R, rest / 2);
            stack.push_back({L, rest - rest / 2});
        } else {
            stack.push_back({R, 0});
            stack.push_back({L, 0});
        }
    }
    for (std::thread &w : workers) w.join();
}

// The root's box (setBoxSize, kdtree.cpp:36-41): the triangles' extremes
// widened by 1e-3 (the front and back of the sorted bounds are the same
// minimum and maximum).
void fit_root_box(const Builder &b, Node &root) {
    for (int d = 0; d < 3; ++d) {
        float lo = FLT_MAX, hi = -FLT_MAX;
        for (int32_t g : root.geo) {
            const int32_t *t = &b.mesh[4 * (int64_t)g];
            for (int k = 0; k < 3; ++k) {
                lo = std::min(lo, b.coord(t[k], d));
                hi = std::max(hi, b.coord(t[k], d));
            }
        }
        root.box_start[d] = lo - 1e-3f;
        root.box_end[d] = hi + 1e-3f;
    }
}

// Counts of a subtree after the one-pass empty-leaf collapse.
struct Counts {
    int64_t nodes = 0, leaves = 0, entries = 0, max_leaf = 0;
    int depth = 0;
    bool empty_leaf = false;
};

Counts leaf_counts(size_t n) {
    Counts c;
    c.nodes = 1;
    c.leaves = n ? 1 : 0;
    c.entries = (int64_t)n;
    c.max_leaf = (int64_t)n;
    c.empty_leaf = n == 0;
    return c;
}

// removeEmptyNode (kdtree.cpp:171-206) on a node whose children's subtrees
// are already collapsed: a child that is an empty leaf makes the node take
// the other child's contents (box, triangles, children).
Counts combine(const Counts &l, const Counts &r) {
    if (l.empty_leaf) return r;
    if (r.empty_leaf) return l;
    Counts c;
    c.nodes = 1 + l.nodes + r.nodes;
    c.leaves = l.leaves + r.leaves;
    c.entries = l.entries + r.entries;
    c.max_leaf = std::max(l.max_leaf, r.max_leaf);
    c.depth = 1 + std::max(l.depth, r.depth);
    return c;
}

// The counts of the tree build_subtree would make under `root`, without
// keeping it: post-order over an explicit stack, each subtree released as
// soon as it is counted, so memory stays at the triangle lists of the
// nodes still pending on the current path.
Counts count_subtree(const Builder &b, std::unique_ptr<Node> root, int threads) {
    struct Frame {
        std::unique_ptr<Node> n;       // stage 0: the node to expand
        int threads = 0;
        int stage = 0;                 // 1: left pending, 2: right pending
        Counts left;
        std::unique_ptr<Node> right;   // counted here after the left
        std::future<Counts> right_async;
    };
    std::vector<Frame> stack;
    stack.emplace_back();
    stack.back().n = std::move(root);
    stack.back().threads = threads;
    Counts done;
    bool have = false;
    for (;;) {
        if (have) {
            if (stack.empty()) return done;
            Frame &p = stack.back();
            if (p.stage == 1) {
                p.left = done;
                have = false;
                p.stage = 2;
                if (p.right_async.valid()) {
                    done = combine(p.left, p.right_async.get());
                    have = true;
                    stack.pop_back();
                } else {
                    std::unique_ptr<Node> r = std::move(p.right);
                    stack.emplace_back();
                    stack.back().n = std::move(r);
                }
            } else {
                done = combine(p.left, done);
                stack.pop_back();
            }
            continue;
        }
        Frame &f = stack.back();
        Node &n = *f.n;
        if (!b.expand(n)) {
            done = leaf_counts(n.geo.size());
            have = true;
            stack.pop_back();
            continue;
        }
        std::unique_ptr<Node> L = std::move(n.left), R = std::move(n.right);
        L->parent = R->parent = nullptr;
        f.n.reset();
        f.stage = 1;
        int child_threads = 0;
        if (f.threads > 0 && L->geo.size() + R->geo.size() > kThreadMin) {
            const int rest = f.threads - 1;
            f.right_async = std::async(std::launch::async, count_subtree, std::cref(b),
                                       std::move(R), rest / 2);
            child_threads = rest - rest / 2;
        } else {
            f.right = std::move(R);
        }
        stack.emplace_back();
        stack.back().n = std::move(L);
        stack.back().threads = child_threads;
    }
}

inline bool is_empty_leaf(const Node *n) {
    return n && n->geo.empty() && !n->left && !n->right;
}

// removeEmptyNode (kdtree.cpp:171-206) at one node whose children were
// already processed: a child that is an empty leaf makes the node take the
// other child's box, triangles and children.
void collapse(Node &n) {
    if (is_empty_leaf(n.left.get())) {
        std::unique_ptr<Node> keep = std::move(n.right);
        std::memcpy(n.box_start, keep->box_start, sizeof(n.box_start));
        std::memcpy(n.box_end, keep->box_end, sizeof(n.box_end));
        n.geo = keep->geo;
        n.left = std::move(keep->left);
        n.right = std::move(keep->right);
    } else if (is_empty_leaf(n.right.get())) {
        std::unique_ptr<Node> keep = std::move(n.left);
        std::memcpy(n.box_start, keep->box_start, sizeof(n.box_start));
        std::memcpy(n.box_end, keep->box_end, sizeof(n.box_end));
        n.geo = keep->geo;
        n.right = std::move(keep->right);
        n.left = std::move(keep->left);
    } else {
        return;
    }
    if (n.right) n.right->parent = &n;
    if (n.left) n.left->parent = &n;
}

// The reference's one recursive pass, children first, over an explicit
// stack: a node is collapsed after both of its subtrees.
void remove_empty(Node *root) {
    std::vector<std::pair<Node *, bool>> stack{{root, false}};
    while (!stack.empty()) {
        const auto [n, children_done] = stack.back();
        stack.pop_back();
        if (!n->left && !n->right) continue;
        if (children_done) {
            collapse(*n);
            continue;
        }
        stack.push_back({n, true});
        stack.push_back({n->right.get(), false});
        stack.push_back({n->left.get(), false});
    }
}

}  // namespace

void build_kdtree(HostScene &scene, int n_threads) {
    const auto t0 = std::chrono::steady_clock::now();
    const int64_t nt = scene.n_triangles();
    if (nt <= 0) throw Error(HPT_ERR_SCENE, "scene has no triangles");
    if (n_threads <= 0) n_threads = (int)std::max(1u, std::thread::hardware_concurrency());

    Builder b{scene.points, scene.mesh};
    std::unique_ptr<Node> root(new Node());
    root->geo.resize(nt);
    for (int64_t i = 0; i < nt; ++i) root->geo[i] = (int32_t)i;
    fit_root_box(b, *root);
    build_subtree(b, root.get(), n_threads - 1);
    remove_empty(root.get());

    // getData (kdtree.cpp:248-292): pre-order numbering, then one header per
    // node.  An explicit stack keeps deep trees off the call stack.
    std::vector<Node *> order;
    {
        std::vector<Node *> stack{root.get()};
        while (!stack.empty()) {
            Node *n = stack.back();
            stack.pop_back();
            order.push_back(n);
            if (n->right) stack.push_back(n->right.get());
            if (n->left) stack.push_back(n->left.get());
        }
    }
    scene.nodes.clear();
    scene.leaf_data.clear();
    scene.nodes.resize(order.size());
    std::unordered_map<const Node *, int32_t> idx;
    idx.reserve(order.size() * 2);
    for (size_t i = 0; i < order.size(); ++i) idx[order[i]] = (int32_t)i;
    for (size_t i = 0; i < order.size(); ++i) {
        const Node *n = order[i];
        hpt_kd_node h;
        std::memset(&h, 0, sizeof(h));
        h.data = -1;
        if (!n->geo.empty()) {
            h.data = (int32_t)scene.leaf_data.size();
            scene.leaf_data.push_back((int32_t)n->geo.size());
            scene.leaf_data.insert(scene.leaf_data.end(), n->geo.begin(), n->geo.end());
        }
        for (int d = 0; d < 3; ++d) {
            h.box_start[d] = n->box_start[d];
            h.box_end[d] = n->box_end[d];
        }
        h.child = n->left ? idx[n->left.get()] : -1;
        h.parent = n->parent ? idx[n->parent] : -1;
        h.sibling = -1;
        if (n->parent && n->parent->right && n->parent->right.get() != n)
            h.sibling = idx[n->parent->right.get()];
        scene.nodes[i] = h;
    }
    // Release the tree iteratively (a recursive unique_ptr chain could be deep).
    {
        std::vector<std::unique_ptr<Node>> stack;
        stack.push_back(std::move(root));
        while (!stack.empty()) {
            std::unique_ptr<Node> n = std::move(stack.back());
            stack.pop_back();
            if (n->left) stack.push_back(std::move(n->left));
            if (n->right) stack.push_back(std::move(n->right));
        }
    }
    scene.build_seconds =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

KdCounts count_kdtree(const std::vector<float> &points, const std::vector<int32_t> &mesh,
                      int n_threads) {
    const auto t0 = std::chrono::steady_clock::now();
    const int64_t nt = (int64_t)mesh.size() / 4;
    if (nt <= 0) throw Error(HPT_ERR_SCENE, "scene has no triangles");
    if (n_threads <= 0) n_threads = (int)std::max(1u, std::thread::hardware_concurrency());
    Builder b{points, mesh};
    std::unique_ptr<Node> root(new Node());
    root->geo.resize(nt);
    for (int64_t i = 0; i < nt; ++i) root->geo[i] = (int32_t)i;
    fit_root_box(b, *root);
    const Counts c = count_subtree(b, std::move(root), n_threads - 1);
    KdCounts out;
    out.nodes = c.nodes;
    out.leaves = c.leaves;
    out.entries = c.entries;
    out.max_leaf = c.max_leaf;
    out.max_depth = c.depth;
    out.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return out;
}

}  // namespace hpt
