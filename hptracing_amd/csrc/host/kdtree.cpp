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
// Changed, without effect on the output: the per-node sorted bound arrays
// are released once a node is split, and the two subtrees of large nodes
// are built on separate threads.
#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cstring>
#include <functional>
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

    // split (kdtree.cpp:93-169); `threads` is the number of extra threads
    // this subtree may still start.
    void split(Node &n, int threads) const {
        if (n.geo.size() < 5) return;
        const float no_split_cost = 3.f * (float)n.geo.size();
        float pos[3], cost[3];
        for (int d = 0; d < 3; ++d) best_split(n, d, pos[d], cost[d]);
        int bd = 0;
        for (int d = 1; d < 3; ++d)
            if (cost[d] < cost[bd]) bd = d;
        if (cost[bd] > no_split_cost) return;
        const float p = pos[bd];
        if (p <= n.box_start[bd] || p >= n.box_end[bd]) return;

        n.left.reset(new Node());
        n.right.reset(new Node());
        Node &L = *n.left, &R = *n.right;
        L.parent = R.parent = &n;
        std::memcpy(L.box_start, n.box_start, sizeof(n.box_start));
        std::memcpy(R.box_start, n.box_start, sizeof(n.box_start));
        std::memcpy(L.box_end, n.box_end, sizeof(n.box_end));
        std::memcpy(R.box_end, n.box_end, sizeof(n.box_end));
        L.box_end[bd] = p;
        R.box_start[bd] = p;
        for (int32_t g : n.geo) {
            const int32_t *t = &mesh[4 * (int64_t)g];
            const float a = coord(t[0], bd), b = coord(t[1], bd), c = coord(t[2], bd);
            if (!(a > p && b > p && c > p)) L.geo.push_back(g);
            if (!(a < p && b < p && c < p)) R.geo.push_back(g);
        }
        n.geo.clear();
        n.geo.shrink_to_fit();
        for (int d = 0; d < 3; ++d) {
            std::vector<float>().swap(n.mins[d]);
            std::vector<float>().swap(n.maxs[d]);
        }
        auto build = [this](Node &c, int t) {
            calc_bounds(c);
            split(c, t);
            for (int d = 0; d < 3; ++d) {
                std::vector<float>().swap(c.mins[d]);
                std::vector<float>().swap(c.maxs[d]);
            }
        };
        if (threads > 0 && L.geo.size() + R.geo.size() > 20000) {
            const int rest = threads - 1;
            std::thread worker(build, std::ref(R), rest / 2);
            build(L, rest - rest / 2);
            worker.join();
        } else {
            build(L, 0);
            build(R, 0);
        }
    }
};

inline bool is_empty_leaf(const Node *n) {
    return n && n->geo.empty() && !n->left && !n->right;
}

// removeEmptyNode (kdtree.cpp:171-206).
void remove_empty(Node &n) {
    if (!n.left && !n.right) return;
    remove_empty(*n.left);
    remove_empty(*n.right);
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
    b.calc_bounds(*root);
    // setBoxSize (kdtree.cpp:36-41).
    for (int d = 0; d < 3; ++d) {
        root->box_start[d] = root->mins[d].front() - 1e-3f;
        root->box_end[d] = root->maxs[d].back() + 1e-3f;
    }
    b.split(*root, n_threads - 1);
    remove_empty(*root);

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

}  // namespace hpt
