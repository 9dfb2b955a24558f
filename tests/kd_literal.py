"""Literal Python restatement of the reference's KD build and flatten
(src/hp/scene/kdtree.cpp:14-292), in float32, for small scenes only.

It follows the reference step by step (per-node sorted bound arrays, sort of
the concatenated candidates, recursive split, one collapse pass, pre-order
flatten) so that the optimised C++ builder can be compared with it node for
node.  Test helper, not product code.
"""
from __future__ import annotations

import sys

import numpy as np

f32 = np.float32
FLT_MAX = f32(np.finfo(np.float32).max)


class Node:
    def __init__(self):
        self.left = None
        self.right = None
        self.parent = None
        self.box_start = None
        self.box_end = None
        self.geo = []
        self.mins = [[], [], []]
        self.maxs = [[], [], []]


def build_literal(points, mesh):
    points = np.asarray(points, np.float32)
    mesh = np.asarray(mesh, np.int32)
    sys.setrecursionlimit(100000)

    def tri_minmax(g, d):  # triangleMinMax (:14-21), first extreme wins
        vals = [points[mesh[g][k]][d] for k in range(3)]
        lo = vals[0]
        hi = vals[0]
        for v in vals[1:]:
            if v < lo:
                lo = v
            if hi < v:
                hi = v
        return lo, hi

    def calc(n):  # calcMinMaxVals (:23-34)
        for d in range(3):
            for g in n.geo:
                lo, hi = tri_minmax(g, d)
                n.mins[d].append(lo)
                n.maxs[d].append(hi)
            n.mins[d].sort()
            n.maxs[d].sort()

    def best_split(n, d):  # findBestSplit (:47-89)
        allv = sorted(n.mins[d] + n.maxs[d])
        best_cost, best_pos = FLT_MAX, f32(0)
        i_min = i_max = 0
        left_t, right_t = 0, len(n.geo)
        for val in allv:
            if val <= n.box_start[d] or val >= n.box_end[d]:
                continue
            while i_min < len(n.mins[d]) and n.mins[d][i_min] <= val:
                i_min += 1
                left_t += 1
            while i_max < len(n.maxs[d]) and n.maxs[d][i_max] < val:
                i_max += 1
                right_t -= 1
            area = f32(n.box_end[d] - n.box_start[d])
            ll = f32(val - n.box_start[d])
            cost = f32(1) + f32(3) * f32(f32(f32(left_t) * ll) / area +
                                         f32(f32(right_t) * f32(f32(1) - f32(ll / area))))
            if cost < best_cost:
                best_cost, best_pos = cost, val
        return best_pos, best_cost

    def split(n):  # split (:93-169)
        if len(n.geo) < 5:
            return
        no_split = f32(3) * f32(len(n.geo))
        res = [best_split(n, d) for d in range(3)]
        bd = 0
        for d in (1, 2):
            if res[d][1] < res[bd][1]:
                bd = d
        pos, cost = res[bd]
        if cost > no_split:
            return
        if pos <= n.box_start[bd] or pos >= n.box_end[bd]:
            return
        L, R = Node(), Node()
        L.parent = R.parent = n
        L.box_start, R.box_start = n.box_start.copy(), n.box_start.copy()
        L.box_end, R.box_end = n.box_end.copy(), n.box_end.copy()
        L.box_end[bd] = pos
        R.box_start[bd] = pos
        for g in n.geo:
            c = [points[mesh[g][k]][bd] for k in range(3)]
            if not all(x > pos for x in c):
                L.geo.append(g)
            if not all(x < pos for x in c):
                R.geo.append(g)
        n.geo = []
        n.left, n.right = L, R
        calc(L)
        split(L)
        calc(R)
        split(R)

    def empty_leaf(n):
        return n is not None and not n.geo and n.left is None and n.right is None

    def remove_empty(n):  # removeEmptyNode (:171-206)
        if n.left is None and n.right is None:
            return
        remove_empty(n.left)
        remove_empty(n.right)
        if empty_leaf(n.left):
            k = n.right
        elif empty_leaf(n.right):
            k = n.left
        else:
            return
        n.box_start, n.box_end, n.geo = k.box_start, k.box_end, list(k.geo)
        n.left, n.right = k.left, k.right
        for c in (n.left, n.right):
            if c is not None:
                c.parent = n

    root = Node()
    root.geo = list(range(len(mesh)))
    calc(root)
    root.box_start = np.array([f32(root.mins[d][0] - f32(1e-3)) for d in range(3)], np.float32)
    root.box_end = np.array([f32(root.maxs[d][-1] + f32(1e-3)) for d in range(3)], np.float32)
    split(root)
    remove_empty(root)

    order, index = [], {}

    def walk(n):  # getData (:248-292)
        index[id(n)] = len(order)
        order.append(n)
        if n.left is not None:
            walk(n.left)
        if n.right is not None:
            walk(n.right)

    walk(root)
    nodes, leaf = [], []
    for n in order:
        data = -1
        if n.geo:
            data = len(leaf)
            leaf.append(len(n.geo))
            leaf.extend(n.geo)
        child = index[id(n.left)] if n.left is not None else -1
        parent = index[id(n.parent)] if n.parent is not None else -1
        sib = -1
        if n.parent is not None and n.parent.right is not None and n.parent.right is not n:
            sib = index[id(n.parent.right)]
        nodes.append((n.box_start.copy(), n.box_end.copy(), child, parent, sib, data))
    return nodes, np.array(leaf, np.int32)
