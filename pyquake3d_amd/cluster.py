"""Cluster tree and block partition of the H-matrix (host side).

Restates the reference's geometric bisection (src/Hmatrix.py:133-182,
``build_block_tree``): a cluster with more than ``leaf`` elements is split at
the midpoint of its bounding box along the longest extent, points equal to the
midpoint going left; a split that leaves one side empty makes a leaf. The
admissibility test is the reference's (src/Hmatrix.py:58-103): a block is
admissible when min(diam_row, diam_col) <= eta * dist(bbox_row, bbox_col).
The block recursion follows ``create_recursive_blocks`` (src/Hmatrix.py:1611-1672).

Differences by design (MI355X storage): the tree is built over a permutation
so that every cluster is a contiguous index range; all device data and
vectors live in that permuted order.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


@dataclass
class ClusterTree:
    perm: np.ndarray          # new index -> original element index
    start: np.ndarray         # node -> first index (permuted order)
    size: np.ndarray          # node -> number of elements
    left: np.ndarray          # node -> left child or -1
    right: np.ndarray         # node -> right child or -1
    lo: np.ndarray            # node -> bbox min (3)
    hi: np.ndarray            # node -> bbox max (3)
    leaf: int = 64

    @property
    def n_nodes(self) -> int:
        return int(self.start.shape[0])

    def is_leaf(self, k: int) -> bool:
        return self.left[k] < 0

    def leaves(self) -> np.ndarray:
        """Leaf nodes in index order."""
        lv = np.where(self.left < 0)[0]
        return lv[np.argsort(self.start[lv], kind="stable")]


def build_cluster_tree(points: np.ndarray, leaf: int = 64) -> ClusterTree:
    pts = np.asarray(points, dtype=np.float64)
    n = pts.shape[0]
    perm = np.arange(n, dtype=np.int64)
    starts, sizes, lefts, rights, los, his = [], [], [], [], [], []

    def new_node(s, m):
        starts.append(s); sizes.append(m); lefts.append(-1); rights.append(-1)
        sub = pts[perm[s:s + m]]
        los.append(sub.min(axis=0)); his.append(sub.max(axis=0))
        return len(starts) - 1

    root = new_node(0, n)
    stack = [root]
    while stack:
        k = stack.pop()
        s, m = starts[k], sizes[k]
        if m <= leaf:
            continue
        idx = perm[s:s + m]
        sub = pts[idx]
        lo, hi = los[k], his[k]
        d = int(np.argmax(hi - lo))
        gamma = (lo[d] + hi[d]) / 2
        mask = sub[:, d] <= gamma
        nl = int(mask.sum())
        if nl == 0 or nl == m:
            continue
        # stable partition keeps the original relative order inside each side
        perm[s:s + m] = np.concatenate([idx[mask], idx[~mask]])
        a = new_node(s, nl)
        b = new_node(s + nl, m - nl)
        lefts[k], rights[k] = a, b
        stack.append(b)
        stack.append(a)
    return ClusterTree(perm, np.array(starts, np.int64), np.array(sizes, np.int64),
                       np.array(lefts, np.int64), np.array(rights, np.int64),
                       np.array(los), np.array(his), leaf)


def admissible(t: ClusterTree, a: int, b: int, eta: float) -> bool:
    da = np.sqrt(((t.hi[a] - t.lo[a]) ** 2).sum())
    db = np.sqrt(((t.hi[b] - t.lo[b]) ** 2).sum())
    gap = np.maximum(0.0, t.lo[b] - t.hi[a]) ** 2 + np.maximum(0.0, t.lo[a] - t.hi[b]) ** 2
    return min(da, db) <= eta * np.sqrt(gap.sum())


@dataclass
class BlockPartition:
    rows: np.ndarray      # leaf block -> row node
    cols: np.ndarray      # leaf block -> col node
    adm: np.ndarray       # leaf block -> admissible (bool)
    tree: ClusterTree = field(repr=False)

    @property
    def n_blocks(self) -> int:
        return int(self.rows.shape[0])


def build_block_partition(t: ClusterTree, eta: float = 2.0,
                          min_size: int | None = None) -> BlockPartition:
    """Leaf blocks of the block tree: recurse over (row, col) node pairs until a
    pair is admissible or either node is a leaf / too small (``min_size``,
    default the tree leaf size)."""
    min_size = t.leaf if min_size is None else min_size
    # precompute diameters for speed
    diam = np.sqrt(((t.hi - t.lo) ** 2).sum(axis=1))
    rows, cols, adm = [], [], []
    stack = [(0, 0)]
    while stack:
        a, b = stack.pop()
        gap = (np.maximum(0.0, t.lo[b] - t.hi[a]) ** 2 +
               np.maximum(0.0, t.lo[a] - t.hi[b]) ** 2).sum()
        is_adm = min(diam[a], diam[b]) <= eta * np.sqrt(gap)
        if (is_adm or t.size[a] <= min_size or t.size[b] <= min_size
                or t.left[a] < 0 or t.left[b] < 0):
            rows.append(a); cols.append(b); adm.append(bool(is_adm))
            continue
        la, ra, lb, rb = t.left[a], t.right[a], t.left[b], t.right[b]
        stack.extend([(ra, rb), (ra, lb), (la, rb), (la, lb)])
    return BlockPartition(np.array(rows, np.int64), np.array(cols, np.int64),
                          np.array(adm, bool), t)
