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

``build_cluster_tree_device`` / ``build_block_partition_device`` build the
same tree and partition on the GPU, one level per launch
(csrc/tree_kernels.hip), bit for bit equal to the host functions here; the
H-matrix uses them. Device trees number their nodes level by level (root 0,
the children of the j-th split node at 2j + 1 and 2j + 2); ranges, boxes and
the block order are those of the host functions.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib


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


# ------------------------------------------------------------------ device
def _tree_level(pts, perm, tmp, nodes, leaf, box, nl):
    _lib.call("pq3d_tree_level", _lib.ptr(pts), int(pts.shape[0]), _lib.ptr(perm),
              _lib.ptr(tmp), _lib.ptr(nodes), int(nodes.shape[0]), int(leaf), _lib.ptr(box),
              _lib.ptr(nl), _lib.stream_handle())


def _block_level(pairs, node, box, eta, min_size, depth, leafs, nxt, counts):
    _lib.call("pq3d_block_level", _lib.ptr(pairs), int(pairs.shape[0]), _lib.ptr(node),
              _lib.ptr(box), float(eta), int(min_size), int(depth), _lib.ptr(leafs),
              _lib.ptr(nxt), _lib.ptr(counts), _lib.stream_handle())


def build_cluster_tree_device(points: np.ndarray, leaf: int = 64, device="cuda") -> ClusterTree:
    """``build_cluster_tree`` on the GPU: one launch per tree level (each
    node's box, split and stable partition of the permutation); the host
    reads back only the level's boxes and left-child sizes."""
    dev = torch.device(device)
    pts = torch.from_numpy(np.ascontiguousarray(points, dtype=np.float64)).to(dev)
    if pts.ndim != 2 or pts.shape[1] != 3:
        raise ValueError("points must be (n, 3)")
    n = int(pts.shape[0])
    if n == 0:
        raise ValueError("no points")
    perm = torch.arange(n, dtype=torch.int64, device=dev)
    tmp = torch.empty_like(perm)
    level = np.array([[0, n]], dtype=np.int64)
    parts = []
    while len(level):
        nodes = torch.from_numpy(level).to(dev)
        box = torch.empty((len(level), 6), dtype=torch.float64, device=dev)
        nl = torch.empty(len(level), dtype=torch.int64, device=dev)
        _tree_level(pts, perm, tmp, nodes, leaf, box, nl)
        b, k = box.cpu().numpy(), nl.cpu().numpy()
        parts.append((level, b, k))
        sp = k > 0
        s, m, l = level[sp, 0], level[sp, 1], k[sp]
        level = np.stack([np.stack([s, l], 1), np.stack([s + l, m - l], 1)],
                         axis=1).reshape(-1, 2)
    start = np.concatenate([p[0][:, 0] for p in parts])
    size = np.concatenate([p[0][:, 1] for p in parts])
    box = np.concatenate([p[1] for p in parts])
    split = np.concatenate([p[2] for p in parts]) > 0
    left = np.full(len(start), -1, dtype=np.int64)
    right = np.full(len(start), -1, dtype=np.int64)
    j = np.arange(int(split.sum()), dtype=np.int64)
    left[split], right[split] = 2 * j + 1, 2 * j + 2
    return ClusterTree(perm.cpu().numpy(), start, size, left, right,
                       np.ascontiguousarray(box[:, :3]), np.ascontiguousarray(box[:, 3:]), leaf)


def build_block_partition_device(t: ClusterTree, eta: float = 2.0, min_size: int | None = None,
                                 device="cuda") -> BlockPartition:
    """``build_block_partition`` on the GPU: one launch per level of the block
    tree; leaf blocks carry their path from the root and are put in the
    recursion's depth-first order by sorting the paths."""
    min_size = t.leaf if min_size is None else min_size
    dev = torch.device(device)
    node = torch.from_numpy(np.ascontiguousarray(
        np.stack([t.start, t.size, t.left, t.right], 1).astype(np.int64))).to(dev)
    box = torch.from_numpy(np.ascontiguousarray(np.concatenate([t.lo, t.hi], 1))).to(dev)
    pairs = torch.zeros((1, 3), dtype=torch.int64, device=dev)   # (root, root), empty path
    counts = torch.zeros(2, dtype=torch.int64, device=dev)
    out = []
    depth = 0
    while pairs.shape[0]:
        P = int(pairs.shape[0])
        leafs = torch.empty((P, 4), dtype=torch.int64, device=dev)
        nxt = torch.empty((4 * P, 3), dtype=torch.int64, device=dev)
        counts.zero_()
        _block_level(pairs, node, box, eta, min_size, depth, leafs, nxt, counts)
        c = counts.cpu().numpy()
        if not (0 <= c[0] <= P and 0 <= c[1] <= 4 * P):
            raise RuntimeError("block level counts out of range")
        out.append(leafs[:int(c[0])])
        pairs = nxt[:int(c[1])]
        depth += 1
    L = torch.cat(out).cpu().numpy()
    L = L[np.argsort(L[:, 3], kind="stable")]
    return BlockPartition(np.ascontiguousarray(L[:, 0]), np.ascontiguousarray(L[:, 1]),
                          L[:, 2].astype(bool), t)
