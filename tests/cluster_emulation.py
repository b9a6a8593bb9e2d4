"""numpy stand-ins for one level of the cluster-tree and block-partition
kernels (csrc/tree_kernels.hip), on CPU tensors, for the host-side tests of
the level-by-level drivers. They restate the host builders' rules per node
and per node pair; the GPU tests check the kernels themselves."""
import numpy as np
import torch


def fake_tree_level(pts, perm, tmp, nodes, leaf, box, nl):
    P, pm = pts.numpy(), perm.numpy()
    for q, (s, m) in enumerate(nodes.numpy().tolist()):
        idx = pm[s:s + m].copy()
        sub = P[idx]
        lo, hi = sub.min(0), sub.max(0)
        box[q, :3], box[q, 3:] = torch.from_numpy(lo), torch.from_numpy(hi)
        nl[q] = -1
        if m <= leaf:
            continue
        d = int(np.argmax(hi - lo))
        mask = sub[:, d] <= (lo[d] + hi[d]) / 2
        k = int(mask.sum())
        if 0 < k < m:
            pm[s:s + m] = np.concatenate([idx[mask], idx[~mask]])
            nl[q] = k


def fake_block_level(pairs, node, box, eta, min_size, depth, leafs, nxt, counts):
    nd, bx = node.numpy(), box.numpy()
    for a, b, path in pairs.numpy().tolist():
        lo_a, hi_a, lo_b, hi_b = bx[a, :3], bx[a, 3:], bx[b, :3], bx[b, 3:]
        gap = (np.maximum(0.0, lo_b - hi_a) ** 2 + np.maximum(0.0, lo_a - hi_b) ** 2).sum()
        da = np.sqrt(((hi_a - lo_a) ** 2).sum())
        db = np.sqrt(((hi_b - lo_b) ** 2).sum())
        adm = min(da, db) <= eta * np.sqrt(gap)
        if adm or nd[a, 1] <= min_size or nd[b, 1] <= min_size or nd[a, 2] < 0 or nd[b, 2] < 0:
            r = int(counts[0])
            leafs[r] = torch.tensor([a, b, int(adm), path])
            counts[0] += 1
            continue
        r = int(counts[1])
        kids = [(nd[a, 2], nd[b, 2]), (nd[a, 2], nd[b, 3]), (nd[a, 3], nd[b, 2]),
                (nd[a, 3], nd[b, 3])]
        for c, (x, y) in enumerate(kids):
            nxt[r + c] = torch.tensor([x, y, path | (c << (2 * (30 - depth)))])
        counts[1] += 4
